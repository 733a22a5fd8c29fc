! harness_basis.f90 -- builds the reference's spd basis descriptor from a text
! orbit list (line 1: nsp; then "n l 2j 2tz con" per orbit), replacing
! read_sp_basis/find_holes (basic_IMSRG.f90:152-400, which read .sps and
! hole_scheme files) with the same side effects: hb4/pb4, holes/parts,
! Jtotal_max, lmax, the 6j table (store_6j) and the (l,j,tz) state lists.
! TEST INFRASTRUCTURE ONLY (used by ref_harness and the drop-in test).
module harness_basis
  use basic_IMSRG
  implicit none
contains
  subroutine setup_basis(fname, jbas)
    character(len=*) :: fname
    type(spd) :: jbas
    integer :: ix, i, r, q, tzi, li, ji, r1, r2
    open(unit=39, file=fname, status='old')
    read(39,*) ix
    jbas%total_orbits = ix
    allocate(jbas%nn(ix), jbas%ll(ix), jbas%jj(ix), jbas%itzp(ix), jbas%nshell(ix))
    allocate(jbas%e(ix), jbas%con(ix), hb4(ix), pb4(ix))
    do i = 1, ix
       read(39,*) jbas%nn(i), jbas%ll(i), jbas%jj(i), jbas%itzp(i), jbas%con(i)
       jbas%nshell(i) = 2*jbas%nn(i) + jbas%ll(i)
       jbas%e(i) = jbas%nshell(i)
    end do
    close(39)
    ! find_holes side effects, basic_IMSRG.f90:370-397
    allocate(jbas%holes(sum(jbas%con)))
    allocate(jbas%parts(jbas%total_orbits - sum(jbas%con)))
    r1 = 1 ; r2 = 1
    do i = 1, jbas%total_orbits
       hb4(i) = sum(jbas%con(1:i-1))
       pb4(i) = i - hb4(i) - 1
       if (jbas%con(i) == 1) then
          jbas%holes(r1) = i ; r1 = r1 + 1
       else
          jbas%parts(r2) = i ; r2 = r2 + 1
       end if
    end do
    ! read_sp_basis tail, basic_IMSRG.f90:304-340
    jbas%Jtotal_max = maxval(jbas%jj)
    jbas%lmax = maxval(jbas%ll)
    call store_6j(jbas, 'n')
    jbas%spblocks = (jbas%Jtotal_max + 1) * (jbas%lmax + 1)
    allocate(jbas%states(jbas%spblocks))
    q = 1
    do tzi = -1, 1, 2
       do li = 0, jbas%lmax
          do ji = 1, jbas%Jtotal_max, 2
             r = 0
             do i = 1, jbas%total_orbits
                if (jbas%jj(i) .ne. ji) cycle
                if (jbas%ll(i) .ne. li) cycle
                if (jbas%itzp(i) .ne. tzi) cycle
                r = r + 1
             end do
             allocate(jbas%states(q)%Z(r))
             r = 0
             do i = 1, jbas%total_orbits
                if (jbas%jj(i) .ne. ji) cycle
                if (jbas%ll(i) .ne. li) cycle
                if (jbas%itzp(i) .ne. tzi) cycle
                r = r + 1
                jbas%states(q)%Z(r) = i
             end do
             q = q + 1
          end do
       end do
    end do
  end subroutine setup_basis
end module harness_basis
