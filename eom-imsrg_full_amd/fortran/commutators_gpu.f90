! commutators_gpu.f90 -- link-time substitute for the reference's
! `module commutators` (n-parzuchowski/EOM-IMSRG_FULL, im-SRG/commutators.f90).
!
! A program that links this object in place of the reference's
! commutators.o, and compiles its callers against this module file, runs
! every scalar commutator on the MI355X through libimsrg_mi355x.so; the
! callers -- IMSRG_MAGNUS.f90 (BCH_EXPAND, MAGNUS_EXPAND, magnus_decouple,
! transform_sq_op_BCH), generators.f90 (build_gs_w2), IMSRG_ODE.f90 (dHds) --
! compile unmodified.  Public names, argument lists and contracts are the
! reference's (commutators.f90:10, 36, 83, 283, 399, 435, 581, 771):
!
!   commutator_110, commutator_220   return the 0-body terms
!   commutator_111, commutator_122   overwrite the one- / two-body part of RES
!   commutator_121, commutator_221   accumulate into the one-body part
!   commutator_222_pp_hh             accumulates into the two-body part; the
!                                    221 one-body term of the same pp/hh
!                                    product is kept for the commutator_221
!                                    call that passes the same w1, w2
!   commutator_222_ph                accumulates the ph term from the Pandya
!                                    matrices LCC, RCC the caller made with
!                                    calculate_cross_coupled
!
! Like the reference module it re-exports module cross_coupled (the Pandya
! storage and calculate_cross_coupled stay the reference's, since other
! modules -- TS_commutators, operators, EOM -- share them).  w1, w2 and WCC are
! kept in the argument lists; the device keeps its own intermediates and
! workspace, and w1/w2 are not filled.  commutator_223_single (the triples
! correction of restore_triples) is not provided: calling it stops.
!
! Operators cross the boundary as the reference's vectorize/repackage streams
! (basic_IMSRG.f90:3794-3918); errors turn into ERROR STOP, as the reference
! STOPs.  The external generators gpu_build_gs_white, gpu_build_gs_atan and
! gpu_build_gs_w2 (after the module) can be passed wherever the reference takes
! `external :: build_generator` (IMSRG_MAGNUS.f90:29).
module commutators
  use iso_c_binding
  use cross_coupled
  implicit none

  public :: commutator_110, commutator_111, commutator_121, commutator_122, commutator_220, &
       commutator_221, commutator_222_pp_hh, commutator_222_ph, commutator_223_single, gpu_generator

  integer(c_int), parameter, private :: T110 = 1, T111 = 2, T121 = 4, T122 = 8, T220 = 16, T221 = 32, &
       T222PPHH = 64, T222PH = 128
  type(c_ptr), save, private :: sub_ctx = c_null_ptr, hA = c_null_ptr, hB = c_null_ptr, hT = c_null_ptr
  integer(c_int64_t), save, private :: sub_neq = 0, sub_ncc = 0
  integer, save, private :: sub_nsp = -1
  real(c_double), allocatable, save, private :: vbuf(:), tbuf(:), rbuf(:), ccl(:), ccr(:)
  ! the 221 one-body term of the last commutator_222_pp_hh call, valid for the
  ! commutator_221 call that passes the same w1 and w2
  real(c_double), allocatable, save, private :: ob221(:)
  type(c_ptr), save, private :: tok_w1 = c_null_ptr, tok_w2 = c_null_ptr
  logical, save, private :: have221 = .false.

  private :: sub_init, sub_check, sub_push, sub_term, sub_merge, sub_flatten_cc

  interface
     function c_last_error() bind(C, name='imsrg_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_last_error
     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
     integer(c_int) function c_ctx_create(nsp, n, l, jj, tz, con, device, out) bind(C, name='imsrg_ctx_create')
       import :: c_int, c_ptr
       integer(c_int), value :: nsp, device
       integer(c_int) :: n(*), l(*), jj(*), tz(*), con(*)
       type(c_ptr) :: out
     end function c_ctx_create
     integer(c_int) function c_ctx_destroy(c) bind(C, name='imsrg_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: c
     end function c_ctx_destroy
     integer(c_int64_t) function c_ctx_neq(c) bind(C, name='imsrg_ctx_neq')
       import :: c_int64_t, c_ptr
       type(c_ptr), value :: c
     end function c_ctx_neq
     integer(c_int64_t) function c_cc_size(c) bind(C, name='imsrg_cross_coupled_size')
       import :: c_int64_t, c_ptr
       type(c_ptr), value :: c
     end function c_cc_size
     integer(c_int) function c_op_create(c, herm, out) bind(C, name='imsrg_op_create')
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       integer(c_int), value :: herm
       type(c_ptr) :: out
     end function c_op_create
     integer(c_int) function c_op_set_herm(op, herm) bind(C, name='imsrg_op_set_herm')
       import :: c_int, c_ptr
       type(c_ptr), value :: op
       integer(c_int), value :: herm
     end function c_op_set_herm
     integer(c_int) function c_op_upload(op, v, n) bind(C, name='imsrg_op_upload')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: op
       real(c_double) :: v(*)
       integer(c_int64_t), value :: n
     end function c_op_upload
     integer(c_int) function c_op_download(op, v, n) bind(C, name='imsrg_op_download')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: op
       real(c_double) :: v(*)
       integer(c_int64_t), value :: n
     end function c_op_download
     integer(c_int) function c_commutator(c, L, R, RES, mask) bind(C, name='imsrg_commutator')
       import :: c_int, c_ptr
       type(c_ptr), value :: c, L, R, RES
       integer(c_int), value :: mask
     end function c_commutator
     integer(c_int) function c_comm110(L, R, out) bind(C, name='imsrg_commutator_110')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: L, R
       real(c_double) :: out
     end function c_comm110
     integer(c_int) function c_comm220(L, R, out) bind(C, name='imsrg_commutator_220')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: L, R
       real(c_double) :: out
     end function c_comm220
     integer(c_int) function c_comm222ph_cc(c, lcc, hl, rcc, hr, n, RES) bind(C, name='imsrg_commutator_222_ph_cc')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: c, RES
       real(c_double) :: lcc(*), rcc(*)
       integer(c_int), value :: hl, hr
       integer(c_int64_t), value :: n
     end function c_comm222ph_cc
     integer(c_int) function c_generator(c, kind, H, ETA) bind(C, name='imsrg_generator')
       import :: c_int, c_ptr
       type(c_ptr), value :: c, H, ETA
       integer(c_int), value :: kind
     end function c_generator
  end interface
  private :: c_last_error, c_strlen, c_ctx_create, c_ctx_destroy, c_ctx_neq, c_cc_size, c_op_create, &
       c_op_set_herm, c_op_upload, c_op_download, c_commutator, c_comm110, c_comm220, c_comm222ph_cc, c_generator

contains

  subroutine sub_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*,*) 'commutators (MI355X): ', what, ' failed (', rc, '): ', msg(1:n)
    error stop 'commutators (MI355X) failure'
  end subroutine sub_check

  ! one device context per basis (made on the first call, remade when the
  ! basis changes) and three device operators reused by every call
  subroutine sub_init(jbas)
    type(spd) :: jbas
    integer(c_int), allocatable :: n(:), l(:), jj(:), tz(:), con(:)
    if (c_associated(sub_ctx) .and. sub_nsp == jbas%total_orbits) return
    if (c_associated(sub_ctx)) call sub_check(c_ctx_destroy(sub_ctx), 'imsrg_ctx_destroy')
    n = jbas%nn ; l = jbas%ll ; jj = jbas%jj ; tz = jbas%itzp ; con = jbas%con
    call sub_check(c_ctx_create(int(jbas%total_orbits, c_int), n, l, jj, tz, con, 0_c_int, sub_ctx), &
         'imsrg_ctx_create')
    sub_nsp = jbas%total_orbits
    sub_neq = c_ctx_neq(sub_ctx)
    sub_ncc = c_cc_size(sub_ctx)
    call sub_check(c_op_create(sub_ctx, 1_c_int, hA), 'imsrg_op_create')
    call sub_check(c_op_create(sub_ctx, 1_c_int, hB), 'imsrg_op_create')
    call sub_check(c_op_create(sub_ctx, 1_c_int, hT), 'imsrg_op_create')
    if (allocated(vbuf)) deallocate(vbuf, tbuf, rbuf, ccl, ccr)
    allocate(vbuf(sub_neq), tbuf(sub_neq), rbuf(sub_neq), ccl(max(sub_ncc, 1_c_int64_t)), &
         ccr(max(sub_ncc, 1_c_int64_t)))
    have221 = .false.
  end subroutine sub_init

  ! op -> device operator h (vectorize stream; op%neq restored afterwards)
  subroutine sub_push(op, h)
    type(sq_op) :: op
    type(c_ptr) :: h
    integer :: keep
    keep = op%neq
    op%neq = int(sub_neq)
    call vectorize(op, vbuf)
    op%neq = keep
    call sub_check(c_op_set_herm(h, int(op%herm, c_int)), 'imsrg_op_set_herm')
    call sub_check(c_op_upload(h, vbuf, sub_neq), 'imsrg_op_upload')
  end subroutine sub_push

  ! tbuf = the vectorize stream of the terms in mask of [L, R] (herm hres)
  subroutine sub_term(L, R, hres, mask)
    type(sq_op) :: L, R
    integer, intent(in) :: hres
    integer(c_int), intent(in) :: mask
    call sub_push(L, hA)
    call sub_push(R, hB)
    call sub_check(c_op_set_herm(hT, int(hres, c_int)), 'imsrg_op_set_herm')
    call sub_check(c_commutator(sub_ctx, hA, hB, hT, mask), 'imsrg_commutator')
    call sub_check(c_op_download(hT, tbuf, sub_neq), 'imsrg_op_download')
  end subroutine sub_term

  ! RES one-body (part 1: fhh, fpp, fph, stream entries 2..1+n1) or two-body
  ! (part 2: the rest) overwritten with / increased by the same entries of t
  subroutine sub_merge(RES, t, part, overwrite)
    type(sq_op) :: RES
    real(c_double), intent(in) :: t(:)
    integer, intent(in) :: part
    logical, intent(in) :: overwrite
    integer(c_int64_t) :: n1, i0, i1
    integer :: keep, nh, np
    nh = RES%belowEF
    np = RES%Nsp - nh
    n1 = int(nh, c_int64_t) * nh + int(np, c_int64_t) * np + int(np, c_int64_t) * nh
    if (part == 1) then
       i0 = 2 ; i1 = 1 + n1
    else
       i0 = 2 + n1 ; i1 = sub_neq
    end if
    keep = RES%neq
    RES%neq = int(sub_neq)
    call vectorize(RES, rbuf)
    if (overwrite) then
       rbuf(i0:i1) = t(i0:i1)
    else
       rbuf(i0:i1) = rbuf(i0:i1) + t(i0:i1)
    end if
    call repackage(RES, rbuf)
    RES%neq = keep
  end subroutine sub_merge

  ! a cc_mat as the stream imsrg_cross_coupled writes: column-major blocks
  subroutine sub_flatten_cc(CC, v)
    type(cc_mat) :: CC
    real(c_double), intent(out) :: v(:)
    integer :: q, m
    integer(c_int64_t) :: k
    k = 0
    do q = 1, CC%nblocks
       m = CC%rlen(q) * CC%nph(q)
       if (m == 0) cycle
       v(k+1:k+m) = reshape(CC%CCX(q)%X, [m])
       k = k + m
    end do
    if (k /= sub_ncc) error stop 'commutators (MI355X): cc_mat does not match the basis'
  end subroutine sub_flatten_cc

  real(8) function commutator_110(L, R, jbas)          ! commutators.f90:10-33
    type(sq_op) :: L, R
    type(spd) :: jbas
    real(c_double) :: x
    call sub_init(jbas)
    call sub_push(L, hA)
    call sub_push(R, hB)
    call sub_check(c_comm110(hA, hB, x), 'imsrg_commutator_110')
    commutator_110 = x
  end function commutator_110

  real(8) function commutator_220(L, R, jbas)          ! commutators.f90:399-432
    type(sq_op) :: L, R
    type(spd) :: jbas
    real(c_double) :: x
    call sub_init(jbas)
    call sub_push(L, hA)
    call sub_push(R, hB)
    call sub_check(c_comm220(hA, hB, x), 'imsrg_commutator_220')
    commutator_220 = x
  end function commutator_220

  subroutine commutator_111(L, R, RES, jbas)            ! commutators.f90:36-80, overwrites one-body
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call sub_init(jbas)
    call sub_term(L, R, RES%herm, T111)
    call sub_merge(RES, tbuf, 1, .true.)
  end subroutine commutator_111

  subroutine commutator_121(L, R, RES, jbas)            ! commutators.f90:83-280, accumulates one-body
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call sub_init(jbas)
    call sub_term(L, R, RES%herm, T121)
    call sub_merge(RES, tbuf, 1, .false.)
  end subroutine commutator_121

  subroutine commutator_122(L, R, RES, jbas)            ! commutators.f90:283-396, overwrites two-body
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call sub_init(jbas)
    call sub_term(L, R, RES%herm, T122)
    call sub_merge(RES, tbuf, 2, .true.)
  end subroutine commutator_122

  ! commutators.f90:581-768: one pp/hh product gives the 222_pp_hh two-body term
  ! (added to RES now) and the 221 one-body term (kept for commutator_221)
  subroutine commutator_222_pp_hh(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES
    type(sq_op), target :: w1, w2
    type(spd) :: jbas
    integer(c_int64_t) :: n1
    call sub_init(jbas)
    call sub_term(L, R, RES%herm, ior(T222PPHH, T221))
    call sub_merge(RES, tbuf, 2, .false.)
    n1 = int(RES%belowEF, c_int64_t)**2 + int(RES%Nsp - RES%belowEF, c_int64_t)**2 + &
         int(RES%belowEF, c_int64_t) * (RES%Nsp - RES%belowEF)
    if (allocated(ob221)) deallocate(ob221)
    allocate(ob221(sub_neq))
    ob221 = 0.d0
    ob221(2:1+n1) = tbuf(2:1+n1)
    tok_w1 = c_loc(w1)
    tok_w2 = c_loc(w2)
    have221 = .true.
  end subroutine commutator_222_pp_hh

  ! commutators.f90:435-578: the one-body term of the pp/hh product of the
  ! commutator_222_pp_hh call that passed these w1, w2; without one, the
  ! product is formed again from L and R
  subroutine commutator_221(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES
    type(sq_op), target :: w1, w2
    type(spd) :: jbas
    call sub_init(jbas)
    if (have221 .and. c_associated(tok_w1, c_loc(w1)) .and. c_associated(tok_w2, c_loc(w2))) then
       call sub_merge(RES, ob221, 1, .false.)
    else
       call sub_term(L, R, RES%herm, T221)
       call sub_merge(RES, tbuf, 1, .false.)
    end if
    have221 = .false.
  end subroutine commutator_221

  ! commutators.f90:771-925 on the caller's Pandya matrices; WCC unused
  subroutine commutator_222_ph(LCC, RCC, RES, WCC, jbas)
    type(cc_mat) :: LCC, RCC, WCC
    type(sq_op) :: RES
    type(spd) :: jbas
    call sub_init(jbas)
    call sub_flatten_cc(LCC, ccl)
    call sub_flatten_cc(RCC, ccr)
    call sub_check(c_op_set_herm(hT, int(RES%herm, c_int)), 'imsrg_op_set_herm')
    call sub_check(c_comm222ph_cc(sub_ctx, ccl, int(LCC%herm, c_int), ccr, int(RCC%herm, c_int), sub_ncc, hT), &
         'imsrg_commutator_222_ph_cc')
    call sub_check(c_op_download(hT, tbuf, sub_neq), 'imsrg_op_download')
    call sub_merge(RES, tbuf, 2, .false.)
  end subroutine commutator_222_ph

  ! commutators.f90:928-1285 (the triples correction of restore_triples) has
  ! no device counterpart
  real(8) function commutator_223_single(L, R, ip, iq, ir, is, it, iu, Jtot, jpq, jst, jbas)
    type(sq_op) :: L, R
    integer, intent(in) :: ip, iq, ir, is, it, iu, Jtot, jpq, jst
    type(spd) :: jbas
    commutator_223_single = 0.d0
    error stop 'commutators (MI355X): commutator_223_single (triples restoration) is not provided'
  end function commutator_223_single

  ! device generator `kind` (0 white, 1 atan, 2 w2; generators.f90:4-160) of H
  subroutine gpu_generator(kind, H, ETA, jbas)
    integer, intent(in) :: kind
    type(sq_op) :: H, ETA
    type(spd) :: jbas
    integer :: keep
    call sub_init(jbas)
    call sub_push(H, hA)
    call sub_check(c_op_set_herm(hT, -1_c_int), 'imsrg_op_set_herm')
    call sub_check(c_generator(sub_ctx, int(kind, c_int), hA, hT), 'imsrg_generator')
    call sub_check(c_op_download(hT, tbuf, sub_neq), 'imsrg_op_download')
    ETA%herm = -1
    keep = ETA%neq
    ETA%neq = int(sub_neq)
    call repackage(ETA, tbuf)
    ETA%neq = keep
  end subroutine gpu_generator

end module commutators

! The device generators as external procedures with the reference's
! build_generator signature (H, ETA, jbas), for magnus_decouple,
! decouple_hamiltonian and dHds (IMSRG_MAGNUS.f90:29; IMSRG_ODE.f90:416).
subroutine gpu_build_gs_white(H, ETA, jbas)               ! build_gs_white, generators.f90:4-106
  use commutators, only: sq_op, spd, gpu_generator
  type(sq_op) :: H, ETA
  type(spd) :: jbas
  call gpu_generator(0, H, ETA, jbas)
end subroutine gpu_build_gs_white

subroutine gpu_build_gs_atan(H, ETA, jbas)                ! build_gs_atan, generators.f90:163-266
  use commutators, only: sq_op, spd, gpu_generator
  type(sq_op) :: H, ETA
  type(spd) :: jbas
  call gpu_generator(1, H, ETA, jbas)
end subroutine gpu_build_gs_atan

subroutine gpu_build_gs_w2(H, ETA, jbas)                  ! build_gs_w2, generators.f90:109-160
  use commutators, only: sq_op, spd, gpu_generator
  type(sq_op) :: H, ETA
  type(spd) :: jbas
  call gpu_generator(2, H, ETA, jbas)
end subroutine gpu_build_gs_w2
