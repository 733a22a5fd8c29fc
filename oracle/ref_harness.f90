! ref_harness.f90 -- drives the reference's own compiled FP64 routines
! (n-parzuchowski/EOM-IMSRG_FULL, im-SRG/*.f90, built from the sources where
! they lie by oracle/ref.mk into oracle/_ref/).
!
! TEST INFRASTRUCTURE ONLY.  It writes golden vectors and times the reference
! CPU path; the product library never links it.
!
! Operators travel as raw FP64 streams in the reference's vectorize order
! (basic_IMSRG.f90:3794-3854).  The basis comes from a text file:
!   line 1: nsp ;  then nsp lines "n l 2j 2tz con"  (con=1 for holes)
! which replaces read_sp_basis/find_holes (basic_IMSRG.f90:152-400) and
! reproduces their side effects (hb4/pb4, holes/parts, states, store_6j).
!
! usage: ref_harness <basis> <command> args...
!   terms  hL L.bin hR R.bin outprefix   -> per-term outputs of [L,R]
!   comm   hL L.bin hR R.bin out.bin     -> full commutator incl. E0
!   cc     h  A.bin out.bin              -> Pandya matrices of A (all blocks)
!   bch    G.bin H.bin out.bin           -> BCH_EXPAND(HS,G,H)
!   magexp DG.bin G.bin out.bin          -> MAGNUS_EXPAND(DG,G,AD)
!   gen    kind H.bin out.bin            -> generator (white|atan|w2)
!   normal H.bin out.bin                 -> normal_order(H)
!   scal   h A.bin                       -> prints frob norm and mbpt2
!   flow   kind HS.bin maxsteps outprefix -> magnus_decouple loop, full precision
!   brute  hA A.bin hB B.bin points.txt out.txt -> brute-force J-coupled sums
!   time   hL L.bin hR R.bin reps        -> wall seconds per full commutator
!   hf     H.bin out.bin D.bin [O.bin Oout.bin] -> calc_HF (no three-body force) on the bare H;
!          D(old orbit, new orbit) as an nsp x nsp column-major stream; prints
!          E_HF; with O, also sqop_to_HF(O) with those coefficients
!   decouple kind HS.bin outprefix      -> magnus_decouple itself (IMSRG_MAGNUS.f90:15-128)
!   dhds   kind HS.bin out.bin          -> dHds (IMSRG_ODE.f90:416-463): vectorize([eta(HS), HS])
!   readint fmt file out.bin hw com_hw beta A htype emax_file lmax_file
!                                        -> the reference's interaction reader fmt (ascii, gz, me2j,
!          me2b) on file into a zero operator (interaction_IO.f90), written as a stream
!   tensor hG G.bin rank dpar hT seed outprefix -> a random tensor T (construct_random_rankX,
!          brute_force_testing.f90:123-265; rank and dpar doubled) written to <outprefix>_T.bin,
!          then each TS_commutator term of [G, T] alone (<outprefix>_<term>.bin) and the
!          whole scalar-tensor commutator in BCH_TENSOR's call order (<outprefix>_full.bin)
!   tbch   hG G.bin rank dpar hT T.bin out.bin -> BCH_TENSOR(G, T) (IMSRG_MAGNUS.f90:389-465)
!   ttime  hG G.bin rank dpar hT T.bin reps   -> wall seconds per scalar-tensor commutator
! Tensor operators travel as raw FP64 streams: E0, fhh, fpp, fph, then for every
! tensor block (allocate_tensor's order, basic_IMSRG.f90:644-907) its nine
! sub-blocks tgam(1..9), each column-major.
!
! Built twice by oracle/ref.mk: ref_harness against the reference's own
! module commutators, and sub_harness (-DSUBSTITUTE) against the product's
! substitute for it (eom-imsrg_full_amd/fortran/commutators_gpu.f90), with
! IMSRG_MAGNUS, generators and IMSRG_ODE compiled unmodified against that;
! sub_harness also takes the device generators gpu_white, gpu_atan, gpu_w2.
program ref_harness
  use basic_IMSRG
  use cross_coupled
  use commutators
  use IMSRG_MAGNUS
  use TS_commutators
  use brute_force_testing
  use HF_mod
  use interaction_IO
  use harness_basis
  implicit none

  type(spd) :: jbas
  type(sq_op) :: HS
  character(len=1024) :: basisfile, cmd
  external :: build_gs_white, build_gs_w2, build_gs_atan, dHds
#ifdef SUBSTITUTE
  external :: gpu_build_gs_white, gpu_build_gs_w2, gpu_build_gs_atan
#endif

  call getarg(1, basisfile)
  call getarg(2, cmd)
  checkpointing = .false.
  OUTPUT_DIR = './'
  prefix = 'ref_harness'

  call setup_basis(trim(basisfile), jbas)
  call allocate_blocks(jbas, HS)
  HS%herm = 1
  HS%neq = compute_neq(HS)
  HS%hospace = 20.d0
  HS%lawson_beta = 0.d0
  HS%com_hw = 0.d0
  HS%eMax = maxval(jbas%nshell)

  select case (trim(cmd))
  case ('terms')
     call do_terms()
  case ('comm')
     call do_comm()
  case ('cc')
     call do_cc()
  case ('bch')
     call do_bch()
  case ('magexp')
     call do_magexp()
  case ('gen')
     call do_gen()
  case ('normal')
     call do_normal()
  case ('scal')
     call do_scal()
  case ('flow')
     call do_flow()
  case ('brute')
     call do_brute()
  case ('time')
     call do_time()
  case ('hf')
     call do_hf()
  case ('decouple')
     call do_decouple()
  case ('dhds')
     call do_dhds()
  case ('readint')
     call do_readint()
  case ('tensor')
     call do_tensor()
  case ('tbch')
     call do_tbch()
  case ('ttime')
     call do_ttime()
  case default
     print*, 'unknown command ', trim(cmd)
     stop 2
  end select

contains

  integer function compute_neq(op)
    ! IMSRG_ODE.f90:36-44
    type(sq_op) :: op
    integer :: A, N, q, nh, np, nb
    A = op%belowEF
    N = op%Nsp
    compute_neq = 1 + A*A + A*(N-A) + (N-A)**2
    do q = 1, op%nblocks
       nh = op%mat(q)%nhh ; np = op%mat(q)%npp ; nb = op%mat(q)%nph
       compute_neq = compute_neq + (nh*nh+nh + nb*nb+nb + np*np+np)/2 + nb*np + nh*np + nh*nb
    end do
  end function compute_neq


  subroutine read_op(fname, op)
    character(len=*) :: fname
    type(sq_op) :: op
    real(8), allocatable :: v(:)
    allocate(v(HS%neq))
    open(unit=41, file=fname, access='stream', form='unformatted', status='old')
    read(41) v
    close(41)
    call repackage(op, v)
  end subroutine read_op

  subroutine write_op(fname, op)
    character(len=*) :: fname
    type(sq_op) :: op
    real(8), allocatable :: v(:)
    allocate(v(HS%neq))
    call vectorize(op, v)
    open(unit=42, file=fname, access='stream', form='unformatted', status='replace')
    write(42) v
    close(42)
  end subroutine write_op

  integer function iarg(k)
    integer :: k
    character(len=64) :: s
    call getarg(k, s)
    read(s,*) iarg
  end function iarg

  character(len=1024) function sarg(k)
    integer :: k
    call getarg(k, sarg)
  end function sarg

  subroutine load_pair(L, R)
    type(sq_op) :: L, R
    call duplicate_sq_op(HS, L)
    call duplicate_sq_op(HS, R)
    L%herm = iarg(3)
    call read_op(trim(sarg(4)), L)
    R%herm = iarg(5)
    call read_op(trim(sarg(6)), R)
  end subroutine load_pair

  subroutine do_terms()
    type(sq_op) :: L, R, RES, w1, w2
    type(cc_mat) :: LCC, RCC, WCC
    character(len=1024) :: pfx
    real(8) :: e110, e220
    call load_pair(L, R)
    pfx = sarg(7)
    call duplicate_sq_op(HS, RES)
    call duplicate_sq_op(HS, w1)
    call duplicate_sq_op(HS, w2)
    RES%herm = -1 * L%herm * R%herm
    call init_ph_mat(L, LCC, jbas)
    call duplicate_ph_mat(LCC, RCC)
    call init_ph_wkspc(LCC, WCC)
    e110 = commutator_110(L, R, jbas)
    e220 = commutator_220(L, R, jbas)
    open(unit=43, file=trim(pfx)//'_e0.txt', status='replace')
    write(43,'(2es26.17)') e110, e220
    close(43)
    call clear_sq_op(RES) ; call commutator_111(L, R, RES, jbas)
    call write_op(trim(pfx)//'_111.bin', RES)
    call clear_sq_op(RES) ; call commutator_121(L, R, RES, jbas)
    call write_op(trim(pfx)//'_121.bin', RES)
    call clear_sq_op(RES) ; call commutator_122(L, R, RES, jbas)
    call write_op(trim(pfx)//'_122.bin', RES)
    call clear_sq_op(RES) ; call commutator_222_pp_hh(L, R, RES, w1, w2, jbas)
    call write_op(trim(pfx)//'_222pphh.bin', RES)
    call clear_sq_op(RES) ; call commutator_221(L, R, RES, w1, w2, jbas)
    call write_op(trim(pfx)//'_221.bin', RES)
    call calculate_cross_coupled(L, LCC, jbas)
    call calculate_cross_coupled(R, RCC, jbas)
    call clear_sq_op(RES) ; call commutator_222_ph(LCC, RCC, RES, WCC, jbas)
    call write_op(trim(pfx)//'_222ph.bin', RES)
  end subroutine do_terms

  subroutine full_comm(L, R, RES)
    ! the call sequence of IMSRG_MAGNUS.f90:207-219
    type(sq_op) :: L, R, RES, w1, w2
    type(cc_mat) :: LCC, RCC, WCC
    call duplicate_sq_op(HS, w1)
    call duplicate_sq_op(HS, w2)
    call init_ph_mat(L, LCC, jbas)
    call duplicate_ph_mat(LCC, RCC)
    call init_ph_wkspc(LCC, WCC)
    call calculate_cross_coupled(R, RCC, jbas)
    call calculate_cross_coupled(L, LCC, jbas)
    RES%E0 = commutator_110(L, R, jbas) + commutator_220(L, R, jbas)
    call commutator_111(L, R, RES, jbas)
    call commutator_121(L, R, RES, jbas)
    call commutator_122(L, R, RES, jbas)
    call commutator_222_pp_hh(L, R, RES, w1, w2, jbas)
    call commutator_221(L, R, RES, w1, w2, jbas)
    call commutator_222_ph(LCC, RCC, RES, WCC, jbas)
  end subroutine full_comm

  subroutine do_comm()
    type(sq_op) :: L, R, RES
    call load_pair(L, R)
    call duplicate_sq_op(HS, RES)
    RES%herm = -1 * L%herm * R%herm
    call full_comm(L, R, RES)
    call write_op(trim(sarg(7)), RES)
  end subroutine do_comm

  subroutine do_cc()
    type(sq_op) :: A
    type(cc_mat) :: ACC
    integer :: q
    call duplicate_sq_op(HS, A)
    A%herm = iarg(3)
    call read_op(trim(sarg(4)), A)
    call init_ph_mat(A, ACC, jbas)
    call calculate_cross_coupled(A, ACC, jbas)
    open(unit=44, file=trim(sarg(5)), access='stream', form='unformatted', status='replace')
    do q = 1, ACC%nblocks
       if (ACC%rlen(q) * ACC%nph(q) == 0) cycle
       write(44) ACC%CCX(q)%X
    end do
    close(44)
  end subroutine do_cc

  subroutine do_bch()
    type(sq_op) :: G, H0, HOUT
    call duplicate_sq_op(HS, G)
    call duplicate_sq_op(HS, H0)
    call duplicate_sq_op(HS, HOUT)
    G%herm = -1
    H0%herm = 1
    call read_op(trim(sarg(3)), G)
    call read_op(trim(sarg(4)), H0)
    call BCH_EXPAND(HOUT, G, H0, jbas, 'n')
    call write_op(trim(sarg(5)), HOUT)
  end subroutine do_bch

  subroutine do_magexp()
    type(sq_op) :: DG, G, AD
    call duplicate_sq_op(HS, DG)
    call duplicate_sq_op(HS, G)
    call duplicate_sq_op(HS, AD)
    DG%herm = -1
    G%herm = -1
    call read_op(trim(sarg(3)), DG)
    call read_op(trim(sarg(4)), G)
    call MAGNUS_EXPAND(DG, G, AD, jbas)
    call write_op(trim(sarg(5)), DG)
  end subroutine do_magexp

  subroutine run_gen(kind, H, ETA)
    character(len=*) :: kind
    type(sq_op) :: H, ETA
    select case (kind)
    case ('white')
       call build_gs_white(H, ETA, jbas)
    case ('atan')
       call build_gs_atan(H, ETA, jbas)
    case ('w2')
       call build_gs_w2(H, ETA, jbas)
#ifdef SUBSTITUTE
    case ('gpu_white')
       call gpu_build_gs_white(H, ETA, jbas)
    case ('gpu_atan')
       call gpu_build_gs_atan(H, ETA, jbas)
    case ('gpu_w2')
       call gpu_build_gs_w2(H, ETA, jbas)
#endif
    case default
       stop 'bad generator'
    end select
  end subroutine run_gen

  subroutine do_decouple()
    ! magnus_decouple (IMSRG_MAGNUS.f90:15-128) with the generator passed as
    ! its external build_generator; writes H(s), Omega(s) and the step lines
    ! it prints into <prefix>_0b_magnus_flow.dat
    type(sq_op) :: G
    integer :: neq0
    neq0 = HS%neq
    HS%herm = 1
    call read_op(trim(sarg(4)), HS)
    prefix = trim(sarg(5))
    select case (trim(sarg(3)))
    case ('white')
       call magnus_decouple(HS, G, jbas, 'n', 'n', build_gs_white)
    case ('w2')
       call magnus_decouple(HS, G, jbas, 'n', 'n', build_gs_w2)
#ifdef SUBSTITUTE
    case ('gpu_white')
       call magnus_decouple(HS, G, jbas, 'n', 'n', gpu_build_gs_white)
    case ('gpu_w2')
       call magnus_decouple(HS, G, jbas, 'n', 'n', gpu_build_gs_w2)
#endif
    case default
       stop 'bad generator'
    end select
    HS%neq = neq0
    call write_op(trim(sarg(5))//'_HS.bin', HS)
    call write_op(trim(sarg(5))//'_G.bin', G)
  end subroutine do_decouple

  subroutine do_readint()
    ! the reference's readers with their globals set: TBME_DIR/intfile name
    ! the file, H carries hw, com_hw, the Lawson beta and A; for me2j the
    ! file's basis jbx is the hk list up to emax_file / lmax_file (what
    ! read_sp_basis builds from the file's .sps), jbx%con its orbit's index
    ! in jbas
    type(spd) :: jbx
    integer :: htype, ef, lf, e, l, jj, tz, ix, i
    TBME_DIR = ''
    intfile = trim(sarg(4))
    HS%herm = 1
    call clear_sq_op(HS)
    HS%hospace = rarg(6)
    HS%com_hw = rarg(7)
    HS%lawson_beta = rarg(8)
    HS%Aprot = iarg(9)
    HS%Aneut = 0
    htype = iarg(10)
    ef = iarg(11)
    lf = iarg(12)
    HS%lmax = maxval(jbas%ll)
    select case (trim(sarg(3)))
    case ('ascii')
       call read_interaction(HS, jbas, htype)
    case ('gz')
       call read_gz(HS, jbas, htype)
    case ('me2j')
       ix = 0
       do e = 0, ef
          do l = mod(e, 2), min(e, lf), 2
             do jj = abs(2*l - 1), 2*l + 1, 2
                do tz = -1, 1, 2
                   ix = ix + 1
                end do
             end do
          end do
       end do
       jbx%total_orbits = ix
       allocate(jbx%nn(ix), jbx%ll(ix), jbx%jj(ix), jbx%itzp(ix), jbx%con(ix))
       ix = 0
       do e = 0, ef
          do l = mod(e, 2), min(e, lf), 2
             do jj = abs(2*l - 1), 2*l + 1, 2
                do tz = -1, 1, 2
                   ix = ix + 1
                   jbx%nn(ix) = (e - l) / 2 ; jbx%ll(ix) = l ; jbx%jj(ix) = jj ; jbx%itzp(ix) = tz
                   jbx%con(ix) = 0
                   do i = 1, jbas%total_orbits
                      if (jbas%nn(i) == jbx%nn(ix) .and. jbas%ll(i) == l .and. jbas%jj(i) == jj .and. &
                           jbas%itzp(i) == tz) jbx%con(ix) = i
                   end do
                end do
             end do
          end do
       end do
       call read_me2j_interaction(HS, jbas, jbx, htype)
    case ('me2b')
       call read_me2b_interaction(HS, jbas, htype)
    case default
       stop 'bad interaction format'
    end select
    call write_op(trim(sarg(5)), HS)
  end subroutine do_readint

  subroutine write_tensor(fname, T)
    character(len=*) :: fname
    type(sq_op) :: T
    integer :: q, g
    open(unit=48, file=fname, access='stream', form='unformatted', status='replace')
    write(48) T%E0
    write(48) T%fhh
    write(48) T%fpp
    write(48) T%fph
    do q = 1, T%nblocks
       do g = 1, 9
          if (size(T%tblck(q)%tgam(g)%X) > 0) write(48) T%tblck(q)%tgam(g)%X
       end do
    end do
    close(48)
  end subroutine write_tensor

  subroutine read_tensor(fname, T)
    character(len=*) :: fname
    type(sq_op) :: T
    integer :: q, g
    open(unit=49, file=fname, access='stream', form='unformatted', status='old')
    read(49) T%E0
    read(49) T%fhh
    read(49) T%fpp
    read(49) T%fph
    do q = 1, T%nblocks
       do g = 1, 9
          if (size(T%tblck(q)%tgam(g)%X) > 0) read(49) T%tblck(q)%tgam(g)%X
       end do
    end do
    close(49)
  end subroutine read_tensor

  subroutine tensor_setup(G, T)
    ! G (scalar, args 3-4) and the shape of T (rank, dpar in args 5-6; herm arg 7);
    ! xmap_tensor / half6j slot 1 as in test_scalar_tensor_commutator
    ! (brute_force_testing.f90:912-960) and its caller
    type(sq_op) :: G, T
    integer :: N
    call duplicate_sq_op(HS, G)
    G%herm = iarg(3)
    call read_op(trim(sarg(4)), G)
    T%rank = iarg(5)
    T%dpar = iarg(6)
    T%xindx = 1
    T%pphh_ph = .false.
    N = jbas%total_orbits
    allocate(jbas%xmap_tensor(1, N*(N+1)/2))
    allocate(half6j(1))
    call allocate_tensor(jbas, T, HS)
    T%herm = iarg(7)
    T%E0 = 0.d0
  end subroutine tensor_setup

  subroutine scalar_tensor_comm(G, T, GCC, TCC, RES, w1, w2)
    ! the call order of BCH_TENSOR (IMSRG_MAGNUS.f90:427-446)
    type(sq_op) :: G, T, RES, w1, w2
    type(cc_mat) :: GCC
    type(pandya_mat) :: TCC
    call TS_commutator_111(G, T, RES, jbas)
    call TS_commutator_121(G, T, RES, jbas)
    call TS_commutator_122(G, T, RES, jbas)
    call TS_commutator_212(G, T, RES, jbas)
    call TS_commutator_222_pp_hh(G, T, RES, w1, w2, jbas)
    call TS_commutator_221(w1, w2, G%herm*T%herm, RES, jbas)
    call TS_commutator_211(GCC, T, RES, jbas)
    call TS_commutator_222_ph(GCC, TCC, T, RES, jbas)
  end subroutine scalar_tensor_comm

  subroutine do_tensor()
    type(sq_op) :: G, T, RES, w1, w2
    type(cc_mat) :: GCC
    type(pandya_mat) :: TCC
    character(len=1024) :: pfx
    integer :: ns, i
    integer, allocatable :: sd(:)
    call tensor_setup(G, T)
    call random_seed(size=ns)
    allocate(sd(ns))
    sd = [(iarg(8) * 7919 + 104729 * i, i = 1, ns)]
    call random_seed(put=sd)
    call construct_random_rankX(T, iarg(7), jbas)
    pfx = sarg(9)
    call write_tensor(trim(pfx)//'_T.bin', T)
    call duplicate_sq_op(T, RES)
    call duplicate_sq_op(T, w1)
    call duplicate_sq_op(T, w2)
    RES%herm = -1 * G%herm * T%herm
    call init_ph_mat(G, GCC, jbas)
    call allocate_small_tensor_CCMAT(T, TCC, jbas)
    call calculate_cross_coupled(G, GCC, jbas)
    call clear_sq_op(RES) ; call TS_commutator_111(G, T, RES, jbas)
    call write_tensor(trim(pfx)//'_111.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_121(G, T, RES, jbas)
    call write_tensor(trim(pfx)//'_121.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_211(GCC, T, RES, jbas)
    call write_tensor(trim(pfx)//'_211.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_122(G, T, RES, jbas)
    call write_tensor(trim(pfx)//'_122.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_212(G, T, RES, jbas)
    call write_tensor(trim(pfx)//'_212.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_222_pp_hh(G, T, RES, w1, w2, jbas)
    call write_tensor(trim(pfx)//'_222pphh.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_221(w1, w2, G%herm*T%herm, RES, jbas)
    call write_tensor(trim(pfx)//'_221.bin', RES)
    call clear_sq_op(RES) ; call TS_commutator_222_ph(GCC, TCC, T, RES, jbas)
    call write_tensor(trim(pfx)//'_222ph.bin', RES)
    call clear_sq_op(RES)
    call scalar_tensor_comm(G, T, GCC, TCC, RES, w1, w2)
    call write_tensor(trim(pfx)//'_full.bin', RES)
  end subroutine do_tensor

  subroutine do_tbch()
    type(sq_op) :: G, T
    call tensor_setup(G, T)
    call read_tensor(trim(sarg(8)), T)
    call BCH_TENSOR(G, T, jbas, 'n')
    call write_tensor(trim(sarg(9)), T)
  end subroutine do_tbch

  subroutine do_ttime()
    type(sq_op) :: G, T, RES, w1, w2
    type(cc_mat) :: GCC
    type(pandya_mat) :: TCC
    integer :: reps, k
    real(8) :: t0, t1, omp_get_wtime
    call tensor_setup(G, T)
    call read_tensor(trim(sarg(8)), T)
    reps = iarg(9)
    call duplicate_sq_op(T, RES)
    call duplicate_sq_op(T, w1)
    call duplicate_sq_op(T, w2)
    RES%herm = -1 * G%herm * T%herm
    call init_ph_mat(G, GCC, jbas)
    call allocate_small_tensor_CCMAT(T, TCC, jbas)
    t0 = omp_get_wtime()
    do k = 1, reps
       call clear_sq_op(RES)
       call calculate_cross_coupled(G, GCC, jbas)
       call scalar_tensor_comm(G, T, GCC, TCC, RES, w1, w2)
    end do
    t1 = omp_get_wtime()
    write(*,'(A,es16.8)') 'SECONDS_PER_TENSOR_COMMUTATOR ', (t1 - t0) / reps
  end subroutine do_ttime

  real(8) function rarg(k)
    integer :: k
    character(len=64) :: s
    call getarg(k, s)
    read(s,*) rarg
  end function rarg

  subroutine do_dhds()
    ! dHds (IMSRG_ODE.f90:416-463) at HS with the generator `kind`
    real(8) :: t
    real(8), allocatable :: yy(:), yp(:)
    HS%herm = 1
    call read_op(trim(sarg(4)), HS)
    allocate(yy(HS%neq), yp(HS%neq))
    call vectorize(HS, yy)
    t = 0.d0
    select case (trim(sarg(3)))
    case ('white')
       call dHds(t, yy, yp, HS, jbas, build_gs_white)
    case ('w2')
       call dHds(t, yy, yp, HS, jbas, build_gs_w2)
#ifdef SUBSTITUTE
    case ('gpu_white')
       call dHds(t, yy, yp, HS, jbas, gpu_build_gs_white)
#endif
    case default
       stop 'bad generator'
    end select
    open(unit=42, file=trim(sarg(5)), access='stream', form='unformatted', status='replace')
    write(42) yp
    close(42)
  end subroutine do_dhds

  subroutine do_gen()
    type(sq_op) :: H, ETA
    call duplicate_sq_op(HS, H)
    call duplicate_sq_op(HS, ETA)
    H%herm = 1
    call read_op(trim(sarg(4)), H)
    ETA%herm = -1
    call run_gen(trim(sarg(3)), H, ETA)
    call write_op(trim(sarg(5)), ETA)
  end subroutine do_gen

  subroutine do_normal()
    type(sq_op) :: H
    call duplicate_sq_op(HS, H)
    H%herm = 1
    call read_op(trim(sarg(3)), H)
    call normal_order(H, jbas)
    call write_op(trim(sarg(4)), H)
  end subroutine do_normal

  subroutine do_scal()
    type(sq_op) :: A
    call duplicate_sq_op(HS, A)
    A%herm = iarg(3)
    call read_op(trim(sarg(4)), A)
    write(*,'(2es26.17)') mat_frob_norm(A), mbpt2(A, jbas)
  end subroutine do_scal

  subroutine do_hf()
    ! calc_HF (HF_mod.f90:14-106) with an unallocated three-body force, and
    ! sqop_to_HF (HF_mod.f90:108-136) for a scalar observable
    type(sq_op) :: H, O
    type(three_body_force) :: tb
    type(full_sp_block_mat) :: coefs
    real(8), allocatable :: dfull(:,:)
    integer :: q, i, j, n
    call duplicate_sq_op(HS, H)
    H%herm = 1
    call read_op(trim(sarg(3)), H)
    call calc_HF(H, tb, jbas, coefs)
    write(*,'(A,es26.17)') 'EHF ', H%E0
    call write_op(trim(sarg(4)), H)
    n = jbas%total_orbits
    allocate(dfull(n, n))
    dfull = 0.d0
    do q = 1, coefs%blocks
       do i = 1, coefs%map(q)
          do j = 1, coefs%map(q)
             dfull(coefs%blkM(q)%states(i), coefs%blkM(q)%states(j)) = coefs%blkM(q)%matrix(i, j)
          end do
       end do
    end do
    open(unit=43, file=trim(sarg(5)), form='unformatted', access='stream', status='replace')
    write(43) dfull
    close(43)
    if (len_trim(sarg(6)) > 0) then
       call duplicate_sq_op(HS, O)
       O%herm = 1
       call read_op(trim(sarg(6)), O)
       call sqop_to_HF(O, coefs, jbas)
       call write_op(trim(sarg(7)), O)
    end if
  end subroutine do_hf

  subroutine do_flow()
    ! body of magnus_decouple, IMSRG_MAGNUS.f90:15-128, with full-precision output
    type(sq_op) :: H, G, AD, DG, H0
    character(len=64) :: kind
    integer :: steps, maxsteps
    real(8) :: s, ds, E_mbpt2, crit
    kind = sarg(3)
    maxsteps = iarg(5)
    call duplicate_sq_op(HS, H0)
    H0%herm = 1
    call read_op(trim(sarg(4)), H0)
    call copy_sq_op(H0, HS)
    call duplicate_sq_op(HS, H)
    call duplicate_sq_op(HS, G)
    call duplicate_sq_op(HS, AD)
    call duplicate_sq_op(HS, DG)
    G%herm = -1
    DG%herm = -1
    call run_gen(trim(kind), HS, DG)
    call copy_sq_op(HS, H)
    s = 0.d0
    ds = 0.5d0
    steps = 0
    E_mbpt2 = mbpt2(HS, jbas)
    crit = abs(E_mbpt2)
    open(unit=45, file=trim(sarg(6))//'_steps.txt', status='replace')
    do while (crit > 1e-6)
       if (steps >= maxsteps) exit
       steps = steps + 1
       call MAGNUS_EXPAND(DG, G, AD, jbas)
       call euler_step(G, DG, s, ds)
       call BCH_EXPAND(HS, G, H, jbas, 'n')
       call run_gen(trim(kind), HS, DG)
       E_mbpt2 = mbpt2(HS, jbas)
       crit = abs(E_mbpt2)
       write(45,'(I6,4es26.17)') steps, s, HS%E0, HS%E0 + E_mbpt2, crit
       flush(45)
    end do
    close(45)
    call write_op(trim(sarg(6))//'_HS.bin', HS)
    call write_op(trim(sarg(6))//'_G.bin', G)
  end subroutine do_flow

  subroutine do_brute()
    type(sq_op) :: A, B
    integer :: u, kind, a1, b1, c1, d1, J, ios
    real(8) :: val
    call load_pair(A, B)
    open(unit=46, file=trim(sarg(7)), status='old')
    open(unit=47, file=trim(sarg(8)), status='replace')
    do
       read(46,*,iostat=ios) kind, a1, b1, c1, d1, J
       if (ios /= 0) exit
       select case (kind)
       case (0)
          val = scalar_scalar_0body_comm(A, B, jbas)
       case (1)
          val = scalar_scalar_1body_comm(A, B, a1, b1, jbas)
       case default
          val = scalar_scalar_2body_comm(A, B, a1, b1, c1, d1, J, jbas)
       end select
       write(47,'(6I6,es26.17)') kind, a1, b1, c1, d1, J, val
    end do
    close(46)
    close(47)
    u = 0
  end subroutine do_brute

  subroutine do_time()
    type(sq_op) :: L, R, RES
    integer :: reps, k
    real(8) :: t0, t1, omp_get_wtime
    call load_pair(L, R)
    reps = iarg(7)
    call duplicate_sq_op(HS, RES)
    RES%herm = -1 * L%herm * R%herm
    t0 = omp_get_wtime()
    do k = 1, reps
       call full_comm(L, R, RES)
    end do
    t1 = omp_get_wtime()
    write(*,'(A,es16.8)') 'SECONDS_PER_COMMUTATOR ', (t1 - t0) / reps
  end subroutine do_time

end program ref_harness
