! drop_in_test.f90 -- the Fortran drop-in (imsrg_gpu) against the reference's
! own routines on the same sq_op inputs.  Built by oracle/ref.mk into
! oracle/_ref/drop_in_test from the reference objects; run by
! tests/test_gpu_fortran_drop_in.py on the GPU.
!   drop_in_test <basis> <L.bin herm -1> <R.bin herm 1> <Hn.bin normal-ordered H> <H0.bin bare H>
! Prints "CHECK <name> <relative error>" lines; exit status 1 on any > 1e-10.
program drop_in_test
  use basic_IMSRG
  use cross_coupled
  use commutators
  use IMSRG_MAGNUS
  use harness_basis
  use HF_mod
  use imsrg_gpu
  implicit none
  type(spd) :: jbas
  type(sq_op) :: HS, L, R, H, A, B, w1, w2, ETA1, ETA2, G, DG1, DG2, AD
  type(cc_mat) :: LCC, RCC, WCC, GCC
  character(len=1024) :: fb, fl, fr, fh, fh0
  type(three_body_force) :: tb
  type(full_sp_block_mat) :: C1, C2
  integer :: nfail, q
  real(8) :: e1, e2
  real(8), allocatable :: yy(:), yp1(:), yp2(:)
  external :: build_gs_white, build_gs_w2, dHds

  call getarg(1, fb) ; call getarg(2, fl) ; call getarg(3, fr) ; call getarg(4, fh) ; call getarg(5, fh0)
  checkpointing = .false.
  OUTPUT_DIR = './'
  prefix = 'drop_in_test'
  call setup_basis(trim(fb), jbas)
  call allocate_blocks(jbas, HS)
  HS%herm = 1
  HS%neq = compute_neq(HS)
  HS%eMax = maxval(jbas%nshell)
  HS%lawson_beta = 0.d0
  nfail = 0
  call gpu_init(jbas)

  call dup(L, -1) ; call load(trim(fl), L)
  call dup(R, 1) ; call load(trim(fr), R)
  call dup(H, 1) ; call load(trim(fh), H)

  ! full commutator, reference call sequence IMSRG_MAGNUS.f90:207-219
  call dup(A, 1) ; call dup(B, 1) ; call dup(w1, 1) ; call dup(w2, 1)
  call init_ph_mat(L, LCC, jbas) ; call duplicate_ph_mat(LCC, RCC) ; call init_ph_wkspc(LCC, WCC)
  call calculate_cross_coupled(R, RCC, jbas) ; call calculate_cross_coupled(L, LCC, jbas)
  A%E0 = commutator_110(L, R, jbas) + commutator_220(L, R, jbas)
  call commutator_111(L, R, A, jbas) ; call commutator_121(L, R, A, jbas)
  call commutator_122(L, R, A, jbas) ; call commutator_222_pp_hh(L, R, A, w1, w2, jbas)
  call commutator_221(L, R, A, w1, w2, jbas) ; call commutator_222_ph(LCC, RCC, A, WCC, jbas)
  call gpu_commutator(L, R, B, .true.)
  call cmp('full_commutator', A, B)
  call scal_cmp('commutator_110', commutator_110(L, R, jbas), gpu_commutator_110(L, R, jbas))
  call scal_cmp('commutator_220', commutator_220(L, R, jbas), gpu_commutator_220(L, R, jbas))

  ! the same sequence routine by routine through the drop-in names
  call clear_sq_op(B)
  B%E0 = gpu_commutator_110(L, R, jbas) + gpu_commutator_220(L, R, jbas)
  call gpu_commutator_111(L, R, B, jbas) ; call gpu_commutator_121(L, R, B, jbas)
  call gpu_commutator_122(L, R, B, jbas) ; call gpu_commutator_222_pp_hh(L, R, B, w1, w2, jbas)
  call gpu_commutator_221(L, R, B, w1, w2, jbas) ; call gpu_commutator_222_ph(L, R, B, jbas)
  call cmp('per_routine_sequence', A, B)

  ! Pandya matrices
  call gpu_calculate_cross_coupled(L, RCC, jbas)
  e1 = 0 ; e2 = 0
  do q = 1, LCC%nblocks
     if (LCC%rlen(q) * LCC%nph(q) == 0) cycle
     e1 = max(e1, maxval(abs(LCC%CCX(q)%X - RCC%CCX(q)%X)))
     e2 = max(e2, maxval(abs(LCC%CCX(q)%X)))
  end do
  call report('cross_coupled', e1 / max(e2, 1d-300))

  ! generators
  call dup(ETA1, -1) ; call dup(ETA2, -1)
  call build_gs_white(H, ETA1, jbas) ; call gpu_build_generator(GEN_WHITE, H, ETA2)
  call cmp('build_gs_white', ETA1, ETA2)
  call build_gs_w2(H, ETA1, jbas) ; call gpu_build_generator(GEN_W2, H, ETA2)
  call cmp('build_gs_w2', ETA1, ETA2)

  ! BCH and Magnus expansions with G = eta_white / 2
  call build_gs_white(H, ETA1, jbas)
  call dup(G, -1) ; call add_sq_op(ETA1, 0.5d0, ETA1, 0.d0, G) ; G%herm = -1
  call BCH_EXPAND(A, G, H, jbas, 'n') ; call gpu_BCH_EXPAND(B, G, H, jbas, 'n')
  call cmp('BCH_EXPAND', A, B)
  call dup(DG1, -1) ; call dup(DG2, -1) ; call dup(AD, -1)
  call copy_sq_op(ETA1, DG1) ; call copy_sq_op(ETA1, DG2)
  call MAGNUS_EXPAND(DG1, G, AD, jbas) ; call gpu_MAGNUS_EXPAND(DG2, G, AD, jbas)
  call cmp('MAGNUS_EXPAND', DG1, DG2)

  ! observable transform (the rank-0 branch of transform_sq_op_BCH) of R
  call copy_sq_op(R, A) ; call copy_sq_op(R, B)
  call transform_sq_op_BCH(A, G, jbas, 'n') ; call gpu_transform_sq_op_BCH(B, G, jbas, 'n')
  call cmp('transform_sq_op_BCH', A, B)

  ! the ODE solver's derivative dHds with the White generator
  allocate(yy(HS%neq), yp1(HS%neq), yp2(HS%neq))
  call vectorize(H, yy)
  call dHds(0.d0, yy, yp1, H, jbas, build_gs_white)
  call gpu_dHds(0.d0, yy, yp2, H, jbas, GEN_WHITE)
  call report('dHds', maxval(abs(yp1 - yp2)) / max(maxval(abs(yp1)), 1d-300))

  ! the whole flow: magnus_decouple vs gpu_magnus_decouple
  call copy_sq_op(H, A) ; call copy_sq_op(H, B)
  call clear_sq_op(ETA1) ; call clear_sq_op(ETA2)   ! Omega(0) = 0
  ETA1%herm = -1 ; ETA2%herm = -1
  call magnus_decouple(A, ETA1, jbas, 'n', 'n', build_gs_white)
  call gpu_magnus_decouple(B, ETA2, jbas, GEN_WHITE)
  A%neq = HS%neq ; B%neq = HS%neq
  call cmp('magnus_decouple_H', A, B)
  call cmp('magnus_decouple_Omega', ETA1, ETA2)

  ! checkpoints: a first reference flow writes Omega every 24 steps; a second
  ! reference flow and the drop-in then both resume from the last checkpoint
  checkpointing = .true.
  prefix = 'dropin_hw20'
  scratch = './'
  resubmitter = ''   ! no batch-job resubmission after a checkpoint (main_IMSRG reads it from argument 2)
  call copy_sq_op(H, A) ; call clear_sq_op(ETA1) ; ETA1%herm = -1
  call magnus_decouple(A, ETA1, jbas, 'n', 'n', build_gs_white)
  call copy_sq_op(H, A) ; call clear_sq_op(ETA1) ; ETA1%herm = -1
  call magnus_decouple(A, ETA1, jbas, 'n', 'n', build_gs_white)
  call copy_sq_op(H, B) ; call clear_sq_op(ETA2) ; ETA2%herm = -1
  call gpu_magnus_decouple(B, ETA2, jbas, GEN_WHITE)
  checkpointing = .false.
  A%neq = HS%neq ; B%neq = HS%neq
  call cmp('checkpoint_restart_H', A, B)
  call cmp('checkpoint_restart_Omega', ETA1, ETA2)

  ! Hartree-Fock: calc_HF vs gpu_calc_HF on the bare H, then sqop_to_HF vs
  ! gpu_observable_to_HF on a hermitian observable (R) with each one's coefficients
  A%herm = 1 ; call load(trim(fh0), A)
  B%herm = 1 ; call load(trim(fh0), B)
  call calc_HF(A, tb, jbas, C1)
  call gpu_calc_HF(B, tb, jbas, C2)
  call cmp('calc_HF', A, B)
  e1 = 0.d0
  do q = 1, C1%blocks
     if (C1%map(q) > 0) e1 = max(e1, maxval(abs(C1%blkM(q)%matrix - C2%blkM(q)%matrix)))
  end do
  call report('calc_HF_coefficients', e1)
  A%herm = 1 ; call load(trim(fr), A)
  B%herm = 1 ; call load(trim(fr), B)
  call sqop_to_HF(A, C1, jbas)
  call gpu_observable_to_HF(B, C2, jbas)
  call cmp('sqop_to_HF', A, B)

  call gpu_finalize()
  if (nfail > 0) then
     write(*,*) 'FAILURES ', nfail
     stop 1
  end if
  write(*,*) 'ALL CHECKS PASSED'

contains

  integer function compute_neq(op)
    type(sq_op) :: op
    integer :: Ax, N, qq, nh, np, nb
    Ax = op%belowEF ; N = op%Nsp
    compute_neq = 1 + Ax*Ax + Ax*(N-Ax) + (N-Ax)**2
    do qq = 1, op%nblocks
       nh = op%mat(qq)%nhh ; np = op%mat(qq)%npp ; nb = op%mat(qq)%nph
       compute_neq = compute_neq + (nh*nh+nh + nb*nb+nb + np*np+np)/2 + nb*np + nh*np + nh*nb
    end do
  end function compute_neq

  subroutine dup(op, herm)
    type(sq_op) :: op
    integer :: herm
    call duplicate_sq_op(HS, op)
    op%herm = herm
    op%neq = HS%neq
  end subroutine dup

  subroutine load(fname, op)
    character(len=*) :: fname
    type(sq_op) :: op
    real(8), allocatable :: v(:)
    allocate(v(HS%neq))
    open(unit=41, file=fname, access='stream', form='unformatted', status='old')
    read(41) v
    close(41)
    call repackage(op, v)
  end subroutine load

  subroutine cmp(name, X, Y)
    character(len=*) :: name
    type(sq_op) :: X, Y
    real(8), allocatable :: vx(:), vy(:)
    X%neq = HS%neq ; Y%neq = HS%neq
    allocate(vx(HS%neq), vy(HS%neq))
    call vectorize(X, vx) ; call vectorize(Y, vy)
    call report(name, maxval(abs(vx - vy)) / max(maxval(abs(vx)), 1d-300))
  end subroutine cmp

  subroutine scal_cmp(name, x, y)
    character(len=*) :: name
    real(8) :: x, y
    call report(name, abs(x - y) / max(abs(x), 1d0))
  end subroutine scal_cmp

  subroutine report(name, err)
    character(len=*) :: name
    real(8) :: err
    write(*,'(A,1X,A,1X,ES12.4)') 'CHECK', name, err
    if (.not. (err <= 1d-10)) nfail = nfail + 1
  end subroutine report
end program drop_in_test
