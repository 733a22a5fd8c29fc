! imsrg_gpu.f90 -- Fortran drop-in for the scalar commutator engine of
! n-parzuchowski/EOM-IMSRG_FULL, forwarding to libimsrg_mi355x.so through
! ISO_C_BINDING (the same pattern as the reference's own zlib binding,
! im-SRG/gzipmod.f90:5-39).
!
! Every public routine takes the reference's own derived types (sq_op, spd,
! cc_mat from basic_IMSRG / cross_coupled) and reproduces the contract of the
! routine it replaces.  Operators cross the boundary as the reference's
! vectorize/repackage streams (basic_IMSRG.f90:3794-3918); errors turn into
! ERROR STOP, as the reference STOPs.
!
!   call gpu_init(jbas)                          ! once, after read_sp_basis
!   call gpu_magnus_decouple(HS, G, jbas, GEN_W2)  ! replaces magnus_decouple
!   call gpu_BCH_EXPAND(HS, G, H, jbas, 'n')     ! replaces BCH_EXPAND
!   call gpu_commutator_111(L, R, RES, jbas) ...  ! per-routine replacements
module imsrg_gpu
  use iso_c_binding
  use basic_IMSRG
  use cross_coupled
  use three_body_routines, only: three_body_force
  implicit none
  private

  integer, parameter, public :: GEN_WHITE = 0, GEN_ATAN = 1, GEN_W2 = 2
  integer(c_int), parameter :: T110 = 1, T111 = 2, T121 = 4, T122 = 8, T220 = 16, T221 = 32, &
       T222PPHH = 64, T222PH = 128, TALL = 255

  public :: gpu_init, gpu_finalize
  public :: gpu_commutator, gpu_commutator_110, gpu_commutator_220, gpu_commutator_111, &
       gpu_commutator_121, gpu_commutator_122, gpu_commutator_222_pp_hh, gpu_commutator_221, &
       gpu_commutator_222_ph, gpu_calculate_cross_coupled
  public :: gpu_BCH_EXPAND, gpu_MAGNUS_EXPAND, gpu_build_generator, gpu_magnus_decouple
  public :: gpu_transform_sq_op_BCH, gpu_dHds, gpu_calc_HF, gpu_observable_to_HF

  type(c_ptr), save :: ctx = c_null_ptr
  integer(c_int64_t), save :: neq = 0
  ! the 221 one-body term of the last gpu_commutator_222_pp_hh call, for the
  ! gpu_commutator_221 call that passes the same w1, w2
  real(8), allocatable, save :: ob_fhh(:,:), ob_fph(:,:), ob_fpp(:,:)
  type(c_ptr), save :: tok_w1 = c_null_ptr, tok_w2 = c_null_ptr
  logical, save :: have221 = .false.

  interface
     function c_last_error() bind(C, name='imsrg_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_last_error
     integer(c_int) function c_ctx_create(nsp, n, l, jj, tz, con, device, out) &
          bind(C, name='imsrg_ctx_create')
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
     integer(c_int) function c_op_create(c, herm, out) bind(C, name='imsrg_op_create')
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       integer(c_int), value :: herm
       type(c_ptr) :: out
     end function c_op_create
     integer(c_int) function c_op_destroy(op) bind(C, name='imsrg_op_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: op
     end function c_op_destroy
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
     integer(c_int64_t) function c_cc_size(c) bind(C, name='imsrg_cross_coupled_size')
       import :: c_int64_t, c_ptr
       type(c_ptr), value :: c
     end function c_cc_size
     integer(c_int) function c_cross_coupled(op, out, n) bind(C, name='imsrg_cross_coupled')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: op
       real(c_double) :: out(*)
       integer(c_int64_t), value :: n
     end function c_cross_coupled
     integer(c_int) function c_generator(c, kind, H, ETA) bind(C, name='imsrg_generator')
       import :: c_int, c_ptr
       type(c_ptr), value :: c, H, ETA
       integer(c_int), value :: kind
     end function c_generator
     integer(c_int) function c_bch(c, HS, G, H, nterms) bind(C, name='imsrg_bch_expand')
       import :: c_int, c_ptr
       type(c_ptr), value :: c, HS, G, H
       integer(c_int) :: nterms
     end function c_bch
     integer(c_int) function c_transform_bch(Op, G, nterms) bind(C, name='imsrg_transform_bch')
       import :: c_int, c_ptr
       type(c_ptr), value :: Op, G
       integer(c_int) :: nterms
     end function c_transform_bch
     integer(c_int) function c_hartree_fock(H, coefs, iterations, e_hf) bind(C, name='imsrg_hartree_fock')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: H
       real(c_double) :: coefs(*), e_hf
       integer(c_int) :: iterations
     end function c_hartree_fock
     integer(c_int) function c_observable_to_hf(O, coefs) bind(C, name='imsrg_observable_to_hf')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: O
       real(c_double) :: coefs(*)
     end function c_observable_to_hf
     integer(c_int) function c_dhds(HS, kind, DH) bind(C, name='imsrg_dhds')
       import :: c_int, c_ptr
       type(c_ptr), value :: HS, DH
       integer(c_int), value :: kind
     end function c_dhds
     integer(c_int) function c_magexp(c, DG, G, nterms) bind(C, name='imsrg_magnus_expand')
       import :: c_int, c_ptr
       type(c_ptr), value :: c, DG, G
       integer(c_int) :: nterms
     end function c_magexp
     integer(c_int) function c_flow_create(c, HS, G, kind, ds, out) bind(C, name='imsrg_flow_create')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c, HS, G
       integer(c_int), value :: kind
       real(c_double), value :: ds
       type(c_ptr) :: out
     end function c_flow_create
     integer(c_int) function c_flow_restart(c, HS, G, kind, ds, s, out) bind(C, name='imsrg_flow_restart')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c, HS, G
       integer(c_int), value :: kind
       real(c_double), value :: ds, s
       type(c_ptr) :: out
     end function c_flow_restart
     integer(c_int) function c_flow_step(f, rec) bind(C, name='imsrg_flow_step')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: f
       real(c_double) :: rec(8)
     end function c_flow_step
     integer(c_int) function c_flow_converged(f, yes) bind(C, name='imsrg_flow_converged')
       import :: c_int, c_ptr
       type(c_ptr), value :: f
       integer(c_int) :: yes
     end function c_flow_converged
     integer(c_int) function c_flow_destroy(f) bind(C, name='imsrg_flow_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: f
     end function c_flow_destroy
     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*,*) 'imsrg_gpu: ', what, ' failed (', rc, '): ', msg(1:n)
    error stop 'imsrg_gpu failure'
  end subroutine check

  ! -- context: read_sp_basis / allocate_blocks / store_6j replacement
  subroutine gpu_init(jbas, device)
    type(spd), intent(in) :: jbas
    integer, intent(in), optional :: device
    integer(c_int) :: dev
    integer(c_int), allocatable :: n(:), l(:), jj(:), tz(:), con(:)
    if (c_associated(ctx)) return
    dev = 0
    if (present(device)) dev = device
    n = jbas%nn ; l = jbas%ll ; jj = jbas%jj ; tz = jbas%itzp ; con = jbas%con
    call check(c_ctx_create(int(jbas%total_orbits, c_int), n, l, jj, tz, con, dev, ctx), 'imsrg_ctx_create')
    neq = c_ctx_neq(ctx)
  end subroutine gpu_init

  subroutine gpu_finalize()
    if (c_associated(ctx)) call check(c_ctx_destroy(ctx), 'imsrg_ctx_destroy')
    ctx = c_null_ptr
  end subroutine gpu_finalize

  ! -- marshalling through the reference's vectorize/repackage
  function push(op) result(h)
    type(sq_op) :: op
    type(c_ptr) :: h
    real(8), allocatable :: v(:)
    op%neq = int(neq)
    allocate(v(neq))
    call vectorize(op, v)
    call check(c_op_create(ctx, int(op%herm, c_int), h), 'imsrg_op_create')
    call check(c_op_upload(h, v, neq), 'imsrg_op_upload')
  end function push

  function fresh(herm) result(h)
    integer, intent(in) :: herm
    type(c_ptr) :: h
    call check(c_op_create(ctx, int(herm, c_int), h), 'imsrg_op_create')
  end function fresh

  subroutine pull(h, op)
    type(c_ptr) :: h
    type(sq_op) :: op
    real(8), allocatable :: v(:)
    op%neq = int(neq)
    allocate(v(neq))
    call check(c_op_download(h, v, neq), 'imsrg_op_download')
    call repackage(op, v)
  end subroutine pull

  subroutine drop(h)
    type(c_ptr) :: h
    call check(c_op_destroy(h), 'imsrg_op_destroy')
  end subroutine drop

  ! -- the full commutator sequence of IMSRG_MAGNUS.f90:207-219 (with_e0) or
  !    :679-685 (MAGNUS_EXPAND, E0 untouched)
  subroutine gpu_commutator(L, R, RES, with_e0)
    type(sq_op) :: L, R, RES
    logical, intent(in) :: with_e0
    type(c_ptr) :: hl, hr, hres
    real(8) :: e0keep
    integer(c_int) :: mask
    e0keep = RES%E0
    hl = push(L) ; hr = push(R) ; hres = fresh(RES%herm)
    mask = TALL
    if (.not. with_e0) mask = TALL - T110 - T220
    call check(c_commutator(ctx, hl, hr, hres, mask), 'imsrg_commutator')
    call pull(hres, RES)
    if (.not. with_e0) RES%E0 = e0keep
    call drop(hl) ; call drop(hr) ; call drop(hres)
  end subroutine gpu_commutator

  real(8) function gpu_commutator_110(L, R, jbas)       ! commutators.f90:10-33
    type(sq_op) :: L, R
    type(spd) :: jbas
    type(c_ptr) :: hl, hr
    real(c_double) :: x
    hl = push(L) ; hr = push(R)
    call check(c_comm110(hl, hr, x), 'imsrg_commutator_110')
    gpu_commutator_110 = x
    call drop(hl) ; call drop(hr)
  end function gpu_commutator_110

  real(8) function gpu_commutator_220(L, R, jbas)       ! commutators.f90:399-432
    type(sq_op) :: L, R
    type(spd) :: jbas
    type(c_ptr) :: hl, hr
    real(c_double) :: x
    hl = push(L) ; hr = push(R)
    call check(c_comm220(hl, hr, x), 'imsrg_commutator_220')
    gpu_commutator_220 = x
    call drop(hl) ; call drop(hr)
  end function gpu_commutator_220

  ! one term of [L,R] into a scratch sq_op shaped like RES
  subroutine term(L, R, RES, mask, T)
    type(sq_op) :: L, R, RES, T
    integer(c_int), intent(in) :: mask
    type(c_ptr) :: hl, hr, ht
    call duplicate_sq_op(RES, T)
    T%herm = RES%herm
    hl = push(L) ; hr = push(R) ; ht = fresh(RES%herm)
    call check(c_commutator(ctx, hl, hr, ht, mask), 'imsrg_commutator')
    call pull(ht, T)
    call drop(hl) ; call drop(hr) ; call drop(ht)
  end subroutine term

  subroutine add_onebody(RES, T, overwrite)
    type(sq_op) :: RES, T
    logical, intent(in) :: overwrite
    if (overwrite) then
       RES%fhh = T%fhh ; RES%fph = T%fph ; RES%fpp = T%fpp
    else
       RES%fhh = RES%fhh + T%fhh ; RES%fph = RES%fph + T%fph ; RES%fpp = RES%fpp + T%fpp
    end if
  end subroutine add_onebody

  subroutine add_twobody(RES, T, overwrite)
    type(sq_op) :: RES, T
    logical, intent(in) :: overwrite
    integer :: q, g
    do q = 1, RES%nblocks
       do g = 1, 6
          if (overwrite) then
             RES%mat(q)%gam(g)%X = T%mat(q)%gam(g)%X
          else
             RES%mat(q)%gam(g)%X = RES%mat(q)%gam(g)%X + T%mat(q)%gam(g)%X
          end if
       end do
    end do
  end subroutine add_twobody

  subroutine gpu_commutator_111(L, R, RES, jbas)        ! overwrites the one-body part
    type(sq_op) :: L, R, RES, T
    type(spd) :: jbas
    call term(L, R, RES, T111, T)
    call add_onebody(RES, T, .true.)
  end subroutine gpu_commutator_111

  subroutine gpu_commutator_121(L, R, RES, jbas)        ! accumulates one-body
    type(sq_op) :: L, R, RES, T
    type(spd) :: jbas
    call term(L, R, RES, T121, T)
    call add_onebody(RES, T, .false.)
  end subroutine gpu_commutator_121

  subroutine gpu_commutator_122(L, R, RES, jbas)        ! overwrites every stored two-body element
    type(sq_op) :: L, R, RES, T
    type(spd) :: jbas
    call term(L, R, RES, T122, T)
    call add_twobody(RES, T, .true.)
  end subroutine gpu_commutator_122

  ! accumulates two-body; one pp/hh product also gives the 221 one-body term,
  ! kept for the gpu_commutator_221 call that passes the same w1, w2 (which are
  ! not filled: the device keeps its own pp/hh intermediates)
  subroutine gpu_commutator_222_pp_hh(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES, T
    type(sq_op), target :: w1, w2
    type(spd) :: jbas
    call term(L, R, RES, ior(T222PPHH, T221), T)
    call add_twobody(RES, T, .false.)
    ob_fhh = T%fhh ; ob_fph = T%fph ; ob_fpp = T%fpp
    tok_w1 = c_loc(w1)
    tok_w2 = c_loc(w2)
    have221 = .true.
  end subroutine gpu_commutator_222_pp_hh

  subroutine gpu_commutator_221(L, R, RES, w1, w2, jbas)  ! accumulates one-body
    type(sq_op) :: L, R, RES, T
    type(sq_op), target :: w1, w2
    type(spd) :: jbas
    if (have221 .and. c_associated(tok_w1, c_loc(w1)) .and. c_associated(tok_w2, c_loc(w2))) then
       RES%fhh = RES%fhh + ob_fhh ; RES%fph = RES%fph + ob_fph ; RES%fpp = RES%fpp + ob_fpp
    else
       call term(L, R, RES, T221, T)
       call add_onebody(RES, T, .false.)
    end if
    have221 = .false.
  end subroutine gpu_commutator_221

  ! the device forms the Pandya matrices of both operands itself, so this takes
  ! the operators rather than their cc_mat (LCC/RCC of the reference)
  subroutine gpu_commutator_222_ph(L, R, RES, jbas)     ! accumulates two-body
    type(sq_op) :: L, R, RES, T
    type(spd) :: jbas
    call term(L, R, RES, T222PH, T)
    call add_twobody(RES, T, .false.)
  end subroutine gpu_commutator_222_ph

  ! calculate_cross_coupled (cross_coupled.f90:1190-1300) into an allocated cc_mat
  subroutine gpu_calculate_cross_coupled(HS, CCME, jbas)
    type(sq_op) :: HS
    type(cc_mat) :: CCME
    type(spd) :: jbas
    type(c_ptr) :: h
    real(8), allocatable :: v(:)
    integer(c_int64_t) :: n, k
    integer :: q, m
    n = c_cc_size(ctx)
    allocate(v(max(n, 1_c_int64_t)))
    h = push(HS)
    call check(c_cross_coupled(h, v, n), 'imsrg_cross_coupled')
    call drop(h)
    k = 0
    do q = 1, CCME%nblocks
       m = CCME%rlen(q) * CCME%nph(q)
       if (m == 0) cycle
       CCME%CCX(q)%X = reshape(v(k+1:k+m), [CCME%rlen(q), CCME%nph(q)])
       k = k + m
    end do
    CCME%herm = HS%herm
  end subroutine gpu_calculate_cross_coupled

  ! BCH_EXPAND (IMSRG_MAGNUS.f90:149-234): HS = e^G H e^-G
  subroutine gpu_BCH_EXPAND(HS, G, H, jbas, quads)
    type(sq_op) :: HS, G, H
    type(spd) :: jbas
    character(1) :: quads
    type(c_ptr) :: hg, hh, hs_
    integer(c_int) :: nterms
    if (quads == 'y') error stop 'imsrg_gpu: quadrupole restoration is not provided'
    hg = push(G) ; hh = push(H) ; hs_ = fresh(1)
    call check(c_bch(ctx, hs_, hg, hh, nterms), 'imsrg_bch_expand')
    HS%herm = H%herm
    call pull(hs_, HS)
    call drop(hg) ; call drop(hh) ; call drop(hs_)
  end subroutine gpu_BCH_EXPAND

  ! transform_sq_op_BCH (IMSRG_MAGNUS.f90:131-146) for rank-0 operators:
  ! Op <- e^G Op e^-G
  subroutine gpu_transform_sq_op_BCH(Op, G, jbas, quads)
    type(sq_op) :: Op, G
    type(spd) :: jbas
    character(1) :: quads
    type(c_ptr) :: ho, hg
    integer(c_int) :: nterms
    if (Op%rank > 0) error stop 'imsrg_gpu: tensor operators (BCH_TENSOR) are not provided'
    if (quads == 'y') error stop 'imsrg_gpu: quadrupole restoration is not provided'
    ho = push(Op) ; hg = push(G)
    call check(c_transform_bch(ho, hg, nterms), 'imsrg_transform_bch')
    call pull(ho, Op)
    call drop(ho) ; call drop(hg)
  end subroutine gpu_transform_sq_op_BCH

  ! calc_HF (HF_mod.f90:14-106) without a three-body force: the bare H becomes
  ! the normal-ordered Hamiltonian in the Hartree-Fock basis, and D (unallocated
  ! on entry, as the reference allocates it) the coefficients in the reference's
  ! block form (allocate_sp_mat), eigval = the Fock matrix's diagonal
  subroutine gpu_calc_HF(H, THREEBOD, jbas, D)
    type(sq_op) :: H
    type(three_body_force) :: THREEBOD
    type(spd) :: jbas
    type(full_sp_block_mat) :: D
    type(c_ptr) :: hh
    real(c_double), allocatable :: cf(:,:)
    real(c_double) :: ehf
    integer(c_int) :: it
    integer :: q, i, j
    if (allocated(THREEBOD%mat)) error stop 'imsrg_gpu: three-body forces in Hartree-Fock are not provided'
    allocate(cf(jbas%total_orbits, jbas%total_orbits))
    hh = push(H)
    call check(c_hartree_fock(hh, cf, it, ehf), 'imsrg_hartree_fock')
    call pull(hh, H)
    call drop(hh)
    write(*,'(A22,f12.7)') ' Hartree Fock Energy: ', ehf
    ! the library's coefs[a*nsp + i] is cf(i+1, a+1) here
    call allocate_sp_mat(jbas, D)
    do q = 1, D%blocks
       do i = 1, D%map(q)
          do j = 1, D%map(q)
             D%blkM(q)%matrix(i, j) = cf(D%blkM(q)%states(j), D%blkM(q)%states(i))
          end do
          D%blkM(q)%eigval(i) = f_elem(D%blkM(q)%states(i), D%blkM(q)%states(i), H, jbas)
       end do
    end do
  end subroutine gpu_calc_HF

  ! sqop_to_HF (HF_mod.f90:108-136) for a scalar observable
  subroutine gpu_observable_to_HF(Op, coefs, jbas)
    type(sq_op) :: Op
    type(full_sp_block_mat) :: coefs
    type(spd) :: jbas
    type(c_ptr) :: ho
    real(c_double), allocatable :: cf(:,:)
    integer :: q, i, j
    if (allocated(Op%tblck)) error stop 'imsrg_gpu: tensor observables (transform_1b_to_HF_tensor) are not provided'
    if (.not. allocated(Op%mat)) return
    allocate(cf(jbas%total_orbits, jbas%total_orbits))
    cf = 0.d0
    do q = 1, coefs%blocks
       do i = 1, coefs%map(q)
          do j = 1, coefs%map(q)
             cf(coefs%blkM(q)%states(j), coefs%blkM(q)%states(i)) = coefs%blkM(q)%matrix(i, j)
          end do
       end do
    end do
    ho = push(Op)
    call check(c_observable_to_hf(ho, cf), 'imsrg_observable_to_hf')
    call pull(ho, Op)
    call drop(ho)
  end subroutine gpu_observable_to_HF

  ! dHds (IMSRG_ODE.f90:416-463), the derivative the ODE solver evaluates: like
  ! the reference it builds the generator from HS (the device generator `kind`
  ! in place of the external build_gen) and returns vectorize([eta, HS]) in yp
  subroutine gpu_dHds(t, yy, yp, HS, jbas, kind)
    real(8) :: t, yy(*), yp(*)
    type(sq_op) :: HS
    type(spd) :: jbas
    integer, intent(in) :: kind
    type(c_ptr) :: hh, hd
    hh = push(HS) ; hd = fresh(HS%herm)
    call check(c_dhds(hh, int(kind, c_int), hd), 'imsrg_dhds')
    call check(c_op_download(hd, yp, neq), 'imsrg_op_download')
    call drop(hh) ; call drop(hd)
  end subroutine gpu_dHds

  ! MAGNUS_EXPAND (IMSRG_MAGNUS.f90:622-693): DG <- dOmega/ds
  subroutine gpu_MAGNUS_EXPAND(DG, G, AD, jbas)
    type(sq_op) :: DG, G, AD
    type(spd) :: jbas
    type(c_ptr) :: hdg, hg
    integer(c_int) :: nterms
    hdg = push(DG) ; hg = push(G)
    call check(c_magexp(ctx, hdg, hg, nterms), 'imsrg_magnus_expand')
    call pull(hdg, DG)
    call drop(hdg) ; call drop(hg)
  end subroutine gpu_MAGNUS_EXPAND

  ! build_gs_white / build_gs_atan / build_gs_w2 (generators.f90)
  subroutine gpu_build_generator(kind, H, ETA)
    integer, intent(in) :: kind
    type(sq_op) :: H, ETA
    type(c_ptr) :: hh, he
    hh = push(H) ; he = fresh(-1)
    call check(c_generator(ctx, int(kind, c_int), hh, he), 'imsrg_generator')
    ETA%herm = -1
    call pull(he, ETA)
    call drop(hh) ; call drop(he)
  end subroutine gpu_build_generator

  ! magnus_decouple (IMSRG_MAGNUS.f90:15-128) with H(s), Omega(s) resident on the
  ! device for the whole flow; prints the reference's step line
  subroutine gpu_magnus_decouple(HS, G, jbas, kind)
    type(sq_op) :: HS, G
    type(spd) :: jbas
    integer, intent(in) :: kind
    type(c_ptr) :: hs_, hg, flow
    real(c_double) :: rec(8)
    real(8) :: s, sx
    integer(c_int) :: done
    integer :: steps, chk
    logical :: restart, first
    if (.not. allocated(G%mat)) call duplicate_sq_op(HS, G)
    ! checkpoints as in magnus_decouple (IMSRG_MAGNUS.f90:55-66, 83-97, 101-108):
    ! resume from the last Omega checkpoint the reference's reader finds, and
    ! write Omega with the reference's writer every chk steps (not on the
    ! first such step after a restart)
    chk = 24
    if (HS%eMax == 14) chk = 4
    if (HS%eMax > 14) chk = 1
    restart = .false.
    if (checkpointing) restart = .not. read_omega_checkpoint(G, sx)
    G%herm = -1
    hs_ = push(HS)
    if (restart) then
       hg = push(G)
       call check(c_flow_restart(ctx, hs_, hg, int(kind, c_int), 0.5d0, sx, flow), 'imsrg_flow_restart')
       s = sx ; steps = nint(sx / 0.5d0) ; first = .true.
    else
       hg = fresh(-1)
       call check(c_flow_create(ctx, hs_, hg, int(kind, c_int), 0.5d0, flow), 'imsrg_flow_create')
       s = 0.d0 ; steps = 0 ; first = .false.
    end if
    call check(c_flow_converged(flow, done), 'imsrg_flow_converged')
    do while (done == 0)
       if (checkpointing .and. mod(steps + 1, chk) == 0) then
          if (first) then
             first = .false.
          else
             call pull(hg, G)
             call write_omega_checkpoint(G, s)
          end if
       end if
       call check(c_flow_step(flow, rec), 'imsrg_flow_step')
       s = rec(1) ; steps = nint(rec(7))
       write(*,'(I6,4(e15.7))') steps, rec(1), rec(2), rec(3), rec(4)
       call check(c_flow_converged(flow, done), 'imsrg_flow_converged')
    end do
    call check(c_flow_destroy(flow), 'imsrg_flow_destroy')
    call pull(hs_, HS)
    call pull(hg, G)
    call drop(hs_) ; call drop(hg)
  end subroutine gpu_magnus_decouple

end module imsrg_gpu
