! ts_commutators_gpu.f90 -- link-time substitute for the reference's
! `module TS_commutators` (n-parzuchowski/EOM-IMSRG_FULL,
! im-SRG/TS_commutators.f90): the scalar-tensor commutator terms on the
! MI355X through libimsrg_mi355x.so.
!
! A program that links this object in place of the reference's
! TS_commutators.o, and compiles its callers against this module file, runs
! every scalar-tensor commutator on the GPU; the callers -- IMSRG_MAGNUS.f90
! (BCH_TENSOR, transform_sq_op_BCH for rank > 0 operators) and
! brute_force_testing.f90 -- compile unmodified.  Public names, argument
! lists and contracts are the reference's (TS_commutators.f90:10, 58, 307,
! 555, 666, 809, 1070, 1459):
!
!   TS_commutator_111                 overwrites the one-body part of RES
!   TS_commutator_122                 overwrites the two-body part of RES
!   TS_commutator_121, _211, _221     add to the one-body part, then rewrite
!                                     each lower entry from its upper one
!   TS_commutator_212, _222_ph        add to the two-body part
!   TS_commutator_222_pp_hh           adds to the two-body part and leaves the
!                                     ladder products in w1 (pp) and w2 (hh)
!                                     for TS_commutator_221
!   TS_commutator_211, _222_ph        take G's Pandya matrices LCC as the
!                                     caller made them with
!                                     calculate_cross_coupled; RCC is workspace
!
! RES is carried through the device (uploaded, updated in place, read back),
! so every term acts on it exactly as the reference routine does.  Like the
! reference module it re-exports module cross_coupled.  TS_commutator_223_single
! (the tensor triples correction) is not provided: calling it stops.
!
! Tensors cross the boundary as the reference's storage flattened in
! declaration order (E0, fhh, fpp, fph, then tgam(1..9) of every block,
! column-major), the scalar operator as its vectorize stream
! (basic_IMSRG.f90:3794-3918); errors turn into ERROR STOP.
module TS_commutators
  use iso_c_binding
  use cross_coupled
  implicit none

  public :: TS_commutator_111, TS_commutator_121, TS_commutator_211, TS_commutator_122, TS_commutator_212, &
       TS_commutator_221, TS_commutator_222_pp_hh, TS_commutator_222_ph, TS_commutator_223_single

  integer(c_int), parameter, private :: M111 = 1, M121 = 2, M211 = 4, M122 = 8, M212 = 16, M222PPHH = 32, &
       M221 = 64, M222PH = 128
  type(c_ptr), save, private :: tctx = c_null_ptr, hG = c_null_ptr, hR = c_null_ptr, hO = c_null_ptr, &
       hW1 = c_null_ptr, hW2 = c_null_ptr
  integer, save, private :: ts_nsp = -1, ts_rank = -1, ts_dpar = -1
  integer(c_int64_t), save, private :: ts_neq = 0, sc_neq = 0, ts_ncc = 0
  real(c_double), allocatable, save, private :: sbuf(:), tbuf(:), ccl(:)

  private :: ts_check, ts_init, ts_free, ts_flat, ts_unflat, ts_push_scalar, ts_push, ts_pull, ts_run, ts_run_cc

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
     integer(c_int) function c_t_create(c, rank, dpar, herm, out) bind(C, name='imsrg_tensor_create')
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       integer(c_int), value :: rank, dpar, herm
       type(c_ptr) :: out
     end function c_t_create
     integer(c_int) function c_t_destroy(t) bind(C, name='imsrg_tensor_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: t
     end function c_t_destroy
     integer(c_int64_t) function c_t_neq(t) bind(C, name='imsrg_tensor_neq')
       import :: c_int64_t, c_ptr
       type(c_ptr), value :: t
     end function c_t_neq
     integer(c_int) function c_t_set_herm(t, herm) bind(C, name='imsrg_tensor_set_herm')
       import :: c_int, c_ptr
       type(c_ptr), value :: t
       integer(c_int), value :: herm
     end function c_t_set_herm
     integer(c_int) function c_t_upload(t, v, n) bind(C, name='imsrg_tensor_upload')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: t
       real(c_double) :: v(*)
       integer(c_int64_t), value :: n
     end function c_t_upload
     integer(c_int) function c_t_download(t, v, n) bind(C, name='imsrg_tensor_download')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: t
       real(c_double) :: v(*)
       integer(c_int64_t), value :: n
     end function c_t_download
     integer(c_int) function c_t_comm(G, T, RES, mask) bind(C, name='imsrg_tensor_commutator')
       import :: c_int, c_ptr
       type(c_ptr), value :: G, T, RES
       integer(c_int), value :: mask
     end function c_t_comm
     integer(c_int) function c_t_comm_cc(T, gcc, hg, n, RES, mask) bind(C, name='imsrg_tensor_commutator_cc')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: T, RES
       real(c_double) :: gcc(*)
       integer(c_int), value :: hg, mask
       integer(c_int64_t), value :: n
     end function c_t_comm_cc
     integer(c_int) function c_t_ladders(G, T, W1, W2) bind(C, name='imsrg_tensor_ladders')
       import :: c_int, c_ptr
       type(c_ptr), value :: G, T, W1, W2
     end function c_t_ladders
     integer(c_int) function c_t_221(W1, W2, pm, RES) bind(C, name='imsrg_tensor_commutator_221')
       import :: c_int, c_ptr
       type(c_ptr), value :: W1, W2, RES
       integer(c_int), value :: pm
     end function c_t_221
  end interface
  private :: c_last_error, c_strlen, c_ctx_create, c_ctx_destroy, c_ctx_neq, c_cc_size, c_op_create, &
       c_op_set_herm, c_op_upload, c_t_create, c_t_destroy, c_t_neq, c_t_set_herm, c_t_upload, c_t_download, &
       c_t_comm, c_t_comm_cc, c_t_ladders, c_t_221

contains

  subroutine ts_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*,*) 'TS_commutators (MI355X): ', what, ' failed (', rc, '): ', msg(1:n)
    error stop 'TS_commutators (MI355X) failure'
  end subroutine ts_check

  subroutine ts_free()
    if (c_associated(hR)) call ts_check(c_t_destroy(hR), 'imsrg_tensor_destroy')
    if (c_associated(hO)) call ts_check(c_t_destroy(hO), 'imsrg_tensor_destroy')
    if (c_associated(hW1)) call ts_check(c_t_destroy(hW1), 'imsrg_tensor_destroy')
    if (c_associated(hW2)) call ts_check(c_t_destroy(hW2), 'imsrg_tensor_destroy')
    hR = c_null_ptr ; hO = c_null_ptr ; hW1 = c_null_ptr ; hW2 = c_null_ptr
    ts_rank = -1
  end subroutine ts_free

  ! one device context per basis and four device tensors per (rank, dpar),
  ! made on first use and remade when the basis or the tensor kind changes
  subroutine ts_init(jbas, rank, dpar)
    type(spd) :: jbas
    integer, intent(in) :: rank, dpar
    integer(c_int), allocatable :: n(:), l(:), jj(:), tz(:), con(:)
    if (.not. (c_associated(tctx) .and. ts_nsp == jbas%total_orbits)) then
       call ts_free()
       if (c_associated(tctx)) call ts_check(c_ctx_destroy(tctx), 'imsrg_ctx_destroy')
       n = jbas%nn ; l = jbas%ll ; jj = jbas%jj ; tz = jbas%itzp ; con = jbas%con
       call ts_check(c_ctx_create(int(jbas%total_orbits, c_int), n, l, jj, tz, con, 0_c_int, tctx), &
            'imsrg_ctx_create')
       ts_nsp = jbas%total_orbits
       sc_neq = c_ctx_neq(tctx)
       ts_ncc = c_cc_size(tctx)
       call ts_check(c_op_create(tctx, 1_c_int, hG), 'imsrg_op_create')
       if (allocated(sbuf)) deallocate(sbuf, ccl)
       allocate(sbuf(sc_neq), ccl(max(ts_ncc, 1_c_int64_t)))
    end if
    if (rank /= ts_rank .or. dpar /= ts_dpar) then
       call ts_free()
       call ts_check(c_t_create(tctx, int(rank, c_int), int(dpar, c_int), 1_c_int, hR), 'imsrg_tensor_create')
       call ts_check(c_t_create(tctx, int(rank, c_int), int(dpar, c_int), 1_c_int, hO), 'imsrg_tensor_create')
       call ts_check(c_t_create(tctx, int(rank, c_int), int(dpar, c_int), 1_c_int, hW1), 'imsrg_tensor_create')
       call ts_check(c_t_create(tctx, int(rank, c_int), int(dpar, c_int), 1_c_int, hW2), 'imsrg_tensor_create')
       ts_neq = c_t_neq(hR)
       if (allocated(tbuf)) deallocate(tbuf)
       allocate(tbuf(ts_neq))
       ts_rank = rank
       ts_dpar = dpar
    end if
  end subroutine ts_init

  ! a tensor sq_op as its storage stream, and back (twobody: tgam only)
  subroutine ts_flat(op, v)
    type(sq_op) :: op
    real(c_double), intent(out) :: v(:)
    integer(c_int64_t) :: k, m
    integer :: q, g
    v(1) = op%E0
    k = 1
    m = size(op%fhh, kind=c_int64_t) ; v(k+1:k+m) = reshape(op%fhh, [m]) ; k = k + m
    m = size(op%fpp, kind=c_int64_t) ; v(k+1:k+m) = reshape(op%fpp, [m]) ; k = k + m
    m = size(op%fph, kind=c_int64_t) ; v(k+1:k+m) = reshape(op%fph, [m]) ; k = k + m
    do q = 1, op%nblocks
       do g = 1, 9
          m = size(op%tblck(q)%tgam(g)%X, kind=c_int64_t)
          if (m == 0) cycle
          v(k+1:k+m) = reshape(op%tblck(q)%tgam(g)%X, [m])
          k = k + m
       end do
    end do
    if (k /= ts_neq) error stop 'TS_commutators (MI355X): tensor does not match the basis'
  end subroutine ts_flat

  subroutine ts_unflat(v, op, twobody)
    real(c_double), intent(in) :: v(:)
    type(sq_op) :: op
    logical, intent(in) :: twobody
    integer(c_int64_t) :: k, m
    integer :: q, g
    k = 1
    m = size(op%fhh, kind=c_int64_t)
    if (.not. twobody) op%fhh = reshape(v(k+1:k+m), shape(op%fhh))
    k = k + m
    m = size(op%fpp, kind=c_int64_t)
    if (.not. twobody) op%fpp = reshape(v(k+1:k+m), shape(op%fpp))
    k = k + m
    m = size(op%fph, kind=c_int64_t)
    if (.not. twobody) op%fph = reshape(v(k+1:k+m), shape(op%fph))
    k = k + m
    if (.not. twobody) op%E0 = v(1)
    do q = 1, op%nblocks
       do g = 1, 9
          m = size(op%tblck(q)%tgam(g)%X, kind=c_int64_t)
          if (m == 0) cycle
          op%tblck(q)%tgam(g)%X = reshape(v(k+1:k+m), shape(op%tblck(q)%tgam(g)%X))
          k = k + m
       end do
    end do
  end subroutine ts_unflat

  subroutine ts_push_scalar(op)
    type(sq_op) :: op
    integer :: keep
    keep = op%neq
    op%neq = int(sc_neq)
    call vectorize(op, sbuf)
    op%neq = keep
    call ts_check(c_op_set_herm(hG, int(op%herm, c_int)), 'imsrg_op_set_herm')
    call ts_check(c_op_upload(hG, sbuf, sc_neq), 'imsrg_op_upload')
  end subroutine ts_push_scalar

  subroutine ts_push(op, h)
    type(sq_op) :: op
    type(c_ptr) :: h
    call ts_flat(op, tbuf)
    call ts_check(c_t_set_herm(h, int(op%herm, c_int)), 'imsrg_tensor_set_herm')
    call ts_check(c_t_upload(h, tbuf, ts_neq), 'imsrg_tensor_upload')
  end subroutine ts_push

  subroutine ts_pull(h, op, twobody)
    type(c_ptr) :: h
    type(sq_op) :: op
    logical, intent(in) :: twobody
    call ts_check(c_t_download(h, tbuf, ts_neq), 'imsrg_tensor_download')
    call ts_unflat(tbuf, op, twobody)
  end subroutine ts_pull

  ! the terms in mask of [L, R] applied to RES on the device
  subroutine ts_run(L, R, RES, jbas, mask)
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    integer(c_int), intent(in) :: mask
    call ts_init(jbas, R%rank, R%dpar)
    call ts_push_scalar(L)
    call ts_push(R, hR)
    call ts_push(RES, hO)
    call ts_check(c_t_comm(hG, hR, hO, mask), 'imsrg_tensor_commutator')
    call ts_pull(hO, RES, .false.)
  end subroutine ts_run

  ! 211 / 222_ph from G's Pandya matrices LCC (cc_mat, column-major blocks)
  subroutine ts_run_cc(LCC, R, RES, jbas, mask)
    type(cc_mat) :: LCC
    type(sq_op) :: R, RES
    type(spd) :: jbas
    integer(c_int), intent(in) :: mask
    integer :: q, m
    integer(c_int64_t) :: k
    call ts_init(jbas, R%rank, R%dpar)
    k = 0
    do q = 1, LCC%nblocks
       m = LCC%rlen(q) * LCC%nph(q)
       if (m == 0) cycle
       ccl(k+1:k+m) = reshape(LCC%CCX(q)%X, [m])
       k = k + m
    end do
    if (k /= ts_ncc) error stop 'TS_commutators (MI355X): cc_mat does not match the basis'
    call ts_push(R, hR)
    call ts_push(RES, hO)
    call ts_check(c_t_comm_cc(hR, ccl, int(LCC%herm, c_int), ts_ncc, hO, mask), 'imsrg_tensor_commutator_cc')
    call ts_pull(hO, RES, .false.)
  end subroutine ts_run_cc

  subroutine TS_commutator_111(L, R, RES, jbas)          ! TS_commutators.f90:10-55
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call ts_run(L, R, RES, jbas, M111)
  end subroutine TS_commutator_111

  subroutine TS_commutator_121(L, R, RES, jbas)          ! TS_commutators.f90:58-303
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call ts_run(L, R, RES, jbas, M121)
  end subroutine TS_commutator_121

  subroutine TS_commutator_211(LCC, R, RES, jbas)        ! TS_commutators.f90:307-552
    type(cc_mat) :: LCC
    type(sq_op) :: R, RES
    type(spd) :: jbas
    call ts_run_cc(LCC, R, RES, jbas, M211)
  end subroutine TS_commutator_211

  subroutine TS_commutator_122(L, R, RES, jbas)          ! TS_commutators.f90:555-663
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call ts_run(L, R, RES, jbas, M122)
  end subroutine TS_commutator_122

  subroutine TS_commutator_212(L, R, RES, jbas)          ! TS_commutators.f90:666-806
    type(sq_op) :: L, R, RES
    type(spd) :: jbas
    call ts_run(L, R, RES, jbas, M212)
  end subroutine TS_commutator_212

  subroutine TS_commutator_221(w1, w2, pm, RES, jbas)    ! TS_commutators.f90:809-1067
    type(sq_op) :: w1, w2, RES
    integer, intent(in) :: pm
    type(spd) :: jbas
    call ts_init(jbas, RES%rank, RES%dpar)
    call ts_push(w1, hW1)
    call ts_push(w2, hW2)
    call ts_push(RES, hO)
    call ts_check(c_t_221(hW1, hW2, int(pm, c_int), hO), 'imsrg_tensor_commutator_221')
    call ts_pull(hO, RES, .false.)
  end subroutine TS_commutator_221

  ! TS_commutators.f90:1070-1456: RES gains W1 - W2, and w1, w2 keep the
  ! ladder products (their two-body parts) for TS_commutator_221
  subroutine TS_commutator_222_pp_hh(L, R, RES, w1, w2, jbas)
    type(sq_op) :: L, R, RES, w1, w2
    type(spd) :: jbas
    call ts_run(L, R, RES, jbas, M222PPHH)
    call ts_check(c_t_ladders(hG, hR, hW1, hW2), 'imsrg_tensor_ladders')
    call ts_pull(hW1, w1, .true.)
    call ts_pull(hW2, w2, .true.)
  end subroutine TS_commutator_222_pp_hh

  subroutine TS_commutator_222_ph(LCC, RCC, R, RES, jbas) ! TS_commutators.f90:1459-1849; RCC is workspace
    type(cc_mat) :: LCC
    type(pandya_mat) :: RCC
    type(sq_op) :: R, RES
    type(spd) :: jbas
    call ts_run_cc(LCC, R, RES, jbas, M222PH)
  end subroutine TS_commutator_222_ph

  ! TS_commutators.f90:1852-2656 (the tensor triples correction) has no device counterpart
  real(8) function TS_commutator_223_single(L, R, ip, iq, ir, is, it, iu, jtot1, jtot2, Jpq, Jst, jbas)
    type(sq_op) :: L, R
    integer, intent(in) :: ip, iq, ir, is, it, iu, jtot1, jtot2, Jpq, Jst
    type(spd) :: jbas
    TS_commutator_223_single = 0.d0
    error stop 'TS_commutators (MI355X): TS_commutator_223_single (triples) is not provided'
  end function TS_commutator_223_single

end module TS_commutators
