/*
 * imsrg_mi355x.h -- C ABI of the MI355X-native IMSRG(2) Magnus commutator engine.
 *
 * Drop-in for the scalar commutator engine of n-parzuchowski/EOM-IMSRG_FULL
 * (im-SRG/commutators.f90, cross_coupled.f90, generators.f90,
 * IMSRG_MAGNUS.f90).  Operators are device-resident; they cross the boundary
 * only as FP64 streams in the reference's own vectorize order
 * (basic_IMSRG.f90:3794-3854), so a Fortran caller moves an sq_op with
 * vectorize/repackage and an ISO_C_BINDING interface (the module
 * eom-imsrg_full_amd/fortran/imsrg_gpu.f90 holds one for every entry point
 * the reference's call sites need).
 *
 * Conventions: angular momenta doubled (2j, 2tz); orbits 0-based on the C side;
 * every function returns 0 on success and a negative IMSRG_E* code on failure,
 * with a message from imsrg_last_error() (the reference STOPs instead).
 * A context is bound to one device and one stream; calls on one context are
 * stream-ordered and must come from one host thread at a time.  Different
 * contexts are independent (no global mutable state besides the per-thread
 * error string and the per-GPU scratch array of imsrg_hartree_fock_noalloc).
 */
#ifndef IMSRG_MI355X_H
#define IMSRG_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct imsrg_ctx_s *imsrg_ctx;
typedef struct imsrg_op_s *imsrg_op;

/* status codes */
#define IMSRG_OK 0
#define IMSRG_EINVAL -1     /* bad argument / size mismatch                  */
#define IMSRG_EDEVICE -2    /* HIP runtime error                              */
#define IMSRG_ENOMEM -3     /* device allocation failed                        */
#define IMSRG_ECLIP -4      /* BCH series did not converge in 40 terms        */
#define IMSRG_ECONV -5      /* Hartree-Fock did not converge in 10000 iterations */

/* term mask bits for imsrg_commutator (one per reference routine) */
#define IMSRG_T110 (1u << 0)      /* commutator_110      commutators.f90:10-33   */
#define IMSRG_T111 (1u << 1)      /* commutator_111      commutators.f90:36-80   */
#define IMSRG_T121 (1u << 2)      /* commutator_121      commutators.f90:83-280  */
#define IMSRG_T122 (1u << 3)      /* commutator_122      commutators.f90:283-396 */
#define IMSRG_T220 (1u << 4)      /* commutator_220      commutators.f90:399-432 */
#define IMSRG_T221 (1u << 5)      /* commutator_221      commutators.f90:435-578 */
#define IMSRG_T222PPHH (1u << 6)  /* commutator_222_pp_hh commutators.f90:581-768 */
#define IMSRG_T222PH (1u << 7)    /* commutator_222_ph   commutators.f90:771-925 */
#define IMSRG_TALL 0xFFu

/* arithmetic of the 222 GEMMs (pp/hh ladder and Pandya products); everything
 * else (E0, norms, gathers, generators) always runs in FP64 */
#define IMSRG_PREC_FP64 0
#define IMSRG_PREC_SPLIT_FP32 1   /* x = hi+lo in FP32; hi*hi+hi*lo+lo*hi on FP32 MFMA, FP64 across K slabs */
#define IMSRG_PREC_SPLIT_BF16 2   /* x = hi+mid+lo in BF16; 6 products on BF16 MFMA, FP32 in-slab, FP64 across */
#define IMSRG_PREC_INT8X5 3       /* x = s_row * sum_{i<=5} a_i 2^-7i in rounded int8 slices (Ozaki scheme);
                                     the 15 slice products with i+j <= 6 on int8 MFMA, exact int32 sums,
                                     FP64 combine */
#define IMSRG_PREC_INT8X4 4       /* the same with four slices and the 10 products with i+j <= 5
                                     (about 2^-29 of max|A row| max|B row| per product) */

/* generator kinds (generators.f90) */
#define IMSRG_GEN_WHITE 0         /* build_gs_white  generators.f90:4-106   */
#define IMSRG_GEN_ATAN 1          /* build_gs_atan   generators.f90:163-266 */
#define IMSRG_GEN_W2 2            /* build_gs_w2     generators.f90:109-160 */

const char *imsrg_last_error(void);

/* ---- basis / context ----------------------------------------------------
 * Replaces read_sp_basis + allocate_blocks + store_6j + allocate_CCMAT
 * (basic_IMSRG.f90:152-641, 2183-2288; cross_coupled.f90:61-177): orbit
 * quantum numbers (n, l, 2j, 2tz with protons -1) and con (1 = hole) in the
 * caller's orbit order; builds the channel layout, index and 6j tables on
 * the device.  device < 0 keeps the current device. */
int imsrg_ctx_create(int nsp, const int *n, const int *l, const int *jj, const int *tz,
                     const int *con, int device, imsrg_ctx *out);
/* Batched context (no counterpart in the reference, which flows one
 * Hamiltonian per process): every operator of the context holds nbatch
 * independent instances on the same basis, and every call acts on all of
 * them with one launch per kernel (the instance is the second grid
 * dimension).  nbatch = 1 is imsrg_ctx_create.  Scalar getters of a batched
 * context are the _batch variants below. */
int imsrg_ctx_create_batch(int nsp, const int *n, const int *l, const int *jj, const int *tz,
                           const int *con, int device, int nbatch, imsrg_ctx *out);
int imsrg_ctx_nbatch(imsrg_ctx ctx);
int imsrg_ctx_destroy(imsrg_ctx ctx);
/* length of the vectorize stream (IMSRG_ODE.f90:36-44) */
int64_t imsrg_ctx_neq(imsrg_ctx ctx);
/* info[0..15]: nsp, nhole, nblocks, ncc, npairs, max block dim, max CC rows,
 * device bytes per operator, vsize, ccsize, cwsize, flop_pphh, flop_ph
 * (as doubles: algorithmic GEMM flop per commutator), ph recoupling J' terms,
 * upper-triangle two-body elements, FP64 tile shapes chosen by the autotuner
 * (4*(pp+1) + (ph+1), shapes -1 untuned, 0 64x64, 1 128x64) */
int imsrg_ctx_info(imsrg_ctx ctx, double *info16);
int imsrg_ctx_set_precision(imsrg_ctx ctx, int prec);
int imsrg_ctx_synchronize(imsrg_ctx ctx);
/* Graph mode (no counterpart in the reference): every commutator the context
 * evaluates -- directly or inside BCH / Magnus series and flows -- is captured
 * once per distinct (operands, hermiticities, term mask, instance set,
 * precision) into a hipGraph and replayed.  Off while profiling is on. */
int imsrg_ctx_set_graphs(imsrg_ctx ctx, int enable);
/* FP64 tile shapes the autotuner chose on the context's first FP64 commutator
 * (-1 not yet, 0 64x64, 1 128x64) and the four timings (ms: pp/hh 64x64,
 * pp/hh 128x64, Pandya product 64x64, 128x64) */
int imsrg_ctx_tuning(imsrg_ctx ctx, int *shape_pp, int *shape_ph, float *ms4);
/* fix the FP64 tile shapes (0 or 1 each), or -1 to tune again on the next commutator */
int imsrg_ctx_set_tile_shapes(imsrg_ctx ctx, int shape_pp, int shape_ph);
int imsrg_ctx_graph_stats(imsrg_ctx ctx, long long *builds, long long *launches);
/* Commutator concurrency (no counterpart in the reference): with enable = 1
 * (the default) the pp/hh chain of a commutator (E0, 122, pp/hh GEMM,
 * one-body partial sums) runs on a second stream of the context, concurrently
 * with the ph chain (Pandya transforms, Z GEMM, J' terms); both join before
 * the ph apply.  Results are identical either way.  Graph capture uses one stream. */
int imsrg_ctx_set_concurrency(imsrg_ctx ctx, int enable);
/* the hipStream_t the context launches on (as void*) */
void *imsrg_ctx_stream(imsrg_ctx ctx);

/* ---- operators (sq_op, basic_IMSRG.f90:50-76) ---------------------------- */
int imsrg_op_create(imsrg_ctx ctx, int herm, imsrg_op *out);     /* duplicate_sq_op :2736 */
int imsrg_op_destroy(imsrg_op op);
int imsrg_op_set_herm(imsrg_op op, int herm);
int imsrg_op_get_herm(imsrg_op op, int *herm);
/* vec: nbatch vectorize-order vectors back to back, n = neq * nbatch */
int imsrg_op_upload(imsrg_op op, const double *vec, int64_t n);     /* repackage :3857-3918 */
int imsrg_op_download(imsrg_op op, double *vec, int64_t n);         /* vectorize :3794-3854 */
int imsrg_op_copy(imsrg_op src, imsrg_op dst);                      /* copy_sq_op :2942-2969 */
int imsrg_op_clear(imsrg_op op);                                    /* clear_sq_op :3537-3562 */
int imsrg_op_scale(imsrg_op op, double a);                          /* scale_sq_op :3510-3534 */
/* C = ax*A + bx*B (C may alias A or B)                                add_sq_op :3450-3478 */
int imsrg_op_add(imsrg_op A, double ax, imsrg_op B, double bx, imsrg_op C);
int imsrg_op_get_e0(imsrg_op op, double *e0);
int imsrg_op_set_e0(imsrg_op op, double e0);                        /* every instance */
int imsrg_op_norm(imsrg_op op, double *out);                        /* mat_frob_norm :3921-3952 */
int imsrg_mbpt2(imsrg_op H, double *out);                           /* mbpt2 :4030-4064 */
/* the same scalars for every instance of a batched context (out[nbatch]) */
int imsrg_op_e0_batch(imsrg_op op, double *e0);
int imsrg_op_norm_batch(imsrg_op op, double *out);
int imsrg_mbpt2_batch(imsrg_op H, double *out);
int imsrg_normal_order(imsrg_op H);                                 /* normal_order :2375-2506 */

/* calc_HF (HF_mod.f90:14-106) for a two-body Hamiltonian (no three-body force):
 * H holds the bare Hamiltonian (before normal ordering) and is replaced by the
 * normal-ordered Hamiltonian in the Hartree-Fock basis: E0 = e_HF (:376-405),
 * one-body = Fock matrix (transform_1b_to_HF :408-491), two-body transformed
 * (transform_2b_to_HF :494-619).  coefs (nsp*nsp doubles, may be NULL)
 * receives D(old orbit a, new orbit i) at coefs[a*nsp + i]; iterations and
 * e_hf may be NULL.  The self-consistency loop, the (l,j,tz)-block
 * eigenproblems (diagonalize_blocks, basic_IMSRG.f90:2163-2180) and the
 * transforms run on the device.  nbatch = 1; groups of at most 16 orbits. */
int imsrg_hartree_fock(imsrg_op H, double *coefs, int *iterations, double *e_hf);
/* imsrg_hartree_fock without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as imsrg_hartree_fock (bit for bit), but every work
 * buffer lives in one static 16 MiB device array per GPU (held until the
 * process ends): the self-consistency state (5 nsp^2 + 3 nsp doubles and a
 * few more) and windows of whole channel blocks for the two-body transform.
 * Returns IMSRG_EINVAL when the state or one channel block does not fit.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two host threads, two contexts on one device) race. */
int imsrg_hartree_fock_noalloc(imsrg_op H, double *coefs, int *iterations, double *e_hf);
/* sqop_to_HF (HF_mod.f90:108-136) for a scalar observable O (bare): the
 * one- and two-body transforms with the coefficients of imsrg_hartree_fock,
 * then normal_order. */
int imsrg_observable_to_hf(imsrg_op O, const double *coefs);

/* ---- commutator engine --------------------------------------------------
 * RES = [L, R] restricted to the terms in mask, in the reference's call
 * contract (IMSRG_MAGNUS.f90:207-219): one-body and two-body parts of RES are
 * fully overwritten with the sum of the selected terms; when mask contains
 * T110 or T220, RES E0 is assigned 110+220 (else left untouched, as in
 * MAGNUS_EXPAND).  RES herm must be -herm(L)*herm(R).  RES must not alias L or R. */
int imsrg_commutator(imsrg_ctx ctx, imsrg_op L, imsrg_op R, imsrg_op RES, unsigned mask);
/* scalar terms (commutators.f90:10-33, 399-432) */
int imsrg_commutator_110(imsrg_op L, imsrg_op R, double *out);
int imsrg_commutator_220(imsrg_op L, imsrg_op R, double *out);
/* Pandya transform of op (calculate_cross_coupled, cross_coupled.f90:1190-1300),
 * copied to the host column-major per non-empty CC block in reference order */
int imsrg_cross_coupled(imsrg_op op, double *host_out, int64_t n);
int64_t imsrg_cross_coupled_size(imsrg_ctx ctx);
/* commutator_222_ph(LCC, RCC, RES, WCC, jbas) (commutators.f90:771-925) on
 * Pandya matrices the caller made (calculate_cross_coupled): lcc and rcc in
 * the layout imsrg_cross_coupled writes (n = imsrg_cross_coupled_size),
 * hL/hR their hermiticities (cc_mat%herm).  RES (herm -hL*hR) receives the
 * 222_ph term alone, as imsrg_commutator with mask IMSRG_T222PH: one- and
 * two-body parts overwritten, E0 untouched.  The workspace WCC of the
 * reference is the context's.  nbatch = 1. */
int imsrg_commutator_222_ph_cc(imsrg_ctx ctx, const double *lcc, int hL, const double *rcc, int hR, int64_t n,
                               imsrg_op RES);

/* Adjoint commutator (no counterpart in the reference; the gradient path):
 * Ybar = d <Cbar, [P, Y]> / dY for Y of hermiticity hY, for the terms in
 * mask, where <,> is the Frobenius product over E0, the full one-body matrix
 * and every full (J,P,Tz) channel matrix.  Cbar (herm -herm(P)*hY) is used
 * after projection onto its hermiticity; Ybar comes out projected onto hY,
 * E0 = 0.  Every term is the exact transpose of its forward kernel.  By
 * antisymmetry d<Cbar,[X,Y]>/dX = -(the same call with P = Y).  nbatch = 1. */
int imsrg_commutator_adjoint(imsrg_ctx ctx, imsrg_op P, imsrg_op Cbar, imsrg_op Ybar, int hY, unsigned mask);

/* Gradient of E0(s_N) after nsteps Magnus steps (magnus_decouple's step loop,
 * IMSRG_MAGNUS.f90:98-126, White or atan generator) with respect to the
 * normal-ordered H0, by the discrete adjoint of the flow: every nested
 * commutator of the BCH and Magnus series is transposed (term counts held at
 * the forward flow's), Omega(s) is kept at ncheck checkpoints and each segment
 * is recomputed.  *e0 = E0(s_N); H0bar = dE0/dH0 (Frobenius over the full
 * layout, hermitian projection).  nbatch = 1.  No counterpart in the reference. */
int imsrg_flow_gradient(imsrg_ctx ctx, imsrg_op H0, int gen_kind, double ds, int nsteps, int ncheck,
                        double *e0, imsrg_op H0bar);
/* transpose of imsrg_normal_order: Hbar = d<Cbar, NO(H)>/dH (E0 of Hbar = 0) */
int imsrg_normal_order_adjoint(imsrg_op Cbar, imsrg_op Hbar);
/* a cotangent as the gradient with respect to the vectorize stream (the
 * transpose of repackage, basic_IMSRG.f90:3857-3918), n = neq * nbatch */
int imsrg_op_download_cotangent(imsrg_op op, double *vec, int64_t n);

/* ---- generators (external build_generator callbacks) ---------------------- */
int imsrg_generator(imsrg_ctx ctx, int kind, imsrg_op H, imsrg_op ETA);
/* build_gs_atanxx (generators.f90:271-374): denominators from Hold, numerators from H */
int imsrg_generator_atanxx(imsrg_ctx ctx, imsrg_op Hold, imsrg_op H, imsrg_op ETA);

/* ---- Magnus driver (IMSRG_MAGNUS.f90) -------------------------------------- */
/* HS = e^G H e^-G by the nested-commutator series (BCH_EXPAND :149-234);
 * nterms[k] = commutators evaluated for instance k (nbatch entries, each
 * instance truncates on its own).  Returns IMSRG_ECLIP after 39 terms. */
int imsrg_bch_expand(imsrg_ctx ctx, imsrg_op HS, imsrg_op G, imsrg_op H, int *nterms);
/* DG <- dOmega/ds series (MAGNUS_EXPAND :622-693); nterms as above */
int imsrg_magnus_expand(imsrg_ctx ctx, imsrg_op DG, imsrg_op G, int *nterms);
/* Op <- e^G Op e^-G for a rank-0 observable (transform_sq_op_BCH :131-146, the
 * BCH_EXPAND branch); G anti-hermitian, Op of either hermiticity; nterms as
 * above */
int imsrg_transform_bch(imsrg_op Op, imsrg_op G, int *nterms);
/* DH = [eta(HS), HS], eta the generator `kind` of HS (dHds, IMSRG_ODE.f90:416-463:
 * the derivative the reference's ODE solver evaluates); DH herm = herm(HS) */
int imsrg_dhds(imsrg_op HS, int kind, imsrg_op DH);
/* full flow (magnus_decouple :15-128, no checkpointing; nbatch = 1): HS in = normal-ordered
 * H(0), out = decoupled H(s); G out = Omega(s).  Per-step records (arrays of
 * length max_steps, any may be NULL): s, E0, E0+MBPT2, |MBPT2|, #Magnus and
 * #BCH commutators.  *steps = steps taken. */
int imsrg_magnus_flow(imsrg_ctx ctx, imsrg_op HS, imsrg_op G, int gen_kind, double ds,
                      int max_steps, double *s_out, double *e0_out, double *e0mbpt2_out,
                      double *crit_out, int *nmag_out, int *nbch_out, int *steps);

/* The same flow one step at a time: set-up of magnus_decouple (:31-96; Omega
 * reset to 0, DG = generator(HS), H(0) kept, crit = |MBPT2(HS)|), then one
 * pass of its step loop (:98-126) per imsrg_flow_step.  rec8 receives, per
 * instance (8 * nbatch doubles), s, E0, E0+MBPT2, |MBPT2|, #Magnus, #BCH
 * commutators, steps taken, 1 if the instance stepped in this call.  In a
 * batched context an instance whose |MBPT2| reached float32(1e-6) stops
 * stepping while the others go on; imsrg_flow_converged reports all done.
 * HS and G stay owned by the caller and must outlive the flow. */
typedef struct imsrg_flow_s *imsrg_flow;
int imsrg_flow_create(imsrg_ctx ctx, imsrg_op HS, imsrg_op G, int gen_kind, double ds, imsrg_flow *out);
/* The checkpoint-restart branch of magnus_decouple (IMSRG_MAGNUS.f90:83-97): HS
 * in = normal-ordered H(0), G in = Omega(s) (for instance read back with
 * imsrg_op_read_stream from a write_omega_checkpoint file); HS becomes
 * BCH(Omega, H(0)), the generator is rebuilt, and the flow goes on from s with
 * nint(s/ds) steps counted, taking at least one more step. */
int imsrg_flow_restart(imsrg_ctx ctx, imsrg_op HS, imsrg_op G, int gen_kind, double ds, double s,
                       imsrg_flow *out);
int imsrg_flow_step(imsrg_flow flow, double *rec8);
int imsrg_flow_converged(imsrg_flow flow, int *yes);   /* |MBPT2| <= float32(1e-6) */
int imsrg_flow_destroy(imsrg_flow flow);

/* ---- operator files ---------------------------------------------------------
 * The reference's `.bin` operator dumps and Omega checkpoints
 * (write_twobody_operator / write_omega_checkpoint, basic_IMSRG.f90:3069-3226;
 * read_twobody_operator / read_omega_checkpoint :3228-3420): the vectorize
 * stream of an operator, neq FP64 values, no header.  Single-instance contexts;
 * a file of another length is rejected with IMSRG_EINVAL. */
int imsrg_op_write_stream(imsrg_op op, const char *path);
int imsrg_op_read_stream(imsrg_op op, const char *path);

/* ---- interaction files (interaction_IO.f90) --------------------------------
 * The two-body matrix elements a file lists are set into H (herm +1, single
 * instance) as the reference's reader sets them: each element and its
 * hermitian mirror, in file order (a repeated element keeps the last value);
 * everything else in H is kept, except that IMSRG_INT_ME2B also sets E0 and
 * the one-body part from its one-body file.  Orbit labels follow each
 * format's convention against the context's orbits (matched by n, l, 2j,
 * 2tz where the file names orbits by quantum numbers). */
#define IMSRG_INT_ASCII 0  /* read_interaction      interaction_IO.f90:8-107: 8 header lines, then
                              "Tz Par J a b c d V" (orbits 1-based, J doubled) */
#define IMSRG_INT_GZ 1     /* read_gz               :110-231: gzip, a count, then one element per line */
#define IMSRG_INT_ME2J 2   /* read_me2j_interaction :282-568: gzip, isospin-coupled values ten per line
                              over the file's basis (hk order up to emax_file, lmax_file), to pn */
#define IMSRG_INT_ME2B 3   /* read_me2b_interaction :570-1151: gzip two-body blocks, then E0 and the
                              one-body part from path_1b */
typedef struct {
  double hw;          /* H%hospace (MeV) */
  double com_hw;      /* H%com_hw */
  double lawson_beta; /* H%lawson_beta: adds beta (p1.p2 + com_hw^2/hw^2 r1.r2) hw/A */
  int A;              /* H%Aneut + H%Aprot */
  int htype;          /* ham_type: 1 subtracts the centre-of-mass kinetic term p1.p2 hw/A */
  int emax_file;      /* IMSRG_INT_ME2J: the file's eMax and lMax */
  int lmax_file;
} imsrg_interaction_params;
int imsrg_read_interaction(imsrg_op H, int format, const char *path, const char *path_1b,
                           const imsrg_interaction_params *p);
/* The same readers on a host vectorize stream (basis_IMSRG.f90:3794-3918
 * order, neq = the basis' stream length) instead of a device operator: vec in
 * = the operator before reading (zeros for a fresh one), out = after.  The
 * basis is given as for imsrg_ctx_create; no device is touched, so a Fortran
 * caller can repackage the result into an sq_op. */
int imsrg_read_interaction_vec(int nsp, const int *n, const int *l, const int *jj, const int *tz, const int *con,
                               int format, const char *path, const char *path_1b,
                               const imsrg_interaction_params *p, double *vec, int64_t neq);

/* ---- spherical tensor operators and the scalar-tensor commutator ---------
 * A rank > 0 operator (allocate_tensor, basic_IMSRG.f90:644-907) on a
 * single-instance context: rank and dpar doubled as the reference stores them
 * (op%rank = 2 lambda, op%dpar = 0 or 2), herm +1 or -1.  Its host stream
 * (imsrg_tensor_upload / download, imsrg_tensor_neq doubles) is the
 * reference's storage flattened in declaration order, since the reference has
 * no vectorize for tensors: E0, fhh (A x A), fpp (P x P), fph (P x A), then for
 * every tensor block in allocate_tensor's order (J1, J2 = max(|J1-rank|, J1)
 * .. J1+rank, Tz = -1..1, P1 = 0..1) its sub-blocks tgam(1..9), each
 * column-major.  The (hole, particle) one-body entries are implied by fph
 * with f_tensor_elem's phase herm * (-1)^((jh-jp)/2). */
typedef struct imsrg_tensor_s *imsrg_tensor;
int imsrg_tensor_create(imsrg_ctx ctx, int rank, int dpar, int herm, imsrg_tensor *out);
int imsrg_tensor_destroy(imsrg_tensor t);
int64_t imsrg_tensor_neq(imsrg_tensor t);
int imsrg_tensor_set_herm(imsrg_tensor t, int herm);
int imsrg_tensor_get_herm(imsrg_tensor t, int *herm);
int imsrg_tensor_upload(imsrg_tensor t, const double *vec, int64_t n);
int imsrg_tensor_download(imsrg_tensor t, double *vec, int64_t n);
int imsrg_tensor_copy(imsrg_tensor src, imsrg_tensor dst);    /* copy_sq_op   basic_IMSRG.f90:2942-2969 */
int imsrg_tensor_clear(imsrg_tensor t);                         /* clear_sq_op  :3537-3562 */
/* C = ax A + bx B                                                 append_operator :3481-3510 */
int imsrg_tensor_add(imsrg_tensor A, double ax, imsrg_tensor B, double bx, imsrg_tensor C);
int imsrg_tensor_norm(imsrg_tensor t, double *out);             /* mat_frob_norm, tensor branch :3921-3952 */

/* Terms of the scalar-tensor commutator [G, T] (TS_commutators.f90; G scalar,
 * T the tensor).  imsrg_tensor_commutator runs the terms in mask in the order
 * BCH_TENSOR calls them (IMSRG_MAGNUS.f90:427-446): 111 and 122 overwrite
 * RES's one- and two-body parts, the others add to RES; RES's herm is the
 * caller's (the reference's tests use -herm(G) herm(T), BCH_TENSOR herm(T)). */
#define IMSRG_TS111 (1u << 0)      /* TS_commutator_111       TS_commutators.f90:10-55     */
#define IMSRG_TS121 (1u << 1)      /* TS_commutator_121       :58-303                      */
#define IMSRG_TS211 (1u << 2)      /* TS_commutator_211       :307-552                     */
#define IMSRG_TS122 (1u << 3)      /* TS_commutator_122       :555-663                     */
#define IMSRG_TS212 (1u << 4)      /* TS_commutator_212       :666-806                     */
#define IMSRG_TS222PPHH (1u << 5)  /* TS_commutator_222_pp_hh :1070-1456                   */
#define IMSRG_TS221 (1u << 6)      /* TS_commutator_221       :809-1067 (ladder products recomputed) */
#define IMSRG_TS222PH (1u << 7)    /* TS_commutator_222_ph    :1459-1849 with calculate_single_pandya,
                                      cross_coupled.f90:1416-1634 */
#define IMSRG_TSALL 0xFFu
int imsrg_tensor_commutator(imsrg_op G, imsrg_tensor T, imsrg_tensor RES, unsigned mask);
/* BCH_TENSOR (IMSRG_MAGNUS.f90:389-465): HS <- sum_n ad_G^n HS / n!, n up to
 * 29, stopping after the first term with ||term|| / n! < float32(1e-6);
 * nterms = commutators evaluated.  Replaces transform_sq_op_BCH for rank > 0
 * operators (:131-146). */
int imsrg_bch_tensor(imsrg_op G, imsrg_tensor HS, int *nterms);
/* The reference's argument lists that carry Pandya matrices or ladder
 * products instead of G, for a caller that holds them:
 * TS_commutator_211(LCC,R,RES) and TS_commutator_222_ph(LCC,RCC,R,RES)
 * (TS_commutators.f90:307, 1459) from G's Pandya matrices in the reference's
 * cc_mat layout (the imsrg_cross_coupled stream, hG = LCC%herm; mask within
 * IMSRG_TS211 | IMSRG_TS222PH; RCC is workspace there); the ladder products
 * w1 (pp) and w2 (hh) that TS_commutator_222_pp_hh leaves for
 * TS_commutator_221 (:1070-1456, two-body parts, the rest zero); and
 * TS_commutator_221(w1,w2,pm,RES) (:809-1067) from given w1, w2. */
int imsrg_tensor_commutator_cc(imsrg_tensor T, const double *gcc, int hG, int64_t n, imsrg_tensor RES,
                               unsigned mask);
int imsrg_tensor_ladders(imsrg_op G, imsrg_tensor T, imsrg_tensor W1, imsrg_tensor W2);
int imsrg_tensor_commutator_221(imsrg_tensor W1, imsrg_tensor W2, int pm, imsrg_tensor RES);

/* ---- measurement (no counterpart in the reference, which only prints
 * omp_get_wtime, main_IMSRG.f90:46,308-309) ---------------------------------
 * While enabled, every engine kernel launch is bracketed by HIP events on the
 * context stream; profile_read synchronizes the stream and returns the summed
 * milliseconds and launch counts per kernel class. */
#define IMSRG_PROF_GEMM_PPHH 0  /* segmented GEMM of 222_pp_hh (+ 221 intermediate) */
#define IMSRG_PROF_GEMM_PH 1    /* segmented GEMM of the Pandya product Z */
#define IMSRG_PROF_PANDYA 2     /* Pandya transforms */
#define IMSRG_PROF_122 3        /* 122 rows */
#define IMSRG_PROF_PH_APPLY 4   /* ph recoupling: 6j-weighted sums into RES */
#define IMSRG_PROF_PH_TERMS 5   /* ph recoupling: gather of Z terms */
#define IMSRG_PROF_MIRROR 6     /* symmetrise + mirror pass */
#define IMSRG_PROF_ONEBODY 7    /* 111 + 121 + 221 */
#define IMSRG_PROF_E0 8
#define IMSRG_PROF_NORM 9
#define IMSRG_PROF_AXPBY 10
#define IMSRG_PROF_GENERATOR 11
#define IMSRG_PROF_SLICE 12      /* int8 slicing of the GEMM operands (IMSRG_PREC_INT8X5 / INT8X4) */
#define IMSRG_PROF_NCLASS 16
int imsrg_ctx_profile(imsrg_ctx ctx, int enable);
/* bracket only the classes in mask (bit k = class k; 0 turns profiling off):
 * fewer events on the stream while a run is timed */
int imsrg_ctx_profile_classes(imsrg_ctx ctx, unsigned mask);
int imsrg_ctx_profile_read(imsrg_ctx ctx, double *ms, long long *launches, int nclass, int reset);

#ifdef __cplusplus
}
#endif
#endif
