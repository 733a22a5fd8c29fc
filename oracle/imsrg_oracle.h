/*
 * CPU FP64 restatement of the reference IMSRG(2) Magnus commutator engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py use it as the checker.  The product
 * library (eom-imsrg_full_amd/csrc) never links or calls it.
 *
 * Storage follows the reference exactly: per (J,P,Tz) block six column-major
 * sub-arrays gam1..gam6 = Vpppp, Vppph, Vpphh, Vphph, Vhhhh, Vphhh
 * (im-SRG/basic_IMSRG.f90:50-76, allocate_blocks :477-641), one-body fhh/fph/fpp.
 * All angular momenta are doubled (2j, 2J, 2tz), orbits are 0-based here.
 */
#ifndef IMSRG_ORACLE_H
#define IMSRG_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_basis or_basis;
typedef struct or_op or_op;
typedef struct or_cc or_cc;

/* basis: orbits in the caller's order; con[i]=1 for holes */
or_basis *or_basis_create(int nsp, const int *n, const int *l, const int *jj,
                          const int *tz, const int *con);
void or_basis_free(or_basis *B);
int64_t or_basis_neq(const or_basis *B);
int or_basis_nblocks(const or_basis *B);
/* block q: J(doubled),P,Tz,npp,nph,nhh */
void or_basis_block_info(const or_basis *B, int q, int *info6);
/* pair list of class c (1=pp,2=ph,3=hh) of block q: 2*n ints (0-based orbits) */
const int *or_basis_block_pairs(const or_basis *B, int q, int c);
int or_basis_ncc(const or_basis *B);
void or_basis_cc_info(const or_basis *B, int qcc, int *info3); /* J, rlen, nph */
double or_sixj(const or_basis *B, int j1, int j2, int j3, int j4, int j5, int j6);
double or_ninej(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8, int j9);

/* operators */
or_op *or_op_create(const or_basis *B, int herm);
void or_op_free(or_op *op);
void or_op_set_herm(or_op *op, int herm);
int or_op_get_herm(const or_op *op);
void or_op_from_vec(or_op *op, const double *v); /* repackage */
void or_op_to_vec(const or_op *op, double *v);   /* vectorize */
double or_op_get_e0(const or_op *op);
void or_op_set_e0(or_op *op, double e0);
double or_f_elem(int a, int b, const or_op *op);
double or_v_elem(int a, int b, int c, int d, int J, const or_op *op);

/* commutator terms (im-SRG/commutators.f90) */
double or_comm110(const or_op *L, const or_op *R);
double or_comm220(const or_op *L, const or_op *R);
void or_comm111(const or_op *L, const or_op *R, or_op *RES);
void or_comm121(const or_op *L, const or_op *R, or_op *RES);
void or_comm122(const or_op *L, const or_op *R, or_op *RES);
void or_comm222_pp_hh(const or_op *L, const or_op *R, or_op *RES, or_op *w1, or_op *w2);
void or_comm221(const or_op *L, const or_op *R, or_op *RES, const or_op *w1, const or_op *w2);
or_cc *or_cc_create(const or_basis *B);
void or_cc_free(or_cc *cc);
void or_calc_cross_coupled(const or_op *op, or_cc *cc);
/* CC matrix of block qcc, column-major rlen x nph */
const double *or_cc_block(const or_cc *cc, int qcc);
void or_comm222_ph(const or_cc *LCC, const or_cc *RCC, or_op *RES, or_cc *WCC_scratch);
/* full commutator in the reference call order (IMSRG_MAGNUS.f90:207-219);
 * with_e0: assign RES E0 from 110+220 (BCH) or leave it (MAGNUS_EXPAND) */
void or_commutator(const or_op *L, const or_op *R, or_op *RES, int with_e0);

/* operator algebra (basic_IMSRG.f90:2942-3562) */
void or_add(const or_op *A, double ax, const or_op *B, double bx, or_op *C);
void or_copy(const or_op *H, or_op *op);
void or_clear(or_op *C);
void or_scale(or_op *A, double ax);
double or_frob_norm(const or_op *op);
double or_mbpt2(const or_op *H);
void or_normal_order(or_op *H);

/* generators (generators.f90) kind: 0=white 1=atan 2=w2 */
void or_build_white(const or_op *H, or_op *ETA);
void or_build_atan(const or_op *H, or_op *ETA);
void or_build_atanxx(const or_op *Hold, const or_op *H, or_op *ETA);
void or_build_w2(const or_op *H, or_op *ETA);
void or_build_generator(int kind, const or_op *H, or_op *ETA);

/* Magnus driver (IMSRG_MAGNUS.f90) */
int or_bch_expand(or_op *HS, const or_op *G, const or_op *H); /* returns #commutators, -1 on clip */
int or_magnus_expand(or_op *DG, const or_op *G);                /* returns #commutators */
/* runs the flow; records per step s, E0, E0+mbpt2, crit and commutator counts.
 * returns number of steps (negative on error). */
int or_magnus_flow(or_op *HS, or_op *G, int gen_kind, double ds, int max_steps,
                   double *s_out, double *e0_out, double *emp2_out, double *crit_out,
                   int *nmag_out, int *nbch_out);

/* brute-force J-coupled sums (brute_force_testing.f90:1765-2076) */
double or_brute_0body(const or_op *A, const or_op *B);
double or_brute_1body(const or_op *A, const or_op *B, int a, int b);
double or_brute_2body(const or_op *A, const or_op *B, int a, int b, int c, int d, int J);

void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
