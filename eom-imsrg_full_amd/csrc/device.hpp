// device.hpp -- device-side view of the layout and the kernel launchers.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace imsrg {

// Work items split into 8 contiguous ranges, one per XCD: workgroup w runs on
// XCD w % 8 and takes item start[w % 8] + w / 8, so items that share data
// (rows of one channel block, terms of one quadruple run) share one L2.
// Launch grids are 8 * len; slots past an XCD's range exit at once.
struct XcdSplit {
  int start[9];
  int len;
};

// Read-only tables resident in HBM, passed by value to every kernel.
struct DevLayout {
  int nsp, nhole, npart, nblocks, ncc, jtot_max, nhalf, nJ, npairs, nob, nphp, nccol;
  const int *jj, *l, *tz, *con;
  const int *grp_of, *grp_start, *grp_cnt, *grp_list;
  const int *bJ, *bnp, *bnb, *bnh, *bdim;
  const long long *boff;
  const int *pbase, *pair_a, *pair_b, *pair_q;
  const int *pidx;
  const int *nb_start, *nb_pos, *nb_x, *nb_i;
  const double *nb_coef;
  const struct Set122 *s122_sets;  // 122 row sets
  const int *s122_rowl;            // per set row: its row in the block
  const int2 *s122_items;          // set kernel work items (set, 64-column chunk)
  const int2 *s122_sx;             // per slot: f-matrix indices of its two entries
  const double *s122_sc;           // per slot: their two coefficients
  int n122_items, n122_slots;
  const int *s122_frows;                                 // rows of the sets past 32 rows (row kernel)
  int n122_frows;
  // Hartree-Fock: groups in the reference's block order, same-group orbit
  // pairs (all; a <= c), per pair row the rows the pair-basis transform couples
  const int *hf_gorder, *hf_b, *hf_d, *hf_a, *hf_c, *hf_nb_start, *hf_nb;
  int ngrp, n_hf_bd, n_hf_ac;
  const int *ob_p, *ob_q;
  const int *php_a, *php_i;  // same-(l,j,tz) (particle, hole) pairs
  int nsame, ob_nchunk;      // one-body outputs with 121/221 terms; 256-term chunks per output
  const int *same_ob, *ob_same;  // their one-body indices; per one-body output its same index or -1
  const int *ccJ, *ccrlen, *ccnph;
  const long long *ccoff, *cwoff;
  const int *ccrow, *ccrbase, *ccrow_a, *ccrow_b, *cccbase, *cccol_p, *cccol_h, *cccol_q;
  const double *sixj;
  const double *ap6_val;       // packed non-zero 6j runs of the ph apply, per row key (ji, jj, J)
  const int *ap6_base;
  int ap6_max;
  const unsigned *ph_quads;  // (i<<24|j<<16|k<<8|l) of every upper (ij,kl), sorted by (i,k,j,l)
  long long nquad;
  const unsigned *ph_terms;  // J' terms: quad<<6 | side<<5 | m, in quadruple order
  long long nphterm;
  const unsigned *ph_zblk;     // per J' term: half of a 2x2 Z block holding its two elements (units of 2 doubles)
  const unsigned char *ph_zsel;  // per J' term: element positions, mirror, phase, side, valid
  const unsigned *ph_two;        // the terms with a second half-block load
  long long nphtwo;
  const unsigned *ph_toff;   // per upper element (row-major per block): first T1 term of its quadruple
  const unsigned long long *ph_desc;  // per upper element: T offset, run lengths, 6j runs, flags (layout.cpp)
  const long long *ph_ubase; // per global pair row r: index of its first upper element
  const int4 *mirror_tiles;  // (block, tile row, tile col) with row <= col, 32x32 tiles
  int nmirror;
  long long f_off, v_off, op_size, vsize, ccsize, cwsize;  // per-instance strides
  const struct PhChunk *ph_chunks;  // k_ph_apply work items, rows in (i, j, J) order
  XcdSplit xs_rows, xs_terms, xs_apply;  // k_122 rows (s122_frows), k_ph_terms chunks of 1024 terms, ph_chunks
};

// A 122 row set: rows s122_rowl[rs .. rs + n) of block q (dimension dim,
// matrix at bo doubles into the two-body part); its n x 2n weights start at
// ab in the weight table (-1: the row kernel takes the set)
struct alignas(16) Set122 {
  int rs, n, dim, q;
  long long bo, ab;
};

// One k_ph_apply work item: up to 64 consecutive upper outputs of one row
// (ij; J), with what the kernel would otherwise look up through four levels
// of dependent table reads.
struct PhChunk {
  long long vrow;   // operator offset of the first output
  long long dbase;  // ph_desc index of the first output
  int n;            // outputs, 1..64
  int b0, len;      // the row key's packed 6j list (ap6_val) start and length
  int half;         // 1 when i == j (the outputs carry 1/sqrt 2)
  long long tbase;  // the row's first J' term: descriptors hold term offsets relative to it
};

// A plain GEMM problem C = A * B (A M x K, B K x N, C M x N, all row-major)
struct GemmProb {
  int M, N, K, lda, ldb, ldc;
  long long a_off, b_off, c_off;
};

// One K segment of a segmented GEMM: sign * A * B with A taken from operand
// `src` and B from the other one (offsets in doubles from the operand base).
struct SegDesc {
  long long a_off, b_off;
  int K, lda, ldb, src;
  double sign;
};
// One square output block: C (M x M) = sum of nseg segments; the running sum
// after the first nmid segments may also be stored (w_off into a side buffer).
struct SegProb {
  int M, nseg, nmid, pad;
  long long c_off, w_off;
  SegDesc seg[4];
};

// ---- int8-sliced products (the Ozaki scheme on the int8 matrix cores).  An
// operand row r is held as x = s_r * sum_{i=1..5} a_i 2^(-7i): a power-of-two
// row scale s_r (max |x| / s_r in [1/4, 1/2)) and five int8 slices a_i in
// [-64, 64], zero-padded to 64 rows / 64 columns.  V rows: the five slices of
// each 64-column chunk of a row lie together, slice i of column k of row r at
// byte r * 5 ld + (k / 64) * 320 + 64 i + k % 64 of the block, so one K chunk
// of a row is 320 contiguous bytes.  Pandya rows: fragment order, slice i of
// columns 32 s .. 32 s + 31 of the 32 rows of group g in one 1 KB piece at
// byte ((g * ld / 32 + s) * 5 + i) * 1024 of the block (lane l of a matrix
// instruction reads the 16 bytes at 16 l: row 32 g + l % 32, columns
// 32 s + 16 (l / 32) .. +15).  A product keeps the slice pairs with
// i + j <= 6 (fifteen int8 products, exact int32 sums per level i + j).  The
// four-plane form (IMSRG_PREC_INT8X4) keeps slices 1..4, laid out the same
// way with 4 in place of 5 (each block in the first four fifths of its
// region), and the ten pairs i + j <= 5.
constexpr int kI8Planes = 5;
struct I8Seg {
  int kofs, K, src;  // K window of the segment in the block's rows (multiples of 64)
  int K32;           // its columns rounded up to 32: a last half chunk skips its second K step
  double sign;
};
// C (M x M) = sum of nseg segments  sign * A_src * B_other^T  over plane rows
// of the block (A rows m, B rows n); c_off, w_off, nmid as in SegProb.
struct I8Prob {
  int M, nseg, nmid, ld;  // ld: K extent of the block's rows (bytes per slice; the row stride is 5 ld)
  long long c_off, w_off;
  long long p_off;        // byte offset of the block's row 0
  long long s_off;        // index of the block's first row scale
  I8Seg seg[4];
};
// Where the slices of an operand go: per channel block (pp/hh columns of the
// V rows: pp at 0, hh at 64*ceil(np/64)) and per Pandya block (all columns).
struct I8Layout {
  const long long *v_poff, *v_soff;
  const int *v_ld;
  const long long *c_poff, *c_soff;
  const int *c_ld;
  const int4 *c_items;   // (Pandya block, first row, first column, partial-max base) per 64 x 256 item
  int n_items;
  long long pm_size;      // partial row maxima per instance (64 per strip and 256-column chunk)
  long long v_bytes, c_bytes;  // slices of one operand's V rows / one Pandya matrix, per instance
  long long v_sr, c_sr;  // row scales per instance
};

// term bits (mirror the C ABI)
enum : unsigned {
  T110 = 1u << 0, T111 = 1u << 1, T121 = 1u << 2, T122 = 1u << 3,
  T220 = 1u << 4, T221 = 1u << 5, T222PPHH = 1u << 6, T222PH = 1u << 7
};

// Kernel-attached timing: a profiling bracket around one GEMM launch sets
// these events, and the launch records them with the kernel dispatch itself
// (hipExtLaunchKernelGGL) instead of as separate markers on the stream, which
// would leave the GPU idle before and after the kernel.  Consumed (cleared)
// by the next launch_k.
struct KernelEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local KernelEvents g_kernel_events;
template <typename F, typename... Args>
inline void launch_k(F kernel, dim3 g, dim3 b, unsigned shm, hipStream_t s, Args... args) {
  const KernelEvents e = g_kernel_events;
  g_kernel_events = KernelEvents();
  if (e.start)
    hipExtLaunchKernelGGL(kernel, g, b, shm, s, e.start, e.stop, 0u, args...);
  else
    hipLaunchKernelGGL(kernel, g, b, shm, s, args...);
}

// ---- launchers (kernels.hip) ----
// Every launcher covers the instances listed in the device table `act`
// (nact entries; act == nullptr with nact == 1 means instance 0).  Instance k
// of an operator starts at k*op_size, of a Pandya matrix at k*ccsize, of Z at
// k*cwsize, of the ph terms at k*nphterm, of the 221 store at k*vsize, of the
// reduction partials at k*kPartials and of the scalar slots at k*kScal.
void launch_fill(hipStream_t s, double *p, long long n, double v, const int *act = nullptr, int nact = 1,
                 long long stride = 0);
void launch_axpby(hipStream_t s, long long n, double a, const double *A, double b, const double *B, double *C,
                  const int *act, int nact, long long stride);
void launch_copy(hipStream_t s, long long n, const double *A, double *C, const int *act, int nact,
                 long long stride);
// mode: 0 FP64 MFMA (64 x 64 tiles), 1 split-FP32 (3 FP32-MFMA products), 2 split-BF16
// (6 BF16-MFMA products), 3 FP64 MFMA on 128 x 64 tiles (tile list built for that shape).
// 222_pp_hh: RES(upper) += L*D*R - R*D*L per channel block; with W, the pp part
// (W1 - pm*W1^T) is also stored to W (upper triangle) for commutator_221.
void launch_seg_gemm_pphh(hipStream_t s, int mode, const SegProb *probs, const int4 *tiles, int ntiles,
                          const double *X0, const double *X1, double *C, double *W, const int *act, int nact,
                          long long sop, long long sw);
// 222_ph: Z = hL*LCC*RCC^T - hR*RCC*LCC^T per Pandya block, both triangles
void launch_seg_gemm_z(hipStream_t s, int mode, const SegProb *probs, const int4 *tiles, int ntiles,
                       const double *LCC, const double *RCC, double *Z, double hL, double hR, const int *act,
                       int nact, long long scc, long long sz);
void launch_pandya(hipStream_t s, const DevLayout &L, const double *op, double *cc, const int *act, int nact);
size_t pandya_lds_bytes(const DevLayout &L);  // per-workgroup LDS of the Pandya kernel
hipError_t pandya_allow_lds(size_t bytes);
// 122: M = A_L R - A_R L over whole channel blocks (diagonal symmetrised);
// launch_sym_mirror later forms RES = M + herm*M^T
void launch_122(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop, double *RES,
                int hres, const int *act, int nact, double *W);
void launch_clear_blocks(hipStream_t s, const DevLayout &L, double *RES, const int *act, int nact);
// ph recoupling: J' terms T(quadruple, J') from Z, then RES(upper) += 6j-weighted sums
void launch_ph_terms(hipStream_t s, const DevLayout &L, const double *Z, double *Tbuf, double hL, double hR,
                     const int *act, int nact);
void launch_ph_apply(hipStream_t s, const DevLayout &L, const double *Tbuf, double *RES, const int *act,
                     int nact);
// upper += herm * (strict lower)^T, then lower = herm * upper (diagonal zeroed for herm = -1).
// Series epilogue (acc non-null): acc += coef * RES over every channel block,
// coef read from *coefp when coefp is non-null (graph replays: the value is
// written before each replay, so one graph serves every term of a series);
// npart non-null: one sum of squares of the final RES per tile pair (nmirror +
// nob partials per instance, the rest written by launch_ob_final)
void launch_sym_mirror(hipStream_t s, const DevLayout &L, double *RES, int herm, const int *act, int nact,
                       double *acc = nullptr, double coef = 0.0, const double *coefp = nullptr,
                       double *npart = nullptr);
// 111 + 121 + 221 one-body outputs in two launches: the 121/221 partial sums
// (Wpp = pp intermediate of launch_seg_gemm_pphh; part: nsame * ob_nchunk
// doubles per instance), then 111 plus the partials into F (and its mirror)
void launch_ob_partial(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop,
                       const double *Wpp, double *part, int hL, int hR, unsigned mask, const int *act, int nact);
void launch_ob_final(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop, const double *part,
                     double *RES, int hres, unsigned mask, const int *act, int nact, double *acc = nullptr,
                     double coef = 0.0, const double *coefp = nullptr, double *npart = nullptr);
// sqrt of the sum of the nmirror + nob partials into out[inst * kScal]
void launch_norm_from_parts(hipStream_t s, const DevLayout &L, const double *npart, double *out, const int *act,
                            int nact);
// E0 = 110 + 220 into RES[inst * rstride]
void launch_e0(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop, double *RES,
               int hL, int hR, unsigned mask, double *partials, const int *act, int nact, long long rstride);
// Frobenius norm (sqrt of the sum of squares of F and V) into out[inst * kScal]
void launch_sumsq(hipStream_t s, const DevLayout &L, const double *op, double *partials, double *out,
                  const int *act, int nact);
void launch_mbpt2(hipStream_t s, const DevLayout &L, const double *op, double *partials, double *out,
                  const int *act, int nact);
void launch_monopole(hipStream_t s, const DevLayout &L, const double *H, double *mono, const int *act,
                     int nact);
void launch_eta(hipStream_t s, const DevLayout &L, const double *Hden, const double *Hnum,
                const double *mono, double *ETA, int kind, const int *act, int nact);
void launch_normal_order(hipStream_t s, const DevLayout &L, const double *T, double *H, const int *act,
                         int nact);
void launch_gather_e0(hipStream_t s, const double *op, long long stride, double *out, const int *act, int nact);
// truncation test of a BCH (mode 0) or Magnus (mode 1) series after a nested
// commutator: instances that stop get -1 in the series table `act`; live[k]
// (host memory) = 1 for the positions that go on
// Hartree-Fock (hf.hip)
constexpr int kHFMaxGroup = 16;
void launch_hf_init(hipStream_t s, const DevLayout &L, const double *op, double *D, double *Tm, double *eig,
                    int reset);
// one self-consistency iteration; state = {converged, iterations}: every kernel
// returns at once when converged, the last one counts and tests crit > conv
void launch_hf_iteration(hipStream_t s, const DevLayout &L, const double *op, double *D, double *rho,
                         const double *Tm, double *Vg, double *Fm, double *eig, double *cpart, double *crit,
                         double *state, double conv);
void launch_hf_finish_1b(hipStream_t s, const DevLayout &L, double *D, const double *M, double *op, int herm,
                         const double *Tm, const double *Vg, double *tv, double *ehf);
void launch_hf_2b(hipStream_t s, const DevLayout &L, const double *D, double *op, double *W, int herm);
// the same transform for the pair rows [r0, r0 + nrows), which must cover whole
// channel blocks: W holds only their blocks, W[0] = element v_off + wofs of a
// full operator layout (wofs = the first block's offset inside V)
void launch_hf_2b_window(hipStream_t s, const DevLayout &L, const double *D, double *op, double *W, int herm,
                         int r0, int nrows, long long wofs);

void launch_series_gate(hipStream_t s, int *act, int nact, double *scal, int mode, double coef, int store,
                        int test, double conv, double tiny, int *live);

// ---- adjoint commutator (adjoint.hip): Ybar += d<C,[P,Y]>/dY term by term
void launch_adj_e0(hipStream_t s, const DevLayout &L, const double *P, const double *C, double *Y, int hP, int hY,
                   unsigned mask);
void launch_adj_onebody(hipStream_t s, const DevLayout &L, const double *P, const double *C, double *Y, int hP,
                        int hY, int hC, unsigned mask);
void launch_adj_122(hipStream_t s, const DevLayout &L, const double *P, const double *C, double *Y, int hC);
void launch_adj_pphh(hipStream_t s, const DevLayout &L, const double *Q, double *Y, int hP, int hC);
void launch_adj_ph_apply(hipStream_t s, const DevLayout &L, const double *C, double *Tbar, int hC);
void launch_adj_ph_terms(hipStream_t s, const DevLayout &L, const double *Tbar, double *Zbar, double sigma);
void launch_adj_ph_sbt(hipStream_t s, const DevLayout &L, const int *ccrlen_host, double *Z, int hP, int hY);
void launch_adj_pandya(hipStream_t s, const DevLayout &L, const double *ccb, double *Y);
// X <- (X + h X^T)/2 on the one-body part and every channel block; E0 <- 0 unless keep_e0
void launch_project(hipStream_t s, const DevLayout &L, double *X, int h, int keep_e0);
void launch_gemm_plain(hipStream_t s, const GemmProb *probs, const int4 *tiles, int ntiles, const double *A,
                       const double *B, double *C);
// transpose of the White (kind 0) / atan (kind 1) generator at H (mono = its
// monopoles): Hb += d<Eb, eta(H)>/dH; mbar: nsp*nsp scratch
void launch_adj_generator(hipStream_t s, const DevLayout &L, const double *H, const double *mono,
                          const double *Eb, double *Hb, double *mbar, int kind);
// transpose of normal ordering: Hb (a copy of the cotangent Cb on entry) becomes d<Cb, NO(H)>/dH
void launch_adj_normal_order(hipStream_t s, const DevLayout &L, const double *Cb, double *Hb, int herm);

// int8 slices (ozaki.hip): the pp/hh columns of every V row of `op`, and the
// rows of the Pandya matrices `cc`, into ns planes (5, or 4) + row scales;
// the planes keep the offsets of a kI8Planes-plane layout
void launch_slice_v(hipStream_t s, const DevLayout &L, const I8Layout &I, int ns, const double *op,
                    signed char *planes, double *scale, const int *act, int nact);
void launch_slice_cc(hipStream_t s, const DevLayout &L, const I8Layout &I, int ns, const double *cc, double *pmax,
                     signed char *planes, double *scale, const int *act, int nact);
// int8-sliced 222_pp_hh GEMM on the 64 x 64 tile lists of the FP64 path
// (C += upper, W after nmid).  Operand 0 / 1 slices P0 / P1 (ns planes), row
// scales S0 / S1; segment factor sign * (src ? g1 : g0).
void launch_seg_gemm_i8(hipStream_t s, int ns, const I8Prob *probs, const int4 *tiles, int ntiles,
                        const signed char *P0, const signed char *P1, const double *S0, const double *S1, double g0,
                        double g1, double *C, double *W, const int *act, int nact, long long sp, long long ss,
                        long long sc, long long sw);
// int8-sliced Z product of the ph channel (2x2-blocked upper Z) on the
// fragment-ordered Pandya planes (ns planes), over a list of 32 x 32 upper
// tiles (x = problem, y / z = row / column group; x = -1 a padding slot),
// four per workgroup.
void launch_zprod_i8(hipStream_t s, int ns, const I8Prob *probs, const int4 *tiles, int ntiles,
                     const signed char *P0, const signed char *P1, const double *S0, const double *S1, double g0,
                     double g1, double *Z, double sigma, const int *act, int nact, long long sp, long long ss,
                     long long sz);

constexpr int kPartials = 1024;  // reduction partials per instance
constexpr int kE0Split = 8;      // E0 (220) workgroups per channel block
constexpr int kScal = 64;        // scalar slots per instance
constexpr int kAdvSlot = 8;      // Magnus series: advals[0..6] in slots 8..14
constexpr int kFullSlot = 15;    // Magnus series: ||Omega||

}  // namespace imsrg
