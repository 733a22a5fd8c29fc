// ozaki.hip -- the 222 GEMMs on the int8 matrix cores of gfx950 (the Ozaki
// scheme: error-free slicing into int8 planes, exact int32 products).
//
// Every operand row r is held as x = s_r * sum_{i=1..NS} a_i 2^(-7i) (NS = 5
// slices, or 4): s_r is a power of two with max|row| / s_r in [1/4, 1/2), and
// a_i in [-64, 64] are produced by rounding x / s_r * 2^7 to nearest and
// carrying the remainder (each step exact in FP64; rounding rather than
// truncation keeps the dropped remainders unbiased, which matters for sums of
// same-sign terms).  A dot product of two rows is then
//   s_A s_B sum_{L=2..NS+1} 2^(-7L) sum_{i+j=L} sum_k a_i[k] b_j[k],
// with the inner sums exact in int32 on v_mfma_i32_32x32x32_i8 (|a_i b_j| <=
// 2^12, at most NS pairs per level and K < 2^17); the slice pairs with
// i + j > NS + 1 are dropped, so the error is about 2^-36 (NS = 5; 2^-29 for
// NS = 4) of max|A row| max|B row| per product, against 2^-24 of each K-slab
// partial sum for split FP32.
// The levels are combined and scaled in FP64, segment by segment, into the
// same FP64 accumulators and epilogues (222_pp_hh C += upper, the 221 W
// intermediate, the 2x2-blocked Z) as the FP64 path.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"
#include "devfun.hpp"

namespace imsrg {

typedef int v4i_t __attribute__((ext_vector_type(4)));
// NS (a template argument): slices per element, 5 or 4

// the NS rounded slices of x / 2^ex (|x| < 2^ex / 2), as bytes
template <int NS>
__device__ __forceinline__ void slices_of(double x, int ex, unsigned b[NS]) {
  double t = ldexp(x, 7 - ex);
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const double a = rint(t);
    b[i] = (unsigned)((int)a & 255);
    t = (t - a) * 128.0;
  }
}
// exponent ex with max / 2^ex in [1/4, 1/2); the row scale is 2^ex (0 for a zero row)
__device__ __forceinline__ int row_exp(double m) {
  int ex = 0;
  (void)frexp(m, &ex);
  return ex + 1;
}

// V rows: one wavefront per pair row (four per workgroup).  The pp columns
// [0, np) go to plane columns [0, np), the hh columns [np+nb, dim) to
// [P, P+nh) with P = 64*ceil(np/64); the padding columns are written as zero.
// Lane l owns plane columns 4l + 256t .. +3; rows up to 1024 plane columns
// keep their values in registers between the maximum and the slicing.
__device__ __forceinline__ int v_col(int j, int np, int nb, int nh, int P) {
  return j < np ? j : (j >= P && j < P + nh) ? np + nb + (j - P) : -1;
}
template <int NS>
__global__ __launch_bounds__(256) void k_slice_v(DevLayout L, I8Layout I, const double *__restrict__ op,
                                                 signed char *__restrict__ planes, double *__restrict__ scale,
                                                 const int *act) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= L.npairs) return;
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    op += inst * L.op_size;
    planes += inst * I.v_bytes;
    scale += inst * I.v_sr;
  }
  const int q = L.pair_q[r], ld = I.v_ld[q];
  if (ld == 0) return;
  const int rl = r - L.pbase[q], dim = L.bdim[q], np = L.bnp[q], nb = L.bnb[q], nh = L.bnh[q];
  const int P = (np + 63) & ~63;
  const double *X = op + L.v_off + L.boff[q] + (long long)rl * dim;
  constexpr int TC = 4;  // cached passes of 256 plane columns
  double v[TC][4];
  double m = 0.0;
#pragma unroll
  for (int t = 0; t < TC; t++)
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = 256 * t + 4 * lane + u;
      const int c = j < ld ? v_col(j, np, nb, nh, P) : -1;
      v[t][u] = X[max(c, 0)];  // off any branch: the sixteen loads are in flight together
      if (c < 0) v[t][u] = 0.0;
      m = fmax(m, fabs(v[t][u]));
    }
  for (int j = 256 * TC + lane; j < ld; j += 64) {
    const int c = v_col(j, np, nb, nh, P);
    if (c >= 0) m = fmax(m, fabs(X[c]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  const int ex = row_exp(m);
  if (lane == 0) scale[I.v_soff[q] + rl] = m > 0.0 ? ldexp(1.0, ex) : 0.0;
  signed char *row = planes + I.v_poff[q] + (long long)rl * NS * ld;
  auto put = [&](int j4, const double x[4]) {
    unsigned w[NS] = {};
#pragma unroll
    for (int u = 0; u < 4; u++) {
      unsigned b[NS] = {};
      if (m > 0.0) slices_of<NS>(x[u], ex, b);
#pragma unroll
      for (int i = 0; i < NS; i++) w[i] |= b[i] << (8 * u);
    }
#pragma unroll
    for (int i = 0; i < NS; i++) *(unsigned *)(row + (j4 >> 6) * (NS * 64) + i * 64 + (j4 & 63)) = w[i];
  };
#pragma unroll
  for (int t = 0; t < TC; t++) {
    const int j4 = 256 * t + 4 * lane;
    if (j4 < ld) put(j4, v[t]);
  }
  for (int j4 = 256 * TC + 4 * lane; j4 < ld; j4 += 256) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int c = v_col(j4 + u, np, nb, nh, P);
      x[u] = c >= 0 ? X[c] : 0.0;
    }
    put(j4, x);
  }
}

// Pandya planes are stored in fragment order: the rows of a block in groups
// of 32; for group g, K step s (32 columns, nks = ld / 32 per row) and slice
// i, 1 KB holds lane l's operand of one v_mfma_i32_32x32x32_i8, the 16 bytes
// of row 32 g + l % 32, columns 32 s + 16 (l / 32) .. +15, at 16 l.
template <int NS>
__device__ __forceinline__ long long frag_off(int g, int s, int i, int nks) {
  return ((long long)(g * nks + s) * NS + i) * 1024;
}

// Pandya rows (column-major r x K per block in `cc`), in items of 64 rows x
// 256 columns.  k_rowmax_cc: the maximum of each row over the item's columns
// (64 consecutive rows per wavefront load) into the strip's partial maxima;
// k_slice_cc: the row maximum from those partials, then each thread slices 16
// consecutive columns of one row per 64-column chunk and stores 16 bytes per
// plane (a wavefront's stores are two runs of 512 contiguous bytes).  Rows
// r..64*ceil(r/64) are written as zero.
__global__ __launch_bounds__(256) void k_rowmax_cc(DevLayout L, I8Layout I, const double *__restrict__ cc,
                                                   double *__restrict__ pmax, const int *act) {
  __shared__ double red[4][64];
  const int4 it = I.c_items[blockIdx.x];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    cc += inst * L.ccsize;
    pmax += inst * I.pm_size;
  }
  const int q = it.x, r0 = it.y, k0 = it.z, tid = threadIdx.x, rloc = tid & 63, kq = tid >> 6;
  const int R = L.ccrlen[q], K = L.ccnph[q], row = r0 + rloc, k1 = min(k0 + 256, K);
  const double *src = cc + L.ccoff[q];
  double m = 0.0;
  if (row < R) {
#pragma unroll 8
    for (int k = k0 + kq; k < k1; k += 4) m = fmax(m, fabs(src[(long long)k * R + row]));
  }
  red[kq][rloc] = m;
  __syncthreads();
  if (kq == 0)
    pmax[it.w + (k0 >> 8) * 64 + rloc] = fmax(fmax(red[0][rloc], red[1][rloc]), fmax(red[2][rloc], red[3][rloc]));
}
template <int NS>
__global__ __launch_bounds__(256) void k_slice_cc(DevLayout L, I8Layout I, const double *__restrict__ cc,
                                                  const double *__restrict__ pmax, signed char *__restrict__ planes,
                                                  double *__restrict__ scale, const int *act) {
  const int4 it = I.c_items[blockIdx.x];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    cc += inst * L.ccsize;
    pmax += inst * I.pm_size;
    planes += inst * I.c_bytes;
    scale += inst * I.c_sr;
  }
  const int q = it.x, r0 = it.y, k0 = it.z, tid = threadIdx.x, rloc = tid & 63, kq = tid >> 6;
  const int R = L.ccrlen[q], K = L.ccnph[q], ld = I.c_ld[q], row = r0 + rloc, nck = (ld + 255) >> 8;
  double m = 0.0;
  for (int c = 0; c < nck; c++) m = fmax(m, pmax[it.w + c * 64 + rloc]);
  const int ex = row_exp(m);
  if (k0 == 0 && kq == 0) scale[I.c_soff[q] + row] = m > 0.0 ? ldexp(1.0, ex) : 0.0;
  const double *src = cc + L.ccoff[q];
  signed char *dst = planes + I.c_poff[q];
  const int nks = ld >> 5;
  for (int kc = k0; kc < min(k0 + 256, ld); kc += 64) {
    const int kb = kc + kq * 16;
    double x[16];  // the sixteen loads first, off any branch, so they are in flight together
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const int k = kb + e;
      x[e] = src[(row < R && k < K) ? (long long)k * R + row : 0];
    }
    unsigned w[NS][4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
#pragma unroll
      for (int i = 0; i < NS; i++) w[i][g] = 0;
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int k = kb + 4 * g + u;
        unsigned b[NS] = {};
        if (row < R && k < K && m > 0.0) slices_of<NS>(x[4 * g + u], ex, b);
#pragma unroll
        for (int i = 0; i < NS; i++) w[i][g] |= b[i] << (8 * u);
      }
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      const v4i_t v = {(int)w[i][0], (int)w[i][1], (int)w[i][2], (int)w[i][3]};
      *(v4i_t *)(dst + frag_off<NS>(row >> 5, kb >> 5, i, nks) + (((row & 31) + 32 * ((kb >> 4) & 1)) << 4)) = v;
    }
  }
}

void launch_slice_v(hipStream_t s, const DevLayout &L, const I8Layout &I, int ns, const double *op,
                    signed char *planes, double *scale, const int *act, int nact) {
  if (nact <= 0 || L.npairs <= 0) return;
  hipLaunchKernelGGL(ns == 4 ? k_slice_v<4> : k_slice_v<5>, dim3((L.npairs + 3) / 4, nact), dim3(256), 0, s, L, I,
                     op, planes, scale, act);
}
void launch_slice_cc(hipStream_t s, const DevLayout &L, const I8Layout &I, int ns, const double *cc, double *pmax,
                     signed char *planes, double *scale, const int *act, int nact) {
  if (nact <= 0 || I.n_items <= 0) return;
  hipLaunchKernelGGL(k_rowmax_cc, dim3(I.n_items, nact), dim3(256), 0, s, L, I, cc, pmax, act);
  hipLaunchKernelGGL(ns == 4 ? k_slice_cc<4> : k_slice_cc<5>, dim3(I.n_items, nact), dim3(256), 0, s, L, I, cc, pmax,
                     planes, scale, act);
}

// ---- the 222_pp_hh product (C += upper, the 221 intermediate W after
// segment nmid) on the V-row planes.  A 64 x 64 output tile per workgroup; each K chunk of 64
// columns of the NS slices of A (64 rows) and B (64 rows, 320 contiguous bytes
// per row) is loaded into registers one chunk ahead and staged in LDS (16-byte column group g of row
// r at g ^ ((r >> 2) & 3), so the 16 rows a quarter-wavefront reads hit
// distinct banks).  Wavefront (wm, wn) owns the 32 x 32 quadrant (rows 32 wm,
// columns 32 wn) as one v_mfma_i32_32x32x32_i8 tile per level (the 16x16x64
// form issues at half its rate).  Fragments: lane l holds row l % 32 and the
// K bytes 16 (l / 32) .. +15 of a 32-byte K step, for A and B alike; result
// element v of lane l is (8 (v / 4) + 4 (l / 32) + v % 4, l % 32).
template <int NS>
struct I8Smem {
  v4i_t a[NS][64][4];
  v4i_t b[NS][64][4];
};
typedef int v16i_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int swz(int r, int g) { return g ^ ((r >> 2) & 3); }

template <int NS>
__global__ __launch_bounds__(256, 2) void k_seg_gemm_i8(const I8Prob *__restrict__ probs,
                                                        const int4 *__restrict__ tiles, int ntiles,
                                                        const signed char *__restrict__ P0,
                                                        const signed char *__restrict__ P1,
                                                        const double *__restrict__ S0, const double *__restrict__ S1,
                                                        double g0, double g1, double *__restrict__ C,
                                                        double *__restrict__ W, const int *act, long long sp,
                                                        long long ss, long long sc, long long sw) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    P0 += inst * sp;
    P1 += inst * sp;
    S0 += inst * ss;
    S1 += inst * ss;
    if (C) C += inst * sc;
    if (W) W += inst * sw;
  }
  const int4 tl = tiles[t];
  if (tl.x < 0) return;  // padding slot of the XCD-balanced schedule
  const I8Prob &p = probs[tl.x];
  __shared__ I8Smem<NS> S2[2];  // double-buffered chunks: 80 KB, two workgroups fill a CU's 160 KB
  const int M = p.M, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = tl.y * 64, n0 = tl.z * 64, wm = wave >> 1, wn = wave & 1;
  const int ra = m0 + wm * 32, cb = n0 + wn * 32;
  // quadrants beyond the block or strictly below the diagonal only help load
  const bool active = ra < M && cb < M && ra < cb + 32;
  const long long rs = (long long)NS * p.ld;  // row stride
  // staging: the chunk of 64 rows x 320 bytes of one operand is 1280 pieces of
  // 16 bytes; thread tid takes pieces tid + 256 q (consecutive threads read
  // consecutive bytes), piece w of a row being slice w / 4, column group w % 4
  int soff[NS], sdst[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) {
    const int pc = tid + 256 * q, r = pc / (4 * NS), w = pc - r * (4 * NS);
    soff[q] = r * (int)rs + w * 16;
    sdst[q] = ((w >> 2) * 64 + r) * 4 + swz(r, w & 3);  // index into a [NS][64][4] array of 16-byte groups
  }
  // fragment reads: row lane % 32 of the quadrant, column group 2 h + lane / 32 of K step h
  const int fr = lane & 31, fh = lane >> 5;
  const int fra = wm * 32 + fr, frb = wn * 32 + fr;
  double acc[16];
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.0;

  auto run = [&](int s0, int s1) {
    for (int s = s0; s < s1; s++) {
      const I8Seg &sgm = p.seg[s];
      const int K = sgm.K;
      if (K == 0) continue;
      const bool o = sgm.src != 0;
      const signed char *A = (o ? P1 : P0) + p.p_off + (long long)m0 * rs + (sgm.kofs >> 6) * (NS * 64);
      const signed char *B = (o ? P0 : P1) + p.p_off + (long long)n0 * rs + (sgm.kofs >> 6) * (NS * 64);
      v4i_t R[2][NS];  // the next chunk, in flight while this one is used
      auto load = [&](int kc) {
        const int co = (kc >> 6) * (NS * 64);
#pragma unroll
        for (int q = 0; q < NS; q++) {
          R[0][q] = *(const v4i_t *)(A + co + soff[q]);
          R[1][q] = *(const v4i_t *)(B + co + soff[q]);
        }
      };
      v16i_t lv[NS];
#pragma unroll
      for (int l = 0; l < NS; l++)
#pragma unroll
        for (int v = 0; v < 16; v++) lv[l][v] = 0;
      auto store = [&](I8Smem<NS> &S) {
#pragma unroll
        for (int q = 0; q < NS; q++) {
          (&S.a[0][0][0])[sdst[q]] = R[0][q];
          (&S.b[0][0][0])[sdst[q]] = R[1][q];
        }
      };
      // chunk kc is read from buffer (kc / 64) % 2 while chunk kc + 64 is
      // stored into the other one; one barrier per chunk.  Every segment ends
      // on a barrier, so both buffers are free when the next one starts.
      load(0);
      store(S2[0]);
      if (64 < K) load(64);
      __syncthreads();
      int buf = 0;
      for (int kc = 0; kc < K; kc += 64) {  // K is workgroup-uniform: every barrier is reached by all
        const I8Smem<NS> &S = S2[buf];
        if (active) {
          const int nh = kc + 32 < sgm.K32 ? 2 : 1;  // workgroup-uniform
#pragma unroll
          for (int h = 0; h < 2; h++) {
            if (h >= nh) break;
            // all B fragments of this K step, then A plane by plane: pair (i, j) adds to level i + j
            v4i_t b[NS];
#pragma unroll
            for (int j = 0; j < NS; j++) b[j] = S.b[j][frb][swz(frb, 2 * h + fh)];
#pragma unroll
            for (int i = 0; i < NS; i++) {
              const v4i_t a = S.a[i][fra][swz(fra, 2 * h + fh)];
#pragma unroll
              for (int j = 0; j + i < NS; j++)
                lv[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b[j], lv[i + j], 0, 0, 0);
            }
          }
        }
        if (kc + 64 < K) {
          store(S2[buf ^ 1]);
          if (kc + 128 < K) load(kc + 128);
        }
        __syncthreads();
        buf ^= 1;
      }
      if (active) {
        const double *SA = (o ? S1 : S0) + p.s_off, *SB = (o ? S0 : S1) + p.s_off;
        const double sb = SB[cb + fr] * (sgm.sign * (o ? g1 : g0));
#pragma unroll
        for (int v = 0; v < 16; v++) {
          const double sa = SA[ra + 8 * (v >> 2) + 4 * fh + (v & 3)];
          double lev = (double)lv[NS - 1][v];
#pragma unroll
          for (int l = NS - 2; l >= 0; l--) lev = (double)lv[l][v] + lev * 0x1p-7;
          acc[v] += sa * sb * (lev * 0x1p-14);
        }
      }
    }
  };
  auto each = [&](auto &&f) {
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int row = ra + 8 * (v >> 2) + 4 * fh + (v & 3), col = cb + fr;
      if (row < M && col < M && row <= col) f(row, col, acc[v]);
    }
  };
  run(0, p.nmid);
  if (W && active) {
    double *Wp = W + p.w_off;
    each([&](int row, int col, double x) { Wp[(long long)row * M + col] = x; });
  }
  run(p.nmid, p.nseg);
  if (C && active) {
    // C's old values are all loaded before the first store (element 0 where
    // the tile leaves the upper triangle), so the sixteen loads are in flight together
    double *Cp = C + p.c_off;
    double old[16];
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int row = ra + 8 * (v >> 2) + 4 * fh + (v & 3), col = cb + fr;
      old[v] = Cp[(row < M && col < M && row <= col) ? (long long)row * M + col : 0];
    }
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int row = ra + 8 * (v >> 2) + 4 * fh + (v & 3), col = cb + fr;
      if (row < M && col < M && row <= col) Cp[(long long)row * M + col] = old[v] + acc[v];
    }
  }
}

void launch_seg_gemm_i8(hipStream_t s, int ns, const I8Prob *probs, const int4 *tiles, int ntiles,
                        const signed char *P0, const signed char *P1, const double *S0, const double *S1, double g0,
                        double g1, double *C, double *W, const int *act, int nact, long long sp, long long ss,
                        long long sc, long long sw) {
  if (ntiles <= 0 || nact <= 0) return;
  launch_k(ns == 4 ? k_seg_gemm_i8<4> : k_seg_gemm_i8<5>, dim3(ntiles, nact), dim3(256), 0, s, probs, tiles, ntiles,
           P0, P1, S0, S1, g0, g1, C, W, act, sp, ss, sc, sw);
}

// ---- the Z product of the ph channel over each Pandya block,
//   Z(m, n) = g0 L_m . R_n + g1 R_m . L_n     (rows of the fragment-ordered planes),
// whose K (the block's ph pairs, at most 116 at emax = 8) is short: the
// product is bound by the latency of its operand loads and by its Z stores,
// not by the matrix cores.  One wavefront per 32 x 32 upper tile, four tiles
// (consecutive column groups of one row group, so they share A through the
// L1 and L2) per workgroup; no barriers.  Every fragment is one coalesced
// 1 KB load; the ten fragments of K step t + 1 are in flight while the
// fifteen v_mfma_i32_32x32x32_i8 of step t run.  The two segments are one
// stream of 2 nk steps; each segment's level sums are scaled into the FP64
// accumulators when its last step is done, with row scales the wavefront
// staged in its own LDS slot at the start (so that no global load of the
// scaling waits behind the operand loads in flight).  Result element v of
// lane l is row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32 of the tile;
// lanes c, c ^ 1 swap one value per row pair so that each writes one 16-byte
// half of a 2x2 Z block (the even lane Z(r,c), Z(r+1,c+1), the odd lane
// Z(r,c+1), Z(r+1,c); see zidx), and the 32 lanes of a row pair write 512
// contiguous bytes.
template <int NS>
__global__ __launch_bounds__(256, 2) void k_zprod_i8(const I8Prob *__restrict__ probs, const int4 *__restrict__ tiles,
                                                  int ntiles, const signed char *__restrict__ P0,
                                                  const signed char *__restrict__ P1, const double *__restrict__ S0,
                                                  const double *__restrict__ S1, double g0, double g1,
                                                  double *__restrict__ Z, double sigma, const int *act, long long sp,
                                                  long long ss, long long sz) {
  __shared__ double scl[4][128];  // per wavefront: A row scales of segments 0, 1; B row scales of segments 0, 1
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int item = (int)blockIdx.x * 4 + wave;
  if (item >= ntiles) return;  // no barriers in this kernel: a wavefront may leave on its own
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    P0 += inst * sp;
    P1 += inst * sp;
    S0 += inst * ss;
    S1 += inst * ss;
    Z += inst * sz;
  }
  const int4 tl = tiles[item];
  if (tl.x < 0) return;  // padding slot of the XCD-balanced schedule
  const I8Prob &p = probs[tl.x];
  const int M = p.M, lane = threadIdx.x & 63, fr = lane & 31, fh = lane >> 5;
  const int ra = tl.y * 32, cb = tl.z * 32;
  const int nks = p.ld >> 5, nk = p.seg[0].K32 >> 5, T = 2 * nk;
  double *sw = scl[wave];
  const double f0 = p.seg[0].sign * g0, f1 = p.seg[1].sign * g1;  // segment factors (uniform)
  // the row scales first (the oldest loads, so that waiting for them leaves
  // the operand loads issued after them in flight)
  double sx, sy;
  {
    const double *sp0 = S0 + p.s_off, *sp1 = S1 + p.s_off;
    sx = fh ? sp1[ra + fr] : sp0[ra + fr];  // A rows: L (segment 0), R (segment 1)
    sy = fh ? sp0[cb + fr] : sp1[cb + fr];  // B rows: R (segment 0), L (segment 1)
  }
  const unsigned lo = 16u * (unsigned)lane;
  const signed char *A0 = P0 + p.p_off + frag_off<NS>(tl.y, 0, 0, nks), *B0 = P1 + p.p_off + frag_off<NS>(tl.z, 0, 0, nks);
  const signed char *A1 = P1 + p.p_off + frag_off<NS>(tl.y, 0, 0, nks), *B1 = P0 + p.p_off + frag_off<NS>(tl.z, 0, 0, nks);
  v16i_t lv[NS];
  double acc[16];
  v4i_t fa[2][NS], fb[2][NS];  // two fragment sets (three spill at two wavefronts per SIMD)
  // step t (clamped to T - 1: the reloads past the last step are never used)
  auto load = [&](int t, v4i_t(&a)[NS], v4i_t(&b)[NS]) {
    t = min(t, T - 1);
    const bool s1 = t >= nk;
    const long long o = (long long)(s1 ? t - nk : t) * (NS * 1024);
    const signed char *pa = (s1 ? A1 : A0) + o, *pb = (s1 ? B1 : B0) + o;
#pragma unroll
    for (int i = 0; i < NS; i++) {
      a[i] = *(const v4i_t *)(pa + i * 1024 + lo);
      b[i] = *(const v4i_t *)(pb + i * 1024 + lo);
    }
    // the sets stay whole and in program order, so a wait for the oldest
    // set leaves the younger ones in flight
    __builtin_amdgcn_sched_barrier(0);
  };
  auto mma = [&](const v4i_t(&a)[NS], const v4i_t(&b)[NS]) {
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int j = 0; i + j < NS; j++) lv[i + j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[i], b[j], lv[i + j], 0, 0, 0);
  };
  auto fold = [&](int sg) {
    const double sb = sw[64 + 32 * sg + fr];
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const double sa = sw[32 * sg + 8 * (v >> 2) + 4 * fh + (v & 3)];
      double lev = (double)lv[NS - 1][v];
#pragma unroll
      for (int l = NS - 2; l >= 0; l--) lev = (double)lv[l][v] + lev * 0x1p-7;
      acc[v] += sa * sb * (lev * 0x1p-14);
    }
#pragma unroll
    for (int l = 0; l < NS; l++)
#pragma unroll
      for (int v = 0; v < 16; v++) lv[l][v] = 0;
  };
#pragma unroll
  for (int l = 0; l < NS; l++)
#pragma unroll
    for (int v = 0; v < 16; v++) lv[l][v] = 0;
#pragma unroll
  for (int v = 0; v < 16; v++) acc[v] = 0.0;
  if (T > 0) load(0, fa[0], fb[0]);  // T = 2 nk >= 2
  sw[lane] = sx;
  sw[64 + lane] = sy * (fh ? f1 : f0);
  if (T > 0) {
    // the two fragment sets in rotation, one step ahead (T is even)
    for (int t = 0; t < T; t += 2) {
      load(t + 1, fa[1], fb[1]);
      mma(fa[0], fb[0]);
      if (t == nk - 1) fold(0);
      load(t + 2, fa[0], fb[0]);
      mma(fa[1], fb[1]);
      if (t + 1 == nk - 1) fold(0);
    }
    fold(1);
  }
  double *Zp = Z + p.c_off;
  const int col = cb + fr;
  const bool odd = fr & 1;
#pragma unroll
  for (int v = 0; v < 16; v += 2) {
    const int row = ra + 8 * (v >> 2) + 4 * fh + (v & 3);  // even: rows row, row + 1 are v, v + 1
    const double y = __shfl_xor(acc[v + 1], 1, 64);  // the partner's row + 1 element
    if (row < M && col < M && (row >> 1) <= (col >> 1)) {
      const long long e = zidx(row, col & ~1, M);
      if (odd)  // Z(row, col), Z(row+1, col-1); a diagonal block takes sigma Z(row, col) for the second
        *(double2 *)(Zp + e + 2) = make_double2(acc[v], (row >> 1) == (col >> 1) ? sigma * acc[v] : y);
      else  // Z(row, col), Z(row+1, col+1)
        *(double2 *)(Zp + e) = make_double2(acc[v], y);
    }
  }
}

void launch_zprod_i8(hipStream_t s, int ns, const I8Prob *probs, const int4 *tiles, int ntiles,
                     const signed char *P0, const signed char *P1, const double *S0, const double *S1, double g0,
                     double g1, double *Z, double sigma, const int *act, int nact, long long sp, long long ss,
                     long long sz) {
  if (ntiles <= 0 || nact <= 0) return;
  launch_k(ns == 4 ? k_zprod_i8<4> : k_zprod_i8<5>, dim3((ntiles + 3) / 4, nact), dim3(256), 0, s, probs, tiles,
           ntiles, P0, P1, S0, S1, g0, g1, Z, sigma, act, sp, ss, sz);
}

}  // namespace imsrg
