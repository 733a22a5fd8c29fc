// kernels.hip -- CDNA4 (gfx950) kernels of the IMSRG(2) commutator engine.
//
// Operators are device buffers [E0 | F (nsp x nsp) | V (full channel blocks)]
// described in layout.hpp.  The two-body terms of a commutator accumulate into
// the upper triangle of each channel block (the 122 kernels also leave the source of
// its transposed half in the lower triangle); the last pass, k_sym_mirror,
// adds that half and writes lower = herm * upper, which reproduces the
// reference's "upper-triangle then mirror" contract for square sub-blocks and
// its implied herm * transpose for the non-square ones.  Launches take a table
// of active instances (batched contexts) indexed with blockIdx.y.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"
#include "devfun.hpp"

namespace imsrg {

// ------------------------------------------------------------------ elementwise
__global__ void k_fill(double *p, long long n, double v, const int *act, long long stride) {
  const long long inst = inst_id(act);
  if (inst < 0) return;
  p += inst * stride;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}
void launch_fill(hipStream_t s, double *p, long long n, double v, const int *act, int nact, long long stride) {
  if (n <= 0 || nact <= 0) return;
  long long nb = (n + 255) / 256;
  int grid = (int)(nb < 8192 ? nb : 8192);
  hipLaunchKernelGGL(k_fill, dim3(grid, nact), dim3(256), 0, s, p, n, v, act, stride);
}

// add_sq_op (basic_IMSRG.f90:3450-3478): C = A*ax + B*bx elementwise
__global__ void k_axpby(long long n, double a, const double *A, double b, const double *B, double *C,
                        const int *act, long long stride) {
  const long long inst = inst_id(act);
  if (inst < 0) return;
  const long long o = inst * stride;
  A += o;
  B += o;
  C += o;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) C[i] = A[i] * a + B[i] * b;
}
void launch_axpby(hipStream_t s, long long n, double a, const double *A, double b, const double *B,
                  double *C, const int *act, int nact, long long stride) {
  if (n <= 0 || nact <= 0) return;
  long long nb = (n + 255) / 256;
  int grid = (int)(nb < 8192 ? nb : 8192);
  hipLaunchKernelGGL(k_axpby, dim3(grid, nact), dim3(256), 0, s, n, a, A, b, B, C, act, stride);
}

// copy_sq_op (basic_IMSRG.f90:2942-2969) on the active instances
__global__ void k_copy(long long n, const double *A, double *C, const int *act, long long stride) {
  const long long inst = inst_id(act);
  if (inst < 0) return;
  const long long o = inst * stride;
  A += o;
  C += o;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long st = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += st) C[i] = A[i];
}
void launch_copy(hipStream_t s, long long n, const double *A, double *C, const int *act, int nact,
                 long long stride) {
  if (n <= 0 || nact <= 0) return;
  long long nb = (n + 255) / 256;
  int grid = (int)(nb < 8192 ? nb : 8192);
  hipLaunchKernelGGL(k_copy, dim3(grid, nact), dim3(256), 0, s, n, A, C, act, stride);
}

// ------------------------------------------------------------------ segmented grouped GEMM
// Every output block C (M x M) is a signed sum of K segments,
//   C = sum_s sign_s * A_s * B_s,
// where segment s takes A from X[src] and B from X[1-src] (X0, X1 = the two
// operands of the commutator, or their Pandya matrices).  A_s is M x K_s,
// row-major (element (m,k) at m*lda + k) or, with A_KMAJOR, K-major (k*lda + m);
// B_s is K_s x M with element (k,n) at k*ldb + n.  The outputs are
// herm-symmetric, so only tiles tm <= tn are launched: one 64x64 tile per
// workgroup, 4 wavefronts with 32x32 quadrants of 2x2 v_mfma_f64_16x16x4
// tiles; in a diagonal tile the strictly-lower quadrant does no MFMA work.
// K is staged through double-buffered LDS slabs of 16 (one barrier per slab,
// the next slab's global loads in flight during the current slab's MFMAs).
#ifndef IMSRG_GEMM_BK
#define IMSRG_GEMM_BK 16
#endif
#ifndef IMSRG_GEMM_PF
#define IMSRG_GEMM_PF 2  // K slabs of global loads in flight ahead of the MFMAs (1 or 2)
#endif
constexpr int GT = 64, GK = IMSRG_GEMM_BK, GP = 1, GJ = GK / 4;
static_assert(GK == 16 || GK == 32, "K slab of 16 or 32");

struct GemmSmem {
  double A[2][GK][GT + GP];
  double B[2][GK][GT + GP];
};
static_assert(sizeof(GemmSmem) >= sizeof(double) * GT * (GT + GP), "transpose buffer must fit");

// `need` bit (2*ti + tj): the 16x16 sub-tile holds outputs that are used
// (inside the block and not strictly below the diagonal); others do no MFMA.
// `need` is wave-uniform (an SGPR).  Off-diagonal tiles inside their block
// run the FULL instantiation (straight-line MFMAs); the kernel picks it once
// per workgroup, so neither path carries the other's registers through the
// K loop.
template <bool FULL>
__device__ __forceinline__ void slab_mfma(const GemmSmem &S, int buf, double4_t acc[2][2], unsigned need) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
  for (int kq = 0; kq < GK / 4; kq++) {
    const int k = kq * 4 + (lane >> 4);
    const double a0 = S.A[buf][k][wm * 32 + (lane & 15)];
    const double a1 = S.A[buf][k][wm * 32 + 16 + (lane & 15)];
    const double b0 = S.B[buf][k][wn * 32 + (lane & 15)];
    const double b1 = S.B[buf][k][wn * 32 + 16 + (lane & 15)];
    if (FULL || (need & 1)) acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
    if (FULL || (need & 2)) acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
    if (FULL || (need & 4)) acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
    if (FULL || (need & 8)) acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
  }
}

// acc += sgn * A * B over one segment, FP64 MFMA.  Each thread owns 4 A and
// 4 B elements of every slab at fixed (row, k-in-slab) positions, kept as
// 32-bit offsets from the segment base that advance by one slab per step.
// Loads are issued unconditionally (an out-of-range element reads offset 0,
// the segment's first element) and the in-range flags are applied, with the
// sign, only when the slab is written to LDS after the current slab's MFMAs,
// so the next slab's loads stay in flight meanwhile.
template <bool A_KMAJOR, bool FULL>
__device__ void seg_f64(const SegDesc &sgd, double sgn, int M, const double *__restrict__ A,
                        const double *__restrict__ B, int m0, int n0, GemmSmem &S, double4_t acc[2][2],
                        unsigned need) {
  const int K = sgd.K, lda = sgd.lda, ldb = sgd.ldb;
  const int nk = (K + GK - 1) / GK;
  if (nk == 0) return;
  const int tid = threadIdx.x;
  int aoff[GJ], boff[GJ], akk[GJ], bkk[GJ];
  unsigned rowok = 0;  // bits 0..GJ-1: A rows in range, bits 16..: B columns in range
  if (A_KMAJOR) {
    const int m = m0 + (tid & 63);
    const bool mv = m < M;
#pragma unroll
    for (int j = 0; j < GJ; j++) {
      akk[j] = (tid >> 6) + 4 * j;
      aoff[j] = akk[j] * lda + (mv ? m : 0);
      rowok |= (unsigned)mv << j;
    }
  } else {
#pragma unroll
    for (int j = 0; j < GJ; j++) {
      const int m = m0 + tid / GK + (256 / GK) * j;
      const bool mv = m < M;
      akk[j] = tid % GK;
      aoff[j] = (mv ? m : 0) * lda + akk[j];
      rowok |= (unsigned)mv << j;
    }
  }
  {
    const int n = n0 + (tid & 63);
    const bool nv = n < M;
#pragma unroll
    for (int j = 0; j < GJ; j++) {
      bkk[j] = (tid >> 6) + 4 * j;
      boff[j] = bkk[j] * ldb + (nv ? n : 0);
      rowok |= (unsigned)nv << (16 + j);
    }
  }
  const int ainc = A_KMAJOR ? GK * lda : GK, binc = GK * ldb;
  // two register sets: the loads of slab kt+2 are issued while slab kt is
  // multiplied, so each load has two slabs of MFMA work to arrive in
  double ra0[GJ], rb0[GJ], ra1[GJ], rb1[GJ];
  unsigned ok0 = 0, ok1 = 0;
  // FULL tiles have every row and column in range, so a slab that lies
  // wholly inside K (all but a partial last one) is loaded and stored
  // without checks; the test is on scalars, one branch per slab.
  auto load = [&](int k0, double (&ra)[GJ], double (&rb)[GJ], unsigned &ok) {
    if (FULL && k0 + GK <= K) {
#pragma unroll
      for (int j = 0; j < GJ; j++) {
        ra[j] = A[aoff[j]];
        rb[j] = B[boff[j]];
        aoff[j] += ainc;
        boff[j] += binc;
      }
      return;
    }
    ok = 0;
#pragma unroll
    for (int j = 0; j < GJ; j++) {
      const bool ga = ((rowok >> j) & 1) && (k0 + akk[j] < K);
      const bool gb = ((rowok >> (16 + j)) & 1) && (k0 + bkk[j] < K);
      ra[j] = A[ga ? aoff[j] : 0];
      rb[j] = B[gb ? boff[j] : 0];
      ok |= ((unsigned)ga << j) | ((unsigned)gb << (16 + j));
      aoff[j] += ainc;
      boff[j] += binc;
    }
  };
  auto store = [&](int buf, const double (&ra)[GJ], const double (&rb)[GJ], unsigned ok, int k0) {
    if (FULL && k0 + GK <= K) {
#pragma unroll
      for (int j = 0; j < GJ; j++) {
        if (A_KMAJOR) S.A[buf][(tid >> 6) + 4 * j][tid & 63] = ra[j];
        else S.A[buf][tid % GK][tid / GK + (256 / GK) * j] = ra[j];
        S.B[buf][(tid >> 6) + 4 * j][tid & 63] = sgn * rb[j];
      }
      return;
    }
#pragma unroll
    for (int j = 0; j < GJ; j++) {
      const double a = ((ok >> j) & 1) ? ra[j] : 0.0;
      const double b = ((ok >> (16 + j)) & 1) ? sgn * rb[j] : 0.0;
      if (A_KMAJOR) S.A[buf][(tid >> 6) + 4 * j][tid & 63] = a;
      else S.A[buf][tid % GK][tid / GK + (256 / GK) * j] = a;
      S.B[buf][(tid >> 6) + 4 * j][tid & 63] = b;
    }
  };
#if IMSRG_GEMM_PF == 2
  load(0, ra0, rb0, ok0);
  store(0, ra0, rb0, ok0, 0);
  if (nk > 1) load(GK, ra1, rb1, ok1);
  __syncthreads();
  for (int kt = 0; kt < nk; kt += 2) {
    // slab kt in LDS 0, slab kt+1 in registers 1, registers 0 free
    if (kt + 2 < nk) load((kt + 2) * GK, ra0, rb0, ok0);
    if (need) slab_mfma<FULL>(S, 0, acc, need);
    if (kt + 1 < nk) store(1, ra1, rb1, ok1, (kt + 1) * GK);
    __syncthreads();
    if (kt + 1 >= nk) break;
    // slab kt+1 in LDS 1, slab kt+2 in registers 0, registers 1 free
    if (kt + 3 < nk) load((kt + 3) * GK, ra1, rb1, ok1);
    if (need) slab_mfma<FULL>(S, 1, acc, need);
    if (kt + 2 < nk) store(0, ra0, rb0, ok0, (kt + 2) * GK);
    __syncthreads();
  }
#else
  (void)ra1;
  (void)rb1;
  (void)ok1;
  load(0, ra0, rb0, ok0);
  store(0, ra0, rb0, ok0, 0);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) load((kt + 1) * GK, ra0, rb0, ok0);
    if (need) slab_mfma<FULL>(S, cur, acc, need);
    if (more) store(cur ^ 1, ra0, rb0, ok0, (kt + 1) * GK);
    __syncthreads();
  }
#endif
}

// local (row, col) of accumulator element (ti, tj, v) of this lane.
// D[i][j] has j = lane%16 and, for v_mfma_f64_16x16x4, i = lane/16 + 4v; for
// v_mfma_f32_16x16x4f32 and v_mfma_f32_16x16x32_bf16, i = 4*(lane/16) + v.
template <int MODE>
__device__ __forceinline__ int acc_row(int ti, int v) {
  const int g = (threadIdx.x & 63) >> 4;
  return ((threadIdx.x >> 6) >> 1) * 32 + ti * 16 + (MODE == 0 ? g + 4 * v : 4 * g + v);
}
__device__ __forceinline__ int acc_col(int tj) {
  return ((threadIdx.x >> 6) & 1) * 32 + tj * 16 + (threadIdx.x & 15);
}


// ---- split-precision segments: the same 64x64 tile and 2x2 quadrant layout;
// products accumulate in FP32 inside one K slab and in FP64 across slabs.
typedef float float4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <bool A_KMAJOR>
__device__ __forceinline__ double seg_a(const SegDesc &sg, int M, const double *A, int m, int k) {
  if (m >= M || k >= sg.K) return 0.0;
  return A_KMAJOR ? A[(long long)k * sg.lda + m] : A[(long long)m * sg.lda + k];
}
__device__ __forceinline__ double seg_b(const SegDesc &sg, int M, const double *B, int k, int n) {
  return (n < M && k < sg.K) ? B[(long long)k * sg.ldb + n] : 0.0;
}

// split-FP32: x = hi + lo (FP32 each); hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x4f32
template <bool A_KMAJOR>
__device__ void seg_s32(const SegDesc &sg, double sgn, int M, const double *__restrict__ A,
                        const double *__restrict__ B, int m0, int n0, double4_t acc[2][2], unsigned need) {
  const bool active = need != 0;
  __shared__ float Ah[16][65], Al[16][65], Bh[16][65], Bl[16][65];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  for (int k0 = 0; k0 < sg.K; k0 += 16) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int idx = tid * 4 + e;  // 0..1023 over a 16 x 64 slab
      const int kk = idx >> 6, mm = idx & 63;
      if (A_KMAJOR) {
        const double a = seg_a<true>(sg, M, A, m0 + mm, k0 + kk);
        const float ah = (float)a;
        Ah[kk][mm] = ah;
        Al[kk][mm] = (float)(a - (double)ah);
      }
      const double b = sgn * seg_b(sg, M, B, k0 + kk, n0 + mm);
      const float bh = (float)b;
      Bh[kk][mm] = bh;
      Bl[kk][mm] = (float)(b - (double)bh);
    }
    if (!A_KMAJOR) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int idx = tid * 4 + e;  // rows m, k contiguous in memory
        const int mm = idx >> 4, kk = idx & 15;
        const double a = seg_a<false>(sg, M, A, m0 + mm, k0 + kk);
        const float ah = (float)a;
        Ah[kk][mm] = ah;
        Al[kk][mm] = (float)(a - (double)ah);
      }
    }
    __syncthreads();
    if (active) {
      float4_t c[2][2];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) c[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kq = 0; kq < 4; kq++) {
        const int k = kq * 4 + (lane >> 4);
        float ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
          ah[t] = Ah[k][wm * 32 + t * 16 + (lane & 15)];
          al[t] = Al[k][wm * 32 + t * 16 + (lane & 15)];
          bh[t] = Bh[k][wn * 32 + t * 16 + (lane & 15)];
          bl[t] = Bl[k][wn * 32 + t * 16 + (lane & 15)];
        }
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
          for (int tj = 0; tj < 2; tj++) {
            c[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[ti], bh[tj], c[ti][tj], 0, 0, 0);
            c[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[ti], bl[tj], c[ti][tj], 0, 0, 0);
            c[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[ti], bh[tj], c[ti][tj], 0, 0, 0);
          }
      }
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) acc[ti][tj][v] += (double)c[ti][tj][v];
    }
    __syncthreads();
  }
}

// split-BF16: x = hi + mid + lo (BF16 each); the six products down to 2^-24
// relative (hh, hm, mh, hl, lh, mm) on v_mfma_f32_16x16x32_bf16, K slab 32
template <bool A_KMAJOR>
__device__ void seg_bf16(const SegDesc &sg, double sgn, int M, const double *__restrict__ A,
                         const double *__restrict__ B, int m0, int n0, double4_t acc[2][2], unsigned need) {
  const bool active = need != 0;
  // [slice][row][k]: a lane reads 8 consecutive k of one row as one 16-byte vector
  __shared__ __attribute__((aligned(16))) __bf16 As[3][64][40];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][64][40];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  for (int k0 = 0; k0 < sg.K; k0 += 32) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int idx = tid * 8 + e;  // 0..2047 over a 64 x 32 slab
      int mm, kk;
      if (A_KMAJOR) { kk = idx >> 6; mm = idx & 63; } else { mm = idx >> 5; kk = idx & 31; }
      const double a = seg_a<A_KMAJOR>(sg, M, A, m0 + mm, k0 + kk);
      const __bf16 h = (__bf16)a;
      const double r1 = a - (double)(float)h;
      const __bf16 md = (__bf16)r1;
      const __bf16 lo = (__bf16)(r1 - (double)(float)md);
      As[0][mm][kk] = h;
      As[1][mm][kk] = md;
      As[2][mm][kk] = lo;
      const int kb = idx >> 6, nb = idx & 63;
      const double b = sgn * seg_b(sg, M, B, k0 + kb, n0 + nb);
      const __bf16 bh = (__bf16)b;
      const double s1 = b - (double)(float)bh;
      const __bf16 bm = (__bf16)s1;
      const __bf16 bo = (__bf16)(s1 - (double)(float)bm);
      Bs[0][nb][kb] = bh;
      Bs[1][nb][kb] = bm;
      Bs[2][nb][kb] = bo;
    }
    __syncthreads();
    if (active) {
      const int kofs = 8 * (lane >> 4);
      bf16x8_t a[3][2], b[3][2];
#pragma unroll
      for (int sl = 0; sl < 3; sl++)
#pragma unroll
        for (int t = 0; t < 2; t++) {
          a[sl][t] = *(const bf16x8_t *)&As[sl][wm * 32 + t * 16 + (lane & 15)][kofs];
          b[sl][t] = *(const bf16x8_t *)&Bs[sl][wn * 32 + t * 16 + (lane & 15)][kofs];
        }
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++) {
          float4_t c = {0.f, 0.f, 0.f, 0.f};
          // smallest terms first so they are not lost against the leading one
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][ti], b[1][tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][ti], b[0][tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ti], b[2][tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][ti], b[0][tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ti], b[1][tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][ti], b[0][tj], c, 0, 0, 0);
#pragma unroll
          for (int v = 0; v < 4; v++) acc[ti][tj][v] += (double)c[v];
        }
    }
    __syncthreads();
  }
}

template <bool A_KMAJOR, int MODE, bool FULL>
__device__ __forceinline__ void run_segments(const SegProb &p, int s0, int s1, const double *X0,
                                             const double *X1, double g0, double g1, int m0, int n0,
                                             GemmSmem &S, double4_t acc[2][2], unsigned need) {
  for (int s = s0; s < s1; s++) {
    const SegDesc &sg = p.seg[s];
    const double *A = (sg.src ? X1 : X0) + sg.a_off;
    const double *B = (sg.src ? X0 : X1) + sg.b_off;
    const double sgn = sg.sign * (sg.src ? g1 : g0);
    if constexpr (MODE == 0) seg_f64<A_KMAJOR, FULL>(sg, sgn, p.M, A, B, m0, n0, S, acc, need);
    else if constexpr (MODE == 1) seg_s32<A_KMAJOR>(sg, sgn, p.M, A, B, m0, n0, acc, need);
    else seg_bf16<A_KMAJOR>(sg, sgn, p.M, A, B, m0, n0, acc, need);
  }
}

// EPI 0 (222_pp_hh): C += acc on the upper triangle; with W, the accumulator
//   after the first nmid segments is also stored to W (upper triangle).
// EPI 1 (Pandya product): Z = acc on the whole tile and sigma * acc^T on the
//   mirrored tile, both written row-contiguously through an LDS transpose.
template <bool A_KMAJOR, int MODE, int EPI, bool FULL>
__device__ __forceinline__ void seg_tile(const SegProb &p, int m0, int n0, const double *X0, const double *X1,
                                         double g0, double g1, double *C, double *W, double sigma,
                                         unsigned need, GemmSmem &S) {
  const int M = p.M;
  const bool active = need != 0;
  double4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  if (EPI == 0) {
    run_segments<A_KMAJOR, MODE, FULL>(p, 0, p.nmid, X0, X1, g0, g1, m0, n0, S, acc, need);
    if (W && active) {
      double *Wp = W + p.w_off;
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + acc_row<MODE>(ti, v), col = n0 + acc_col(tj);
            if (row < M && col < M && row <= col) Wp[(long long)row * M + col] = acc[ti][tj][v];
          }
    }
    run_segments<A_KMAJOR, MODE, FULL>(p, p.nmid, p.nseg, X0, X1, g0, g1, m0, n0, S, acc, need);
    if (C && active) {
      // every old value of C first (element 0 outside the upper triangle),
      // so the loads are in flight together, then the stores
      double *Cp = C + p.c_off;
      double old[2][2][4];
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + acc_row<MODE>(ti, v), col = n0 + acc_col(tj);
            old[ti][tj][v] = Cp[(row < M && col < M && row <= col) ? (long long)row * M + col : 0];
          }
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + acc_row<MODE>(ti, v), col = n0 + acc_col(tj);
            if (row < M && col < M && row <= col) Cp[(long long)row * M + col] = old[ti][tj][v] + acc[ti][tj][v];
          }
    }
  } else {
    run_segments<A_KMAJOR, MODE, FULL>(p, 0, p.nseg, X0, X1, g0, g1, m0, n0, S, acc, need);
    // Z in 2x2 blocks, upper blocks only: every computed element on or above
    // the diagonal directly, and the strictly lower element of a diagonal block
    // as sigma * its mirror
    if (need) {
      double *Zp = C + p.c_off;
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + acc_row<MODE>(ti, v), col = n0 + acc_col(tj);
            if (row < M && col < M && row <= col) {
              Zp[zidx(row, col, M)] = acc[ti][tj][v];
              if (row < col && (row >> 1) == (col >> 1)) Zp[zidx(col, row, M)] = sigma * acc[ti][tj][v];
            }
          }
    }
  }
}

template <bool A_KMAJOR, int MODE, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 0 ? 4 : 1))) void k_seg_gemm(const SegProb *__restrict__ probs,
                                                  const int4 *__restrict__ tiles, int ntiles,
                                                  const double *__restrict__ X0,
                                                  const double *__restrict__ X1, double g0, double g1,
                                                  double *__restrict__ C, double *__restrict__ W,
                                                  double sigma, const int *act, long long sx,
                                                  long long sc, long long sw) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    X0 += inst * sx;
    X1 += inst * sx;
    if (C) C += inst * sc;
    if (W) W += inst * sw;
  }
  const int4 tl = tiles[t];
  if (tl.x < 0) return;  // padding slot of the XCD-balanced schedule
  const SegProb &p = probs[tl.x];
  const int m0 = tl.y * GT, n0 = tl.z * GT, M = p.M;
  const bool diag = tl.y == tl.z;
  const int wave = threadIdx.x >> 6;
  unsigned need = 0;
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++) {
      const int r16 = m0 + (wave >> 1) * 32 + ti * 16, c16 = n0 + (wave & 1) * 32 + tj * 16;
      if (r16 < M && c16 < M && !(diag && r16 > c16)) need |= 1u << (2 * ti + tj);
    }
  need = __builtin_amdgcn_readfirstlane(need);  // uniform per wavefront
  __shared__ GemmSmem S;
  // workgroup-uniform choice: an off-diagonal tile inside the block needs every sub-tile
  if (MODE == 0 && !diag && m0 + GT <= M && n0 + GT <= M)
    seg_tile<A_KMAJOR, MODE, EPI, true>(p, m0, n0, X0, X1, g0, g1, C, W, sigma, need, S);
  else
    seg_tile<A_KMAJOR, MODE, EPI, false>(p, m0, n0, X0, X1, g0, g1, C, W, sigma, need, S);
}

// ---- FP64 path on 128 x 64 tiles: 4 wavefronts, each 32 rows x 64 columns
// (2 x 4 v_mfma_f64_16x16x4 tiles), so every k-step reads 6 fragments for 8
// MFMAs and a K slab moves 10.7 flop per byte.  Tile (I, J) covers rows
// [128I, 128I+128) and columns [64J, 64J+64); tiles with no element on or
// above the diagonal are not launched.  Writes: direct for row <= col, and
// (EPI 1) mirrored sigma * value for row < col, which covers every element of
// Z exactly once whatever the tile shape.
constexpr int WM = 128, WN = 64;
struct GemmSmemW {
  double A[2][GK][WM + GP];
  double B[2][GK][WN + GP];
};
static_assert(sizeof(GemmSmemW) >= sizeof(double) * 64 * (64 + GP), "transpose buffer must fit");

template <bool FULL>
__device__ __forceinline__ void slab_mfma_w(const GemmSmemW &S, int buf, double4_t acc[2][4], unsigned need) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int kq = 0; kq < GK / 4; kq++) {
    const int k = kq * 4 + (lane >> 4);
    double a[2], b[4];
#pragma unroll
    for (int t = 0; t < 2; t++) a[t] = S.A[buf][k][wave * 32 + t * 16 + (lane & 15)];
#pragma unroll
    for (int t = 0; t < 4; t++) b[t] = S.B[buf][k][t * 16 + (lane & 15)];
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int tj = 0; tj < 4; tj++)
        if (FULL || (need & (1u << (4 * ti + tj))))
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
  }
}

template <bool A_KMAJOR, bool FULL>
__device__ void seg_f64_w(const SegDesc &sgd, double sgn, int M, const double *__restrict__ A,
                          const double *__restrict__ B, int m0, int n0, GemmSmemW &S, double4_t acc[2][4],
                          unsigned need) {
  constexpr int JA = WM * GK / 256, JB = WN * GK / 256;  // elements per thread and slab
  const int K = sgd.K, lda = sgd.lda, ldb = sgd.ldb;
  const int nk = (K + GK - 1) / GK;
  if (nk == 0) return;
  const int tid = threadIdx.x;
  int aoff[JA], akk[JA], boff[JB], bkk[JB];
  unsigned rowok = 0;  // bits 0..JA-1: A rows in range, bits 16..: B columns in range
  if (A_KMAJOR) {
    const int m = m0 + (tid & 127);
    const bool mv = m < M;
#pragma unroll
    for (int j = 0; j < JA; j++) {
      akk[j] = (tid >> 7) + 2 * j;
      aoff[j] = akk[j] * lda + (mv ? m : 0);
      rowok |= (unsigned)mv << j;
    }
  } else {
#pragma unroll
    for (int j = 0; j < JA; j++) {
      const int m = m0 + tid / GK + (256 / GK) * j;
      const bool mv = m < M;
      akk[j] = tid % GK;
      aoff[j] = (mv ? m : 0) * lda + akk[j];
      rowok |= (unsigned)mv << j;
    }
  }
  {
    const int n = n0 + (tid & 63);
    const bool nv = n < M;
#pragma unroll
    for (int j = 0; j < JB; j++) {
      bkk[j] = (tid >> 6) + 4 * j;
      boff[j] = bkk[j] * ldb + (nv ? n : 0);
      rowok |= (unsigned)nv << (16 + j);
    }
  }
  const int ainc = A_KMAJOR ? GK * lda : GK, binc = GK * ldb;
  double ra[JA], rb[JB];
  unsigned ok = 0;
  auto load = [&](int k0) {
    ok = 0;
#pragma unroll
    for (int j = 0; j < JA; j++) {
      const bool g = ((rowok >> j) & 1) && (k0 + akk[j] < K);
      ra[j] = A[g ? aoff[j] : 0];
      ok |= (unsigned)g << j;
      aoff[j] += ainc;
    }
#pragma unroll
    for (int j = 0; j < JB; j++) {
      const bool g = ((rowok >> (16 + j)) & 1) && (k0 + bkk[j] < K);
      rb[j] = B[g ? boff[j] : 0];
      ok |= (unsigned)g << (16 + j);
      boff[j] += binc;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < JA; j++) {
      const double a = ((ok >> j) & 1) ? ra[j] : 0.0;
      if (A_KMAJOR) S.A[buf][(tid >> 7) + 2 * j][tid & 127] = a;
      else S.A[buf][tid % GK][tid / GK + (256 / GK) * j] = a;
    }
#pragma unroll
    for (int j = 0; j < JB; j++) S.B[buf][(tid >> 6) + 4 * j][tid & 63] = ((ok >> (16 + j)) & 1) ? sgn * rb[j] : 0.0;
  };
  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < nk; kt++) {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) load((kt + 1) * GK);
    if (need) slab_mfma_w<FULL>(S, cur, acc, need);
    if (more) store(cur ^ 1);
    __syncthreads();
  }
}

template <bool A_KMAJOR, int EPI, bool FULL>
__device__ __forceinline__ void seg_tile_w(const SegProb &p, int m0, int n0, const double *X0, const double *X1,
                                           double g0, double g1, double *C, double *W, double sigma,
                                           unsigned need, GemmSmemW &S) {
  const int M = p.M;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
  double4_t acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  auto run = [&](int s0, int s1) {
    for (int s = s0; s < s1; s++) {
      const SegDesc &sg = p.seg[s];
      seg_f64_w<A_KMAJOR, FULL>(sg, sg.sign * (sg.src ? g1 : g0), M, (sg.src ? X1 : X0) + sg.a_off,
                          (sg.src ? X0 : X1) + sg.b_off, m0, n0, S, acc, need);
    }
  };
  // accumulator element (ti, tj, v): row lane/16 + 4v of 16x16 tile (ti, tj)
  if (EPI == 0) {
    run(0, p.nmid);
    if (W && need) {
      double *Wp = W + p.w_off;
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + wave * 32 + ti * 16 + g + 4 * v, col = n0 + tj * 16 + (lane & 15);
            if (row < M && col < M && row <= col) Wp[(long long)row * M + col] = acc[ti][tj][v];
          }
    }
    run(p.nmid, p.nseg);
    if (C && need) {
      // every old value of C first (element 0 outside the upper triangle),
      // so the loads are in flight together, then the stores
      double *Cp = C + p.c_off;
      double old[2][4][4];
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + wave * 32 + ti * 16 + g + 4 * v, col = n0 + tj * 16 + (lane & 15);
            old[ti][tj][v] = Cp[(row < M && col < M && row <= col) ? (long long)row * M + col : 0];
          }
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + wave * 32 + ti * 16 + g + 4 * v, col = n0 + tj * 16 + (lane & 15);
            if (row < M && col < M && row <= col) Cp[(long long)row * M + col] = old[ti][tj][v] + acc[ti][tj][v];
          }
    }
  } else {
    run(0, p.nseg);
    if (need) {  // Z in 2x2 blocks, upper blocks only (see zidx)
      double *Zp = C + p.c_off;
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int row = m0 + wave * 32 + ti * 16 + g + 4 * v, col = n0 + tj * 16 + (lane & 15);
            if (row < M && col < M && row <= col) {
              Zp[zidx(row, col, M)] = acc[ti][tj][v];
              if (row < col && (row >> 1) == (col >> 1)) Zp[zidx(col, row, M)] = sigma * acc[ti][tj][v];
            }
          }
    }
  }
}
template <bool A_KMAJOR, int EPI>
__global__ __launch_bounds__(256) void k_seg_gemm_w(const SegProb *__restrict__ probs, const int4 *__restrict__ tiles,
                                                    int ntiles, const double *__restrict__ X0,
                                                    const double *__restrict__ X1, double g0, double g1,
                                                    double *__restrict__ C, double *__restrict__ W, double sigma,
                                                    const int *act, long long sx, long long sc, long long sw) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    X0 += inst * sx;
    X1 += inst * sx;
    if (C) C += inst * sc;
    if (W) W += inst * sw;
  }
  const int4 tl = tiles[t];
  if (tl.x < 0) return;
  const SegProb &p = probs[tl.x];
  const int m0 = tl.y * WM, n0 = tl.z * WN, M = p.M;
  const int wave = threadIdx.x >> 6;
  unsigned need = 0;
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
      const int r16 = m0 + wave * 32 + ti * 16, c16 = n0 + tj * 16;
      if (r16 < M && c16 < M && r16 <= c16) need |= 1u << (4 * ti + tj);
    }
  need = __builtin_amdgcn_readfirstlane(need);  // uniform per wavefront
  __shared__ GemmSmemW S;
  // workgroup-uniform choice: every 16x16 sub-tile inside the block and on or above the diagonal
  if (m0 + WM <= M && n0 + WN <= M && m0 + WM - 16 <= n0)
    seg_tile_w<A_KMAJOR, EPI, true>(p, m0, n0, X0, X1, g0, g1, C, W, sigma, need, S);
  else
    seg_tile_w<A_KMAJOR, EPI, false>(p, m0, n0, X0, X1, g0, g1, C, W, sigma, need, S);
}

void launch_seg_gemm_pphh(hipStream_t s, int mode, const SegProb *probs, const int4 *tiles, int ntiles,
                          const double *X0, const double *X1, double *C, double *W, const int *act, int nact,
                          long long sop, long long sw) {
  if (ntiles <= 0 || nact <= 0) return;
  const dim3 g(ntiles, nact), b(256);
  if (mode == 3) {  // FP64 on 128 x 64 tiles (tiles built for that shape)
    launch_k(k_seg_gemm_w<false, 0>, g, b, 0, s, probs, tiles, ntiles, X0, X1, 1.0, 1.0, C, W, 0.0, act,
                       sop, sop, sw);
    return;
  }
  if (mode == 1)
    launch_k(k_seg_gemm<false, 1, 0>, g, b, 0, s, probs, tiles, ntiles, X0, X1, 1.0, 1.0, C, W, 0.0,
                       act, sop, sop, sw);
  else if (mode == 2)
    launch_k(k_seg_gemm<false, 2, 0>, g, b, 0, s, probs, tiles, ntiles, X0, X1, 1.0, 1.0, C, W, 0.0,
                       act, sop, sop, sw);
  else
    launch_k(k_seg_gemm<false, 0, 0>, g, b, 0, s, probs, tiles, ntiles, X0, X1, 1.0, 1.0, C, W, 0.0,
                       act, sop, sop, sw);
}
void launch_seg_gemm_z(hipStream_t s, int mode, const SegProb *probs, const int4 *tiles, int ntiles,
                       const double *LCC, const double *RCC, double *Z, double hL, double hR, const int *act,
                       int nact, long long scc, long long sz) {
  if (ntiles <= 0 || nact <= 0) return;
  const dim3 g(ntiles, nact), b(256);
  const double g0 = hL, g1 = -hR, sigma = -hL * hR;
  if (mode == 3) {  // FP64 on 128 x 64 tiles
    launch_k(k_seg_gemm_w<true, 1>, g, b, 0, s, probs, tiles, ntiles, LCC, RCC, g0, g1, Z, nullptr, sigma,
                       act, scc, sz, 0LL);
    return;
  }
  if (mode == 1)
    launch_k(k_seg_gemm<true, 1, 1>, g, b, 0, s, probs, tiles, ntiles, LCC, RCC, g0, g1, Z, nullptr,
                       sigma, act, scc, sz, 0LL);
  else if (mode == 2)
    launch_k(k_seg_gemm<true, 2, 1>, g, b, 0, s, probs, tiles, ntiles, LCC, RCC, g0, g1, Z, nullptr,
                       sigma, act, scc, sz, 0LL);
  else
    launch_k(k_seg_gemm<true, 0, 1>, g, b, 0, s, probs, tiles, ntiles, LCC, RCC, g0, g1, Z, nullptr,
                       sigma, act, scc, sz, 0LL);
}

// ------------------------------------------------------------------ Pandya transform
// calculate_cross_coupled (cross_coupled.f90:1190-1300): one workgroup per
// Pandya column (p,h), threads over rows (a,b); output column-major per block.
// Everything a term needs except the operator element itself is staged in
// LDS once per column: the 6j slice {jp jh JC; ja jb J'}, the pair positions
// of (h,a) and (p,b) in every J' block, block offsets and sizes, and the
// orbit quantum numbers, so each J' term does a single global load.
__global__ __launch_bounds__(256) void k_pandya(DevLayout L, const double *__restrict__ op,
                                                double *__restrict__ cc, const int *act) {
  extern __shared__ double pan_lds[];
  const int col = blockIdx.x;
  if (col >= L.nccol) return;
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    op += inst * L.op_size;
    cc += inst * L.ccsize;
  }
  const int q1 = L.cccol_q[col];
  const int c = col - L.cccbase[q1];
  const int p = L.cccol_p[col], h = L.cccol_h[col];
  const int JC = L.ccJ[q1], rlen = L.ccrlen[q1];
  const int jp = L.jj[p], jh = L.jj[h], lp = L.l[p], lh = L.l[h], tp = L.tz[p], th = L.tz[h];
  const double *V = op + L.v_off;
  double *out = cc + L.ccoff[q1] + (long long)c * rlen;
  const double sq = sqrt(JC + 1.0);
  const int N = L.nsp, h6 = L.nhalf, nJ = L.nJ, nsl = h6 * h6 * nJ, NB = L.nblocks;
  const size_t NN = (size_t)N * N;
  double *S6 = pan_lds;
  long long *BO = reinterpret_cast<long long *>(S6 + nsl);
  int *BD = reinterpret_cast<int *>(BO + NB);
  int *OQ = BD + NB;                           // jj | l << 8 | (2tz + 1) << 16
  // [J'/2][a]: position of pair (h,a) in block J'; [J'/2][b]: of pair (p,b)
  // (16 bits: positions below 2^15, -1 absent; the smaller list lets eight
  // wavefronts per SIMD stay resident)
  short *PR = reinterpret_cast<short *>(OQ + N);
  short *PS = PR + nJ * N;
  {
    // every staging load of a batch is issued before the first LDS store, so
    // a thread waits for one memory latency per batch, not one per element
    const double *six0 = L.sixj + (((size_t)(jp >> 1) * h6 + (jh >> 1)) * nJ + (JC >> 1)) * (size_t)nsl;
    for (int e0 = 0; e0 < nsl; e0 += 256 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = six0[min(e0 + (int)threadIdx.x + 256 * u, nsl - 1)];  // off any branch
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + (int)threadIdx.x + 256 * u;
        if (e < nsl) S6[e] = v[u];
      }
    }
    for (int e = threadIdx.x; e < NB; e += 256) {
      BO[e] = L.boff[e];
      BD[e] = L.bdim[e];
    }
    for (int e0 = 0; e0 < nJ * N; e0 += 256 * 8) {
      int vr[8], vs[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = min(e0 + (int)threadIdx.x + 256 * u, nJ * N - 1);  // off any branch
        const int J2 = e / N, x = e - J2 * N;
        const int *pj = L.pidx + (size_t)J2 * NN;
        vr[u] = pj[h > x ? x * N + h : h * N + x];
        vs[u] = pj[p > x ? x * N + p : p * N + x];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + (int)threadIdx.x + 256 * u;
        if (e < nJ * N) {
          PR[e] = (short)vr[u];
          PS[e] = (short)vs[u];
        }
      }
    }
    for (int e = threadIdx.x; e < N; e += 256) OQ[e] = L.jj[e] | (L.l[e] << 8) | ((L.tz[e] + 1) << 16);
  }
  __syncthreads();
  // the next row's orbits are loaded while the current row's terms are summed
  const int *rowa = L.ccrow_a + L.ccrbase[q1], *rowb = L.ccrow_b + L.ccrbase[q1];
  int a_next = -1, b_next = -1;
  if ((int)threadIdx.x < rlen) {
    a_next = rowa[threadIdx.x];
    b_next = rowb[threadIdx.x];
  }
  for (int r = threadIdx.x; r < rlen; r += blockDim.x) {
    const int a = a_next, b = b_next;
    if (r + (int)blockDim.x < rlen) {
      a_next = rowa[r + blockDim.x];
      b_next = rowb[r + blockDim.x];
    }
    if (a < 0) {  // zero row that makes the row count even
      out[r] = 0.0;
      continue;
    }
    const int qa = OQ[a], qbb = OQ[b];
    const int ja = qa & 255, jb = qbb & 255, la = (qa >> 8) & 255, lb = (qbb >> 8) & 255;
    const int ta = (qa >> 16) - 1, tb = (qbb >> 16) - 1;
    double val = 0.0;
    if ((((la + lh) & 1) == ((lb + lp) & 1)) && (ta + th == tb + tp)) {
      const int Jmin = max(abs(jp - jb), abs(ja - jh));
      const int Jmax = min(jp + jb, ja + jh);
      // v_elem(h,a,p,b;J'): bra pair (h,a), ket pair (p,b), block (J', T2, P)
      const int qb = 2 * ((th + ta) / 2 + 1) + ((lh + la) & 1);
      const bool sw1 = h > a, sw2 = p > b;
      const double *six = S6 + ((ja >> 1) * h6 + (jb >> 1)) * nJ;
      double sm = 0.0;
      // eight J' terms a batch; the element offsets and weights first (a J'
      // past Jmax, or an element the block does not hold, reads element 0
      // with weight 0), then the eight gathers, none of them on a branch, so
      // they are in flight together
      for (int J0 = Jmin; J0 <= Jmax; J0 += 16) {
        long long off[8];
        double w[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int JT = min(J0 + 2 * u, Jmax);
          const int rr = PR[(JT >> 1) * N + a], ss = PS[(JT >> 1) * N + b], q = 3 * JT + qb;
          double pre = 1.0;
          if (sw1) pre = sgnp(1 + ((jh + ja - JT) >> 1));
          else if (h == a) pre = pre * M_SQRT2;
          if (sw2) pre = pre * sgnp(1 + ((jp + jb - JT) >> 1));
          else if (p == b) pre = pre * M_SQRT2;
          const bool ok = J0 + 2 * u <= Jmax && (rr | ss) >= 0;
          off[u] = ok ? BO[q] + (long long)rr * BD[q] + ss : 0;
          w[u] = ok ? pre * (double)(sgnp(JT >> 1) * (JT + 1)) * six[JT >> 1] : 0.0;
        }
        double vv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) vv[u] = V[off[u]];
#pragma unroll
        for (int u = 0; u < 8; u++) sm = sm + w[u] * vv[u];
      }
      val = sm * sgnp((jh + jb + JC) >> 1) * sq;
    }
    out[r] = val;
  }
}
size_t pandya_lds_bytes(const DevLayout &L) {
  return sizeof(double) * L.nhalf * L.nhalf * L.nJ + sizeof(long long) * L.nblocks +
         sizeof(int) * (L.nblocks + L.nsp) + sizeof(short) * 2 * L.nJ * L.nsp;
}
// lets k_pandya take more than 64 KiB of dynamic LDS (up to the device's per-workgroup limit)
hipError_t pandya_allow_lds(size_t bytes) {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pandya), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}
void launch_pandya(hipStream_t s, const DevLayout &L, const double *op, double *cc, const int *act, int nact) {
  if (L.nccol <= 0 || nact <= 0) return;
  hipLaunchKernelGGL(k_pandya, dim3(L.nccol, nact), dim3(256), pandya_lds_bytes(L), s, L, op, cc, act);
}

// ------------------------------------------------------------------ 122
// commutator_122 (commutators.f90:283-396).  With M = A_L R - A_R L, where
// A_X is the pair-space action of X's one-body part,
//   M(r,s) = sum_{e in nb(r)} c_e [fL(x,i) R(r_e,s) - fR(x,i) L(r_e,s)],
// the right-multiplication half equals -hL*hR*M(s,r) (hermiticity of f and
// V), so RES(r,s) = M(r,s) + herm(RES)*M(s,r).  The 122 kernels write M over
// the whole block; the diagonal is written already symmetrised,
// M(r,r)*(1+herm).  k_sym_mirror adds the transpose at the end.
// The row form, for the rows of the sets the set form does not take (past 32
// rows, or a slot with more than two entries): one wavefront per row,
// coalesced 16-byte row reads, wave-uniform neighbour entries.
__global__ __launch_bounds__(64) void k_122(DevLayout L, const double *__restrict__ Lop,
                                            const double *__restrict__ Rop, double *__restrict__ RES,
                                            double hres, const int *act) {
  const int ri = xcd_item(L.xs_rows);
  if (ri < 0) return;
  const int r = L.s122_frows[ri];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    const long long o = inst * L.op_size;
    Lop += o;
    Rop += o;
    RES += o;
  }
  const int q = L.pair_q[r], rl = r - L.pbase[q], dim = L.bdim[q], N = L.nsp;
  const long long bo = L.v_off + L.boff[q];
  const double *LV = Lop + bo, *RV = Rop + bo;
  double *OV = RES + bo + (long long)rl * dim;
  const double *FL = Lop + L.f_off, *FR = Rop + L.f_off;
  const int e0 = L.nb_start[r], e1 = L.nb_start[r + 1];
  // eight outputs per lane, as four pairs s, s + 1 (s = s0 + 128 u + 2 lane,
  // 16-byte accesses); each neighbour entry issues 8 independent row loads
  // and its scalars are read once for all eight
  for (int s0 = 0; s0 < dim; s0 += 512) {
    dpair_t val[4];
#pragma unroll
    for (int u = 0; u < 4; u++) val[u] = (dpair_t){0.0, 0.0};
    for (int e = e0; e < e1; e++) {
      const long long pe = L.nb_pos[e];
      const int xi = L.nb_x[e] * N + L.nb_i[e];
      const double ce = L.nb_coef[e], fl = FL[xi], fr = FR[xi];
      const double *rrow = RV + pe * dim, *lrow = LV + pe * dim;
      dpair_t rv[4], lv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int s = s0 + 128 * u + 2 * (int)threadIdx.x;
        if (s + 1 < dim) {
          rv[u] = *(const dpair_t *)(rrow + s);
          lv[u] = *(const dpair_t *)(lrow + s);
        } else {  // the last element of an odd row, or past the row
          rv[u] = (dpair_t){s < dim ? rrow[s] : 0.0, 0.0};
          lv[u] = (dpair_t){s < dim ? lrow[s] : 0.0, 0.0};
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) val[u] += ce * (fl * rv[u] - fr * lv[u]);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int s = s0 + 128 * u + 2 * (int)threadIdx.x;
      dpair_t v = val[u];
      if (s == rl) v.x *= 1.0 + hres;
      if (s + 1 == rl) v.y *= 1.0 + hres;
      if (s + 1 < dim) *(dpair_t *)(OV + s) = v;
      else if (s < dim) OV[s] = v.x;
    }
  }
}
// The set form.  A set holds n <= 32 rows of one block whose neighbour rows
// all lie in the set, so over any column the set's 122 outputs are a small
// dense product: M = A R - B L (n x dim), A(k,k') = sum of c_e fL(e) over the
// entries of row k whose neighbour is row k' (B with fR).  Written as one
// product M = W X with W(k, 2k'+t) = (A, -B)(k,k') and X(2k'+t, c) = (R, L)
// rows, it runs on v_mfma_f64_16x16x4.  k_122w forms every set's W (n x 2n,
// row-major) from its slot entries, one thread per slot.
__global__ __launch_bounds__(256) void k_122w(DevLayout L, const double *__restrict__ Lop,
                                              const double *__restrict__ Rop, double *__restrict__ W,
                                              const int *act) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= L.n122_slots) return;
  const long long inst = inst_id(act);
  if (inst < 0) return;
  const double *FL = Lop + inst * L.op_size + L.f_off, *FR = Rop + inst * L.op_size + L.f_off;
  const int2 x = L.s122_sx[i];
  const dpair_t c = *(const dpair_t *)(L.s122_sc + 2 * (long long)i);
  const dpair_t w = {c.x * FL[x.x] + c.y * FL[x.y], -(c.x * FR[x.x] + c.y * FR[x.y])};
  *(dpair_t *)(W + inst * 2 * (long long)L.n122_slots + 2 * (long long)i) = w;
}

// One wavefront per (set, 64 columns) of k_122w's W times the set's R and L
// rows.  A k-step of 4 covers rows 2kq, 2kq+1, R and L each (k = 4kq +
// lane/16 reads row 2kq + lane/32, L for odd k); the set is padded to 2*NK
// rows (pad rows read the last row again with zero weight).  A chunk of 64
// columns inside the block runs as two 32-column halves, each lane reading
// columns 2(lane%16), +1 of a half as one 16-byte access: its even and odd
// columns are two 16-column tiles.  The block's last chunk, when shorter,
// runs as four 16-column tiles with clamped columns.  Sets past 16 rows run
// as two 16-row output tiles over two 16-row halves of the rows read.
template <int NK, int TU, bool FULL>
__device__ __forceinline__ void s122_wave(const int *__restrict__ RL, const double *__restrict__ Wt, int n, int dim,
                                          const double *src, int c0, double hres, double *OV, int nhalf) {
  const int lane = threadIdx.x & 63, cc = lane & 15, g = lane >> 4, n2 = 2 * n;
  for (int mt = 0; mt < nhalf; mt++) {
    double4_t acc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) acc[u] = (double4_t){0.0, 0.0, 0.0, 0.0};
    for (int kh = 0; kh < nhalf; kh++) {
      int pr[NK];
      double a[NK];
#pragma unroll
      for (int kq = 0; kq < NK; kq++) {
        const int r0 = RL[min(16 * kh + 2 * kq, n - 1)], r1 = RL[min(16 * kh + 2 * kq + 1, n - 1)];
        pr[kq] = (g >> 1) ? r1 : r0;
        const int m = 16 * mt + cc, j = 32 * kh + 4 * kq + g;
        const double w = Wt[min(m, n - 1) * n2 + min(j, n2 - 1)];
        a[kq] = (m < n && j < n2) ? w : 0.0;
      }
      if (FULL) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
          dpair_t x[NK];
#pragma unroll
          for (int kq = 0; kq < NK; kq++) x[kq] = *(const dpair_t *)(src + (long long)pr[kq] * dim + c0 + 32 * h + 2 * cc);
#pragma unroll
          for (int kq = 0; kq < NK; kq++) {
            acc[2 * h] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kq], x[kq].x, acc[2 * h], 0, 0, 0);
            acc[2 * h + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kq], x[kq].y, acc[2 * h + 1], 0, 0, 0);
          }
        }
      } else {
        int col[4];
#pragma unroll
        for (int u = 0; u < 4; u++) col[u] = min(c0 + 16 * u + cc, dim - 1);
#pragma unroll
        for (int u0 = 0; u0 < 4; u0 += TU) {
          double b[NK][TU];
#pragma unroll
          for (int kq = 0; kq < NK; kq++)
#pragma unroll
            for (int u = 0; u < TU; u++) b[kq][u] = src[(long long)pr[kq] * dim + col[u0 + u]];
#pragma unroll
          for (int kq = 0; kq < NK; kq++)
#pragma unroll
            for (int u = 0; u < TU; u++)
              acc[u0 + u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kq], b[kq][u], acc[u0 + u], 0, 0, 0);
        }
      }
    }
    // accumulator element v: row 16 mt + lane/16 + 4v, column lane%16 of its tile
    const double dg = 1.0 + hres;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int m = 16 * mt + g + 4 * v;
      if (m < n) {
        const int rk = RL[m];
        double *orow = OV + (long long)rk * dim;
        if (FULL) {
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const int c = c0 + 32 * h + 2 * cc;
            dpair_t y = {acc[2 * h][v], acc[2 * h + 1][v]};
            if (c == rk) y.x *= dg;
            if (c + 1 == rk) y.y *= dg;
            *(dpair_t *)(orow + c) = y;
          }
        } else {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int c = c0 + 16 * u + cc;
            if (c < dim) orow[c] = c == rk ? acc[u][v] * dg : acc[u][v];
          }
        }
      }
    }
  }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_122s(
    DevLayout L, const double *__restrict__ Lop, const double *__restrict__ Rop, double *__restrict__ RES,
    const double *__restrict__ W, double hres, const int *act) {
  const int2 it = L.s122_items[blockIdx.x];
  const long long inst = inst_id(act);
  if (inst < 0) return;
  const long long o = inst * L.op_size;
  const Set122 S = L.s122_sets[it.x];
  const int n = S.n, dim = S.dim, c0 = 64 * it.y;
  const long long bo = o + L.v_off + S.bo;
  const int *RL = L.s122_rowl + S.rs;
  const double *Wt = W + inst * 2 * (long long)L.n122_slots + S.ab;
  // k = 4kq + lane/16: odd k read L
  const double *src = ((threadIdx.x >> 4) & 1) ? Lop + bo : Rop + bo;
  double *OV = RES + bo;
  if (c0 + 64 <= dim) {
    if (n <= 2) s122_wave<1, 4, true>(RL, Wt, n, dim, src, c0, hres, OV, 1);
    else if (n <= 4) s122_wave<2, 4, true>(RL, Wt, n, dim, src, c0, hres, OV, 1);
    else if (n <= 8) s122_wave<4, 2, true>(RL, Wt, n, dim, src, c0, hres, OV, 1);
    else s122_wave<8, 1, true>(RL, Wt, n, dim, src, c0, hres, OV, n > 16 ? 2 : 1);
  } else {
    if (n <= 2) s122_wave<1, 4, false>(RL, Wt, n, dim, src, c0, hres, OV, 1);
    else if (n <= 4) s122_wave<2, 4, false>(RL, Wt, n, dim, src, c0, hres, OV, 1);
    else if (n <= 8) s122_wave<4, 2, false>(RL, Wt, n, dim, src, c0, hres, OV, 1);
    else s122_wave<8, 1, false>(RL, Wt, n, dim, src, c0, hres, OV, n > 16 ? 2 : 1);
  }
}

void launch_122(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop, double *RES,
                int hres, const int *act, int nact, double *W) {
  if (L.npairs <= 0 || nact <= 0) return;
  if (L.n122_items > 0) {
    hipLaunchKernelGGL(k_122w, dim3((L.n122_slots + 255) / 256, nact), dim3(256), 0, s, L, Lop, Rop, W, act);
    hipLaunchKernelGGL(k_122s, dim3(L.n122_items, nact), dim3(64), 0, s, L, Lop, Rop, RES, (const double *)W,
                       (double)hres, act);
  }
  if (L.n122_frows > 0)
    hipLaunchKernelGGL(k_122, dim3(8 * L.xs_rows.len, nact), dim3(64), 0, s, L, Lop, Rop, RES, (double)hres, act);
}

// zero every channel block (both triangles): the start state when 122 is off
__global__ __launch_bounds__(64) void k_clear_rows(DevLayout L, double *RES, const int *act) {
  const int r = blockIdx.x;
  if (r >= L.npairs) return;
  const long long inst = inst_id(act);
  if (inst < 0) return;
  RES += inst * L.op_size;
  const int q = L.pair_q[r], rl = r - L.pbase[q], dim = L.bdim[q];
  double *OV = RES + L.v_off + L.boff[q] + (long long)rl * dim;
  for (int s = threadIdx.x; s < dim; s += 64) OV[s] = 0.0;
}
void launch_clear_blocks(hipStream_t s, const DevLayout &L, double *RES, const int *act, int nact) {
  if (L.npairs <= 0 || nact <= 0) return;
  hipLaunchKernelGGL(k_clear_rows, dim3(L.npairs, nact), dim3(64), 0, s, L, RES, act);
}

// ------------------------------------------------------------------ 222_ph recoupling gather
// commutators.f90:804-923 rewritten on Z = hL*W - hR*W^T (Z^T = sigma*Z,
// sigma = -hL*hR).  For an upper element (ij,kl;J), with Z(ab,cd) =
// Z[row(a,b)][row(c,d)] of Pandya block J' and ph = (-1)^((ji+jj+jk+jl)/2):
//   sum1 = -(-1)^((jl+jk+J)/2) sum_J' {ji jj J; jl jk J'} T1(J'),
//          T1 = Z(ki,jl) + ph*sigma*Z(ik,lj)
//   sum2 =  sum_J' {ji jj J; jk jl J'} T2(J'),
//          T2 = sigma*Z(jk,li) + ph*Z(kj,il)           (symmetry applied)
// T1/T2 depend on the (i,j,k,l) quadruple only, not on J.  Phase A gathers
// every T once (one thread per J' term, terms ordered by quadruple sorted by
// (i,k,j,l) so that runs of threads share Z rows); phase B adds the 6j-weighted
// sums to every upper element, one wavefront per RES row.
// Z is stored in 2x2 blocks, upper blocks only: Z(R1,C1) and Z(R2,C2), where
// (R1,R2) and (C1,C2) are the two orders of one orbit pair each, lie in one
// aligned 16-byte half of block (R1/2, C1/2), or of its mirror (C1/2, R1/2)
// times sigma (see zidx), unless a pair is (a,a): then R2 = R1 or C2 = C1 and
// the second element is in the other half of the block.
// Four J' terms per thread (t + 256 u of a chunk of 1024): T = z1 + ph*sigma*z2
// (side 0) or sigma*z1 + ph*z2 (side 1), both Z elements read with one
// 16-byte load of the half the host-built address table names (a second one
// for the (a,a) terms); the four terms' table reads, then their four
// half-block reads, are in flight together.
__device__ __forceinline__ double ph_term(unsigned sel, double2 b, double2 c, double sigma) {
  const double2 h2 = (sel & 2) ? c : b;
  double z1 = (sel & 1) ? b.y : b.x, z2 = (sel & 4) ? h2.y : h2.x;
  if (sel & 16) {
    z1 = sigma * z1;
    z2 = sigma * z2;
  }
  const double ph = (sel & 32) ? -1.0 : 1.0;
  const double v = (sel & 64) ? sigma * z1 + ph * z2 : z1 + ph * sigma * z2;
  return (sel & 128) ? v : 0.0;
}
// Work items past the XCD split's are the (a,a) terms, one per thread; the
// chunks skip them.  Every load is unconditional (past the end: the last
// term; invalid terms: half 0), so the four terms' loads stay in flight together.
__global__ __launch_bounds__(256) void k_ph_terms(DevLayout L, const double *__restrict__ Z,
                                                  double *__restrict__ Tbuf, double sigma, const int *act) {
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    Z += inst * L.cwsize;
    Tbuf += inst * L.nphterm;
  }
  const double2 *Z2 = reinterpret_cast<const double2 *>(Z);
  const int nchunk = 8 * L.xs_terms.len;
  if ((int)blockIdx.x >= nchunk) {
    const long long k = (long long)(blockIdx.x - nchunk) * 256 + threadIdx.x;
    if (k >= L.nphtwo) return;
    const unsigned t = L.ph_two[k], sel = L.ph_zsel[t], blk = L.ph_zblk[t];
    Tbuf[t] = ph_term(sel, Z2[blk], Z2[blk ^ 1u], sigma);
    return;
  }
  const int chunk = xcd_item(L.xs_terms);
  if (chunk < 0) return;
  const long long t0 = (long long)chunk * 1024 + threadIdx.x;
  unsigned sel[4], blk[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const long long t = min(t0 + 256 * u, L.nphterm - 1);
    sel[u] = L.ph_zsel[t];
    blk[u] = L.ph_zblk[t];
  }
  double2 b[4];
#pragma unroll
  for (int u = 0; u < 4; u++) b[u] = Z2[blk[u]];
  // the four half-block reads are all issued here (a value used only under
  // the store's condition would otherwise be loaded inside it, one at a time)
  asm volatile("" ::"v"(b[0].x), "v"(b[0].y), "v"(b[1].x), "v"(b[1].y), "v"(b[2].x), "v"(b[2].y), "v"(b[3].x),
               "v"(b[3].y));
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const long long t = t0 + 256 * u;
    if (t < L.nphterm && (sel[u] & 130) != 130) Tbuf[t] = ph_term(sel[u], b[u], b[u], sigma);
  }
}
void launch_ph_terms(hipStream_t s, const DevLayout &L, const double *Z, double *Tbuf, double hL, double hR,
                     const int *act, int nact) {
  if (L.nphterm <= 0 || nact <= 0) return;
  const unsigned g = 8 * L.xs_terms.len + (unsigned)((L.nphtwo + 255) / 256);
  hipLaunchKernelGGL(k_ph_terms, dim3(g, nact), dim3(256), 0, s, L, Z, Tbuf, -hL * hR, act);
}

// The 6j symbols a row can use, {ji jj J; x y J'} with triangle (x, y, J) and
// J' in the triangle ranges of the two J' sums, are one packed list per
// (ji, jj, J), staged in LDS once per work item.  A work item is up to 64
// outputs of one row (one wavefront); rows run in (i, j, J) order, split into
// 8 contiguous ranges (one per XCD): the rows of one orbit pair read the same
// quadruples' J' terms, so the repeats hit L2.  What each output needs beyond
// its row (where its J' terms start, how many, where its 6j runs start in the
// list, signs) is one 64-bit host-built descriptor; what the row needs is one
// PhChunk header.  An output's two J' ranges are one contiguous run in Tbuf
// that starts at an even offset (runs are padded to even lengths with zero
// terms).  Four lanes read an output's run together, each two terms with one
// 16-byte load, so the four read 64 contiguous bytes per instruction (one or
// two cache lines, where one lane per output touched a line per term); the
// wavefront covers its 64 outputs as four groups of 16, with every group's
// descriptor, old value and first two loads issued before any is used, and
// two lane exchanges add each output's four partial sums.
__device__ __forceinline__ void ph_decode(unsigned long long d, int &toff, int &n1, int &n, int &o1, int &o2,
                                          double &g1, double &g2, double &pre2) {
  toff = (int)(d & 0x7FFFFFull);
  n1 = (int)((d >> 23) & 31);
  n = n1 + (int)((d >> 28) & 31);
  o1 = (int)((d >> 33) & 4095);
  o2 = (int)((d >> 45) & 4095) - n1;
  g1 = ((d >> 57) & 1) ? (((d >> 58) & 1) ? -1.0 : 1.0) : 0.0;
  g2 = ((d >> 59) & 1) ? 1.0 : 0.0;
  pre2 = ((d >> 60) & 1) ? .70710678118 : 1.0;
}
// 6j weight of term m of an output (0 past its run); index 0 when unused
__device__ __forceinline__ double ph_coef(const double *S6, int m, int n, int n1, int o1, int o2, double g1,
                                          double g2) {
  const bool s1 = m < n1, in = m < n;
  const int e = in ? (s1 ? o1 + m : o2 + m) : 0;
  return (in ? (s1 ? g1 : g2) : 0.0) * S6[e];
}
__global__ __launch_bounds__(64) void k_ph_apply(DevLayout L, const double *__restrict__ Tbuf,
                                                 double *__restrict__ RES, const int *act) {
  extern __shared__ double S6[];  // max(ap6_max, 64) doubles
  const int it = xcd_item(L.xs_apply);
  if (it < 0) return;
  const PhChunk h = L.ph_chunks[it];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    Tbuf += inst * L.nphterm + h.tbase;  // descriptors hold offsets relative to the row's first term
    RES += inst * L.op_size;
  }
  double *OV = RES + h.vrow;
  const unsigned long long *desc = L.ph_desc + h.dbase;
  const int g = (int)(threadIdx.x >> 2), sub = (int)(threadIdx.x & 3);
  // descriptors and old values of outputs g, g + 16, g + 32, g + 48
  unsigned long long d[4];
  double old[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int o = g + 16 * u;
    d[u] = o < h.n ? desc[o] : 0ull;
    old[u] = (o < h.n && sub == 0) ? OV[o] : 0.0;
  }
  // the list, zero-padded to 64 entries: an output whose J' sum is absent
  // reads (weight 0) entries below its run length, never above 62
  {
    const int ns = max(h.len, 64);
    for (int e0 = 0; e0 < ns; e0 += 256) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + (int)threadIdx.x + 64 * u;
        v[u] = e < h.len ? L.ap6_val[h.b0 + e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + (int)threadIdx.x + 64 * u;
        if (e < ns) S6[e] = v[u];
      }
    }
  }
  // terms m, m + 1 for m = 2 sub + 8 w: two loads per output in flight cover n <= 16
  // (only the run offset and length are decoded here; the rest at use, so
  // that few registers stay live across the loads)
  const int m0 = 2 * sub;
  double2 t[4][2];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int toff = (int)(d[u] & 0x7FFFFFull), n = (int)((d[u] >> 23) & 31) + (int)((d[u] >> 28) & 31);
    const double2 *T = reinterpret_cast<const double2 *>(Tbuf + toff);
#pragma unroll
    for (int w = 0; w < 2; w++) t[u][w] = (m0 + 8 * w < n) ? T[(m0 >> 1) + 4 * w] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double pre = h.half ? .70710678118 : 1.0;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    int toff, n1, n, o1, o2;
    double g1, g2, pre2;
    ph_decode(d[u], toff, n1, n, o1, o2, g1, g2, pre2);
    double acc = 0.0;
#pragma unroll
    for (int w = 0; w < 2; w++) {
      const int m = m0 + 8 * w;
      acc += ph_coef(S6, m, n, n1, o1, o2, g1, g2) * t[u][w].x + ph_coef(S6, m + 1, n, n1, o1, o2, g1, g2) * t[u][w].y;
    }
    const double2 *T = reinterpret_cast<const double2 *>(Tbuf + toff);
    for (int m = m0 + 16; m < n; m += 8) {
      const double2 tt = T[m >> 1];
      acc += ph_coef(S6, m, n, n1, o1, o2, g1, g2) * tt.x + ph_coef(S6, m + 1, n, n1, o1, o2, g1, g2) * tt.y;
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    const int o = g + 16 * u;
    if (sub == 0 && o < h.n) OV[o] = old[u] + acc * pre2 * pre;
  }
}
void launch_ph_apply(hipStream_t s, const DevLayout &L, const double *Tbuf, double *RES, const int *act,
                     int nact) {
  if (L.nphterm <= 0 || nact <= 0) return;
  const size_t lds = sizeof(double) * std::max(L.ap6_max, 64);  // the list is staged zero-padded to >= 64
  hipLaunchKernelGGL(k_ph_apply, dim3(8 * L.xs_apply.len, nact), dim3(64), lds, s, L, Tbuf, RES, act);
}

// ------------------------------------------------------------------ symmetrise + mirror
// Final pass over every channel block: the upper triangle U holds M + the
// accumulated 222 terms, the strict lower triangle still holds M^T's source
// M(s,r) from the 122 kernels, so for r < s
//   RES(r,s) = U(r,s) + herm*lower(s,r),  RES(s,r) = herm*RES(r,s);
// the diagonal is kept (zeroed for herm = -1).  This reproduces the
// reference's "upper triangle then mirror x herm" contract (commutators.f90
// 388-395, 915-923).  32x32 tile pairs go through LDS so all reads and
// writes are row-contiguous.
__global__ __launch_bounds__(256) void k_sym_mirror(DevLayout L, const int4 *__restrict__ tiles, double *RES,
                                                    int herm, const int *act, double *acc, double coef,
                                                    const double *coefp, double *npart, long long nstride) {
  __shared__ double sh[256];
  const long long inst = inst_id(act);
  if (inst < 0) return;
  if (coefp) coef = *coefp;
  RES += inst * L.op_size;
  const int4 t = tiles[blockIdx.x];  // x = block, y = tile row, z = tile col (y <= z)
  const int q = t.x, dim = L.bdim[q];
  double *OV = RES + L.v_off + L.boff[q];
  double *AV = acc ? acc + inst * L.op_size + L.v_off + L.boff[q] : nullptr;
  __shared__ double U[32][33], W[32][33];
  const int r0 = t.y * 32, c0 = t.z * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const double h = (double)herm;
  // a thread writes exactly the elements it loads (upper (r0+k, c0+tx) and
  // lower (c0+k, r0+tx)), so the accumulator's old values are loaded here
  // too: every load of the tile pair is in flight at once, the positions past
  // the block read element 0 and are never stored
  const double *AR = AV ? AV : OV;
  double au[4], al[4], x[4], y[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int k = ty + 8 * m;
    const int r = r0 + k, c = c0 + tx, r2 = c0 + k, c2 = r0 + tx;
    const long long o1 = (r < dim && c < dim) ? (long long)r * dim + c : 0;
    const long long o2 = (r2 < dim && c2 < dim) ? (long long)r2 * dim + c2 : 0;
    x[m] = OV[o1];
    y[m] = OV[o2];
    au[m] = AR[o1];  // (RES itself when there is no accumulator: a load off any branch)
    al[m] = AR[o2];
  }
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int k = ty + 8 * m;
    const bool in1 = r0 + k < dim && c0 + tx < dim, in2 = c0 + k < dim && r0 + tx < dim;
    U[k][tx] = in1 ? x[m] : 0.0;  // upper tile (r0, c0)
    W[k][tx] = in2 ? y[m] : 0.0;  // lower tile (c0, r0)
  }
  __syncthreads();
  const bool diag = t.y == t.z;
  double sq = 0.0;
  // upper: element (r0+k, c0+tx)
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int k = ty + 8 * m;
    const int r = r0 + k, c = c0 + tx;
    if (r < dim && c < dim && (!diag || k <= tx)) {
      double v = U[k][tx];
      if (!diag || k < tx) v += h * W[tx][k];
      else if (herm == -1) v = 0.0;
      const long long o = (long long)r * dim + c;
      OV[o] = v;
      sq += v * v;
      if (AV) AV[o] = au[m] + coef * v;
    }
  }
  // lower: element (c0+k, r0+tx) = herm * upper (r0+tx, c0+k)
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int k = ty + 8 * m;
    const int r = c0 + k, c = r0 + tx;
    if (r < dim && c < dim && (!diag || tx < k)) {
      const double v = h * (U[tx][k] + h * W[k][tx]);
      const long long o = (long long)r * dim + c;
      OV[o] = v;
      sq += v * v;
      if (AV) AV[o] = al[m] + coef * v;
    }
  }
  if (npart) {
    sq = block_sum256(sq, sh);
    if (threadIdx.x == 0) npart[inst * nstride + blockIdx.x] = sq;
  }
}
void launch_sym_mirror(hipStream_t s, const DevLayout &L, double *RES, int herm, const int *act, int nact,
                       double *acc, double coef, const double *coefp, double *npart) {
  if (L.nmirror <= 0 || nact <= 0) return;
  hipLaunchKernelGGL(k_sym_mirror, dim3(L.nmirror, nact), dim3(256), 0, s, L, L.mirror_tiles, RES, herm, act, acc,
                     coef, coefp, npart, (long long)(L.nmirror + L.nob));
}

// ------------------------------------------------------------------ one-body (111 + 121 + 221)
// (Y - pm*Y^T) at pair rows (c,x),(c,y) of block J, the 221 intermediate, with
// v_elem's ordering phase and identical-orbit sqrt2 applied to both pairs.
// Hole c: Y = W1 = L(:,pp) R(pp,:), read from the upper-triangle store Wpp of
// the 222_pp_hh GEMM ((W1 - pm W1^T)^T = -pm (W1 - pm W1^T)).  Particle c:
// Y = W2 = L(:,hh) R(hh,:), summed here (hh ranges are a few pairs long).
__device__ __forceinline__ double w221(const DevLayout &L, const double *Wpp, const double *VL,
                                       const double *VR, int c, int x, int y, int J, double pm) {
  const int jc = L.jj[c];
  const int P = (L.l[c] + L.l[x]) & 1, T2 = L.tz[c] + L.tz[x];
  const int q = 3 * J + 2 * (T2 / 2 + 1) + P;
  double pre = 1.0;
  int r, s;
  if (c > x) {
    r = dpos(L, J, x, c);
    pre = sgnp(1 + ((jc + L.jj[x] - J) >> 1));
  } else {
    if (c == x) pre = pre * M_SQRT2;
    r = dpos(L, J, c, x);
  }
  if (c > y) {
    s = dpos(L, J, y, c);
    pre = pre * sgnp(1 + ((jc + L.jj[y] - J) >> 1));
  } else {
    if (c == y) pre = pre * M_SQRT2;
    s = dpos(L, J, c, y);
  }
  if ((r | s) < 0) return 0.0;
  const long long bo = L.boff[q];
  const int dim = L.bdim[q];
  double val;
  if (L.con[c]) {
    val = (r <= s) ? Wpp[bo + (long long)r * dim + s] : -pm * Wpp[bo + (long long)s * dim + r];
  } else {
    const double *Lb = VL + bo, *Rb = VR + bo;
    val = 0.0;
    // the hh columns four at a time, reads clamped into the block (weight 0
    // past its end) so that the sixteen reads are in flight together
    for (int k0 = L.bnp[q] + L.bnb[q]; k0 < dim; k0 += 4) {
      double lr[4], rk[4], rr[4], lk[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int k = min(k0 + u, dim - 1);
        lr[u] = Lb[(long long)r * dim + k];
        rk[u] = Rb[(long long)k * dim + s];
        rr[u] = Rb[(long long)r * dim + k];
        lk[u] = Lb[(long long)k * dim + s];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (k0 + u < dim) val += lr[u] * rk[u] - rr[u] * lk[u];
    }
  }
  return val * pre;
}


// 121 + 221 sums of the one-body outputs (p,q) whose orbits share (l,j,tz)
// (the others get 111 only).  The (orbit, J) terms of one such output (first
// the nphp*nJ terms of 121, then the nsp*nJ terms of 221; MJ = nJ slots per
// orbit, the most J values two orbits can couple to) are cut into chunks of
// 256, one workgroup per chunk, each writing its (2J+1)-weighted partial sum;
// k_ob_final adds the partials in chunk order, so results do not depend on
// scheduling.
__global__ __launch_bounds__(256) void k_ob_partial(DevLayout L, const double *__restrict__ Lop,
                                                    const double *__restrict__ Rop,
                                                    const double *__restrict__ Wpp, double *__restrict__ part,
                                                    int hL, int hR, unsigned mask, const int *act) {
  __shared__ double red[4];
  const int sp = blockIdx.x / L.ob_nchunk, ch = blockIdx.x - sp * L.ob_nchunk;
  if (sp >= L.nsame) return;
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    Lop += inst * L.op_size;
    Rop += inst * L.op_size;
    Wpp += inst * L.vsize;
    part += inst * (long long)L.nsame * L.ob_nchunk;
  }
  const int tid = threadIdx.x, t = L.same_ob[sp];
  const int p = L.ob_p[t], qo = L.ob_q[t], N = L.nsp, MJ = L.nJ;
  const double *FL = Lop + L.f_off, *FR = Rop + L.f_off;
  const double *VL = Lop + L.v_off, *VR = Rop + L.v_off;
  const int n121 = (mask & T121) ? L.nphp * MJ : 0, n221 = (mask & T221) ? N * MJ : 0;
  const int it = ch * 256 + tid;
  double v = 0.0;
  if (it < n121) {
    const int e = it / MJ, m = it - e * MJ;
    const int a = L.php_a[e], i = L.php_i[e], ji = L.jj[i], jq = L.jj[qo];
    const int JT = abs(ji - jq) + 2 * m;
    if (JT <= ji + jq) {
      // the two elements' positions (the same in L and R), then the four
      // reads together (position 0 and factor 0 for a structural zero)
      double fx, fy;
      const long long ox = dv_pos(L, a, p, i, qo, JT, fx), oy = dv_pos(L, i, p, a, qo, JT, fy);
      const long long oxc = ox < 0 ? 0 : ox, oyc = oy < 0 ? 0 : oy;
      const double smx = VR[oxc] * fx, smx2 = VL[oxc] * fx, smy = VR[oyc] * fy, smy2 = VL[oyc] * fy;
      v = (JT + 1) * (FL[a * N + i] * (hL * smx - smy) - FR[a * N + i] * (hR * smx2 - smy2));
    }
  } else if (it < n121 + n221) {
    const int c = (it - n121) / MJ, m = (it - n121) - c * MJ;
    const int jc = L.jj[c], ji = L.jj[p], JT = abs(jc - ji) + 2 * m;
    if (JT <= jc + ji) v = w221(L, Wpp, VL, VR, c, p, qo, JT, (double)(hL * hR)) * (JT + 1);
  }
  v = wave_sum(v);
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}
// one wavefront per one-body output (p,q): 111 (full F products), plus the
// 121 + 221 partials / (2j_p + 1); writes F(p,q) and F(q,p) = herm F(p,q)
__global__ __launch_bounds__(64) void k_ob_final(DevLayout L, const double *__restrict__ Lop,
                                                 const double *__restrict__ Rop, const double *__restrict__ part,
                                                 double *__restrict__ RES, int hres, unsigned mask, const int *act,
                                                 double *acc, double coef, const double *coefp, double *npart) {
  const int t = blockIdx.x;
  if (t >= L.nob) return;
  const long long inst = inst_id(act);
  if (inst < 0) return;
  if (coefp) coef = *coefp;
  Lop += inst * L.op_size;
  Rop += inst * L.op_size;
  RES += inst * L.op_size;
  part += inst * (long long)L.nsame * L.ob_nchunk;
  const int lane = threadIdx.x, p = L.ob_p[t], qo = L.ob_q[t], N = L.nsp;
  const double *FL = Lop + L.f_off, *FR = Rop + L.f_off;
  double val = 0.0;
  if (mask & T111)
    for (int k = lane; k < N; k += 64) val += FL[p * N + k] * FR[k * N + qo] - FR[p * N + k] * FL[k * N + qo];
  val = wave_sum(val);
  if (lane == 0) {
    const int sp = L.ob_same[t];
    if (sp >= 0 && (mask & (T121 | T221))) {
      double sm = 0.0;
      for (int c = 0; c < L.ob_nchunk; c++) sm += part[(long long)sp * L.ob_nchunk + c];
      val += sm / (L.jj[p] + 1.0);
    }
    double *F = RES + L.f_off;
    const double vt = (p != qo) ? hres * val : 0.0;  // F(q,p)
    if (p == qo && hres == -1) val = 0.0;
    F[p * N + qo] = val;
    if (p != qo) F[qo * N + p] = vt;
    if (acc) {
      double *A = acc + inst * L.op_size;
      double *AF = A + L.f_off;
      AF[p * N + qo] = AF[p * N + qo] + coef * val;
      if (p != qo) AF[qo * N + p] = AF[qo * N + p] + coef * vt;
      if (t == 0) A[0] = A[0] + coef * RES[0];  // E0 (as add_sq_op)
    }
    if (npart) npart[inst * (long long)(L.nmirror + L.nob) + L.nmirror + t] = val * val + vt * vt;
  }
}
void launch_ob_partial(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop,
                       const double *Wpp, double *part, int hL, int hR, unsigned mask, const int *act, int nact) {
  if (L.nob <= 0 || nact <= 0 || L.nsame <= 0 || !(mask & (T121 | T221))) return;
  hipLaunchKernelGGL(k_ob_partial, dim3(L.nsame * L.ob_nchunk, nact), dim3(256), 0, s, L, Lop, Rop, Wpp, part,
                     hL, hR, mask, act);
}
void launch_ob_final(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop, const double *part,
                     double *RES, int hres, unsigned mask, const int *act, int nact, double *acc, double coef,
                     const double *coefp, double *npart) {
  if (L.nob <= 0 || nact <= 0) return;
  hipLaunchKernelGGL(k_ob_final, dim3(L.nob, nact), dim3(64), 0, s, L, Lop, Rop, part, RES, hres, mask, act, acc,
                     coef, coefp, npart);
}

// ------------------------------------------------------------------ reductions

// 220 per block (commutators.f90:399-432) into partials[q * nsplit + split]; 110 into
// partials[nblocks * nsplit] (nsplit workgroups per block, as many as kPartials allows, at most kE0Split)
__global__ __launch_bounds__(256) void k_e0_partial(DevLayout L, const double *Lop, const double *Rop,
                                                     int hL, int hR, double *partials, const int *act, int nsplit) {
  __shared__ double sh[256];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    Lop += inst * L.op_size;
    Rop += inst * L.op_size;
    partials += inst * kPartials;
  }
  const int b = blockIdx.x;
  if (b < L.nblocks * nsplit) {
    // nsplit workgroups per block, each a strided share of its pphh elements
    const int q = b / nsplit, sp = b - q * nsplit;
    const int np = L.bnp[q], nb = L.bnb[q], nh = L.bnh[q], dim = L.bdim[q];
    const double *A = Lop + L.v_off + L.boff[q], *B = Rop + L.v_off + L.boff[q];
    double s = 0.0;
    for (int e = sp * 256 + (int)threadIdx.x; e < np * nh; e += 256 * nsplit) {
      const int ix = e / nh, jx = e % nh;
      const long long o = (long long)ix * dim + np + nb + jx;
      s = s - A[o] * B[o] * hR;
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) partials[b] = s * (L.bJ[q] + 1);
  } else {
    // the (particle, hole) pairs: the last P*A entries of the one-body output list
    const int N = L.nsp, A = L.nhole, P = L.npart, t0 = A * (A + 1) / 2 + P * (P + 1) / 2;
    const double *FL = Lop + L.f_off, *FR = Rop + L.f_off;
    double s = 0.0;
    for (int e = threadIdx.x; e < P * A; e += 256) {
      const int a = L.ob_p[t0 + e], i = L.ob_q[t0 + e];
      const double lf = FL[a * N + i], rf = FR[a * N + i];
      s = s + (lf * rf * hL - rf * lf * hR) * (L.jj[i] + 1);
    }
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) partials[b] = s;
  }
}
__global__ __launch_bounds__(256) void k_e0_final(int nparts, const double *partials, double *RES,
                                                   int use110, int use220, const int *act, long long rstride) {
  __shared__ double sh[256];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    partials += inst * kPartials;
    RES += inst * rstride;
  }
  double s = 0.0;
  for (int q = threadIdx.x; q < nparts; q += 256) s += partials[q];
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) {
    double e = 0.0;
    if (use110) e = partials[nparts];
    if (use220) e = e + s * 2.0;
    RES[0] = e;
  }
}
void launch_e0(hipStream_t s, const DevLayout &L, const double *Lop, const double *Rop, double *RES,
               int hL, int hR, unsigned mask, double *partials, const int *act, int nact, long long rstride) {
  if (nact <= 0) return;
  const int nsplit = std::max(1, std::min(kE0Split, (kPartials - 1) / std::max(L.nblocks, 1)));
  hipLaunchKernelGGL(k_e0_partial, dim3(L.nblocks * nsplit + 1, nact), dim3(256), 0, s, L, Lop, Rop, hL, hR,
                     partials, act, nsplit);
  const int use220 = (mask & T220) && (hL * hR == -1);
  hipLaunchKernelGGL(k_e0_final, dim3(1, nact), dim3(256), 0, s, L.nblocks * nsplit, partials, RES,
                     (mask & T110) ? 1 : 0, use220 ? 1 : 0, act, rstride);
}

// sum of squares of F and V = mat_frob_norm^2 (basic_IMSRG.f90:3921-3952)
__global__ __launch_bounds__(256) void k_sumsq_partial(const double *x, long long n, double *partials,
                                                        const int *act, long long stride) {
  __shared__ double sh[256];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    x += inst * stride;
    partials += inst * kPartials;
  }
  // eight strided elements per round, read off any branch (clamped to the
  // last element, discarded past it): the same summation order as one at a time
  double s = 0.0;
  const long long st = (long long)gridDim.x * 256;
  for (long long i0 = (long long)blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 8 * st) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = x[min(i0 + u * st, n - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (i0 + u * st < n) s += v[u] * v[u];
  }
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_final_sum(const double *partials, int n, double *out, int do_sqrt,
                                                    const int *act, long long pstride) {
  __shared__ double sh[256];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    partials += inst * pstride;
    out += inst * kScal;
  }
  double s = 0.0;  // eight partials per round off any branch, in the one-at-a-time order
  for (int i0 = threadIdx.x; i0 < n; i0 += 8 * 256) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = partials[min(i0 + 256 * u, n - 1)];
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (i0 + 256 * u < n) s += v[u];
  }
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) out[0] = do_sqrt ? sqrt(s) : s;
}
void launch_sumsq(hipStream_t s, const DevLayout &L, const double *op, double *partials, double *out,
                  const int *act, int nact) {
  if (nact <= 0) return;
  const long long n = L.op_size - L.f_off;
  hipLaunchKernelGGL(k_sumsq_partial, dim3(kPartials, nact), dim3(256), 0, s, op + L.f_off, n, partials, act,
                     L.op_size);
  hipLaunchKernelGGL(k_final_sum, dim3(1, nact), dim3(256), 0, s, partials, kPartials, out, 1, act,
                     (long long)kPartials);
}

// the Frobenius norm from the partial sums of squares that k_sym_mirror and
// k_ob_final wrote (nmirror + nob per instance), into out[inst * kScal]
void launch_norm_from_parts(hipStream_t s, const DevLayout &L, const double *npart, double *out, const int *act,
                            int nact) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_final_sum, dim3(1, nact), dim3(256), 0, s, npart, L.nmirror + L.nob, out, 1, act,
                     (long long)(L.nmirror + L.nob));
}

// mbpt2 (basic_IMSRG.f90:4030-4064): one workgroup per block
__global__ __launch_bounds__(256) void k_mbpt2_partial(DevLayout L, const double *H, double *partials,
                                                        const int *act) {
  __shared__ double sh[256];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    H += inst * L.op_size;
    partials += inst * kPartials;
  }
  const int q = blockIdx.x;
  const int np = L.bnp[q], nb = L.bnb[q], nh = L.bnh[q], dim = L.bdim[q], base = L.pbase[q], N = L.nsp;
  const double *V = H + L.v_off + L.boff[q];
  const double *F = H + L.f_off;
  double s = 0.0;
  for (int e = threadIdx.x; e < np * nh; e += 256) {
    const int ii = e / nh, jx = e % nh;
    const int i = L.pair_a[base + ii], j = L.pair_b[base + ii];
    const int k = L.pair_a[base + np + nb + jx], l = L.pair_b[base + np + nb + jx];
    const double eden = F[k * N + k] + F[l * N + l] - F[i * N + i] - F[j * N + j];
    const double x = V[(long long)ii * dim + np + nb + jx];
    s = s + x * x / eden;
  }
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) partials[q] = (L.bJ[q] + 1.0) * s;
}
void launch_mbpt2(hipStream_t s, const DevLayout &L, const double *op, double *partials, double *out,
                  const int *act, int nact) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_mbpt2_partial, dim3(L.nblocks, nact), dim3(256), 0, s, L, op, partials, act);
  hipLaunchKernelGGL(k_final_sum, dim3(1, nact), dim3(256), 0, s, partials, L.nblocks, out, 0, act,
                     (long long)kPartials);
}

// ------------------------------------------------------------------ generators
// twobody_monopole (basic_IMSRG.f90:2031-2046) for every orbit pair
__global__ void k_monopole(DevLayout L, const double *H, double *mono, const int *act) {
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    H += inst * L.op_size;
    mono += inst * L.nsp * L.nsp;
  }
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = L.nsp;
  if (t >= N * N) return;
  const int a = t / N, b = t % N, ja = L.jj[a], jb = L.jj[b];
  const double *V = H + L.v_off;
  double sm = 0.0;
  for (int JT = abs(ja - jb); JT <= ja + jb; JT += 2) sm = sm + (JT + 1) * dv(L, V, a, b, a, b, JT);
  mono[t] = sm / (ja + 1.0) / (jb + 1.0);
}
void launch_monopole(hipStream_t s, const DevLayout &L, const double *H, double *mono, const int *act,
                     int nact) {
  if (nact <= 0) return;
  const int n = L.nsp * L.nsp;
  hipLaunchKernelGGL(k_monopole, dim3((n + 255) / 256, nact), dim3(256), 0, s, L, H, mono, act);
}

// build_gs_white / build_gs_atan / build_gs_atanxx (generators.f90:4-374):
// ETA must be zero on entry; kind 0 = H/Eden, 1 = atan(2H/Eden)/2
__global__ void k_eta_onebody(DevLayout L, const double *Hden, const double *Hnum, double *ETA, int kind,
                              const int *act) {
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    const long long o = inst * L.op_size;
    Hden += o;
    Hnum += o;
    ETA += o;
  }
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L.nphp) return;
  const int a = L.php_a[e], i = L.php_i[e], N = L.nsp, ji = L.jj[i];
  const double *V = Hden + L.v_off, *Fd = Hden + L.f_off;
  double Eden = 0.0;
  for (int JT = 0; JT <= 2 * ji; JT += 2) Eden = Eden - (JT + 1) * dv(L, V, a, i, a, i, JT);
  Eden = Eden / (ji + 1.0) / (ji + 1.0);
  Eden = Eden + Fd[a * N + a] - Fd[i * N + i];
  const double h = Hnum[L.f_off + a * N + i];
  const double eta = (kind == 0) ? h / Eden : atan(2 * h / Eden) * 0.5;
  ETA[L.f_off + a * N + i] = eta;
  ETA[L.f_off + i * N + a] = -eta;
}
__global__ __launch_bounds__(256) void k_eta_twobody(DevLayout L, const double *Hden, const double *Hnum,
                                                      const double *mono, double *ETA, int kind, const int *act) {
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    Hden += inst * L.op_size;
    Hnum += inst * L.op_size;
    ETA += inst * L.op_size;
    mono += inst * L.nsp * L.nsp;
  }
  const int q = blockIdx.x;
  const int np = L.bnp[q], nb = L.bnb[q], nh = L.bnh[q], dim = L.bdim[q], base = L.pbase[q], N = L.nsp;
  const double *Fd = Hden + L.f_off;
  const double *Vn = Hnum + L.v_off + L.boff[q];
  double *Ve = ETA + L.v_off + L.boff[q];
  for (int e = threadIdx.x; e < np * nh; e += 256) {
    const int ix = e / nh, jx = e % nh;
    const int a = L.pair_a[base + ix], b = L.pair_b[base + ix];
    const int i = L.pair_a[base + np + nb + jx], j = L.pair_b[base + np + nb + jx];
    double Eden = 0.0;
    Eden = Eden + mono[a * N + b];
    Eden = Eden + mono[i * N + j];
    Eden = Eden - mono[a * N + i];
    Eden = Eden - mono[a * N + j];
    Eden = Eden - mono[i * N + b];
    Eden = Eden - mono[j * N + b];
    Eden = Eden + Fd[a * N + a] + Fd[b * N + b] - Fd[i * N + i] - Fd[j * N + j];
    const long long o = (long long)ix * dim + np + nb + jx;
    const double h = Vn[o];
    const double eta = (kind == 0) ? h / Eden : 0.5 * atan(2 * h / Eden);
    Ve[o] = eta;
    Ve[(long long)(np + nb + jx) * dim + ix] = -eta;
  }
}
void launch_eta(hipStream_t s, const DevLayout &L, const double *Hden, const double *Hnum,
                const double *mono, double *ETA, int kind, const int *act, int nact) {
  if (nact <= 0) return;
  if (L.nphp > 0)
    hipLaunchKernelGGL(k_eta_onebody, dim3((L.nphp + 63) / 64, nact), dim3(64), 0, s, L, Hden, Hnum, ETA, kind,
                       act);
  hipLaunchKernelGGL(k_eta_twobody, dim3(L.nblocks, nact), dim3(256), 0, s, L, Hden, Hnum, mono, ETA, kind,
                     act);
}

// ------------------------------------------------------------------ normal ordering
// normal_order (basic_IMSRG.f90:2375-2506); T is an unmodified copy of H
__global__ __launch_bounds__(256) void k_no_e0(DevLayout L, const double *T, double *H, const int *act) {
  __shared__ double sh[256];
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    const long long o = inst * L.op_size;
    T += o;
    H += o;
  }
  const int N = L.nsp;
  const double *F = T + L.f_off, *V = T + L.v_off;
  double s1 = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) s1 += F[i * N + i] * L.con[i] * (L.jj[i] + 1);
  s1 = block_sum256(s1, sh);
  double s2 = 0.0;
  for (int e = threadIdx.x; e < N * N; e += 256) {
    const int i = e / N, j = e % N;
    if (L.con[i] * L.con[j] == 0) continue;
    for (int JT = abs(L.jj[i] - L.jj[j]); JT <= L.jj[i] + L.jj[j]; JT += 2)
      s2 = s2 + dv(L, V, i, j, i, j, JT) * (JT + 1);
  }
  s2 = block_sum256(s2, sh);
  if (threadIdx.x == 0) H[0] = (s1 * 2.0 + s2) * 0.5;
}
__global__ void k_no_f(DevLayout L, const double *T, double *H, int herm, const int *act) {
  {
    const long long inst = inst_id(act);
    if (inst < 0) return;
    const long long o = inst * L.op_size;
    T += o;
    H += o;
  }
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.nob) return;
  const int a = L.ob_p[t], b = L.ob_q[t], N = L.nsp;
  if (!same_ljt(L, a, b)) return;
  const double *V = T + L.v_off;
  double s2 = 0.0;
  for (int ii = 0; ii < N; ii++) {
    if (L.con[ii] != 1) continue;
    for (int JT = abs(L.jj[ii] - L.jj[a]); JT <= L.jj[ii] + L.jj[a]; JT += 2)
      s2 = s2 + (JT + 1.0) / (L.jj[a] + 1) * dv(L, V, a, ii, b, ii, JT);
  }
  const double val = T[L.f_off + a * N + b] + s2;
  H[L.f_off + a * N + b] = val;
  if (a != b) H[L.f_off + b * N + a] = herm * val;
}
void launch_normal_order(hipStream_t s, const DevLayout &L, const double *T, double *H, const int *act,
                         int nact) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_no_e0, dim3(1, nact), dim3(256), 0, s, L, T, H, act);
  if (L.nob > 0) hipLaunchKernelGGL(k_no_f, dim3((L.nob + 63) / 64, nact), dim3(64), 0, s, L, T, H, 1, act);
}

// E0 of every active instance of an operator into out[inst * kScal]
__global__ void k_gather_e0(const double *op, long long stride, double *out, const int *act, int nact) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nact) return;
  const long long inst = act ? act[k] : 0;
  if (inst < 0) return;
  out[inst * kScal] = op[inst * stride];
}
void launch_gather_e0(hipStream_t s, const double *op, long long stride, double *out, const int *act, int nact) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_gather_e0, dim3((nact + 63) / 64), dim3(64), 0, s, op, stride, out, act, nact);
}

// The truncation tests of BCH_EXPAND (IMSRG_MAGNUS.f90:229-230) and
// MAGNUS_EXPAND (IMSRG_MAGNUS.f90:656-674, 689) made on the device after each
// nested commutator, so the host can enqueue the next commutator before the
// norm is known.  Position k of the series table `act` holds instance act[k],
// or a negative number once that instance has left the series; a leaving
// instance gets -1, which every later kernel of the series skips.
//   BCH (mode 0): goes on while !(||ad^n|| * coef < conv).
//   Magnus (mode 1): advals[store] = ||ad^n|| * coef when store >= 0 (coef =
//   |cof|; advals in scalar slots kAdvSlot.., ||Omega|| in kFullSlot), then goes
//   on while !(fullnorm < tiny) and !(|advals[test] / fullnorm| < conv).
// ||ad^n|| is scalar slot 0.  live[k] (host memory) gets 1 or 0 per position.
__global__ void k_series_gate(int *act, int nact, double *scal, int mode, double coef, int store, int test,
                              double conv, double tiny, int *live) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nact) return;
  const int inst = act[k];
  int on = 0;
  if (inst >= 0) {
    double *sc = scal + (long long)inst * kScal;
    if (mode == 0) {
      on = !(sc[0] * coef < conv);
    } else {
      if (store >= 0) sc[kAdvSlot + store] = sc[0] * coef;
      const double full = sc[kFullSlot];
      on = !(full < tiny) && !(fabs(sc[kAdvSlot + test] / full) < conv);
    }
    if (!on) act[k] = -1;
  }
  live[k] = on;
  __threadfence_system();
}
void launch_series_gate(hipStream_t s, int *act, int nact, double *scal, int mode, double coef, int store,
                        int test, double conv, double tiny, int *live) {
  if (nact <= 0) return;
  hipLaunchKernelGGL(k_series_gate, dim3((nact + 63) / 64), dim3(64), 0, s, act, nact, scal, mode, coef, store,
                     test, conv, tiny, live);
}
}  // namespace imsrg
