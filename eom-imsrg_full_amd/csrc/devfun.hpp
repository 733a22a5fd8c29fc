// devfun.hpp -- device helpers shared by the kernel files: instance and XCD
// indexing, angular-momentum phases, the 6j table, v_elem on the full layout.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"

namespace imsrg {

typedef double double4_t __attribute__((ext_vector_type(4)));
// two doubles at any 8-byte boundary: one 16-byte access (global_load/
// store_dwordx4; the hardware takes them unaligned), half the vector memory
// instructions of two 8-byte accesses
typedef double dpair_t __attribute__((ext_vector_type(2), aligned(8)));

#define HIP_KERNEL_CHECK() ((void)0)

// Z of one Pandya block (M x M, M even) is stored in 2x2 blocks: block
// (R/2, C/2) at ((R/2)*(M/2) + C/2)*4, and inside it the diagonal pair
// first, Z(2u,2w), Z(2u+1,2w+1), then Z(2u,2w+1), Z(2u+1,2w).  Rows 2u, 2u+1
// are the pairs (a,b), (b,a) (or two (a,a) pairs), so the two Z elements
// Z(R,C), Z(R^1,C^1) of one ph J' term are one aligned 16-byte half.
// Only blocks u <= v are stored (Z(R,C) = sigma Z(C,R) gives the rest).
__device__ __forceinline__ long long zidx(int R, int C, int M) {
  return ((long long)(R >> 1) * (M >> 1) + (C >> 1)) * 4 + (((R ^ C) & 1) << 1) + (R & 1);
}

// Instance of a batched launch: blockIdx.y indexes the active-instance table
// (null table: one instance, number 0).  Instance k of an operator, Pandya
// matrix or workspace starts k strides after the buffer base.  A negative
// entry marks an instance that has left a BCH or Magnus series (the series
// gate kernels write it): every kernel returns at once for such a workgroup
// row, before any barrier.
__device__ __forceinline__ long long inst_id(const int *act) { return act ? act[blockIdx.y] : 0; }

__device__ __forceinline__ int xcd_item(const XcdSplit &x) {
  const int xcd = blockIdx.x & 7, it = x.start[xcd] + (int)(blockIdx.x >> 3);
  return it < x.start[xcd + 1] ? it : -1;
}

__device__ __forceinline__ int sgnp(int e) { return (e & 1) ? -1 : 1; }  // (-1)^e
__device__ __forceinline__ bool dtri(int a, int b, int J) { return J <= a + b && J >= abs(a - b); }
__device__ __forceinline__ int dpos(const DevLayout &L, int J, int a, int b) {
  return L.pidx[(size_t)(J >> 1) * L.nsp * L.nsp + (size_t)a * L.nsp + b];
}
__device__ __forceinline__ double dsixj(const DevLayout &L, int j1, int j2, int j3, int j4, int j5,
                                        int j6) {
  // table holds 0 for any failed triangle, matching sixj (basic_IMSRG.f90:2509-2555)
  const int h = L.nhalf, nJ = L.nJ;
  size_t idx = (((((size_t)(j1 >> 1) * h + (j2 >> 1)) * nJ + (j3 >> 1)) * h + (j4 >> 1)) * h +
                (j5 >> 1)) * nJ + (j6 >> 1);
  return L.sixj[idx];
}

// v_elem (basic_IMSRG.f90:1270-1385) on the full layout; V = operator + v_off
__device__ inline double dv(const DevLayout &L, const double *V, int a, int b, int c, int d, int J) {
  const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
  if (!(dtri(ja, jb, J) && dtri(jc, jd, J))) return 0.0;
  const int P = (L.l[a] + L.l[b]) & 1;
  if (((L.l[c] + L.l[d]) & 1) != P) return 0.0;
  const int T2 = L.tz[a] + L.tz[b];
  if (L.tz[c] + L.tz[d] != T2) return 0.0;
  const int q = 3 * J + 2 * (T2 / 2 + 1) + P;
  double pre = 1.0;
  int r, s;
  if (a > b) {
    r = dpos(L, J, b, a);
    pre = sgnp(1 + ((ja + jb - J) >> 1));
  } else {
    if (a == b) pre = pre * M_SQRT2;
    r = dpos(L, J, a, b);
  }
  if (c > d) {
    s = dpos(L, J, d, c);
    pre = pre * sgnp(1 + ((jc + jd - J) >> 1));
  } else {
    if (c == d) pre = pre * M_SQRT2;
    s = dpos(L, J, c, d);
  }
  if ((r | s) < 0) return 0.0;  // unreachable for valid tables; never read out of bounds
  return V[L.boff[q] + (long long)r * L.bdim[q] + s] * pre;
}

// position of v_elem(a,b,c,d;J) inside the V region of an operator (-1 if it
// is structurally zero) and the factor v_elem applies to the stored value
__device__ inline long long dv_pos(const DevLayout &L, int a, int b, int c, int d, int J, double &pre) {
  const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
  pre = 0.0;
  if (!(dtri(ja, jb, J) && dtri(jc, jd, J))) return -1;
  const int P = (L.l[a] + L.l[b]) & 1;
  if (((L.l[c] + L.l[d]) & 1) != P) return -1;
  const int T2 = L.tz[a] + L.tz[b];
  if (L.tz[c] + L.tz[d] != T2) return -1;
  const int q = 3 * J + 2 * (T2 / 2 + 1) + P;
  double f = 1.0;
  int r, s;
  if (a > b) {
    r = dpos(L, J, b, a);
    f = sgnp(1 + ((ja + jb - J) >> 1));
  } else {
    if (a == b) f = f * M_SQRT2;
    r = dpos(L, J, a, b);
  }
  if (c > d) {
    s = dpos(L, J, d, c);
    f = f * sgnp(1 + ((jc + jd - J) >> 1));
  } else {
    if (c == d) f = f * M_SQRT2;
    s = dpos(L, J, c, d);
  }
  if ((r | s) < 0) return -1;
  pre = f;
  return L.boff[q] + (long long)r * L.bdim[q] + s;
}

__device__ __forceinline__ bool same_ljt(const DevLayout &L, int a, int b) {
  return L.jj[a] == L.jj[b] && L.l[a] == L.l[b] && L.tz[a] == L.tz[b];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline double block_sum256(double v, double *sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) sh[tid] += sh[tid + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace imsrg
