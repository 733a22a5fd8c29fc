// hf.hip -- Hartree-Fock without three-body forces on device-resident
// operators: the self-consistent loop of calc_HF (HF_mod.f90:14-106), the
// one- and two-body transforms to the Hartree-Fock basis (transform_1b_to_HF,
// HF_mod.f90:408-491; transform_2b_to_HF, :494-619) and e_HF (:376-405).
//
// Single-particle matrices (D, rho, the gamma and Fock matrices, the bare
// one-body T) are nsp x nsp row-major with zeros outside the (l,j,tz) groups;
// D(a, i) = D[a * nsp + i] is the component of old orbit a in new orbit i,
// where the k-th lowest eigenvector of a group belongs to the group's k-th
// orbit (orbit order), as in the reference's blocks.
#include "devfun.hpp"

namespace imsrg {

// T = the bare one-body part within groups (write_kin_matrix, HF_mod.f90:
// 138-167: upper element mirrored) and, with reset, D = 1 and eigenvalues 0
__global__ void k_hf_init(DevLayout L, const double *__restrict__ op, double *__restrict__ D,
                          double *__restrict__ Tm, double *__restrict__ eig, int reset) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n_hf_bd) return;
  const int b = L.hf_b[t], d = L.hf_d[t], N = L.nsp;
  Tm[b * N + d] = op[L.f_off + (long long)min(b, d) * N + max(b, d)];
  if (reset) {
    D[b * N + d] = (b == d) ? 1.0 : 0.0;
    if (b == d) eig[b] = 0.0;
  }
}

// density_matrix (HF_mod.f90:346-373) without the (2j+1) that gamma_matrix
// divides out again: rho(b,d) = sum over hole orbits i of the group of D(b,i) D(d,i)
// state[0] != 0: the loop has converged and every kernel of a further
// iteration returns at once (the host enqueues iterations in groups)
__global__ void k_hf_density(DevLayout L, const double *__restrict__ D, double *__restrict__ rho,
                             const double *__restrict__ state) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n_hf_bd || state[0] != 0.0) return;
  const int b = L.hf_b[t], d = L.hf_d[t], N = L.nsp, g = L.grp_of[b];
  const int *mem = L.grp_list + L.grp_start[g];
  double s = 0.0;
  for (int k = 0; k < L.grp_cnt[g]; k++) {
    const int i = mem[k];
    if (L.con[i]) s += D[b * N + i] * D[d * N + i];
  }
  rho[b * N + d] = s;
}

// gamma_matrix (HF_mod.f90:169-232): for a <= c of one group,
//   gam(a,c) = 1/(ja+1) sum_{groups, b, d} rho(d,b) sum_J (2J+1) v(a,b,c,d;J)
// one workgroup per (a,c), threads over the (b,d) pairs; the Fock matrix
// T + gam is written with it (both triangles)
__global__ __launch_bounds__(256) void k_hf_gamma(DevLayout L, const double *__restrict__ op,
                                                  const double *__restrict__ rho, const double *__restrict__ Tm,
                                                  double *__restrict__ Vg, double *__restrict__ Fm,
                                                  const double *__restrict__ state) {
  __shared__ double sh[256];
  const int t = blockIdx.x;
  if (t >= L.n_hf_ac || state[0] != 0.0) return;
  const int a = L.hf_a[t], c = L.hf_c[t], ja = L.jj[a], N = L.nsp;
  const double *V = op + L.v_off;
  double s = 0.0;
  for (int e = threadIdx.x; e < L.n_hf_bd; e += 256) {
    const int b = L.hf_b[e], d = L.hf_d[e], jr = L.jj[b];
    const double r = rho[d * N + b];
    if (r == 0.0) continue;
    double sj = 0.0;
    for (int J = abs(jr - ja); J <= jr + ja; J += 2) sj += (J + 1.0) * dv(L, V, a, b, c, d, J);
    s += r * sj;
  }
  s = block_sum256(s, sh);
  if (threadIdx.x == 0) {
    const double g = s / (ja + 1.0);
    Vg[a * N + c] = Vg[c * N + a] = g;
    Fm[a * N + c] = Fm[c * N + a] = Tm[a * N + c] + g;
  }
}

// diagonalize_blocks (basic_IMSRG.f90:2163-2180) for one group per
// workgroup: the cyclic Jacobi method (Jacobi 1846; the threshold and
// rotation formulas of Rutishauser's procedure, Handbook for Automatic
// Computation II, 1971) on the group's Fock matrix in LDS, eigenvalues sorted
// ascending with their vectors.  Rotations run one after another; lane k
// updates row / column k of each.  cpart[g] = |eig_old - eig_new|_2 of the group.
__global__ __launch_bounds__(64) void k_hf_diag(DevLayout L, const double *__restrict__ Fm, double *__restrict__ D,
                                                double *__restrict__ eig, double *__restrict__ cpart,
                                                const double *__restrict__ state) {
  __shared__ double A[kHFMaxGroup][kHFMaxGroup], Q[kHFMaxGroup][kHFMaxGroup];
  __shared__ double off_s;
  const int g = blockIdx.x;
  if (g >= L.ngrp || state[0] != 0.0) return;
  const int m = L.grp_cnt[g], N = L.nsp, k = threadIdx.x;
  const int *mem = L.grp_list + L.grp_start[g];
  if (k < m)
    for (int y = 0; y < m; y++) {
      A[k][y] = Fm[mem[k] * N + mem[y]];
      Q[k][y] = (k == y) ? 1.0 : 0.0;
    }
  __syncthreads();
  for (int sweep = 0; sweep < 100; sweep++) {
    if (k == 0) {
      double off = 0.0;
      for (int p = 0; p < m; p++)
        for (int q = p + 1; q < m; q++) off += fabs(A[p][q]);
      off_s = off;
    }
    __syncthreads();
    if (off_s == 0.0) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = A[p][q], app = A[p][p], aqq = A[q][q];
        if (apq == 0.0) continue;
        const double big = 100.0 * fabs(apq);
        if (sweep > 3 && fabs(app) + big == fabs(app) && fabs(aqq) + big == fabs(aqq)) {
          __syncthreads();
          if (k == 0) A[p][q] = A[q][p] = 0.0;
          __syncthreads();
          continue;
        }
        const double h = aqq - app;
        double tn;
        if (fabs(h) + big == fabs(h)) {
          tn = apq / h;
        } else {
          const double th = 0.5 * h / apq;
          tn = 1.0 / (fabs(th) + sqrt(1.0 + th * th));
          if (th < 0.0) tn = -tn;
        }
        const double cs = 1.0 / sqrt(1.0 + tn * tn), sn = tn * cs;
        __syncthreads();  // every lane has read A[p][q], A[p][p], A[q][q]
        if (k < m) {      // columns p, q of A P
          const double u = A[k][p], v = A[k][q];
          A[k][p] = cs * u - sn * v;
          A[k][q] = sn * u + cs * v;
        }
        __syncthreads();
        if (k < m) {  // rows p, q of P^T (A P)
          const double u = A[p][k], v = A[q][k];
          A[p][k] = cs * u - sn * v;
          A[q][k] = sn * u + cs * v;
          const double w = Q[k][p], z = Q[k][q];
          Q[k][p] = cs * w - sn * z;
          Q[k][q] = sn * w + cs * z;
        }
        __syncthreads();
        if (k == 0) A[p][q] = A[q][p] = 0.0;
        __syncthreads();
      }
  }
  if (k != 0) return;
  for (int x = 0; x < m; x++) {  // ascending eigenvalues (selection sort)
    int j = x;
    for (int y = x + 1; y < m; y++)
      if (A[y][y] < A[j][j]) j = y;
    if (j != x) {
      const double w = A[x][x];
      A[x][x] = A[j][j];
      A[j][j] = w;
      for (int r = 0; r < m; r++) {
        const double u = Q[r][x];
        Q[r][x] = Q[r][j];
        Q[r][j] = u;
      }
    }
  }
  double cr = 0.0;
  for (int j = 0; j < m; j++) {
    const double dl = eig[mem[j]] - A[j][j];
    cr += dl * dl;
    eig[mem[j]] = A[j][j];
    for (int x = 0; x < m; x++) D[mem[x] * N + mem[j]] = Q[x][j];
  }
  cpart[g] = sqrt(cr);
}

// crit = sum of the groups' eigenvalue changes, in the reference's block
// order; counts the iteration (state[1]) and ends the loop (state[0] = 1)
// once crit > conv fails, as the reference's `do while (crit > 1e-6)`
__global__ void k_hf_crit(DevLayout L, const double *__restrict__ cpart, double *__restrict__ crit,
                          double *__restrict__ state, double conv) {
  if (blockIdx.x != 0 || threadIdx.x != 0 || state[0] != 0.0) return;
  double s = 0.0;
  for (int k = 0; k < L.ngrp; k++) s += cpart[L.hf_gorder[k]];
  crit[0] = s;
  state[1] += 1.0;
  if (!(s > conv)) state[0] = 1.0;
}

// eigenvectors with a negative diagonal component change sign (HF_mod.f90:84-89)
__global__ void k_hf_signfix(DevLayout L, double *__restrict__ D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.nsp) return;
  const int N = L.nsp, g = L.grp_of[i];
  if (D[i * N + i] >= 0.0) return;
  const int *mem = L.grp_list + L.grp_start[g];
  for (int x = 0; x < L.grp_cnt[g]; x++) D[mem[x] * N + i] = -D[mem[x] * N + i];
}

// one-body transform (transform_1b_to_HF, HF_mod.f90:408-491): M' = D^T M D
// per group, stored as the reference stores it: for a <= c, pp and hh take
// M'(a,c) in both orders; a particle-hole entry takes M'(particle, hole), its
// mirror herm times that.  With Tm/Vg, also the diagonals of D^T T D and
// D^T gam D that e_HF sums (tv[2 i], tv[2 i + 1]).
__global__ void k_hf_1b(DevLayout L, const double *__restrict__ D, const double *__restrict__ M,
                        double *__restrict__ op, int herm, const double *__restrict__ Tm,
                        const double *__restrict__ Vg, double *__restrict__ tv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n_hf_ac) return;
  const int a = L.hf_a[t], c = L.hf_c[t], N = L.nsp, g = L.grp_of[a];
  const int *mem = L.grp_list + L.grp_start[g];
  const int m = L.grp_cnt[g];
  auto tr = [&](const double *X, int i, int k) {  // (D^T X D)(i, k)
    double s = 0.0;
    for (int u = 0; u < m; u++) {
      double r = 0.0;
      for (int w = 0; w < m; w++) r += X[mem[u] * N + mem[w]] * D[mem[w] * N + k];
      s += D[mem[u] * N + i] * r;
    }
    return s;
  };
  double *F = op + L.f_off;
  if (L.con[a] == L.con[c]) {
    F[a * N + c] = F[c * N + a] = tr(M, a, c);
  } else if (L.con[a] == 0) {
    const double v = tr(M, a, c);
    F[a * N + c] = v;
    F[c * N + a] = herm * v;
  } else {
    const double v = tr(M, c, a);
    F[c * N + a] = v;
    F[a * N + c] = herm * v;
  }
  if (tv && a == c) {
    tv[2 * a] = tr(Tm, a, a);
    tv[2 * a + 1] = tr(Vg, a, a);
  }
}

// e_HF (HF_mod.f90:376-405): sum over hole orbits of (2j+1)(t + v/2), groups in
// the reference's order, written as E0
__global__ void k_hf_ehf(DevLayout L, const double *__restrict__ tv, double *__restrict__ op,
                         double *__restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int k = 0; k < L.ngrp; k++) {
    const int g = L.hf_gorder[k];
    for (int x = 0; x < L.grp_cnt[g]; x++) {
      const int i = L.grp_list[L.grp_start[g] + x];
      if (L.con[i]) s += (tv[2 * i] + 0.5 * tv[2 * i + 1]) * (L.jj[i] + 1.0);
    }
  }
  op[0] = s;
  out[0] = s;
}

// Element (ab, ij) of the pair-basis transform U = C - (-1)^J Crev of
// transform_2b_to_HF: [D(a,i) D(b,j) sqrt(1+d_ab) - (-1)^J (1-d_ab)
// (-1)^((ja+jb)/2) D(b,i) D(a,j)] / sqrt(1+d_ij), so that V' = U^T V U.
__device__ __forceinline__ double hf_u(const DevLayout &L, const double *D, int a, int b, int i, int j, int J) {
  const int N = L.nsp;
  double u = D[a * N + i] * D[b * N + j];
  if (a == b) {
    u *= M_SQRT2;
  } else {
    u -= (double)(sgnp(J >> 1) * sgnp((L.jj[a] + L.jj[b]) >> 1)) * D[b * N + i] * D[a * N + j];
  }
  return (i == j) ? u * .70710678118654752 : u;
}

// pass 1: W = U^T V, row by row: W(ij, s) = sum over the rows (ab) that U
// couples to (ij) of U(ab,ij) V(ab, s); one wavefront per pair row
__global__ __launch_bounds__(64) void k_hf_rows(DevLayout L, const double *__restrict__ D,
                                                const double *__restrict__ op, double *__restrict__ W, int r0,
                                                long long wofs) {
  const int r = r0 + (int)blockIdx.x;
  if (r >= L.npairs) return;
  const int q = L.pair_q[r], base = L.pbase[q], rl = r - base, dim = L.bdim[q], J = L.bJ[q];
  const int i = L.pair_a[r], j = L.pair_b[r];
  const double *VB = op + L.v_off + L.boff[q];
  double *WR = W + (L.v_off + L.boff[q] - wofs) + (long long)rl * dim;
  const int e0 = L.hf_nb_start[r], e1 = L.hf_nb_start[r + 1];
  for (int s0 = 0; s0 < dim; s0 += 256) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int e = e0; e < e1; e++) {
      const int p = L.hf_nb[e];
      const double u = hf_u(L, D, L.pair_a[base + p], L.pair_b[base + p], i, j, J);
      const double *row = VB + (long long)p * dim;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int s = s0 + 64 * k + (int)threadIdx.x;
        if (s < dim) acc[k] += u * row[s];
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int s = s0 + 64 * k + (int)threadIdx.x;
      if (s < dim) WR[s] = acc[k];
    }
  }
}

// pass 2: V' = W U on the upper triangle, V'(ij, kl) = sum over the rows (ab)
// U couples to (kl) of W(ij, ab) U(ab, kl); the lower triangle is herm times
// the upper (diagonal zero for herm = -1); one wavefront per pair row (ij)
__global__ __launch_bounds__(64) void k_hf_cols(DevLayout L, const double *__restrict__ D,
                                                const double *__restrict__ W, double *__restrict__ op, int herm,
                                                int r0, long long wofs) {
  const int r = r0 + (int)blockIdx.x;
  if (r >= L.npairs) return;
  const int q = L.pair_q[r], base = L.pbase[q], rl = r - base, dim = L.bdim[q], J = L.bJ[q];
  const double *WR = W + (L.v_off + L.boff[q] - wofs) + (long long)rl * dim;
  double *VB = op + L.v_off + L.boff[q];
  for (int s = rl + (int)threadIdx.x; s < dim; s += 64) {
    const int k = L.pair_a[base + s], l = L.pair_b[base + s];
    double acc = 0.0;
    for (int e = L.hf_nb_start[base + s]; e < L.hf_nb_start[base + s + 1]; e++) {
      const int p = L.hf_nb[e];
      acc += WR[p] * hf_u(L, D, L.pair_a[base + p], L.pair_b[base + p], k, l, J);
    }
    if (s == rl) {
      VB[(long long)rl * dim + s] = (herm == 1) ? acc : 0.0;
    } else {
      VB[(long long)rl * dim + s] = acc;
      VB[(long long)s * dim + rl] = herm * acc;
    }
  }
}

static inline unsigned nblk(int n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_hf_init(hipStream_t s, const DevLayout &L, const double *op, double *D, double *Tm, double *eig,
                    int reset) {
  hipLaunchKernelGGL(k_hf_init, dim3(nblk(L.n_hf_bd, 256)), dim3(256), 0, s, L, op, D, Tm, eig, reset);
}
void launch_hf_iteration(hipStream_t s, const DevLayout &L, const double *op, double *D, double *rho,
                         const double *Tm, double *Vg, double *Fm, double *eig, double *cpart, double *crit,
                         double *state, double conv) {
  hipLaunchKernelGGL(k_hf_density, dim3(nblk(L.n_hf_bd, 256)), dim3(256), 0, s, L, D, rho, state);
  hipLaunchKernelGGL(k_hf_gamma, dim3(L.n_hf_ac), dim3(256), 0, s, L, op, rho, Tm, Vg, Fm, state);
  hipLaunchKernelGGL(k_hf_diag, dim3(L.ngrp), dim3(64), 0, s, L, Fm, D, eig, cpart, state);
  hipLaunchKernelGGL(k_hf_crit, dim3(1), dim3(64), 0, s, L, cpart, crit, state, conv);
}
void launch_hf_finish_1b(hipStream_t s, const DevLayout &L, double *D, const double *M, double *op, int herm,
                         const double *Tm, const double *Vg, double *tv, double *ehf) {
  if (tv) hipLaunchKernelGGL(k_hf_signfix, dim3(nblk(L.nsp, 64)), dim3(64), 0, s, L, D);
  hipLaunchKernelGGL(k_hf_1b, dim3(nblk(L.n_hf_ac, 64)), dim3(64), 0, s, L, D, M, op, herm, Tm, Vg, tv);
  if (tv) hipLaunchKernelGGL(k_hf_ehf, dim3(1), dim3(64), 0, s, L, tv, op, ehf);
}
void launch_hf_2b(hipStream_t s, const DevLayout &L, const double *D, double *op, double *W, int herm) {
  if (L.npairs <= 0) return;
  hipLaunchKernelGGL(k_hf_rows, dim3(L.npairs), dim3(64), 0, s, L, D, op, W, 0, 0LL);
  hipLaunchKernelGGL(k_hf_cols, dim3(L.npairs), dim3(64), 0, s, L, D, W, op, herm, 0, 0LL);
}
void launch_hf_2b_window(hipStream_t s, const DevLayout &L, const double *D, double *op, double *W, int herm,
                         int r0, int nrows, long long wofs) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_hf_rows, dim3(nrows), dim3(64), 0, s, L, D, op, W, r0, wofs);
  hipLaunchKernelGGL(k_hf_cols, dim3(nrows), dim3(64), 0, s, L, D, W, op, herm, r0, wofs);
}

}  // namespace imsrg
