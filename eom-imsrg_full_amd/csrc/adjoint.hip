// adjoint.hip -- transpose of the IMSRG(2) commutator in its second operand.
//
// For fixed P and a cotangent C (herm hC = -hP*hY) these kernels accumulate
//   Ybar = d <C, [P, Y]> / dY
// where <A, B> is the Frobenius product over the full device layout
// (E0, the full one-body matrix, every channel block in full).  Each term is
// the exact transpose of its forward kernel in kernels.hip, term by term:
// gathers become scatters (FP64 atomics), the 222 products become products
// with the transposed operands.  The caller projects Ybar onto operators of
// hermiticity hY afterwards.  By antisymmetry of the commutator,
// d <C, [X, Y]> / dX = -Ybar(P = Y).
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "devfun.hpp"

namespace imsrg {

// cotangent of one-body output (p,q) of the ob list: the forward kernel writes
// F(p,q) = val and F(q,p) = hC*val (diagonal: val, or 0 for hC = -1)
__device__ __forceinline__ double ob_cot(const DevLayout &L, const double *CF, int p, int q, int hC) {
  const int N = L.nsp;
  if (p != q) return CF[p * N + q] + hC * CF[q * N + p];
  return hC == 1 ? CF[p * N + p] : 0.0;
}

// ---- E0 terms: 110 (commutators.f90:10-33) and 220 (:399-432)
__global__ __launch_bounds__(256) void k_adj_e0(DevLayout L, const double *__restrict__ P,
                                                const double *__restrict__ C, double *__restrict__ Y, int hP,
                                                int hY, unsigned mask) {
  const double e = C[0];
  const int b = blockIdx.x;
  if (b < L.nblocks) {
    if (!(mask & T220) || hP * hY != -1) return;
    const int np = L.bnp[b], nb = L.bnb[b], nh = L.bnh[b], dim = L.bdim[b];
    const double *A = P + L.v_off + L.boff[b];
    double *YV = Y + L.v_off + L.boff[b];
    const double w = -2.0 * (L.bJ[b] + 1) * hY * e;
    for (int t = threadIdx.x; t < np * nh; t += 256) {
      const int ix = t / nh, jx = t % nh;
      const long long o = (long long)ix * dim + np + nb + jx;
      YV[o] += w * A[o];
    }
  } else {
    if (!(mask & T110)) return;
    const int N = L.nsp;
    const double *FP = P + L.f_off;
    double *YF = Y + L.f_off;
    for (int t = threadIdx.x; t < N * N; t += 256) {
      const int a = t / N, i = t % N;
      if (L.con[a] != 0 || L.con[i] != 1) continue;
      YF[t] += e * FP[t] * (double)(hP - hY) * (L.jj[i] + 1);
    }
  }
}

// ---- one-body outputs: 111 (:36-80), 121 (:83-280), 221 (:435-578)
__global__ __launch_bounds__(256) void k_adj_onebody(DevLayout L, const double *__restrict__ P,
                                                     const double *__restrict__ C, double *__restrict__ Y,
                                                     int hP, int hY, int hC, unsigned mask) {
  const int t = blockIdx.x;
  if (t >= L.nob) return;
  const int p = L.ob_p[t], q = L.ob_q[t], N = L.nsp, MJ = L.nJ, tid = threadIdx.x;
  const double c = ob_cot(L, C + L.f_off, p, q, hC);
  if (c == 0.0) return;
  const double *FP = P + L.f_off, *VP = P + L.v_off;
  double *YF = Y + L.f_off, *YV = Y + L.v_off;
  if (mask & T111)  // val = sum_k P(p,k) Y(k,q) - Y(p,k) P(k,q)
    for (int k = tid; k < N; k += 256) {
      atomicAdd(YF + k * N + q, c * FP[p * N + k]);
      atomicAdd(YF + p * N + k, -c * FP[k * N + q]);
    }
  if (!same_ljt(L, p, q)) return;
  const double w0 = c / (L.jj[p] + 1.0);
  if (mask & T121) {
    const int jq = L.jj[q];
    for (int it = tid; it < L.nphp * MJ; it += 256) {
      const int e = it / MJ, m = it - e * MJ;
      const int a = L.php_a[e], i = L.php_i[e], ji = L.jj[i];
      const int JT = abs(ji - jq) + 2 * m;
      if (JT > ji + jq) continue;
      const double g = w0 * (JT + 1);
      // + FP(a,i) (hP v_Y(a,p,i,q) - v_Y(i,p,a,q))
      double pre;
      long long o = dv_pos(L, a, p, i, q, JT, pre);
      if (o >= 0) atomicAdd(YV + o, g * FP[a * N + i] * hP * pre);
      o = dv_pos(L, i, p, a, q, JT, pre);
      if (o >= 0) atomicAdd(YV + o, -g * FP[a * N + i] * pre);
      // - FY(a,i) (hY v_P(a,p,i,q) - v_P(i,p,a,q))
      atomicAdd(YF + a * N + i, -g * (hY * dv(L, VP, a, p, i, q, JT) - dv(L, VP, i, p, a, q, JT)));
    }
  }
  if (mask & T221) {
    // val += sum_c,J (J+1)/(2j_p+1) pre * sum_{k} (P(r,k) Y(k,s) - Y(r,k) P(k,s)),
    // r = (c,p), s = (c,q); k over pp pairs for a hole c, hh pairs for a particle c
    const int ji = L.jj[p];
    for (int it = tid; it < N * MJ; it += 256) {
      const int cc = it / MJ, m = it - cc * MJ;
      const int jc = L.jj[cc], JT = abs(jc - ji) + 2 * m;
      if (JT > jc + ji) continue;
      const int Pq = (L.l[cc] + L.l[p]) & 1, T2 = L.tz[cc] + L.tz[p];
      const int qb = 3 * JT + 2 * (T2 / 2 + 1) + Pq;
      double pre = 1.0;
      int r, s;
      if (cc > p) {
        r = dpos(L, JT, p, cc);
        pre = sgnp(1 + ((jc + L.jj[p] - JT) >> 1));
      } else {
        if (cc == p) pre = pre * M_SQRT2;
        r = dpos(L, JT, cc, p);
      }
      if (cc > q) {
        s = dpos(L, JT, q, cc);
        pre = pre * sgnp(1 + ((jc + L.jj[q] - JT) >> 1));
      } else {
        if (cc == q) pre = pre * M_SQRT2;
        s = dpos(L, JT, cc, q);
      }
      if ((r | s) < 0) continue;
      const long long bo = L.boff[qb];
      const int dim = L.bdim[qb];
      const int k0 = L.con[cc] ? 0 : L.bnp[qb] + L.bnb[qb];
      const int k1 = L.con[cc] ? L.bnp[qb] : dim;
      const double g = w0 * (JT + 1) * pre;
      const double *Pb = VP + bo;
      double *Yb = YV + bo;
      for (int k = k0; k < k1; k++) {
        atomicAdd(Yb + (long long)k * dim + s, g * Pb[(long long)r * dim + k]);
        atomicAdd(Yb + (long long)r * dim + k, -g * Pb[(long long)k * dim + s]);
      }
    }
  }
}

// ---- 122: RES = M + hC M^T, M(r,s) = sum_{e in nb(r)} c_e [fP(x,i) Y(r_e,s) - fY(x,i) P(r_e,s)].
// With C of hermiticity hC the cotangent of M is Mbar = C + hC C^T = 2C (diagonal (1+hC) C).
__global__ __launch_bounds__(64) void k_adj_122(DevLayout L, const double *__restrict__ P,
                                                const double *__restrict__ C, double *__restrict__ Y, int hC) {
  const int r = blockIdx.x;
  if (r >= L.npairs) return;
  const int q = L.pair_q[r], rl = r - L.pbase[q], dim = L.bdim[q], N = L.nsp;
  const long long bo = L.v_off + L.boff[q];
  const double *PV = P + bo, *CV = C + bo + (long long)rl * dim;
  double *YV = Y + bo;
  const double *FP = P + L.f_off;
  double *YF = Y + L.f_off;
  for (int e = L.nb_start[r]; e < L.nb_start[r + 1]; e++) {
    const long long pe = L.nb_pos[e];
    const int xi = L.nb_x[e] * N + L.nb_i[e];
    const double ce = L.nb_coef[e], w = ce * FP[xi];
    double dot = 0.0;
    for (int s = threadIdx.x; s < dim; s += 64) {
      const double mb = (s == rl) ? (1.0 + hC) * CV[s] : 2.0 * CV[s];
      if (w != 0.0) atomicAdd(YV + pe * dim + s, w * mb);
      dot += PV[pe * dim + s] * mb;
    }
    dot = wave_sum(dot);
    if (threadIdx.x == 0 && dot != 0.0) atomicAdd(YF + xi, -ce * dot);
  }
}

// ---- 222_pp_hh: C = P D Y - Y D P  =>  Ybar(r,s) += hP d_r Q(r,s) - hC d_s Q(s,r), Q = P C
__global__ __launch_bounds__(64) void k_adj_pphh(DevLayout L, const double *__restrict__ Q, double *__restrict__ Y,
                                                 int hP, int hC) {
  const int r = blockIdx.x;
  if (r >= L.npairs) return;
  const int q = L.pair_q[r], rl = r - L.pbase[q], dim = L.bdim[q], np = L.bnp[q], nb = L.bnb[q];
  const double *Qb = Q + L.boff[q];
  double *Yb = Y + L.v_off + L.boff[q] + (long long)rl * dim;
  const double dr = rl < np ? 1.0 : (rl >= np + nb ? -1.0 : 0.0);
  for (int s = threadIdx.x; s < dim; s += 64) {
    const double ds = s < np ? 1.0 : (s >= np + nb ? -1.0 : 0.0);
    Yb[s] += hP * dr * Qb[(long long)rl * dim + s] - hC * ds * Qb[(long long)s * dim + rl];
  }
}

// ---- 222_ph, step 1: cotangents of the J' terms (transpose of k_ph_apply)
__global__ __launch_bounds__(64) void k_adj_ph_apply(DevLayout L, const double *__restrict__ C,
                                                     double *__restrict__ Tbar, int hC) {
  const int r = blockIdx.x;
  if (r >= L.npairs) return;
  const int q = L.pair_q[r], base = L.pbase[q], rl = r - base, dim = L.bdim[q], J = L.bJ[q];
  const int i = L.pair_a[r], j = L.pair_b[r];
  const int ji = L.jj[i], jj = L.jj[j], li = L.l[i], lj = L.l[j], ti = L.tz[i], tj = L.tz[j];
  const double pre = (i == j) ? .70710678118 : 1.0;
  const double *CV = C + L.v_off + L.boff[q] + (long long)rl * dim;
  const unsigned *toff = L.ph_toff + L.ph_ubase[r] - rl;
  for (int s = rl + (int)threadIdx.x; s < dim; s += 64) {
    // the forward writes the upper element and its mirror herm * value
    const double ob = (s == rl) ? ((hC == 1) ? CV[s] : 0.0) : 2.0 * CV[s];
    if (ob == 0.0) continue;
    const int k = L.pair_a[base + s], l = L.pair_b[base + s];
    const int jk = L.jj[k], jl = L.jj[l], lk = L.l[k], ll = L.l[l], tk = L.tz[k], tl = L.tz[l];
    double *T1 = Tbar + toff[s];
    const int a1 = max(abs(jj - jl), abs(ji - jk)), b1 = min(jj + jl, ji + jk);
    const int a2 = max(abs(ji - jl), abs(jj - jk)), b2 = min(ji + jl, jj + jk);
    double *T2 = T1 + (b1 >= a1 ? (b1 - a1) / 2 + 1 : 0);
    const double w = ob * pre * ((k == l) ? .70710678118 : 1.0);
    if (abs(tl - tj) == abs(ti - tk) && (ll + lj) % 2 == (li + lk) % 2) {
      const double w1 = -(double)sgnp((jl + jk + J) >> 1) * w;
      for (int JP = a1, m = 0; JP <= b1; JP += 2, m++) atomicAdd(T1 + m, w1 * dsixj(L, ji, jj, J, jl, jk, JP));
    }
    if (abs(tk - tj) == abs(ti - tl) && (lk + lj) % 2 == (li + ll) % 2)
      for (int JP = a2, m = 0; JP <= b2; JP += 2, m++) atomicAdd(T2 + m, w * dsixj(L, ji, jj, J, jk, jl, JP));
  }
}

// ---- 222_ph, step 2: cotangent of Z (transpose of k_ph_terms)
__global__ __launch_bounds__(256) void k_adj_ph_terms(DevLayout L, const double *__restrict__ Tbar,
                                                      double *__restrict__ Zbar, double sigma) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= L.nphterm) return;
  const double v = Tbar[t];
  if (v == 0.0) return;
  const unsigned tc = L.ph_terms[t];
  if (tc == 0xFFFFFFFFu) return;  // padding term
  const unsigned code = L.ph_quads[tc >> 6];
  const int side = (tc >> 5) & 1, m = tc & 31;
  const int i = code >> 24, j = (code >> 16) & 255, k = (code >> 8) & 255, l = code & 255;
  const int N = L.nsp, JTM1 = L.jtot_max + 1;
  const int ji = L.jj[i], jj = L.jj[j], jk = L.jj[k], jl = L.jj[l];
  const double ph = sgnp((ji + jj + jk + jl) >> 1);
  if (side == 0) {
    const int Tz = abs(L.tz[i] - L.tz[k]) / 2, P = (L.l[i] + L.l[k]) % 2;
    const int JP = max(abs(jj - jl), abs(ji - jk)) + 2 * m;
    const int qx = (JP >> 1) + Tz * JTM1 + 2 * P * JTM1;
    const int *rowt = L.ccrow + (size_t)qx * N * N;
    const int gik = rowt[k * N + i], rjl = rowt[j * N + l], rik = rowt[i * N + k], gjl = rowt[l * N + j];
    if ((gik | rjl | rik | gjl) >= 0) {
      const long long rl = L.ccrlen[qx];
      double *Zq = Zbar + L.cwoff[qx];
      atomicAdd(Zq + gik * rl + rjl, v);
      atomicAdd(Zq + rik * rl + gjl, ph * sigma * v);
    }
  } else {
    const int Tz = abs(L.tz[i] - L.tz[l]) / 2, P = (L.l[i] + L.l[l]) % 2;
    const int JP = max(abs(ji - jl), abs(jj - jk)) + 2 * m;
    const int qx = (JP >> 1) + Tz * JTM1 + 2 * P * JTM1;
    const int *rowt = L.ccrow + (size_t)qx * N * N;
    const int rjk = rowt[j * N + k], gil = rowt[l * N + i], gjk = rowt[k * N + j], ril = rowt[i * N + l];
    if ((rjk | gil | gjk | ril) >= 0) {
      const long long rl = L.ccrlen[qx];
      double *Zq = Zbar + L.cwoff[qx];
      atomicAdd(Zq + rjk * rl + gil, sigma * v);
      atomicAdd(Zq + gjk * rl + ril, ph * v);
    }
  }
}

// ---- 222_ph, step 3: Sbar^T = hP Zbar - hY Zbar^T in place, one thread per pair (a, b), a <= b
__global__ __launch_bounds__(256) void k_adj_ph_sbt(DevLayout L, double *__restrict__ Z, int qx, int hP, int hY) {
  const long long r = L.ccrlen[qx];
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= r * r) return;
  const long long a = t / r, b = t % r;
  if (a > b) return;
  double *Zq = Z + L.cwoff[qx];
  const double x = Zq[a * r + b], y = Zq[b * r + a];
  Zq[a * r + b] = hP * x - hY * y;
  if (a != b) Zq[b * r + a] = hP * y - hY * x;
}

// ---- 222_ph, step 5: transpose of the Pandya transform of Y (cross_coupled.f90:1190-1300)
__global__ __launch_bounds__(256) void k_adj_pandya(DevLayout L, const double *__restrict__ ccb,
                                                    double *__restrict__ Y) {
  const int col = blockIdx.x;
  if (col >= L.nccol) return;
  const int q1 = L.cccol_q[col];
  const int c = col - L.cccbase[q1];
  const int p = L.cccol_p[col], h = L.cccol_h[col];
  const int JC = L.ccJ[q1], rlen = L.ccrlen[q1];
  const int jp = L.jj[p], jh = L.jj[h], lp = L.l[p], lh = L.l[h], tp = L.tz[p], th = L.tz[h];
  double *YV = Y + L.v_off;
  const double *in = ccb + L.ccoff[q1] + (long long)c * rlen;
  const double sq = sqrt(JC + 1.0);
  for (int r = threadIdx.x; r < rlen; r += blockDim.x) {
    const double g = in[r];
    const int a = L.ccrow_a[L.ccrbase[q1] + r], b = L.ccrow_b[L.ccrbase[q1] + r];
    if (g == 0.0 || a < 0) continue;
    const int ja = L.jj[a], jb = L.jj[b];
    if (!((((L.l[a] + lh) & 1) == ((L.l[b] + lp) & 1)) && (L.tz[a] + th == L.tz[b] + tp))) continue;
    const double g0 = g * sgnp((jh + jb + JC) >> 1) * sq;
    const int Jmin = max(abs(jp - jb), abs(ja - jh)), Jmax = min(jp + jb, ja + jh);
    for (int JT = Jmin; JT <= Jmax; JT += 2) {
      double pre;
      const long long o = dv_pos(L, h, a, p, b, JT, pre);
      if (o < 0) continue;
      atomicAdd(YV + o, g0 * (double)(sgnp(JT >> 1) * (JT + 1)) * dsixj(L, jp, jh, JC, ja, jb, JT) * pre);
    }
  }
}

// ---- projection onto hermiticity h: X <- (X + h X^T)/2 on F and every block;
// E0 <- 0 unless keep_e0
__global__ __launch_bounds__(256) void k_project(DevLayout L, const int4 *__restrict__ tiles, double *X, int h,
                                                 int keep_e0) {
  if (blockIdx.x == L.nmirror) {  // one-body part and E0
    const int N = L.nsp;
    double *F = X + L.f_off;
    for (int t = threadIdx.x; t < N * N; t += 256) {
      const int a = t / N, b = t % N;
      if (a < b) {
        const double u = F[a * N + b], v = F[b * N + a];
        F[a * N + b] = 0.5 * (u + h * v);
        F[b * N + a] = 0.5 * (v + h * u);
      } else if (a == b && h == -1) {
        F[t] = 0.0;
      }
    }
    if (threadIdx.x == 0 && !keep_e0) X[0] = 0.0;
    return;
  }
  const int4 t = tiles[blockIdx.x];
  const int q = t.x, dim = L.bdim[q];
  double *OV = X + L.v_off + L.boff[q];
  __shared__ double U[32][33], W[32][33];
  const int r0 = t.y * 32, c0 = t.z * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;
    U[k][tx] = (r < dim && c < dim) ? OV[(long long)r * dim + c] : 0.0;
    const int r2 = c0 + k, c2 = r0 + tx;
    W[k][tx] = (r2 < dim && c2 < dim) ? OV[(long long)r2 * dim + c2] : 0.0;
  }
  __syncthreads();
  const bool diag = t.y == t.z;
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + tx;  // upper element (r, c); its mirror is W[tx][k]
    if (r < dim && c < dim && (!diag || k <= tx)) {
      const double v = 0.5 * (U[k][tx] + h * W[tx][k]);
      OV[(long long)r * dim + c] = (diag && k == tx && h == -1) ? 0.0 : v;
      if (!diag || k < tx) OV[(long long)c * dim + r] = h * v;
    }
  }
}

// ---- plain FP64 GEMM C = A * B (A M x K row-major, B K x N row-major), all tiles
__global__ __launch_bounds__(256) void k_gemm_plain(const GemmProb *__restrict__ probs, const int4 *__restrict__ tiles,
                                                    int ntiles, const double *__restrict__ A,
                                                    const double *__restrict__ B, double *__restrict__ C) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int4 tl = tiles[t];
  const GemmProb p = probs[tl.x];
  const int m0 = tl.y * 64, n0 = tl.z * 64;
  __shared__ double As[16][65], Bs[16][65];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
  double4_t acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const double *Ap = A + p.a_off, *Bp = B + p.b_off;
  for (int k0 = 0; k0 < p.K; k0 += 16) {
    for (int e = tid; e < 16 * 64; e += 256) {
      const int kk = e & 15, mm = e >> 4;  // A rows: 16 consecutive k of row m0 + mm
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < p.M && k < p.K) ? Ap[(long long)m * p.lda + k] : 0.0;
      const int kb = e >> 6, nn = e & 63;
      const int n = n0 + nn, k2 = k0 + kb;
      Bs[kb][nn] = (n < p.N && k2 < p.K) ? Bp[(long long)k2 * p.ldb + n] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < 4; kq++) {
      const int k = kq * 4 + (lane >> 4);
      const double a0 = As[k][wm * 32 + (lane & 15)], a1 = As[k][wm * 32 + 16 + (lane & 15)];
      const double b0 = Bs[k][wn * 32 + (lane & 15)], b1 = Bs[k][wn * 32 + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  double *Cp = C + p.c_off;
  const int g = lane >> 4;
  for (int ti = 0; ti < 2; ti++)
    for (int tj = 0; tj < 2; tj++)
      for (int v = 0; v < 4; v++) {  // v_mfma_f64_16x16x4: row lane/16 + 4v, column lane%16
        const int row = m0 + wm * 32 + ti * 16 + g + 4 * v, col = n0 + wn * 32 + tj * 16 + (lane & 15);
        if (row < p.M && col < p.N) Cp[(long long)row * p.ldc + col] = acc[ti][tj][v];
      }
}


// ---- generator transpose: build_gs_white / build_gs_atan (generators.f90:4-106, 163-266)
// eta = g(h, Eden) with g = h/E (kind 0) or atan(2h/E)/2 (kind 1), written at
// (a,i) and (i,a) = -eta; with the cotangent ebar of that pair the pieces are
//   Hbar(h) += ebar dg/dh,  Ebar = ebar dg/dE  -> f diagonals and V elements of E.
__device__ __forceinline__ void gen_derivs(double h, double E, int kind, double &dh, double &dE) {
  if (kind == 0) {
    dh = 1.0 / E;
    dE = -h / (E * E);
  } else {
    const double x = 2.0 * h / E, w = 1.0 / (1.0 + x * x);
    dh = w / E;
    dE = -w * h / (E * E);
  }
}
__global__ void k_adj_eta_onebody(DevLayout L, const double *__restrict__ H, const double *__restrict__ Eb,
                                  double *__restrict__ Hb, int kind) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= L.nphp) return;
  const int a = L.php_a[e], i = L.php_i[e], N = L.nsp, ji = L.jj[i];
  const double *V = H + L.v_off, *F = H + L.f_off;
  double Eden = 0.0;
  for (int JT = 0; JT <= 2 * ji; JT += 2) Eden = Eden - (JT + 1) * dv(L, V, a, i, a, i, JT);
  Eden = Eden / (ji + 1.0) / (ji + 1.0);
  Eden = Eden + F[a * N + a] - F[i * N + i];
  const double h = F[a * N + i];
  const double eb = Eb[L.f_off + a * N + i] - Eb[L.f_off + i * N + a];
  if (eb == 0.0) return;
  double dh, dE;
  gen_derivs(h, Eden, kind, dh, dE);
  const double EbarD = eb * dE;
  double *HF = Hb + L.f_off, *HV = Hb + L.v_off;
  atomicAdd(HF + a * N + i, eb * dh);
  atomicAdd(HF + a * N + a, EbarD);
  atomicAdd(HF + i * N + i, -EbarD);
  for (int JT = 0; JT <= 2 * ji; JT += 2) {
    double pre;
    const long long o = dv_pos(L, a, i, a, i, JT, pre);
    if (o >= 0) atomicAdd(HV + o, -EbarD * (JT + 1) / ((ji + 1.0) * (ji + 1.0)) * pre);
  }
}
// two-body part: Ebar of every pphh element goes to the f diagonals and, through
// the J-averaged monopoles, into mbar (nsp x nsp) for k_adj_monopole
__global__ __launch_bounds__(256) void k_adj_eta_twobody(DevLayout L, const double *__restrict__ H,
                                                         const double *__restrict__ mono,
                                                         const double *__restrict__ Eb, double *__restrict__ Hb,
                                                         double *__restrict__ mbar, int kind) {
  const int q = blockIdx.x;
  const int np = L.bnp[q], nb = L.bnb[q], nh = L.bnh[q], dim = L.bdim[q], base = L.pbase[q], N = L.nsp;
  const double *F = H + L.f_off;
  const double *Vh = H + L.v_off + L.boff[q];
  const double *Vb = Eb + L.v_off + L.boff[q];
  double *HV = Hb + L.v_off + L.boff[q], *HF = Hb + L.f_off;
  for (int e = threadIdx.x; e < np * nh; e += 256) {
    const int ix = e / nh, jx = e % nh;
    const int a = L.pair_a[base + ix], b = L.pair_b[base + ix];
    const int i = L.pair_a[base + np + nb + jx], j = L.pair_b[base + np + nb + jx];
    const long long o = (long long)ix * dim + np + nb + jx, oT = (long long)(np + nb + jx) * dim + ix;
    const double eb = Vb[o] - Vb[oT];
    if (eb == 0.0) continue;
    const double Eden = mono[a * N + b] + mono[i * N + j] - mono[a * N + i] - mono[a * N + j] - mono[i * N + b] -
                        mono[j * N + b] + F[a * N + a] + F[b * N + b] - F[i * N + i] - F[j * N + j];
    double dh, dE;
    gen_derivs(Vh[o], Eden, kind, dh, dE);
    const double EbarD = eb * dE;
    HV[o] += eb * dh;  // each pphh element is written by one thread
    atomicAdd(HF + a * N + a, EbarD);
    atomicAdd(HF + b * N + b, EbarD);
    atomicAdd(HF + i * N + i, -EbarD);
    atomicAdd(HF + j * N + j, -EbarD);
    atomicAdd(mbar + a * N + b, EbarD);
    atomicAdd(mbar + i * N + j, EbarD);
    atomicAdd(mbar + a * N + i, -EbarD);
    atomicAdd(mbar + a * N + j, -EbarD);
    atomicAdd(mbar + i * N + b, -EbarD);
    atomicAdd(mbar + j * N + b, -EbarD);
  }
}
// mono(a,b) = sum_J (2J+1) v(ab,ab;J) / ((2ja+1)(2jb+1))  (twobody_monopole, basic_IMSRG.f90:2031-2046)
__global__ void k_adj_monopole(DevLayout L, const double *__restrict__ mbar, double *__restrict__ Hb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = L.nsp;
  if (t >= N * N) return;
  const double m = mbar[t];
  if (m == 0.0) return;
  const int a = t / N, b = t % N, ja = L.jj[a], jb = L.jj[b];
  for (int JT = abs(ja - jb); JT <= ja + jb; JT += 2) {
    double pre;
    const long long o = dv_pos(L, a, b, a, b, JT, pre);
    if (o >= 0) atomicAdd(Hb + L.v_off + o, m * (JT + 1) / ((ja + 1.0) * (jb + 1.0)) * pre);
  }
}

// ---- normal-ordering transpose (normal_order, basic_IMSRG.f90:2375-2506)
// forward: E0 = sum_h (2j_h+1) f_hh + 1/2 sum_{hh',J} (2J+1) v(hh',hh';J);
// f(a,b) += sum_{h,J} (2J+1)/(2j_a+1) v(ah,bh;J) for the (a,b) of the one-body
// list with equal (l,j,tz) (f(b,a) mirrored); V unchanged.  Hb starts as a
// copy of the output cotangent Cb.
__global__ void k_adj_no_f(DevLayout L, const double *__restrict__ Cb, double *__restrict__ Hb, int herm) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.nob) return;
  const int a = L.ob_p[t], b = L.ob_q[t], N = L.nsp;
  if (!same_ljt(L, a, b)) return;
  const double *CF = Cb + L.f_off;
  double *HF = Hb + L.f_off;
  // output (a,b) = T(a,b) + sum; output (b,a) = herm * output(a,b) overwrites T(b,a)
  const double c = (a != b) ? CF[a * N + b] + herm * CF[b * N + a] : CF[a * N + a];
  HF[a * N + b] = c;
  if (a != b) HF[b * N + a] = 0.0;
  for (int ii = 0; ii < N; ii++) {
    if (L.con[ii] != 1) continue;
    for (int JT = abs(L.jj[ii] - L.jj[a]); JT <= L.jj[ii] + L.jj[a]; JT += 2) {
      double pre;
      const long long o = dv_pos(L, a, ii, b, ii, JT, pre);
      if (o >= 0) atomicAdd(Hb + L.v_off + o, c * (JT + 1.0) / (L.jj[a] + 1) * pre);
    }
  }
}
__global__ __launch_bounds__(256) void k_adj_no_e0(DevLayout L, const double *__restrict__ Cb, double *__restrict__ Hb) {
  const double e = Cb[0];
  const int N = L.nsp;
  if (e != 0.0)
    for (int t = threadIdx.x; t < N * N; t += 256) {
      const int i = t / N, j = t % N;
      if (L.con[i] * L.con[j] == 0) continue;
      if (i == j) atomicAdd(Hb + L.f_off + i * N + i, e * (L.jj[i] + 1));
      for (int JT = abs(L.jj[i] - L.jj[j]); JT <= L.jj[i] + L.jj[j]; JT += 2) {
        double pre;
        const long long o = dv_pos(L, i, j, i, j, JT, pre);
        if (o >= 0) atomicAdd(Hb + L.v_off + o, 0.5 * e * (JT + 1) * pre);
      }
    }
}

void launch_adj_generator(hipStream_t s, const DevLayout &L, const double *H, const double *mono,
                          const double *Eb, double *Hb, double *mbar, int kind) {
  launch_fill(s, mbar, (long long)L.nsp * L.nsp, 0.0);
  if (L.nphp > 0)
    hipLaunchKernelGGL(k_adj_eta_onebody, dim3((L.nphp + 63) / 64), dim3(64), 0, s, L, H, Eb, Hb, kind);
  hipLaunchKernelGGL(k_adj_eta_twobody, dim3(L.nblocks), dim3(256), 0, s, L, H, mono, Eb, Hb, mbar, kind);
  const int n = L.nsp * L.nsp;
  hipLaunchKernelGGL(k_adj_monopole, dim3((n + 255) / 256), dim3(256), 0, s, L, mbar, Hb);
}
void launch_adj_normal_order(hipStream_t s, const DevLayout &L, const double *Cb, double *Hb, int herm) {
  if (L.nob > 0) hipLaunchKernelGGL(k_adj_no_f, dim3((L.nob + 63) / 64), dim3(64), 0, s, L, Cb, Hb, herm);
  hipLaunchKernelGGL(k_adj_no_e0, dim3(1), dim3(256), 0, s, L, Cb, Hb);
}

// ---------------------------------------------------------------- launchers
void launch_adj_e0(hipStream_t s, const DevLayout &L, const double *P, const double *C, double *Y, int hP, int hY,
                   unsigned mask) {
  hipLaunchKernelGGL(k_adj_e0, dim3(L.nblocks + 1), dim3(256), 0, s, L, P, C, Y, hP, hY, mask);
}
void launch_adj_onebody(hipStream_t s, const DevLayout &L, const double *P, const double *C, double *Y, int hP,
                        int hY, int hC, unsigned mask) {
  if (L.nob > 0) hipLaunchKernelGGL(k_adj_onebody, dim3(L.nob), dim3(256), 0, s, L, P, C, Y, hP, hY, hC, mask);
}
void launch_adj_122(hipStream_t s, const DevLayout &L, const double *P, const double *C, double *Y, int hC) {
  if (L.npairs > 0) hipLaunchKernelGGL(k_adj_122, dim3(L.npairs), dim3(64), 0, s, L, P, C, Y, hC);
}
void launch_adj_pphh(hipStream_t s, const DevLayout &L, const double *Q, double *Y, int hP, int hC) {
  if (L.npairs > 0) hipLaunchKernelGGL(k_adj_pphh, dim3(L.npairs), dim3(64), 0, s, L, Q, Y, hP, hC);
}
void launch_adj_ph_apply(hipStream_t s, const DevLayout &L, const double *C, double *Tbar, int hC) {
  if (L.npairs > 0 && L.nphterm > 0) hipLaunchKernelGGL(k_adj_ph_apply, dim3(L.npairs), dim3(64), 0, s, L, C, Tbar, hC);
}
void launch_adj_ph_terms(hipStream_t s, const DevLayout &L, const double *Tbar, double *Zbar, double sigma) {
  if (L.nphterm <= 0) return;
  hipLaunchKernelGGL(k_adj_ph_terms, dim3((unsigned)((L.nphterm + 255) / 256)), dim3(256), 0, s, L, Tbar, Zbar,
                     sigma);
}
void launch_adj_ph_sbt(hipStream_t s, const DevLayout &L, const int *ccrlen_host, double *Z, int hP, int hY) {
  for (int q = 0; q < L.ncc; q++) {
    const long long r = ccrlen_host[q];
    if (r == 0) continue;
    hipLaunchKernelGGL(k_adj_ph_sbt, dim3((unsigned)((r * r + 255) / 256)), dim3(256), 0, s, L, Z, q, hP, hY);
  }
}
void launch_adj_pandya(hipStream_t s, const DevLayout &L, const double *ccb, double *Y) {
  if (L.nccol > 0) hipLaunchKernelGGL(k_adj_pandya, dim3(L.nccol), dim3(256), 0, s, L, ccb, Y);
}
void launch_project(hipStream_t s, const DevLayout &L, double *X, int h, int keep_e0) {
  hipLaunchKernelGGL(k_project, dim3(L.nmirror + 1), dim3(256), 0, s, L, L.mirror_tiles, X, h, keep_e0);
}
void launch_gemm_plain(hipStream_t s, const GemmProb *probs, const int4 *tiles, int ntiles, const double *A,
                       const double *B, double *C) {
  if (ntiles > 0) hipLaunchKernelGGL(k_gemm_plain, dim3(ntiles), dim3(256), 0, s, probs, tiles, ntiles, A, B, C);
}

}  // namespace imsrg
