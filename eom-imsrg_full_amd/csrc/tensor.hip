// tensor.hip -- spherical tensor operators (rank > 0) and the scalar-tensor
// commutator [G, T] with its BCH series on the device: allocate_tensor and
// tensor_elem (basic_IMSRG.f90:644-907, 1504-1641), TS_commutator_111 ..
// _222_ph (TS_commutators.f90:10-1849), the tensor Pandya transform
// (calculate_single_pandya, cross_coupled.f90:1416-1634) and BCH_TENSOR
// (IMSRG_MAGNUS.f90:389-465).
//
// Layout.  A tensor operator of rank R (doubled) and parity change dpar is one
// FP64 buffer [E0 | pad][F: nsp x nsp row-major][V]: F holds f_tensor_elem(a, b)
// for every orbit pair (the (hole, particle) entries are the herm-phased
// (particle, hole) ones), V holds, per tensor block (J1 <= J2, Tz, P1) in the
// reference's order, the FULL n1 x n2 row-major matrix between the scalar
// channels (J1, Tz, P1) and (J2, Tz, P1 + dpar/2): rows and columns are the
// channels' pair lists [pp | ph | hh] of the scalar layout, so the nine
// reference sub-blocks tgam(1..9) tile it and every 222_pp_hh product is a
// plain GEMM against the scalar full channel blocks of G.
//
// Kernels.  One thread per output element for the gathers (111, 121, 211, 221,
// 122, 212, tensor Pandya, ph recoupling), FP64 MFMA for the products (the
// pp/hh ladders and the two Pandya products of each cross-coupled block).  The
// ph recoupling is a gather: every stored output element sums, over the
// sixteen index patterns through which TS_commutator_222_ph scatters into it
// and over the cross-coupled (J3, J4) blocks, the 9j-weighted product entries
// that pass the reference's |W| > 1e-6 test.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <utility>

#include "devfun.hpp"
#include "engine_internal.hpp"

using namespace imsrg;

namespace {

// ---------------------------------------------------------------- device view
struct DevTensor {
  int rank, dpar, nblk, ncc, nJ, nh, hx, nhole, npart;
  long long nelem, nx, nr;      // tensor V elements; CCX / CCR elements
  const int *holes, *parts;
  const int *J1, *J2, *q1, *q2, *n1, *n2;
  const long long *off, *tstart;  // per tensor block: V offset; first flat element
  const int *bidx;                // [((J1/2)*nJ + J2/2)*6 + (Tz+1)*2 + P1] -> block (J1 <= J2) or -1
  const int *cJ3, *cJ4, *cq3, *cq4;
  const long long *xoff, *roff, *woff, *xstart, *rstart;
  const int *cidx;                // [((J3/2)*nJ + J4/2)*4 + Tz*2 + P] -> cross-coupled block or -1
  const int *ccol;                // [(scalar Pandya block * nsp + p) * nsp + h] -> column of (p,h)
  const double *x6, *s6;
  const double *n9;               // 9j table of the ph recoupling and the tensor Pandya (k_fill9j)
  int nd;                         // rank/2 + 1: J4 - J3 and J2 - J1 steps the table holds
  const unsigned *quads;          // orbit quadruples x0|x1<<8|x2<<16|x3<<24 (x0 <= x1, x2 <= x3) with outputs
  long long nquad;
  int nslotmax;                   // most gather slots of any quadruple (k_t222ph_q's LDS list)
  // the quadruples in four runs by slot count (at most 256, 384, 512,
  // nslotmax), one launch each with an LDS list of that size: run k starts
  // at qrun[k]
  long long qrun[5];
  int qcap[4];
  // the orbits i the tensor's one-body part couples to orbit x (same tz, l + dpar/2
  // of the same parity, triangle (ji, jx, R)): part[pstart[x] .. pstart[x+1]),
  // each i | ji << 8
  const int *pstart, *part;
  // thread orders of k_tpandya inside each scalar Pandya block: columns by
  // (2j of p, 2j of h), rows by the 2j of their two orbits (zero rows last),
  // so that the lanes of a wavefront sum (J3', J4') runs of equal length;
  // entry i is the block-local column / row thread i computes
  const int *cperm, *rperm;
  // thread order of the 122 and 212 gathers inside each tensor block's rows:
  // the ket pairs of every scalar block by the (l,j,tz) groups of their two
  // orbits (the lengths of their partner loops); entry pbase[q] + i is the
  // column thread i of a row computes
  const int *sperm;
  long long f_off, v_off;
};

// A product C (M x N, row-major, ldc) = sum of nseg segments alpha * A * B
// with A(m,k) = X[asrc][a_off + m*a_rs + k*a_cs], B(k,n) = X[bsrc][b_off + k*b_rs + n*b_cs].
struct TSeg {
  long long a_off, b_off;
  int K, a_rs, a_cs, b_rs, b_cs, asrc, bsrc;
  double alpha;
};
struct TProb {
  int M, N, ldc, nseg;
  long long c_off;
  TSeg seg[2];
};


// {J1 J2 R; j4 j5 j6} with J1, J2 integer (doubled), j4 up to jmax + R, j5, j6 of the basis
__device__ __forceinline__ double tx6(const DevTensor &T, int J1, int J2, int j4, int j5, int j6) {
  return T.x6[((((size_t)(J1 >> 1) * T.nJ + (J2 >> 1)) * T.hx + (j4 >> 1)) * T.nh + (j5 >> 1)) * T.nh + (j6 >> 1)];
}
// {c d J2; b x J4}: c, d, b of the basis, x up to jmax + R
__device__ __forceinline__ double ts6(const DevTensor &T, int c, int d, int J2, int b, int x, int J4) {
  return T.s6[(((((size_t)(c >> 1) * T.nh + (d >> 1)) * T.nJ + (J2 >> 1)) * T.nh + (b >> 1)) * T.hx + (x >> 1)) *
                  T.nJ + (J4 >> 1)];
}
// ninej (basic_IMSRG.f90:4290-4308): {a b J1; c d J2; J3 J4 R} from three 6j per x
__device__ inline double t9j(const DevTensor &T, int a, int b, int J1, int c, int d, int J2, int J3, int J4) {
  const int R = T.rank;
  const int xmin = max(max(abs(J2 - b), abs(a - R)), abs(J4 - c));
  const int xmax = min(min(J2 + b, a + R), J4 + c);
  double sm = 0.0;
  for (int x = xmin; x <= xmax; x += 2)
    sm -= (x + 1.0) * tx6(T, J1, J2, x, a, b) * ts6(T, c, d, J2, b, x, J4) * tx6(T, J3, J4, x, a, c);
  return sm;
}

// The 9j symbols of the ph recoupling, {j1 j2 J3; j3 j4 J4; J1 J2 R} with
// J3 <= J4 <= J3 + R and J1 <= J2 <= J1 + R, tabulated once per tensor space:
// index (((((((k1*nh + k2)*nh + k3)*nh + k4)*nJ + J3/2)*nd + (J4-J3)/2)*nJ + J1/2)*nd + (J2-J1)/2
__device__ __forceinline__ long long n9_index(const DevTensor &T, int j1, int j2, int j3, int j4, int J3, int J4,
                                              int J1, int J2) {
  const int h = T.nh;
  return ((((((((long long)(j1 >> 1) * h + (j2 >> 1)) * h + (j3 >> 1)) * h + (j4 >> 1)) * T.nJ + (J3 >> 1)) * T.nd +
            ((J4 - J3) >> 1)) * T.nJ + (J1 >> 1)) * T.nd + ((J2 - J1) >> 1));
}
// {a b J3; c d J4; J3p J4p R} with J3 <= J4 <= J3 + R (a cross-coupled block) and any (J3p, J4p):
// from the table, through the exchange of the first two columns (a factor (-1)^(sum of all nine))
// when J4p < J3p
__device__ __forceinline__ double t9j_cc(const DevTensor &T, int a, int b, int J3, int c, int d, int J4, int J3p,
                                         int J4p) {
  if (J4p - J3p > T.rank || J3p - J4p > T.rank) return 0.0;  // no triangle (J3p, J4p, R)
  if (J3p <= J4p) return T.n9[n9_index(T, a, b, c, d, J3, J4, J3p, J4p)];
  const double v = T.n9[n9_index(T, b, a, d, c, J3, J4, J4p, J3p)];
  return ((a + b + J3 + c + d + J4 + J3p + J4p + T.rank) >> 1) & 1 ? -v : v;
}
__global__ void k_fill9j(DevTensor T, double *__restrict__ n9, long long n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  long long u = e;
  const int d2 = (int)(u % T.nd); u /= T.nd;
  const int J1 = 2 * (int)(u % T.nJ); u /= T.nJ;
  const int d4 = (int)(u % T.nd); u /= T.nd;
  const int J3 = 2 * (int)(u % T.nJ); u /= T.nJ;
  const int k4 = (int)(u % T.nh); u /= T.nh;
  const int k3 = (int)(u % T.nh); u /= T.nh;
  const int k2 = (int)(u % T.nh); u /= T.nh;
  const int k1 = (int)u;
  const int J4 = J3 + 2 * d4, J2 = J1 + 2 * d2, JM2 = 2 * (T.nJ - 1);
  n9[e] = (J4 <= JM2 && J2 <= JM2) ? t9j(T, 2 * k1 + 1, 2 * k2 + 1, J3, 2 * k3 + 1, 2 * k4 + 1, J4, J1, J2) : 0.0;
}

// tensor_elem (basic_IMSRG.f90:1504-1641): position in V of T^{J1x J2x}_{ax bx, cx dx}
// and the factor applied to the stored value (switch phase, exchange phase, sqrt 2)
__device__ inline long long t_pos(const DevLayout &L, const DevTensor &T, int herm, int ax, int bx, int cx, int dx,
                                  int J1x, int J2x, double &pre) {
  pre = 0.0;
  int a = ax, b = bx, c = cx, d = dx, J1 = J1x, J2 = J2x;
  bool sw = false;
  if (J1x > J2x) {
    a = cx; b = dx; c = ax; d = bx; J1 = J2x; J2 = J1x;
    sw = true;
  }
  if (!dtri(J1, J2, T.rank)) return -1;
  const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
  if (!(dtri(ja, jb, J1) && dtri(jc, jd, J2))) return -1;
  const int P = (L.l[a] + L.l[b]) & 1;
  if (((L.l[c] + L.l[d]) & 1) != (P ^ ((T.dpar >> 1) & 1))) return -1;
  const int T2 = L.tz[a] + L.tz[b];
  if (L.tz[c] + L.tz[d] != T2) return -1;
  const int q = T.bidx[((J1 >> 1) * T.nJ + (J2 >> 1)) * 6 + (T2 / 2 + 1) * 2 + P];
  if (q < 0) return -1;
  double f = sw ? (double)(sgnp((J2 - J1) >> 1) * herm) : 1.0;
  int r, s;
  if (a > b) {
    r = dpos(L, J1, b, a);
    f *= sgnp(1 + ((ja + jb - J1) >> 1));
  } else {
    if (a == b) f *= M_SQRT2;
    r = dpos(L, J1, a, b);
  }
  if (c > d) {
    s = dpos(L, J2, d, c);
    f *= sgnp(1 + ((jc + jd - J2) >> 1));
  } else {
    if (c == d) f *= M_SQRT2;
    s = dpos(L, J2, c, d);
  }
  if ((r | s) < 0) return -1;
  pre = f;
  return T.off[q] + (long long)r * T.n2[q] + s;
}
// t_pos with every table read made (indices in range whatever the tests
// say) and no early return: an element the tensor does not hold gets
// position 0 and factor 0.  For batches whose loads should be in flight
// together.  chan: the caller's test that the pairs' parities (with the
// tensor's) and Tz match; j*x the orbits' 2j.
__device__ inline long long t_pos_all(const DevLayout &L, const DevTensor &T, int herm, int ax, int bx, int cx, int dx,
                                      int jax, int jbx, int jcx, int jdx, int J1x, int J2x, bool chan, double &pre) {
  const bool sw = J1x > J2x;
  const int a = sw ? cx : ax, b = sw ? dx : bx, c = sw ? ax : cx, d = sw ? bx : dx;
  const int ja = sw ? jcx : jax, jb = sw ? jdx : jbx, jc = sw ? jax : jcx, jd = sw ? jbx : jdx;
  const int J1 = sw ? J2x : J1x, J2 = sw ? J1x : J2x;
  const int P = (L.l[a] + L.l[b]) & 1, T2 = L.tz[a] + L.tz[b];
  const int q = T.bidx[((J1 >> 1) * T.nJ + (J2 >> 1)) * 6 + (T2 / 2 + 1) * 2 + P];
  const int r = dpos(L, J1, min(a, b), max(a, b)), s = dpos(L, J2, min(c, d), max(c, d));
  const int qc = max(q, 0);
  const long long o = T.off[qc];
  const int n2 = T.n2[qc];
  double f = sw ? (double)(sgnp((J2 - J1) >> 1) * herm) : 1.0;
  f *= a > b ? (double)sgnp(1 + ((ja + jb - J1) >> 1)) : (a == b ? M_SQRT2 : 1.0);
  f *= c > d ? (double)sgnp(1 + ((jc + jd - J2) >> 1)) : (c == d ? M_SQRT2 : 1.0);
  const bool ok = chan && dtri(J1, J2, T.rank) && dtri(ja, jb, J1) && dtri(jc, jd, J2) && q >= 0 && (r | s) >= 0;
  pre = ok ? f : 0.0;
  return ok ? o + (long long)r * n2 + s : 0;
}
__device__ __forceinline__ double t_elem(const DevLayout &L, const DevTensor &T, const double *V, int herm, int a,
                                         int b, int c, int d, int J1, int J2) {
  double f;
  const long long p = t_pos(L, T, herm, a, b, c, d, J1, J2, f);
  return p < 0 ? 0.0 : V[p] * f;
}

// flat element e of a prefix table (n + 1 entries) -> item
__device__ __forceinline__ int find_item(const long long *start, int n, long long e) {
  int lo = 0, hi = n;  // start[lo] <= e < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------- one-body terms
// TS_commutator_111 (TS_commutators.f90:10-55): overwrites fhh, fph, fpp of RES.
// LF: the scalar operator's full one-body matrix, RF the tensor's.
__global__ void k_t111(DevLayout L, DevTensor T, const double *__restrict__ LF, const double *__restrict__ RF,
                       double *__restrict__ F, int hL, int hR, int hres) {
  const int A = T.nhole, P = T.npart, n = L.nsp;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)A * A + (long long)P * A + (long long)P * P) return;
  auto lf = [&](int a, int b) { return LF[(size_t)a * n + b]; };
  auto rf = [&](int a, int b) { return RF[(size_t)a * n + b]; };
  if (t < (long long)A * A) {
    const int i = T.holes[t / A], j = T.holes[t % A];
    // th1 = Lhh Rhh, th2 = Lph^T Rph at (i,j) and (j,i)
    double th1 = 0, th2 = 0, th1t = 0, th2t = 0;
    for (int k = 0; k < A; k++) {
      const int hk = T.holes[k];
      th1 += lf(i, hk) * rf(hk, j);
      th1t += lf(j, hk) * rf(hk, i);
    }
    for (int k = 0; k < P; k++) {
      const int p = T.parts[k];
      th2 += lf(p, i) * rf(p, j);
      th2t += lf(p, j) * rf(p, i);
    }
    const double ph = sgnp((L.jj[i] - L.jj[j]) >> 1);
    F[(size_t)i * n + j] = th1 + th2 * hL - (th1t * hL + th2t) * hR * ph;
  } else if (t < (long long)A * A + (long long)P * A) {
    const long long u = t - (long long)A * A;
    const int p = T.parts[u / A], h = T.holes[u % A];
    double sm = 0;
    for (int k = 0; k < A; k++) {
      const int hk = T.holes[k];
      sm += lf(p, hk) * rf(hk, h);
    }
    for (int k = 0; k < P; k++) {
      const int q = T.parts[k];
      sm += lf(p, q) * rf(q, h);
    }
    double sm2 = 0;
    for (int k = 0; k < A; k++) {
      const int hk = T.holes[k];
      sm2 += rf(p, hk) * lf(hk, h);
    }
    for (int k = 0; k < P; k++) {
      const int q = T.parts[k];
      sm2 += rf(p, q) * lf(q, h);
    }
    const double v = sm - sm2;
    F[(size_t)p * n + h] = v;
    F[(size_t)h * n + p] = v * hres * sgnp((L.jj[h] - L.jj[p]) >> 1);
  } else {
    const long long u = t - (long long)A * A - (long long)P * A;
    const int p = T.parts[u / P], q = T.parts[u % P];
    // tp1 = Lph Rph^T, tp2 = Lpp Rpp at (p,q) and (q,p)
    double tp1 = 0, tp1t = 0, tp2 = 0, tp2t = 0;
    for (int k = 0; k < A; k++) {
      const int h = T.holes[k];
      tp1 += lf(p, h) * rf(q, h);
      tp1t += lf(q, h) * rf(p, h);
    }
    for (int k = 0; k < P; k++) {
      const int r = T.parts[k];
      tp2 += lf(p, r) * rf(r, q);
      tp2t += lf(q, r) * rf(r, p);
    }
    const double ph = sgnp((L.jj[p] - L.jj[q]) >> 1);
    F[(size_t)p * n + q] = tp1 * hR * ph + tp2 - (tp1t + tp2t * hR * ph) * hL;
  }
}

// TS_commutator_121 (kind 0, :58-303), _211 (kind 1, :307-552) and _221
// (kind 2, :809-1067): per upper output (p,q) allowed by the tensor's
// selection rules, F(p,q) += sum, then the lower entry becomes
// herm * (-1)^((jp-jq)/2) F(p,q) (diagonal included), exactly as each
// reference routine leaves it; (particle, hole) outputs also refresh the
// derived (hole, particle) entry.  One wavefront per output: the lanes split
// the intermediate orbits and reduce.
__global__ __launch_bounds__(64) void k_t_onebody(DevLayout L, DevTensor T, int kind, const double *__restrict__ LF,
                                                  const double *__restrict__ RF, const double *__restrict__ RV,
                                                  const double *__restrict__ gcc, const double *__restrict__ W1,
                                                  const double *__restrict__ W2, double *__restrict__ F, int hL,
                                                  int hR, int hres) {
  const int t = blockIdx.x, lane = threadIdx.x;
  if (t >= L.nob) return;
  const int p = L.ob_p[t], q = L.ob_q[t], n = L.nsp, R = T.rank;
  const int jp = L.jj[p], jq = L.jj[q], lp = L.l[p], lq = L.l[q];
  if ((lq & 1) != ((lp + (T.dpar >> 1)) & 1)) return;
  if (L.tz[q] != L.tz[p]) return;
  if (!dtri(jq, jp, R)) return;
  double sm = 0.0;
  if (kind == 0) {
    for (int u = lane; u < L.nphp; u += 64) {
      const int a = L.php_a[u], i = L.php_i[u];
      const int ja = L.jj[a], ji = L.jj[i];
      double smx = 0.0, smy = 0.0;
      for (int J2 = abs(ji - jq); J2 <= ji + jq; J2 += 2)
        for (int J1 = abs(ja - jp); J1 <= ja + jp; J1 += 2)
          smx += t_elem(L, T, RV, hR, a, p, i, q, J1, J2) * sqrt((J1 + 1.0) * (J2 + 1.0)) * sgnp(J1 >> 1) *
                 tx6(T, J1, J2, jq, jp, ji);
      for (int J1 = abs(ji - jp); J1 <= ji + jp; J1 += 2)
        for (int J2 = abs(ja - jq); J2 <= ja + jq; J2 += 2)
          smy += t_elem(L, T, RV, hR, i, p, a, q, J1, J2) * sqrt((J1 + 1.0) * (J2 + 1.0)) * sgnp(J1 >> 1) *
                 tx6(T, J1, J2, jq, jp, ji);
      sm += LF[(size_t)a * n + i] * (hL * smx - smy) * sgnp((R + jq + ji) >> 1);
    }
  } else if (kind == 1) {
    const int PAR = (lp + lq) & 1;
    const int qx = (R >> 1) + 2 * PAR * (L.jtot_max + 1);  // Pandya block (J = R, Tz = 0, PAR)
    const int rpq = L.ccrow[((size_t)qx * n + p) * n + q], rqp = L.ccrow[((size_t)qx * n + q) * n + p];
    const long long base = L.ccoff[qx];
    const int rlen = L.ccrlen[qx];
    const double inv = 1.0 / sqrt(R + 1.0);
    const double c1 = sgnp((jp - jq + R) >> 1), c2 = (double)(hR * hL * sgnp(R >> 1));
    for (int u = lane; u < T.npart * T.nhole; u += 64) {
      const int a = T.parts[u / T.nhole], i = T.holes[u % T.nhole];
      const int ja = L.jj[a], la = L.l[a], ji = L.jj[i], li = L.l[i];
      if ((li & 1) != ((la + (T.dpar >> 1)) & 1)) continue;
      if (L.tz[i] != L.tz[a]) continue;
      if (!dtri(ja, ji, R)) continue;
      if (((la + li) & 1) != PAR) continue;
      const int col = T.ccol[((size_t)qx * n + a) * n + i];
      if (col < 0 || rpq < 0 || rqp < 0) continue;  // unreachable for allowed outputs
      sm += RF[(size_t)a * n + i] *
            (c1 * gcc[base + (long long)col * rlen + rpq] - c2 * gcc[base + (long long)col * rlen + rqp]) * inv;
    }
  } else {
    const int pm = hL * hR, i = p, j = q, ji = jp, jj = jq;
    for (int u = lane; u < n; u += 64) {  // c over holes with W1, over particles with W2
      const int c = u < T.nhole ? T.holes[u] : T.parts[u - T.nhole], jc = L.jj[c];
      const double *W = u < T.nhole ? W1 : W2;
      double s1 = 0.0;
      // J2 over its whole range, four at a time with every load off any
      // branch: J2 < J1 reads W(cj, ci; J2 J1) (the lower half, by herm), J2 >= J1
      // W(ci, cj; J1 J2).  The pairs' channel matches the output's, which
      // passed the tensor's parity and Tz tests.
      const int J2lo = abs(jc - jj), n2 = (jc + jj - J2lo) / 2 + 1;
      for (int J1 = abs(jc - ji); J1 <= jc + ji; J1 += 2) {
        for (int k0 = 0; k0 < n2; k0 += 4) {
          long long pos[4];
          double w[4];
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int J2 = J2lo + 2 * min(k0 + v, n2 - 1);
            const bool lo = J2 < J1;
            double f;
            pos[v] = lo ? t_pos_all(L, T, hR, c, j, c, i, jc, jj, jc, ji, J2, J1, true, f)
                        : t_pos_all(L, T, hR, c, i, c, j, jc, ji, jc, jj, J1, J2, true, f);
            const double co = sqrt((J1 + 1.0) * (J2 + 1.0)) * tx6(T, J1, J2, jj, ji, jc);
            w[v] = k0 + v < n2 ? (lo ? -co * sgnp(J2 >> 1) * pm : co * sgnp(J1 >> 1)) * f : 0.0;
          }
          double g[4];
#pragma unroll
          for (int v = 0; v < 4; v++) g[v] = W[pos[v]];
#pragma unroll
          for (int v = 0; v < 4; v++) s1 += w[v] * g[v];
        }
      }
      sm += s1 * sgnp((jc + jj) >> 1);
    }
    sm *= sgnp(R >> 1);
  }
  sm = wave_sum(sm);
  if (lane != 0) return;
  const size_t pq = (size_t)p * n + q, qp = (size_t)q * n + p;
  const double v = F[pq] + sm;
  F[pq] = v;
  if (L.con[p] == L.con[q])
    F[qp] = v * hres * sgnp((jp - jq) >> 1);
  else
    F[qp] = v * hres * sgnp((jq - jp) >> 1);  // (hole, particle): f_tensor_elem's phase
}

// ---------------------------------------------------------------- two-body gathers
// flat element e -> block q, orbits, position; with perm, the columns of a
// row in the sperm order (r, s: row and column in the block)
__device__ __forceinline__ bool elem_of(const DevLayout &L, const DevTensor &T, long long e, int &q, int &a, int &b,
                                        int &c, int &d, long long &vpos, bool perm = false, int *rs = nullptr) {
  if (e >= T.nelem) return false;
  q = find_item(T.tstart, T.nblk, e);
  const long long u = e - T.tstart[q];
  const int n2 = T.n2[q];
  const int r = (int)(u / n2), s0 = (int)(u % n2), s = perm ? T.sperm[L.pbase[T.q2[q]] + s0] : s0;
  if (rs) {
    rs[0] = r;
    rs[1] = s;
  }
  const int g1 = L.pbase[T.q1[q]] + r, g2 = L.pbase[T.q2[q]] + s;
  a = L.pair_a[g1]; b = L.pair_b[g1]; c = L.pair_a[g2]; d = L.pair_b[g2];
  vpos = T.off[q] + (long long)r * n2 + s;
  return true;
}

// TS_commutator_122 (TS_commutators.f90:555-663): overwrites every element.
// Per output (ab|cd; J1 J2) four sums over the orbits i of the (l,j,tz) group
// of a, b, c, d: +LF(a,i) T(ib,cd), +LF(b,i) T(ai,cd), -LF(i,c) T(ab,id),
// -LF(i,d) T(ab,ci).  Each T element lies in the output's own tensor block
// with its column (kinds 0, 1) or row (kinds 2, 3) fixed, so it costs one
// pair lookup and one load; the partners go four at a time with every load
// off any branch (clamped to a valid entry, weight 0 past the group), as in
// k_t212.
__global__ void k_t122(DevLayout L, DevTensor T, const double *__restrict__ LF, const double *__restrict__ RV,
                       double *__restrict__ OV, int hR) {
  (void)hR;  // every element read lies in the output's block (J1 <= J2): no hermitian swap
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int q, a, b, c, d, rs[2];
  long long vp;
  if (!elem_of(L, T, e, q, a, b, c, d, vp, true, rs)) return;
  const int J1 = T.J1[q], J2 = T.J2[q], n = L.nsp, n2 = T.n2[q];
  const int r1 = rs[0], s2 = rs[1];  // row of (ab), column of (cd) in the block
  const double *Tq = RV + T.off[q];
  const int *P1 = L.pidx + (size_t)(J1 >> 1) * n * n, *P2 = L.pidx + (size_t)(J2 >> 1) * n * n;
  const double rab = a == b ? M_SQRT2 : 1.0, rcd = c == d ? M_SQRT2 : 1.0;
  double sm = 0.0;
  // kind 0: x = a, pair (i,b) at J1; 1: x = b, pair (a,i); 2: x = c, pair (i,d) at J2; 3: x = d, pair (c,i)
#pragma unroll
  for (int kind = 0; kind < 4; kind++) {
    const int x = kind == 0 ? a : kind == 1 ? b : kind == 2 ? c : d;
    const int o = kind == 0 ? b : kind == 1 ? a : kind == 2 ? d : c;  // the pair's other orbit
    const int jo = L.jj[o], J = kind < 2 ? J1 : J2, g = L.grp_of[x];
    const int k0 = L.grp_start[g], cnt = L.grp_cnt[g], ji = L.jj[x];  // the group shares x's 2j
    for (int k = 0; k < cnt; k += 4) {
      int pi[4];
#pragma unroll
      for (int v = 0; v < 4; v++) pi[v] = L.grp_list[k0 + min(k + v, cnt - 1)];
      long long pos[4];
      double w[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = pi[v], lo = min(i, o), hi = max(i, o);
        const int rr = (kind < 2 ? P1 : P2)[lo * n + hi];
        // i first in kinds 0 and 2, second in 1 and 3: the stored pair is (lo, hi)
        const bool swp = (kind == 0 || kind == 2) ? i > o : o > i;
        const double f = swp ? (double)sgnp(1 + ((ji + jo - J) >> 1)) : (i == o ? M_SQRT2 : 1.0);
        const double lf = kind == 0 ? LF[(size_t)a * n + i] : kind == 1 ? LF[(size_t)b * n + i]
                        : kind == 2 ? LF[(size_t)i * n + c] : LF[(size_t)i * n + d];
        const bool ok = k + v < cnt && rr >= 0 && dtri(ji, jo, J);
        pos[v] = !ok ? 0 : kind < 2 ? (long long)rr * n2 + s2 : (long long)r1 * n2 + rr;
        w[v] = ok ? f * lf * (kind < 2 ? rcd : rab) : 0.0;
      }
      double t[4];
#pragma unroll
      for (int v = 0; v < 4; v++) t[v] = Tq[pos[v]];
#pragma unroll
      for (int v = 0; v < 4; v++) sm += (kind < 2 ? w[v] : -w[v]) * t[v];
    }
  }
  sm = sm / sqrt(1.0 + (a == b)) / sqrt(1.0 + (c == d));
  OV[vp] = sm;
}

// TS_commutator_212 (TS_commutators.f90:666-806): accumulates into every element.
// GV: the scalar operator's V region (v_elem), RF: the tensor's one-body matrix.
// Per output (ab|cd; J1 J2) four sums over the orbits i the tensor's one-body
// part couples to a, b, d, c (the partner lists):
//   -{J1 J2 R; ji ja jb} RF(a,i) v(ib,cd;J2) (-1)^((ja+jb-J2)/2)
//   +{J1 J2 R; ji jb ja} RF(b,i) v(ia,cd;J2) (-1)^((J1+J2)/2)
//   +{J1 J2 R; jd ji jc} RF(i,d) v(ab,ci;J1) (-1)^((jc+jd-J1)/2)
//   -{J1 J2 R; jc ji jd} RF(i,c) v(ab,di;J1) (-1)^((J1+J2)/2)
// Each v is an element of the scalar channel block of (cd) at J2 or of (ab) at
// J1 whose fixed pair is the output's own column or row, so it costs one pair
// lookup and one load; the partners go four at a time with every load off any
// branch (clamped to a valid entry, weight 0 past the list), so a batch makes
// three dependent round trips instead of three per partner.
__global__ void k_t212(DevLayout L, DevTensor T, const double *__restrict__ GV, const double *__restrict__ RF,
                       double *__restrict__ OV) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int q, a, b, c, d, rs[2];
  long long vp;
  if (!elem_of(L, T, e, q, a, b, c, d, vp, true, rs)) return;
  const int J1 = T.J1[q], J2 = T.J2[q], n = L.nsp, R = T.rank;
  const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
  const int q1 = T.q1[q], q2 = T.q2[q];
  const int r1 = rs[0], s2 = rs[1];  // row of (ab) in block q1, column of (cd) in q2
  const double *V1 = GV + L.boff[q1], *V2 = GV + L.boff[q2];
  const int d1 = L.bdim[q1], d2 = L.bdim[q2];
  const int *P1 = L.pidx + (size_t)(J1 >> 1) * n * n, *P2 = L.pidx + (size_t)(J2 >> 1) * n * n;
  const double rab = a == b ? M_SQRT2 : 1.0, rcd = c == d ? M_SQRT2 : 1.0;
  const double *x6 = T.x6 + (((size_t)(J1 >> 1) * T.nJ + (J2 >> 1)) * T.hx) * T.nh * T.nh;  // {J1 J2 R; . . .}
  auto six = [&](int j4, int j5, int j6) { return x6[((size_t)(j4 >> 1) * T.nh + (j5 >> 1)) * T.nh + (j6 >> 1)]; };
  double sm = 0.0;
  // kind 0: x = a, pair (i,b) at J2; 1: x = b, pair (i,a) at J2; 2: x = d, pair (c,i) at J1; 3: x = c, pair (d,i) at J1
#pragma unroll
  for (int kind = 0; kind < 4; kind++) {
    const int x = kind == 0 ? a : kind == 1 ? b : kind == 2 ? d : c;
    const int o = kind == 0 ? b : kind == 1 ? a : kind == 2 ? c : d;  // the pair's other orbit
    const int jo = L.jj[o], J = kind < 2 ? J2 : J1;
    const int k0 = T.pstart[x], k1 = T.pstart[x + 1];
    const double sg = kind == 0 ? -(double)sgnp((ja + jb - J2) >> 1)
                      : kind == 1 ? (double)sgnp((J1 + J2) >> 1)
                      : kind == 2 ? (double)sgnp((jc + jd - J1) >> 1)
                                  : -(double)sgnp((J1 + J2) >> 1);
    for (int k = k0; k < k1; k += 4) {
      int pi[4];
#pragma unroll
      for (int v = 0; v < 4; v++) pi[v] = T.part[min(k + v, k1 - 1)];
      long long pos[4];
      double w[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = pi[v] & 255, ji = pi[v] >> 8;
        // the pair's position in its block and v_elem's factor: kinds 0, 1 put i
        // first (bra (i,o) against the fixed ket (cd)), kinds 2, 3 second (ket (o,i))
        const int lo = min(i, o), hi = max(i, o);
        const int rr = (kind < 2 ? P2 : P1)[lo * n + hi];
        const bool swp = kind < 2 ? i > o : o > i;
        double f = swp ? (double)sgnp(1 + ((ji + jo - J) >> 1)) : (i == o ? M_SQRT2 : 1.0);
        const double rf = kind == 0 ? RF[(size_t)a * n + i] : kind == 1 ? RF[(size_t)b * n + i]
                        : kind == 2 ? RF[(size_t)i * n + d] : RF[(size_t)i * n + c];
        const double sx = kind == 0 ? six(ji, ja, jb) : kind == 1 ? six(ji, jb, ja)
                        : kind == 2 ? six(jd, ji, jc) : six(jc, ji, jd);
        const bool ok = k + v < k1 && rr >= 0;
        pos[v] = !ok ? 0 : kind < 2 ? (long long)rr * d2 + s2 : (long long)r1 * d1 + rr;
        w[v] = ok ? f * rf * sx : 0.0;
      }
      double g[4];
#pragma unroll
      for (int v = 0; v < 4; v++) g[v] = (kind < 2 ? V2 : V1)[pos[v]];
#pragma unroll
      for (int v = 0; v < 4; v++) sm += sg * w[v] * g[v] * (kind < 2 ? rcd : rab);
    }
  }
  sm *= sqrt((J1 + 1.0) * (J2 + 1.0) / (1.0 + (a == b)) / (1.0 + (c == d))) * sgnp(R >> 1);
  OV[vp] += sm;
}

// Tensor Pandya matrices of one cross-coupled block (calculate_single_pandya,
// cross_coupled.f90:1416-1634).  mode 0: CCX (r3 x nb4, rows = ordered pairs at
// J3 in the scalar Pandya row order, columns = (p,h) at J4); mode 1: CCR
// (nb3 x r4, rows = (p,h) at J3, columns = ordered pairs (b,a) at J4).
__global__ void k_tpandya(DevLayout L, DevTensor T, int mode, const double *__restrict__ RV, double *__restrict__ out,
                          int hR) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ntot = mode == 0 ? T.nx : T.nr;
  if (e >= ntot) return;
  const long long *st = mode == 0 ? T.xstart : T.rstart;
  const int x = find_item(st, T.ncc, e);
  const long long u = e - st[x];
  const int J3 = T.cJ3[x], J4 = T.cJ4[x], q3 = T.cq3[x], q4 = T.cq4[x];
  int a, b, p, h;
  long long o;  // the entry's position in its matrix
  if (mode == 0) {
    const int nb4 = L.ccnph[q4];
    const int R1 = (int)(u / nb4), k = T.cperm[L.cccbase[q4] + (int)(u % nb4)];
    o = (long long)R1 * nb4 + k;
    a = L.ccrow_a[L.ccrbase[q3] + R1];
    b = L.ccrow_b[L.ccrbase[q3] + R1];
    p = L.cccol_p[L.cccbase[q4] + k];
    h = L.cccol_h[L.cccbase[q4] + k];
  } else {
    const int r4 = L.ccrlen[q4];
    const int k = (int)(u / r4), G = T.rperm[L.ccrbase[q4] + (int)(u % r4)];
    o = (long long)k * r4 + G;
    p = L.cccol_p[L.cccbase[q3] + k];
    h = L.cccol_h[L.cccbase[q3] + k];
    b = L.ccrow_a[L.ccrbase[q4] + G];  // the row pair is (b, a)
    a = L.ccrow_b[L.ccrbase[q4] + G];
  }
  double v = 0.0;
  if (a >= 0) {
    const int ja = L.jj[a], jb = L.jj[b], jp = L.jj[p], jh = L.jj[h];
    if (((L.l[a] + L.l[h]) & 1) == ((L.l[b] + L.l[p] + (T.dpar >> 1)) & 1) && L.tz[a] + L.tz[h] == L.tz[b] + L.tz[p]) {
      // the (J3', J4') terms four at a time: positions, factors and 9j of a
      // batch first (every table read off any branch), then its four tensor
      // reads, so a batch makes three dependent round trips instead of one
      // chain per term.  The channel test is the same for every term.
      const int a3 = abs(ja - jh), n3 = (ja + jh - a3) / 2 + 1;
      const int a4 = abs(jp - jb), n4 = (jp + jb - a4) / 2 + 1, nt = n3 * n4;
      const bool chan = true;  // the parity and Tz tests above hold for both pairs of every term
      double sm = 0.0;
      for (int t0 = 0; t0 < nt; t0 += 4) {
        long long pos[4];
        double w[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int t = min(t0 + k, nt - 1);
          const int J3p = a3 + 2 * (t / n4), J4p = a4 + 2 * (t % n4);
          double f;
          pos[k] = mode == 0 ? t_pos_all(L, T, hR, a, h, p, b, ja, jh, jp, jb, J3p, J4p, chan, f)
                             : t_pos_all(L, T, hR, h, a, b, p, jh, ja, jb, jp, J3p, J4p, chan, f);
          const double nj = mode == 0 ? t9j_cc(T, ja, jb, J3, jh, jp, J4, J3p, J4p)
                                      : t9j_cc(T, jh, jp, J3, ja, jb, J4, J3p, J4p);
          w[k] = t0 + k < nt ? -sgnp(J4p >> 1) * sqrt((J3p + 1.0) * (J4p + 1.0)) * nj * f : 0.0;
        }
        double g[4];
#pragma unroll
        for (int k = 0; k < 4; k++) g[k] = RV[pos[k]];
#pragma unroll
        for (int k = 0; k < 4; k++) sm += w[k] * g[k];
      }
      const double s34 = sqrt((J3 + 1.0) * (J4 + 1.0));
      v = mode == 0 ? sm * sgnp((jh + jb + J4) >> 1) * s34
                    : sm * sgnp((jp + ja + J4) >> 1) * s34 * sgnp((jp + jh) >> 1);
    }
  }
  out[(mode == 0 ? T.xoff[x] : T.roff[x]) + o] = v;
}

// The ph recoupling of TS_commutator_222_ph (TS_commutators.f90:1459-1849) as
// a gather by orbit quadruple (x0 <= x1 | x2 <= x3), one wavefront each.  The
// reference adds every product entry W(IX = (d,a) at J3, JX = (c,b) at J4)
// (Wx or Wy) into up to 16 x (J1, J2) outputs; read backwards, an output
// (x0 x1 | x2 x3; J1 J2) receives entries through sixteen patterns: the source
// orbits (a,b,c,d) as a permutation of x, the orbit-order condition of the
// reference's branch, one of four 9j argument orders, a phase and a herm
// factor.  The entries a quadruple's outputs read do not depend on the
// outputs' (J1, J2): they are gathered once, over (pattern, J3, J4), with the
// parts of their weights that do not depend on (J1, J2) (sign, herm, (J3,J4)
// phase and norm, the |W| test) and the (J1, J2)-independent part of their 9j
// table index, into a compacted LDS list; each output (J1, J2) of the
// quadruple then sums the list across the lanes.
//
// The sixteen patterns (TS_commutators.f90:1562-1849): the output receives
// W(d,a;c,b) times P * 9j * sign * herm, P = sqrt((2J1+1)(2J2+1)(2J3+1)(2J4+1)),
// ph34 = (-1)^((J3+J4)/2), abcd = (-1)^((ja+jb+jc+jd)/2), all angular momenta doubled.
//  k  W   output = V^{J1 J2} of   branch          9j {. . J3; . . J4; J1 J2 R}   factor
//  0  Wx  (a,b,c,d)               b>=a, d>=c       (ja jd | jb jc)   (-1)^((jb+jd+J2)/2) ph34 hL
//  1  Wx  (c,d,a,b)               b>=a, d>=c       (jd ja | jc jb)  -(-1)^((jc+ja+J1+R)/2) hR
//  2  Wx  (a,b,d,c)               b>=a, d<c        (ja jd | jb jc)   (-1)^((jb+jc)/2) ph34 hL
//  3  Wx  (d,c,a,b)               b>=a, d<c        (jd ja | jc jb)   (-1)^((ja-jd+R)/2) hR
//  4  Wx  (b,a,c,d)               b<a, d>=c        (ja jd | jb jc)   (-1)^((ja+jd+J1+J2)/2) ph34 hL
//  5  Wx  (c,d,b,a)               b<a, d>=c        (jd ja | jc jb)   (-1)^((jc-jb+J1+J2+R)/2) hR
//  6  Wx  (b,a,d,c)               b<a, d<c         (ja jd | jb jc)   (-1)^((ja+jc+J1)/2) ph34 hL
//  7  Wx  (d,c,b,a)               b<a, d<c         (jd ja | jc jb)   (-1)^((jb-jd+J2+R)/2) hR
//  8  Wy  (d,b,c,a)               b>=d, a>=c       (jd ja | jb jc)   (-1)^((jc+ja+J2)/2) hL
//  9  Wy  (c,a,d,b)               b>=d, a>=c       (ja jd | jc jb)   (-1)^((jd-jb+J1+R)/2) ph34 hR
// 10  Wy  (d,b,a,c)               b>=d, a<c        (jd ja | jb jc)  -hL
// 11  Wy  (a,c,d,b)               b>=d, a<c        (ja jd | jc jb)   abcd ph34 (-1)^(R/2) hR
// 12  Wy  (b,d,c,a)               b<d, a>=c        (jd ja | jb jc)  -abcd (-1)^((J1+J2)/2) hL
// 13  Wy  (c,a,b,d)               b<d, a>=c        (ja jd | jc jb)   (-1)^((J1+J2+R)/2) ph34 hR
// 14  Wy  (b,d,a,c)               b<d, a<c         (jd ja | jb jc)   (-1)^((jb+jd+J1)/2) hL
// 15  Wy  (a,c,b,d)               b<d, a<c         (ja jd | jc jb)   (-1)^((ja-jc+J2+R)/2) ph34 hR
// Matching the output (x0 x1 | x2 x3) to the k-th form gives (a,b,c,d) as a
// permutation of x (the packed tables ia..id below) and turns the branch into
// x0 < x1 and / or x2 < x3; no Wy where J3 = J4 (the reference skips it).
// the patterns' (a, b, c, d) as indices into (x0, x1, x2, x3), two bits each,
// pattern k in bits 2k..2k+1, and their branch conditions (1: x0 < x1,
// 2: x2 < x3, 3: both)
constexpr unsigned kPhIa = 0x2727DD88u, kPhIb = 0x88DD88DDu, kPhIc = 0x72727272u, kPhId = 0xDD882727u;
constexpr unsigned kPhCond = 0xF960F960u;
struct PhSlot {
  double c;      // sign * herm * phase(J3,J4) * sqrt((2J3+1)(2J4+1)) * W
  unsigned b9;   // 9j table index without the (J1, J2) part
  int ab;        // bit 0: (-1)^(J1/2) applies, bit 1: (-1)^(J2/2) applies
};
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(7))) void k_t222ph_q(
    DevLayout L, DevTensor T, const double *__restrict__ Wx, const double *__restrict__ Wy, double *__restrict__ OV,
    int hL, int hR, int run) {
  extern __shared__ double ph_lds[];
  PhSlot *slot = reinterpret_cast<PhSlot *>(ph_lds);
  __shared__ int pa[16], pb[16], pc[16], pd[16], pok[16], pab[16], pph[16], pj[16][4], pj3[16], pstart[17];
  __shared__ double pcs[16];
  const long long qd = T.qrun[run] + blockIdx.x;
  const int lane = threadIdx.x, cap = T.qcap[run];
  if (qd >= T.qrun[run + 1]) return;
  const unsigned code = T.quads[qd];
  const int x0 = code & 255, x1 = (code >> 8) & 255, x2 = (code >> 16) & 255, x3 = (int)(code >> 24);
  const int R = T.rank, n = L.nsp, JM = L.jtot_max, nd = T.nd, nJ = T.nJ, h = T.nh, dpb = (T.dpar >> 1) & 1;
  if (lane < 16) {
    const int k = lane;
    auto pick = [&](unsigned tbl) {
      const int i = (tbl >> (2 * k)) & 3;
      return i == 0 ? x0 : i == 1 ? x1 : i == 2 ? x2 : x3;
    };
    const int a = pick(kPhIa), b = pick(kPhIb), c = pick(kPhIc), d = pick(kPhId);
    // the branch condition of the reference
    const unsigned cond = (kPhCond >> (2 * k)) & 3;
    bool ok = (!(cond & 1) || x0 < x1) && (!(cond & 2) || x2 < x3);
    const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
    const int Tz = abs(L.tz[d] - L.tz[a]) >> 1, PAR = (L.l[d] + L.l[a]) & 1, PAR2 = PAR ^ ((T.dpar >> 1) & 1);
    if (((L.l[c] + L.l[b]) & 1) != PAR2 || (abs(L.tz[c] - L.tz[b]) >> 1) != Tz) ok = false;
    // constant sign exponent (halves), J1/J2 flags and the (J3,J4) phase of each pattern
    // (TS_commutators.f90:1562-1849)
    int e = 0, ab = 0;
    double sg = 1.0;
    const int abcd = (ja + jb + jc + jd) >> 1;
    switch (k) {
      case 0: e = (jb + jd) >> 1; ab = 2; break;
      case 1: e = (jc + ja + R) >> 1; ab = 1; sg = -1.0; break;
      case 2: e = (jb + jc) >> 1; break;
      case 3: e = (ja - jd + R) >> 1; break;
      case 4: e = (ja + jd) >> 1; ab = 3; break;
      case 5: e = (jc - jb + R) >> 1; ab = 3; break;
      case 6: e = (ja + jc) >> 1; ab = 1; break;
      case 7: e = (jb - jd + R) >> 1; ab = 2; break;
      case 8: e = (jc + ja) >> 1; ab = 2; break;
      case 9: e = (jd - jb + R) >> 1; ab = 1; break;
      case 10: sg = -1.0; break;
      case 11: e = abcd + (R >> 1); break;
      case 12: e = abcd; ab = 3; sg = -1.0; break;
      case 13: e = R >> 1; ab = 3; break;
      case 14: e = (jb + jd) >> 1; ab = 1; break;
      default: e = (ja - jc + R) >> 1; ab = 2; break;
    }
    const bool ph34 = (k < 8) ? !(k & 1) : (k & 1);
    // the 9j argument order of the pattern: (ja,jd | jb,jc), (jd,ja | jc,jb), (jd,ja | jb,jc), (ja,jd | jc,jb)
    const int var = (k < 8) ? (k & 1) : 2 + (k & 1);
    pa[k] = a; pb[k] = b; pc[k] = c; pd[k] = d;
    pok[k] = ok ? (Tz | (PAR << 1) | 4) : 0;
    pj3[k] = abs(jd - ja);                                   // the pattern's J3 run: |jd-ja| .. jd+ja
    pstart[k + 1] = ok ? (((jd + ja - abs(jd - ja)) >> 1) + 1) * nd : 0;  // its slots (J3, J4 - J3)
    pab[k] = ab;
    pph[k] = ph34;
    pcs[k] = sg * sgnp(e) * ((k & 1) ? hR : hL);
    pj[k][0] = (var == 0 || var == 3) ? ja : jd;
    pj[k][1] = (var == 0 || var == 3) ? jd : ja;
    pj[k][2] = (var == 0 || var == 2) ? jb : jc;
    pj[k][3] = (var == 0 || var == 2) ? jc : jb;
  }
  __syncthreads();
  if (lane == 0) {
    pstart[0] = 0;
    for (int k = 0; k < 16; k++) pstart[k + 1] += pstart[k];
  }
  __syncthreads();
  // gather: slots (pattern, J3 of its run, J4 - J3), compacted to the entries that pass
  const int nslot = pstart[16];
  const double thr = (double)1e-6f;
  int count = 0, k = 0;  // a lane's pattern: its slot only grows, so k carries over
  for (int s0 = 0; s0 < nslot; s0 += 64) {
    const int s = s0 + lane;
    bool live = false;
    double cval = 0.0;
    unsigned b9 = 0;
    int abk = 0;
    {
      // every table read is made for every lane, with indices in range
      // whatever the tests say, so that they are in flight together: the
      // entry's row IX = (d,a) in the J3 Pandya block, its column JX = (c,b)
      // in the J4 one (the block index is arithmetic: J4, Tz and the parity of
      // the J3 pairs with the tensor's), the product's block, then W
      const int sc = min(s, nslot - 1);
      while (pstart[k + 1] <= sc) k++;
      const int r = sc - pstart[k], J3 = pj3[k] + 2 * (r / nd), J4 = J3 + 2 * (r % nd);
      const int okc = pok[k];
      const int a = pa[k], b = pb[k], c = pc[k], d = pd[k], Tz = okc & 1, PAR = (okc >> 1) & 1;
      const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
      const bool isx = k < 8;
      const bool tri = s < nslot && okc && dtri(jd, ja, J3) && J4 <= 2 * JM && dtri(J3, J4, R) &&
                       dtri(jc, jb, J4) && (isx || J4 != J3);
      const int J4c = min(J4, 2 * JM);
      const int q3 = (J3 >> 1) + Tz * (JM + 1) + 2 * PAR * (JM + 1);
      const int q4 = (J4c >> 1) + Tz * (JM + 1) + 2 * (PAR ^ dpb) * (JM + 1);
      const int IX = L.ccrow[((size_t)q3 * n + d) * n + a];
      const int JX = L.ccrow[((size_t)q4 * n + c) * n + b];
      const int qx = T.cidx[((J3 >> 1) * nJ + (J4c >> 1)) * 4 + Tz * 2 + PAR];
      const int rl4 = L.ccrlen[q4];
      const long long wo = T.woff[max(qx, 0)];
      const bool ok = tri && IX >= 0 && JX >= 0 && qx >= 0;
      const long long wi = ok ? wo + (long long)IX * rl4 + JX : 0;
      const double wv = (isx ? Wx : Wy)[wi];
      const double w = ok ? (isx ? hL * wv : wv) : 0.0;  // Wx carries LCC%herm (TS_commutators.f90:1506)
      if (fabs(w) > thr) {
        live = true;
        cval = pcs[k] * (pph[k] ? (double)sgnp((J3 + J4) >> 1) : 1.0) * sqrt((J3 + 1.0) * (J4 + 1.0)) * w;
        b9 = (unsigned)((((((pj[k][0] >> 1) * h + (pj[k][1] >> 1)) * h + (pj[k][2] >> 1)) * h +
                          (pj[k][3] >> 1)) * nJ + (J3 >> 1)) * nd + ((J4 - J3) >> 1)) * (unsigned)(nJ * nd);
        abk = pab[k];
      }
    }
    const unsigned long long m = __ballot(live);
    if (live) {
      // pos < this quadruple's slot count <= its run's cap (the host counts the same slots)
      const int pos = min(count + __popcll(m & ((1ull << lane) - 1ull)), cap - 1);
      slot[pos].c = cval;
      slot[pos].b9 = b9;
      slot[pos].ab = abk;
    }
    count += __popcll(m);
  }
  __syncthreads();
  if (count == 0) return;
  // outputs (x0 x1 | x2 x3; J1 <= J2) of the quadruple, 64 at a time: lane o
  // takes J1 = |j0 - j1| + 2 (o / nd), J2 = J1 + 2 (o % nd), finds its element
  // and reads its old value (every table read off any branch, for all the
  // outputs at once); the outputs' sums over the slots then go four at a time
  // (their 9j reads in flight together), each lane keeps its own output's sum,
  // and the lanes write their elements together.
  const int j0 = L.jj[x0], j1 = L.jj[x1], j2 = L.jj[x2], j3 = L.jj[x3];
  const int Tzq = (L.tz[x0] + L.tz[x1]) / 2 + 1, P1 = (L.l[x0] + L.l[x1]) & 1;
  const double pre = (x0 == x1 ? M_SQRT2 : 1.0) * (x2 == x3 ? M_SQRT2 : 1.0);
  const int Jlo = abs(j0 - j1), nout = (min(j0, j1) + 1) * nd;
  for (int o0 = 0; o0 < nout; o0 += 64) {
    const int o = o0 + lane;
    const int J1 = Jlo + 2 * (o / nd), J2 = J1 + 2 * (o % nd);
    bool val = o < nout && J1 <= j0 + j1 && J2 >= abs(j2 - j3) && J2 <= j2 + j3 && J1 + J2 >= R &&
               !(x0 == x1 && ((J1 >> 1) & 1)) && !(x2 == x3 && ((J2 >> 1) & 1));
    const int J1c = val ? J1 : Jlo, J2c = val ? J2 : Jlo;  // in the tables' range
    const int q = T.bidx[((J1c >> 1) * nJ + (J2c >> 1)) * 6 + Tzq * 2 + P1];
    const int r = dpos(L, J1c, x0, x1), sc = dpos(L, J2c, x2, x3);
    const int qc = max(q, 0);
    val = val && q >= 0 && (r | sc) >= 0;
    const long long po = val ? T.off[qc] + (long long)r * T.n2[qc] + sc : 0;
    const double old = OV[po];
    unsigned long long vm = __ballot(val);
    double res = 0.0;
    while (vm) {
      int ok_[4];
      unsigned offk[4];
      int ek[4][2];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        ok_[k] = vm != 0;
        const int ol = vm ? __ffsll((long long)vm) - 1 : 0;
        vm &= vm - 1;
        const int J1k = __shfl(J1, ol, 64), J2k = __shfl(J2, ol, 64);
        offk[k] = (unsigned)((J1k >> 1) * nd + ((J2k - J1k) >> 1));
        ek[k][0] = (J1k >> 1) & 1;
        ek[k][1] = (J2k >> 1) & 1;
        ok_[k] = ok_[k] ? ol : -1;
      }
      double acc[4] = {0.0, 0.0, 0.0, 0.0};
      for (int t = lane; t < count; t += 64) {
        const PhSlot sl = slot[t];
        double v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = T.n9[sl.b9 + offk[k]];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int e = ((sl.ab & 1) ? ek[k][0] : 0) + ((sl.ab & 2) ? ek[k][1] : 0);
          const double x = sl.c * v[k];
          acc[k] += (e & 1) ? -x : x;
        }
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const double a = wave_sum(acc[k]);
        if (lane == ok_[k]) res = a;
      }
    }
    if (val) OV[po] = old + res * sqrt((J1 + 1.0) * (J2 + 1.0)) * pre;
  }
}

// ---------------------------------------------------------------- products
// 64 x 64 tiles, 4 waves of 32 x 32 on v_mfma_f64_16x16x4; tiles (problem, tile row, tile column)
__global__ __launch_bounds__(256) void k_tgemm(const TProb *__restrict__ probs, const int4 *__restrict__ tiles,
                                               int ntiles, const double *__restrict__ X0,
                                               const double *__restrict__ X1, double *__restrict__ C) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int4 tl = tiles[t];
  const TProb &p = probs[tl.x];
  const int m0 = tl.y * 64, n0 = tl.z * 64;
  __shared__ double As[16][65], Bs[16][65];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
  double4_t acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (double4_t){0.0, 0.0, 0.0, 0.0};
  // each thread stages 4 elements of A and 4 of B per K step of 16; the next
  // step's elements are loaded into registers (clamped to the operand, so
  // every load is unconditional and the eight stay in flight together) while
  // the current step runs on the matrix cores
  for (int sgi = 0; sgi < p.nseg; sgi++) {
    const TSeg sg = p.seg[sgi];
    if (sg.K <= 0) continue;
    const double *Ap = (sg.asrc ? X1 : X0) + sg.a_off, *Bp = (sg.bsrc ? X1 : X0) + sg.b_off;
    const bool arow = sg.a_cs == 1, brow = sg.b_cs == 1;
    int am[4], ak[4], bn[4], bk[4];  // element (row/column, k) of each staged value in the tile
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = tid + 256 * u;
      am[u] = arow ? e >> 4 : e & 63;
      ak[u] = arow ? e & 15 : e >> 6;
      bn[u] = brow ? e & 63 : e >> 4;
      bk[u] = brow ? e >> 6 : e & 15;
    }
    double ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int m = min(m0 + am[u], p.M - 1), k = min(k0 + ak[u], sg.K - 1);
        const int n = min(n0 + bn[u], p.N - 1), k2 = min(k0 + bk[u], sg.K - 1);
        ra[u] = Ap[(long long)m * sg.a_rs + (long long)k * sg.a_cs];
        rb[u] = Bp[(long long)k2 * sg.b_rs + (long long)n * sg.b_cs];
      }
    };
    gload(0);
    for (int k0 = 0; k0 < sg.K; k0 += 16) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        As[ak[u]][am[u]] = (m0 + am[u] < p.M && k0 + ak[u] < sg.K) ? sg.alpha * ra[u] : 0.0;
        Bs[bk[u]][bn[u]] = (n0 + bn[u] < p.N && k0 + bk[u] < sg.K) ? rb[u] : 0.0;
      }
      __syncthreads();
      gload(k0 + 16);  // past the end: clamped reloads, never used
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

// mat_frob_norm of a tensor (basic_IMSRG.f90:3921-3952): fhh^2 + 2 fph^2 + fpp^2
// + sum over blocks (1 if J1 = J2, else 2) x block^2; one partial per workgroup
__global__ void k_tnorm(DevLayout L, DevTensor T, const double *__restrict__ op, double *__restrict__ part) {
  __shared__ double sh[256];
  const int n = L.nsp;
  const long long nf = (long long)n * n, ntot = nf + T.nelem;
  double s = 0.0;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < ntot; e += (long long)gridDim.x * blockDim.x) {
    if (e < nf) {
      const int a = (int)(e / n), b = (int)(e % n);
      const double v = op[T.f_off + e];
      if (L.con[a] == L.con[b]) s += v * v;
      else if (L.con[a] == 0) s += 2.0 * v * v;  // (particle, hole); (hole, particle) is derived
    } else {
      int q, a, b, c, d;
      long long vp;
      elem_of(L, T, e - nf, q, a, b, c, d, vp);
      const double v = op[T.v_off + vp];
      s += (T.J1[q] == T.J2[q] ? 1.0 : 2.0) * v * v;
    }
  }
  const double r = block_sum256(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

constexpr int kNormParts = 512;

}  // namespace

// ---------------------------------------------------------------- host side
struct TensorSpace {
  int rank = 0, dpar = 0, nblk = 0, ncc = 0, hx = 0;
  std::vector<int> J1, J2, q1, q2, n1, n2, bidx;
  std::vector<long long> off, tstart;
  long long vsize = 0, v_off = 0, op_size = 0, neq = 0, nelem = 0;
  std::vector<int> cJ3, cJ4, cq3, cq4, cidx, ccol;
  std::vector<long long> xoff, roff, woff, xstart, rstart;
  long long xsize = 0, rsize = 0, wsize = 0;
  std::vector<double> x6, s6;
  long long n9size = 0;
  DevTensor dt{};
  double *gcc = nullptr, *ccx = nullptr, *ccr = nullptr, *wx = nullptr, *wy = nullptr;
  double *w1 = nullptr, *w2 = nullptr, *part = nullptr, *hpart = nullptr;
  const TProb *pp_probs[2] = {nullptr, nullptr}, *cc_probs[2] = {nullptr, nullptr};
  const int4 *pp_tiles = nullptr, *cc_tiles[2] = {nullptr, nullptr};
  int n_pp_tiles = 0, n_cc_tiles[2] = {0, 0};
  ~TensorSpace() {
    if (hpart) (void)hipHostFree(hpart);
  }
};

struct imsrg_tensor_s {
  imsrg_ctx_s *ctx = nullptr;
  TensorSpace *ts = nullptr;
  double *d = nullptr;
  int herm = 1;
};

static long long align32(long long x) { return (x + 31) & ~31LL; }

static void add_tiles(int pi, int M, int N, std::vector<int4> &tiles) {
  for (int i = 0; i < (M + 63) / 64; i++)
    for (int j = 0; j < (N + 63) / 64; j++) tiles.push_back(make_int4(pi, i, j, 0));
}

static int tspace_build(imsrg_ctx_s *c, int rank, int dpar, TensorSpace &S) {
  const Layout &Ly = c->lay;
  const int JM = Ly.jtot_max, nJ = Ly.nJ, nsp = Ly.nsp, nh = Ly.nhalf;
  S.rank = rank;
  S.dpar = dpar;
  const int dp = (dpar >> 1) & 1;
  // tensor blocks, allocate_tensor's order (basic_IMSRG.f90:690-770)
  S.bidx.assign((size_t)nJ * nJ * 6, -1);
  long long off = 0, el = 0;
  for (int J1 = 0; J1 <= 2 * JM; J1 += 2)
    for (int J2 = std::max(std::abs(J1 - rank), J1); J2 <= J1 + rank; J2 += 2)
      for (int Tz = -1; Tz <= 1; Tz++)
        for (int P1 = 0; P1 <= 1; P1++) {
          const int q = (int)S.J1.size();
          const bool live = J2 <= 2 * JM;
          const int b1 = live ? 3 * J1 + 2 * (Tz + 1) + P1 : -1;
          const int b2 = live ? 3 * J2 + 2 * (Tz + 1) + (P1 ^ dp) : -1;
          S.J1.push_back(J1);
          S.J2.push_back(J2);
          S.q1.push_back(b1);
          S.q2.push_back(b2);
          const int m1 = live ? Ly.bdim[b1] : 0, m2 = live ? Ly.bdim[b2] : 0;
          S.n1.push_back(m1);
          S.n2.push_back(m2);
          S.off.push_back(off);
          S.tstart.push_back(el);
          off = align32(off + (long long)m1 * m2);
          el += (long long)m1 * m2;
          if (live) S.bidx[((size_t)(J1 >> 1) * nJ + (J2 >> 1)) * 6 + (Tz + 1) * 2 + P1] = q;
        }
  S.nblk = (int)S.J1.size();
  S.tstart.push_back(el);
  S.nelem = el;
  S.vsize = off;
  S.v_off = align32(Ly.f_off + (long long)nsp * nsp);
  S.op_size = S.v_off + S.vsize;
  S.neq = 1 + (long long)Ly.nhole * Ly.nhole + (long long)Ly.npart * Ly.npart + (long long)Ly.npart * Ly.nhole + el;
  // cross-coupled tensor blocks (allocate_small_tensor_CCMAT, cross_coupled.f90:498-700):
  // J3, J4 >= J3 within rank, |Tz|, parity of the J3 pairs
  S.cidx.assign((size_t)nJ * nJ * 4, -1);
  long long xo = 0, ro = 0, wo = 0, xe = 0, re = 0;
  for (int J3 = 0; J3 <= 2 * JM; J3 += 2)
    for (int J4 = std::max(std::abs(J3 - rank), J3); J4 <= std::min(J3 + rank, 2 * JM); J4 += 2)
      for (int Tz = 0; Tz <= 1; Tz++)
        for (int P = 0; P <= 1; P++) {
          const int x = (int)S.cJ3.size();
          const int q3 = (J3 >> 1) + Tz * (JM + 1) + 2 * P * (JM + 1);
          const int q4 = (J4 >> 1) + Tz * (JM + 1) + 2 * (P ^ dp) * (JM + 1);
          S.cJ3.push_back(J3);
          S.cJ4.push_back(J4);
          S.cq3.push_back(q3);
          S.cq4.push_back(q4);
          S.cidx[((size_t)(J3 >> 1) * nJ + (J4 >> 1)) * 4 + Tz * 2 + P] = x;
          const long long r3 = Ly.ccrlen[q3], r4 = Ly.ccrlen[q4], nb3 = Ly.ccnph[q3], nb4 = Ly.ccnph[q4];
          S.xoff.push_back(xo);
          S.roff.push_back(ro);
          S.woff.push_back(wo);
          S.xstart.push_back(xe);
          S.rstart.push_back(re);
          xo = align32(xo + r3 * nb4);
          ro = align32(ro + nb3 * r4);
          wo = align32(wo + r3 * r4);
          xe += r3 * nb4;
          re += nb3 * r4;
        }
  S.ncc = (int)S.cJ3.size();
  S.xstart.push_back(xe);
  S.rstart.push_back(re);
  S.xsize = xo;
  S.rsize = ro;
  S.wsize = wo;
  // column of (p,h) inside each scalar Pandya block
  S.ccol.assign((size_t)Ly.ncc * nsp * nsp, -1);
  for (int q = 0; q < Ly.ncc; q++)
    for (int k = Ly.cccbase[q]; k < Ly.cccbase[q + 1]; k++)
      S.ccol[((size_t)q * nsp + Ly.cccol_p[k]) * nsp + Ly.cccol_h[k]] = k - Ly.cccbase[q];
  // 6j tables of the tensor terms (xxxsixj, store_6j_3halfint, basic_IMSRG.f90:2291-2370, 2558-2595)
  S.hx = (JM + rank + 1) / 2;
  const int hx = S.hx;
  S.x6.assign((size_t)nJ * nJ * hx * nh * nh, 0.0);
  for (int a1 = 0; a1 < nJ; a1++)
    for (int a2 = 0; a2 < nJ; a2++) {
      if (!(2 * a2 <= 2 * a1 + rank && 2 * a2 >= std::abs(2 * a1 - rank))) continue;
      for (int i4 = 0; i4 < hx; i4++)
        for (int i5 = 0; i5 < nh; i5++)
          for (int i6 = 0; i6 < nh; i6++)
            S.x6[((((size_t)a1 * nJ + a2) * hx + i4) * nh + i5) * nh + i6] =
                racah6j(2 * a1, 2 * a2, rank, 2 * i4 + 1, 2 * i5 + 1, 2 * i6 + 1);
    }
  S.s6.assign((size_t)nh * nh * nJ * nh * hx * nJ, 0.0);
  for (int i1 = 0; i1 < nh; i1++)
    for (int i2 = 0; i2 < nh; i2++)
      for (int a3 = 0; a3 < nJ; a3++) {
        if (!(2 * a3 <= 2 * i1 + 1 + 2 * i2 + 1 && 2 * a3 >= std::abs(2 * i1 - 2 * i2))) continue;
        for (int i4 = 0; i4 < nh; i4++)
          for (int i5 = 0; i5 < hx; i5++)
            for (int a6 = 0; a6 < nJ; a6++)
              S.s6[(((((size_t)i1 * nh + i2) * nJ + a3) * nh + i4) * hx + i5) * nJ + a6] =
                  racah6j(2 * i1 + 1, 2 * i2 + 1, 2 * a3, 2 * i4 + 1, 2 * i5 + 1, 2 * a6);
      }
  // device tables
  DevTensor &D = S.dt;
  D.rank = rank; D.dpar = dpar; D.nblk = S.nblk; D.ncc = S.ncc; D.nJ = nJ; D.nh = nh; D.hx = hx;
  D.nhole = Ly.nhole; D.npart = Ly.npart; D.nelem = S.nelem; D.nx = xe; D.nr = re;
  D.f_off = Ly.f_off; D.v_off = S.v_off;
#define TUP(vec, field) \
  if (upload_table(c, vec, &D.field) != 0) return IMSRG_ENOMEM
  TUP(Ly.holes, holes); TUP(Ly.parts, parts);
  TUP(S.J1, J1); TUP(S.J2, J2); TUP(S.q1, q1); TUP(S.q2, q2); TUP(S.n1, n1); TUP(S.n2, n2);
  TUP(S.off, off); TUP(S.tstart, tstart); TUP(S.bidx, bidx);
  TUP(S.cJ3, cJ3); TUP(S.cJ4, cJ4); TUP(S.cq3, cq3); TUP(S.cq4, cq4);
  TUP(S.xoff, xoff); TUP(S.roff, roff); TUP(S.woff, woff); TUP(S.xstart, xstart); TUP(S.rstart, rstart);
  TUP(S.cidx, cidx); TUP(S.ccol, ccol); TUP(S.x6, x6); TUP(S.s6, s6);
  {
    std::vector<int> cperm(Ly.cccbase[Ly.ncc]), rperm(Ly.ccrbase[Ly.ncc]);
    for (int q = 0; q < Ly.ncc; q++) {
      const int c0 = Ly.cccbase[q], nc = Ly.cccbase[q + 1] - c0, r0 = Ly.ccrbase[q], nr = Ly.ccrlen[q];
      std::vector<int> ix(nc);
      for (int k = 0; k < nc; k++) ix[k] = k;
      std::stable_sort(ix.begin(), ix.end(), [&](int x, int y) {
        return std::make_pair(Ly.jj[Ly.cccol_p[c0 + x]], Ly.jj[Ly.cccol_h[c0 + x]]) <
               std::make_pair(Ly.jj[Ly.cccol_p[c0 + y]], Ly.jj[Ly.cccol_h[c0 + y]]);
      });
      std::copy(ix.begin(), ix.end(), cperm.begin() + c0);
      auto rkey = [&](int g) {
        const int a = Ly.ccrow_a[r0 + g], b = Ly.ccrow_b[r0 + g];
        return a < 0 ? std::make_pair(1 << 20, 0) : std::make_pair(Ly.jj[a], Ly.jj[b]);
      };
      ix.resize(nr);
      for (int g = 0; g < nr; g++) ix[g] = g;
      std::stable_sort(ix.begin(), ix.end(), [&](int x, int y) { return rkey(x) < rkey(y); });
      std::copy(ix.begin(), ix.end(), rperm.begin() + r0);
    }
    TUP(cperm, cperm); TUP(rperm, rperm);
    std::vector<int> sperm(Ly.npairs);
    for (int q = 0; q < Ly.nblocks; q++) {
      const int p0 = Ly.pbase[q], np = Ly.pbase[q + 1] - p0;
      std::vector<int> ix(np);
      for (int k = 0; k < np; k++) ix[k] = k;
      std::stable_sort(ix.begin(), ix.end(), [&](int x, int y) {
        return std::make_pair(Ly.grp_of[Ly.pair_a[p0 + x]], Ly.grp_of[Ly.pair_b[p0 + x]]) <
               std::make_pair(Ly.grp_of[Ly.pair_a[p0 + y]], Ly.grp_of[Ly.pair_b[p0 + y]]);
      });
      std::copy(ix.begin(), ix.end(), sperm.begin() + p0);
    }
    TUP(sperm, sperm);
  }
  {
    std::vector<int> ps(nsp + 1, 0), pl;
    for (int x = 0; x < nsp; x++) {
      ps[x] = (int)pl.size();
      for (int i = 0; i < nsp; i++)
        if (Ly.tz[i] == Ly.tz[x] && ((Ly.l[i] + dp) & 1) == (Ly.l[x] & 1) && std::abs(Ly.jj[i] - Ly.jj[x]) <= rank &&
            rank <= Ly.jj[i] + Ly.jj[x])
          pl.push_back(i | Ly.jj[i] << 8);
    }
    ps[nsp] = (int)pl.size();
    if (pl.empty()) pl.push_back(0);  // a readable entry for the clamped loads
    TUP(ps, pstart);
    TUP(pl, part);
  }
#undef TUP
  // 9j table of the recoupling and the tensor Pandya transform (up to 4 GiB:
  // rank 2 to emax=14, rank 3 to emax=12)
  D.nd = rank / 2 + 1;
  {
    const long long n9 = (long long)nh * nh * nh * nh * nJ * nJ * D.nd * D.nd;
    if (n9 > (1LL << 29)) return fail(IMSRG_EINVAL, "tensor rank too large for the 9j table at this basis");
    double *t9 = nullptr;
    if (dalloc(c, n9, &t9)) return IMSRG_ENOMEM;
    hipLaunchKernelGGL(k_fill9j, dim3((unsigned)((n9 + 255) / 256)), dim3(256), 0, c->stream, D, t9, n9);
    D.n9 = t9;
    S.n9size = n9;
  }
  // orbit quadruples of the recoupling gather (k_t222ph_q): row pair x0 <= x1,
  // column pair x2 <= x3, with at least one stored output (orbits packed in 8 bits)
  if (nsp > 256) return fail(IMSRG_EINVAL, "tensor operators need at most 256 orbits");
  {
    std::vector<unsigned> qrun[4];  // by slot count: at most 256, 384, 512, more
    std::vector<int> pa_, pb_;
    int nslotmax = 1;
    for (int a = 0; a < nsp; a++)
      for (int b = a; b < nsp; b++) { pa_.push_back(a); pb_.push_back(b); }
    for (size_t u = 0; u < pa_.size(); u++) {
      const int x0 = pa_[u], x1 = pb_[u], j0 = Ly.jj[x0], j1 = Ly.jj[x1];
      const int t1 = Ly.tz[x0] + Ly.tz[x1], p1 = (Ly.l[x0] + Ly.l[x1]) & 1;
      for (size_t v = 0; v < pa_.size(); v++) {
        const int x2 = pa_[v], x3 = pb_[v];
        if (Ly.tz[x2] + Ly.tz[x3] != t1 || ((Ly.l[x2] + Ly.l[x3]) & 1) != (p1 ^ dp)) continue;
        const int j2 = Ly.jj[x2], j3 = Ly.jj[x3];
        bool any = false;
        for (int J1 = std::abs(j0 - j1); J1 <= j0 + j1 && !any; J1 += 2) {
          if (x0 == x1 && ((J1 >> 1) & 1)) continue;
          for (int J2 = std::max(J1, std::abs(j2 - j3)); J2 <= std::min(j2 + j3, J1 + rank); J2 += 2) {
            if (x2 == x3 && ((J2 >> 1) & 1)) continue;
            if (J1 + J2 < rank) continue;
            any = true;
            break;
          }
        }
        if (!any) continue;
        const unsigned code = (unsigned)x0 | (unsigned)x1 << 8 | (unsigned)x2 << 16 | (unsigned)x3 << 24;
        // the quadruple's gather slots (k_t222ph_q): a run of J3 times the
        // J4 - J3 steps per pattern that passes the branch, parity and Tz tests
        const int xs[4] = {x0, x1, x2, x3};
        int ns = 0;
        for (int k = 0; k < 16; k++) {
          const int a = xs[(kPhIa >> (2 * k)) & 3], b = xs[(kPhIb >> (2 * k)) & 3];
          const int cc = xs[(kPhIc >> (2 * k)) & 3], dd = xs[(kPhId >> (2 * k)) & 3];
          const unsigned cond = (kPhCond >> (2 * k)) & 3;
          if (((cond & 1) && !(x0 < x1)) || ((cond & 2) && !(x2 < x3))) continue;
          const int Tz = std::abs(Ly.tz[dd] - Ly.tz[a]) >> 1, PAR2 = ((Ly.l[dd] + Ly.l[a]) & 1) ^ dp;
          if (((Ly.l[cc] + Ly.l[b]) & 1) != PAR2 || (std::abs(Ly.tz[cc] - Ly.tz[b]) >> 1) != Tz) continue;
          const int jd = Ly.jj[dd], ja = Ly.jj[a];
          ns += ((jd + ja - std::abs(jd - ja)) / 2 + 1) * (rank / 2 + 1);
        }
        nslotmax = std::max(nslotmax, ns);
        qrun[ns <= 256 ? 0 : ns <= 384 ? 1 : ns <= 512 ? 2 : 3].push_back(code);
      }
    }
    std::vector<unsigned> quads;
    D.qrun[0] = 0;
    for (int k = 0; k < 4; k++) {
      quads.insert(quads.end(), qrun[k].begin(), qrun[k].end());
      D.qrun[k + 1] = (long long)quads.size();
    }
    if (upload_table(c, quads, &D.quads) != 0) return IMSRG_ENOMEM;
    D.nquad = (long long)quads.size();
    D.nslotmax = nslotmax;
    D.qcap[0] = std::min(256, nslotmax);
    D.qcap[1] = std::min(384, nslotmax);
    D.qcap[2] = std::min(512, nslotmax);
    D.qcap[3] = nslotmax;
  }
  // products: the pp and hh ladders per tensor block (W1, W2)
  {
    std::vector<TProb> pr[2];
    std::vector<int4> tiles;
    for (int q = 0; q < S.nblk; q++) {
      const int m1 = S.n1[q], m2 = S.n2[q];
      if (m1 == 0 || m2 == 0) continue;
      const int b1 = S.q1[q], b2 = S.q2[q];
      const int np1 = Ly.bnp[b1], nb1 = Ly.bnb[b1], nh1 = Ly.bnh[b1];
      const int np2 = Ly.bnp[b2], nb2 = Ly.bnb[b2], nh2 = Ly.bnh[b2];
      for (int w = 0; w < 2; w++) {  // w 0: pp intermediate, 1: hh
        const int k1 = w ? np1 + nb1 : 0, K1 = w ? nh1 : np1, k2 = w ? np2 + nb2 : 0, K2 = w ? nh2 : np2;
        TProb p{};
        p.M = m1; p.N = m2; p.ldc = m2; p.nseg = 2; p.c_off = S.off[q];
        // G(J1)[:, k] T[k, :]  (operand 0: G's V, 1: T's V)
        p.seg[0] = TSeg{Ly.boff[b1] + k1, S.off[q] + (long long)k1 * m2, K1, m1, 1, m2, 1, 0, 1, 1.0};
        // - T[:, k] G(J2)[k, :]
        p.seg[1] = TSeg{S.off[q] + k2, Ly.boff[b2] + (long long)k2 * m2, K2, m2, 1, m2, 1, 1, 0, -1.0};
        pr[w].push_back(p);
      }
      add_tiles((int)pr[0].size() - 1, m1, m2, tiles);
    }
    for (int w = 0; w < 2; w++)
      if (upload_table(c, pr[w], &S.pp_probs[w]) != 0) return IMSRG_ENOMEM;
    if (upload_table(c, tiles, &S.pp_tiles) != 0) return IMSRG_ENOMEM;
    S.n_pp_tiles = (int)tiles.size();
  }
  // the two Pandya products per cross-coupled block (TS_commutator_222_ph, TS_commutators.f90:1496-1521):
  // alpha carries 1/sqrt(J3+1) and 1/sqrt(J4+1); the herm of G that the reference also folds into Wx
  // is applied where k_t222ph_q reads Wx, so these tables serve every G
  {
    std::vector<TProb> pr[2];
    std::vector<int4> tiles[2];
    for (int x = 0; x < S.ncc; x++) {
      const int q3 = S.cq3[x], q4 = S.cq4[x];
      const int r3 = Ly.ccrlen[q3], r4 = Ly.ccrlen[q4], nb3 = Ly.ccnph[q3], nb4 = Ly.ccnph[q4];
      if (r3 == 0 || r4 == 0) continue;
      TProb px{};  // Wx = LCC(J3) CCR / sqrt(J3+1): operand 0 the Pandya matrices of G, 1 CCR
      px.M = r3; px.N = r4; px.ldc = r4; px.c_off = S.woff[x];
      px.nseg = nb3 > 0 ? 1 : 0;
      px.seg[0] = TSeg{Ly.ccoff[q3], S.roff[x], nb3, 1, r3, r4, 1, 0, 1, 1.0 / std::sqrt(S.cJ3[x] + 1.0)};
      pr[0].push_back(px);
      add_tiles((int)pr[0].size() - 1, r3, r4, tiles[0]);
      TProb py{};  // Wy = CCX LCC(J4)^T / sqrt(J4+1): operand 0 CCX, 1 the Pandya matrices of G
      py.M = r3; py.N = r4; py.ldc = r4; py.c_off = S.woff[x];
      py.nseg = (nb4 > 0 && S.cJ3[x] != S.cJ4[x]) ? 1 : 0;
      py.seg[0] = TSeg{S.xoff[x], Ly.ccoff[q4], nb4, nb4, 1, r4, 1, 0, 1, 1.0 / std::sqrt(S.cJ4[x] + 1.0)};
      pr[1].push_back(py);
      add_tiles((int)pr[1].size() - 1, r3, r4, tiles[1]);
    }
    for (int w = 0; w < 2; w++) {
      if (upload_table(c, pr[w], &S.cc_probs[w]) != 0) return IMSRG_ENOMEM;
      if (upload_table(c, tiles[w], &S.cc_tiles[w]) != 0) return IMSRG_ENOMEM;
      S.n_cc_tiles[w] = (int)tiles[w].size();
    }
  }
  // workspaces
  if (dalloc(c, Ly.ccsize, &S.gcc) || dalloc(c, S.xsize, &S.ccx) || dalloc(c, S.rsize, &S.ccr) ||
      dalloc(c, S.wsize, &S.wx) || dalloc(c, S.wsize, &S.wy) || dalloc(c, S.vsize, &S.w1) ||
      dalloc(c, S.vsize, &S.w2) || dalloc(c, kNormParts, &S.part))
    return IMSRG_ENOMEM;
  launch_fill(c->stream, S.w1, S.vsize, 0.0);
  launch_fill(c->stream, S.w2, S.vsize, 0.0);
  launch_fill(c->stream, S.wx, S.wsize, 0.0);
  launch_fill(c->stream, S.wy, S.wsize, 0.0);
  HIPCHK(hipHostMalloc((void **)&S.hpart, kNormParts * sizeof(double)));
  return 0;
}

static int tspace_get(imsrg_ctx_s *c, int rank, int dpar, TensorSpace **out) {
  auto key = std::make_pair(rank, dpar);
  auto it = c->tspaces.find(key);
  if (it != c->tspaces.end()) {
    *out = it->second.get();
    return 0;
  }
  auto S = std::make_shared<TensorSpace>();
  const int rc = tspace_build(c, rank, dpar, *S);
  if (rc != 0) return rc;
  c->tspaces[key] = S;
  *out = S.get();
  return 0;
}

// ---------------------------------------------------------------- C ABI
static unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

int imsrg_tensor_create(imsrg_ctx c, int rank, int dpar, int herm, imsrg_tensor *out) {
  if (!c || !out) return fail(IMSRG_EINVAL, "null argument");
  if (c->nbatch != 1) return fail(IMSRG_EINVAL, "tensor operators need a single-instance context");
  if (rank <= 0 || (rank & 1)) return fail(IMSRG_EINVAL, "rank must be doubled and positive (2 lambda)");
  if (dpar != 0 && dpar != 2) return fail(IMSRG_EINVAL, "dpar must be 0 or 2 (doubled)");
  if (herm != 1 && herm != -1) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  TensorSpace *S = nullptr;
  if (int rc = tspace_get(c, rank, dpar, &S)) return rc;
  imsrg_tensor_s *t = new imsrg_tensor_s;
  t->ctx = c;
  t->ts = S;
  t->herm = herm;
  if (hipMalloc(&t->d, sizeof(double) * S->op_size) != hipSuccess) {
    delete t;
    return fail(IMSRG_ENOMEM, "hipMalloc tensor operator failed");
  }
  launch_fill(c->stream, t->d, S->op_size, 0.0);
  *out = t;
  return 0;
}

int imsrg_tensor_destroy(imsrg_tensor t) {
  if (!t) return 0;
  (void)hipStreamSynchronize(t->ctx->stream);
  (void)hipFree(t->d);
  delete t;
  return 0;
}

int64_t imsrg_tensor_neq(imsrg_tensor t) { return t ? t->ts->neq : -1; }
int imsrg_tensor_set_herm(imsrg_tensor t, int herm) {
  if (!t || (herm != 1 && herm != -1)) return fail(IMSRG_EINVAL, "bad argument");
  t->herm = herm;
  return 0;
}
int imsrg_tensor_get_herm(imsrg_tensor t, int *herm) {
  if (!t || !herm) return fail(IMSRG_EINVAL, "null argument");
  *herm = t->herm;
  return 0;
}

// stream <-> device image (see the header for the stream order)
static void stream_to_image(const Layout &Ly, const TensorSpace &S, int herm, const double *v, double *img) {
  std::fill(img, img + S.op_size, 0.0);
  const int A = Ly.nhole, P = Ly.npart, n = Ly.nsp;
  long long k = 0;
  img[0] = v[k++];
  double *F = img + Ly.f_off;
  for (int j = 0; j < A; j++)
    for (int i = 0; i < A; i++) F[(size_t)Ly.holes[i] * n + Ly.holes[j]] = v[k + i + (long long)j * A];
  k += (long long)A * A;
  for (int j = 0; j < P; j++)
    for (int i = 0; i < P; i++) F[(size_t)Ly.parts[i] * n + Ly.parts[j]] = v[k + i + (long long)j * P];
  k += (long long)P * P;
  for (int j = 0; j < A; j++)
    for (int i = 0; i < P; i++) {
      const int p = Ly.parts[i], h = Ly.holes[j];
      const double x = v[k + i + (long long)j * P];
      F[(size_t)p * n + h] = x;
      F[(size_t)h * n + p] = x * herm * (((Ly.jj[h] - Ly.jj[p]) / 2) & 1 ? -1.0 : 1.0);
    }
  k += (long long)P * A;
  double *V = img + S.v_off;
  for (int q = 0; q < S.nblk; q++) {
    const int m1 = S.n1[q], m2 = S.n2[q];
    if (m1 == 0 || m2 == 0) continue;
    const int b1 = S.q1[q], b2 = S.q2[q];
    const int s1[4] = {0, Ly.bnp[b1], Ly.bnp[b1] + Ly.bnb[b1], m1};
    const int s2[4] = {0, Ly.bnp[b2], Ly.bnp[b2] + Ly.bnb[b2], m2};
    static const int sea1[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2}, sea2[9] = {0, 1, 2, 1, 2, 2, 0, 0, 1};
    for (int g = 0; g < 9; g++) {
      const int r0 = s1[sea1[g]], nr = s1[sea1[g] + 1] - r0, c0 = s2[sea2[g]], nc = s2[sea2[g] + 1] - c0;
      for (int jc = 0; jc < nc; jc++)
        for (int ir = 0; ir < nr; ir++) V[S.off[q] + (long long)(r0 + ir) * m2 + c0 + jc] = v[k++];
    }
  }
}
static void image_to_stream(const Layout &Ly, const TensorSpace &S, const double *img, double *v) {
  const int A = Ly.nhole, P = Ly.npart, n = Ly.nsp;
  long long k = 0;
  v[k++] = img[0];
  const double *F = img + Ly.f_off;
  for (int j = 0; j < A; j++)
    for (int i = 0; i < A; i++) v[k++] = F[(size_t)Ly.holes[i] * n + Ly.holes[j]];
  for (int j = 0; j < P; j++)
    for (int i = 0; i < P; i++) v[k++] = F[(size_t)Ly.parts[i] * n + Ly.parts[j]];
  for (int j = 0; j < A; j++)
    for (int i = 0; i < P; i++) v[k++] = F[(size_t)Ly.parts[i] * n + Ly.holes[j]];
  const double *V = img + S.v_off;
  for (int q = 0; q < S.nblk; q++) {
    const int m1 = S.n1[q], m2 = S.n2[q];
    if (m1 == 0 || m2 == 0) continue;
    const int b1 = S.q1[q], b2 = S.q2[q];
    const int s1[4] = {0, Ly.bnp[b1], Ly.bnp[b1] + Ly.bnb[b1], m1};
    const int s2[4] = {0, Ly.bnp[b2], Ly.bnp[b2] + Ly.bnb[b2], m2};
    static const int sea1[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2}, sea2[9] = {0, 1, 2, 1, 2, 2, 0, 0, 1};
    for (int g = 0; g < 9; g++) {
      const int r0 = s1[sea1[g]], nr = s1[sea1[g] + 1] - r0, c0 = s2[sea2[g]], nc = s2[sea2[g] + 1] - c0;
      for (int jc = 0; jc < nc; jc++)
        for (int ir = 0; ir < nr; ir++) v[k++] = V[S.off[q] + (long long)(r0 + ir) * m2 + c0 + jc];
    }
  }
}

int imsrg_tensor_upload(imsrg_tensor t, const double *vec, int64_t n) {
  if (!t || !vec) return fail(IMSRG_EINVAL, "null argument");
  const TensorSpace &S = *t->ts;
  if (n != S.neq) return fail(IMSRG_EINVAL, "tensor stream length " + std::to_string(n) + " != " + std::to_string(S.neq));
  std::vector<double> img(S.op_size);
  stream_to_image(t->ctx->lay, S, t->herm, vec, img.data());
  HIPCHK(hipMemcpyAsync(t->d, img.data(), sizeof(double) * S.op_size, hipMemcpyHostToDevice, t->ctx->stream));
  HIPCHK(hipStreamSynchronize(t->ctx->stream));
  return 0;
}

int imsrg_tensor_download(imsrg_tensor t, double *vec, int64_t n) {
  if (!t || !vec) return fail(IMSRG_EINVAL, "null argument");
  const TensorSpace &S = *t->ts;
  if (n != S.neq) return fail(IMSRG_EINVAL, "tensor stream length " + std::to_string(n) + " != " + std::to_string(S.neq));
  std::vector<double> img(S.op_size);
  HIPCHK(hipMemcpyAsync(img.data(), t->d, sizeof(double) * S.op_size, hipMemcpyDeviceToHost, t->ctx->stream));
  HIPCHK(hipStreamSynchronize(t->ctx->stream));
  image_to_stream(t->ctx->lay, S, img.data(), vec);
  return 0;
}

static int same_space(imsrg_tensor a, imsrg_tensor b) {
  if (!a || !b) return fail(IMSRG_EINVAL, "null argument");
  if (a->ts != b->ts || a->ctx != b->ctx) return fail(IMSRG_EINVAL, "tensors of different rank, parity or context");
  return 0;
}

int imsrg_tensor_copy(imsrg_tensor src, imsrg_tensor dst) {
  if (int rc = same_space(src, dst)) return rc;
  HIPCHK(hipMemcpyAsync(dst->d, src->d, sizeof(double) * src->ts->op_size, hipMemcpyDeviceToDevice, src->ctx->stream));
  dst->herm = src->herm;
  return 0;
}

int imsrg_tensor_clear(imsrg_tensor t) {
  if (!t) return fail(IMSRG_EINVAL, "null argument");
  launch_fill(t->ctx->stream, t->d, t->ts->op_size, 0.0);
  return 0;
}

int imsrg_tensor_add(imsrg_tensor A, double ax, imsrg_tensor B, double bx, imsrg_tensor C) {
  if (int rc = same_space(A, B)) return rc;
  if (int rc = same_space(A, C)) return rc;
  launch_axpby(A->ctx->stream, A->ts->op_size, ax, A->d, bx, B->d, C->d, nullptr, 1, 0);
  return 0;
}

static int tnorm(imsrg_tensor t, double *out) {
  imsrg_ctx_s *c = t->ctx;
  TensorSpace &S = *t->ts;
  hipLaunchKernelGGL(k_tnorm, dim3(kNormParts), dim3(256), 0, c->stream, c->dl, S.dt, (const double *)t->d, S.part);
  HIPCHK(hipMemcpyAsync(S.hpart, S.part, kNormParts * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double s = 0.0;
  for (int k = 0; k < kNormParts; k++) s += S.hpart[k];
  *out = std::sqrt(s);
  return 0;
}

int imsrg_tensor_norm(imsrg_tensor t, double *out) {
  if (!t || !out) return fail(IMSRG_EINVAL, "null argument");
  return tnorm(t, out);
}

// [G, T] into RES in BCH_TENSOR's call order (IMSRG_MAGNUS.f90:427-446):
// 111 (overwrites the one-body part), 121, 122 (overwrites the two-body
// part), 212, 222_pp_hh, 221, 211, 222_ph; the others accumulate.  221 reads
// the pp/hh ladder products, computed here whenever 221 is asked for.  Gd:
// the scalar operator's buffer (may be null when only 211 / 222_ph run on
// Pandya matrices already in S.gcc, gcc_ready).
static int ts_comm(imsrg_ctx_s *c, const double *Gd, int hG, imsrg_tensor T, imsrg_tensor RES, unsigned mask,
                   bool gcc_ready) {
  TensorSpace &S = *T->ts;
  const DevLayout &L = c->dl;
  const DevTensor &D = S.dt;
  hipStream_t s = c->stream;
  const int hL = hG, hR = T->herm, hres = RES->herm;
  const double *GF = Gd ? Gd + c->lay.f_off : nullptr, *GV = Gd ? Gd + c->lay.v_off : nullptr;
  const double *RF = T->d + c->lay.f_off, *RV = T->d + S.v_off;
  double *OF = RES->d + c->lay.f_off, *OV = RES->d + S.v_off;
  if (!Gd && (mask & ~(IMSRG_TS211 | IMSRG_TS222PH))) return fail(IMSRG_EINVAL, "terms need the scalar operator");
  const long long nf = (long long)D.nhole * D.nhole + (long long)D.npart * D.nhole + (long long)D.npart * D.npart;
  if (mask & IMSRG_TS111)
    hipLaunchKernelGGL(k_t111, dim3(nblk(nf)), dim3(256), 0, s, L, D, GF, RF, OF, hL, hR, hres);
  if (mask & IMSRG_TS121)
    hipLaunchKernelGGL(k_t_onebody, dim3(L.nob), dim3(64), 0, s, L, D, 0, GF, RF, RV, (const double *)S.gcc,
                       (const double *)S.w1, (const double *)S.w2, OF, hL, hR, hres);
  if (S.nelem > 0 && (mask & IMSRG_TS122))
    hipLaunchKernelGGL(k_t122, dim3(nblk(S.nelem)), dim3(256), 0, s, L, D, GF, RV, OV, hR);
  if (S.nelem > 0 && (mask & IMSRG_TS212))
    hipLaunchKernelGGL(k_t212, dim3(nblk(S.nelem)), dim3(256), 0, s, L, D, GV, RF, OV);
  if ((mask & (IMSRG_TS222PPHH | IMSRG_TS221)) && S.n_pp_tiles > 0) {
    hipLaunchKernelGGL(k_tgemm, dim3(S.n_pp_tiles), dim3(256), 0, s, S.pp_probs[0], S.pp_tiles, S.n_pp_tiles, GV, RV,
                       S.w1);
    hipLaunchKernelGGL(k_tgemm, dim3(S.n_pp_tiles), dim3(256), 0, s, S.pp_probs[1], S.pp_tiles, S.n_pp_tiles, GV, RV,
                       S.w2);
  }
  if (mask & IMSRG_TS222PPHH) {
    launch_axpby(s, S.vsize, 1.0, OV, 1.0, S.w1, OV, nullptr, 1, 0);
    launch_axpby(s, S.vsize, 1.0, OV, -1.0, S.w2, OV, nullptr, 1, 0);
  }
  if (mask & IMSRG_TS221)
    hipLaunchKernelGGL(k_t_onebody, dim3(L.nob), dim3(64), 0, s, L, D, 2, GF, RF, RV, (const double *)S.gcc,
                       (const double *)S.w1, (const double *)S.w2, OF, hL, hR, hres);
  if ((mask & (IMSRG_TS211 | IMSRG_TS222PH)) && !gcc_ready) launch_pandya(s, L, Gd, S.gcc, nullptr, 1);
  if (mask & IMSRG_TS211)
    hipLaunchKernelGGL(k_t_onebody, dim3(L.nob), dim3(64), 0, s, L, D, 1, GF, RF, RV, (const double *)S.gcc,
                       (const double *)S.w1, (const double *)S.w2, OF, hL, hR, hres);
  if ((mask & IMSRG_TS222PH) && S.nelem > 0) {
    if (D.nx > 0) hipLaunchKernelGGL(k_tpandya, dim3(nblk(D.nx)), dim3(256), 0, s, L, D, 0, RV, S.ccx, hR);
    if (D.nr > 0) hipLaunchKernelGGL(k_tpandya, dim3(nblk(D.nr)), dim3(256), 0, s, L, D, 1, RV, S.ccr, hR);
    if (S.n_cc_tiles[0] > 0)
      hipLaunchKernelGGL(k_tgemm, dim3(S.n_cc_tiles[0]), dim3(256), 0, s, S.cc_probs[0], S.cc_tiles[0],
                         S.n_cc_tiles[0], (const double *)S.gcc, (const double *)S.ccr, S.wx);
    if (S.n_cc_tiles[1] > 0)
      hipLaunchKernelGGL(k_tgemm, dim3(S.n_cc_tiles[1]), dim3(256), 0, s, S.cc_probs[1], S.cc_tiles[1],
                         S.n_cc_tiles[1], (const double *)S.ccx, (const double *)S.gcc, S.wy);
    for (int k = 0; k < 4; k++)
      if (D.qrun[k + 1] > D.qrun[k])
        hipLaunchKernelGGL(k_t222ph_q, dim3((unsigned)(D.qrun[k + 1] - D.qrun[k])), dim3(64),
                           (unsigned)(sizeof(PhSlot) * D.qcap[k]), s, L, D, (const double *)S.wx,
                           (const double *)S.wy, OV, hL, hR, k);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_tensor_commutator(imsrg_op G, imsrg_tensor T, imsrg_tensor RES, unsigned mask) {
  if (!G || !T || !RES) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = same_space(T, RES)) return rc;
  if (G->ctx != T->ctx) return fail(IMSRG_EINVAL, "G and T belong to different contexts");
  if (T == RES) return fail(IMSRG_EINVAL, "RES must differ from T");
  return ts_comm(G->ctx, G->d, G->herm, T, RES, mask, false);
}

int imsrg_tensor_commutator_cc(imsrg_tensor T, const double *gcc, int hG, int64_t n, imsrg_tensor RES,
                               unsigned mask) {
  if (!T || !gcc || !RES) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = same_space(T, RES)) return rc;
  if (T == RES) return fail(IMSRG_EINVAL, "RES must differ from T");
  if (hG != 1 && hG != -1) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  if (mask & ~(IMSRG_TS211 | IMSRG_TS222PH)) return fail(IMSRG_EINVAL, "only 211 and 222_ph run on Pandya matrices");
  imsrg_ctx_s *c = T->ctx;
  if (n != imsrg_cross_coupled_size(c)) return fail(IMSRG_EINVAL, "Pandya stream size mismatch");
  // the reference's row order (all ordered pairs, first orbit major) into the
  // device's paired rows; the zero rows stay zero
  const Layout &Ly = c->lay;
  std::vector<double> img(Ly.ccsize, 0.0);
  int64_t o = 0;
  for (int q = 0; q < Ly.ncc; q++) {
    const int rr = Ly.ccrlen_ref[q], rd = Ly.ccrlen[q], nb = Ly.ccnph[q];
    const int *ref = Ly.ccref.data() + Ly.ccrefbase[q];
    for (int col = 0; col < nb; col++)
      for (int r = 0; r < rr; r++) img[Ly.ccoff[q] + (int64_t)col * rd + ref[r]] = gcc[o++];
  }
  HIPCHK(hipMemcpyAsync(T->ts->gcc, img.data(), sizeof(double) * Ly.ccsize, hipMemcpyHostToDevice, c->stream));
  const int rc = ts_comm(c, nullptr, hG, T, RES, mask, true);
  HIPCHK(hipStreamSynchronize(c->stream));  // img is a local buffer
  return rc;
}

int imsrg_tensor_ladders(imsrg_op G, imsrg_tensor T, imsrg_tensor W1, imsrg_tensor W2) {
  if (!G || !T || !W1 || !W2) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = same_space(T, W1)) return rc;
  if (int rc = same_space(T, W2)) return rc;
  if (G->ctx != T->ctx) return fail(IMSRG_EINVAL, "G and T belong to different contexts");
  imsrg_ctx_s *c = T->ctx;
  TensorSpace &S = *T->ts;
  hipStream_t s = c->stream;
  const double *GV = G->d + c->lay.v_off, *RV = T->d + S.v_off;
  launch_fill(s, W1->d, S.op_size, 0.0);
  launch_fill(s, W2->d, S.op_size, 0.0);
  if (S.n_pp_tiles > 0) {
    hipLaunchKernelGGL(k_tgemm, dim3(S.n_pp_tiles), dim3(256), 0, s, S.pp_probs[0], S.pp_tiles, S.n_pp_tiles, GV, RV,
                       W1->d + S.v_off);
    hipLaunchKernelGGL(k_tgemm, dim3(S.n_pp_tiles), dim3(256), 0, s, S.pp_probs[1], S.pp_tiles, S.n_pp_tiles, GV, RV,
                       W2->d + S.v_off);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_tensor_commutator_221(imsrg_tensor W1, imsrg_tensor W2, int pm, imsrg_tensor RES) {
  if (!W1 || !W2 || !RES) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = same_space(W1, RES)) return rc;
  if (int rc = same_space(W2, RES)) return rc;
  if (pm != 1 && pm != -1) return fail(IMSRG_EINVAL, "pm must be +1 or -1");
  imsrg_ctx_s *c = RES->ctx;
  TensorSpace &S = *RES->ts;
  double *OF = RES->d + c->lay.f_off;
  const double *RV = RES->d + S.v_off;  // not read by the 221 kernel
  hipLaunchKernelGGL(k_t_onebody, dim3(c->dl.nob), dim3(64), 0, c->stream, c->dl, S.dt, 2, (const double *)OF,
                     (const double *)OF, RV, (const double *)S.gcc, (const double *)(W1->d + S.v_off),
                     (const double *)(W2->d + S.v_off), OF, pm, 1, RES->herm);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_bch_tensor(imsrg_op G, imsrg_tensor HS, int *nterms) {
  if (!G || !HS) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = HS->ctx;
  if (G->ctx != c) return fail(IMSRG_EINVAL, "G and HS belong to different contexts");
  TensorSpace &S = *HS->ts;
  imsrg_tensor_s INT2, AD;
  INT2.ctx = AD.ctx = c;
  INT2.ts = AD.ts = &S;
  double *buf = nullptr;
  if (hipMalloc(&buf, sizeof(double) * S.op_size * 2) != hipSuccess)
    return fail(IMSRG_ENOMEM, "hipMalloc BCH_TENSOR workspace failed");
  INT2.d = buf;
  AD.d = buf + S.op_size;
  // INT2 = HS (copy_sq_op keeps HS's herm, IMSRG_MAGNUS.f90:419-422)
  int rc = imsrg_tensor_copy(HS, &INT2);
  const double conv = (double)1e-6f;  // real(8), parameter :: conv = 1e-6
  double coef = 1.0;
  int n = 0;
  for (int iw = 2; iw <= 30 && rc == 0; iw++) {
    coef = coef / (iw - 1.0);
    if ((rc = imsrg_tensor_copy(&INT2, &AD))) break;
    launch_fill(c->stream, INT2.d, S.op_size, 0.0);
    if ((rc = ts_comm(c, G->d, G->herm, &AD, &INT2, IMSRG_TSALL, false))) break;
    n++;
    launch_axpby(c->stream, S.op_size, 1.0, HS->d, coef, INT2.d, HS->d, nullptr, 1, 0);
    double nrm = 0.0;
    if ((rc = tnorm(&INT2, &nrm))) break;
    if (nrm * coef < conv) break;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  if (nterms) *nterms = n;
  return rc;
}
