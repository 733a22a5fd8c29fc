// interaction_io.cpp -- two-body interaction files into an operator's host
// image (the device layout of layout.hpp: [E0 | pad][F nsp x nsp][V blocks]).
//
// The four readers of im-SRG/interaction_IO.f90, each setting the matrix
// elements its file lists the way the reference's reader sets them (the
// element and its hermitian mirror, last write wins, everything else kept):
//   ascii  read_interaction        interaction_IO.f90:8-107
//   gz     read_gz                 :110-231
//   me2j   read_me2j_interaction   :282-568 (isospin-coupled file -> pn)
//   me2b   read_me2b_interaction   :570-1151 (two-body file, then E0 and the
//                                  one-body part from the matching .me1b.gz)
// and the centre-of-mass terms r1_r2 / p1_p2 (basic_IMSRG.f90:4310-4530) that
// the first three add when the Lawson beta or ham_type 1 asks for them.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "layout.hpp"

namespace imsrg {

namespace {

// one line of a gzip (or plain) text file; false at the end
struct GzLines {
  gzFile f = nullptr;
  explicit GzLines(const char *path) { f = gzopen(path, "rb"); }
  ~GzLines() {
    if (f) gzclose(f);
  }
  bool next(std::string &s) {
    char buf[4096];
    if (!f || !gzgets(f, buf, sizeof(buf))) return false;
    s.assign(buf);
    while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
    return true;
  }
};

// a fixed-width integer field as Fortran's I edit descriptor reads it
// (blanks ignored, an all-blank field is 0)
int fixed_int(const std::string &s, size_t pos, size_t len, bool &ok) {
  ok = false;
  std::string f = pos < s.size() ? s.substr(pos, len) : std::string();
  size_t k = 0;
  while (k < f.size() && f[k] == ' ') k++;
  if (k == f.size()) {
    ok = true;
    return 0;
  }
  char *end = nullptr;
  const long v = std::strtol(f.c_str() + k, &end, 10);
  for (const char *c = end; *c; c++)
    if (*c != ' ') return 0;
  ok = end != f.c_str() + k;
  return (int)v;
}

// the operator's host image and the reference's element placement: orbits
// 0-based, J doubled; `pre` collects the caller's factors, the exchange phase
// of a swapped pair is applied here (interaction_IO.f90:59-85)
struct Target {
  const Layout &L;
  double *img;
  std::string err;
  Target(const Layout &l, double *h) : L(l), img(h) {}
  bool set(int a, int b, int c, int d, int J, double V) {
    double pre = 1.0;
    const int N = L.nsp;
    if (a < 0 || b < 0 || c < 0 || d < 0 || a >= N || b >= N || c >= N || d >= N) {
      err = "orbit index out of range";
      return false;
    }
    if (a > b) {
      pre *= ((1 + (L.jj[a] + L.jj[b] - J) / 2) & 1) ? -1.0 : 1.0;
      std::swap(a, b);
    }
    if (c > d) {
      pre *= ((1 + (L.jj[c] + L.jj[d] - J) / 2) & 1) ? -1.0 : 1.0;
      std::swap(c, d);
    }
    const int T = (L.tz[a] + L.tz[b]) / 2, P = (L.l[a] + L.l[b]) & 1;
    if ((L.tz[c] + L.tz[d]) / 2 != T || ((L.l[c] + L.l[d]) & 1) != P || J < 0 || J / 2 > L.jtot_max) {
      err = "matrix element between different (J, parity, Tz) channels";
      return false;
    }
    const size_t nn2 = (size_t)N * N;
    const int r1 = L.pidx[(size_t)(J / 2) * nn2 + (size_t)a * N + b];
    const int r2 = L.pidx[(size_t)(J / 2) * nn2 + (size_t)c * N + d];
    if (r1 < 0 || r2 < 0) {
      err = "pair not coupled to this J";
      return false;
    }
    const int q = (J / 2) * 6 + 2 * (T + 1) + P;
    const int dim = L.bdim[q];
    double *B = img + L.v_off + L.boff[q];
    B[(size_t)r1 * dim + r2] = V * pre;
    B[(size_t)r2 * dim + r1] = V * pre;
    return true;
  }
};

// <a || r || b> and <a || p || b> of the harmonic oscillator in units of b
// (reduced_r / reduced_p, basic_IMSRG.f90:4436-4530)
double reduced_rp(const Layout &L, int a, int b, bool p) {
  const int la = L.l[a], lb = L.l[b], na = L.n[a], nb = L.n[b];
  double c1 = 0.0, v = 0.0;
  if (la == lb + 1) {
    c1 = std::sqrt(lb + 1.0);
    if (na == nb) v = (p ? -1.0 : 1.0) * std::sqrt(nb + lb + 1.5);
    else if (na == nb - 1) v = -std::sqrt(nb * 1.0);
  } else if (la == lb - 1) {
    c1 = -std::sqrt(1.0 * lb);
    if (na == nb) v = std::sqrt(nb + lb + 0.5);
    else if (na == nb + 1) v = (p ? 1.0 : -1.0) * std::sqrt(nb + 1.0);
  }
  return c1 * v;
}

// the one-body factor of orbits x, y for the rank-1 operator (r or p)
double rank1(const Layout &L, int x, int y, bool p) {
  if (L.tz[x] != L.tz[y]) return 0.0;
  const int jx = L.jj[x], jy = L.jj[y], lx = L.l[x], ly = L.l[y];
  const double ph = (((3 + 2 * lx + jy) / 2) & 1) ? -1.0 : 1.0;
  return std::sqrt((jx + 1.0) * (jy + 1.0)) * ph * racah6j(jx, 2, jy, 2 * ly, 1, 2 * lx) * reduced_rp(L, x, y, p);
}

// r1.r2 (times b^-2) or p1.p2 (times b^2) between normalised antisymmetric pair
// states (r1_r2 / p1_p2, basic_IMSRG.f90:4310-4434); the exchange phase
// (-1)^((jb -/+ jc - J)/2) carries the sign difference of the two
double one_dot_two(const Layout &L, int a, int b, int c, int d, int J, bool p) {
  const int ja = L.jj[a], jb = L.jj[b], jc = L.jj[c], jd = L.jj[d];
  const double jac = rank1(L, a, c, p), jbd = rank1(L, b, d, p), jad = rank1(L, a, d, p), jbc = rank1(L, b, c, p);
  auto sgn = [](int x) { return ((x / 2) & 1) ? -1.0 : 1.0; };
  const int s = p ? -1 : 1;
  const double direct = sgn(jb + s * jc - J + 4 * 64) * racah6j(ja, jb, J, jd, jc, 2) * jac * jbd;
  const double exch = sgn(jc + jd - J + 4 * 64) * sgn(jb + s * jd - J + 4 * 64) * racah6j(ja, jb, J, jc, jd, 2) * jad * jbc;
  return (direct - exch) / std::sqrt((1.0 + (a == b)) * (1.0 + (c == d)));
}

// V + (g1 - g3*COM) hw / A with g1 = (g3 + com_hw^2/hw^2 g2) beta (the
// readers' centre-of-mass and Lawson terms, interaction_IO.f90:44-51)
double com_term(const Layout &L, const ReadParams &p, int a, int b, int c, int d, int J) {
  if (p.lawson_beta == 0.0 && p.htype != 1) return 0.0;
  const double g2 = one_dot_two(L, a, b, c, d, J, false), g3 = one_dot_two(L, a, b, c, d, J, true);
  const double g1 = (g3 + p.com_hw * p.com_hw / (p.hw * p.hw) * g2) * p.lawson_beta;
  const double com = p.htype == 1 ? 1.0 : 0.0;
  return (g1 - g3 * com) * p.hw / p.A;
}

// ---------------------------------------------------------------- ascii / gz
// A line "Tz Par J a b c d V" (orbits 1-based in the basis order, J doubled):
// the channel labels are checked against the orbits, V gets the COM terms.
bool morten_line(Target &t, const ReadParams &p, int Tz, int Par, int J, int a, int b, int c, int d, double V) {
  const Layout &L = t.L;
  a--, b--, c--, d--;
  if (a < 0 || b < 0 || c < 0 || d < 0 || a >= L.nsp || b >= L.nsp || c >= L.nsp || d >= L.nsp) {
    t.err = "orbit label out of range";
    return false;
  }
  if ((L.tz[a] + L.tz[b]) / 2 != Tz || ((L.l[a] + L.l[b]) & 1) != Par) {
    t.err = "Tz / parity label does not match the orbits";
    return false;
  }
  V += com_term(L, p, a, b, c, d, J);
  return t.set(a, b, c, d, J, V);
}

int read_ascii(Target &t, const ReadParams &p, const char *path) {
  FILE *f = std::fopen(path, "r");
  if (!f) {
    t.err = std::string(path) + ": cannot open";
    return -1;
  }
  char line[4096];
  for (int k = 0; k < 8; k++)  // eight header lines (interaction_IO.f90:22-23)
    if (!std::fgets(line, sizeof(line), f)) break;
  int rc = 0;
  while (std::fgets(line, sizeof(line), f)) {
    int Tz, Par, J, a, b, c, d;
    double V;
    if (std::sscanf(line, "%d %d %d %d %d %d %d %lf", &Tz, &Par, &J, &a, &b, &c, &d, &V) != 8) {
      bool blank = true;
      for (char *s = line; *s; s++) blank = blank && (*s == ' ' || *s == '\n' || *s == '\r' || *s == '\t');
      if (blank) continue;
      t.err = std::string(path) + ": malformed line";
      rc = -1;
      break;
    }
    if (!morten_line(t, p, Tz, Par, J, a, b, c, d, V)) {
      rc = -1;
      break;
    }
  }
  std::fclose(f);
  return rc;
}

// gz variant (read_gz :110-231), parsed by position as the reference does:
// the first line holds the number of elements ending two characters before
// the marker XXXX; each element line (its first 52 characters) is left-
// adjusted, cut two characters before its marker (searched from column 20)
// and re-padded with two blanks if it starts with '-', else three; then
// Tz, Par, J, a, b, c, d are I4 fields at columns 1-28 and V an E20.6 field
// at columns 29-48
int read_gzip(Target &t, const ReadParams &p, const char *path) {
  GzLines in(path);
  std::string s;
  if (!in.f || !in.next(s)) {
    t.err = std::string(path) + ": cannot open";
    return -1;
  }
  auto marker = [](const std::string &x, size_t from, size_t to) -> size_t {  // 0-based start, or npos
    for (size_t k = from; k <= to && k + 4 <= x.size(); k++)
      if (x.compare(k, 4, "XXXX") == 0) return k;
    return std::string::npos;
  };
  const size_t m0 = marker(s, 2, 19);
  if (m0 == std::string::npos) {
    t.err = std::string(path) + ": no count on the first line";
    return -1;
  }
  bool ok;
  const long long ntot = fixed_int(s.substr(0, m0 - 1), 0, m0 - 1, ok);
  if (!ok) {
    t.err = std::string(path) + ": bad count";
    return -1;
  }
  for (long long k = 0; k < ntot; k++) {
    if (!in.next(s)) {
      t.err = std::string(path) + ": fewer matrix elements than announced";
      return -1;
    }
    std::string a = s.substr(0, 52);
    const size_t lead = a.find_first_not_of(' ');
    a = lead == std::string::npos ? std::string() : a.substr(lead);
    const size_t m = marker(a, 19, 79);
    if (m == std::string::npos || m < 1) {
      t.err = std::string(path) + ": no marker on an element line";
      return -1;
    }
    const std::string fixed = (a[0] == '-' ? "  " : "   ") + a.substr(0, m - 1);
    int f[7];
    bool good = true;
    for (int q = 0; q < 7; q++) {
      f[q] = fixed_int(fixed, 4 * q, 4, ok);
      good = good && ok;
    }
    const std::string vf = fixed.size() > 28 ? fixed.substr(28, 20) : std::string();
    char *end = nullptr;
    const double V = std::strtod(vf.c_str(), &end);
    if (!good || end == vf.c_str()) {
      t.err = std::string(path) + ": malformed element line";
      return -1;
    }
    if (!morten_line(t, p, f[0], f[1], f[2], f[3], f[4], f[5], f[6], V)) return -1;
  }
  return 0;
}

// ---------------------------------------------------------------- me2j
// The file's basis: harmonic-oscillator orbits in the hk order up to its own
// eMax and lMax, each (n, l, j) as a proton then a neutron (the .sps file of
// get_me2j_spfile, interaction_IO.f90:234-279).  The matrix elements are
// isospin coupled, four (T=0, then T=1 with M_T = -1, 0, +1 in the file's
// labelling) per (nlj1 >= nlj2, nnlj1 >= nnlj2, J) in the reader's loop
// order, ten values per line.
int read_me2j(Target &t, const ReadParams &p, const char *path) {
  const Layout &L = t.L;
  struct Orb { int n, l, jj, tz, to; };
  std::vector<Orb> fb;
  for (int e = 0; e <= p.emax_file; e++)
    for (int l = e % 2; l <= std::min(e, p.lmax_file); l += 2)
      for (int jj = std::abs(2 * l - 1); jj <= 2 * l + 1; jj += 2)
        for (int tz = -1; tz <= 1; tz += 2) {
          Orb o{(e - l) / 2, l, jj, tz, -1};
          for (int i = 0; i < L.nsp; i++)
            if (L.n[i] == o.n && L.l[i] == o.l && L.jj[i] == o.jj && L.tz[i] == o.tz) o.to = i;
          fb.push_back(o);
        }
  int lmax = 0, emax = 0;  // the calculation's cuts: the context's basis
  for (int i = 0; i < L.nsp; i++) {
    lmax = std::max(lmax, L.l[i]);
    emax = std::max(emax, 2 * L.n[i] + L.l[i]);
  }
  const int nf = (int)fb.size() / 2;
  // the number of values the file holds for this basis (the reader's first
  // loop), then the values: ten per line, in an f12.7 field and nine f13.7
  long long imax = 0;
  for (int k1 = 0; k1 < nf; k1++)
    for (int k2 = 0; k2 <= k1; k2++)
      for (int k3 = 0; k3 <= k1; k3++)
        for (int k4 = 0; k4 <= (k1 == k3 ? k2 : k3); k4++) {
          const Orb &o1 = fb[2 * k1], &o2 = fb[2 * k2], &o3 = fb[2 * k3], &o4 = fb[2 * k4];
          if (((o1.l + o2.l) & 1) != ((o3.l + o4.l) & 1)) continue;
          const int jmin = std::max(std::abs(o1.jj - o2.jj), std::abs(o3.jj - o4.jj));
          const int jmax = std::min(o1.jj + o2.jj, o3.jj + o4.jj);
          if (jmin <= jmax) imax += 4 * ((jmax - jmin) / 2 + 1);
        }
  std::vector<double> me(imax);
  {
    GzLines in(path);
    std::string s;
    if (!in.f || !in.next(s)) {  // header line
      t.err = std::string(path) + ": cannot open";
      return -1;
    }
    for (long long i = 0; i < imax; i += 10) {
      if (!in.next(s)) {
        t.err = std::string(path) + ": fewer matrix elements than the basis needs";
        return -1;
      }
      const int nv = (int)std::min<long long>(10, imax - i);
      for (int k = 0; k < nv; k++) {
        const size_t pos = k == 0 ? 0 : 12 + 13 * (k - 1), len = k == 0 ? 12 : 13;
        if (pos >= s.size()) {
          t.err = std::string(path) + ": short line";
          return -1;
        }
        me[i + k] = std::atof(s.substr(pos, len).c_str());
      }
    }
  }
  long long i = 0;
  for (int k1 = 0; k1 < nf; k1++)
    for (int k2 = 0; k2 <= k1; k2++)
      for (int k3 = 0; k3 <= k1; k3++)
        for (int k4 = 0; k4 <= (k1 == k3 ? k2 : k3); k4++) {
          const Orb &o1 = fb[2 * k1], &o2 = fb[2 * k2], &o3 = fb[2 * k3], &o4 = fb[2 * k4];
          if (((o1.l + o2.l) & 1) != ((o3.l + o4.l) & 1)) continue;
          const int jmin = std::max(std::abs(o1.jj - o2.jj), std::abs(o3.jj - o4.jj));
          const int jmax = std::min(o1.jj + o2.jj, o3.jj + o4.jj);
          if (jmin > jmax) continue;
          for (int J = jmin; J <= jmax; J += 2) {
            const double m0 = me[i], m1 = me[i + 1], m2 = me[i + 2], m3 = me[i + 3];
            i += 4;
            if (o1.l > lmax || o2.l > lmax || o3.l > lmax || o4.l > lmax) continue;
            if (2 * o1.n + o1.l > emax || 2 * o2.n + o2.l > emax) continue;
            if (2 * o3.n + o3.l > emax || 2 * o4.n + o4.l > emax) continue;
            for (int ax = 0; ax < 2; ax++)
              for (int bx = 0; bx < 2; bx++)
                for (int cx = 0; cx < 2; cx++)
                  for (int dx = 0; dx < 2; dx++) {
                    const int a = fb[2 * k1 + ax].to, b = fb[2 * k2 + bx].to;
                    const int c = fb[2 * k3 + cx].to, d = fb[2 * k4 + dx].to;
                    if (a < 0 || b < 0 || c < 0 || d < 0) {
                      t.err = "me2j: an orbit within the cuts is missing from the basis";
                      return -1;
                    }
                    double pre2 = 1.0;
                    if (a == b) pre2 *= std::sqrt(0.5);
                    if (c == d) pre2 *= std::sqrt(0.5);
                    // the file's isospin labels are the negatives of the basis'
                    const int ta = -L.tz[a], tb = -L.tz[b], tc = -L.tz[c], td = -L.tz[d];
                    if (ta + tb != tc + td) continue;
                    const double V0 = 0.125 * (ta - tb) * (tc - td) * m0 +
                                      ((ta + tb == -2 && tc + td == -2) ? m1 : 0.0) +
                                      ((ta + tb == 2 && tc + td == 2) ? m3 : 0.0) +
                                      0.125 * std::abs((ta - tb) * (tc - td)) * m2;
                    const double V = V0 * pre2 + com_term(L, p, a, b, c, d, J);
                    if (!t.set(a, b, c, d, J, V)) return -1;
                  }
          }
        }
  return 0;
}

// ---------------------------------------------------------------- me2b
// Two-body part: three header lines, one line per block (its size at
// columns 10-16), then per block a blank and a label line and its matrix
// elements "A  B  V" (0-based positions in the block's pair list, pairs
// hh, then ph, then pp, each in the file's (l,j) x n order) up to the last
// pair's diagonal element.  The file's Tz = +1 / 0 / -1 blocks hold orbit
// pairs of 2tz (+1,+1) / (+1,-1) / (-1,-1).  Then E0 and the one-body part
// from me1b (interaction_IO.f90:1054-1151).
int read_me2b(Target &t, const char *path, const char *path1b) {
  const Layout &L = t.L;
  int Lmax = 0, eMax = 0;
  for (int i = 0; i < L.nsp; i++) {
    Lmax = std::max(Lmax, L.l[i]);
    eMax = std::max(eMax, 2 * L.n[i] + L.l[i]);
  }
  std::vector<int> ljl, ljj, nmax;  // the (l, j) list and its largest n
  for (int twol = 0; twol <= 2 * Lmax; twol += 2)
    for (int twoj = std::abs(twol - 1); twoj <= twol + 1; twoj += 2) {
      ljl.push_back(twol / 2);
      ljj.push_back(twoj);
      nmax.push_back((eMax - twol / 2) / 2);
    }
  const int nlj = (int)ljl.size();
  auto find = [&](int n, int l, int jj, int tz) {
    for (int i = 0; i < L.nsp; i++)
      if (L.jj[i] == jj && L.n[i] == n && L.l[i] == l && L.tz[i] == tz) return i;
    return -1;
  };
  struct Blk { int J, P, Tz; std::vector<int> a, b; };
  std::vector<Blk> blocks;
  const int JTM = L.jtot_max;
  for (int Tz = 1; Tz >= -1; Tz--)
    for (int Pi = 0; Pi <= 1; Pi++)
      for (int JT = 0; JT <= 2 * JTM; JT += 2) {
        if (Lmax == eMax && JT == 2 * JTM && std::abs(Tz) == 1) continue;
        if (JT == 2 * JTM && Pi == 1) continue;
        const int t1 = Tz == -1 ? -1 : 1, t2 = Tz == 1 ? 1 : -1;
        Blk B{JT, Pi, Tz, {}, {}};
        std::vector<int> hh[2], ph[2], pp[2];
        for (int lj1 = 0; lj1 < nlj; lj1++)
          for (int lj2 = 0; lj2 < nlj; lj2++) {
            const int j1 = ljj[lj1], j2 = ljj[lj2], l1 = ljl[lj1], l2 = ljl[lj2];
            if (JT < std::abs(j1 - j2) || JT > j1 + j2) continue;
            if (((l1 + l2) & 1) != Pi) continue;
            for (int n1 = 0; n1 <= nmax[lj1]; n1++) {
              const int idx = lj1 * (nmax[0] + 1) + n1;
              for (int n2 = 0; n2 <= nmax[lj2]; n2++) {
                const int idxx = lj2 * (nmax[0] + 1) + n2;
                if (Tz != 0 && idx > idxx) continue;
                if (((JT / 2) & 1) && lj1 == lj2 && n1 == n2 && Tz != 0) continue;
                const int i = find(n1, l1, j1, t1), j = find(n2, l2, j2, t2);
                if (i < 0 || j < 0) {
                  t.err = "me2b: an orbit of the file's basis is missing from the context";
                  return -1;
                }
                const int cls = L.con[i] + L.con[j];
                std::vector<int> *dst = cls == 2 ? hh : cls == 1 ? ph : pp;
                dst[0].push_back(i);
                dst[1].push_back(j);
              }
            }
          }
        for (auto *v : {hh, ph, pp}) {
          B.a.insert(B.a.end(), v[0].begin(), v[0].end());
          B.b.insert(B.b.end(), v[1].begin(), v[1].end());
        }
        blocks.push_back(std::move(B));
      }
  GzLines in(path);
  std::string s;
  if (!in.f) {
    t.err = std::string(path) + ": cannot open";
    return -1;
  }
  for (int k = 0; k < 3 + (int)blocks.size(); k++)
    if (!in.next(s)) {
      t.err = std::string(path) + ": truncated header";
      return -1;
    }
  bool skip_label = true;
  for (const Blk &B : blocks) {
    const int ntot = (int)B.a.size();
    if (!in.next(s) || (skip_label && !in.next(s))) {
      t.err = std::string(path) + ": truncated";
      return -1;
    }
    skip_label = true;
    for (;;) {
      if (!in.next(s)) {
        if (ntot == 0) break;
        t.err = std::string(path) + ": truncated block";
        return -1;
      }
      if (s.empty() || s[0] < '0' || s[0] > '9') {  // the block ended without its last element
        skip_label = false;
        break;
      }
      // "A  B  V": A up to the first space, B from two past it to the next space
      const size_t s1 = s.find(' ');
      const size_t s2 = s.find(' ', s1 + 2);
      if (s1 == std::string::npos || s2 == std::string::npos) {
        t.err = std::string(path) + ": malformed line";
        return -1;
      }
      const int AA = std::atoi(s.substr(0, s1).c_str());
      const int BB = std::atoi(s.substr(s1 + 2, s2 - s1 - 2).c_str());
      const double V = std::atof(s.substr(s2 + 2, s[s2 + 2] == '-' ? 11 : 10).c_str());
      if (AA < 0 || BB < 0 || AA >= ntot || BB >= ntot) {
        t.err = std::string(path) + ": pair index beyond the block";
        return -1;
      }
      int a = B.a[AA], b = B.b[AA], c = B.a[BB], d = B.b[BB];
      double pre = 1.0;
      if (a <= b && a == b) pre /= std::sqrt(2.0);
      if (c <= d && c == d) pre /= std::sqrt(2.0);
      if (!t.set(a, b, c, d, B.J, V * pre)) return -1;
      if (AA == ntot - 1 && BB == ntot - 1) break;
    }
  }
  // me1b: version line, count, E0, then one line per one-body element
  GzLines in1(path1b);
  if (!in1.f || !in1.next(s) || !in1.next(s)) {
    t.err = std::string(path1b) + ": cannot open";
    return -1;
  }
  const int amax = std::atoi(s.substr(0, 4).c_str());
  if (!in1.next(s)) {
    t.err = std::string(path1b) + ": truncated";
    return -1;
  }
  {
    size_t w = 0;
    if (s.size() > 3 && s[3] == '.') w = 10;
    else if (s.size() > 4 && s[4] == '.') w = 11;
    else if (s.size() > 5 && s[5] == '.') w = 12;
    if (!w) {
      t.err = std::string(path1b) + ": E0 line not understood";
      return -1;
    }
    t.img[0] = std::atof(s.substr(0, w).c_str());
  }
  const int N = L.nsp;
  double *F = t.img + L.f_off;
  for (int k = 0; k < amax; k++) {
    if (!in1.next(s)) {
      t.err = std::string(path1b) + ": truncated";
      return -1;
    }
    bool ok1, ok2, ok3, ok4;
    const int tt = fixed_int(s, 1, 1, ok1), lj = fixed_int(s, 3, 2, ok2);
    const int n1 = fixed_int(s, 7, 2, ok3), n2 = fixed_int(s, 10, 2, ok4);
    if (!ok1 || !ok2 || !ok3 || !ok4 || lj < 0 || lj >= nlj) {
      t.err = std::string(path1b) + ": malformed line";
      return -1;
    }
    const double V = std::atof(s.substr(14, 14).c_str());
    const int t1 = -2 * tt + 1, l1 = ljl[lj], j1 = ljj[lj];
    const int i = find(n1, l1, j1, t1), j = find(n2, l1, j1, t1);
    if (i < 0 || j < 0) {
      t.err = "me1b: orbit missing from the context";
      return -1;
    }
    if (L.con[i] + L.con[j] == 2 || L.con[i] + L.con[j] == 0) {
      F[(size_t)i * N + j] = V;
    } else if (L.con[i] == 0 && L.con[j] == 1) {  // fph(p, h) and its mirror
      F[(size_t)i * N + j] = V;
      F[(size_t)j * N + i] = V;
    }
  }
  return 0;
}

}  // namespace

int read_interaction_file(const Layout &L, int format, const char *path, const char *path2, const ReadParams &p,
                          double *img, std::string &err) {
  Target t(L, img);
  int rc = -1;
  if ((format == 0 || format == 1 || format == 2) && (p.hw <= 0.0 || p.A <= 0) &&
      (p.lawson_beta != 0.0 || p.htype == 1)) {
    err = "centre-of-mass terms need hw > 0 and A > 0";
    return -1;
  }
  switch (format) {
    case 0: rc = read_ascii(t, p, path); break;
    case 1: rc = read_gzip(t, p, path); break;
    case 2: rc = read_me2j(t, p, path); break;
    case 3: rc = read_me2b(t, path, path2); break;
    default: t.err = "unknown interaction format";
  }
  err = t.err;
  return rc;
}

}  // namespace imsrg
