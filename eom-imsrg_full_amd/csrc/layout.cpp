// layout.cpp -- builds the channel layout, index tables and 6j table.
// Orbit/block/Pandya enumeration follows the reference's rules
// (basic_IMSRG.f90:477-641, cross_coupled.f90:61-177); the storage itself is
// this library's own (full square channel matrices, dense lookup tables).
#include "layout.hpp"

#include <algorithm>
#include <map>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace imsrg {

static inline bool tri(int j1, int j2, int J) { return J <= j1 + j2 && J >= std::abs(j1 - j2); }
static inline int64_t align32(int64_t x) { return (x + 31) & ~int64_t(31); }

// ---------------------------------------------------------------- 6j
static long double g_fact[256];
static bool g_fact_ready = false;
static void init_fact() {
  if (g_fact_ready) return;
  g_fact[0] = 1.0L;
  for (int i = 1; i < 256; i++) g_fact[i] = g_fact[i - 1] * (long double)i;
  g_fact_ready = true;
}
static bool tri_ok(int a, int b, int c) {
  return a >= 0 && b >= 0 && c >= 0 && c <= a + b && c >= std::abs(a - b) && ((a + b + c) & 1) == 0;
}
static long double tri_delta(int a, int b, int c) {
  return sqrtl(g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] * g_fact[(-a + b + c) / 2] /
               g_fact[(a + b + c) / 2 + 1]);
}
// Racah's single-sum formula for {j1 j2 j3; j4 j5 j6}, arguments doubled.
double racah6j(int j1, int j2, int j3, int j4, int j5, int j6) {
  init_fact();
  if (!tri_ok(j1, j2, j3) || !tri_ok(j1, j5, j6) || !tri_ok(j4, j2, j6) || !tri_ok(j4, j5, j3))
    return 0.0;
  const int a[4] = {(j1 + j2 + j3) / 2, (j1 + j5 + j6) / 2, (j4 + j2 + j6) / 2, (j4 + j5 + j3) / 2};
  const int b[3] = {(j1 + j2 + j4 + j5) / 2, (j2 + j3 + j5 + j6) / 2, (j3 + j1 + j6 + j4) / 2};
  int tmin = a[0], tmax = b[0];
  for (int k = 1; k < 4; k++) tmin = a[k] > tmin ? a[k] : tmin;
  for (int k = 1; k < 3; k++) tmax = b[k] < tmax ? b[k] : tmax;
  long double s = 0.0L;
  for (int t = tmin; t <= tmax; t++) {
    long double den = g_fact[t - a[0]] * g_fact[t - a[1]] * g_fact[t - a[2]] * g_fact[t - a[3]] *
                      g_fact[b[0] - t] * g_fact[b[1] - t] * g_fact[b[2] - t];
    long double term = g_fact[t + 1] / den;
    s += (t & 1) ? -term : term;
  }
  s *= tri_delta(j1, j2, j3) * tri_delta(j1, j5, j6) * tri_delta(j4, j2, j6) * tri_delta(j4, j5, j3);
  return (double)s;
}

// ---------------------------------------------------------------- layout
std::string Layout::build(int nsp_, const int *n_, const int *l_, const int *jj_, const int *tz_,
                          const int *con_) {
  if (nsp_ <= 0 || nsp_ > 4096) return "nsp out of range";
  nsp = nsp_;
  n.assign(n_, n_ + nsp);
  l.assign(l_, l_ + nsp);
  jj.assign(jj_, jj_ + nsp);
  tz.assign(tz_, tz_ + nsp);
  con.assign(con_, con_ + nsp);
  for (int i = 0; i < nsp; i++) {
    if (jj[i] <= 0 || (jj[i] & 1) == 0) return "2j must be a positive odd integer";
    if (tz[i] != -1 && tz[i] != 1) return "2tz must be -1 or +1";
    if (con[i] != 0 && con[i] != 1) return "con must be 0 or 1";
    if (l[i] < 0 || n[i] < 0) return "n and l must be non-negative";
  }
  holes.clear();
  parts.clear();
  for (int i = 0; i < nsp; i++) (con[i] ? holes : parts).push_back(i);
  nhole = (int)holes.size();
  npart = (int)parts.size();
  jtot_max = 0;
  lmax = 0;
  for (int i = 0; i < nsp; i++) {
    jtot_max = jj[i] > jtot_max ? jj[i] : jtot_max;
    lmax = l[i] > lmax ? l[i] : lmax;
  }
  if (jtot_max > 31) return "2j > 31 is not supported (ph term codes hold 5-bit J' indices)";
  // groups
  grp_of.assign(nsp, -1);
  grp_start.clear();
  grp_cnt.clear();
  grp_list.clear();
  for (int i = 0; i < nsp; i++) {
    if (grp_of[i] >= 0) continue;
    int g = (int)grp_start.size();
    grp_start.push_back((int)grp_list.size());
    grp_cnt.push_back(0);
    for (int k = i; k < nsp; k++)
      if (l[k] == l[i] && jj[k] == jj[i] && tz[k] == tz[i]) {
        grp_of[k] = g;
        grp_list.push_back(k);
        grp_cnt[g]++;
      }
  }
  // two-body blocks
  const int JTM = jtot_max;
  nblocks = (JTM + 1) * 6;
  bJ.assign(nblocks, 0); bP.assign(nblocks, 0); bTz.assign(nblocks, 0);
  bnp.assign(nblocks, 0); bnb.assign(nblocks, 0); bnh.assign(nblocks, 0); bdim.assign(nblocks, 0);
  boff.assign(nblocks, 0);
  pbase.assign(nblocks + 1, 0);
  const size_t nn2 = (size_t)nsp * nsp;
  pidx.assign((size_t)(JTM + 1) * nn2, -1);
  pair_a.clear(); pair_b.clear(); pair_q.clear();
  int q = 0;
  int64_t off = 0;
  for (int J = 0; J <= 2 * JTM; J += 2)
    for (int T = -1; T <= 1; T++)
      for (int P = 0; P <= 1; P++) {
        bJ[q] = J; bP[q] = P; bTz[q] = T;
        std::vector<int> lists[3];
        for (int i = 0; i < nsp; i++)
          for (int j = i; j < nsp; j++) {
            if (T != (tz[i] + tz[j]) / 2) continue;
            if (P != (l[i] + l[j]) % 2) continue;
            if (!tri(jj[i], jj[j], J)) continue;
            lists[con[i] + con[j]].push_back(i * nsp + j);
          }
        bnp[q] = (int)lists[0].size();
        bnb[q] = (int)lists[1].size();
        bnh[q] = (int)lists[2].size();
        bdim[q] = bnp[q] + bnb[q] + bnh[q];
        pbase[q] = (int)pair_a.size();
        int pos = 0;
        for (int c = 0; c < 3; c++)
          for (int code : lists[c]) {
            int a = code / nsp, b = code % nsp;
            pidx[(size_t)(J / 2) * nn2 + (size_t)a * nsp + b] = pos++;
            pair_a.push_back(a);
            pair_b.push_back(b);
            pair_q.push_back(q);
          }
        boff[q] = off;
        off = align32(off + (int64_t)bdim[q] * bdim[q]);
        q++;
      }
  pbase[nblocks] = (int)pair_a.size();
  npairs = (int)pair_a.size();
  vsize = off;
  // one-body outputs (p,q) with p<=q in hh and pp, every (particle, hole)
  ob_p.clear(); ob_q.clear();
  for (int a = 0; a < nhole; a++)
    for (int b = a; b < nhole; b++) { ob_p.push_back(holes[a]); ob_q.push_back(holes[b]); }
  for (int a = 0; a < npart; a++)
    for (int b = a; b < npart; b++) { ob_p.push_back(parts[a]); ob_q.push_back(parts[b]); }
  for (int a = 0; a < npart; a++)
    for (int b = 0; b < nhole; b++) { ob_p.push_back(parts[a]); ob_q.push_back(holes[b]); }
  // 122 neighbour lists
  nb_start.assign(npairs + 1, 0);
  nb_pos.clear(); nb_x.clear(); nb_i.clear(); nb_coef.clear();
  for (int r = 0; r < npairs; r++) {
    nb_start[r] = (int)nb_pos.size();
    int a = pair_a[r], b = pair_b[r], J = bJ[pair_q[r]];
    double inv_old = (a == b) ? 1.0 / std::sqrt(2.0) : 1.0;
    for (int side = 0; side < 2; side++) {
      int x = side == 0 ? a : b;
      int g = grp_of[x];
      for (int t = 0; t < grp_cnt[g]; t++) {
        int i = grp_list[grp_start[g] + t];
        int u = side == 0 ? i : a, v = side == 0 ? b : i;  // new ordered pair (u,v)
        double pre = 1.0;
        int lo = u, hi = v;
        if (u > v) {
          lo = v; hi = u;
          pre = (((jj[u] + jj[v] - J) / 2 + 1) % 2 == 0) ? 1.0 : -1.0;
        } else if (u == v) {
          pre = std::sqrt(2.0);
        }
        int p = pidx[(size_t)(J / 2) * nn2 + (size_t)lo * nsp + hi];
        if (p < 0) return "internal: 122 neighbour outside block";
        nb_pos.push_back(p);
        nb_x.push_back(x);
        nb_i.push_back(i);
        nb_coef.push_back(pre * inv_old);
      }
    }
  }
  nb_start[npairs] = (int)nb_pos.size();
  // 122 row sets: rows grouped by block and unordered group pair, each
  // neighbour entry's index inside its set, and for the sets the set kernel
  // takes (at most 32 rows, at most two entries per (row, neighbour) slot)
  // the slots' entries and the work items
  {
    s122_q.clear(); s122_start.clear(); s122_rows.clear(); s122_frows.clear(); s122_items.clear();
    s122_ab.clear(); s122_sx.clear(); s122_sc.clear();
    nb_loc.assign(nb_pos.size(), -1);
    std::vector<int> loc(npairs, -1);
    for (int qq = 0; qq < nblocks; qq++) {
      std::map<std::pair<int, int>, std::vector<int>> sets;
      for (int r = pbase[qq]; r < pbase[qq + 1]; r++) {
        const int ga = grp_of[pair_a[r]], gb = grp_of[pair_b[r]];
        sets[{std::min(ga, gb), std::max(ga, gb)}].push_back(r);
      }
      for (auto &kv : sets) {
        s122_q.push_back(qq);
        s122_start.push_back((int)s122_rows.size());
        for (int k = 0; k < (int)kv.second.size(); k++) {
          loc[kv.second[k]] = k;
          s122_rows.push_back(kv.second[k]);
        }
      }
    }
    s122_start.push_back((int)s122_rows.size());
    for (int r = 0; r < npairs; r++) {
      const int base = pbase[pair_q[r]];
      for (int e = nb_start[r]; e < nb_start[r + 1]; e++) {
        const int rn = base + nb_pos[e];
        const int ga = grp_of[pair_a[r]], gb = grp_of[pair_b[r]], na = grp_of[pair_a[rn]], nb2 = grp_of[pair_b[rn]];
        if (std::min(ga, gb) != std::min(na, nb2) || std::max(ga, gb) != std::max(na, nb2))
          return "internal: 122 neighbour outside its row set";
        nb_loc[e] = loc[rn];
      }
    }
    std::vector<std::pair<int, int>> items;
    const int nsets = (int)s122_q.size();
    for (int st = 0; st < nsets; st++) {
      const int rs = s122_start[st], ns = s122_start[st + 1] - rs, qq = s122_q[st];
      // slot (k, k'): the entries of row k whose neighbour is row k', in entry order
      std::vector<int> sx(2 * (size_t)ns * ns, 0), cnt((size_t)ns * ns, 0);
      std::vector<double> sc(2 * (size_t)ns * ns, 0.0);
      bool ok = ns <= 32;
      for (int k = 0; k < ns && ok; k++) {
        const int r = s122_rows[rs + k];
        for (int e = nb_start[r]; e < nb_start[r + 1]; e++) {
          const int sl = k * ns + nb_loc[e];
          if (cnt[sl] == 2) {
            ok = false;
            break;
          }
          sx[2 * sl + cnt[sl]] = nb_x[e] * nsp + nb_i[e];
          sc[2 * sl + cnt[sl]] = nb_coef[e];
          cnt[sl]++;
        }
      }
      if (!ok) {
        s122_ab.push_back(-1);
        for (int k = 0; k < ns; k++) s122_frows.push_back(s122_rows[rs + k]);
        continue;
      }
      s122_ab.push_back((long long)s122_sx.size());  // two values (A, -B) per slot
      s122_sx.insert(s122_sx.end(), sx.begin(), sx.end());
      s122_sc.insert(s122_sc.end(), sc.begin(), sc.end());
      for (int c0 = 0; c0 < bdim[qq]; c0 += 64) items.push_back({st, c0 >> 6});
    }
    std::stable_sort(items.begin(), items.end(), [&](const std::pair<int, int> &x, const std::pair<int, int> &y) {
      return s122_start[x.first + 1] - s122_start[x.first] > s122_start[y.first + 1] - s122_start[y.first];
    });
    for (auto &it : items) {
      s122_items.push_back(it.first);
      s122_items.push_back(it.second);
    }
  }
  // Hartree-Fock tables
  {
    const int ng = (int)grp_start.size();
    hf_gorder.resize(ng);
    for (int g = 0; g < ng; g++) hf_gorder[g] = g;
    std::sort(hf_gorder.begin(), hf_gorder.end(), [&](int x, int y) {
      const int a = grp_list[grp_start[x]], b = grp_list[grp_start[y]];
      if (tz[a] != tz[b]) return tz[a] < tz[b];
      if (l[a] != l[b]) return l[a] < l[b];
      return jj[a] < jj[b];
    });
    hf_b.clear(); hf_d.clear(); hf_a.clear(); hf_c.clear();
    for (int g : hf_gorder)
      for (int x = 0; x < grp_cnt[g]; x++)
        for (int y = 0; y < grp_cnt[g]; y++) {
          const int a = grp_list[grp_start[g] + x], c = grp_list[grp_start[g] + y];
          hf_b.push_back(a);
          hf_d.push_back(c);
          if (x <= y) {
            hf_a.push_back(a);
            hf_c.push_back(c);
          }
        }
    hf_nb_start.assign(npairs + 1, 0);
    hf_nb.clear();
    std::vector<int> cand;
    for (int r = 0; r < npairs; r++) {
      hf_nb_start[r] = (int)hf_nb.size();
      const int i = pair_a[r], j = pair_b[r], J = bJ[pair_q[r]];
      cand.clear();
      for (int side = 0; side < 2; side++) {
        const int g1 = grp_of[side == 0 ? i : j], g2 = grp_of[side == 0 ? j : i];
        for (int x = 0; x < grp_cnt[g1]; x++)
          for (int y = 0; y < grp_cnt[g2]; y++) {
            const int a = grp_list[grp_start[g1] + x], b = grp_list[grp_start[g2] + y];
            const int p = pidx[(size_t)(J / 2) * nn2 + (size_t)std::min(a, b) * nsp + std::max(a, b)];
            if (p < 0) return "internal: Hartree-Fock pair outside block";
            cand.push_back(p);
          }
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      hf_nb.insert(hf_nb.end(), cand.begin(), cand.end());
    }
    hf_nb_start[npairs] = (int)hf_nb.size();
  }
  // Pandya blocks
  ncc = (JTM + 1) * 4;
  ccJ.assign(ncc, 0); ccrlen.assign(ncc, 0); ccnph.assign(ncc, 0); ccTz.assign(ncc, 0); ccP.assign(ncc, 0);
  ccoff.assign(ncc, 0); cwoff.assign(ncc, 0);
  ccrow.assign((size_t)ncc * nn2, -1);
  ccrbase.assign(ncc + 1, 0); cccbase.assign(ncc + 1, 0); ccrefbase.assign(ncc + 1, 0); ccrlen_ref.assign(ncc, 0);
  ccrow_a.clear(); ccrow_b.clear(); cccol_p.clear(); cccol_h.clear(); ccref.clear();
  int64_t coff = 0, woff = 0;
  for (int q1 = 0; q1 < ncc; q1++) {
    int JC = (q1 % (JTM + 1)) * 2;
    int PAR = q1 / (2 * JTM + 2);
    int TZ = (q1 % (2 * JTM + 2)) / (JTM + 1);
    ccJ[q1] = JC; ccTz[q1] = TZ; ccP[q1] = PAR;
    ccrbase[q1] = (int)ccrow_a.size();
    cccbase[q1] = (int)cccol_p.size();
    ccrefbase[q1] = (int)ccref.size();
    auto valid = [&](int i, int j) {
      return tri(jj[i], jj[j], JC) && (l[i] + l[j]) % 2 == PAR && std::abs(tz[i] - tz[j]) / 2 == TZ;
    };
    // rows: the ordered pairs (a,b), (b,a) of every a < b at 2u, 2u+1, then the
    // (a,a) rows two to a row pair, then one zero row if the count is odd.  A
    // ph J' term combines Z(R1,C1) and Z(R2,C2) with R2 the reverse of R1's pair
    // (the same row when it is an (a,a) pair), so both lie in one 2x2 block.
    int r = 0, nb = 0;
    std::vector<int> same;
    for (int a = 0; a < nsp; a++)
      for (int b = a; b < nsp; b++) {
        if (!valid(a, b)) continue;
        if (a == b) {
          same.push_back(a);
          continue;
        }
        ccrow[((size_t)q1 * nsp + a) * nsp + b] = r;
        ccrow[((size_t)q1 * nsp + b) * nsp + a] = r + 1;
        ccrow_a.push_back(a);
        ccrow_b.push_back(b);
        ccrow_a.push_back(b);
        ccrow_b.push_back(a);
        r += 2;
      }
    for (int a : same) {
      ccrow[((size_t)q1 * nsp + a) * nsp + a] = r++;
      ccrow_a.push_back(a);
      ccrow_b.push_back(a);
    }
    if (r & 1) {
      ccrow_a.push_back(-1);
      ccrow_b.push_back(-1);
      r++;
    }
    int rref = 0;
    for (int i = 0; i < nsp; i++)
      for (int j = 0; j < nsp; j++) {
        if (!valid(i, j)) continue;
        ccref.push_back(ccrow[((size_t)q1 * nsp + i) * nsp + j]);
        rref++;
        if (con[i] == 0 && con[j] == 1) {
          nb++;
          cccol_p.push_back(i);
          cccol_h.push_back(j);
        }
      }
    ccrlen_ref[q1] = rref;
    ccrlen[q1] = r;
    ccnph[q1] = nb;
    ccoff[q1] = coff;
    cwoff[q1] = woff;
    coff = align32(coff + (int64_t)r * nb);
    woff = align32(woff + (int64_t)r * r);
  }
  ccrbase[ncc] = (int)ccrow_a.size();
  cccbase[ncc] = (int)cccol_p.size();
  ccrefbase[ncc] = (int)ccref.size();
  ccsize = coff;
  cwsize = woff;
  // quadruples of the ph recoupling kernel: every (i<=j, k<=l) whose pairs
  // share a block and whose (ij) precedes (kl) there (that order does not
  // depend on J: class and in-class order of a pair are J-independent)
  {
    if (nsp > 256) return "more than 256 orbits: ph quadruple codes need 8-bit orbits";
    std::vector<uint64_t> keys;
    const uint64_t N = nsp;
    for (int qq = 0; qq < nblocks; qq++) {
      const int dim = bdim[qq], base = pbase[qq], J = bJ[qq];
      for (int r = 0; r < dim; r++) {
        const int i = pair_a[base + r], j = pair_b[base + r];
        const int Jlo_ij = std::abs(jj[i] - jj[j]);
        for (int s2 = r; s2 < dim; s2++) {
          const int k = pair_a[base + s2], l = pair_b[base + s2];
          // emit once per quadruple: at its lowest J
          const int Jlo = std::max(Jlo_ij, std::abs(jj[k] - jj[l]));
          if (J != Jlo) continue;
          keys.push_back(((((uint64_t)i * N + k) * N + j) * N + l));
        }
      }
    }
    std::sort(keys.begin(), keys.end());
    ph_quads.resize(keys.size());
    for (size_t t = 0; t < keys.size(); t++) {
      uint64_t x = keys[t];
      const uint32_t l = x % N; x /= N;
      const uint32_t j = x % N; x /= N;
      const uint32_t k = x % N; x /= N;
      const uint32_t i = (uint32_t)x;
      ph_quads[t] = (i << 24) | (j << 16) | (k << 8) | l;
    }
    if (ph_quads.size() >= (1u << 26)) return "too many ph quadruples for the term codes";
    // J' terms in quadruple order; toff of each quadruple
    std::vector<uint32_t> qtoff(ph_quads.size());
    ph_terms.clear();
    for (size_t t = 0; t < ph_quads.size(); t++) {
      const uint32_t c = ph_quads[t];
      const int i = c >> 24, j = (c >> 16) & 255, k = (c >> 8) & 255, l = c & 255;
      const int a1 = std::max(std::abs(jj[j] - jj[l]), std::abs(jj[i] - jj[k])), b1 = std::min(jj[j] + jj[l], jj[i] + jj[k]);
      const int a2 = std::max(std::abs(jj[i] - jj[l]), std::abs(jj[j] - jj[k])), b2 = std::min(jj[i] + jj[l], jj[j] + jj[k]);
      qtoff[t] = (uint32_t)ph_terms.size();
      for (int JP = a1, m = 0; JP <= b1; JP += 2, m++) ph_terms.push_back(((uint32_t)t << 6) | (uint32_t)m);
      for (int JP = a2, m = 0; JP <= b2; JP += 2, m++) ph_terms.push_back(((uint32_t)t << 6) | 32u | (uint32_t)m);
      if (ph_terms.size() & 1) ph_terms.push_back(kPhPadTerm);  // runs start at even offsets
    }
    if (ph_terms.size() >= 0xFFFFFFFFull) return "too many ph terms";
    // per upper element (block-major, row-major upper triangle): its quadruple's toff
    ph_ubase.assign(npairs + 1, 0);
    ph_toff.clear();
    for (int qq = 0; qq < nblocks; qq++) {
      const int dim = bdim[qq], base = pbase[qq];
      for (int r = 0; r < dim; r++) {
        ph_ubase[base + r] = (int64_t)ph_toff.size();
        const uint64_t i = pair_a[base + r], j = pair_b[base + r];
        for (int s2 = r; s2 < dim; s2++) {
          const uint64_t k = pair_a[base + s2], l = pair_b[base + s2];
          const uint64_t key = ((i * N + k) * N + j) * N + l;
          const size_t t = std::lower_bound(keys.begin(), keys.end(), key) - keys.begin();
          if (t >= keys.size() || keys[t] != key) return "internal: ph quadruple missing";
          ph_toff.push_back(qtoff[t]);
        }
      }
    }
    ph_ubase[npairs] = (int64_t)ph_toff.size();
    // Z address of every J' term: the 16-byte half of a 2x2 block of Z
    // (index in units of 2 doubles from the Z base) holding the first Z
    // element the term combines, and a selector: bit 0 that element's
    // position in the half, bit 1 the second element lies in the other half
    // of the block (only where a pair is (a,a): otherwise the two elements
    // Z(R1,C1), Z(R1^1,C1^1) share a half), bit 2 its position there, 4 the
    // block is the mirror (x sigma), 5 the phase (-1)^((ji+jj+jk+jl)/2) is
    // -1, 6 side (T2), 7 valid (all four Pandya rows exist)
    ph_zblk.assign(ph_terms.size(), 0);
    ph_zsel.assign(ph_terms.size(), 0);
    const int JT1 = JTM + 1;
    for (size_t t = 0; t < ph_terms.size(); t++) {
      const uint32_t tc = ph_terms[t];
      if (tc == kPhPadTerm) continue;  // a zero term (selector 0)
      const uint32_t c = ph_quads[tc >> 6];
      const int side = (tc >> 5) & 1, m = tc & 31;
      const int oi = c >> 24, oj = (c >> 16) & 255, ok = (c >> 8) & 255, ol = c & 255;
      int qx, R1, C1, R2, C2;
      auto row = [&](int q, int a, int b) { return ccrow[((size_t)q * nsp + a) * nsp + b]; };
      if (side == 0) {
        const int TZ = std::abs(tz[oi] - tz[ok]) / 2, PAR = (l[oi] + l[ok]) % 2;
        const int JP = std::max(std::abs(jj[oj] - jj[ol]), std::abs(jj[oi] - jj[ok])) + 2 * m;
        qx = (JP >> 1) + TZ * JT1 + 2 * PAR * JT1;
        R1 = row(qx, ok, oi); C1 = row(qx, oj, ol); R2 = row(qx, oi, ok); C2 = row(qx, ol, oj);
      } else {
        const int TZ = std::abs(tz[oi] - tz[ol]) / 2, PAR = (l[oi] + l[ol]) % 2;
        const int JP = std::max(std::abs(jj[oi] - jj[ol]), std::abs(jj[oj] - jj[ok])) + 2 * m;
        qx = (JP >> 1) + TZ * JT1 + 2 * PAR * JT1;
        R1 = row(qx, oj, ok); C1 = row(qx, ol, oi); R2 = row(qx, ok, oj); C2 = row(qx, oi, ol);
      }
      uint8_t sel = (uint8_t)(side << 6);
      if (((jj[oi] + jj[oj] + jj[ok] + jj[ol]) >> 1) & 1) sel |= 32;
      if ((R1 | C1 | R2 | C2) >= 0) {
        const int u = R1 >> 1, w = C1 >> 1, h2 = ccrlen[qx] >> 1;
        if ((R2 >> 1) != u || (C2 >> 1) != w) return "internal: ph term elements in two Z blocks";
        // the stored element (a, b) of the block (a = its row, b = its column
        // parity, transposed in a mirror block) sits at 2 (a ^ b) + a
        auto pos = [&](int R, int C) { return u <= w ? (((R ^ C) & 1) << 1) | (R & 1) : (((R ^ C) & 1) << 1) | (C & 1); };
        const int e1 = pos(R1, C1), e2 = pos(R2, C2);
        if (u > w) sel |= 16;
        const int64_t half = (cwoff[qx] / 4 + (int64_t)std::min(u, w) * h2 + std::max(u, w)) * 2 + (e1 >> 1);
        if (half > 0xFFFFFFFFll) return "Z too large for 32-bit half-block indices";
        ph_zblk[t] = (uint32_t)half;
        sel |= (uint8_t)(128 | (e1 & 1) | (((e1 ^ e2) >> 1) << 1) | ((e2 & 1) << 2));
      }
      ph_zsel[t] = sel;
    }
    ph_two.clear();
    for (size_t t = 0; t < ph_terms.size(); t++)
      if ((ph_zsel[t] & 130) == 130) ph_two.push_back((uint32_t)t);
  }
  // 6j table
  nhalf = (JTM + 1) / 2;
  nJ = JTM + 1;
  {
    size_t h = nhalf, J = nJ;
    sixj.assign(h * h * h * h * J * J, 0.0);
    for (size_t i1 = 0; i1 < h; i1++)
      for (size_t i2 = 0; i2 < h; i2++)
        for (size_t i4 = 0; i4 < h; i4++)
          for (size_t i5 = 0; i5 < h; i5++)
            for (size_t a3 = 0; a3 < J; a3++)
              for (size_t a6 = 0; a6 < J; a6++)
                sixj[((((i1 * h + i2) * J + a3) * h + i4) * h + i5) * J + a6] =
                    racah6j((int)(2 * i1 + 1), (int)(2 * i2 + 1), (int)(2 * a3), (int)(2 * i4 + 1),
                            (int)(2 * i5 + 1), (int)(2 * a6));
  }
  // Packed 6j rows of the ph recoupling apply: for a row key (ji, jj, J) the
  // symbols {ji jj J; x y J'} that can be non-zero, i.e. triangle (x, y, J)
  // and J' from max(|jj-x|, |ji-y|) to min(jj+x, ji+y); ap6_off[key*h*h +
  // (x/2)*h + y/2] is the position of the (x, y) run inside the key's list, -1
  // when the run is empty.
  {
    const int h = nhalf, nk = nhalf * nhalf * nJ;
    ap6_base.assign(nk + 1, 0);
    ap6_off.assign((size_t)nk * h * h, -1);
    ap6_val.clear();
    ap6_max = 0;
    for (int a = 0; a < h; a++)
      for (int b = 0; b < h; b++)
        for (int c = 0; c < nJ; c++) {
          const int key = (a * h + b) * nJ + c, ji = 2 * a + 1, jjv = 2 * b + 1, J = 2 * c;
          ap6_base[key] = (int)ap6_val.size();
          if (tri(ji, jjv, J)) {
            for (int x = 0; x < h; x++)
              for (int y = 0; y < h; y++) {
                const int X = 2 * x + 1, Y = 2 * y + 1;
                const int lo = std::max(std::abs(jjv - X), std::abs(ji - Y)), hi = std::min(jjv + X, ji + Y);
                if (!tri(X, Y, J) || hi < lo) continue;
                ap6_off[(size_t)key * h * h + x * h + y] = (int)ap6_val.size() - ap6_base[key];
                for (int JP = lo; JP <= hi; JP += 2)
                  ap6_val.push_back(sixj[(((((size_t)a * h + b) * nJ + c) * h + x) * h + y) * nJ + (JP >> 1)]);
              }
          }
          ap6_max = std::max(ap6_max, (int)ap6_val.size() - ap6_base[key]);
        }
    ap6_base[nk] = (int)ap6_val.size();
    ap6_max = std::max(ap6_max, 2 * nJ);  // room for an empty run read from position 0
  }
  // Per upper element (ij,kl;J) of the ph apply, everything its output needs
  // beyond the row: bits 0-22 first T1 term of its quadruple (relative to the
  // row's first term), 23-27 n1 (T1 terms), 28-32 n2 (T2 terms), 33-44 / 45-56
  // start of the {ji jj J; jl jk J'} / {ji jj J; jk jl J'} runs in the row
  // key's packed 6j list, 57 T1 sum used, 58 its sign is -1, 59 T2 sum used,
  // 60 k == l (1/sqrt 2).
  {
    const int h = nhalf;
    ph_desc.assign(ph_toff.size(), 0);
    // term offsets are stored relative to the row's smallest one: the terms of
    // one row (i, j) all have quadruples (i, k, j, l), which the (i, k, j, l)
    // order keeps within the terms of orbit i
    ph_tbase.assign(npairs, 0);
    for (int r = 0; r < npairs; r++) {
      uint64_t mn = ~0ull;
      for (int64_t u = ph_ubase[r]; u < ph_ubase[r + 1]; u++) mn = std::min<uint64_t>(mn, ph_toff[u]);
      ph_tbase[r] = ph_ubase[r + 1] > ph_ubase[r] ? (int64_t)mn : 0;
    }
    for (int qq = 0; qq < nblocks; qq++) {
      const int dim = bdim[qq], base = pbase[qq], J = bJ[qq];
      for (int r = 0; r < dim; r++) {
        const int i = pair_a[base + r], j = pair_b[base + r];
        const int ji = jj[i], jjv = jj[j];
        const int key = ((ji >> 1) * h + (jjv >> 1)) * nJ + (J >> 1);
        for (int s2 = r; s2 < dim; s2++) {
          const int k = pair_a[base + s2], l2 = pair_b[base + s2];
          const int jk = jj[k], jl = jj[l2];
          const int a1 = std::max(std::abs(jjv - jl), std::abs(ji - jk)), b1 = std::min(jjv + jl, ji + jk);
          const int a2 = std::max(std::abs(ji - jl), std::abs(jjv - jk)), b2 = std::min(ji + jl, jjv + jk);
          const uint64_t n1 = b1 >= a1 ? (b1 - a1) / 2 + 1 : 0, n2 = b2 >= a2 ? (b2 - a2) / 2 + 1 : 0;
          int o1 = ap6_off[(size_t)key * h * h + (jl >> 1) * h + (jk >> 1)];
          int o2 = ap6_off[(size_t)key * h * h + (jk >> 1) * h + (jl >> 1)];
          const bool v1 = o1 >= 0 && std::abs(tz[l2] - tz[j]) == std::abs(tz[i] - tz[k]) &&
                          (l[l2] + l[j]) % 2 == (l[i] + l[k]) % 2;
          const bool v2 = o2 >= 0 && std::abs(tz[k] - tz[j]) == std::abs(tz[i] - tz[l2]) &&
                          (l[k] + l[j]) % 2 == (l[i] + l[l2]) % 2;
          o1 = v1 ? o1 : 0;
          o2 = v2 ? o2 : 0;
          const size_t u = ph_ubase[base + r] + (s2 - r);
          const uint64_t toff = ph_toff[u] - (uint64_t)ph_tbase[base + r];
          if (toff >= (1ull << 23) || n1 > 31 || n2 > 31 || o1 > 4095 || o2 > 4095)
            return "basis too large for the ph apply descriptors";
          const bool neg1 = ((((jl + jk + J) >> 1) & 1) == 0);  // -(-1)^((jl+jk+J)/2) = -1
          ph_desc[u] = toff | (n1 << 23) | (n2 << 28) | ((uint64_t)o1 << 33) | ((uint64_t)o2 << 45) |
                       ((uint64_t)v1 << 57) | ((uint64_t)neg1 << 58) | ((uint64_t)v2 << 59) |
                       ((uint64_t)(k == l2) << 60);
        }
      }
    }
  }
  // neq: the length of the vectorize stream (IMSRG_ODE.f90:36-44)
  {
    int64_t A = nhole, Pp = npart;
    neq = 1 + A * A + A * Pp + Pp * Pp;
    for (int qq = 0; qq < nblocks; qq++) {
      int64_t np = bnp[qq], nb = bnb[qq], nh = bnh[qq];
      neq += (nh * nh + nh + nb * nb + nb + np * np + np) / 2 + nb * np + nh * np + nh * nb;
    }
  }
  f_off = 32;
  v_off = align32(f_off + (int64_t)nsp * nsp);
  op_size = v_off + vsize;
  flop_pphh = 0;
  for (int qq = 0; qq < nblocks; qq++) flop_pphh += 2.0 * bdim[qq] * bdim[qq] * (bnp[qq] + bnh[qq]);
  flop_ph = 0;
  for (int q1 = 0; q1 < ncc; q1++) flop_ph += 2.0 * ccrlen_ref[q1] * ccrlen_ref[q1] * ccnph[q1];
  return "";
}

// sub-block geometry of gam1..gam6 (basic_IMSRG.f90:118-134)
static const int SEA1[6] = {1, 1, 1, 2, 3, 2};
static const int SEA2[6] = {1, 2, 3, 2, 3, 3};
static const bool SQS[6] = {true, false, false, true, true, false};

// repackage (basic_IMSRG.f90:3857-3918) into the full layout
void Layout::unpack_vec(const double *v, int herm, double *buf) const {
  std::memset(buf, 0, sizeof(double) * op_size);
  size_t k = 0;
  buf[0] = v[k++];
  double *F = buf + f_off;
  for (int i = 0; i < nhole; i++)
    for (int j = 0; j < nhole; j++) F[(size_t)holes[j] * nsp + holes[i]] = v[k++];  // fhh(j,i)
  for (int i = 0; i < npart; i++)
    for (int j = 0; j < npart; j++) F[(size_t)parts[j] * nsp + parts[i]] = v[k++];  // fpp(j,i)
  for (int i = 0; i < nhole; i++)
    for (int j = 0; j < npart; j++) {                                                 // fph(j,i)
      double x = v[k++];
      F[(size_t)parts[j] * nsp + holes[i]] = x;
      F[(size_t)holes[i] * nsp + parts[j]] = herm * x;
    }
  double *V = buf + v_off;
  for (int q = 0; q < nblocks; q++) {
    int dim = bdim[q];
    const int cls_off[4] = {0, bnp[q], bnp[q] + bnb[q], dim};
    double *M = V + boff[q];
    for (int g = 0; g < 6; g++) {
      int r0 = cls_off[SEA1[g] - 1], nr = cls_off[SEA1[g]] - r0;
      int c0 = cls_off[SEA2[g] - 1], nc = cls_off[SEA2[g]] - c0;
      for (int i = 0; i < nc; i++)
        for (int j = SQS[g] ? i : 0; j < nr; j++) {
          double x = v[k++];
          M[(size_t)(r0 + j) * dim + (c0 + i)] = x;          // X(j,i)
          if (SQS[g]) M[(size_t)(r0 + i) * dim + (c0 + j)] = x * herm;   // X(i,j)
          else M[(size_t)(c0 + i) * dim + (r0 + j)] = x * herm;          // implied lower block
        }
    }
  }
}

void Layout::unpack_vec_transpose(const double *buf, int herm, double *v) const {
  size_t k = 0;
  v[k++] = buf[0];
  const double *F = buf + f_off;
  for (int i = 0; i < nhole; i++)
    for (int j = 0; j < nhole; j++) v[k++] = F[(size_t)holes[j] * nsp + holes[i]];
  for (int i = 0; i < npart; i++)
    for (int j = 0; j < npart; j++) v[k++] = F[(size_t)parts[j] * nsp + parts[i]];
  for (int i = 0; i < nhole; i++)
    for (int j = 0; j < npart; j++)
      v[k++] = F[(size_t)parts[j] * nsp + holes[i]] + herm * F[(size_t)holes[i] * nsp + parts[j]];
  const double *V = buf + v_off;
  for (int q = 0; q < nblocks; q++) {
    int dim = bdim[q];
    const int cls_off[4] = {0, bnp[q], bnp[q] + bnb[q], dim};
    const double *M = V + boff[q];
    for (int g = 0; g < 6; g++) {
      int r0 = cls_off[SEA1[g] - 1], nr = cls_off[SEA1[g]] - r0;
      int c0 = cls_off[SEA2[g] - 1], nc = cls_off[SEA2[g]] - c0;
      for (int i = 0; i < nc; i++)
        for (int j = SQS[g] ? i : 0; j < nr; j++) {
          const double a = M[(size_t)(r0 + j) * dim + (c0 + i)];
          if (SQS[g]) {
            // X(j,i) = x then X(i,j) = herm*x; on the diagonal the second write wins
            v[k++] = (j == i) ? herm * a : a + herm * M[(size_t)(r0 + i) * dim + (c0 + j)];
          } else {
            v[k++] = a + herm * M[(size_t)(c0 + i) * dim + (r0 + j)];
          }
        }
    }
  }
}

// vectorize (basic_IMSRG.f90:3794-3854) from the full layout
void Layout::pack_vec(const double *buf, double *v) const {
  size_t k = 0;
  v[k++] = buf[0];
  const double *F = buf + f_off;
  for (int i = 0; i < nhole; i++)
    for (int j = 0; j < nhole; j++) v[k++] = F[(size_t)holes[j] * nsp + holes[i]];
  for (int i = 0; i < npart; i++)
    for (int j = 0; j < npart; j++) v[k++] = F[(size_t)parts[j] * nsp + parts[i]];
  for (int i = 0; i < nhole; i++)
    for (int j = 0; j < npart; j++) v[k++] = F[(size_t)parts[j] * nsp + holes[i]];
  const double *V = buf + v_off;
  for (int q = 0; q < nblocks; q++) {
    int dim = bdim[q];
    const int cls_off[4] = {0, bnp[q], bnp[q] + bnb[q], dim};
    const double *M = V + boff[q];
    for (int g = 0; g < 6; g++) {
      int r0 = cls_off[SEA1[g] - 1], nr = cls_off[SEA1[g]] - r0;
      int c0 = cls_off[SEA2[g] - 1], nc = cls_off[SEA2[g]] - c0;
      for (int i = 0; i < nc; i++)
        for (int j = SQS[g] ? i : 0; j < nr; j++) v[k++] = M[(size_t)(r0 + j) * dim + (c0 + i)];
    }
  }
}

}  // namespace imsrg
