// engine.hip -- contexts, device-resident operators, the commutator
// orchestration and the Magnus/BCH drivers behind the C ABI (imsrg_mi355x.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "engine_internal.hpp"

namespace imsrg {
thread_local KernelEvents g_kernel_events;
}

// Records the launch(es) in `stmt` as one timed interval of class `cls`
// when profiling is on; events go on the context stream, nothing syncs.
static hipEvent_t prof_event(imsrg_ctx_s *c) {
  if (c->evused == c->evpool.size()) {
    if (c->evpool.size() >= 400000) return nullptr;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->evpool.push_back(e);
  }
  return c->evpool[c->evused++];
}
#define PROFS(c, cls, strm, stmt)                                            \
  do {                                                                       \
    hipEvent_t ea_ = nullptr, eb_ = nullptr;                                 \
    if ((c)->prof_on && (((c)->prof_mask >> (cls)) & 1u)) {                 \
      ea_ = prof_event(c);                                                   \
      eb_ = prof_event(c);                                                   \
      if (ea_ && eb_) (void)hipEventRecord(ea_, (strm));                     \
    }                                                                        \
    stmt;                                                                    \
    if (ea_ && eb_) {                                                        \
      (void)hipEventRecord(eb_, (strm));                                     \
      (c)->pend.push_back({(cls), ea_, eb_});                                \
    }                                                                        \
  } while (0)
#define PROF(c, cls, stmt) PROFS(c, cls, (c)->stream, stmt)
// The same for a statement that makes exactly one launch_k launch (the
// GEMMs): the events are recorded with that kernel (KernelEvents), so the
// bracket adds no markers to the stream.
#define PROFK(c, cls, stmt)                                                  \
  do {                                                                       \
    hipEvent_t ea_ = nullptr, eb_ = nullptr;                                 \
    if ((c)->prof_on && (((c)->prof_mask >> (cls)) & 1u)) {                 \
      ea_ = prof_event(c);                                                   \
      eb_ = prof_event(c);                                                   \
      if (ea_ && eb_) g_kernel_events = KernelEvents{ea_, eb_};              \
    }                                                                        \
    stmt;                                                                    \
    g_kernel_events = KernelEvents();                                        \
    if (ea_ && eb_) (c)->pend.push_back({(cls), ea_, eb_});                  \
  } while (0)

const char *imsrg_last_error(void) { return g_err.c_str(); }

// 8 contiguous item ranges of about equal total weight (one per XCD)
static XcdSplit make_split(const std::vector<double> &w) {
  XcdSplit x{};
  double tot = 0;
  for (double v : w) tot += v;
  const int n = (int)w.size();
  double acc = 0;
  int i = 0;
  x.start[0] = 0;
  for (int k = 1; k < 8; k++) {
    const double target = tot * k / 8.0;
    while (i < n && acc + 0.5 * w[i] < target) acc += w[i++];
    x.start[k] = i;
  }
  x.start[8] = n;
  x.len = 0;
  for (int k = 0; k < 8; k++) x.len = std::max(x.len, x.start[k + 1] - x.start[k]);
  return x;
}

// Upper-triangle 64x64 tile schedule (tm <= tn) of one segmented-GEMM launch.
// Workgroup i is dispatched to XCD i % 8, so problems are spread over the 8
// XCDs by flop weight (heaviest first onto the least-loaded XCD): every tile
// of one channel block then runs on one XCD and shares its L2, and each XCD's
// list runs heaviest first.  Short lists are padded with empty slots (x = -1).
static void upper_tiles(int M, int TM, int TN, int pi, std::vector<int4> &out) {
  for (int i = 0; i * TM < M; i++)
    for (int j = 0; j * TN < M; j++)
      if (i * TM <= j * TN + TN - 1) out.push_back(make_int4(pi, i, j, 0));
}
// The int8 Z product's schedule: 32 x 32 upper tiles, one per wavefront and
// four per workgroup.  Problems go to the XCD bins by weight (K x tiles) as
// above; each bin lists its tiles row group by row group and is padded to
// whole workgroups, and workgroup 8 i + x takes tiles 4 i .. 4 i + 3 of bin x,
// so the wavefronts of a workgroup share a row group of A.
static void build_wave_tiles(const std::vector<I8Prob> &probs, std::vector<int4> &tiles) {
  const int NX = 8;
  std::vector<double> w(probs.size());
  std::vector<int> order(probs.size());
  for (size_t i = 0; i < probs.size(); i++) {
    std::vector<int4> t;
    upper_tiles(probs[i].M, 32, 32, 0, t);
    w[i] = (double)(probs[i].seg[0].K32 + 32) * (double)t.size();  // + 32: the stores of a K = 0 block
    order[i] = (int)i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
  std::vector<std::vector<int4>> bins(NX);
  std::vector<double> load(NX, 0.0);
  for (int pi : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    upper_tiles(probs[pi].M, 32, 32, pi, bins[x]);
    load[x] += w[pi];
  }
  size_t len = 0;
  for (auto &b : bins) len = std::max(len, (b.size() + 3) / 4);
  tiles.assign(len * NX * 4, make_int4(-1, 0, 0, 0));
  for (int x = 0; x < NX; x++)
    for (size_t i = 0; i < bins[x].size(); i++) tiles[((i / 4) * NX + x) * 4 + i % 4] = bins[x][i];
}
static void build_seg_tiles(const std::vector<SegProb> &probs, std::vector<int4> &tiles, int TM = 64, int TN = 64) {
  const int NX = 8;
  std::vector<double> w(probs.size());
  std::vector<int> order(probs.size());
  for (size_t i = 0; i < probs.size(); i++) {
    double kt = 0;
    for (int s = 0; s < probs[i].nseg; s++) kt += probs[i].seg[s].K;
    std::vector<int4> t;
    upper_tiles(probs[i].M, TM, TN, 0, t);
    w[i] = kt * (double)t.size() * TM * TN;
    order[i] = (int)i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
  std::vector<std::vector<int4>> bins(NX);
  std::vector<double> load(NX, 0.0);
  for (int pi : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    upper_tiles(probs[pi].M, TM, TN, pi, bins[x]);
    load[x] += w[pi];
  }
  size_t len = 0;
  for (auto &b : bins) len = std::max(len, b.size());
  tiles.assign(len * NX, make_int4(-1, 0, 0, 0));
  for (int x = 0; x < NX; x++)
    for (size_t i = 0; i < bins[x].size(); i++) tiles[i * NX + x] = bins[x][i];
}

// A set of active instances: device table (what the kernels index with
// blockIdx.y) and its host copy.  In graph mode a flow's table is padded with
// -1 entries (which every kernel skips) to the context's full batch, so the
// captured launches stay valid while instances converge one after another:
// the kernels launch over g() >= n() entries, the host reads the first n().
struct ActSet {
  const int *d = nullptr;
  std::vector<int> h;
  int ng = -1;  // table entries the kernels launch over; -1: n()
  int n() const { return (int)h.size(); }
  int g() const { return ng < 0 ? n() : ng; }
};
static ActSet all_set(imsrg_ctx_s *c) {
  ActSet a;
  a.d = c->act_d[0];
  a.h.resize(c->nbatch);
  for (int k = 0; k < c->nbatch; k++) a.h[k] = k;
  return a;
}
// Writes `list` into table `slot` (an asynchronous copy on the context
// stream).  Drivers rewrite a slot only after a stream synchronization, so
// the pinned source is never overwritten while a copy from it is pending.
// Slots 1 and 2 are the BCH and Magnus series tables (series_set).
// With pad > list size the table is padded with -1 entries to pad entries.
static ActSet make_set(imsrg_ctx_s *c, int slot, const std::vector<int> &list, int pad = 0) {
  ActSet a;
  a.d = c->act_d[slot];
  a.h = list;
  const int n = (int)list.size();
  if (pad > n) a.ng = pad;
  if (a.g() > 0) {
    std::memcpy(c->act_h[slot], list.data(), n * sizeof(int));
    for (int k = n; k < a.g(); k++) c->act_h[slot][k] = -1;
    (void)hipMemcpyAsync(c->act_d[slot], c->act_h[slot], a.g() * sizeof(int), hipMemcpyHostToDevice, c->stream);
  }
  return a;
}

static int op_new(imsrg_ctx_s *c, int herm, imsrg_op_s **out) {
  imsrg_op_s *op = new imsrg_op_s;
  op->ctx = c;
  op->herm = herm;
  if (hipMalloc(&op->d, sizeof(double) * c->lay.op_size * c->nbatch) != hipSuccess) {
    delete op;
    return fail(IMSRG_ENOMEM, "hipMalloc operator failed");
  }
  launch_fill(c->stream, op->d, c->lay.op_size * c->nbatch, 0.0);
  *out = op;
  return 0;
}
static imsrg_op_s *tmp_op(imsrg_ctx_s *c, int slot, int herm) {
  if (!c->tmp[slot]) {
    if (op_new(c, herm, &c->tmp[slot]) != 0) return nullptr;
  }
  c->tmp[slot]->herm = herm;
  return c->tmp[slot];
}
static void op_copy(imsrg_op_s *src, imsrg_op_s *dst, const ActSet &a) {
  imsrg_ctx_s *c = src->ctx;
  launch_copy(c->stream, c->lay.op_size, src->d, dst->d, a.d, a.n(), c->lay.op_size);
  dst->herm = src->herm;
}
// scalar slot `idx` of every instance in `a`, in the order of a.h
static int read_scalars(imsrg_ctx_s *c, int idx, const ActSet &a, std::vector<double> &out) {
  HIPCHK(hipMemcpyAsync(c->hscal, c->scal, sizeof(double) * kScal * c->nbatch, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  out.resize(a.n());
  for (int k = 0; k < a.n(); k++) out[k] = c->hscal[(size_t)a.h[k] * kScal + idx];
  return 0;
}
// mat_frob_norm of every instance in `a` (scalar slot 0)
static int op_norms(imsrg_op_s *op, const ActSet &a, std::vector<double> &out) {
  imsrg_ctx_s *c = op->ctx;
  PROF(c, IMSRG_PROF_NORM, launch_sumsq(c->stream, c->dl, op->d, c->partials, c->scal, a.d, a.n()));
  return read_scalars(c, 0, a, out);
}
static int single_only(imsrg_ctx_s *c, const char *what) {
  return c->nbatch == 1 ? 0 : fail(IMSRG_EINVAL, std::string(what) + ": batched context, use the _batch variant");
}

// ---------------------------------------------------------------- static scratch
// Device memory of imsrg_hartree_fock_noalloc: the self-consistency state
// (five nsp x nsp matrices, eigenvalues, e_HF diagonals, per-group changes,
// four scalars) and, in the rest, one window of channel blocks for the
// two-body transform.  Each GPU the library runs on holds its own copy of the
// array until the process ends; host code uses the address that
// hipGetSymbolAddress returns for the current device, never the array's name.
static constexpr long long kHFScratchDoubles = 2LL << 20;  // 16 MiB
static __device__ __align__(256) double scratch[kHFScratchDoubles];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (context creation): the device
// address of `scratch`.
static int scratch_setup(imsrg_ctx_s *c) {
  if (c->device < 0 || c->device >= kMaxDevices) return fail(IMSRG_EDEVICE, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(IMSRG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[c->device] = static_cast<double *>(p);
  }
  c->hf_scratch = g_scratch_ptr[c->device];
  return 0;
}

// ---------------------------------------------------------------- context
int imsrg_ctx_create(int nsp, const int *n, const int *l, const int *jj, const int *tz, const int *con,
                     int device, imsrg_ctx *out) {
  return imsrg_ctx_create_batch(nsp, n, l, jj, tz, con, device, 1, out);
}
int imsrg_ctx_create_batch(int nsp, const int *n, const int *l, const int *jj, const int *tz, const int *con,
                           int device, int nbatch, imsrg_ctx *out) {
  if (!out || !n || !l || !jj || !tz || !con) return fail(IMSRG_EINVAL, "null argument");
  if (nbatch < 1 || nbatch > 4096) return fail(IMSRG_EINVAL, "nbatch must be 1..4096");
  imsrg_ctx_s *c = new imsrg_ctx_s;
  c->nbatch = nbatch;
  std::string err = c->lay.build(nsp, n, l, jj, tz, con);
  if (!err.empty()) {
    delete c;
    return fail(IMSRG_EINVAL, err);
  }
  if (device >= 0) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
      delete c;
      return fail(IMSRG_EDEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
  }
  (void)hipGetDevice(&c->device);
  if (int rc = scratch_setup(c)) {
    delete c;
    return rc;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(IMSRG_EDEVICE, "hipStreamCreate failed");
  }
  if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    imsrg_ctx_destroy(c);
    return fail(IMSRG_EDEVICE, "second stream / events failed");
  }
  for (int k = 0; k < imsrg_ctx_s::kGateRing; k++)
    if (hipEventCreateWithFlags(&c->ev_gate[k], hipEventDisableTiming) != hipSuccess) {
      imsrg_ctx_destroy(c);
      return fail(IMSRG_EDEVICE, "series gate events failed");
    }
  const Layout &Ly = c->lay;
  DevLayout &D = c->dl;
  D.nsp = Ly.nsp; D.nhole = Ly.nhole; D.npart = Ly.npart; D.nblocks = Ly.nblocks; D.ncc = Ly.ncc;
  D.jtot_max = Ly.jtot_max; D.nhalf = Ly.nhalf; D.nJ = Ly.nJ; D.npairs = Ly.npairs;
  D.nob = (int)Ly.ob_p.size();
  D.f_off = Ly.f_off; D.v_off = Ly.v_off; D.op_size = Ly.op_size; D.vsize = Ly.vsize;
  D.ccsize = Ly.ccsize; D.cwsize = Ly.cwsize;
  std::vector<int> php_a, php_i;
  for (int a : Ly.parts)
    for (int i : Ly.holes)
      if (Ly.jj[a] == Ly.jj[i] && Ly.l[a] == Ly.l[i] && Ly.tz[a] == Ly.tz[i]) {
        php_a.push_back(a);
        php_i.push_back(i);
      }
  D.nphp = (int)php_a.size();
  D.nsp = Ly.nsp; D.nJ = Ly.nJ; D.nhalf = Ly.nhalf; D.nblocks = Ly.nblocks;
  {
    // the Pandya kernel stages its tables in LDS: up to the device's
    // per-workgroup limit (160 KiB on gfx950), above 64 KiB by request
    int lds_max = 65536;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
    const size_t need = pandya_lds_bytes(D);
    if (!Ly.bdim.empty() && *std::max_element(Ly.bdim.begin(), Ly.bdim.end()) >= 32768) {
      imsrg_ctx_destroy(c);
      return fail(IMSRG_EINVAL, "basis too large: channel blocks of 2^15 pairs or more");
    }
    if (need > (size_t)lds_max) {
      imsrg_ctx_destroy(c);
      return fail(IMSRG_EINVAL, "basis too large: the Pandya kernel stages " + std::to_string(need) +
                                    " bytes of tables per workgroup (device limit " + std::to_string(lds_max) + ")");
    }
    if (need > 65536 && pandya_allow_lds(need) != hipSuccess) {
      imsrg_ctx_destroy(c);
      return fail(IMSRG_EDEVICE, "the Pandya kernel cannot take " + std::to_string(need) + " bytes of LDS");
    }
  }
  std::vector<int> same_ob, ob_same(Ly.ob_p.size(), -1);
  for (size_t t = 0; t < Ly.ob_p.size(); t++) {
    const int p = Ly.ob_p[t], q = Ly.ob_q[t];
    if (Ly.jj[p] == Ly.jj[q] && Ly.l[p] == Ly.l[q] && Ly.tz[p] == Ly.tz[q]) {
      ob_same[t] = (int)same_ob.size();
      same_ob.push_back((int)t);
    }
  }
  D.nsame = (int)same_ob.size();
  D.ob_nchunk = (D.nphp * Ly.nJ + Ly.nsp * Ly.nJ + 255) / 256;
  std::vector<int> cccol_q(Ly.cccol_p.size());
  for (int q = 0; q < Ly.ncc; q++)
    for (int k = Ly.cccbase[q]; k < Ly.cccbase[q + 1]; k++) cccol_q[k] = q;
  D.nccol = (int)cccol_q.size();
  std::vector<long long> boff(Ly.boff.begin(), Ly.boff.end()), ccoff(Ly.ccoff.begin(), Ly.ccoff.end()),
      cwoff(Ly.cwoff.begin(), Ly.cwoff.end());
  int rc = 0;
#define UP(vec, field) \
  if (!rc) rc = upload_table(c, vec, &D.field)
  UP(Ly.jj, jj); UP(Ly.l, l); UP(Ly.tz, tz); UP(Ly.con, con);
  UP(Ly.grp_of, grp_of); UP(Ly.grp_start, grp_start); UP(Ly.grp_cnt, grp_cnt); UP(Ly.grp_list, grp_list);
  UP(Ly.s122_frows, s122_frows);
  D.n122_frows = (int)Ly.s122_frows.size();
  {
    std::vector<Set122> sets;
    std::vector<int> rowl(Ly.s122_rows.size());
    for (size_t k = 0; k < Ly.s122_q.size(); k++) {
      const int q = Ly.s122_q[k], rs = Ly.s122_start[k], n = Ly.s122_start[k + 1] - rs;
      sets.push_back(Set122{rs, n, Ly.bdim[q], q, (long long)Ly.boff[q], Ly.s122_ab[k]});
      for (int j = rs; j < rs + n; j++) rowl[j] = Ly.s122_rows[j] - Ly.pbase[q];
    }
    std::vector<int2> items(Ly.s122_items.size() / 2), sx(Ly.s122_sx.size() / 2);
    for (size_t k = 0; k < items.size(); k++) items[k] = make_int2(Ly.s122_items[2 * k], Ly.s122_items[2 * k + 1]);
    for (size_t k = 0; k < sx.size(); k++) sx[k] = make_int2(Ly.s122_sx[2 * k], Ly.s122_sx[2 * k + 1]);
    D.n122_items = (int)items.size();
    D.n122_slots = (int)sx.size();
    UP(sets, s122_sets); UP(rowl, s122_rowl); UP(items, s122_items); UP(sx, s122_sx); UP(Ly.s122_sc, s122_sc);
  }
  UP(Ly.bJ, bJ); UP(Ly.bnp, bnp); UP(Ly.bnb, bnb); UP(Ly.bnh, bnh); UP(Ly.bdim, bdim); UP(boff, boff);
  UP(Ly.pbase, pbase); UP(Ly.pair_a, pair_a); UP(Ly.pair_b, pair_b); UP(Ly.pair_q, pair_q);
  UP(Ly.pidx, pidx);
  UP(Ly.nb_start, nb_start); UP(Ly.nb_pos, nb_pos); UP(Ly.nb_x, nb_x); UP(Ly.nb_i, nb_i);
  UP(Ly.hf_gorder, hf_gorder); UP(Ly.hf_b, hf_b); UP(Ly.hf_d, hf_d); UP(Ly.hf_a, hf_a); UP(Ly.hf_c, hf_c);
  UP(Ly.hf_nb_start, hf_nb_start); UP(Ly.hf_nb, hf_nb);
  D.ngrp = (int)Ly.grp_start.size();
  D.n_hf_bd = (int)Ly.hf_b.size();
  D.n_hf_ac = (int)Ly.hf_a.size();
  UP(Ly.nb_coef, nb_coef);
  UP(Ly.ob_p, ob_p); UP(Ly.ob_q, ob_q); UP(php_a, php_a); UP(php_i, php_i);
  UP(same_ob, same_ob); UP(ob_same, ob_same);
  UP(Ly.ccJ, ccJ); UP(Ly.ccrlen, ccrlen); UP(Ly.ccnph, ccnph); UP(ccoff, ccoff); UP(cwoff, cwoff);
  UP(Ly.ccrow, ccrow); UP(Ly.ccrbase, ccrbase); UP(Ly.ccrow_a, ccrow_a); UP(Ly.ccrow_b, ccrow_b);
  UP(Ly.cccbase, cccbase); UP(Ly.cccol_p, cccol_p); UP(Ly.cccol_h, cccol_h); UP(cccol_q, cccol_q);
  UP(Ly.sixj, sixj);
  UP(Ly.ap6_val, ap6_val); UP(Ly.ap6_base, ap6_base);
  D.ap6_max = Ly.ap6_max;
  UP(Ly.ph_quads, ph_quads);
  D.nquad = (long long)Ly.ph_quads.size();
  UP(Ly.ph_terms, ph_terms);
  D.nphterm = (long long)Ly.ph_terms.size();
  UP(Ly.ph_toff, ph_toff);
  UP(Ly.ph_zblk, ph_zblk); UP(Ly.ph_zsel, ph_zsel); UP(Ly.ph_two, ph_two);
  D.nphtwo = (long long)Ly.ph_two.size();
  {
    std::vector<unsigned long long> pd(Ly.ph_desc.begin(), Ly.ph_desc.end());
    UP(pd, ph_desc);
  }
  {
    std::vector<long long> ub(Ly.ph_ubase.begin(), Ly.ph_ubase.end());
    UP(ub, ph_ubase);
  }
  {
    std::vector<double> wr(Ly.s122_frows.size()), wt((Ly.ph_terms.size() + 1023) / 1024, 1.0);
    for (size_t k = 0; k < wr.size(); k++) {
      const int r = Ly.s122_frows[k];
      wr[k] = Ly.bdim[Ly.pair_q[r]] * (double)(Ly.nb_start[r + 1] - Ly.nb_start[r] + 1);
    }
    D.xs_rows = make_split(wr);
    D.xs_terms = make_split(wt);
    std::vector<int> arow(Ly.npairs);
    for (int r = 0; r < Ly.npairs; r++) arow[r] = r;
    std::sort(arow.begin(), arow.end(), [&](int x, int y) {
      if (Ly.pair_a[x] != Ly.pair_a[y]) return Ly.pair_a[x] < Ly.pair_a[y];
      if (Ly.pair_b[x] != Ly.pair_b[y]) return Ly.pair_b[x] < Ly.pair_b[y];
      return Ly.bJ[Ly.pair_q[x]] < Ly.bJ[Ly.pair_q[y]];
    });
    // one work item per 64 upper outputs of a row, rows in (i, j, J) order
    std::vector<PhChunk> chunks;
    std::vector<double> wa;
    for (int k = 0; k < Ly.npairs; k++) {
      const int r = arow[k], q = Ly.pair_q[r], rl = r - Ly.pbase[q], dim = Ly.bdim[q];
      const int i = Ly.pair_a[r], j = Ly.pair_b[r];
      const int key = ((Ly.jj[i] >> 1) * Ly.nhalf + (Ly.jj[j] >> 1)) * Ly.nJ + (Ly.bJ[q] >> 1);
      for (int s0 = rl; s0 < dim; s0 += 64) {
        PhChunk h;
        h.vrow = Ly.v_off + boff[q] + (long long)rl * dim + s0;
        h.dbase = Ly.ph_ubase[r] - rl + s0;
        h.n = std::min(64, dim - s0);
        h.b0 = Ly.ap6_base[key];
        h.len = Ly.ap6_base[key + 1] - Ly.ap6_base[key];
        h.half = i == j;
        h.tbase = Ly.ph_tbase[r];
        chunks.push_back(h);
        wa.push_back(16.0 + h.n);
      }
    }
    D.xs_apply = make_split(wa);
    UP(chunks, ph_chunks);
  }
  std::vector<int4> mtiles;
  for (int q = 0; q < Ly.nblocks; q++) {
    const int nt = (Ly.bdim[q] + 31) / 32;
    for (int a = 0; a < nt; a++)
      for (int b = a; b < nt; b++) mtiles.push_back(make_int4(q, a, b, 0));
  }
  D.nmirror = (int)mtiles.size();
  UP(mtiles, mirror_tiles);
#undef UP
  const size_t B = (size_t)nbatch;
  if (!rc) rc = dalloc(c, (size_t)Ly.vsize * B, &c->wpp);
  if (!rc) rc = dalloc(c, Ly.ccsize * B, &c->cc[0]);
  if (!rc) rc = dalloc(c, Ly.ccsize * B, &c->cc[1]);
  if (!rc) rc = dalloc(c, Ly.cwsize * B, &c->wph);
  if (!rc) rc = dalloc(c, Ly.s122_sx.size() * B, &c->w122);
  if (!rc) rc = dalloc(c, Ly.ph_terms.size() * B, &c->tph);
  if (!rc) rc = dalloc(c, kPartials * B, &c->partials);
  if (!rc) rc = dalloc(c, (size_t)std::max(D.nsame, 1) * D.ob_nchunk * B, &c->obpart);
  if (!rc) rc = dalloc(c, (size_t)(D.nmirror + D.nob) * B, &c->npart);
  if (!rc) rc = dalloc(c, kScal * B, &c->scal);
  if (!rc) rc = dalloc(c, 1, &c->coef_d);
  if (!rc) rc = dalloc(c, (size_t)Ly.nsp * Ly.nsp * B, &c->mono);
  if (!rc) rc = dalloc(c, (size_t)Ly.nsp * Ly.nsp, &c->mbar);
  for (int k = 0; k < imsrg_ctx_s::kSlots && !rc; k++) {
    void *p = nullptr;
    if (hipMalloc(&p, B * sizeof(int)) != hipSuccess) {
      rc = fail(IMSRG_ENOMEM, "hipMalloc instance table failed");
      break;
    }
    c->dallocs.push_back(p);
    c->act_d[k] = (int *)p;
    if (hipHostMalloc((void **)&c->act_h[k], B * sizeof(int), hipHostMallocDefault) != hipSuccess)
      rc = fail(IMSRG_ENOMEM, "hipHostMalloc failed");
  }
  if (!rc) {
    for (size_t k = 0; k < B; k++) c->act_h[0][k] = (int)k;
    if (hipMemcpy(c->act_d[0], c->act_h[0], B * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(IMSRG_EDEVICE, "instance table upload failed");
  }
  if (!rc && hipHostMalloc((void **)&c->hscal, kScal * B * sizeof(double), hipHostMallocDefault) != hipSuccess)
    rc = fail(IMSRG_ENOMEM, "hipHostMalloc failed");
  if (!rc && (hipHostMalloc((void **)&c->live_h, imsrg_ctx_s::kGateRing * B * sizeof(int),
                            hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
              hipHostGetDevicePointer((void **)&c->live_d, c->live_h, 0) != hipSuccess))
    rc = fail(IMSRG_ENOMEM, "hipHostMalloc series gate ring failed");
  if (!rc &&
      hipHostMalloc((void **)&c->hstage, std::max<size_t>(Ly.op_size * B, Ly.ccsize) * sizeof(double),
                    hipHostMallocDefault) != hipSuccess)
    rc = fail(IMSRG_ENOMEM, "hipHostMalloc staging failed");
  // GEMM problem tables.  222_pp_hh per channel block (pairs ordered pp|ph|hh):
  //   RES += L_pp R_pp - R_pp L_pp - L_hh R_hh + R_hh L_hh, midpoint after the pp part.
  // 222_ph per Pandya block (column-major r x nb): Z = g0*LCC*RCC^T + g1*RCC*LCC^T.
  if (!rc) {
    std::vector<SegProb> pp;
    for (int q = 0; q < Ly.nblocks; q++) {
      const int dim = Ly.bdim[q], np = Ly.bnp[q], nb = Ly.bnb[q], nh = Ly.bnh[q];
      if (dim == 0) continue;  // blocks without pp/hh pairs still write their zeros
      const long long bo = Ly.boff[q];
      const long long ah = bo + np + nb, bh = bo + (long long)(np + nb) * dim;
      SegProb P{};
      P.M = dim;
      P.c_off = bo;
      P.w_off = bo;
      P.nmid = 2;
      P.nseg = nh > 0 ? 4 : 2;
      P.seg[0] = SegDesc{bo, bo, np, dim, dim, 0, 1.0};
      P.seg[1] = SegDesc{bo, bo, np, dim, dim, 1, -1.0};
      P.seg[2] = SegDesc{ah, bh, nh, dim, dim, 0, -1.0};
      P.seg[3] = SegDesc{ah, bh, nh, dim, dim, 1, 1.0};
      pp.push_back(P);
    }
    std::vector<int4> tiles, tw;
    build_seg_tiles(pp, tiles);
    build_seg_tiles(pp, tw, 128, 64);
    c->n_pp_tiles = (int)tiles.size();
    c->n_pp_tiles_w = (int)tw.size();
    rc = upload_table(c, pp, (const SegProb **)&c->pp_probs);
    if (!rc) rc = upload_table(c, tiles, (const int4 **)&c->pp_tiles);
    if (!rc) rc = upload_table(c, tw, (const int4 **)&c->pp_tiles_w);
  }
  if (!rc) {
    std::vector<SegProb> ph;
    for (int q = 0; q < Ly.ncc; q++) {
      const int r = Ly.ccrlen[q], nbq = Ly.ccnph[q];
      if (r == 0) continue;  // blocks without ph columns still write Z = 0 on every call
      SegProb P{};
      P.M = r;
      P.c_off = Ly.cwoff[q];
      P.w_off = -1;
      P.nmid = P.nseg = 2;
      P.seg[0] = SegDesc{Ly.ccoff[q], Ly.ccoff[q], nbq, r, r, 0, 1.0};
      P.seg[1] = SegDesc{Ly.ccoff[q], Ly.ccoff[q], nbq, r, r, 1, 1.0};
      ph.push_back(P);
    }
    std::vector<int4> tiles, tw;
    build_seg_tiles(ph, tiles);
    build_seg_tiles(ph, tw, 128, 64);
    c->n_ph_tiles = (int)tiles.size();
    c->n_ph_tiles_w = (int)tw.size();
    rc = upload_table(c, ph, (const SegProb **)&c->ph_probs);
    if (!rc) rc = upload_table(c, tiles, (const int4 **)&c->ph_tiles);
    if (!rc) rc = upload_table(c, tw, (const int4 **)&c->ph_tiles_w);
  }
  // adjoint products: Q = P * Cbar per channel block; Ybar_CC^T = P_CC^T * Sbar^T per Pandya block
  if (!rc) {
    std::vector<GemmProb> pp, ph;
    std::vector<int4> tp, th;
    for (int q = 0; q < Ly.nblocks; q++) {
      const int dim = Ly.bdim[q];
      if (dim == 0) continue;
      const int id = (int)pp.size();
      pp.push_back(GemmProb{dim, dim, dim, dim, dim, dim, Ly.boff[q], Ly.boff[q], Ly.boff[q]});
      for (int i = 0; i < (dim + 63) / 64; i++)
        for (int j = 0; j < (dim + 63) / 64; j++) tp.push_back(make_int4(id, i, j, 0));
    }
    for (int q = 0; q < Ly.ncc; q++) {
      const int r = Ly.ccrlen[q], nbq = Ly.ccnph[q];
      if (r == 0 || nbq == 0) continue;
      const int id = (int)ph.size();
      ph.push_back(GemmProb{nbq, r, r, r, r, r, Ly.ccoff[q], Ly.cwoff[q], Ly.ccoff[q]});
      for (int i = 0; i < (nbq + 63) / 64; i++)
        for (int j = 0; j < (r + 63) / 64; j++) th.push_back(make_int4(id, i, j, 0));
    }
    c->n_adj_pp_tiles = (int)tp.size();
    c->n_adj_ph_tiles = (int)th.size();
    rc = upload_table(c, pp, (const GemmProb **)&c->adj_pp_probs);
    if (!rc) rc = upload_table(c, tp, (const int4 **)&c->adj_pp_tiles);
    if (!rc) rc = upload_table(c, ph, (const GemmProb **)&c->adj_ph_probs);
    if (!rc) rc = upload_table(c, th, (const int4 **)&c->adj_ph_tiles);
  }
  if (!rc) launch_fill(c->stream, c->wpp, Ly.vsize * B, 0.0);
  if (!rc && Ly.cwsize > 0) launch_fill(c->stream, c->wph, Ly.cwsize * B, 0.0);
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(IMSRG_EDEVICE, "context setup failed");
  if (rc) {
    imsrg_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int imsrg_ctx_destroy(imsrg_ctx c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  for (auto *t : c->tmp)
    if (t) {
      (void)hipFree(t->d);
      delete t;
    }
  for (auto &kv : c->gexec) (void)hipGraphExecDestroy(kv.second);
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  for (void *p : c->dallocs) (void)hipFree(p);
  if (c->hscal) (void)hipHostFree(c->hscal);
  for (int k = 0; k < imsrg_ctx_s::kSlots; k++)
    if (c->act_h[k]) (void)hipHostFree(c->act_h[k]);
  if (c->hstage) (void)hipHostFree(c->hstage);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  for (hipEvent_t e : c->ev_gate)
    if (e) (void)hipEventDestroy(e);
  if (c->live_h) (void)hipHostFree(c->live_h);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int64_t imsrg_ctx_neq(imsrg_ctx c) { return c ? c->lay.neq : -1; }
int imsrg_ctx_nbatch(imsrg_ctx c) { return c ? c->nbatch : -1; }

int imsrg_ctx_info(imsrg_ctx c, double *info) {
  if (!c || !info) return fail(IMSRG_EINVAL, "null argument");
  const Layout &L = c->lay;
  int maxdim = 0, maxr = 0;
  for (int d : L.bdim) maxdim = std::max(maxdim, d);
  for (int r : L.ccrlen) maxr = std::max(maxr, r);
  double nup = 0;
  for (int d : L.bdim) nup += 0.5 * d * (d + 1.0);
  double v[16] = {(double)L.nsp, (double)L.nhole, (double)L.nblocks, (double)L.ncc, (double)L.npairs,
                  (double)maxdim, (double)maxr, 8.0 * L.op_size, (double)L.vsize, (double)L.ccsize,
                  (double)L.cwsize, L.flop_pphh, L.flop_ph, (double)L.ph_terms.size(), nup,
                  (double)(c->shape_pp + 1) * 4 + (double)(c->shape_ph + 1)};
  std::memcpy(info, v, sizeof(v));
  return 0;
}

// Tables and buffers of the int8-sliced GEMMs (ozaki.hip).  Channel block q:
// 64*ceil(dim/64) rows of ld = 64*ceil(np/64) + 64*ceil(nh/64) columns (pp
// columns, then hh columns), 5 ld bytes each; Pandya block q: 64*ceil(r/64)
// rows of 64*ceil(nph/64) columns.  The problem tables follow the FP64 ones entry for
// entry, so the FP64 64 x 64 tile lists serve both.  Padding stays zero: the
// slice kernels rewrite every byte of a used row and never touch pad rows.
static int i8_setup(imsrg_ctx_s *c) {
  const Layout &Ly = c->lay;
  const size_t B = c->nbatch;
  auto pad = [](long long x) { return (x + 63) & ~63LL; };
  auto pad32 = [](long long x) { return (x + 31) & ~31LL; };
  std::vector<long long> vpoff(Ly.nblocks, 0), vsoff(Ly.nblocks, 0), cpoff(Ly.ncc, 0), csoff(Ly.ncc, 0);
  std::vector<int> vld(Ly.nblocks, 0), cld(Ly.ncc, 0);
  std::vector<int4> items;
  std::vector<I8Prob> pp, ph;
  long long pb = 0, rb = 0, pmb = 0;
  int kmax = 0;
  for (int q = 0; q < Ly.nblocks; q++) {
    const int dim = Ly.bdim[q], np = Ly.bnp[q], nh = Ly.bnh[q];
    if (dim == 0) continue;
    const int ld = (int)(pad(np) + pad(nh));
    vld[q] = ld;
    vpoff[q] = pb;
    vsoff[q] = rb;
    I8Prob P{};
    P.M = dim;
    P.ld = ld;
    P.c_off = P.w_off = Ly.boff[q];
    P.p_off = pb;
    P.s_off = rb;
    P.nmid = 2;
    P.nseg = nh > 0 ? 4 : 2;
    P.seg[0] = I8Seg{0, (int)pad(np), 0, (int)pad32(np), 1.0};
    P.seg[1] = I8Seg{0, (int)pad(np), 1, (int)pad32(np), -1.0};
    P.seg[2] = I8Seg{(int)pad(np), (int)pad(nh), 0, (int)pad32(nh), -1.0};
    P.seg[3] = I8Seg{(int)pad(np), (int)pad(nh), 1, (int)pad32(nh), 1.0};
    pp.push_back(P);
    kmax = std::max(kmax, ld);
    pb += pad(dim) * ld * kI8Planes;
    rb += pad(dim);
  }
  c->i8.v_bytes = (pb + 255) & ~255LL;
  c->i8.v_sr = std::max<long long>(rb, 1);
  pb = rb = 0;
  for (int q = 0; q < Ly.ncc; q++) {
    const int r = Ly.ccrlen[q], nbq = Ly.ccnph[q];
    if (r == 0) continue;
    const int ld = (int)pad(nbq);
    cld[q] = ld;
    cpoff[q] = pb;
    csoff[q] = rb;
    I8Prob P{};
    P.M = r;
    P.ld = ld;
    P.c_off = Ly.cwoff[q];
    P.w_off = -1;
    P.p_off = pb;
    P.s_off = rb;
    P.nmid = P.nseg = 2;
    P.seg[0] = I8Seg{0, ld, 0, (int)pad32(nbq), 1.0};
    P.seg[1] = I8Seg{0, ld, 1, (int)pad32(nbq), 1.0};
    ph.push_back(P);
    if (nbq > 0)
      for (int r0 = 0; r0 < r; r0 += 64) {
        const int nck = (ld + 255) / 256;
        for (int k0 = 0; k0 < ld; k0 += 256) items.push_back(make_int4(q, r0, k0, (int)pmb));
        pmb += 64LL * nck;
      }
    kmax = std::max(kmax, ld);
    pb += pad(r) * ld * kI8Planes;
    rb += pad(r);
  }
  c->i8.c_bytes = (pb + 255) & ~255LL;
  c->i8.c_sr = std::max<long long>(rb, 1);
  c->i8.n_items = (int)items.size();
  c->i8.pm_size = std::max<long long>(pmb, 1);
  // a level sums at most kI8Planes products of |a_i b_j| <= 64^2 over K
  if ((long long)kmax * kI8Planes * 64 * 64 >= (1LL << 31)) return fail(IMSRG_EINVAL, "int8 slices: K too long");
  int rc = upload_table(c, vpoff, &c->i8.v_poff);
  if (!rc) rc = upload_table(c, vsoff, &c->i8.v_soff);
  if (!rc) rc = upload_table(c, vld, &c->i8.v_ld);
  if (!rc) rc = upload_table(c, cpoff, &c->i8.c_poff);
  if (!rc) rc = upload_table(c, csoff, &c->i8.c_soff);
  if (!rc) rc = upload_table(c, cld, &c->i8.c_ld);
  if (!rc) rc = upload_table(c, items, &c->i8.c_items);
  if (!rc) rc = dalloc(c, c->i8.pm_size * B, &c->i8pm);
  if (!rc) rc = upload_table(c, pp, &c->pp_i8);
  if (!rc) rc = upload_table(c, ph, &c->ph_i8);
  if (!rc) {
    std::vector<int4> zt;
    build_wave_tiles(ph, zt);
    c->n_ph_i8_tiles = (int)zt.size();
    rc = upload_table(c, zt, &c->ph_i8_tiles);
  }
  for (int k = 0; k < 2 && !rc; k++) {
    void *p = nullptr;
    const size_t vb = (size_t)c->i8.v_bytes * B, cb = (size_t)c->i8.c_bytes * B;
    if (hipMalloc(&p, std::max<size_t>(vb, 256)) != hipSuccess) return fail(IMSRG_ENOMEM, "hipMalloc int8 planes failed");
    c->dallocs.push_back(p);
    c->i8v[k] = (signed char *)p;
    HIPCHK(hipMemset(p, 0, std::max<size_t>(vb, 256)));
    if (hipMalloc(&p, std::max<size_t>(cb, 256)) != hipSuccess) return fail(IMSRG_ENOMEM, "hipMalloc int8 planes failed");
    c->dallocs.push_back(p);
    c->i8c[k] = (signed char *)p;
    HIPCHK(hipMemset(p, 0, std::max<size_t>(cb, 256)));
    rc = dalloc(c, c->i8.v_sr * B, &c->i8vs[k]);
    if (!rc) rc = dalloc(c, c->i8.c_sr * B, &c->i8cs[k]);
    if (!rc) HIPCHK(hipMemset(c->i8vs[k], 0, sizeof(double) * c->i8.v_sr * B));
    if (!rc) HIPCHK(hipMemset(c->i8cs[k], 0, sizeof(double) * c->i8.c_sr * B));
  }
  if (rc) return rc;
  HIPCHK(hipDeviceSynchronize());
  c->i8_ready = true;
  return 0;
}

int imsrg_ctx_set_precision(imsrg_ctx c, int prec) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  if (prec != IMSRG_PREC_FP64 && prec != IMSRG_PREC_SPLIT_FP32 && prec != IMSRG_PREC_SPLIT_BF16 &&
      prec != IMSRG_PREC_INT8X5 && prec != IMSRG_PREC_INT8X4)
    return fail(IMSRG_EINVAL, "unknown precision");
  if ((prec == IMSRG_PREC_INT8X5 || prec == IMSRG_PREC_INT8X4) && !c->i8_ready) {
    int rc = i8_setup(c);
    if (rc) return rc;
  }
  c->prec = prec;
  return 0;
}
int imsrg_ctx_tuning(imsrg_ctx c, int *shape_pp, int *shape_ph, float *ms4) {
  if (!c || !shape_pp || !shape_ph) return fail(IMSRG_EINVAL, "null argument");
  *shape_pp = c->shape_pp;
  *shape_ph = c->shape_ph;
  if (ms4) std::memcpy(ms4, c->tune_ms, sizeof(c->tune_ms));
  return 0;
}
int imsrg_ctx_set_tile_shapes(imsrg_ctx c, int shape_pp, int shape_ph) {
  if (!c || shape_pp < -1 || shape_pp > 1 || shape_ph < -1 || shape_ph > 1)
    return fail(IMSRG_EINVAL, "shapes: -1 (tune), 0 (64x64) or 1 (128x64)");
  if (shape_pp < 0 || shape_ph < 0) shape_pp = shape_ph = -1;
  c->shape_pp = shape_pp;
  c->shape_ph = shape_ph;
  return 0;
}
int imsrg_ctx_set_graphs(imsrg_ctx c, int enable) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->graphs = enable != 0;
  return 0;
}
int imsrg_ctx_graph_stats(imsrg_ctx c, long long *builds, long long *launches) {
  if (!c || !builds || !launches) return fail(IMSRG_EINVAL, "null argument");
  *builds = c->graph_builds;
  *launches = c->graph_launches;
  return 0;
}
int imsrg_ctx_synchronize(imsrg_ctx c) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}
void *imsrg_ctx_stream(imsrg_ctx c) { return c ? (void *)c->stream : nullptr; }

// ---------------------------------------------------------------- operators
int imsrg_op_create(imsrg_ctx c, int herm, imsrg_op *out) {
  if (!c || !out) return fail(IMSRG_EINVAL, "null argument");
  if (herm != 1 && herm != -1) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  return op_new(c, herm, out);
}
static void drop_graphs_of(imsrg_ctx_s *c, const void *d) {
  for (auto it = c->gexec.begin(); it != c->gexec.end();) {
    if (std::get<0>(it->first) == d || std::get<1>(it->first) == d || std::get<2>(it->first) == d) {
      (void)hipGraphExecDestroy(it->second);
      it = c->gexec.erase(it);
    } else {
      ++it;
    }
  }
}
int imsrg_op_destroy(imsrg_op op) {
  if (!op) return 0;
  (void)hipStreamSynchronize(op->ctx->stream);
  drop_graphs_of(op->ctx, op->d);
  (void)hipFree(op->d);
  delete op;
  return 0;
}
int imsrg_op_set_herm(imsrg_op op, int herm) {
  if (!op || (herm != 1 && herm != -1)) return fail(IMSRG_EINVAL, "bad herm");
  op->herm = herm;
  return 0;
}
int imsrg_op_get_herm(imsrg_op op, int *herm) {
  if (!op || !herm) return fail(IMSRG_EINVAL, "null argument");
  *herm = op->herm;
  return 0;
}
// vec holds nbatch vectorize-order vectors back to back
int imsrg_op_upload(imsrg_op op, const double *vec, int64_t n) {
  if (!op || !vec) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = op->ctx;
  if (n != c->lay.neq * c->nbatch) return fail(IMSRG_EINVAL, "length must be neq * nbatch");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nbatch; k++)
    c->lay.unpack_vec(vec + (size_t)k * c->lay.neq, op->herm, c->hstage + (size_t)k * c->lay.op_size);
  HIPCHK(hipMemcpyAsync(op->d, c->hstage, sizeof(double) * c->lay.op_size * c->nbatch, hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}
int imsrg_op_download(imsrg_op op, double *vec, int64_t n) {
  if (!op || !vec) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = op->ctx;
  if (n != c->lay.neq * c->nbatch) return fail(IMSRG_EINVAL, "length must be neq * nbatch");
  HIPCHK(hipMemcpyAsync(c->hstage, op->d, sizeof(double) * c->lay.op_size * c->nbatch, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nbatch; k++)
    c->lay.pack_vec(c->hstage + (size_t)k * c->lay.op_size, vec + (size_t)k * c->lay.neq);
  return 0;
}
int imsrg_op_copy(imsrg_op src, imsrg_op dst) {
  if (!src || !dst || src->ctx != dst->ctx) return fail(IMSRG_EINVAL, "bad operands");
  op_copy(src, dst, all_set(src->ctx));
  return 0;
}
int imsrg_op_clear(imsrg_op op) {
  if (!op) return fail(IMSRG_EINVAL, "null operand");
  launch_fill(op->ctx->stream, op->d, op->ctx->lay.op_size * op->ctx->nbatch, 0.0);
  return 0;
}
int imsrg_op_scale(imsrg_op op, double a) {
  if (!op) return fail(IMSRG_EINVAL, "null operand");
  imsrg_ctx_s *c = op->ctx;
  const ActSet all = all_set(c);
  launch_axpby(c->stream, c->lay.op_size, a, op->d, 0.0, op->d, op->d, all.d, all.n(), c->lay.op_size);
  return 0;
}
int imsrg_op_add(imsrg_op A, double ax, imsrg_op B, double bx, imsrg_op C) {
  if (!A || !B || !C || A->ctx != B->ctx || A->ctx != C->ctx) return fail(IMSRG_EINVAL, "bad operands");
  imsrg_ctx_s *c = A->ctx;
  const ActSet all = all_set(c);
  launch_axpby(c->stream, c->lay.op_size, ax, A->d, bx, B->d, C->d, all.d, all.n(), c->lay.op_size);
  return 0;
}
int imsrg_op_e0_batch(imsrg_op op, double *e0) {
  if (!op || !e0) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = op->ctx;
  const ActSet all = all_set(c);
  launch_gather_e0(c->stream, op->d, c->lay.op_size, c->scal + 1, all.d, all.n());
  std::vector<double> v;
  int rc = read_scalars(c, 1, all, v);
  if (!rc) std::copy(v.begin(), v.end(), e0);
  return rc;
}
int imsrg_op_get_e0(imsrg_op op, double *e0) {
  if (!op || !e0) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = single_only(op->ctx, "imsrg_op_get_e0")) return rc;
  return imsrg_op_e0_batch(op, e0);
}
int imsrg_op_set_e0(imsrg_op op, double e0) {
  if (!op) return fail(IMSRG_EINVAL, "null operand");
  imsrg_ctx_s *c = op->ctx;
  const ActSet all = all_set(c);
  launch_fill(c->stream, op->d, 1, e0, all.d, all.n(), c->lay.op_size);
  return 0;
}
int imsrg_op_norm_batch(imsrg_op op, double *out) {
  if (!op || !out) return fail(IMSRG_EINVAL, "null argument");
  std::vector<double> v;
  int rc = op_norms(op, all_set(op->ctx), v);
  if (!rc) std::copy(v.begin(), v.end(), out);
  return rc;
}
int imsrg_op_norm(imsrg_op op, double *out) {
  if (!op || !out) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = single_only(op->ctx, "imsrg_op_norm")) return rc;
  return imsrg_op_norm_batch(op, out);
}
static int mbpt2_set(imsrg_op_s *H, const ActSet &a, std::vector<double> &out) {
  imsrg_ctx_s *c = H->ctx;
  PROF(c, IMSRG_PROF_NORM, launch_mbpt2(c->stream, c->dl, H->d, c->partials, c->scal + 2, a.d, a.n()));
  return read_scalars(c, 2, a, out);
}
int imsrg_mbpt2_batch(imsrg_op H, double *out) {
  if (!H || !out) return fail(IMSRG_EINVAL, "null argument");
  std::vector<double> v;
  int rc = mbpt2_set(H, all_set(H->ctx), v);
  if (!rc) std::copy(v.begin(), v.end(), out);
  return rc;
}
int imsrg_mbpt2(imsrg_op H, double *out) {
  if (!H || !out) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = single_only(H->ctx, "imsrg_mbpt2")) return rc;
  return imsrg_mbpt2_batch(H, out);
}
int imsrg_normal_order(imsrg_op H) {
  if (!H) return fail(IMSRG_EINVAL, "null operand");
  imsrg_ctx_s *c = H->ctx;
  imsrg_op_s *T = tmp_op(c, 5, H->herm);
  if (!T) return IMSRG_ENOMEM;
  const ActSet all = all_set(c);
  op_copy(H, T, all);
  launch_normal_order(c->stream, c->dl, T->d, H->d, all.d, all.n());
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- Hartree-Fock
static int hf_setup(imsrg_ctx_s *c) {
  if (c->hf_D) return 0;
  const Layout &Ly = c->lay;
  for (int n : Ly.grp_cnt)
    if (n > kHFMaxGroup) return fail(IMSRG_EINVAL, "Hartree-Fock: an (l,j,tz) group has more than 16 orbits");
  const size_t n2 = (size_t)Ly.nsp * Ly.nsp;
  double **mats[5] = {&c->hf_D, &c->hf_rho, &c->hf_Vg, &c->hf_Fm, &c->hf_Tm};
  for (double **m : mats) {
    if (int rc = dalloc(c, n2, m)) return rc;
    HIPCHK(hipMemsetAsync(*m, 0, n2 * sizeof(double), c->stream));
  }
  if (int rc = dalloc(c, Ly.nsp, &c->hf_eig)) return rc;
  if (int rc = dalloc(c, 2 * (size_t)Ly.nsp, &c->hf_tv)) return rc;
  if (int rc = dalloc(c, Ly.grp_start.size(), &c->hf_cp)) return rc;
  if (int rc = dalloc(c, 4, &c->hf_sc)) return rc;  // crit, e_HF, converged, iterations
  return 0;
}

int imsrg_hartree_fock(imsrg_op H, double *coefs, int *iterations, double *e_hf) {
  if (!H) return fail(IMSRG_EINVAL, "null operand");
  imsrg_ctx_s *c = H->ctx;
  if (int rc = single_only(c, "imsrg_hartree_fock")) return rc;
  if (H->herm != 1) return fail(IMSRG_EINVAL, "imsrg_hartree_fock: H must be hermitian");
  if (int rc = hf_setup(c)) return rc;
  const DevLayout &D = c->dl;
  launch_hf_init(c->stream, D, H->d, c->hf_D, c->hf_Tm, c->hf_eig, 1);
  // the reference's loop test `crit > 1e-6` compares with a default-real
  // literal.  Iterations are enqueued eight at a time; the device counts them
  // and, once converged, skips the rest, so the host reads the state once per
  // eight iterations.
  const double conv = (double)1e-6f;
  HIPCHK(hipMemsetAsync(c->hf_sc + 2, 0, 2 * sizeof(double), c->stream));
  int it = 0;
  for (;;) {
    for (int k = 0; k < 8; k++)
      launch_hf_iteration(c->stream, D, H->d, c->hf_D, c->hf_rho, c->hf_Tm, c->hf_Vg, c->hf_Fm, c->hf_eig,
                          c->hf_cp, c->hf_sc, c->hf_sc + 2, conv);
    HIPCHK(hipMemcpyAsync(c->hscal, c->hf_sc, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    it = (int)c->hscal[3];
    if (c->hscal[2] != 0.0) break;
    if (it >= 10000) return fail(IMSRG_ECONV, "Hartree-Fock did not converge in 10000 iterations");
  }
  launch_hf_finish_1b(c->stream, D, c->hf_D, c->hf_Fm, H->d, 1, c->hf_Tm, c->hf_Vg, c->hf_tv, c->hf_sc + 1);
  imsrg_op_s *W = tmp_op(c, 5, 1);
  if (!W) return IMSRG_ENOMEM;
  launch_hf_2b(c->stream, D, c->hf_D, H->d, W->d, 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->hscal, c->hf_sc + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (coefs)
    HIPCHK(hipMemcpyAsync(coefs, c->hf_D, sizeof(double) * c->lay.nsp * c->lay.nsp, hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (e_hf) *e_hf = c->hscal[0];
  if (iterations) *iterations = it;
  return 0;
}

// imsrg_hartree_fock with every buffer in the static scratch array: the
// self-consistency state at the front, then the two-body transform block by
// block through windows of whole channel blocks that fit the rest.  Nothing is
// allocated; the state is cleared by a fill kernel on every call.
int imsrg_hartree_fock_noalloc(imsrg_op H, double *coefs, int *iterations, double *e_hf) {
  if (!H) return fail(IMSRG_EINVAL, "null operand");
  imsrg_ctx_s *c = H->ctx;
  if (c->nbatch != 1) return fail(IMSRG_EINVAL, "imsrg_hartree_fock_noalloc: batched context");
  if (H->herm != 1) return fail(IMSRG_EINVAL, "imsrg_hartree_fock_noalloc: H must be hermitian");
  if (!c->hf_scratch) return fail(IMSRG_EDEVICE, "imsrg_hartree_fock_noalloc: scratch not set up");
  const Layout &Ly = c->lay;
  for (int n : Ly.grp_cnt)
    if (n > kHFMaxGroup) return fail(IMSRG_EINVAL, "Hartree-Fock: an (l,j,tz) group has more than 16 orbits");
  auto a32 = [](long long x) { return (x + 31) & ~31LL; };
  const long long n2 = (long long)Ly.nsp * Ly.nsp;
  double *const base = c->hf_scratch;
  long long o = 0;
  double *m[5];
  for (double *&p : m) {
    p = base + o;
    o += a32(n2);
  }
  double *eig = base + o;
  o += a32(Ly.nsp);
  double *tv = base + o;
  o += a32(2LL * Ly.nsp);
  double *cp = base + o;
  o += a32((long long)Ly.grp_start.size());
  double *sc = base + o;  // crit, e_HF, converged flag, iterations
  o += 32;
  const long long state = o, wcap = kHFScratchDoubles - state;
  if (wcap <= 0)
    return fail(IMSRG_EINVAL, "imsrg_hartree_fock_noalloc: " + std::to_string(Ly.nsp) +
                                  " orbits need more than the 16 MiB scratch (" + std::to_string(state * 8) +
                                  " bytes of state)");
  for (int q = 0; q < Ly.nblocks; q++)
    if ((long long)Ly.bdim[q] * Ly.bdim[q] > wcap)
      return fail(IMSRG_EINVAL, "imsrg_hartree_fock_noalloc: channel block " + std::to_string(q) + " (dim " +
                                    std::to_string(Ly.bdim[q]) + ") exceeds the scratch window of " +
                                    std::to_string(wcap) + " doubles");
  double *D = m[0], *rho = m[1], *Vg = m[2], *Fm = m[3], *Tm = m[4], *W = base + state;
  const DevLayout &DL = c->dl;
  launch_fill(c->stream, base, state, 0.0);
  launch_hf_init(c->stream, DL, H->d, D, Tm, eig, 1);
  const double conv = (double)1e-6f;  // the reference's default-real loop test
  int it = 0;
  for (;;) {
    for (int k = 0; k < 8; k++)
      launch_hf_iteration(c->stream, DL, H->d, D, rho, Tm, Vg, Fm, eig, cp, sc, sc + 2, conv);
    HIPCHK(hipMemcpyAsync(c->hscal, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    it = (int)c->hscal[3];
    if (c->hscal[2] != 0.0) break;
    if (it >= 10000) return fail(IMSRG_ECONV, "Hartree-Fock did not converge in 10000 iterations");
  }
  launch_hf_finish_1b(c->stream, DL, D, Fm, H->d, 1, Tm, Vg, tv, sc + 1);
  // windows of consecutive channel blocks [q0, q1) whose blocks fit wcap
  for (int q0 = 0; q0 < Ly.nblocks;) {
    int q1 = q0 + 1;
    while (q1 < Ly.nblocks &&
           Ly.boff[q1] + (long long)Ly.bdim[q1] * Ly.bdim[q1] - Ly.boff[q0] <= wcap)
      q1++;
    const int r0 = Ly.pbase[q0];
    const int nrows = Ly.pbase[q1 - 1] + Ly.bdim[q1 - 1] - r0;
    launch_hf_2b_window(c->stream, DL, D, H->d, W, 1, r0, nrows, (long long)Ly.boff[q0]);
    q0 = q1;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(c->hscal, sc + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (coefs)
    HIPCHK(hipMemcpyAsync(coefs, D, sizeof(double) * n2, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (e_hf) *e_hf = c->hscal[0];
  if (iterations) *iterations = it;
  return 0;
}

int imsrg_observable_to_hf(imsrg_op O, const double *coefs) {
  if (!O || !coefs) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = O->ctx;
  if (int rc = single_only(c, "imsrg_observable_to_hf")) return rc;
  if (int rc = hf_setup(c)) return rc;
  const DevLayout &D = c->dl;
  HIPCHK(hipMemcpyAsync(c->hf_D, coefs, sizeof(double) * c->lay.nsp * c->lay.nsp, hipMemcpyHostToDevice,
                        c->stream));
  launch_hf_init(c->stream, D, O->d, c->hf_D, c->hf_Tm, nullptr, 0);
  launch_hf_finish_1b(c->stream, D, c->hf_D, c->hf_Tm, O->d, O->herm, nullptr, nullptr, nullptr, nullptr);
  imsrg_op_s *W = tmp_op(c, 5, O->herm);
  if (!W) return IMSRG_ENOMEM;
  launch_hf_2b(c->stream, D, c->hf_D, O->d, W->d, O->herm);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));  // the caller's coefs may go away
  return imsrg_normal_order(O);
}

// ---------------------------------------------------------------- commutator
// Series epilogue of a nested commutator in BCH_EXPAND / MAGNUS_EXPAND:
// acc += coef * RES (add_sq_op, basic_IMSRG.f90:3450-3478, fused into the
// final passes over RES) and, with norm, mat_frob_norm(RES) into scalar slot 0.
struct CommEpi {
  double *acc = nullptr;
  double coef = 0.0;
  const double *coefp = nullptr;  // coef in device memory (graph replays)
  bool norm = false;
};
// cc_ready bits: kLPrepped -- cc[0] already holds the Pandya matrices of L,
// and with int8 slices the slices of L too (prep_L; drivers reuse them across
// the nested commutators of one BCH / Magnus series, where L = Omega);
// kLPandya / kRPandya -- cc[0] / cc[1] hold the Pandya matrices of L / R
// (uploaded by the caller), their int8 slices still to be made.
// Every launch covers the instances of `a`.
enum { kLPrepped = 1, kLPandya = 2, kRPandya = 4 };
static int comm_launches(imsrg_ctx_s *c, imsrg_op_s *L, imsrg_op_s *R, imsrg_op_s *RES, unsigned mask,
                         int cc_ready, const ActSet &a, const CommEpi &epi = CommEpi());

// Tile-shape autotuner: on the first FP64 commutator of a context, each 222
// GEMM is timed (HIP events, best of 3) in both FP64 tile shapes on the
// context's real channel/Pandya blocks and the faster one is kept.  The
// pp/hh timing writes nothing (no C, no W); the Pandya timing writes the Z
// workspace, which every commutator recomputes before reading.  The timing
// covers every instance of the context (a series table may already mark
// instances as stopped, and skipped workgroups would time nothing).
static int autotune(imsrg_ctx_s *c, imsrg_op_s *L, imsrg_op_s *R) {
  const ActSet a = all_set(c);
  hipStream_t s = c->stream;
  const DevLayout &D = c->dl;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto timed = [&](auto &&launch) {
    float best = 1e30f;
    for (int r = 0; r < 4; r++) {
      (void)hipEventRecord(e0, s);
      launch();
      (void)hipEventRecord(e1, s);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) best = std::min(best, ms);  // first run warms up
    }
    return best;
  };
  const double *X0 = L->d + D.v_off, *X1 = R->d + D.v_off;
  c->tune_ms[0] = timed([&] {
    launch_seg_gemm_pphh(s, 0, c->pp_probs, c->pp_tiles, c->n_pp_tiles, X0, X1, nullptr, nullptr, a.d, a.n(), D.op_size,
                         D.vsize);
  });
  c->tune_ms[1] = timed([&] {
    launch_seg_gemm_pphh(s, 3, c->pp_probs, c->pp_tiles_w, c->n_pp_tiles_w, X0, X1, nullptr, nullptr, a.d, a.n(),
                         D.op_size, D.vsize);
  });
  launch_pandya(s, D, L->d, c->cc[0], a.d, a.n());
  launch_pandya(s, D, R->d, c->cc[1], a.d, a.n());
  c->tune_ms[2] = timed([&] {
    launch_seg_gemm_z(s, 0, c->ph_probs, c->ph_tiles, c->n_ph_tiles, c->cc[0], c->cc[1], c->wph, (double)L->herm,
                      (double)R->herm, a.d, a.n(), D.ccsize, D.cwsize);
  });
  c->tune_ms[3] = timed([&] {
    launch_seg_gemm_z(s, 3, c->ph_probs, c->ph_tiles_w, c->n_ph_tiles_w, c->cc[0], c->cc[1], c->wph, (double)L->herm,
                      (double)R->herm, a.d, a.n(), D.ccsize, D.cwsize);
  });
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->shape_pp = c->tune_ms[1] < c->tune_ms[0] ? 1 : 0;
  c->shape_ph = c->tune_ms[3] < c->tune_ms[2] ? 1 : 0;
  HIPCHK(hipGetLastError());
  return 0;
}
// One commutator: the launch sequence of comm_launches, replayed from a cached
// hipGraph when graph mode is on (and profiling off).  The graph holds kernel
// nodes only; operands, workspaces and the instance table are fixed device
// addresses, so a replay is the same work as the eager sequence.
static int comm_impl(imsrg_ctx_s *c, imsrg_op_s *L, imsrg_op_s *R, imsrg_op_s *RES, unsigned mask,
                     int cc_ready, const ActSet &a, const CommEpi &epi = CommEpi()) {
  if (c->prec == IMSRG_PREC_FP64 && c->shape_pp < 0 && a.n() > 0 && (mask & (T222PPHH | T221 | T222PH))) {
    c->fork_mark = nullptr;  // the tuning launches follow the mark on the stream
    int rc = autotune(c, L, R);
    if (rc) return rc;
  }
  if (!c->graphs || c->prof_on || a.n() == 0) return comm_launches(c, L, R, RES, mask, cc_ready, a, epi);
  c->fork_mark = nullptr;  // a replayed graph runs on one stream
  // the series coefficient goes through device memory, written before the
  // replay, so the terms of a series share one graph
  CommEpi ge = epi;
  if (ge.acc) {
    launch_fill(c->stream, c->coef_d, 1, epi.coef);
    ge.coef = 0.0;
    ge.coefp = c->coef_d;
  }
  const imsrg_ctx_s::GraphKey key(L->d, R->d, RES->d, L->herm, R->herm, RES->herm, mask, cc_ready, a.d,
                                  a.g(), c->prec, ge.acc, 0.0, ge.norm ? 1 : 0, c->shape_pp, c->shape_ph);
  auto it = c->gexec.find(key);
  if (it == c->gexec.end()) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = comm_launches(c, L, R, RES, mask, cc_ready, a, ge);
    c->capturing = false;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(IMSRG_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(IMSRG_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    it = c->gexec.emplace(key, x).first;
    c->graph_builds++;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  c->graph_launches++;
  return 0;
}

static int comm_launches(imsrg_ctx_s *c, imsrg_op_s *L, imsrg_op_s *R, imsrg_op_s *RES, unsigned mask,
                         int cc_ready, const ActSet &a, const CommEpi &epi) {
  const bool lprep = cc_ready & kLPrepped, lcc = lprep || (cc_ready & kLPandya), rcc = cc_ready & kRPandya;
  hipStream_t s = c->stream;
  const DevLayout &D = c->dl;
  const int hL = L->herm, hR = R->herm, hres = RES->herm;
  const int *ad = a.d, na = a.g();
  const long long OS = D.op_size;
  if (na == 0) {
    c->fork_mark = nullptr;
    return 0;
  }
  // The pp/hh chain (122 or clear, pp/hh GEMM, one-body partials) and the
  // ph chain (Pandya, Z GEMM, J' terms), with E0 on one of them, touch
  // disjoint outputs and scratch, so they run on two streams; both end before
  // the ph apply accumulates into the blocks the pp/hh chain wrote.  A graph
  // capture records one stream.
  const bool dual = c->dual && !c->capturing;
  hipStream_t s2 = dual ? c->stream2 : s;
  if (dual) {
    if (!c->fork_mark) HIPCHK(hipEventRecord(c->ev_fork, s));
    HIPCHK(hipStreamWaitEvent(s2, c->fork_mark ? c->fork_mark : c->ev_fork, 0));
  }
  c->fork_mark = nullptr;
  // E0 (110, 220) reads only the operands; it joins the chain that ends
  // first at emax=8: the ph chain with four-slice products, the pp/hh chain
  // otherwise (the ph chain's Z product is then the longer one)
  const int ns = i8_planes(c->prec);
  const bool e0_ph = ns == 4;
  if ((mask & (T110 | T220)) && !e0_ph)
    PROFS(c, IMSRG_PROF_E0, s2, launch_e0(s2, D, L->d, R->d, RES->d, hL, hR, mask, c->partials, ad, na, OS));
  if (mask & T122) PROFS(c, IMSRG_PROF_122, s2, launch_122(s2, D, L->d, R->d, RES->d, hres, ad, na, c->w122));
  else launch_clear_blocks(s2, D, RES->d, ad, na);
  const bool i8 = ns > 0;
  if ((mask & (T222PPHH | T221)) && i8) {
    PROFS(c, IMSRG_PROF_SLICE, s2, {
      if (!lprep) launch_slice_v(s2, D, c->i8, ns, L->d, c->i8v[0], c->i8vs[0], ad, na);
      launch_slice_v(s2, D, c->i8, ns, R->d, c->i8v[1], c->i8vs[1], ad, na);
    });
    // B rows are rows of the other operand: X[k][n] = herm * X[n][k]
    PROFK(c, IMSRG_PROF_GEMM_PPHH,
          launch_seg_gemm_i8(s2, ns, c->pp_i8, c->pp_tiles, c->n_pp_tiles, c->i8v[0], c->i8v[1], c->i8vs[0],
                             c->i8vs[1], (double)hR, (double)hL, (mask & T222PPHH) ? RES->d + D.v_off : nullptr,
                             (mask & T221) ? c->wpp : nullptr, ad, na, c->i8.v_bytes, c->i8.v_sr, OS, D.vsize));
  } else if (mask & (T222PPHH | T221)) {
    const bool w = c->prec == IMSRG_PREC_FP64 && c->shape_pp == 1;
    PROFK(c, IMSRG_PROF_GEMM_PPHH,
          launch_seg_gemm_pphh(s2, w ? 3 : c->prec, c->pp_probs, w ? c->pp_tiles_w : c->pp_tiles,
                               w ? c->n_pp_tiles_w : c->n_pp_tiles, L->d + D.v_off, R->d + D.v_off,
                               (mask & T222PPHH) ? RES->d + D.v_off : nullptr, (mask & T221) ? c->wpp : nullptr,
                               ad, na, OS, D.vsize));
  }
  PROFS(c, IMSRG_PROF_ONEBODY, s2,
        launch_ob_partial(s2, D, L->d, R->d, c->wpp, c->obpart, hL, hR, mask, ad, na));
  if (mask & T222PH) {
    if (!lcc) PROF(c, IMSRG_PROF_PANDYA, launch_pandya(s, D, L->d, c->cc[0], ad, na));
    if (!rcc) PROF(c, IMSRG_PROF_PANDYA, launch_pandya(s, D, R->d, c->cc[1], ad, na));
    if (i8) {
      PROF(c, IMSRG_PROF_SLICE, {
        if (!lprep) launch_slice_cc(s, D, c->i8, ns, c->cc[0], c->i8pm, c->i8c[0], c->i8cs[0], ad, na);
        launch_slice_cc(s, D, c->i8, ns, c->cc[1], c->i8pm, c->i8c[1], c->i8cs[1], ad, na);
      });
      PROFK(c, IMSRG_PROF_GEMM_PH,
            launch_zprod_i8(s, ns, c->ph_i8, c->ph_i8_tiles, c->n_ph_i8_tiles, c->i8c[0], c->i8c[1], c->i8cs[0],
                            c->i8cs[1], (double)hL, -(double)hR, c->wph, -(double)hL * (double)hR, ad, na,
                            c->i8.c_bytes, c->i8.c_sr, D.cwsize));
    } else {
      const bool w = c->prec == IMSRG_PREC_FP64 && c->shape_ph == 1;
      PROFK(c, IMSRG_PROF_GEMM_PH,
            launch_seg_gemm_z(s, w ? 3 : c->prec, c->ph_probs, w ? c->ph_tiles_w : c->ph_tiles,
                             w ? c->n_ph_tiles_w : c->n_ph_tiles, c->cc[0], c->cc[1], c->wph, (double)hL,
                             (double)hR, ad, na, D.ccsize, D.cwsize));
    }
    PROF(c, IMSRG_PROF_PH_TERMS, launch_ph_terms(s, D, c->wph, c->tph, (double)hL, (double)hR, ad, na));
  }
  if ((mask & (T110 | T220)) && e0_ph)
    PROF(c, IMSRG_PROF_E0, launch_e0(s, D, L->d, R->d, RES->d, hL, hR, mask, c->partials, ad, na, OS));
  if (dual) {
    HIPCHK(hipEventRecord(c->ev_join, s2));
    HIPCHK(hipStreamWaitEvent(s, c->ev_join, 0));
  }
  if (mask & T222PH) PROF(c, IMSRG_PROF_PH_APPLY, launch_ph_apply(s, D, c->tph, RES->d, ad, na));
  double *np = epi.norm ? c->npart : nullptr;
  PROF(c, IMSRG_PROF_MIRROR, launch_sym_mirror(s, D, RES->d, hres, ad, na, epi.acc, epi.coef, epi.coefp, np));
  PROF(c, IMSRG_PROF_ONEBODY,
       launch_ob_final(s, D, L->d, R->d, c->obpart, RES->d, hres, mask, ad, na, epi.acc, epi.coef, epi.coefp, np));
  if (epi.norm) PROF(c, IMSRG_PROF_NORM, launch_norm_from_parts(s, D, c->npart, c->scal, ad, na));
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator(imsrg_ctx c, imsrg_op L, imsrg_op R, imsrg_op RES, unsigned mask) {
  if (!c || !L || !R || !RES) return fail(IMSRG_EINVAL, "null argument");
  if (L->ctx != c || R->ctx != c || RES->ctx != c) return fail(IMSRG_EINVAL, "operands of another context");
  if (RES == L || RES == R) return fail(IMSRG_EINVAL, "RES must not alias L or R");
  if (RES->herm != -L->herm * R->herm) return fail(IMSRG_EINVAL, "RES herm must be -herm(L)*herm(R)");
  return comm_impl(c, L, R, RES, mask & IMSRG_TALL, 0, all_set(c));
}

static int scalar_term(imsrg_op L, imsrg_op R, unsigned mask, double *out) {
  if (!L || !R || !out || L->ctx != R->ctx) return fail(IMSRG_EINVAL, "bad operands");
  imsrg_ctx_s *c = L->ctx;
  if (int rc = single_only(c, "commutator_110/220")) return rc;
  const ActSet all = all_set(c);
  launch_e0(c->stream, c->dl, L->d, R->d, c->scal + 3, L->herm, R->herm, mask, c->partials, all.d, all.n(), kScal);
  std::vector<double> v;
  int rc = read_scalars(c, 3, all, v);
  if (!rc) *out = v[0];
  return rc;
}
int imsrg_commutator_110(imsrg_op L, imsrg_op R, double *out) { return scalar_term(L, R, T110, out); }
int imsrg_commutator_220(imsrg_op L, imsrg_op R, double *out) { return scalar_term(L, R, T220, out); }

int64_t imsrg_cross_coupled_size(imsrg_ctx c) {
  if (!c) return -1;
  int64_t n = 0;
  for (int q = 0; q < c->lay.ncc; q++) n += (int64_t)c->lay.ccrlen_ref[q] * c->lay.ccnph[q];
  return n;
}
int imsrg_cross_coupled(imsrg_op op, double *host_out, int64_t n) {
  if (!op || !host_out) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = op->ctx;
  if (int rc = single_only(c, "imsrg_cross_coupled")) return rc;
  if (n != imsrg_cross_coupled_size(c)) return fail(IMSRG_EINVAL, "size mismatch");
  const ActSet all = all_set(c);
  launch_pandya(c->stream, c->dl, op->d, c->cc[0], all.d, all.n());
  // device rows are ordered (a,b)/(b,a) pairs, then (a,a); the host copy is in the reference's row order
  const Layout &Ly = c->lay;
  HIPCHK(hipMemcpyAsync(c->hstage, c->cc[0], sizeof(double) * Ly.ccsize, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int64_t k = 0;
  for (int q = 0; q < Ly.ncc; q++) {
    const int rr = Ly.ccrlen_ref[q], rd = Ly.ccrlen[q], nb = Ly.ccnph[q];
    const double *blk = c->hstage + Ly.ccoff[q];
    const int *ref = Ly.ccref.data() + Ly.ccrefbase[q];
    for (int col = 0; col < nb; col++)
      for (int r = 0; r < rr; r++) host_out[k++] = blk[(int64_t)col * rd + ref[r]];
  }
  return 0;
}

// commutator_222_ph (commutators.f90:771-925) from Pandya matrices the caller
// made with calculate_cross_coupled: the two streams (reference row order,
// as imsrg_cross_coupled writes them) go into the device's paired-row layout,
// then the Z product, the J' terms and the 6j-weighted sums run as in
// imsrg_commutator with mask IMSRG_T222PH.
int imsrg_commutator_222_ph_cc(imsrg_ctx c, const double *lcc, int hL, const double *rcc, int hR, int64_t n,
                               imsrg_op RES) {
  if (!c || !lcc || !rcc || !RES) return fail(IMSRG_EINVAL, "null argument");
  if (RES->ctx != c) return fail(IMSRG_EINVAL, "RES belongs to another context");
  if (int rc = single_only(c, "imsrg_commutator_222_ph_cc")) return rc;
  if ((hL != 1 && hL != -1) || (hR != 1 && hR != -1)) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  if (RES->herm != -hL * hR) return fail(IMSRG_EINVAL, "RES herm must be -herm(LCC)*herm(RCC)");
  if (n != imsrg_cross_coupled_size(c)) return fail(IMSRG_EINVAL, "Pandya stream size mismatch");
  const Layout &Ly = c->lay;
  const double *src[2] = {lcc, rcc};
  for (int k = 0; k < 2; k++) {
    std::memset(c->hstage, 0, sizeof(double) * Ly.ccsize);
    int64_t o = 0;
    for (int q = 0; q < Ly.ncc; q++) {
      const int rr = Ly.ccrlen_ref[q], rd = Ly.ccrlen[q], nb = Ly.ccnph[q];
      double *blk = c->hstage + Ly.ccoff[q];
      const int *ref = Ly.ccref.data() + Ly.ccrefbase[q];
      for (int col = 0; col < nb; col++)
        for (int r = 0; r < rr; r++) blk[(int64_t)col * rd + ref[r]] = src[k][o++];
    }
    HIPCHK(hipMemcpyAsync(c->cc[k], c->hstage, sizeof(double) * Ly.ccsize, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));  // hstage is rewritten next
  }
  // operands stand-ins carrying the hermiticities: with mask T222PH and both
  // Pandya matrices in place no kernel reads their storage
  imsrg_op_s Lh = *RES, Rh = *RES;
  Lh.herm = hL;
  Rh.herm = hR;
  c->fork_mark = nullptr;
  return comm_launches(c, &Lh, &Rh, RES, T222PH, kLPandya | kRPandya, all_set(c));
}

// ---------------------------------------------------------------- adjoint commutator
// Ybar = d <Cbar, [P, Y]> / dY for Y of hermiticity hY, restricted to the terms
// in mask, projected onto hermiticity hY (E0 of Ybar is 0: the commutator does
// not depend on E0 of its operands).  <,> is the Frobenius product over the
// full device layout.  Single-instance contexts only.
static int adjoint_impl(imsrg_ctx_s *c, imsrg_op_s *P, imsrg_op_s *Cin, imsrg_op_s *Ybar, int hY, unsigned mask) {
  hipStream_t s = c->stream;
  const DevLayout &D = c->dl;
  const Layout &Ly = c->lay;
  const int hP = P->herm, hC = -hP * hY;
  if (Cin->herm != hC) return fail(IMSRG_EINVAL, "cotangent herm must be -herm(P)*hY");
  const ActSet all = all_set(c);
  imsrg_op_s *Ct = tmp_op(c, 5, hC);
  if (!Ct) return IMSRG_ENOMEM;
  op_copy(Cin, Ct, all);
  launch_project(s, D, Ct->d, hC, 1);
  launch_fill(s, Ybar->d, Ly.op_size, 0.0);
  if (mask & (T110 | T220)) launch_adj_e0(s, D, P->d, Ct->d, Ybar->d, hP, hY, mask);
  launch_adj_onebody(s, D, P->d, Ct->d, Ybar->d, hP, hY, hC, mask);
  if (mask & T122) launch_adj_122(s, D, P->d, Ct->d, Ybar->d, hC);
  if (mask & T222PPHH) {
    launch_gemm_plain(s, c->adj_pp_probs, c->adj_pp_tiles, c->n_adj_pp_tiles, P->d + D.v_off, Ct->d + D.v_off,
                      c->wpp);
    launch_adj_pphh(s, D, c->wpp, Ybar->d, hP, hC);
  }
  if (mask & T222PH) {
    launch_pandya(s, D, P->d, c->cc[0], all.d, all.n());
    launch_fill(s, c->tph, (long long)Ly.ph_terms.size(), 0.0);
    launch_adj_ph_apply(s, D, Ct->d, c->tph, hC);
    launch_fill(s, c->wph, Ly.cwsize, 0.0);
    launch_adj_ph_terms(s, D, c->tph, c->wph, (double)(-hP * hY));
    launch_adj_ph_sbt(s, D, Ly.ccrlen.data(), c->wph, hP, hY);
    launch_gemm_plain(s, c->adj_ph_probs, c->adj_ph_tiles, c->n_adj_ph_tiles, c->cc[0], c->wph, c->cc[1]);
    launch_adj_pandya(s, D, c->cc[1], Ybar->d);
  }
  launch_project(s, D, Ybar->d, hY, 0);
  Ybar->herm = hY;
  HIPCHK(hipGetLastError());
  return 0;
}
int imsrg_commutator_adjoint(imsrg_ctx c, imsrg_op P, imsrg_op Cbar, imsrg_op Ybar, int hY, unsigned mask) {
  if (!c || !P || !Cbar || !Ybar) return fail(IMSRG_EINVAL, "null argument");
  if (P->ctx != c || Cbar->ctx != c || Ybar->ctx != c) return fail(IMSRG_EINVAL, "operands of another context");
  if (Ybar == P || Ybar == Cbar) return fail(IMSRG_EINVAL, "Ybar must not alias P or Cbar");
  if (hY != 1 && hY != -1) return fail(IMSRG_EINVAL, "hY must be +1 or -1");
  if (int rc = single_only(c, "imsrg_commutator_adjoint")) return rc;
  return adjoint_impl(c, P, Cbar, Ybar, hY, mask & IMSRG_TALL);
}

// ---------------------------------------------------------------- generators
static void gen_simple(imsrg_ctx_s *c, imsrg_op_s *Hden, imsrg_op_s *Hnum, imsrg_op_s *ETA, int kind,
                       const ActSet &a) {
  hipStream_t s = c->stream;
  PROF(c, IMSRG_PROF_GENERATOR, {
    launch_fill(s, ETA->d + c->lay.f_off, c->lay.op_size - c->lay.f_off, 0.0, a.d, a.n(), c->lay.op_size);
    launch_monopole(s, c->dl, Hden->d, c->mono, a.d, a.n());
    launch_eta(s, c->dl, Hden->d, Hnum->d, c->mono, ETA->d, kind, a.d, a.n());
  });
  ETA->herm = -1;
}
static int gen_impl(imsrg_ctx_s *c, int kind, imsrg_op_s *H, imsrg_op_s *ETA, const ActSet &a) {
  if (kind == IMSRG_GEN_WHITE || kind == IMSRG_GEN_ATAN) {
    gen_simple(c, H, H, ETA, kind == IMSRG_GEN_WHITE ? 0 : 1, a);
  } else if (kind == IMSRG_GEN_W2) {
    // build_gs_w2 (generators.f90:109-160)
    imsrg_op_s *E1 = tmp_op(c, 2, -1), *H2 = tmp_op(c, 3, 1), *E2 = tmp_op(c, 4, -1);
    if (!E1 || !H2 || !E2) return IMSRG_ENOMEM;
    gen_simple(c, H, H, E1, 1, a);
    int rc = comm_impl(c, E1, H, H2, IMSRG_TALL, 0, a);
    if (rc) return rc;
    gen_simple(c, H, H2, E2, 1, a);
    launch_axpby(c->stream, c->lay.op_size, 1.5, E1->d, 0.5, E2->d, ETA->d, a.d, a.n(), c->lay.op_size);
    ETA->herm = -1;
  } else {
    return fail(IMSRG_EINVAL, "unknown generator kind");
  }
  HIPCHK(hipGetLastError());
  return 0;
}
int imsrg_generator(imsrg_ctx c, int kind, imsrg_op H, imsrg_op ETA) {
  if (!c || !H || !ETA || H == ETA) return fail(IMSRG_EINVAL, "bad operands");
  return gen_impl(c, kind, H, ETA, all_set(c));
}
int imsrg_generator_atanxx(imsrg_ctx c, imsrg_op Hold, imsrg_op H, imsrg_op ETA) {
  if (!c || !Hold || !H || !ETA || ETA == H || ETA == Hold) return fail(IMSRG_EINVAL, "bad operands");
  gen_simple(c, Hold, H, ETA, 1, all_set(c));
  HIPCHK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------- Magnus / BCH
static const double kConv = (double)1e-6f;   // real(8), parameter :: conv = 1e-6
static const double kTiny = (double)1e-9f;   // fullnorm < 1e-9 (default-real literal)

// A BCH or Magnus series works on its own copy of the caller's instance
// table (slot `slot`), which the gate kernels edit: an instance that stops
// gets -1 there and every later kernel of the series skips it.  The copy is
// made on the stream, behind the launches that read the slot before.
static ActSet series_set(imsrg_ctx_s *c, int slot, const ActSet &a) {
  ActSet s;
  s.d = c->act_d[slot];
  s.h = a.h;
  s.ng = a.ng;
  (void)hipMemcpyAsync(c->act_d[slot], a.d, a.g() * sizeof(int), hipMemcpyDeviceToDevice, c->stream);
  return s;
}
// The operand L = Omega of every nested commutator of a series, prepared once:
// its Pandya matrices and, with int8 slices, the slices of its V rows and of
// those matrices.
static void prep_L(imsrg_ctx_s *c, imsrg_op_s *G, const ActSet &set) {
  launch_pandya(c->stream, c->dl, G->d, c->cc[0], set.d, set.n());
  if (const int ns = i8_planes(c->prec)) {
    launch_slice_v(c->stream, c->dl, c->i8, ns, G->d, c->i8v[0], c->i8vs[0], set.d, set.n());
    launch_slice_cc(c->stream, c->dl, c->i8, ns, c->cc[0], c->i8pm, c->i8c[0], c->i8cs[0], set.d, set.n());
  }
}
// The series gate after nested commutator n: the test on the device, its
// live flags into ring slot n % kGateRing, marked by an event.
static int series_gate(imsrg_ctx_s *c, int slot, const ActSet &set, int n, int mode, double coef, int store,
                       int test) {
  const int r = n % imsrg_ctx_s::kGateRing;
  launch_series_gate(c->stream, c->act_d[slot], set.g(), c->scal, mode, coef, store, test, kConv, kTiny,
                     c->live_d + (size_t)r * c->nbatch);
  HIPCHK(hipEventRecord(c->ev_gate[r], c->stream));
  return 0;
}
// Waits for the live flags of gate n; live[k] = 1 for positions that go on.
static int series_live(imsrg_ctx_s *c, const ActSet &set, int n, std::vector<int> &live, int &nlive) {
  const int r = n % imsrg_ctx_s::kGateRing;
  HIPCHK(hipEventSynchronize(c->ev_gate[r]));
  const volatile int *f = c->live_h + (size_t)r * c->nbatch;
  live.assign(set.n(), 0);
  nlive = 0;
  for (int k = 0; k < set.n(); k++) nlive += (live[k] = f[k] ? 1 : 0);
  return 0;
}

// BCH_EXPAND (IMSRG_MAGNUS.f90:149-234) on the instances of `a`.  Every
// instance keeps its own truncation: after each nested commutator the gate
// kernel drops the instances whose ||ad^n||/n! fell below conv, so one launch
// per kernel serves all instances still expanding.  The host enqueues nested
// commutator n+1 before it reads the outcome of the test after n-1 (its
// flags are ready while n runs), so the device never waits for the host;
// when every instance has stopped, the one commutator already enqueued
// beyond the end is skipped by all its workgroups.  nterms[k] is the number
// of nested commutators instance a.h[k] used.
static int bch_impl(imsrg_ctx_s *c, imsrg_op_s *HS, imsrg_op_s *G, imsrg_op_s *H, const ActSet &a,
                    std::vector<int> &nterms) {
  imsrg_op_s *AD = tmp_op(c, 0, 1), *INT2 = tmp_op(c, 1, 1);
  if (!AD || !INT2) return IMSRG_ENOMEM;
  nterms.assign(a.n(), 0);
  if (a.n() == 0) return 0;
  op_copy(H, HS, a);
  op_copy(H, AD, a);  // AD_1 = H
  // [G, X] with G anti-hermitian keeps the hermiticity of X
  AD->herm = H->herm;
  INT2->herm = H->herm;
  const ActSet cur = series_set(c, 1, a);
  prep_L(c, G, cur);
  std::vector<int> live(a.n(), 1);  // positions in the series for nested commutator iw
  int nlive = a.n();
  double coef = 1.0;
  for (int iw = 2; iw <= 40; iw++) {
    coef = coef / (iw - 1.0);
    // INT2 = [G, AD]; HS += coef * INT2 and ||INT2|| in the commutator's final passes
    CommEpi epi;
    epi.acc = HS->d;
    epi.coef = coef;
    epi.norm = true;
    // nothing was enqueued since the last gate: the second stream can fork from its event
    c->fork_mark = iw > 2 ? c->ev_gate[(iw - 1) % imsrg_ctx_s::kGateRing] : nullptr;
    int rc = comm_impl(c, G, AD, INT2, IMSRG_TALL, kLPrepped, cur, epi);
    if (!rc) rc = series_gate(c, 1, cur, iw, 0, coef, -1, 0);
    if (!rc && iw > 2) rc = series_live(c, cur, iw - 1, live, nlive);
    if (rc) return rc;
    if (nlive == 0) return 0;  // the test after iw-1 stopped every instance
    for (int k = 0; k < cur.n(); k++) nterms[k] += live[k];
    std::swap(c->tmp[0], c->tmp[1]);
    AD = c->tmp[0];
    INT2 = c->tmp[1];
  }
  int rc = series_live(c, cur, 40, live, nlive);
  if (rc) return rc;
  if (nlive > 0) return fail(IMSRG_ECLIP, "clipping BCH");
  return 0;
}
// transform_sq_op_BCH (IMSRG_MAGNUS.f90:131-146) for a rank-0 operator:
// BCH_EXPAND into a scratch operator, copied back over Op.
int imsrg_transform_bch(imsrg_op Op, imsrg_op G, int *nterms) {
  if (!Op || !G || Op == G || Op->ctx != G->ctx) return fail(IMSRG_EINVAL, "bad operands");
  if (G->herm != -1) return fail(IMSRG_EINVAL, "Omega must be anti-hermitian");
  imsrg_ctx_s *c = Op->ctx;
  imsrg_op_s *T = tmp_op(c, 6, Op->herm);
  if (!T) return IMSRG_ENOMEM;
  const ActSet all = all_set(c);
  std::vector<int> nt;
  int rc = bch_impl(c, T, G, Op, all, nt);
  if (rc) return rc;
  op_copy(T, Op, all);
  if (nterms) std::copy(nt.begin(), nt.end(), nterms);
  HIPCHK(hipGetLastError());
  return 0;
}

// dHds (IMSRG_ODE.f90:416-463): DH = [eta, HS] with eta the generator of
// `kind` built from HS; every term, E0 from 110 + 220.
int imsrg_dhds(imsrg_op HS, int kind, imsrg_op DH) {
  if (!HS || !DH || HS == DH || HS->ctx != DH->ctx) return fail(IMSRG_EINVAL, "bad operands");
  if (DH->herm != HS->herm) return fail(IMSRG_EINVAL, "DH herm must equal herm(HS)");
  imsrg_ctx_s *c = HS->ctx;
  imsrg_op_s *ETA = tmp_op(c, 6, -1);
  if (!ETA) return IMSRG_ENOMEM;
  const ActSet all = all_set(c);
  int rc = gen_impl(c, kind, HS, ETA, all);
  if (!rc) rc = comm_impl(c, ETA, HS, DH, IMSRG_TALL, 0, all);
  return rc;
}

int imsrg_bch_expand(imsrg_ctx c, imsrg_op HS, imsrg_op G, imsrg_op H, int *nterms) {
  if (!c || !HS || !G || !H || HS == G || HS == H) return fail(IMSRG_EINVAL, "bad operands");
  std::vector<int> nt;
  int rc = bch_impl(c, HS, G, H, all_set(c), nt);
  if (!rc && nterms) std::copy(nt.begin(), nt.end(), nterms);
  return rc;
}

// MAGNUS_EXPAND (IMSRG_MAGNUS.f90:622-693) on the instances of `a`, each with
// its own early exits, tested on the device and read back one nested
// commutator late as in bch_impl; nterms as in bch_impl.
static int magexp_impl(imsrg_ctx_s *c, imsrg_op_s *DG, imsrg_op_s *G, const ActSet &a, std::vector<int> &nterms) {
  static const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  nterms.assign(a.n(), 0);
  if (a.n() == 0) return 0;
  imsrg_op_s *cur = tmp_op(c, 0, DG->herm), *oth = tmp_op(c, 1, DG->herm);
  if (!cur || !oth) return IMSRG_ENOMEM;
  op_copy(DG, cur, a);
  op_copy(DG, oth, a);
  const ActSet set = series_set(c, 2, a);
  // advals(1) = ||DG|| and fullnorm = ||G|| into the series scalar slots; the
  // first gate: fullnorm < 1e-9 returns, then the test on advals(1)
  PROF(c, IMSRG_PROF_NORM,
       launch_sumsq(c->stream, c->dl, cur->d, c->partials, c->scal + kAdvSlot, set.d, set.n()));
  PROF(c, IMSRG_PROF_NORM,
       launch_sumsq(c->stream, c->dl, G->d, c->partials, c->scal + kFullSlot, set.d, set.n()));
  int rc = series_gate(c, 2, set, 1, 1, 0.0, -1, 0);
  if (rc) return rc;
  prep_L(c, G, set);
  std::vector<int> live;
  int nlive = 0;
  for (int i = 2; i <= 7; i++) {
    // q of the early exit before nested commutator i+1 (advals(i+1-q) is tested)
    const int q = (std::fabs(cof[i - 1]) > kConv) ? 1 : 2;
    // oth = [G, cur]; DG += cof * oth and ||oth|| in the commutator's final passes
    CommEpi epi;
    epi.acc = DG->d;
    epi.coef = cof[i - 1];
    epi.norm = true;
    c->fork_mark = i > 2 ? c->ev_gate[(i - 1) % imsrg_ctx_s::kGateRing] : nullptr;
    rc = comm_impl(c, G, cur, oth, IMSRG_TALL & ~(T110 | T220), kLPrepped, set, epi);
    // advals(i) = ||oth|| |cof(i)|, then the test for commutator i+1
    if (!rc && i < 7) rc = series_gate(c, 2, set, i, 1, std::fabs(cof[i - 1]), i - 1, i - q);
    if (!rc) rc = series_live(c, set, i - 1, live, nlive);
    if (rc) return rc;
    if (nlive == 0) break;
    for (int k = 0; k < set.n(); k++) nterms[k] += live[k];
    std::swap(cur, oth);
  }
  return 0;
}
int imsrg_magnus_expand(imsrg_ctx c, imsrg_op DG, imsrg_op G, int *nterms) {
  if (!c || !DG || !G || DG == G) return fail(IMSRG_EINVAL, "bad operands");
  std::vector<int> nt;
  int rc = magexp_impl(c, DG, G, all_set(c), nt);
  if (!rc && nterms) std::copy(nt.begin(), nt.end(), nterms);
  return rc;
}

// magnus_decouple (IMSRG_MAGNUS.f90:15-128) without checkpointing, split into
// set-up (:31-96) and one pass of the step loop (:98-126).  With a batched
// context every instance flows on its own: an instance whose |MBPT2| has
// fallen to conv stops stepping while the others go on.
struct imsrg_flow_s {
  imsrg_ctx_s *ctx = nullptr;
  imsrg_op_s *HS = nullptr, *G = nullptr, *H0 = nullptr, *DG = nullptr;
  int gen_kind = 0;
  double ds = 0.5;
  std::vector<double> s, crit, emp2;
  std::vector<int> steps;
};

// Set-up of magnus_decouple (IMSRG_MAGNUS.f90:31-97).  Fresh start: Omega = 0,
// DG = generator(H(0)), crit = |MBPT2|.  Checkpoint restart (:83-97): Omega(s0)
// given in G, HS = BCH(Omega, H(0)), DG = generator(HS), s = s0, steps =
// nint(s0/ds) and crit = 10 (the reference's initial value, so a step follows).
static int flow_setup(imsrg_ctx c, imsrg_op HS, imsrg_op G, int gen_kind, double ds, bool restart, double s0,
                      imsrg_flow *out) {
  if (!c || !HS || !G || HS == G || !out) return fail(IMSRG_EINVAL, "bad operands");
  if (HS->ctx != c || G->ctx != c) return fail(IMSRG_EINVAL, "operands of another context");
  if (restart && G->herm != -1) return fail(IMSRG_EINVAL, "Omega must be anti-hermitian");
  imsrg_flow_s *f = new imsrg_flow_s;
  f->ctx = c; f->HS = HS; f->G = G; f->gen_kind = gen_kind; f->ds = ds;
  const int B = c->nbatch;
  f->s.assign(B, restart ? s0 : 0.0); f->crit.assign(B, 0.0); f->emp2.assign(B, 0.0);
  f->steps.assign(B, restart ? (int)std::lround(s0 / ds) : 0);
  const ActSet all = all_set(c);
  int rc = op_new(c, 1, &f->H0);
  if (!rc) rc = op_new(c, -1, &f->DG);
  if (!rc && restart) {
    op_copy(HS, f->H0, all);
    std::vector<int> nt;
    rc = bch_impl(c, HS, G, f->H0, all, nt);
    if (!rc) rc = gen_impl(c, gen_kind, HS, f->DG, all);
    for (int k = 0; k < B; k++) f->crit[k] = 10.0;
  } else if (!rc) {
    launch_fill(c->stream, G->d, c->lay.op_size * B, 0.0);
    G->herm = -1;
    rc = gen_impl(c, gen_kind, HS, f->DG, all);
    if (!rc) {
      op_copy(HS, f->H0, all);
      rc = mbpt2_set(HS, all, f->emp2);
      for (int k = 0; k < B; k++) f->crit[k] = std::fabs(f->emp2[k]);
    }
  }
  if (rc) {
    imsrg_flow_destroy(f);
    return rc;
  }
  *out = f;
  return 0;
}
int imsrg_flow_create(imsrg_ctx c, imsrg_op HS, imsrg_op G, int gen_kind, double ds, imsrg_flow *out) {
  return flow_setup(c, HS, G, gen_kind, ds, false, 0.0, out);
}
int imsrg_flow_restart(imsrg_ctx c, imsrg_op HS, imsrg_op G, int gen_kind, double ds, double s, imsrg_flow *out) {
  return flow_setup(c, HS, G, gen_kind, ds, true, s, out);
}

// write_twobody_operator / write_omega_checkpoint and read_twobody_operator /
// read_omega_checkpoint (basic_IMSRG.f90:3069-3420) store an operator as its
// vectorize stream, neq FP64 values in an unformatted stream file, no header.
int imsrg_op_write_stream(imsrg_op op, const char *path) {
  if (!op || !path) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = single_only(op->ctx, "imsrg_op_write_stream")) return rc;
  const int64_t n = op->ctx->lay.neq;
  std::vector<double> v(n);
  int rc = imsrg_op_download(op, v.data(), n);
  if (rc) return rc;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return fail(IMSRG_EINVAL, std::string("cannot open ") + path + " for writing");
  const size_t w = std::fwrite(v.data(), sizeof(double), (size_t)n, fp);
  const int cl = std::fclose(fp);
  if (w != (size_t)n || cl != 0) return fail(IMSRG_EINVAL, std::string("short write to ") + path);
  return 0;
}
int imsrg_op_read_stream(imsrg_op op, const char *path) {
  if (!op || !path) return fail(IMSRG_EINVAL, "null argument");
  if (int rc = single_only(op->ctx, "imsrg_op_read_stream")) return rc;
  const int64_t n = op->ctx->lay.neq;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return fail(IMSRG_EINVAL, std::string("cannot open ") + path);
  std::vector<double> v(n);
  const size_t r = std::fread(v.data(), sizeof(double), (size_t)n, fp);
  const bool more = std::fgetc(fp) != EOF;
  std::fclose(fp);
  if (r != (size_t)n || more)
    return fail(IMSRG_EINVAL, std::string(path) + ": not a stream of " + std::to_string(n) + " FP64 values");
  return imsrg_op_upload(op, v.data(), n);
}

// interaction files (interaction_IO.f90): the operator's image to the host,
// the reader's writes into it, back to the device
int imsrg_read_interaction(imsrg_op H, int format, const char *path, const char *path_1b,
                           const imsrg_interaction_params *p) {
  if (!H || !path || !p) return fail(IMSRG_EINVAL, "null argument");
  if (format == IMSRG_INT_ME2B && !path_1b) return fail(IMSRG_EINVAL, "me2b needs its one-body file");
  imsrg_ctx_s *c = H->ctx;
  if (int rc = single_only(c, "imsrg_read_interaction")) return rc;
  if (H->herm != 1) return fail(IMSRG_EINVAL, "imsrg_read_interaction: H must be hermitian");
  ReadParams rp;
  rp.hw = p->hw;
  rp.com_hw = p->com_hw;
  rp.lawson_beta = p->lawson_beta;
  rp.A = p->A;
  rp.htype = p->htype;
  rp.emax_file = p->emax_file;
  rp.lmax_file = p->lmax_file;
  const Layout &Ly = c->lay;
  HIPCHK(hipMemcpyAsync(c->hstage, H->d, sizeof(double) * Ly.op_size, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::string err;
  if (read_interaction_file(Ly, format, path, path_1b, rp, c->hstage, err) != 0)
    return fail(IMSRG_EINVAL, "imsrg_read_interaction: " + err);
  HIPCHK(hipMemcpyAsync(H->d, c->hstage, sizeof(double) * Ly.op_size, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// the same readers on a host vectorize stream (no device involved)
int imsrg_read_interaction_vec(int nsp, const int *n, const int *l, const int *jj, const int *tz, const int *con,
                               int format, const char *path, const char *path_1b,
                               const imsrg_interaction_params *p, double *vec, int64_t neq) {
  if (!n || !l || !jj || !tz || !con || !path || !p || !vec) return fail(IMSRG_EINVAL, "null argument");
  if (format == IMSRG_INT_ME2B && !path_1b) return fail(IMSRG_EINVAL, "me2b needs its one-body file");
  Layout Ly;
  std::string err = Ly.build(nsp, n, l, jj, tz, con);
  if (!err.empty()) return fail(IMSRG_EINVAL, err);
  if (neq != Ly.neq) return fail(IMSRG_EINVAL, "stream length must be the basis' neq");
  ReadParams rp;
  rp.hw = p->hw;
  rp.com_hw = p->com_hw;
  rp.lawson_beta = p->lawson_beta;
  rp.A = p->A;
  rp.htype = p->htype;
  rp.emax_file = p->emax_file;
  rp.lmax_file = p->lmax_file;
  std::vector<double> img(Ly.op_size, 0.0);
  Ly.unpack_vec(vec, 1, img.data());
  if (read_interaction_file(Ly, format, path, path_1b, rp, img.data(), err) != 0)
    return fail(IMSRG_EINVAL, "imsrg_read_interaction_vec: " + err);
  Ly.pack_vec(img.data(), vec);
  return 0;
}

// rec: 8 doubles per instance (s, E0, E0+MBPT2, crit, Magnus terms, BCH
// terms, steps taken, 1 if the instance stepped in this call else 0)
int imsrg_flow_step(imsrg_flow f, double *rec) {
  if (!f) return fail(IMSRG_EINVAL, "null flow");
  imsrg_ctx_s *c = f->ctx;
  const int B = c->nbatch;
  std::vector<int> live;
  for (int k = 0; k < B; k++)
    if (f->crit[k] > kConv) live.push_back(k);
  std::vector<int> nm(B, 0), nb(B, 0);
  std::vector<double> e0(B, 0.0);
  int rc = 0;
  if (!live.empty()) {
    const ActSet a = ((int)live.size() == B) ? all_set(c) : make_set(c, 3, live, c->graphs ? B : 0);
    std::vector<int> tm, tb;
    rc = magexp_impl(c, f->DG, f->G, a, tm);
    if (rc) return rc;
    PROF(c, IMSRG_PROF_AXPBY,
         launch_axpby(c->stream, c->lay.op_size, 1.0, f->G->d, f->ds, f->DG->d, f->G->d, a.d, a.n(),
                      c->lay.op_size));  // euler_step :756-766
    rc = bch_impl(c, f->HS, f->G, f->H0, a, tb);
    if (rc) return rc;
    rc = gen_impl(c, f->gen_kind, f->HS, f->DG, a);
    if (rc) return rc;
    std::vector<double> mp;
    rc = mbpt2_set(f->HS, a, mp);
    if (rc) return rc;
    for (int m = 0; m < a.n(); m++) {
      const int k = a.h[m];
      f->s[k] += f->ds;
      f->emp2[k] = mp[m];
      f->crit[k] = std::fabs(mp[m]);
      f->steps[k]++;
      nm[k] = tm[m];
      nb[k] = tb[m];
    }
  }
  if (rec) {
    rc = imsrg_op_e0_batch(f->HS, e0.data());
    if (rc) return rc;
    std::vector<char> stepped(B, 0);
    for (int k : live) stepped[k] = 1;
    for (int k = 0; k < B; k++) {
      const double r[8] = {f->s[k], e0[k], e0[k] + f->emp2[k], f->crit[k], (double)nm[k], (double)nb[k],
                           (double)f->steps[k], stepped[k] ? 1.0 : 0.0};
      std::memcpy(rec + 8 * (size_t)k, r, sizeof(r));
    }
  }
  return 0;
}

int imsrg_flow_converged(imsrg_flow f, int *yes) {
  if (!f || !yes) return fail(IMSRG_EINVAL, "null argument");
  int all = 1;
  for (double cr : f->crit)
    if (cr > kConv) all = 0;  // do while (crit > 1e-6)
  *yes = all;
  return 0;
}

int imsrg_flow_destroy(imsrg_flow f) {
  if (!f) return 0;
  if (f->H0) imsrg_op_destroy(f->H0);
  if (f->DG) imsrg_op_destroy(f->DG);
  delete f;
  return 0;
}

int imsrg_magnus_flow(imsrg_ctx c, imsrg_op HS, imsrg_op G, int gen_kind, double ds, int max_steps,
                      double *s_out, double *e0_out, double *e0mbpt2_out, double *crit_out,
                      int *nmag_out, int *nbch_out, int *steps_out) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  if (int rc = single_only(c, "imsrg_magnus_flow")) return rc;
  imsrg_flow f = nullptr;
  int rc = imsrg_flow_create(c, HS, G, gen_kind, ds, &f);
  if (rc) return rc;
  int steps = 0, done = 0;
  imsrg_flow_converged(f, &done);
  while (!done && steps < max_steps) {
    double r[8];
    rc = imsrg_flow_step(f, r);
    if (rc) break;
    int k = steps++;
    if (s_out) s_out[k] = r[0];
    if (e0_out) e0_out[k] = r[1];
    if (e0mbpt2_out) e0mbpt2_out[k] = r[2];
    if (crit_out) crit_out[k] = r[3];
    if (nmag_out) nmag_out[k] = (int)r[4];
    if (nbch_out) nbch_out[k] = (int)r[5];
    imsrg_flow_converged(f, &done);
  }
  if (steps_out) *steps_out = steps;
  imsrg_flow_destroy(f);
  return rc;
}


// ---------------------------------------------------------------- gradient through the Magnus flow
// L = E0 of H(s_N) after nsteps steps of magnus_decouple (IMSRG_MAGNUS.f90:98-126)
// from the normal-ordered H0, differentiated by the discrete adjoint: every
// nested commutator of every BCH and Magnus series is transposed with
// adjoint_impl, the Euler step and the generator by hand; the series keep the
// term counts the forward flow chose.  Omega(s) is stored at ncheck
// checkpoints and each segment is recomputed from its checkpoint.
namespace {
struct OpSet {
  imsrg_ctx_s *c;
  std::vector<imsrg_op_s *> ops;
  explicit OpSet(imsrg_ctx_s *cc) : c(cc) {}
  ~OpSet() {
    for (auto *o : ops) imsrg_op_destroy(o);
  }
  imsrg_op_s *get(int herm) {
    imsrg_op_s *o = nullptr;
    if (op_new(c, herm, &o) != 0) return nullptr;
    ops.push_back(o);
    return o;
  }
};
}  // namespace

static void axpby_all(imsrg_ctx_s *c, double a, imsrg_op_s *A, double b, imsrg_op_s *B, imsrg_op_s *C) {
  launch_axpby(c->stream, c->lay.op_size, a, A->d, b, B->d, C->d, nullptr, 1, 0);
}

// A[0] = X0 and A[k] = [Om, A[k-1]], k = 1..K (the nested commutators of one series)
static int nested_chain(imsrg_ctx_s *c, imsrg_op_s *Om, imsrg_op_s *X0, int K, unsigned mask,
                        std::vector<imsrg_op_s *> &A) {
  const ActSet all = all_set(c);
  op_copy(X0, A[0], all);
  prep_L(c, Om, all);
  for (int k = 1; k <= K; k++) {
    A[k]->herm = -Om->herm * A[k - 1]->herm;
    int rc = comm_impl(c, Om, A[k - 1], A[k], mask, kLPrepped, all);
    if (rc) return rc;
  }
  return 0;
}

// Backward through S = sum_k w[k] A[k], A[k] = [Om, A[k-1]]: given Sbar, adds
// dL/dA[0] to X0bar and dL/dOm to Ombar.
static int nested_backward(imsrg_ctx_s *c, imsrg_op_s *Om, const std::vector<imsrg_op_s *> &A, int K,
                           const std::vector<double> &w, unsigned mask, imsrg_op_s *Sbar, imsrg_op_s *X0bar,
                           imsrg_op_s *Ombar, OpSet &tmp) {
  const int hA = A[0]->herm;
  imsrg_op_s *Ab = tmp.get(hA), *T1 = tmp.get(hA), *T2 = tmp.get(Om->herm);
  if (!Ab || !T1 || !T2) return IMSRG_ENOMEM;
  axpby_all(c, w[K], Sbar, 0.0, Sbar, Ab);
  Ab->herm = hA;
  for (int k = K; k >= 1; k--) {
    int rc = adjoint_impl(c, Om, Ab, T1, hA, mask);  // d/dA[k-1]
    if (!rc) rc = adjoint_impl(c, A[k - 1], Ab, T2, Om->herm, mask);  // d/dOm (with a minus sign)
    if (rc) return rc;
    axpby_all(c, 1.0, Ombar, -1.0, T2, Ombar);
    axpby_all(c, w[k - 1], Sbar, 1.0, T1, Ab);
  }
  axpby_all(c, 1.0, X0bar, 1.0, Ab, X0bar);
  return 0;
}

// Hbar += d<Eb, generator(H)>/dH (White or atan)
static int generator_backward(imsrg_ctx_s *c, int kind, imsrg_op_s *H, imsrg_op_s *Eb, imsrg_op_s *Hbar,
                              OpSet &tmp) {
  imsrg_op_s *G = tmp.get(1);
  if (!G) return IMSRG_ENOMEM;
  const ActSet all = all_set(c);
  launch_fill(c->stream, G->d, c->lay.op_size, 0.0);
  launch_monopole(c->stream, c->dl, H->d, c->mono, all.d, all.n());
  launch_adj_generator(c->stream, c->dl, H->d, c->mono, Eb->d, G->d, c->mbar, kind == IMSRG_GEN_WHITE ? 0 : 1);
  launch_project(c->stream, c->dl, G->d, 1, 0);
  axpby_all(c, 1.0, Hbar, 1.0, G, Hbar);
  return 0;
}

static int flow_grad_impl(imsrg_ctx_s *c, imsrg_op_s *H0, int kind, double ds, int nsteps, int ncheck,
                          double *e0_out, imsrg_op_s *H0b) {
  static const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  const ActSet all = all_set(c);
  const unsigned maskB = IMSRG_TALL, maskM = IMSRG_TALL & ~(T110 | T220);
  OpSet keep(c);
  const int K = std::max(1, (nsteps + ncheck - 1) / ncheck);
  const int nseg = (nsteps + K - 1) / K;
  imsrg_op_s *Om = keep.get(-1), *eta = keep.get(-1), *DG = keep.get(-1), *HS = keep.get(1);
  if (!Om || !eta || !DG || !HS) return IMSRG_ENOMEM;
  std::vector<imsrg_op_s *> ck(nseg, nullptr);
  std::vector<int> nm(nsteps, 0), nb(nsteps, 0);
  // forward with checkpoints (the same sequence as imsrg_flow_step)
  int rc = gen_impl(c, kind, H0, eta, all);
  for (int n = 0; n < nsteps && !rc; n++) {
    if (n % K == 0) {
      ck[n / K] = keep.get(-1);
      if (!ck[n / K]) return IMSRG_ENOMEM;
      op_copy(Om, ck[n / K], all);
    }
    std::vector<int> t;
    op_copy(eta, DG, all);
    rc = magexp_impl(c, DG, Om, all, t);
    if (rc) break;
    nm[n] = t[0];
    axpby_all(c, 1.0, Om, ds, DG, Om);
    rc = bch_impl(c, HS, Om, H0, all, t);
    if (rc) break;
    nb[n] = t[0];
    rc = gen_impl(c, kind, HS, eta, all);
  }
  if (rc) return rc;
  if (e0_out) {
    rc = imsrg_op_e0_batch(HS, e0_out);
    if (rc) return rc;
  }
  // backward
  imsrg_op_s *Hb = keep.get(1), *Omb = keep.get(-1), *DGb = keep.get(-1), *etab = keep.get(-1);
  if (!Hb || !Omb || !DGb || !etab) return IMSRG_ENOMEM;
  std::vector<imsrg_op_s *> chain(41);
  for (auto &a : chain)
    if (!(a = keep.get(1))) return IMSRG_ENOMEM;
  launch_fill(c->stream, H0b->d, c->lay.op_size, 0.0);
  H0b->herm = 1;
  launch_fill(c->stream, Hb->d, c->lay.op_size, 0.0);
  launch_fill(c->stream, Hb->d, 1, 1.0);  // dL/dE0(H_N) = 1
  launch_fill(c->stream, Omb->d, c->lay.op_size, 0.0);
  std::vector<double> wB(41), wM(7);
  for (int k = 0; k < 7; k++) wM[k] = cof[k];
  wB[0] = 1.0;
  for (int k = 1; k < 41; k++) wB[k] = wB[k - 1] / k;
  for (int seg = nseg - 1; seg >= 0; seg--) {
    const int n0 = seg * K, n1 = std::min(n0 + K, nsteps);
    OpSet segops(c);
    std::vector<imsrg_op_s *> Oms(n1 - n0 + 1), etas(n1 - n0), Hs(n1 - n0);
    for (auto &o : Oms)
      if (!(o = segops.get(-1))) return IMSRG_ENOMEM;
    for (int i = 0; i < n1 - n0; i++) {
      etas[i] = segops.get(-1);
      Hs[i] = segops.get(1);
      if (!etas[i] || !Hs[i]) return IMSRG_ENOMEM;
    }
    // recompute Omega, H and eta over the segment
    op_copy(ck[seg], Oms[0], all);
    std::vector<int> t;
    if (n0 == 0) op_copy(H0, Hs[0], all);
    else if ((rc = bch_impl(c, Hs[0], Oms[0], H0, all, t))) return rc;
    for (int i = 0; i < n1 - n0; i++) {
      if ((rc = gen_impl(c, kind, Hs[i], etas[i], all))) return rc;
      op_copy(etas[i], DG, all);
      if ((rc = magexp_impl(c, DG, Oms[i], all, t))) return rc;
      axpby_all(c, 1.0, Oms[i], ds, DG, Oms[i + 1]);
      if (i + 1 < n1 - n0 && (rc = bch_impl(c, Hs[i + 1], Oms[i + 1], H0, all, t))) return rc;
    }
    for (int n = n1; n > n0; n--) {
      const int i = n - n0;  // Oms[i] = Omega_n
      OpSet tmp(c);
      // H_n = BCH(Omega_n, H0)
      for (int k = 0; k <= nb[n - 1]; k++) chain[k]->herm = 1;
      if ((rc = nested_chain(c, Oms[i], H0, nb[n - 1], maskB, chain))) return rc;
      if ((rc = nested_backward(c, Oms[i], chain, nb[n - 1], wB, maskB, Hb, H0b, Omb, tmp))) return rc;
      // Omega_n = Omega_{n-1} + ds dOmega_{n-1}
      axpby_all(c, ds, Omb, 0.0, Omb, DGb);
      // dOmega_{n-1} = MAGNUS_EXPAND(eta_{n-1}, Omega_{n-1})
      launch_fill(c->stream, etab->d, c->lay.op_size, 0.0);
      for (int k = 0; k <= nm[n - 1]; k++) chain[k]->herm = -1;
      if ((rc = nested_chain(c, Oms[i - 1], etas[i - 1], nm[n - 1], maskM, chain))) return rc;
      if ((rc = nested_backward(c, Oms[i - 1], chain, nm[n - 1], wM, maskM, DGb, etab, Omb, tmp))) return rc;
      // eta_{n-1} = generator(H_{n-1}); H_{n-1}'s cotangent starts here
      launch_fill(c->stream, Hb->d, c->lay.op_size, 0.0);
      if ((rc = generator_backward(c, kind, Hs[i - 1], etab, (n - 1 == 0) ? H0b : Hb, tmp))) return rc;
      if ((rc = (int)hipStreamSynchronize(c->stream) == hipSuccess ? 0 : fail(IMSRG_EDEVICE, "gradient step")))
        return rc;
    }
  }
  launch_project(c->stream, c->dl, H0b->d, 1, 1);  // dE0(s_N)/dE0(0) = 1 stays
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int imsrg_flow_gradient(imsrg_ctx c, imsrg_op H0, int gen_kind, double ds, int nsteps, int ncheck, double *e0,
                        imsrg_op H0bar) {
  if (!c || !H0 || !H0bar || H0 == H0bar) return fail(IMSRG_EINVAL, "bad operands");
  if (int rc = single_only(c, "imsrg_flow_gradient")) return rc;
  if (gen_kind != IMSRG_GEN_WHITE && gen_kind != IMSRG_GEN_ATAN)
    return fail(IMSRG_EINVAL, "gradient: White or atan generator");
  if (nsteps < 1 || ncheck < 1) return fail(IMSRG_EINVAL, "nsteps and ncheck must be >= 1");
  if (H0->herm != 1) return fail(IMSRG_EINVAL, "H0 must be hermitian");
  return flow_grad_impl(c, H0, gen_kind, ds, nsteps, ncheck, e0, H0bar);
}

int imsrg_normal_order_adjoint(imsrg_op Cbar, imsrg_op Hbar) {
  if (!Cbar || !Hbar || Cbar == Hbar || Cbar->ctx != Hbar->ctx) return fail(IMSRG_EINVAL, "bad operands");
  imsrg_ctx_s *c = Cbar->ctx;
  if (int rc = single_only(c, "imsrg_normal_order_adjoint")) return rc;
  op_copy(Cbar, Hbar, all_set(c));
  launch_adj_normal_order(c->stream, c->dl, Cbar->d, Hbar->d, Cbar->herm);
  launch_project(c->stream, c->dl, Hbar->d, Cbar->herm, 0);
  HIPCHK(hipGetLastError());
  return 0;
}

// gradient with respect to the vectorize stream: transpose of repackage
int imsrg_op_download_cotangent(imsrg_op op, double *vec, int64_t n) {
  if (!op || !vec) return fail(IMSRG_EINVAL, "null argument");
  imsrg_ctx_s *c = op->ctx;
  if (n != c->lay.neq * c->nbatch) return fail(IMSRG_EINVAL, "length must be neq * nbatch");
  HIPCHK(hipMemcpyAsync(c->hstage, op->d, sizeof(double) * c->lay.op_size * c->nbatch, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nbatch; k++)
    c->lay.unpack_vec_transpose(c->hstage + (size_t)k * c->lay.op_size, op->herm, vec + (size_t)k * c->lay.neq);
  return 0;
}

// ---------------------------------------------------------------- profiling
int imsrg_ctx_profile(imsrg_ctx c, int enable) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  c->prof_on = enable != 0;
  c->prof_mask = ~0u;
  return 0;
}

int imsrg_ctx_set_concurrency(imsrg_ctx c, int enable) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipStreamSynchronize(c->stream2));
  c->dual = enable != 0;
  return 0;
}

int imsrg_ctx_profile_classes(imsrg_ctx c, unsigned mask) {
  if (!c) return fail(IMSRG_EINVAL, "null context");
  c->prof_on = mask != 0;
  c->prof_mask = mask;
  return 0;
}

int imsrg_ctx_profile_read(imsrg_ctx c, double *ms, long long *launches, int nclass, int reset) {
  if (!c || !ms || !launches) return fail(IMSRG_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto &r : c->pend) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      c->prof_ms[r.cls] += t;
      c->prof_n[r.cls] += 1;
    }
  }
  c->pend.clear();
  c->evused = 0;
  for (int k = 0; k < nclass && k < IMSRG_PROF_NCLASS; k++) {
    ms[k] = c->prof_ms[k];
    launches[k] = c->prof_n[k];
  }
  if (reset)
    for (int k = 0; k < IMSRG_PROF_NCLASS; k++) {
      c->prof_ms[k] = 0.0;
      c->prof_n[k] = 0;
    }
  return 0;
}
