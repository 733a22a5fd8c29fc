// engine_internal.hpp -- the context and operator objects behind the C ABI
// handles, shared by the translation units that implement entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/imsrg_mi355x.h"
#include "device.hpp"
#include "layout.hpp"

using namespace imsrg;

struct TensorSpace;  // tensor.hip

inline thread_local std::string g_err;
inline int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(IMSRG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

struct imsrg_op_s;

struct imsrg_ctx_s {
  Layout lay;
  DevLayout dl{};
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of a commutator: the pp/hh chain (122, pp/hh GEMM, one-body
  // partials) runs there while the ph chain runs on `stream`; fork/join events
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // an event recorded on `stream` after its last work (a series gate), which
  // the next commutator's fork waits on instead of recording ev_fork; cleared
  // by that commutator
  hipEvent_t fork_mark = nullptr;
  bool dual = true;        // use stream2 (imsrg_ctx_set_concurrency)
  // BCH / Magnus series: the gate kernel after nested commutator n writes one
  // live flag per series position into ring slot n % kGateRing of pinned host
  // memory and ev_gate marks it; the host reads slot n-1 while commutator n runs
  static constexpr int kGateRing = 4;
  int *live_h = nullptr, *live_d = nullptr;  // host and device views, kGateRing * nbatch
  hipEvent_t ev_gate[kGateRing] = {};
  bool capturing = false;  // graph capture records one stream only
  std::vector<void *> dallocs;
  int nbatch = 1;           // independent instances per operator
  double *wpp = nullptr;    // 221 intermediate W1 - pm*W1^T (upper triangles), vsize per instance
  double *cc[2] = {nullptr, nullptr};
  double *wph = nullptr;
  double *w122 = nullptr;   // 122 set weights (A, -B per slot), per instance
  double *tph = nullptr;    // ph J' terms
  double *partials = nullptr;
  double *obpart = nullptr;  // one-body 121/221 partial sums, nsame * ob_nchunk per instance
  double *npart = nullptr;   // sums of squares of a commutator result, nmirror + nob per instance
  double *scal = nullptr;   // device scalars, kScal per instance
  double *coef_d = nullptr; // series coefficient of the next graph replay
  double *hscal = nullptr;  // pinned host mirror of scal
  double *mono = nullptr;
  double *mbar = nullptr;   // generator-transpose scratch (nsp * nsp)
  double *hstage = nullptr; // pinned staging: max(op_size * nbatch, ccsize) doubles
  SegProb *pp_probs = nullptr, *ph_probs = nullptr;
  GemmProb *adj_pp_probs = nullptr, *adj_ph_probs = nullptr;  // adjoint products
  int4 *adj_pp_tiles = nullptr, *adj_ph_tiles = nullptr;
  int n_adj_pp_tiles = 0, n_adj_ph_tiles = 0;
  int4 *pp_tiles = nullptr, *ph_tiles = nullptr;        // 64 x 64 tiles
  int n_pp_tiles = 0, n_ph_tiles = 0;
  int4 *pp_tiles_w = nullptr, *ph_tiles_w = nullptr;    // 128 x 64 tiles (FP64 path)
  int n_pp_tiles_w = 0, n_ph_tiles_w = 0;
  // FP64 tile shape per GEMM, chosen by timing both on the first FP64
  // commutator of the context: -1 not yet tuned, 0 64x64, 1 128x64
  int shape_pp = -1, shape_ph = -1;
  float tune_ms[4] = {0, 0, 0, 0};                      // pp 64, pp 128x64, ph 64, ph 128x64
  int prec = IMSRG_PREC_FP64;
  // int8-sliced GEMM operands (IMSRG_PREC_INT8X5 / INT8X4), built on first use:
  // planes and row scales of operand 0 (L) and 1 (R), V rows and Pandya rows
  bool i8_ready = false;
  I8Layout i8{};
  signed char *i8v[2] = {nullptr, nullptr}, *i8c[2] = {nullptr, nullptr};
  double *i8vs[2] = {nullptr, nullptr}, *i8cs[2] = {nullptr, nullptr};
  double *i8pm = nullptr;  // partial row maxima of a Pandya matrix (k_rowmax_cc)
  const I8Prob *pp_i8 = nullptr, *ph_i8 = nullptr;
  const int4 *ph_i8_tiles = nullptr;  // 32 x 32 tiles of the int8 Z product, four per workgroup
  int n_ph_i8_tiles = 0;
  // active-instance tables: slot 0 = all instances (constant), 1.. = driver work sets
  static constexpr int kSlots = 4;
  int *act_d[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  int *act_h[kSlots] = {nullptr, nullptr, nullptr, nullptr};
  imsrg_op_s *tmp[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // Hartree-Fock workspace (allocated on first use): D, rho, gamma, Fock and
  // bare one-body matrices (nsp x nsp), eigenvalues, e_HF diagonals,
  // per-group eigenvalue changes, and [crit, e_HF, converged flag, iterations]
  double *hf_D = nullptr, *hf_rho = nullptr, *hf_Vg = nullptr, *hf_Fm = nullptr, *hf_Tm = nullptr;
  double *hf_eig = nullptr, *hf_tv = nullptr, *hf_cp = nullptr, *hf_sc = nullptr;
  // this device's copy of the static Hartree-Fock scratch array (the
  // allocation-free entry point imsrg_hartree_fock_noalloc)
  double *hf_scratch = nullptr;
  // per-kernel-class HIP event timing on this context's stream
  bool prof_on = false;
  unsigned prof_mask = ~0u;  // kernel classes bracketed while profiling
  // hipGraph replay of whole commutators: one executable graph per distinct
  // (operands, herms, mask, Pandya reuse, instance table, precision, series
  // epilogue, FP64 tile shapes)
  bool graphs = false;
  typedef std::tuple<const void *, const void *, const void *, int, int, int, unsigned, int, const void *, int, int,
                     const void *, double, int, int, int>
      GraphKey;
  std::map<GraphKey, hipGraphExec_t> gexec;
  long long graph_launches = 0, graph_builds = 0;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  struct Rec { int cls; hipEvent_t a, b; };
  std::vector<Rec> pend;
  double prof_ms[IMSRG_PROF_NCLASS] = {};
  long long prof_n[IMSRG_PROF_NCLASS] = {};
  // tensor operator spaces (layout, tables, workspaces) per (rank, parity change)
  std::map<std::pair<int, int>, std::shared_ptr<TensorSpace>> tspaces;
};

struct imsrg_op_s {
  imsrg_ctx_s *ctx = nullptr;
  double *d = nullptr;
  int herm = 1;
};

// ---------------------------------------------------------------- helpers
// int8 slices per element of a precision (0: not an int8 precision)
inline int i8_planes(int prec) {
  return prec == IMSRG_PREC_INT8X5 ? 5 : prec == IMSRG_PREC_INT8X4 ? 4 : 0;
}
template <class T>
inline int upload_table(imsrg_ctx_s *c, const std::vector<T> &v, const T **out) {
  size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return fail(IMSRG_ENOMEM, "hipMalloc table failed");
  c->dallocs.push_back(p);
  if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (const T *)p;
  return 0;
}
inline int dalloc(imsrg_ctx_s *c, size_t ndouble, double **out) {
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(ndouble, 1) * sizeof(double)) != hipSuccess)
    return fail(IMSRG_ENOMEM, "hipMalloc workspace failed (" + std::to_string(ndouble * 8) + " bytes)");
  c->dallocs.push_back(p);
  *out = (double *)p;
  return 0;
}

