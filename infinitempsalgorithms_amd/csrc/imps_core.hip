// libimps runtime: device-resident VUMPS state, the effective-Hamiltonian and
// transfer operators built on the MFMA GEMM, Krylov solvers whose vectors stay in HBM
// (FP64 Gram-Schmidt, host-side projected problems), QRpos/LQpos gauge fixing, the
// Schur-MPO environment solves and the VUMPS driver, behind the C ABI of imps.h.
//
// Reference files cited below are relative to the InfiniteMPSAlgorithms.jl root.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <tuple>
#include <string>
#include <thread>
#include <vector>

#include "imps.h"
#include "imps_internal.h"

using namespace imps;

namespace {

struct ImpsError : std::runtime_error {
  int code;
  ImpsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ImpsError(IMPS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

#define RBCHK(expr)                                                                          \
  do {                                                                                       \
    rocblas_status s_ = (expr);                                                              \
    if (s_ != rocblas_status_success)                                                        \
      throw ImpsError(IMPS_E_SOLVER, std::string(#expr) + ": rocblas status " +              \
                                         std::to_string((int)s_));                           \
  } while (0)

static void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw ImpsError(code, msg);
}

struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct imps_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle rb = nullptr;
  std::string err;
  // model: square Schur tensors (w, d, w, d) per site
  int N = 0, w = 0, d = 0;
  // physical dimension of each site; sites with pdim[l] < d carry empty sectors s >= pdim[l]
  // (imps_set_mpo_chain), which no state may populate
  std::vector<int> pdim;
  std::vector<std::vector<double>> Wh;
  std::vector<double*> Wm;  // per-site H_AC mixing matrix (w d) x (w d)
  // state
  int D = 0;
  std::vector<double*> AL, AR, AC, C, pAC, pC, pAL;
  // the state is complex: AL, AR, AC, C (and the environments) hold the real embedding of the
  // complex tensors of bond dimension D / 2 (launch_cplx_embed), C and AC scaled by 1/sqrt(2) so
  // that every trace-normalised quantity (energy, Galerkin error, expectation values) equals the
  // complex one
  bool cplx = false;
  // active bond dimensions (bond l right of site l) when they differ from bond to bond: the
  // tensors are held at the largest one, D, with zeros outside the active blocks; empty for a
  // uniform state
  std::vector<int> chi;
  // environments
  bool envs_valid = false;
  std::vector<double*> GL, GR;
  // per site: level 0 of GL / level w-1 of GR is exactly the identity (the Schur structure of
  // DMRGCache pins them); the effective operators then apply those levels as copies
  std::vector<char> gl_id, gr_id;
  // site whose GL/GR split-BF16 planes are current for the eigensolve in progress (-1: none;
  // then a split-BF16 H_AC application splits them itself)
  int planes_site = -1;
  int hc_planes_site = -1;  // the same for H_C (leftenv(site + 1), rightenv(site))
  // workspaces and pinned host scratch for reductions
  std::map<std::string, DevBuf> ws;
  double* hpart = nullptr;
  // pinned buffer of kSmallPinned doubles for small readbacks outside the Gram-Schmidt steps
  // (trace_prod, Cholesky breakdown flags, gradient sums)
  double* hscal = nullptr;
  // pinned staging of host<->device copies of caller arrays (allocated on first use), so that no
  // DMA ever targets pageable memory
  double* hstage = nullptr;
  double* dpart = nullptr;
  hipEvent_t ev_coef = nullptr;  // marks the arrival of Gram-Schmidt coefficients in hpart
  // counters
  long hac_applies = 0, hc_applies = 0;
  // workers for the concurrent per-site eigensolves of the local update: each has its own
  // stream, workspaces and reduction buffers and borrows this context's model, state and
  // environments (device pointers); workers have no workers of their own
  std::vector<imps_ctx*> workers;
  hipEvent_t ev_ready = nullptr;  // parent: state and environments enqueued
  hipEvent_t ev_done = nullptr;   // worker: its solve enqueued
  std::set<long> pretuned;        // (D, w, d, prec) whose H_AC/H_C GEMM classes were tuned uncontended
  // captured kernel sequences (hipGraph) keyed by the sequence tag and the buffers it touches;
  // exec == nullptr: seen once, captured on the next run
  struct Graph {
    std::vector<uintptr_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<Graph> graphs;
};

static thread_local std::string g_create_error;

// Device scratch of imps_hac_apply_noalloc (T1, T2 and the split-K workspace of its GEMMs), for
// callers that cannot allocate: every GPU the library runs on holds one copy, 128 MiB, until the
// process ends.  Host code never uses the symbol's name as an address: the per-device address
// comes from hipGetSymbolAddress (scratch_ptr) and only that pointer reaches kernels.
constexpr size_t kScratchDoubles = (size_t)1 << 24;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace imps {
// message for imps_last_error(NULL): errors of calls that have no context
void set_global_error(const std::string& msg) { g_create_error = msg; }
// message for imps_last_error(c)
void set_context_error(imps_ctx* c, const std::string& msg) { c->err = msg; }
bool context_model(const imps_ctx* c, int* nsites, int* d) {
  if (c->N <= 0 || c->w <= 0) return false;
  *nsites = c->N;
  *d = c->d;
  return true;
}
}  // namespace imps

namespace {

// ------------------------------------------------------------------ memory helpers

constexpr size_t kSmallPinned = 4096;        // doubles in imps_ctx::hscal
constexpr size_t kStageDoubles = 1u << 19;   // 4 MiB pinned staging chunk

double* dalloc(size_t n) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)));
  return (double*)p;
}

void dfree(double*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void clear_graphs(imps_ctx* c) {
  for (auto& g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

double* ws(imps_ctx* c, const std::string& name, size_t n) {
  DevBuf& b = c->ws[name];
  if (b.n < n) {
    HIPCHK(hipStreamSynchronize(c->stream));
    clear_graphs(c);  // captured sequences may hold the old buffer
    dfree(b.p);
    b.p = dalloc(n);
    b.n = n;
  }
  return b.p;
}

double* stage(imps_ctx* c) {
  if (!c->hstage) HIPCHK(hipHostMalloc((void**)&c->hstage, kStageDoubles * sizeof(double), hipHostMallocDefault));
  return c->hstage;
}

// Caller (pageable) host memory <-> device, through the pinned staging chunk: the host copies
// into / out of pinned memory and the DMA engine only ever sees pinned pages.  Both return with
// the copy complete (the staging chunk is reused).
void h2d_bytes(imps_ctx* c, void* dst, const void* src, size_t bytes) {
  char* st = (char*)stage(c);
  const size_t chunk = kStageDoubles * sizeof(double);
  for (size_t off = 0; off < bytes; off += chunk) {
    const size_t m = std::min(chunk, bytes - off);
    HIPCHK(hipStreamSynchronize(c->stream));  // the previous chunk has left the staging buffer
    std::memcpy(st, (const char*)src + off, m);
    HIPCHK(hipMemcpyAsync((char*)dst + off, st, m, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
}

void d2h_bytes(imps_ctx* c, void* dst, const void* src, size_t bytes) {
  char* st = (char*)stage(c);
  const size_t chunk = kStageDoubles * sizeof(double);
  for (size_t off = 0; off < bytes; off += chunk) {
    const size_t m = std::min(chunk, bytes - off);
    HIPCHK(hipMemcpyAsync(st, (const char*)src + off, m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memcpy((char*)dst + off, st, m);
  }
}

void h2d(imps_ctx* c, double* dst, const double* src, size_t n) { h2d_bytes(c, dst, src, n * sizeof(double)); }
void d2h(imps_ctx* c, double* dst, const double* src, size_t n) { d2h_bytes(c, dst, src, n * sizeof(double)); }

// n <= kSmallPinned doubles from the device into the context's pinned scalar buffer (synchronous);
// returns the pinned copy
const double* d2h_small(imps_ctx* c, const void* src, size_t bytes) {
  require(bytes <= kSmallPinned * sizeof(double), IMPS_E_ARG, "d2h_small: readback too large");
  HIPCHK(hipMemcpyAsync(c->hscal, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return c->hscal;
}

void d2d(imps_ctx* c, double* dst, const double* src, size_t n) {
  if (dst == src || n == 0) return;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}

// Graph capture of fixed kernel sequences (process-wide switch, on by default; 0 runs every
// sequence eagerly, for tests and measurement).
std::atomic<int> g_graphs{1};

// Runs `enqueue` (kernel launches and device-to-device copies on c->stream only: no host
// synchronisation, no allocation, no memset) through a hipGraph cached under `key`: the first run
// of a key is eager (the GEMM tuner and the workspaces see it), the second captures the sequence
// and launches the graph, later runs replay it.  The key must name every buffer the sequence
// touches that can change between runs; workspaces are covered because a workspace that grows
// drops every captured graph of the context.
void run_graphed(imps_ctx* c, const std::vector<uintptr_t>& key, const std::function<void()>& enqueue) {
  if (g_graphs.load(std::memory_order_relaxed) == 0) {
    enqueue();
    return;
  }
  for (auto& g : c->graphs)
    if (g.key == key) {
      if (g.exec) {
        HIPCHK(hipGraphLaunch(g.exec, c->stream));
        return;
      }
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
      try {
        enqueue();
      } catch (...) {
        (void)hipStreamEndCapture(c->stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      HIPCHK(hipStreamEndCapture(c->stream, &graph));
      hipGraphExec_t exec = nullptr;
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      HIPCHK(e);
      g.exec = exec;
      HIPCHK(hipGraphLaunch(g.exec, c->stream));
      return;
    }
  if (c->graphs.size() >= 32) clear_graphs(c);
  c->graphs.push_back(imps_ctx::Graph{key, nullptr});
  enqueue();
}

// ------------------------------------------------------------------ BLAS-1 (FP64)

constexpr int kMaxDotBlocks = 1024;
constexpr int kMaxVec = 64;

void multidot(imps_ctx* c, size_t n, int nv, const double* X, size_t ldx, const double* y,
              double* out) {
  require(nv <= kMaxVec, IMPS_E_ARG, "multidot: too many vectors");
  const int nb = dot_blocks(n);
  HIPCHK(launch_multidot(n, nv, X, ldx, y, c->dpart, c->stream));
  HIPCHK(hipMemcpyAsync(c->hpart, c->dpart, sizeof(double) * nb * nv, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int v = 0; v < nv; ++v) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += c->hpart[(size_t)v * nb + b];
    out[v] = s;
  }
}

double dot(imps_ctx* c, size_t n, const double* x, const double* y) {
  double r;
  multidot(c, n, 1, x, n, y, &r);
  return r;
}

double nrm2(imps_ctx* c, size_t n, const double* x) { return std::sqrt(std::max(0.0, dot(c, n, x, x))); }

void multiaxpy(imps_ctx* c, size_t n, int nv, const double* X, size_t ldx, const double* coef,
               double alpha, double* y) {
  HIPCHK(launch_multiaxpy(n, nv, X, ldx, coef, alpha, y, c->stream));
}

void scal(imps_ctx* c, size_t n, double a, double* x) { HIPCHK(launch_scale(n, a, x, c->stream)); }

void axpby(imps_ctx* c, size_t n, double a, const double* x, double b, double* y) {
  HIPCHK(launch_axpby(n, a, x, b, y, c->stream));
}

void fill(imps_ctx* c, size_t n, double v, double* x) { HIPCHK(launch_fill(n, v, x, c->stream)); }

// Classical Gram-Schmidt (two passes) of w against V[0..nv) followed by the
// normalisation of w, with the coefficients reduced and applied on the device.
// cgs2_enqueue only enqueues (the coefficients travel to pinned host memory behind an
// event), so that work depending on the normalised w can be enqueued before the host
// reads them; cgs2_wait blocks until they arrived: h[0..nv) receives the summed
// coefficients and *beta the norm of w before normalisation (w is left as is when that norm
// is zero).
void cgs2_enqueue(imps_ctx* c, size_t n, int nv, const double* V, double* w) {
  require(nv <= kMaxVec, IMPS_E_ARG, "cgs2: too many vectors");
  double* coef = ws(c, "ks_coef", 2 * kMaxVec + 2);
  const int nb = dot_blocks(n);
  HIPCHK(launch_multidot(n, nv, V, n, w, c->dpart, c->stream));
  HIPCHK(launch_partials_reduce(nv, nb, c->dpart, coef, c->stream));
  if (nv <= 32) {
    // fused sweeps: three reads of the basis instead of four
    HIPCHK(launch_axpy_reduce(n, nv, V, n, coef, w, c->dpart, true, c->stream));
    HIPCHK(launch_partials_reduce(nv, nb, c->dpart, coef + kMaxVec, c->stream));
    HIPCHK(launch_axpy_reduce(n, nv, V, n, coef + kMaxVec, w, c->dpart, false, c->stream));
  } else {
    HIPCHK(launch_multiaxpy_dev(n, nv, V, n, coef, -1.0, 1.0, w, c->stream));
    HIPCHK(launch_multidot(n, nv, V, n, w, c->dpart, c->stream));
    HIPCHK(launch_partials_reduce(nv, nb, c->dpart, coef + kMaxVec, c->stream));
    HIPCHK(launch_multiaxpy_dev(n, nv, V, n, coef + kMaxVec, -1.0, 1.0, w, c->stream));
    HIPCHK(launch_multidot(n, 1, w, n, w, c->dpart, c->stream));
  }
  HIPCHK(launch_partials_reduce(1, nb, c->dpart, coef + 2 * kMaxVec, c->stream));
  HIPCHK(launch_scale_inv_sqrt(n, w, coef + 2 * kMaxVec, c->stream));
  HIPCHK(hipMemcpyAsync(c->hpart, coef, sizeof(double) * (2 * kMaxVec + 1), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipEventRecord(c->ev_coef, c->stream));
}

void cgs2_wait(imps_ctx* c, int nv, double* h, double* beta) {
  HIPCHK(hipEventSynchronize(c->ev_coef));
  for (int i = 0; i < nv; ++i) h[i] = c->hpart[i] + c->hpart[kMaxVec + i];
  *beta = std::sqrt(std::max(0.0, c->hpart[2 * kMaxVec]));
}

void cgs2_normalize(imps_ctx* c, size_t n, int nv, const double* V, double* w, double* h, double* beta) {
  cgs2_enqueue(c, n, nv, V, w);
  cgs2_wait(c, nv, h, beta);
}

// tr(l v) = sum_{a,b} l[a + n b] v[b + n a] for n x n contiguous l and v: block partials
// reduced on the device in a fixed order, one double read back into the pinned scalar buffer
// (not hpart, where a speculative Gram-Schmidt readback may be in flight)
double trace_prod(imps_ctx* c, int n, const double* l, const double* v) {
  int np = 0;
  const size_t maxp = (size_t)((n + 31) / 32) * ((n + 31) / 32);
  double* part = ws(c, "tp_part", maxp + 1);
  HIPCHK(launch_trace_prod_parts(n, l, n, v, n, part, &np, c->stream));
  HIPCHK(launch_partials_reduce(1, np, part, part + maxp, c->stream));
  return d2h_small(c, part + maxp, sizeof(double))[0];
}

// regularize!(v, l, I) (envs.jl:28-32) with the identity as rvec: v -= tr(l v) I, entirely on the
// device (two launches, no host round trip), v (n x n, leading dimension ldv), l contiguous
// (nact < n: the identity of the bond's active block, for bonds held below the largest dimension)
void regularize_dev(imps_ctx* c, int n, double* v, long ldv, const double* l, int nact = -1) {
  int np = 0;
  const size_t maxp = (size_t)((n + 31) / 32) * ((n + 31) / 32);
  double* part = ws(c, "rgd_part", maxp);
  HIPCHK(launch_trace_prod_parts(n, l, n, v, ldv, part, &np, c->stream));
  HIPCHK(launch_reduce_add_diag(np, part, -1.0, v, ldv, (nact >= 0 && nact < n) ? nact : n, nullptr, c->stream));
}

// A named roctx range over a scope: rocprofv3 --marker-trace records the phases of the solvers'
// iterations; without a tool attached the calls return at once.
struct RoctxRange {
  explicit RoctxRange(const char* name) { roctxRangePushA(name); }
  ~RoctxRange() { roctxRangePop(); }
  RoctxRange(const RoctxRange&) = delete;
  RoctxRange& operator=(const RoctxRange&) = delete;
};

// ------------------------------------------------------------------ GEMM wrappers

// GEMM tile autotuning (process-wide, on by default): the first GEMM of each shape, stride and
// alignment class times every bit-identical tile variant (one warm-up launch each, then the best of
// three single launches, into a scratch output with beta = 0) on the caller's stream and keeps the
// fastest for the rest of the process.  Shapes for which launch_gemm splits K keep its choice.
// Mode: 0 fixed heuristic, 1 autotune, 2 + v tile variant v wherever it is valid (measurement
// and tests; other shapes autotune).
std::atomic<int> g_gemm_autotune{1};

typedef std::tuple<int, int, int, int, int, long, long, long, long, long, long, int> TuneKey;

struct TuneTable {
  std::mutex mu;
  std::map<TuneKey, int> best;
};

TuneTable& tune_table() {
  static TuneTable t;
  return t;
}

int tuned_variant(imps_ctx* c, const GemmDesc& g, int prec, const int* ids, int nv) {
  const int align = (((uintptr_t)g.A & 15) ? 1 : 0) | (((uintptr_t)g.B & 15) ? 2 : 0);
  const TuneKey key(prec, g.M, g.N, g.K, g.Kc, g.sAm, g.sAk, g.sAc, g.sBk, g.sBn, g.sBc, align);
  TuneTable& tt = tune_table();
  {
    std::lock_guard<std::mutex> lk(tt.mu);
    auto it = tt.best.find(key);
    if (it != tt.best.end()) return it->second;
  }
  GemmDesc t = g;
  t.C = ws(c, "tune_C", (size_t)g.M * g.N);
  t.ldc = g.M;
  t.beta = 0.0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // one warm-up launch each, then three rounds over the variants in turn, one timed launch per
  // variant and round; the minimum per variant decides (rounds in turn so that clock changes
  // during tuning affect every variant alike)
  float vms[8];
  for (int k = 0; k < nv; ++k) {
    vms[k] = std::numeric_limits<float>::infinity();
    HIPCHK(launch_gemm_variant(t, prec, ids[k], c->stream));
  }
  for (int round = 0; round < 3; ++round)
    for (int k = 0; k < nv; ++k) {
      HIPCHK(hipEventRecord(e0, c->stream));
      HIPCHK(launch_gemm_variant(t, prec, ids[k], c->stream));
      HIPCHK(hipEventRecord(e1, c->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      vms[k] = std::min(vms[k], ms);
    }
  int best = ids[0];
  float best_ms = vms[0];
  for (int k = 1; k < nv; ++k)
    if (vms[k] < best_ms) {
      best_ms = vms[k];
      best = ids[k];
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(tt.mu);
  tt.best.emplace(key, best);
  return best;
}

// the autotuned variant of a GEMM class when one is already known, else -1 (never tunes)
int cached_variant(const GemmDesc& g, int prec) {
  const int align = (((uintptr_t)g.A & 15) ? 1 : 0) | (((uintptr_t)g.B & 15) ? 2 : 0);
  const TuneKey key(prec, g.M, g.N, g.K, g.Kc, g.sAm, g.sAk, g.sAc, g.sBk, g.sBn, g.sBc, align);
  TuneTable& tt = tune_table();
  std::lock_guard<std::mutex> lk(tt.mu);
  auto it = tt.best.find(key);
  return it == tt.best.end() ? -1 : it->second;
}

// GEMM without any allocation: the split-K workspace (when the shape takes one) is W, sized by
// the caller with gemm_workspace; the tile is a forced variant, the autotuned one when it is already
// known, or the fixed heuristic -- all of which sum over K in the same order, so the result equals
// gemm()'s bit for bit.
void gemm_prealloc(imps_ctx* c, const GemmDesc& g0, int prec, double* W) {
  GemmDesc g = g0;
  g.W = gemm_workspace(g.M, g.N, g.K) > 0 ? W : nullptr;
  const int mode = g_gemm_autotune.load(std::memory_order_relaxed);
  if (mode != 0) {
    int ids[8];
    const int nv = gemm_tile_variants(g, prec, ids);
    if (nv > 1) {
      int v = -1;
      for (int k = 0; k < nv && mode >= 2; ++k)
        if (ids[k] == mode - 2) v = ids[k];
      if (v < 0) v = cached_variant(g, prec);
      if (v >= 0) {
        HIPCHK(launch_gemm_variant(g, prec, v, c->stream));
        return;
      }
    }
  }
  HIPCHK(launch_gemm(g, prec, c->stream));
}

void gemm(imps_ctx* c, const GemmDesc& g0, int prec) {
  GemmDesc g = g0;
  const size_t nw = gemm_workspace(g.M, g.N, g.K);
  if (nw > 0 && g.W == nullptr) g.W = ws(c, "gemm_splitk", nw);
  const int mode = g_gemm_autotune.load(std::memory_order_relaxed);
  if (mode != 0) {
    int ids[8];
    const int nv = gemm_tile_variants(g, prec, ids);
    if (nv > 1) {
      int v = -1;
      for (int k = 0; k < nv && mode >= 2; ++k)
        if (ids[k] == mode - 2) v = ids[k];
      if (v < 0) v = tuned_variant(c, g, prec, ids, nv);
      HIPCHK(launch_gemm_variant(g, prec, v, c->stream));
      return;
    }
  }
  HIPCHK(launch_gemm(g, prec, c->stream));
}

// C(MxN, ldc) = alpha op(A) op(B) + beta C, column-major
void gemm2(imps_ctx* c, char ta, char tb, int M, int N, int K, double alpha, const double* A,
           long lda, const double* B, long ldb, double beta, double* Cm, long ldc, int prec = PREC_FP64) {
  GemmDesc g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.Kc = K;
  g.A = A;
  g.B = B;
  g.C = Cm;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  if (ta == 'N') {
    g.sAm = 1;
    g.sAk = lda;
  } else {
    g.sAm = lda;
    g.sAk = 1;
  }
  if (tb == 'N') {
    g.sBk = 1;
    g.sBn = ldb;
  } else {
    g.sBk = ldb;
    g.sBn = 1;
  }
  gemm(c, g, prec);
}

// ------------------------------------------------------------------ context checks

void need_mpo(imps_ctx* c) { require(c->N > 0 && c->w > 0, IMPS_E_STATE, "MPO not set"); }
void need_state(imps_ctx* c) { require(c->D > 0 && !c->AL.empty(), IMPS_E_STATE, "state not set"); }
// entry points whose host arrays or arithmetic are real-only refuse a complex state
void require_real_state(imps_ctx* c, const char* what) {
  require(!c->cplx, IMPS_E_UNSUPPORTED,
          std::string(what) + ": not available for a complex state (real host arrays or real arithmetic only)");
}
// on a complex state the real kernels return the real part of an expectation value, which is the
// whole value for a symmetric (Hermitian) real operator; other operators are refused there
void require_symmetric_if_complex(imps_ctx* c, const double* O, const char* what) {
  if (!c->cplx) return;
  const int d = c->d;
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < j; ++i)
      require(O[i + (size_t)d * j] == O[j + (size_t)d * i], IMPS_E_UNSUPPORTED,
              std::string(what) + ": a complex state takes symmetric (Hermitian) real operators only");
}
void need_envs(imps_ctx* c) { require(c->envs_valid, IMPS_E_STATE, "environments not computed"); }
void need_match(imps_ctx* c) {
  need_mpo(c);
  need_state(c);
  require((int)c->AL.size() == c->N, IMPS_E_ARG, "unit cell of the state and of the MPO differ");
}
int wrap(imps_ctx* c, int site) {
  const int N = (int)c->AL.size() > 0 ? (int)c->AL.size() : c->N;
  return ((site % N) + N) % N;
}

// ------------------------------------------------------------------ operators

// Split-BF16 planes path of H_AC: environments with identity structural levels, D % 128 == 0,
// w <= 8, d <= 4.  Other shapes use the split-on-load GEMM of launch_gemm.
bool hac_planes_ok(imps_ctx* c, int site) {
  return c->D % 128 == 0 && c->w >= 2 && c->w <= 8 && c->d <= 4 && c->gl_id[site] && c->gr_id[site];
}

uint16_t* ws_bf16(imps_ctx* c, const std::string& name, size_t n) {
  return (uint16_t*)ws(c, name, (n + 7) / 4);  // 16-byte multiple
}

// GL levels as [m = b + D i][k = a'] rows and GR as [n = b'][k = bb + D j] rows, both planes
void hac_planes_prepare(imps_ctx* c, int site) {
  const int D = c->D, w = c->w;
  const size_t ne = (size_t)D * w * D;
  HIPCHK(launch_split_planes(D * w, D, c->GL[site], (long)D * w, true, ws_bf16(c, "hp_GLh", ne),
                             ws_bf16(c, "hp_GLl", ne), D, c->stream));
  HIPCHK(launch_split_planes(D, D * w, c->GR[site], (long)D * w, false, ws_bf16(c, "hp_GRh", ne),
                             ws_bf16(c, "hp_GRl", ne), (long)D * w, c->stream));
}

// H_AC with identity structural levels on planes: split x, GEMM 1 (levels 1..w-1 of GL, FP32
// out), mix (level 0 from x, level w-1 of GR straight into y, other levels as planes), GEMM 2
// (levels 0..w-2 of GR) accumulating onto y.
void hac_apply_planes(imps_ctx* c, int site, const double* x, double* y, hipEvent_t* ev) {
  const int D = c->D, w = c->w, d = c->d;
  if (c->planes_site != site) hac_planes_prepare(c, site);
  const size_t ne = (size_t)D * w * D, nx = (size_t)D * d * D;
  const uint16_t* GLh = ws_bf16(c, "hp_GLh", ne);
  const uint16_t* GLl = ws_bf16(c, "hp_GLl", ne);
  const uint16_t* GRh = ws_bf16(c, "hp_GRh", ne);
  const uint16_t* GRl = ws_bf16(c, "hp_GRl", ne);
  uint16_t* xh = ws_bf16(c, "hp_xh", nx);
  uint16_t* xl = ws_bf16(c, "hp_xl", nx);
  float* T1 = (float*)ws(c, "hp_T1", (nx * w + 1) / 2);
  const size_t nt = nx * (w - 1);
  uint16_t* Th = ws_bf16(c, "hp_T2h", nt);
  uint16_t* Tl = ws_bf16(c, "hp_T2l", nt);
  if (ev) HIPCHK(hipEventRecord(ev[0], c->stream));
  HIPCHK(launch_split_planes(d * D, D, x, D, false, xh, xl, D, c->stream));
  const size_t lvl = (size_t)D * D;  // one level of GL planes
  HIPCHK(launch_gemm_planes(D * (w - 1), d * D, D, GLh + lvl, GLl + lvl, D, xh, xl, D, T1 + D, (long)D * w, false,
                            1.0, 0.0, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
  HIPCHK(launch_wmix_planes(D, w, d, c->Wm[site], T1, x, Th, Tl, y, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
  const long KT = (long)D * (w - 1);
  HIPCHK(launch_gemm_planes(D * d, D, (int)KT, Th, Tl, KT, GRh, GRl, (long)D * w, y, (long)D * d, true, 1.0, 1.0,
                            c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[3], c->stream));
  c->hac_applies++;
}

// H_AC (effective.jl:90-102) as GEMM (GL x) -> W mix -> GEMM (. GR).  When GL's level 0 and
// GR's level w-1 are the identity (environments built by DMRGCache), their products are the
// operands themselves: GEMM 1 skips level 0 (the mix reads x instead), the mix writes the
// level-(w-1) channel straight into y and GEMM 2 accumulates the other levels onto it, so
// both GEMMs shrink by one level, 4 (w-1) d D^3 flop instead of 4 w d D^3, same result.
// `ev` (optional, 4 events) brackets the three kernels for timing.
// Buffers of the GEMM path of H_AC: T1 and T2 (D w d D doubles each) and, when `fixed`, the
// split-K workspace W (gemm_workspace doubles of the larger GEMM, may be null when neither takes
// one) with GEMMs that allocate nothing (gemm_prealloc); otherwise the GEMMs are gemm()'s.
struct HacBufs {
  double* T1;
  double* T2;
  double* W;
  bool fixed;
};

void hac_apply_bufs(imps_ctx* c, int site, const double* x, double* y, int prec, const HacBufs& b,
                    hipEvent_t* ev);

// For a complex state (the real embedding, D even) the outputs of the effective operators and of
// the transfer pushes are projected back onto the embedded form: the operators preserve it
// exactly, and the projection keeps rounding from seeding the conjugate sector, whose transfer
// spectrum reaches the same modulus when the state is close to its own conjugate (a real state
// after a phase), which would otherwise stall the Krylov solvers.
void project_embedded(imps_ctx* c, int m, double* p) {
  if (c->cplx) HIPCHK(launch_cplx_extract(c->D / 2, m, c->D / 2, p, 1.0, nullptr, true, c->stream));
}

void hac_apply(imps_ctx* c, int site, const double* x, double* y, int prec, hipEvent_t* ev = nullptr) {
  const int D = c->D, w = c->w, d = c->d;
  site = wrap(c, site);
  if (prec == PREC_BF16X3 && hac_planes_ok(c, site)) {
    hac_apply_planes(c, site, x, y, ev);
  } else {
    HacBufs b{ws(c, "hac_T1", (size_t)D * w * d * D), ws(c, "hac_T2", (size_t)D * w * d * D), nullptr, false};
    hac_apply_bufs(c, site, x, y, prec, b, ev);
  }
  project_embedded(c, d, y);
}

// the GEMM -> W mix -> GEMM chain of hac_apply on the buffers b (site already wrapped)
void hac_apply_bufs(imps_ctx* c, int site, const double* x, double* y, int prec, const HacBufs& b,
                    hipEvent_t* ev) {
  const int D = c->D, w = c->w, d = c->d;
  const double* GL = c->GL[site];
  const double* GR = c->GR[site];
  double* T1 = b.T1;
  double* T2 = b.T2;
  const bool skip = w >= 2 && c->gl_id[site] && c->gr_id[site];
  const int lv = skip ? 1 : 0;  // levels handled as copies at each end
  GemmDesc g1;
  g1.M = D * (w - lv);
  g1.N = d * D;
  g1.K = D;
  g1.Kc = D;
  g1.A = GL + (size_t)D * lv;
  g1.sAm = 1;
  g1.sAk = (long)D * w;
  g1.B = x;
  g1.sBk = 1;
  g1.sBn = D;
  g1.C = T1 + (size_t)D * lv;
  g1.ldc = (long)D * w;
  if (ev) HIPCHK(hipEventRecord(ev[0], c->stream));
  if (b.fixed)
    gemm_prealloc(c, g1, prec, b.W);
  else
    gemm(c, g1, prec);
  if (ev) HIPCHK(hipEventRecord(ev[1], c->stream));
  HIPCHK(launch_wmix(D, w, d, c->Wm[site], T1, T2, skip ? x : nullptr, skip ? y : nullptr, c->stream));
  if (ev) HIPCHK(hipEventRecord(ev[2], c->stream));
  GemmDesc g2;
  g2.M = D * d;
  g2.N = D;
  g2.K = D * (w - lv);
  g2.Kc = g2.K;
  g2.A = T2;
  g2.sAm = 1;
  g2.sAk = (long)D * d;
  g2.B = GR;
  g2.sBk = 1;
  g2.sBn = (long)D * w;
  g2.C = y;
  g2.ldc = (long)D * d;
  g2.beta = skip ? 1.0 : 0.0;
  if (b.fixed)
    gemm_prealloc(c, g2, prec, b.W);
  else
    gemm(c, g2, prec);
  if (ev) HIPCHK(hipEventRecord(ev[3], c->stream));
  c->hac_applies++;
}

// Device address of this GPU's copy of `scratch`, taken once per device with that device current
// (imps_create, or the first no-allocation call on a device) and kept; -- the symbol's name is a
// host-side stand-in and is never passed to a kernel.
std::mutex g_scratch_mu;
double* g_scratch_ptr[64] = {};

double* scratch_ptr(int device) {
  require(device >= 0 && device < 64, IMPS_E_ARG, "scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, IMPS_E_HIP, "scratch: hipGetSymbolAddress returned no address");
    g_scratch_ptr[device] = (double*)p;
  }
  return g_scratch_ptr[device];
}

// Split-BF16 planes path of H_C (identity structural levels, D % 128 == 0): with T1_k = GL_k x,
// y = T1_{w-1} + sum_{k < w-1} T1_k GR_k.  GEMM 1 makes T1_k for k >= 1 in FP32; T1_0 = x and the
// FP32 levels 1..w-2 are split (transposed) into the K-contiguous planes of the second GEMM,
// which runs split-K into FP32 parts; the reduction adds the parts and T1_{w-1} into y.
bool hc_planes_ok(imps_ctx* c, int site) {
  return c->D % 128 == 0 && c->w >= 2 && c->gl_id[wrap(c, site + 1)] && c->gr_id[site];
}

void hc_planes_prepare(imps_ctx* c, int site) {
  const int D = c->D, w = c->w;
  const size_t ne = (size_t)D * w * D;
  HIPCHK(launch_split_planes(D * w, D, c->GL[wrap(c, site + 1)], (long)D * w, true, ws_bf16(c, "hcp_GLh", ne),
                             ws_bf16(c, "hcp_GLl", ne), D, c->stream));
  HIPCHK(launch_split_planes(D, D * w, c->GR[site], (long)D * w, false, ws_bf16(c, "hcp_GRh", ne),
                             ws_bf16(c, "hcp_GRl", ne), (long)D * w, c->stream));
}

void hc_apply_planes(imps_ctx* c, int site, const double* x, double* y) {
  const int D = c->D, w = c->w;
  if (c->hc_planes_site != site) hc_planes_prepare(c, site);
  const size_t ne = (size_t)D * w * D, nd = (size_t)D * D;
  const long KT = (long)D * (w - 1);
  const uint16_t* GLh = ws_bf16(c, "hcp_GLh", ne);
  const uint16_t* GLl = ws_bf16(c, "hcp_GLl", ne);
  const uint16_t* GRh = ws_bf16(c, "hcp_GRh", ne);
  const uint16_t* GRl = ws_bf16(c, "hcp_GRl", ne);
  uint16_t* xh = ws_bf16(c, "hcp_xh", nd);
  uint16_t* xl = ws_bf16(c, "hcp_xl", nd);
  uint16_t* Ah = ws_bf16(c, "hcp_Ah", nd * (w - 1));
  uint16_t* Al = ws_bf16(c, "hcp_Al", nd * (w - 1));
  float* T1 = (float*)ws(c, "hcp_T1", (ne + 1) / 2);
  // split-K depth of GEMM 2: about 512 workgroups of 128 x 64
  const long tiles = (long)(D / 128) * (D / 64);
  int S = 1;
  while (S < 8 && tiles * S * 2 <= 512 && KT % (64L * S * 2) == 0) S *= 2;
  float* P = (float*)ws(c, "hcp_P", (nd * S + 1) / 2);
  // B of GEMM 1: x as [b][a'] rows; level 0 of GEMM 2's A: x transposed into [a][b]
  HIPCHK(launch_split_planes(D, D, x, D, false, xh, xl, D, c->stream));
  HIPCHK(launch_split_planes(D, D, x, D, true, Ah, Al, KT, c->stream));
  HIPCHK(launch_gemm_planes(D * (w - 1), D, D, GLh + nd, GLl + nd, D, xh, xl, D, T1 + D, (long)D * w, false, 1.0, 0.0,
                            c->stream));
  if (w > 2)
    HIPCHK(launch_split_planes_levels(D, D, w - 2, T1 + D, (long)D * w, D, Ah + D, Al + D, KT, D, c->stream));
  HIPCHK(launch_gemm_planes_splitk(D, D, (int)KT, S, Ah, Al, KT, GRh, GRl, (long)D * w, P, c->stream));
  HIPCHK(launch_planes_reduce(D, S, P, (long)nd, T1 + (size_t)D * (w - 1), (long)D * w, y, c->stream));
  c->hc_applies++;
}

// H_C (effective.jl:72-76): y = sum_k GL_k x GR_k with GL = leftenv(site+1), GR = rightenv(site).
// With identity structural levels, GL_0 x = x (copied into level 0 of T1) and the level-(w-1)
// term is T1_{w-1} itself (copied into y, onto which the GEMM over the other levels adds).
void hc_apply_impl(imps_ctx* c, int site, const double* x, double* y, int prec);

void hc_apply(imps_ctx* c, int site, const double* x, double* y, int prec) {
  hc_apply_impl(c, site, x, y, prec);
  project_embedded(c, 1, y);
}

void hc_apply_impl(imps_ctx* c, int site, const double* x, double* y, int prec) {
  const int D = c->D, w = c->w;
  site = wrap(c, site);
  const int sl = wrap(c, site + 1);
  const double* GL = c->GL[sl];
  const double* GR = c->GR[site];
  if (prec == PREC_BF16X3 && hc_planes_ok(c, site)) {
    hc_apply_planes(c, site, x, y);
    return;
  }
  double* T1 = ws(c, "hc_T1", (size_t)D * w * D);
  const bool skip = w >= 2 && c->gl_id[sl] && c->gr_id[site];
  const int lv = skip ? 1 : 0;
  gemm2(c, 'N', 'N', D * (w - lv), D, D, 1.0, GL + (size_t)D * lv, (long)D * w, x, D, 0.0, T1 + (size_t)D * lv,
        (long)D * w, prec);
  if (skip) {
    HIPCHK(launch_copy2d(D, D, x, D, T1, (long)D * w, c->stream));
    HIPCHK(launch_copy2d(D, D, T1 + (size_t)D * (w - 1), (long)D * w, y, D, c->stream));
  }
  const int nk = w - lv;  // levels 0 .. nk-1 go through the second GEMM
  if (D % 32 == 0) {
    GemmDesc g;
    g.M = D;
    g.N = D;
    g.K = D * nk;
    g.Kc = D;
    g.A = T1;
    g.sAm = 1;
    g.sAk = (long)D * w;
    g.sAc = D;
    g.B = GR;
    g.sBk = 1;
    g.sBc = D;
    g.sBn = (long)D * w;
    g.C = y;
    g.ldc = D;
    g.beta = skip ? 1.0 : 0.0;
    gemm(c, g, prec);
  } else {
    double* T1p = ws(c, "hc_T1p", (size_t)D * w * D);
    for (int k = 0; k < nk; ++k)
      HIPCHK(launch_copy2d(D, D, T1 + (size_t)D * k, (long)D * w, T1p + (size_t)D * D * k, D, c->stream));
    gemm2(c, 'N', 'N', D, D, D * nk, 1.0, T1p, D, GR, (long)D * w, skip ? 1.0 : 0.0, y, D, prec);
  }
  c->hc_applies++;
}

// generalized identity-channel pushes; X (D x D, ld ldx) -> out (D x D, ld ldo)
// left:  out[a',b'] = sum below[a,s,a'] X[a,b] above[b,s,b']
void push_left2(imps_ctx* c, const double* X, long ldx, const double* above, const double* below,
                double* out, long ldo, int prec = PREC_FP64) {
  const int D = c->D, d = c->d;
  double* Y = ws(c, "pl_Y", (size_t)D * d * D);
  gemm2(c, 'N', 'N', D, d * D, D, 1.0, X, ldx, above, D, 0.0, Y, D, prec);
  gemm2(c, 'T', 'N', D, D, d * D, 1.0, below, (long)d * D, Y, (long)d * D, 0.0, out, ldo, prec);
  if (ldo == D) project_embedded(c, 1, out);
}

// right: out[a',b'] = sum above[a',s,a] X[a,b] below[b',s,b]
void push_right2(imps_ctx* c, const double* X, long ldx, const double* above, const double* below,
                 double* out, long ldo, int prec = PREC_FP64) {
  const int D = c->D, d = c->d;
  double* Y = ws(c, "pr_Y", (size_t)D * d * D);
  gemm2(c, 'N', 'N', d * D, D, D, 1.0, above, (long)d * D, X, ldx, 0.0, Y, (long)d * D, prec);
  gemm2(c, 'N', 'T', D, D, d * D, 1.0, Y, D, below, D, 0.0, out, ldo, prec);
  if (ldo == D) project_embedded(c, 1, out);
}

// block W[l,:,i,:] of a (w, d, w, d) tensor as a (d x d) column-major matrix M[u + d*s]
std::vector<double> wblock_of(const std::vector<double>& W, int w, int d, int l, int i) {
  std::vector<double> M((size_t)d * d);
  for (int s = 0; s < d; ++s)
    for (int u = 0; u < d; ++u) M[u + (size_t)d * s] = W[l + (size_t)w * (u + (size_t)d * (i + (size_t)w * s))];
  return M;
}

// MPO block W[l,:,i,:] of site
std::vector<double> wblock(imps_ctx* c, int site, int l, int i) { return wblock_of(c->Wh[site], c->w, c->d, l, i); }

bool is_zero(const std::vector<double>& M) {
  for (double v : M)
    if (v != 0.0) return false;
  return true;
}

bool is_eye(const std::vector<double>& M, int d, double atol = 1e-14) {
  for (int s = 0; s < d; ++s)
    for (int u = 0; u < d; ++u)
      if (std::fabs(M[u + (size_t)d * s] - (u == s ? 1.0 : 0.0)) > atol) return false;
  return true;
}

// Left slice sum (envs.jl:104-107, 121-136): out = sum_terms push_slice_left(L_t, W_t, A)
struct SliceTerm {
  const double* L;  // nullptr: identity
  long ld;
  std::vector<double> Wb;
};

// With distinct layers (the defect environments of the fixed-point response):
// out = sum_terms bra^T mix(W_t, L_t ket) + beta out.
void slice_sum_left_bk(imps_ctx* c, const std::vector<SliceTerm>& terms, const double* bra, const double* ket,
                       double* out, long ldo, double beta) {
  const int D = c->D, d = c->d;
  if (terms.empty()) {
    if (beta == 0.0) HIPCHK(launch_fill2d(D, D, 0.0, out, ldo, c->stream));
    return;
  }
  double* Z = ws(c, "sl_Z", (size_t)D * d * D);
  double* Y = ws(c, "sl_Y", (size_t)D * d * D);
  bool first = true;
  for (const SliceTerm& t : terms) {
    const double* src = ket;
    if (t.L) {
      gemm2(c, 'N', 'N', D, d * D, D, 1.0, t.L, t.ld, ket, D, 0.0, Y, D);
      src = Y;
    }
    HIPCHK(launch_mix_mid(D, d, D, t.Wb.data(), src, Z, first ? 0.0 : 1.0, c->stream));
    first = false;
  }
  gemm2(c, 'T', 'N', D, D, d * D, 1.0, bra, (long)d * D, Z, (long)d * D, beta, out, ldo);
}

void slice_sum_left(imps_ctx* c, const std::vector<SliceTerm>& terms, const double* A, double* out,
                    long ldo) {
  slice_sum_left_bk(c, terms, A, A, out, ldo, 0.0);
}

// Right slice sum (envs.jl:110-114, 162-177)
// out = sum_terms mix(W_t, ket R_t) bra^T + beta out (distinct layers)
void slice_sum_right_bk(imps_ctx* c, const std::vector<SliceTerm>& terms, const double* bra, const double* ket,
                        double* out, long ldo, double beta) {
  const int D = c->D, d = c->d;
  if (terms.empty()) {
    if (beta == 0.0) HIPCHK(launch_fill2d(D, D, 0.0, out, ldo, c->stream));
    return;
  }
  double* Z = ws(c, "sr_Z", (size_t)D * d * D);
  double* Y = ws(c, "sr_Y", (size_t)D * d * D);
  bool first = true;
  for (const SliceTerm& t : terms) {
    const double* src = ket;
    if (t.L) {
      gemm2(c, 'N', 'N', d * D, D, D, 1.0, ket, (long)d * D, t.L, t.ld, 0.0, Y, (long)d * D);
      src = Y;
    }
    HIPCHK(launch_mix_mid(D, d, D, t.Wb.data(), src, Z, first ? 0.0 : 1.0, c->stream));
    first = false;
  }
  gemm2(c, 'N', 'T', D, D, d * D, 1.0, Z, D, bra, D, beta, out, ldo);
}

void slice_sum_right(imps_ctx* c, const std::vector<SliceTerm>& terms, const double* A, double* out,
                     long ldo) {
  slice_sum_right_bk(c, terms, A, A, out, ldo, 0.0);
}

// MPO-channel pushes (transfermatrix.jl:11-13, 31-35) as sums of slice pushes over the nonzero
// d x d blocks of W: out[:, v, :] = sum_k push_slice_left(L[:, k, :], W[k, :, v, :], A) and
// out[:, v, :] = sum_k push_slice_right(R[:, k, :], W[v, :, k, :], A).  out must not alias L / R.
// W: a (w, d, w, d) tensor on the host; L, R, out: (D, w, D) on the device.
void push_w_left(imps_ctx* c, const std::vector<double>& W, int w, const double* L, const double* A, double* out) {
  const int D = c->D;
  for (int v = 0; v < w; ++v) {
    std::vector<SliceTerm> terms;
    for (int k = 0; k < w; ++k) {
      std::vector<double> M = wblock_of(W, w, c->d, k, v);
      if (!is_zero(M)) terms.push_back(SliceTerm{L + (size_t)D * k, (long)D * w, M});
    }
    slice_sum_left(c, terms, A, out + (size_t)D * v, (long)D * w);
  }
}

void push_w_right(imps_ctx* c, const std::vector<double>& W, int w, const double* R, const double* A, double* out) {
  const int D = c->D;
  for (int v = 0; v < w; ++v) {
    std::vector<SliceTerm> terms;
    for (int k = 0; k < w; ++k) {
      std::vector<double> M = wblock_of(W, w, c->d, v, k);
      if (!is_zero(M)) terms.push_back(SliceTerm{R + (size_t)D * k, (long)D * w, M});
    }
    slice_sum_right(c, terms, A, out + (size_t)D * v, (long)D * w);
  }
}

void push_mpo_left(imps_ctx* c, int site, const double* L, const double* A, double* out) {
  push_w_left(c, c->Wh[site], c->w, L, A, out);
}

void push_mpo_right(imps_ctx* c, int site, const double* R, const double* A, double* out) {
  push_w_right(c, c->Wh[site], c->w, R, A, out);
}

// ------------------------------------------------------------------ Krylov solvers

typedef std::function<void(const double*, double*)> Op;

struct KInfo {
  int converged = 0, numiter = 0, numops = 0;
  double normres = 0.0;
};

// Lanczos eigsolve, lowest eigenpair (KrylovKit.eigsolve(..., :SR, Lanczos) via
// fixedpoint, utility.jl:180-186): thick restart keep = div(3 kdim + 2 conv, 5),
// eager convergence test |beta u_K| <= tol, CGS2 full reorthogonalisation.
double lanczos(imps_ctx* c, const Op& op, size_t n, const double* x0, double* xout, double tol,
               int kdim, int maxiter, bool eager, KInfo* info) {
  const int kmax = (int)std::min<size_t>(kdim, n);
  require(kmax >= 1 && kmax < kMaxVec, IMPS_E_ARG, "lanczos: krylovdim out of range");
  double* V = ws(c, "lz_V", (size_t)(kmax + 1) * n);
  const int ldt = kmax + 1;
  std::vector<double> T((size_t)ldt * ldt, 0.0);
  auto Vc = [&](int k) { return V + (size_t)k * n; };
  const double nx = nrm2(c, n, x0);
  require(nx > 0 && std::isfinite(nx), IMPS_E_ARG, "lanczos: zero or non-finite start vector");
  axpby(c, n, 1.0 / nx, x0, 0.0, Vc(0));
  int numops = 0, numiter = 1, K = 0;
  double beta = 0.0;
  std::vector<double> h(kmax + 1);
  // Expansion k: V_{k+1} = normalised (op V_k) orthogonalised against V_0..V_k.  The next
  // application op(V_{k+1}) is enqueued right behind the Gram-Schmidt kernels, before the host
  // waits for the coefficients and runs the convergence test, so the device does not idle
  // on the host; when the test converges (or a restart follows) that application is simply
  // not used.
  bool pending = false;  // op(V_K) already enqueued
  auto expand = [&](int k) {
    if (!pending) op(Vc(k), Vc(k + 1));
    pending = false;
    numops++;
    cgs2_enqueue(c, n, k + 1, V, Vc(k + 1));
    if (k + 1 < kmax) {
      op(Vc(k + 1), Vc(k + 2));
      pending = true;
    }
    cgs2_wait(c, k + 1, h.data(), &beta);
    for (int i = 0; i <= k; ++i) {
      T[i + (size_t)ldt * k] = h[i];
      T[k + (size_t)ldt * i] = h[i];
    }
  };
  expand(0);
  K = 1;
  std::vector<double> theta, U;
  int converged = 0;
  while (true) {
    converged = 0;
    const bool check = (K == kmax || beta <= tol || eager);
    if (check) {
      std::vector<double> A((size_t)K * K);
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i)
          A[i + (size_t)K * j] = 0.5 * (T[i + (size_t)ldt * j] + T[j + (size_t)ldt * i]);
      theta.assign(K, 0.0);
      U.assign((size_t)K * K, 0.0);
      sym_eig(K, A.data(), theta.data(), U.data());
      if (std::fabs(beta * U[(K - 1) + (size_t)K * 0]) <= tol) {
        converged = 1;
        break;
      }
    }
    if (K < kmax) {
      expand(K);
      K += 1;
    } else {
      if (numiter == maxiter) break;
      const int keep = (3 * kmax + 2 * converged) / 5;
      double* tmp = ws(c, "lz_tmp", (size_t)keep * n);
      double* dU = ws(c, "lz_U", (size_t)K * keep);
      h2d(c, dU, U.data(), (size_t)K * keep);
      gemm2(c, 'N', 'N', (int)n, keep, K, 1.0, V, (long)n, dU, K, 0.0, tmp, (long)n);
      d2d(c, V, tmp, (size_t)keep * n);
      d2d(c, Vc(keep), Vc(K), n);
      std::vector<double> f(keep);
      for (int j = 0; j < keep; ++j) f[j] = beta * U[(K - 1) + (size_t)K * j];
      std::fill(T.begin(), T.end(), 0.0);
      for (int j = 0; j < keep; ++j) {
        T[j + (size_t)ldt * j] = theta[j];
        T[keep + (size_t)ldt * j] = f[j];
        T[j + (size_t)ldt * keep] = f[j];
      }
      K = keep;
      expand(K);
      K += 1;
      numiter++;
    }
  }
  std::vector<double> u0(K);
  for (int i = 0; i < K; ++i) u0[i] = U[i];
  multiaxpy(c, n, K, V, n, u0.data(), 0.0, xout);
  if (dot(c, n, xout, x0) < 0) scal(c, n, -1.0, xout);
  if (info) {
    info->converged = converged;
    info->numiter = numiter;
    info->numops = numops;
    info->normres = std::fabs(beta * U[K - 1]);
  }
  return theta[0];
}

// x /= ||x|| on the device (the norm reduced in a fixed order, no host round trip)
void normalize_dev(imps_ctx* c, size_t n, double* x) {
  double* sq = ws(c, "nd_sq", 1);
  HIPCHK(launch_multidot(n, 1, x, n, x, c->dpart, c->stream));
  HIPCHK(launch_partials_reduce(1, dot_blocks(n), c->dpart, sq, c->stream));
  HIPCHK(launch_scale_inv_sqrt(n, x, sq, c->stream));
}

// Lowest eigenpair refined from a good start vector with a reduced-precision operator (the FP64
// refinement of the mixed-precision VUMPS eigensolves).  Outer loop in FP64: theta = <x, H x>,
// r = H x - theta x, stop when |r| <= tol (the residual Lanczos' eager test bounds); else solve the
// Jacobi-Davidson correction equation P (H_low - theta) P t = -r, t _|_ x (P = I - x x^T) with
// MINRES to relative accuracy max(tol / (2 |r|), 1e-6), and x <- (x + t) / |x + t|.  The reduced
// precision only enters the correction, whose relative error it bounds, so the FP64 residual keeps
// falling below the reduced operator's own rounding level.  MINRES runs entirely on the device
// (launch_minres_*: no Gram-Schmidt against a basis, no host round trip per step); the host reads
// its residual estimate every kCheck steps.  Returns false (xout then holds the last iterate) when
// tol is not reached within max_outer corrections or a correction does not converge in max_inner
// steps; the caller falls back to FP64 Lanczos.  *lambda receives theta of the last FP64 residual.
bool jd_refine(imps_ctx* c, const Op& op64, const Op& oplow, size_t n, const double* x0, double* xout, double tol,
               int max_outer, int max_inner, double* lambda, KInfo* info) {
  constexpr int kCheck = 8;
  double* y = ws(c, "jd_y", n);
  double* r = ws(c, "jd_r", n);
  double* t = ws(c, "jd_t", n);
  double* vb[3] = {ws(c, "jd_v0", n), ws(c, "jd_v1", n), ws(c, "jd_v2", n)};
  double* wb[2] = {ws(c, "jd_w0", n), ws(c, "jd_w1", n)};
  double* S = ws(c, "jd_S", MR_NS);
  const int nb = dot_blocks(n);
  double* part = ws(c, "jd_part", 2 * (size_t)nb);
  const double nx = nrm2(c, n, x0);
  require(nx > 0 && std::isfinite(nx), IMPS_E_ARG, "refine: zero or non-finite start vector");
  axpby(c, n, 1.0 / nx, x0, 0.0, xout);
  int nops = 0, outer = 0;
  double res = 0.0, theta = 0.0;
  bool ok = false;
  for (outer = 0; outer <= max_outer; ++outer) {
    op64(xout, y);
    nops++;
    theta = dot(c, n, xout, y);
    axpby(c, n, 1.0, y, 0.0, r);
    axpby(c, n, -theta, xout, 1.0, r);
    res = nrm2(c, n, r);
    if (res <= tol) {
      ok = true;
      break;
    }
    if (outer == max_outer || !std::isfinite(res)) break;
    // MINRES on P (H_low - theta) P t = -r from t = 0: v_1 = -r / |r|
    const double rtol = std::max(0.5 * tol / res, 1e-6);
    double* vold = vb[0];
    double* v = vb[1];
    double* z = vb[2];
    double* wold = wb[0];
    double* w = wb[1];
    fill(c, n, 0.0, vold);
    fill(c, n, 0.0, wold);
    fill(c, n, 0.0, w);
    fill(c, n, 0.0, t);
    axpby(c, n, -1.0 / res, r, 0.0, v);
    HIPCHK(launch_minres_init(S, res, c->stream));
    bool inner_ok = false;
    for (int k = 1; k <= max_inner; ++k) {
      oplow(v, y);
      nops++;
      HIPCHK(launch_dot2(n, v, xout, y, part, c->stream));
      HIPCHK(launch_partials_reduce(2, nb, part, S + MR_DELTA, c->stream));
      HIPCHK(launch_minres_z(n, y, xout, v, vold, S, z, part, c->stream));
      HIPCHK(launch_partials_reduce(1, nb, part, S + MR_ZZ, c->stream));
      HIPCHK(launch_minres_scalar(S, theta, c->stream));
      HIPCHK(launch_minres_update(n, v, wold, w, t, z, S, c->stream));
      // rotate: w_new sits in wold; the normalised z is the next v
      std::swap(wold, w);
      double* vn = z;
      z = vold;
      vold = v;
      v = vn;
      if (k % kCheck == 0 || k == max_inner) {
        const double* hs = d2h_small(c, S, sizeof(double) * MR_NS);
        if (std::fabs(hs[MR_ETA]) <= rtol * res || hs[MR_DONE] != 0.0) {
          inner_ok = true;
          break;
        }
        if (!std::isfinite(hs[MR_ETA])) break;
      }
    }
    if (!inner_ok) break;
    axpby(c, n, 1.0, t, 1.0, xout);
    normalize_dev(c, n, xout);
  }
  if (lambda) *lambda = theta;
  if (info) {
    info->converged = ok ? 1 : 0;
    info->numiter = outer + 1;
    info->numops = nops;
    info->normres = res;
  }
  return ok;
}

// exp(t op) x0 for a symmetric operator and real t: KrylovKit.exponentiate with Lanczos, the
// integrator of TDVP (tdvp.jl:14-19).  A Lanczos basis (two-pass Gram-Schmidt on the device) grows
// from the current vector; with T = U diag(theta) U^T on the host, y(tau) = U exp(tau theta) U^T e1
// is accepted for the whole remaining time when |beta_K y_K(tau)| <= tol (or the space is
// invariant); when the basis reaches krylovdim first, the largest tau = remaining / 2^j with an
// accepted estimate advances the vector to V y(tau) and the basis restarts for the rest.  The
// running norm is kept apart from the unit vector, so exp(t op) may grow or shrink the result.
KInfo lanczos_expm(imps_ctx* c, const Op& op, size_t n, const double* x0, double* xout, double t, double tol,
                   int kdim, int maxiter) {
  const int kmax = (int)std::min<size_t>(kdim, n);
  require(kmax >= 1 && kmax < kMaxVec, IMPS_E_ARG, "exponentiate: krylovdim out of range");
  double* V = ws(c, "lx_V", (size_t)(kmax + 1) * n);
  double* tmp = ws(c, "lx_tmp", n);
  auto Vc = [&](int k) { return V + (size_t)k * n; };
  const int ldt = kmax + 1;
  std::vector<double> T((size_t)ldt * ldt), h(kmax + 1);
  double scale = nrm2(c, n, x0);
  require(scale > 0 && std::isfinite(scale), IMPS_E_ARG, "exponentiate: zero or non-finite start vector");
  axpby(c, n, 1.0 / scale, x0, 0.0, Vc(0));
  double remaining = t;
  KInfo info;
  bool converged = true;
  int steps = 0;
  while (true) {
    std::fill(T.begin(), T.end(), 0.0);
    int K = 0;
    double beta = 0.0, tau = remaining;
    std::vector<double> y;
    bool accepted = false;
    while (true) {
      op(Vc(K), Vc(K + 1));
      info.numops++;
      cgs2_enqueue(c, n, K + 1, V, Vc(K + 1));
      cgs2_wait(c, K + 1, h.data(), &beta);
      for (int i = 0; i <= K; ++i) {
        T[i + (size_t)ldt * K] = h[i];
        T[K + (size_t)ldt * i] = h[i];
      }
      K += 1;
      std::vector<double> A((size_t)K * K), theta(K), U((size_t)K * K);
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) A[i + (size_t)K * j] = 0.5 * (T[i + (size_t)ldt * j] + T[j + (size_t)ldt * i]);
      sym_eig(K, A.data(), theta.data(), U.data());
      auto y_of = [&](double tt) {
        std::vector<double> r(K, 0.0);
        for (int j = 0; j < K; ++j) {
          const double f = std::exp(tt * theta[j]) * U[(size_t)K * j];
          for (int i = 0; i < K; ++i) r[i] += U[i + (size_t)K * j] * f;
        }
        return r;
      };
      y = y_of(remaining);
      if (std::fabs(beta * y[K - 1]) <= tol || beta == 0.0) {
        accepted = true;
        tau = remaining;
        break;
      }
      if (K == kmax) {
        double tt = remaining;
        for (int j = 0; j < 60; ++j) {
          tt *= 0.5;
          std::vector<double> yt = y_of(tt);
          if (std::fabs(beta * yt[K - 1]) <= tol) {
            accepted = true;
            tau = tt;
            y = yt;
            break;
          }
        }
        break;
      }
    }
    steps++;
    if (!accepted) converged = false;
    multiaxpy(c, n, K, V, n, y.data(), 0.0, tmp);
    const double ny = nrm2(c, n, tmp);
    scale *= ny;
    axpby(c, n, 1.0 / ny, tmp, 0.0, Vc(0));
    remaining -= tau;
    if (remaining == 0.0 || !converged || steps >= maxiter) {
      converged = converged && remaining == 0.0;
      break;
    }
  }
  axpby(c, n, scale, Vc(0), 0.0, xout);
  info.converged = converged ? 1 : 0;
  info.numiter = steps;
  return info;
}

// exp(t op) x0 for complex t on the real embedding of a complex vector (rows in re/im pairs, see
// launch_cplx_embed): op is the embedding of a Hermitian operator, so the Lanczos basis and the
// tridiagonal T are real and exp(t T) e1 has complex coefficients y; the result is
// V Re(y) + i V Im(y), with i applied by launch_cplx_axpy_i.  Same acceptance and time-splitting
// rules as lanczos_expm (the integrate of tdvp.jl:14-19 with a complex dt).
KInfo lanczos_expm_c(imps_ctx* c, const Op& op, size_t n, const double* x0, double* xout, std::complex<double> t,
                     double tol, int kdim, int maxiter) {
  using cd = std::complex<double>;
  const int kmax = (int)std::min<size_t>(kdim, n);
  require(kmax >= 1 && kmax < kMaxVec, IMPS_E_ARG, "exponentiate: krylovdim out of range");
  require(n % 2 == 0, IMPS_E_ARG, "exponentiate: complex step on an odd-length vector");
  double* V = ws(c, "lx_V", (size_t)(kmax + 1) * n);
  double* tmp = ws(c, "lx_tmp", n);
  double* tmpi = ws(c, "lx_tmpi", n);
  auto Vc = [&](int k) { return V + (size_t)k * n; };
  const int ldt = kmax + 1;
  std::vector<double> T((size_t)ldt * ldt), h(kmax + 1);
  double scale = nrm2(c, n, x0);
  require(scale > 0 && std::isfinite(scale), IMPS_E_ARG, "exponentiate: zero or non-finite start vector");
  axpby(c, n, 1.0 / scale, x0, 0.0, Vc(0));
  cd remaining = t;
  KInfo info;
  bool converged = true;
  int steps = 0;
  while (true) {
    std::fill(T.begin(), T.end(), 0.0);
    int K = 0;
    double beta = 0.0;
    cd tau = remaining;
    std::vector<cd> y;
    bool accepted = false;
    while (true) {
      op(Vc(K), Vc(K + 1));
      info.numops++;
      cgs2_enqueue(c, n, K + 1, V, Vc(K + 1));
      cgs2_wait(c, K + 1, h.data(), &beta);
      for (int i = 0; i <= K; ++i) {
        T[i + (size_t)ldt * K] = h[i];
        T[K + (size_t)ldt * i] = h[i];
      }
      K += 1;
      std::vector<double> A((size_t)K * K), theta(K), U((size_t)K * K);
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) A[i + (size_t)K * j] = 0.5 * (T[i + (size_t)ldt * j] + T[j + (size_t)ldt * i]);
      sym_eig(K, A.data(), theta.data(), U.data());
      auto y_of = [&](cd tt) {
        std::vector<cd> r(K, cd(0.0, 0.0));
        for (int j = 0; j < K; ++j) {
          const cd f = std::exp(tt * theta[j]) * U[(size_t)K * j];
          for (int i = 0; i < K; ++i) r[i] += U[i + (size_t)K * j] * f;
        }
        return r;
      };
      y = y_of(remaining);
      if (std::abs(beta * y[K - 1]) <= tol || beta == 0.0) {
        accepted = true;
        tau = remaining;
        break;
      }
      if (K == kmax) {
        cd tt = remaining;
        for (int j = 0; j < 60; ++j) {
          tt *= 0.5;
          std::vector<cd> yt = y_of(tt);
          if (std::abs(beta * yt[K - 1]) <= tol) {
            accepted = true;
            tau = tt;
            y = yt;
            break;
          }
        }
        break;
      }
    }
    steps++;
    if (!accepted) converged = false;
    std::vector<double> yr(K), yi(K);
    for (int i = 0; i < K; ++i) {
      yr[i] = y[i].real();
      yi[i] = y[i].imag();
    }
    multiaxpy(c, n, K, V, n, yr.data(), 0.0, tmp);
    multiaxpy(c, n, K, V, n, yi.data(), 0.0, tmpi);
    HIPCHK(launch_cplx_axpy_i(n, 1.0, tmpi, tmp, c->stream));
    const double ny = nrm2(c, n, tmp);
    scale *= ny;
    axpby(c, n, 1.0 / ny, tmp, 0.0, Vc(0));
    remaining -= tau;
    if (remaining == cd(0.0, 0.0) || !converged || steps >= maxiter) {
      converged = converged && remaining == cd(0.0, 0.0);
      break;
    }
  }
  axpby(c, n, scale, Vc(0), 0.0, xout);
  info.converged = converged ? 1 : 0;
  info.numiter = steps;
  return info;
}

// Dominant eigenpair (which = :LM) of a real operator by Arnoldi (KrylovKit.schursolve
// via gauge_fixedpoint, utility.jl:204-211; eigsolve(..., :LM) via dot, canonicalmps.jl:
// 262-269); restarts from the dominant Ritz vector.  The residual is |beta x_K| with x the unit
// (complex, for a complex pair) eigenvector of the Hessenberg matrix; for a complex dominant pair
// xout is the real part of the phase-fixed Ritz vector (a vector of the pair's real invariant
// plane, from which the next restart spans it), the return value Re(lambda) and *lam_im Im(lambda).
// Arnoldi in complex arithmetic for a complex state (real embedding, pairs of rows): the basis
// holds v_k and i v_k side by side (slots 2k, 2k+1), so that Gram-Schmidt against every slot is
// complex Gram-Schmidt and the projection is the complex Hessenberg matrix with entries
// <v_j, w> + i <i v_j, w>.  On the embedding a complex eigenvalue lambda of the operator is the
// pair (lambda, conj lambda) of a real operator, which a real Krylov space cannot separate when
// Im lambda is tiny (the gauge fixed points of a state close to a real one); in complex
// arithmetic it is one simple eigenvalue.  Same restart and convergence rules as the real case.
double arnoldi_dominant_cplx(imps_ctx* c, const Op& op, size_t n, const double* x0, double* xout, double tol, int kdim,
                             int maxiter, bool eager, KInfo* info, bool* complex_dominant, double* lam_im) {
  using cd = std::complex<double>;
  // the embedded vectors of n doubles hold n / 4 complex numbers
  const int kmax = (int)std::min<size_t>(kdim, std::max<size_t>(1, n / 4));
  require(kmax >= 1 && 2 * (kmax + 1) <= kMaxVec, IMPS_E_ARG, "arnoldi: krylovdim out of range");
  double* V = ws(c, "arz_V", (size_t)2 * (kmax + 1) * n);
  auto Vc = [&](int slot) { return V + (size_t)slot * n; };
  const int ldh = kmax + 1;
  std::vector<cd> H((size_t)ldh * kmax), x(kmax);
  std::vector<double> h(2 * (kmax + 1)), coef(2 * kmax);
  int numiter = 1, numops = 0, K = 0;
  bool conv = false;
  cd lam(0.0, 0.0);
  double res = 0.0;
  const double nx = nrm2(c, n, x0);
  require(nx > 0 && std::isfinite(nx), IMPS_E_ARG, "arnoldi: zero or non-finite start vector");
  axpby(c, n, 1.0 / nx, x0, 0.0, Vc(0));
  auto with_i = [&](int k) {  // slot 2k + 1 = i * slot 2k
    fill(c, n, 0.0, Vc(2 * k + 1));
    HIPCHK(launch_cplx_axpy_i(n, 1.0, Vc(2 * k), Vc(2 * k + 1), c->stream));
  };
  while (true) {
    std::fill(H.begin(), H.end(), cd(0.0, 0.0));
    K = 0;
    with_i(0);
    while (true) {
      op(Vc(2 * K), Vc(2 * K + 2));
      numops++;
      double beta = 0.0;
      cgs2_normalize(c, n, 2 * K + 2, V, Vc(2 * K + 2), h.data(), &beta);
      for (int j = 0; j <= K; ++j) H[j + (size_t)ldh * K] = cd(h[2 * j], h[2 * j + 1]);
      H[(K + 1) + (size_t)ldh * K] = cd(beta, 0.0);
      K++;
      with_i(K);
      if (K == kmax || beta <= tol || eager) {
        double last = 0.0;
        require(hess_dominant_z(K, H.data(), ldh, &lam, x.data(), &last) == 0, IMPS_E_SOLVER,
                "arnoldi: Hessenberg eigenvalue iteration failed");
        res = std::fabs(beta) * last;
        if (res <= tol) {
          conv = true;
          break;
        }
      }
      if (K == kmax) break;
    }
    for (int k = 0; k < K; ++k) {
      coef[2 * k] = x[k].real();
      coef[2 * k + 1] = x[k].imag();
    }
    multiaxpy(c, n, 2 * K, V, n, coef.data(), 0.0, xout);
    if (conv || numiter == maxiter) break;
    const double nr = nrm2(c, n, xout);
    axpby(c, n, 1.0 / nr, xout, 0.0, Vc(0));
    numiter++;
  }
  if (info) {
    info->converged = conv ? 1 : 0;
    info->numiter = numiter;
    info->numops = numops;
    info->normres = res;
  }
  if (complex_dominant) *complex_dominant = std::fabs(lam.imag()) > 1e-12 * std::max(1.0, std::abs(lam));
  if (lam_im) *lam_im = lam.imag();
  return lam.real();
}

double arnoldi_dominant(imps_ctx* c, const Op& op, size_t n, const double* x0, double* xout,
                        double tol, int kdim, int maxiter, bool eager, KInfo* info, bool* complex_dominant = nullptr,
                        double* lam_im = nullptr) {
  const int kmax = (int)std::min<size_t>(kdim, n);
  if (c->cplx && n % 2 == 0)
    return arnoldi_dominant_cplx(c, op, n, x0, xout, tol, kdim, maxiter, eager, info, complex_dominant, lam_im);
  require(kmax >= 1 && kmax < kMaxVec, IMPS_E_ARG, "arnoldi: krylovdim out of range");
  double* V = ws(c, "ar_V", (size_t)(kmax + 1) * n);
  auto Vc = [&](int k) { return V + (size_t)k * n; };
  const int ldh = kmax + 1;
  std::vector<double> H((size_t)ldh * kmax), h(kmax + 1), u(kmax);
  int numiter = 1, numops = 0, K = 0;
  bool conv = false;
  double lam = 0.0, lim = 0.0, res = 0.0;
  const double nx = nrm2(c, n, x0);
  require(nx > 0 && std::isfinite(nx), IMPS_E_ARG, "arnoldi: zero or non-finite start vector");
  axpby(c, n, 1.0 / nx, x0, 0.0, Vc(0));
  while (true) {
    std::fill(H.begin(), H.end(), 0.0);
    K = 0;
    // the next application is enqueued behind the Gram-Schmidt kernels, before the host waits
    // for the coefficients and computes the Ritz pair; unused if that test converges
    bool pending = false;  // op(V_K) already enqueued
    while (true) {
      if (!pending) op(Vc(K), Vc(K + 1));
      pending = false;
      numops++;
      double beta = 0.0;
      cgs2_enqueue(c, n, K + 1, V, Vc(K + 1));
      if (K + 1 < kmax) {
        op(Vc(K + 1), Vc(K + 2));
        pending = true;
      }
      cgs2_wait(c, K + 1, h.data(), &beta);
      for (int i = 0; i <= K; ++i) H[i + (size_t)ldh * K] = h[i];
      H[(K + 1) + (size_t)ldh * K] = beta;
      K++;
      if (K == kmax || beta <= tol || eager) {
        double last = 0.0;
        const int hrc = hess_dominant_c(K, H.data(), ldh, &lam, &lim, u.data(), &last);
        require(hrc != 2, IMPS_E_SOLVER, "arnoldi: Hessenberg eigenvalue iteration failed");
        if (complex_dominant) *complex_dominant = (hrc == 1);
        if (hrc == 0) lim = 0.0;  // a real eigenvalue: drop the solver's rounding-level imaginary part
        res = std::fabs(beta) * last;
        if (res <= tol) {
          conv = true;
          break;
        }
      }
      if (K == kmax) break;
    }
    multiaxpy(c, n, K, V, n, u.data(), 0.0, xout);
    if (conv || numiter == maxiter) break;
    const double nr = nrm2(c, n, xout);
    axpby(c, n, 1.0 / nr, xout, 0.0, Vc(0));
    numiter++;
  }
  if (info) {
    info->converged = conv ? 1 : 0;
    info->numiter = numiter;
    info->numops = numops;
    info->normres = res;
  }
  if (lam_im) *lam_im = lim;
  return lam;
}

// Restarted GMRES for (a0 + a1 A) x = b (KrylovKit.linsolve via envs.jl:14-19); x holds
// the starting vector on entry and the solution on exit.  Absolute tolerance.
void gmres(imps_ctx* c, const Op& op, size_t n, const double* b, double* x, double tol, int kdim,
           int maxiter, double a0, double a1, KInfo* info) {
  const int kmax = (int)std::min<size_t>(kdim, n);
  require(kmax >= 1 && kmax < kMaxVec, IMPS_E_ARG, "gmres: krylovdim out of range");
  double* V = ws(c, "gm_V", (size_t)(kmax + 1) * n);
  double* r = ws(c, "gm_r", n);
  double* t = ws(c, "gm_t", n);
  auto Vc = [&](int k) { return V + (size_t)k * n; };
  auto applyA = [&](const double* v, double* out) {
    op(v, out);
    axpby(c, n, a0, v, a1, out);
  };
  int numops = 1, numiter = 0;
  applyA(x, t);
  d2d(c, r, b, n);
  axpby(c, n, -1.0, t, 1.0, r);
  double beta = nrm2(c, n, r);
  if (beta < tol) {
    if (info) *info = KInfo{1, 0, numops, beta};
    return;
  }
  const int ldh = kmax + 1;
  std::vector<double> H((size_t)ldh * kmax), cs(kmax), sn(kmax), g(kmax + 1), h(kmax + 1), y(kmax);
  while (numiter < maxiter) {
    numiter++;
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    axpby(c, n, 1.0 / beta, r, 0.0, Vc(0));
    int k = 0;
    while (true) {
      applyA(Vc(k), Vc(k + 1));
      numops++;
      double hn = 0.0;
      cgs2_normalize(c, n, k + 1, V, Vc(k + 1), h.data(), &hn);
      for (int i = 0; i <= k; ++i) H[i + (size_t)ldh * k] = h[i];
      H[(k + 1) + (size_t)ldh * k] = hn;
      for (int i = 0; i < k; ++i) {
        const double a = H[i + (size_t)ldh * k], bb = H[(i + 1) + (size_t)ldh * k];
        H[i + (size_t)ldh * k] = cs[i] * a + sn[i] * bb;
        H[(i + 1) + (size_t)ldh * k] = -sn[i] * a + cs[i] * bb;
      }
      const double a = H[k + (size_t)ldh * k], bb = H[(k + 1) + (size_t)ldh * k];
      const double rho = std::hypot(a, bb);
      cs[k] = a / rho;
      sn[k] = bb / rho;
      H[k + (size_t)ldh * k] = rho;
      H[(k + 1) + (size_t)ldh * k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      k++;
      if (std::fabs(g[k]) <= tol || k == kmax) break;
    }
    for (int i = k - 1; i >= 0; --i) {
      double s = g[i];
      for (int j = i + 1; j < k; ++j) s -= H[i + (size_t)ldh * j] * y[j];
      y[i] = s / H[i + (size_t)ldh * i];
    }
    multiaxpy(c, n, k, V, n, y.data(), 1.0, x);
    applyA(x, t);
    numops++;
    d2d(c, r, b, n);
    axpby(c, n, -1.0, t, 1.0, r);
    beta = nrm2(c, n, r);
    if (beta < tol) {
      if (info) *info = KInfo{1, numiter, numops, beta};
      return;
    }
  }
  if (info) *info = KInfo{0, numiter, numops, beta};
}

// ------------------------------------------------------------------ QRpos / LQpos

// Householder QR (rocSOLVER geqrf/orgqr) with the sign fix of QRpos.
void qrpos_householder(imps_ctx* c, int m, int n, double* A, double* R) {
  double* tau = ws(c, "qr_tau", n);
  RBCHK(rocsolver_dgeqrf(c->rb, m, n, A, m, tau));
  HIPCHK(launch_copy_upper(n, A, m, R, c->stream));
  RBCHK(rocsolver_dorgqr(c->rb, m, n, n, A, m, tau));
  HIPCHK(launch_qr_signfix(m, n, A, m, R, n, ws(c, "qr_sign", n), c->stream));
}

// Cholesky factor of the n x n G in place (upper: G = R^T R, lower: G = L L^T); false on
// breakdown.  Blocked right-looking: 64 x 64 diagonal blocks factored in LDS, panel solves,
// trailing updates on the MFMA GEMM.  The lower factor is always computed; for `upper`
// it is mirrored into the upper triangle (R = L^T).  The other triangle is left holding
// scratch values, as with LAPACK's potrf.
// Blocked Cholesky G = L L^T (lower factor in place; mirrored to the upper triangle when
// `upper`, so that G then also holds R = L^T).  Enqueued only: a non-positive pivot sets
// *fail (substituting 1 so the factorisation continues with finite values); the caller
// clears *fail beforehand and reads it afterwards.
// Ut receives the inverted diagonal blocks U_j = L_jj^{-T} (64 x 64 each, ld 64, block j at
// Ut + 4096 j), which the triangular solves below reuse.
void chol_async(imps_ctx* c, int n, double* G, bool upper, int* fail, double* Ut_all) {
  double* X = ws(c, "chol_panel", (size_t)n * 64);
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int nb = std::min(64, n - j0);
    double* Djj = G + j0 + (size_t)n * j0;
    double* Ut = Ut_all + (size_t)4096 * (j0 / 64);
    HIPCHK(launch_chol_diag(nb, Djj, n, Ut, fail, c->stream));
    const int rows = n - j0 - nb;
    if (rows > 0) {
      // panel: X = P L^{-T} = P U (MFMA GEMM with the inverted diagonal block), written back
      // to P, then the trailing update G22 -= X X^T
      double* P = G + (j0 + nb) + (size_t)n * j0;
      gemm2(c, 'N', 'N', rows, nb, nb, 1.0, P, n, Ut, 64, 0.0, X, rows);
      HIPCHK(launch_copy2d(rows, nb, X, rows, P, n, c->stream));
      gemm2(c, 'N', 'T', rows, rows, nb, -1.0, X, rows, X, rows, 1.0, G + (j0 + nb) + (size_t)n * (j0 + nb), n);
    }
  }
  if (upper) HIPCHK(launch_mirror_lower(n, G, n, c->stream));
}

// Triangular solves of the Cholesky QR on the library's GEMM, recursive halving so that the
// updates are large GEMMs and only the 64-column leaves use the inverted diagonal blocks of
// chol_async (Ut_all, U_j = L_jj^{-T}, block j at Ut_all + 4096 j); results go out of place.
//
// Q (m x n, ld m) <- A R^{-1}, R = L^T with L the lower factor in G (ld ldg).  A's columns are
// updated in place (A is destroyed).  Columns [c0, c0 + nc):
//   X_1 = A_1 R_11^{-1};  A_2 -= X_1 R_12 (R_12[k, j] = L[j, k]);  X_2 = A_2 R_22^{-1}.
void trsm_right_upper(imps_ctx* c, int m, double* A, double* Q, const double* G, long ldg, const double* Ut_all,
                      int c0, int nc) {
  const int blocks = (nc + 63) / 64;
  if (blocks <= 1) {
    gemm2(c, 'N', 'N', m, nc, nc, 1.0, A + (size_t)m * c0, m, Ut_all + (size_t)4096 * (c0 / 64), 64, 0.0,
          Q + (size_t)m * c0, m);
    return;
  }
  const int h = 64 * ((blocks + 1) / 2);
  trsm_right_upper(c, m, A, Q, G, ldg, Ut_all, c0, h);
  gemm2(c, 'N', 'T', m, nc - h, h, -1.0, Q + (size_t)m * c0, m, G + (c0 + h) + ldg * c0, ldg, 1.0,
        A + (size_t)m * (c0 + h), m);
  trsm_right_upper(c, m, A, Q, G, ldg, Ut_all, c0 + h, nc - h);
}

// Q (m x n, ld m) <- L^{-1} M with L the lower factor in G (ld ldg); M's rows are updated in place.
// Rows [r0, r0 + nr): Y_1 = L_11^{-1} M_1;  M_2 -= L_21 Y_1;  Y_2 = L_22^{-1} M_2;  L_ii^{-1} = U_i^T.
void trsm_left_lower(imps_ctx* c, int m, int n, double* M, double* Q, const double* G, long ldg,
                     const double* Ut_all, int r0, int nr) {
  const int blocks = (nr + 63) / 64;
  if (blocks <= 1) {
    gemm2(c, 'T', 'N', nr, n, nr, 1.0, Ut_all + (size_t)4096 * (r0 / 64), 64, M + r0, m, 0.0, Q + r0, m);
    return;
  }
  const int h = 64 * ((blocks + 1) / 2);
  trsm_left_lower(c, m, n, M, Q, G, ldg, Ut_all, r0, h);
  gemm2(c, 'N', 'N', nr - h, n, h, -1.0, G + (r0 + h) + ldg * r0, ldg, Q + r0, m, 1.0, M + r0 + h, m);
  trsm_left_lower(c, m, n, M, Q, G, ldg, Ut_all, r0 + h, nr - h);
}

int read_flag(imps_ctx* c, const int* flag) {
  int h = 0;
  std::memcpy(&h, d2h_small(c, flag, sizeof(int)), sizeof(int));
  return h;
}

// Shifted Cholesky QR, three passes (one shifted pass, two plain passes): A (m x n, m >= n)
// is overwritten by Q, R (n x n) gets the upper factor with a positive diagonal, which is
// the QRpos factorisation.  Returns false (A restored) if a Cholesky factorisation breaks
// down.
// With `defer` the breakdown flag is the caller's (cleared by the caller, read later) and A is not
// backed up: the call returns true at once and the caller redoes its work on a breakdown.
bool cholqr3(imps_ctx* c, int m, int n, double* A, double* R, int* defer = nullptr, int nact = -1) {
  if (nact < 0 || nact > n) nact = n;
  const size_t nn = (size_t)n * n, mn = (size_t)m * n;
  double* G = ws(c, "cq_G", nn);
  double* R2 = ws(c, "cq_R2", nn);
  double* bak = defer ? nullptr : ws(c, "cq_bak", mn);
  int* fail = defer ? defer : (int*)ws(c, "chol_info", 1);
  double* Ut = ws(c, "chol_ut", (size_t)4096 * ((n + 63) / 64));
  double* Rp = ws(c, "cq_Rp", nn);
  double* W = ws(c, "cq_W", mn);
  if (!defer) {
    d2d(c, bak, A, mn);
    HIPCHK(launch_fill_int(1, 0, fail, c->stream));
  }
  // the passes alternate between A and W (each triangular solve writes out of place)
  double* cur = A;
  double* nxt = W;
  for (int pass = 0; pass < 3; ++pass) {
    gemm2(c, 'T', 'N', n, n, m, 1.0, cur, m, cur, m, 0.0, G, n);
    if (pass == 0)
      HIPCHK(launch_add_diag_trace(n, 11.0 * ((double)m * n + (double)n * (n + 1)) * 1.1102230246251565e-16, G, n,
                                   c->stream));
    HIPCHK(launch_pad_unit_diag(n, nact, G, n, c->stream));
    chol_async(c, n, G, true, fail, Ut);
    trsm_right_upper(c, m, cur, nxt, G, n, Ut, 0, n);  // nxt <- cur R_pass^{-1}
    std::swap(cur, nxt);
    // R_pass = L^T, the upper triangle of G (its strict lower triangle holds L)
    HIPCHK(launch_copy2d(n, n, G, n, pass == 0 ? R : Rp, n, c->stream));
    HIPCHK(launch_tri_keep(n, pass == 0 ? R : Rp, n, true, c->stream));
    if (pass > 0) {
      gemm2(c, 'N', 'N', n, n, n, 1.0, Rp, n, R, n, 0.0, R2, n);  // R <- R_pass R
      d2d(c, R, R2, nn);
    }
  }
  d2d(c, A, cur, mn);  // no-op when the result is already in A
  if (defer) {
    HIPCHK(launch_tri_keep(n, R, n, true, c->stream));
    HIPCHK(launch_mask_block(n, n, nact, nact, R, n, c->stream));
    return true;
  }
  // a breakdown in any pass: restore A (the passes after it ran on substituted pivots)
  if (read_flag(c, fail) != 0) {
    d2d(c, A, bak, mn);
    return false;
  }
  HIPCHK(launch_tri_keep(n, R, n, true, c->stream));
  HIPCHK(launch_mask_block(n, n, nact, nact, R, n, c->stream));
  return true;
}

// LQ counterpart: M (m x n, ld m, m <= n) = L Q computed as Q (in place of M) and L (m x m,
// lower, positive diagonal).
bool cholqr3_lq(imps_ctx* c, int m, int n, double* M, double* L, int* defer = nullptr, int nact = -1) {
  if (nact < 0 || nact > m) nact = m;
  const size_t mm = (size_t)m * m, mn = (size_t)m * n;
  double* G = ws(c, "cq_G", mm);
  double* L2 = ws(c, "cq_R2", mm);
  double* bak = defer ? nullptr : ws(c, "cq_bak", mn);
  int* fail = defer ? defer : (int*)ws(c, "chol_info", 1);
  double* Ut = ws(c, "chol_ut", (size_t)4096 * ((m + 63) / 64));
  double* Lp = ws(c, "cq_Rp", mm);
  double* W = ws(c, "cq_W", mn);
  if (!defer) {
    d2d(c, bak, M, mn);
    HIPCHK(launch_fill_int(1, 0, fail, c->stream));
  }
  double* cur = M;
  double* nxt = W;
  for (int pass = 0; pass < 3; ++pass) {
    gemm2(c, 'N', 'T', m, m, n, 1.0, cur, m, cur, m, 0.0, G, m);
    if (pass == 0)
      HIPCHK(launch_add_diag_trace(m, 11.0 * ((double)m * n + (double)m * (m + 1)) * 1.1102230246251565e-16, G, m,
                                   c->stream));
    HIPCHK(launch_pad_unit_diag(m, nact, G, m, c->stream));
    chol_async(c, m, G, false, fail, Ut);
    trsm_left_lower(c, m, n, cur, nxt, G, m, Ut, 0, m);  // nxt <- L_pass^{-1} cur
    std::swap(cur, nxt);
    // L_pass: the lower triangle of G
    HIPCHK(launch_copy2d(m, m, G, m, pass == 0 ? L : Lp, m, c->stream));
    HIPCHK(launch_tri_keep(m, pass == 0 ? L : Lp, m, false, c->stream));
    if (pass > 0) {
      gemm2(c, 'N', 'N', m, m, m, 1.0, L, m, Lp, m, 0.0, L2, m);  // L <- L L_pass
      d2d(c, L, L2, mm);
    }
  }
  d2d(c, M, cur, mn);
  if (defer) {
    HIPCHK(launch_tri_keep(m, L, m, false, c->stream));
    HIPCHK(launch_mask_block(m, m, nact, nact, L, m, c->stream));
    return true;
  }
  if (read_flag(c, fail) != 0) {
    d2d(c, M, bak, mn);
    return false;
  }
  HIPCHK(launch_tri_keep(m, L, m, false, c->stream));
  HIPCHK(launch_mask_block(m, m, nact, nact, L, m, c->stream));
  return true;
}

std::atomic<long> g_qr_fallbacks(0);

// A (m x n, lda = m, m >= n) is overwritten by Q; R (n x n) receives the factor with a
// non-negative diagonal (FiniteMPSAlgorithms' QRpos).  nact < n: only the first nact columns
// are active (the rest are zero, a bond held below the largest bond dimension); Q keeps zero
// columns there and R zero rows and columns.
void qrpos(imps_ctx* c, int m, int n, double* A, double* R, int* defer = nullptr, int nact = -1) {
  if (defer) {
    cholqr3(c, m, n, A, R, defer, nact);
    return;
  }
  if (cholqr3(c, m, n, A, R, nullptr, nact)) return;
  g_qr_fallbacks++;
  qrpos_householder(c, m, n, A, R);
  if (nact >= 0 && nact < n) {
    HIPCHK(launch_mask_block(m, n, m, nact, A, m, c->stream));
    HIPCHK(launch_mask_block(n, n, nact, nact, R, n, c->stream));
  }
}

// M (m x n, ld m, m <= n) = L Q with L (m x m) lower, non-negative diagonal and Q (m x n)
// row-orthonormal (LQpos).  Q may alias M.  nact < m: only the first nact rows are active.
void lqpos(imps_ctx* c, int m, int n, const double* M, double* L, double* Q, int* defer = nullptr, int nact = -1) {
  d2d(c, Q, M, (size_t)m * n);
  if (defer) {
    cholqr3_lq(c, m, n, Q, L, defer, nact);
    return;
  }
  if (cholqr3_lq(c, m, n, Q, L, nullptr, nact)) return;
  g_qr_fallbacks++;
  double* Mt = ws(c, "lq_Mt", (size_t)m * n);
  double* R = ws(c, "lq_R", (size_t)m * m);
  HIPCHK(launch_transpose(m, n, Q, m, Mt, n, c->stream));
  qrpos_householder(c, n, m, Mt, R);
  HIPCHK(launch_transpose(m, m, R, m, L, m, c->stream));
  HIPCHK(launch_transpose(n, m, Mt, n, Q, m, c->stream));
  if (nact >= 0 && nact < m) {
    HIPCHK(launch_mask_block(m, n, nact, n, Q, m, c->stream));
    HIPCHK(launch_mask_block(m, m, nact, nact, L, m, c->stream));
  }
}

// active dimension of the bond right of `site` / left of `site` (-1: the whole D)
int act_right(const imps_ctx* c, int site) {
  if (c->chi.empty()) return -1;
  const int N = (int)c->chi.size();
  return c->chi[((site % N) + N) % N];
}
int act_left(const imps_ctx* c, int site) { return act_right(c, site - 1); }

// ------------------------------------------------------------------ gauge fixing


// The end of one gauge sweep (ortho.jl:176-201 / 135-163): Cn /= ||Cn||, then eps = ||Cn - Cpre||
// (Cpre is overwritten), read back together with the sweep's Cholesky breakdown flag in one
// host round trip.  Returns eps; *broke receives the flag.
void sweep_end_enqueue(imps_ctx* c, size_t n, double* Cn, double* Cpre) {
  normalize_dev(c, n, Cn);
  HIPCHK(launch_axpby(n, 1.0, Cn, -1.0, Cpre, c->stream));
  double* e2 = ws(c, "sw_e2", 1);
  HIPCHK(launch_multidot(n, 1, Cpre, n, Cpre, c->dpart, c->stream));
  HIPCHK(launch_partials_reduce(1, dot_blocks(n), c->dpart, e2, c->stream));
}

// the readback half of sweep_end (after sweep_end_enqueue)
double sweep_end_read(imps_ctx* c, const int* flag, bool* broke) {
  double* e2 = ws(c, "sw_e2", 1);
  HIPCHK(hipMemcpyAsync(c->hscal, e2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->hscal + 1, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  int f = 0;
  std::memcpy(&f, c->hscal + 1, sizeof(int));
  *broke = f != 0;
  return std::sqrt(std::max(0.0, c->hscal[0]));
}

double sweep_end(imps_ctx* c, size_t n, double* Cn, double* Cpre, const int* flag, bool* broke) {
  sweep_end_enqueue(c, n, Cn, Cpre);
  return sweep_end_read(c, flag, broke);
}

constexpr double kTolMin = 1.0e-14, kTolMax = 1.0e-4;

double updatetol(double factor, int it, double eps, double tmin = kTolMin, double tmax = kTolMax) {
  it = std::max(it, 1);
  return std::min(std::max(eps * factor / std::sqrt((double)it), tmin), tmax);
}

// uniform_rightorth! (ortho.jl:166-202): AR, C from A and C0 (C0 is read before C is written)
int uniform_rightorth(imps_ctx* c, const std::vector<double*>& A, const double* C0, double tol,
                      int maxiter, int eig_miniter = 10) {
  const int N = (int)A.size(), D = c->D, d = c->d;
  const size_t DD = (size_t)D * D;
  double* Cn = c->C[N - 1];
  double* Cpre = ws(c, "ro_Cpre", DD);
  double* Csave = ws(c, "ro_Csave", DD);
  double* M = ws(c, "ro_M", (size_t)D * d * D);
  double* ev = ws(c, "ro_ev", DD);
  int* sflag = (int*)ws(c, "ro_flag", 1);
  d2d(c, Cpre, C0, DD);
  d2d(c, Cn, Cpre, DD);
  normalize_dev(c, DD, Cn);
  int it = 0;
  double eps = std::numeric_limits<double>::infinity();
  while (true) {
    if (it >= eig_miniter) {
      RoctxRange range("gauge_eigvec_refresh");
      const double tol_e = updatetol(1.0, 1, eps * eps);
      Op tm = [&](const double* v, double* out) {
        double* t0 = ws(c, "ro_t0", DD);
        double* t1 = ws(c, "ro_t1", DD);
        const double* cur = v;
        for (int l = N - 1; l >= 0; --l) {
          double* dst = (l == 0) ? out : ((N - 1 - l) % 2 == 0 ? t0 : t1);
          push_right2(c, cur, D, A[l], c->AR[l], dst, D);
          cur = dst;
        }
      };
      arnoldi_dominant(c, tm, DD, Cn, ev, tol_e, 30, 300, true, nullptr);
      double* Ldum = ws(c, "ro_Ldum", DD);
      lqpos(c, D, D, ev, Cn, Ldum, nullptr, act_left(c, 0));
    }
    // the sweep runs with the Cholesky breakdown checks deferred to its end (one host round trip
    // per sweep), as one captured kernel sequence (run_graphed); on a breakdown it is redone from
    // its input with per-site checks and the Householder fallback
    RoctxRange sweep_range("gauge_lq_sweep");
    std::vector<uintptr_t> key{1, (uintptr_t)N, (uintptr_t)D, (uintptr_t)d};
    for (int i = 0; i < N; ++i)
      for (const void* q : {(const void*)A[i], (const void*)c->C[i], (const void*)c->AR[i]}) key.push_back((uintptr_t)q);
    run_graphed(c, key, [&] {
      d2d(c, Cpre, Cn, DD);
      d2d(c, Csave, Cn, DD);
      HIPCHK(launch_fill_int(1, 0, sflag, c->stream));
      for (int i = N - 1; i >= 0; --i) {
        gemm2(c, 'N', 'N', D * d, D, D, 1.0, A[i], (long)D * d, c->C[i], D, 0.0, M, (long)D * d);
        lqpos(c, D, d * D, M, c->C[(i - 1 + N) % N], c->AR[i], sflag, act_left(c, i));
      }
      sweep_end_enqueue(c, DD, Cn, Cpre);
    });
    bool broke = false;
    eps = sweep_end_read(c, sflag, &broke);
    if (broke) {
      d2d(c, Cn, Csave, DD);
      d2d(c, Cpre, Csave, DD);
      for (int i = N - 1; i >= 0; --i) {
        gemm2(c, 'N', 'N', D * d, D, D, 1.0, A[i], (long)D * d, c->C[i], D, 0.0, M, (long)D * d);
        lqpos(c, D, d * D, M, c->C[(i - 1 + N) % N], c->AR[i], nullptr, act_left(c, i));
      }
      HIPCHK(launch_fill_int(1, 0, sflag, c->stream));
      eps = sweep_end(c, DD, Cn, Cpre, sflag, &broke);
    }
    it++;
    if (eps < tol) return it;
    if (it > maxiter) {
      fprintf(stderr, "imps: uniform_rightorth not converged (eps = %.3e)\n", eps);
      return it;
    }
  }
}

// uniform_leftorth! (ortho.jl:125-164)
int uniform_leftorth(imps_ctx* c, const std::vector<double*>& A, const double* C0, double tol,
                     int maxiter, int eig_miniter = 10) {
  const int N = (int)A.size(), D = c->D, d = c->d;
  const size_t DD = (size_t)D * D;
  double* Cn = c->C[N - 1];
  double* Cpre = ws(c, "lo_Cpre", DD);
  double* Csave = ws(c, "lo_Csave", DD);
  double* ev = ws(c, "lo_ev", DD);
  int* sflag = (int*)ws(c, "lo_flag", 1);
  d2d(c, Cpre, C0, DD);
  d2d(c, Cn, Cpre, DD);
  normalize_dev(c, DD, Cn);
  int it = 0;
  double eps = std::numeric_limits<double>::infinity();
  while (true) {
    if (it >= eig_miniter) {
      const double tol_e = updatetol(1.0, 1, eps * eps);
      Op tm = [&](const double* v, double* out) {
        double* t0 = ws(c, "lo_t0", DD);
        double* t1 = ws(c, "lo_t1", DD);
        const double* cur = v;
        for (int l = 0; l < N; ++l) {
          double* dst = (l == N - 1) ? out : (l % 2 == 0 ? t0 : t1);
          push_left2(c, cur, D, A[l], c->AL[l], dst, D);
          cur = dst;
        }
      };
      arnoldi_dominant(c, tm, DD, Cn, ev, tol_e, 30, 300, true, nullptr);
      double* Rf = ws(c, "lo_R", DD);
      qrpos(c, D, D, ev, Rf, nullptr, act_left(c, 0));
      d2d(c, Cn, Rf, DD);
    }
    d2d(c, Cpre, Cn, DD);
    d2d(c, Csave, Cn, DD);
    HIPCHK(launch_fill_int(1, 0, sflag, c->stream));
    for (int i = 0; i < N; ++i) {
      // M = C[i-1] A[i] written straight into AL[i], then QR in place (checks deferred as in
      // uniform_rightorth)
      gemm2(c, 'N', 'N', D, d * D, D, 1.0, c->C[(i - 1 + N) % N], D, A[i], D, 0.0, c->AL[i], D);
      qrpos(c, D * d, D, c->AL[i], c->C[i], sflag, act_right(c, i));
    }
    bool broke = false;
    eps = sweep_end(c, DD, Cn, Cpre, sflag, &broke);
    if (broke) {
      d2d(c, Cn, Csave, DD);
      d2d(c, Cpre, Csave, DD);
      for (int i = 0; i < N; ++i) {
        gemm2(c, 'N', 'N', D, d * D, D, 1.0, c->C[(i - 1 + N) % N], D, A[i], D, 0.0, c->AL[i], D);
        qrpos(c, D * d, D, c->AL[i], c->C[i], nullptr, act_right(c, i));
      }
      HIPCHK(launch_fill_int(1, 0, sflag, c->stream));
      eps = sweep_end(c, DD, Cn, Cpre, sflag, &broke);
    }
    it++;
    if (eps < tol) return it;
    if (it > maxiter) {
      fprintf(stderr, "imps: uniform_leftorth not converged (eps = %.3e)\n", eps);
      return it;
    }
  }
}

void mul_AL_C(imps_ctx* c, const double* AL, const double* Cm, double* AC) {
  const int D = c->D, d = c->d;
  gemm2(c, 'N', 'N', D * d, D, D, 1.0, AL, (long)D * d, Cm, D, 0.0, AC, (long)D * d);
}

// ------------------------------------------------------------------ environments

bool level_identity(imps_ctx* c, int i) {
  if (i == 0) return true;
  for (int s = 0; s < c->N; ++s) {
    if (i == c->w - 1) continue;
    if (!is_eye(wblock(c, s, i, i), c->d)) return false;
  }
  return true;
}

bool level_empty(imps_ctx* c, int i) {
  if (i == 0 || i == c->w - 1) return false;
  for (int s = 0; s < c->N; ++s) {
    if (!is_zero(wblock(c, s, i, i)) || !is_zero(wblock(c, s, 0, i)) || !is_zero(wblock(c, s, i, c->w - 1)))
      return false;
  }
  return true;
}

struct EnvStats {
  int solves = 0, ops = 0;
};

// DMRGCache(psi, H::SparseIMPO; tol, maxiter, krylovdim, init_lefts, init_rights)
// (envs.jl:321-518) for square Schur tensors; warm start from the current levels.
void envs_build(imps_ctx* c, double tol, int maxiter, int krylovdim, EnvStats* st) {
  need_match(c);
  const int N = c->N, D = c->D, w = c->w;
  const size_t DD = (size_t)D * D;
  const long ldg = (long)D * w;
  const bool warm = c->envs_valid;
  // old levels (GMRES starting vectors) are kept in separate buffers
  std::vector<double*> oldL(N), oldR(N);
  for (int l = 0; l < N; ++l) {
    oldL[l] = ws(c, "env_oldL" + std::to_string(l), (size_t)D * w * D);
    oldR[l] = ws(c, "env_oldR" + std::to_string(l), (size_t)D * w * D);
    if (warm) {
      d2d(c, oldL[l], c->GL[l], (size_t)D * w * D);
      d2d(c, oldR[l], c->GR[l], (size_t)D * w * D);
    }
  }
  for (int l = 0; l < N; ++l) {
    fill(c, (size_t)D * w * D, 0.0, c->GL[l]);
    fill(c, (size_t)D * w * D, 0.0, c->GR[l]);
    HIPCHK(launch_set_identity(D, c->GL[l], ldg, c->stream));
    HIPCHK(launch_set_identity(D, c->GR[l] + (size_t)D * (w - 1), ldg, c->stream));
    if (!c->chi.empty()) {  // the identity of each bond's active block
      HIPCHK(launch_mask_block(D, D, act_left(c, l), act_left(c, l), c->GL[l], ldg, c->stream));
      HIPCHK(launch_mask_block(D, D, act_right(c, l), act_right(c, l), c->GR[l] + (size_t)D * (w - 1), ldg,
                               c->stream));
    }
    c->gl_id[l] = 1;
    c->gr_id[l] = 1;
  }
  const int kdim = (int)std::min<long>((long)D * D, krylovdim);
  double* rhs = ws(c, "env_rhs", DD);
  double* x = ws(c, "env_x", DD);
  double* rho = ws(c, "env_rho", DD);
  auto slice = [&](double* G, int k) { return G + (size_t)D * k; };

  auto left_terms = [&](int site, int i, int lmax) {
    std::vector<SliceTerm> terms;
    for (int l = 0; l <= lmax; ++l) {
      std::vector<double> Wb = wblock(c, site, l, i);
      if (is_zero(Wb)) continue;
      terms.push_back(SliceTerm{l == 0 ? nullptr : slice(c->GL[site], l), ldg, Wb});
    }
    return terms;
  };
  auto right_terms = [&](int site, int i) {
    std::vector<SliceTerm> terms;
    for (int l = i; l < w; ++l) {
      std::vector<double> Wb = wblock(c, site, i, l);
      if (is_zero(Wb)) continue;
      terms.push_back(SliceTerm{l == w - 1 ? nullptr : slice(c->GR[site], l), ldg, Wb});
    }
    return terms;
  };
  auto left_cycle = [&](int i, bool closing) {
    for (int site = 0; site < N; ++site) {
      const int sn = (site + 1) % N;
      slice_sum_left(c, left_terms(site, i, closing ? w - 1 : i), c->AL[site], slice(c->GL[sn], i), ldg);
    }
  };
  auto right_cycle = [&](int i) {
    for (int site = N - 1; site >= 0; --site) {
      const int sp = (site - 1 + N) % N;
      slice_sum_right(c, right_terms(site, i), c->AR[site], slice(c->GR[sp], i), ldg);
    }
  };
  auto solve = [&](const Op& op, int which_side, int i) {
    // x0: previous level (warm start) or the zero level
    double* G = which_side == 0 ? c->GL[0] : c->GR[N - 1];
    double* old = which_side == 0 ? oldL[0] : oldR[N - 1];
    HIPCHK(launch_copy2d(D, D, slice(G, i), ldg, rhs, D, c->stream));
    if (warm)
      HIPCHK(launch_copy2d(D, D, slice(old, i), ldg, x, D, c->stream));
    else
      fill(c, DD, 0.0, x);
    KInfo info;
    gmres(c, op, DD, rhs, x, tol, kdim, maxiter, 1.0, -1.0, &info);
    if (st) {
      st->solves++;
      st->ops += info.numops;
    }
    HIPCHK(launch_copy2d(D, D, x, D, slice(G, i), ldg, c->stream));
  };
  auto regularize_slice = [&](double* v, long ld, const double* l, int nact) {
    regularize_dev(c, D, v, ld, l, nact);
  };
  auto ring_left = [&](const double* v, double* out, double* rh) {
    double* t0 = ws(c, "env_t0", DD);
    double* t1 = ws(c, "env_t1", DD);
    const double* cur = v;
    for (int l = 0; l < N; ++l) {
      double* dst = (l == N - 1) ? out : (l % 2 == 0 ? t0 : t1);
      push_left2(c, cur, D, c->AL[l], c->AL[l], dst, D);
      cur = dst;
    }
    if (rh) regularize_slice(out, D, rh, act_left(c, 0));
  };
  auto ring_right = [&](const double* v, double* out, double* rh) {
    double* t0 = ws(c, "env_t0", DD);
    double* t1 = ws(c, "env_t1", DD);
    const double* cur = v;
    for (int l = N - 1; l >= 0; --l) {
      double* dst = (l == 0) ? out : ((N - 1 - l) % 2 == 0 ? t0 : t1);
      push_right2(c, cur, D, c->AR[l], c->AR[l], dst, D);
      cur = dst;
    }
    if (rh) regularize_slice(out, D, rh, act_right(c, N - 1));
  };

  // ---- left interior levels (envs.jl:364-421)
  for (int i = 1; i <= w - 2; ++i) {
    left_cycle(i, false);
    if (level_identity(c, i)) {
      gemm2(c, 'N', 'T', D, D, D, 1.0, c->C[N - 1], D, c->C[N - 1], D, 0.0, rho, D);
      solve([&](const double* v, double* o) { ring_left(v, o, rho); }, 0, i);
      if (N > 1) left_cycle(i, false);
      for (int l = 0; l < N; ++l) {
        double* rl = ws(c, "env_rhol", DD);
        gemm2(c, 'N', 'T', D, D, D, 1.0, c->C[(l - 1 + N) % N], D, c->C[(l - 1 + N) % N], D, 0.0, rl, D);
        regularize_slice(slice(c->GL[l], i), ldg, rl, act_left(c, l));
      }
    } else {
      if (!level_empty(c, i)) {
        bool zero_diag = true;
        for (int s = 0; s < N; ++s) zero_diag = zero_diag && is_zero(wblock(c, s, i, i));
        solve([&](const double* v, double* o) {
          if (zero_diag) {
            fill(c, DD, 0.0, o);
            return;
          }
          double* t0 = ws(c, "env_t0", DD);
          double* t1 = ws(c, "env_t1", DD);
          const double* cur = v;
          for (int l = 0; l < N; ++l) {
            double* dst = (l == N - 1) ? o : (l % 2 == 0 ? t0 : t1);
            std::vector<SliceTerm> tt{SliceTerm{cur, D, wblock(c, l, i, i)}};
            slice_sum_left(c, tt, c->AL[l], dst, D);
            cur = dst;
          }
        }, 0, i);
      }
      if (N > 1) left_cycle(i, false);
    }
  }
  // ---- left closing level (envs.jl:424-455)
  {
    const int i = w - 1;
    left_cycle(i, true);
    gemm2(c, 'N', 'T', D, D, D, 1.0, c->C[N - 1], D, c->C[N - 1], D, 0.0, rho, D);
    solve([&](const double* v, double* o) { ring_left(v, o, rho); }, 0, i);
    if (N > 1) left_cycle(i, true);
    for (int l = 0; l < N; ++l) {
      double* rl = ws(c, "env_rhol", DD);
      gemm2(c, 'N', 'T', D, D, D, 1.0, c->C[(l - 1 + N) % N], D, c->C[(l - 1 + N) % N], D, 0.0, rl, D);
      regularize_slice(slice(c->GL[l], i), ldg, rl, act_left(c, l));
    }
  }
  // ---- right levels (envs.jl:461-515)
  for (int i = w - 2; i >= 0; --i) {
    right_cycle(i);
    if (level_identity(c, i)) {
      gemm2(c, 'T', 'N', D, D, D, 1.0, c->C[N - 1], D, c->C[N - 1], D, 0.0, rho, D);
      solve([&](const double* v, double* o) { ring_right(v, o, rho); }, 1, i);
      if (N > 1) right_cycle(i);
      for (int l = 0; l < N; ++l) {
        double* rl = ws(c, "env_rhol", DD);
        gemm2(c, 'T', 'N', D, D, D, 1.0, c->C[l], D, c->C[l], D, 0.0, rl, D);
        regularize_slice(slice(c->GR[l], i), ldg, rl, act_right(c, l));
      }
    } else {
      if (!level_empty(c, i)) {
        bool zero_diag = true;
        for (int s = 0; s < N; ++s) zero_diag = zero_diag && is_zero(wblock(c, s, i, i));
        solve([&](const double* v, double* o) {
          if (zero_diag) {
            fill(c, DD, 0.0, o);
            return;
          }
          double* t0 = ws(c, "env_t0", DD);
          double* t1 = ws(c, "env_t1", DD);
          const double* cur = v;
          for (int l = N - 1; l >= 0; --l) {
            double* dst = (l == 0) ? o : ((N - 1 - l) % 2 == 0 ? t0 : t1);
            std::vector<SliceTerm> tt{SliceTerm{cur, D, wblock(c, l, i, i)}};
            slice_sum_right(c, tt, c->AR[l], dst, D);
            cur = dst;
          }
        }, 1, i);
      }
      if (N > 1) right_cycle(i);
    }
  }
  c->envs_valid = true;
}

// ------------------------------------------------------------------ observables

double dot_psi(imps_ctx* c) {
  const int N = c->N, D = c->D;
  const size_t DD = (size_t)D * D;
  double* v0 = ws(c, "dp_v0", DD);
  double* ev = ws(c, "dp_ev", DD);
  HIPCHK(launch_set_identity(D, v0, D, c->stream));
  if (!c->chi.empty()) HIPCHK(launch_mask_block(D, D, act_left(c, 0), act_left(c, 0), v0, D, c->stream));
  Op tm = [&](const double* v, double* out) {
    double* t0 = ws(c, "dp_t0", DD);
    double* t1 = ws(c, "dp_t1", DD);
    const double* cur = v;
    for (int l = 0; l < N; ++l) {
      double* dst = (l == N - 1) ? out : (l % 2 == 0 ? t0 : t1);
      push_left2(c, cur, D, c->AL[l], c->AL[l], dst, D);
      cur = dst;
    }
  };
  return arnoldi_dominant(c, tm, DD, v0, ev, 1.0e-10, 30, 300, true, nullptr);
}

double energy(imps_ctx* c) {
  need_envs(c);
  const int N = c->N, D = c->D, w = c->w, d = c->d;
  double E = 0.0;
  for (int site = 0; site < N; ++site) {
    double* T1 = ws(c, "hac_T1", (size_t)D * w * d * D);
    double* T2 = ws(c, "hac_T2", (size_t)D * w * d * D);
    gemm2(c, 'N', 'N', D * w, d * D, D, 1.0, c->GL[site], (long)D * w, c->AC[site], D, 0.0, T1, (long)D * w);
    HIPCHK(launch_wmix(D, w, d, c->Wm[site], T1, T2, nullptr, nullptr, c->stream));
    E += dot(c, (size_t)D * d * D, c->AC[site], T2 + (size_t)D * d * D * (w - 1));
  }
  return E / dot_psi(c);
}

// d e_cell / d W[site][i,u,j,s] at fixed psi: sum GL[b,i,a] AC[a,s,bb] GR[bb,j,c] AC[b,u,c],
// divided by dot(psi, psi).  At a variational fixed point this is the derivative of the
// converged energy for every non-structural Schur block.
void energy_gradient(imps_ctx* c, int site, double* out) {
  const int D = c->D, w = c->w, d = c->d, wd = w * d;
  double* T1 = ws(c, "hac_T1", (size_t)D * w * d * D);
  double* T3 = ws(c, "hac_T2", (size_t)D * w * d * D);
  double* Gp = ws(c, "eg_part", (size_t)wd * wd * D);
  double* G = ws(c, "eg_sum", (size_t)wd * wd);
  gemm2(c, 'N', 'N', D * w, d * D, D, 1.0, c->GL[site], (long)D * w, c->AC[site], D, 0.0, T1, (long)D * w);
  gemm2(c, 'N', 'T', D * d, D * w, D, 1.0, c->AC[site], (long)D * d, c->GR[site], (long)D * w, 0.0, T3, (long)D * d);
  HIPCHK(launch_gsand(D, w, d, T1, T3, Gp, c->stream));
  HIPCHK(launch_partials_reduce(wd * wd, D, Gp, G, c->stream));
  const double* hp = d2h_small(c, G, sizeof(double) * wd * wd);
  std::vector<double> h(hp, hp + (size_t)wd * wd);
  const double nrm = dot_psi(c);
  for (int i = 0; i < w; ++i)
    for (int u = 0; u < d; ++u)
      for (int j = 0; j < w; ++j)
        for (int s2 = 0; s2 < d; ++s2)
          out[i + (size_t)w * (u + (size_t)d * (j + (size_t)w * s2))] =
              h[(i + (size_t)w * s2) + (size_t)wd * (u + (size_t)d * j)] / nrm;
}

double galerkin(imps_ctx* c, int prec) {
  need_envs(c);
  const int N = c->N, D = c->D, d = c->d;
  const size_t n = (size_t)D * d * D;
  double eps = 0.0;
  for (int site = 0; site < N; ++site) {
    double* x = ws(c, "gk_x", n);
    double* q = ws(c, "gk_q", (size_t)D * D);
    hac_apply(c, site, c->AC[site], x, prec);
    scal(c, n, 1.0 / nrm2(c, n, x), x);
    gemm2(c, 'T', 'N', D, D, d * D, 1.0, c->AL[site], (long)d * D, x, (long)d * D, 0.0, q, D);
    gemm2(c, 'N', 'N', d * D, D, D, -1.0, c->AL[site], (long)d * D, q, D, 1.0, x, (long)d * D);
    eps = std::max(eps, nrm2(c, n, x));
  }
  return eps;
}

// ------------------------------------------------------------------ fixed-point response
//
// Tangent-space linear response of a converged one-site state.  The reference has no derivative
// of its fixed point; this one is built from its operators: H_AC (effective.jl:90-102), the slice
// pushes (envs.jl:104-114), regularize! (envs.jl:28-32), the GMRES linsolve (envs.jl:14-19) and
// the level order of DMRGCache (envs.jl:321-518).
//
// Tangent vectors are tensors B of the MPS shape with AL^T B = 0; P = 1 - AL AL^T.  The Hessian
// of the energy density along AL + t B C^-1 is H_eff + K with
//   H_eff B = P[(H_AC - lambda) B + H_AC[LK, GR](AR) + H_AC[GL, RK](AL)]   (lambda = <AC, H_AC AC>)
//   K B     = P[H_AC[GL, RKb](AC) + C^-T H_AC[LKb, GR](AC)]                 (manifold curvature)
// whose one-defect environments are sums over the defect's distance, solved level by level; an
// identity diagonal level solves x - T x + tr(dual x) fix = rhs - tr(dual rhs) fix with the fixed
// point `fix` of the mixed transfer matrix and its dual:
//   LK   ket B on the left, then ket AR / bra AL          fix C,     dual C^T
//   RK   ket B on the right, then ket AL / bra AR         fix C,     dual C^T
//   LKb  bra B on the left, then bra AR / ket AL          fix C^T,   dual C
//   RKb  bra C^-1 B on the right, then AR / AR            fix I,     dual C^T C
// For an objective MPO O with tangent gradient g = P[H_AC^O(AC)] the response of <O> to the
// Hamiltonian tensor is dO/dW = -2 d/dW <Phi(L)| H[W] |psi> with (H_eff + K) L = g: the sandwich
// of the reverse H_AC contraction at the defect site plus the adjoint environments (bra defect,
// ket psi) on either side of it.  C^-1 is the pseudo-inverse with relative cutoff rcond.

struct Tangent {
  const double *AL, *AR, *C, *AC, *GL, *GR;
  bool gl_id, gr_id;                 // level 0 of GL / level w-1 of GR is the identity
  double lam = 0.0;
  double *Cinv, *Ct, *CtC, *CCt, *Id;  // D x D on the device
  double tol = 1e-13;                // relative tolerance of the level solves
  long solves = 0, ops = 0, hessian_applies = 0, unconverged = 0;
};

// v -= tr(dual v) fix for n x n matrices (v with leading dimension ldv; dual, fix contiguous),
// the trace reduced and applied on the device
void sub_trace_fix(imps_ctx* c, int n, double* v, long ldv, const double* dual, const double* fix) {
  int np = 0;
  const size_t maxp = (size_t)((n + 31) / 32) * ((n + 31) / 32);
  double* part = ws(c, "stf_part", maxp);
  HIPCHK(launch_trace_prod_parts(n, dual, n, v, ldv, part, &np, c->stream));
  HIPCHK(launch_reduce_axpy_mat(np, part, -1.0, fix, n, v, ldv, n, c->stream));
}

// H_AC[GL, GR] x (effective.jl:90-102) with every level through the GEMMs: environments of any
// level structure (the defect environments have no identity level), the model of site 0
void hac_apply_env(imps_ctx* c, const double* GL, const double* GR, const double* x, double* y) {
  const int D = c->D, w = c->w, d = c->d;
  double* T1 = ws(c, "rsp_T1", (size_t)D * w * d * D);
  double* T2 = ws(c, "rsp_T2", (size_t)D * w * d * D);
  gemm2(c, 'N', 'N', D * w, d * D, D, 1.0, GL, (long)D * w, x, D, 0.0, T1, (long)D * w);
  HIPCHK(launch_wmix(D, w, d, c->Wm[0], T1, T2, nullptr, nullptr, c->stream));
  gemm2(c, 'N', 'N', D * d, D, D * w, 1.0, T2, (long)D * d, GR, (long)D * w, 0.0, y, (long)D * d);
}

// E <- P E = E - AL (AL^T E) (the projector of calc_galerkin, vumps.jl:111-124)
void tangent_project(imps_ctx* c, const double* AL, double* E) {
  const int D = c->D, d = c->d;
  double* q = ws(c, "tg_q", (size_t)D * D);
  gemm2(c, 'T', 'N', D, D, d * D, 1.0, AL, (long)d * D, E, (long)d * D, 0.0, q, D);
  gemm2(c, 'N', 'N', d * D, D, D, -1.0, AL, (long)d * D, q, D, 1.0, E, (long)d * D);
}

// One level of a defect environment: x = rhs + push(x, M) with the diagonal block M; rhs (D x D,
// contiguous) is overwritten; the solution goes to out (leading dimension ldo).
void defect_level_solve(imps_ctx* c, Tangent& t, bool left, double* rhs, const std::vector<double>& M,
                        const double* bra, const double* ket, const double* fix, const double* dual, double* out,
                        long ldo) {
  const int D = c->D;
  const size_t DD = (size_t)D * D;
  if (is_zero(M)) {
    HIPCHK(launch_copy2d(D, D, rhs, D, out, ldo, c->stream));
    return;
  }
  const bool eye = is_eye(M, c->d, 0.0);
  if (eye) sub_trace_fix(c, D, rhs, D, dual, fix);
  const double nb = nrm2(c, DD, rhs);
  if (nb == 0.0) {
    HIPCHK(launch_fill2d(D, D, 0.0, out, ldo, c->stream));
    return;
  }
  Op op = [&](const double* X, double* o) {
    std::vector<SliceTerm> tt{SliceTerm{X, D, M}};
    if (left)
      slice_sum_left_bk(c, tt, bra, ket, o, D, 0.0);
    else
      slice_sum_right_bk(c, tt, bra, ket, o, D, 0.0);
    if (eye) sub_trace_fix(c, D, o, D, dual, fix);
  };
  double* x = ws(c, "df_x", DD);
  fill(c, DD, 0.0, x);
  KInfo info;
  gmres(c, op, DD, rhs, x, t.tol * nb, (int)std::min<size_t>(DD, 30), 300, 1.0, -1.0, &info);
  t.solves++;
  t.ops += info.numops;
  if (!info.converged) t.unconverged++;
  HIPCHK(launch_copy2d(D, D, x, D, out, ldo, c->stream));
}

// Levels lo..hi of a defect environment E (D, w, D) of the Hamiltonian of site 0; levels outside
// lo..hi are zero.  Left (level j ascending):
//   E_j = sum_{i<=j} push(G_i, W[i,:,j,:]; sbra, sket) + sum_{lo<=i<j} push(E_i, W[i,:,j,:]; bra, ket)
//         + push(E_j, W[j,:,j,:]; bra, ket)
// right (level i descending) with W[i,:,j,:], j >= i, and the right pushes.  g_id: the level of
// G that is the identity (-1: none), applied without a GEMM.
void defect_env(imps_ctx* c, Tangent& t, bool left, const double* G, int g_id, const double* sbra, const double* sket,
                const double* bra, const double* ket, const double* fix, const double* dual, int lo, int hi,
                double* E) {
  const int D = c->D, w = c->w, d = c->d;
  const long ldg = (long)D * w;
  const std::vector<double>& W = c->Wh[0];
  fill(c, (size_t)D * w * D, 0.0, E);
  double* rhs = ws(c, "df_rhs", (size_t)D * D);
  for (int step = 0; step <= hi - lo; ++step) {
    const int j = left ? lo + step : hi - step;  // the level being solved
    std::vector<SliceTerm> ts, te;
    if (left) {
      for (int i = 0; i <= j; ++i) {
        std::vector<double> M = wblock_of(W, w, d, i, j);
        if (!is_zero(M)) ts.push_back(SliceTerm{i == g_id ? nullptr : G + (size_t)D * i, ldg, M});
        if (i >= lo && i < j && !is_zero(M)) te.push_back(SliceTerm{E + (size_t)D * i, ldg, M});
      }
      slice_sum_left_bk(c, ts, sbra, sket, rhs, D, 0.0);
      slice_sum_left_bk(c, te, bra, ket, rhs, D, 1.0);
    } else {
      for (int k = j; k < w; ++k) {
        std::vector<double> M = wblock_of(W, w, d, j, k);
        if (!is_zero(M)) ts.push_back(SliceTerm{k == g_id ? nullptr : G + (size_t)D * k, ldg, M});
        if (k > j && k <= hi && !is_zero(M)) te.push_back(SliceTerm{E + (size_t)D * k, ldg, M});
      }
      slice_sum_right_bk(c, ts, sbra, sket, rhs, D, 0.0);
      slice_sum_right_bk(c, te, bra, ket, rhs, D, 1.0);
    }
    defect_level_solve(c, t, left, rhs, wblock_of(W, w, d, j, j), bra, ket, fix, dual, E + (size_t)D * j, ldg);
  }
}

// out = P (H_eff + K) P B (projecting the input keeps the operator symmetric on the whole space)
void tangent_hessian(imps_ctx* c, Tangent& t, const double* Bin, double* out) {
  const int D = c->D, w = c->w, d = c->d;
  const size_t n = (size_t)D * d * D;
  double* B = ws(c, "tg_B", n);
  double* X = ws(c, "tg_X", n);
  double* CiB = ws(c, "tg_CiB", n);
  double* Ek = ws(c, "tg_env", (size_t)D * w * D);
  const int gl0 = t.gl_id ? 0 : -1, grw = t.gr_id ? w - 1 : -1;
  d2d(c, B, Bin, n);
  tangent_project(c, t.AL, B);
  // H_eff
  hac_apply_env(c, t.GL, t.GR, B, out);
  axpby(c, n, -t.lam, B, 1.0, out);
  defect_env(c, t, true, t.GL, gl0, t.AL, B, t.AL, t.AR, t.C, t.Ct, 1, w - 1, Ek);
  hac_apply_env(c, Ek, t.GR, t.AR, X);
  axpby(c, n, 1.0, X, 1.0, out);
  defect_env(c, t, false, t.GR, grw, t.AR, B, t.AR, t.AL, t.C, t.Ct, 1, w - 1, Ek);
  hac_apply_env(c, t.GL, Ek, t.AL, X);
  axpby(c, n, 1.0, X, 1.0, out);
  // curvature
  gemm2(c, 'N', 'N', D, d * D, D, 1.0, t.Cinv, D, B, D, 0.0, CiB, D);
  defect_env(c, t, false, t.GR, grw, CiB, t.AR, t.AR, t.AR, t.Id, t.CtC, 1, w - 1, Ek);
  hac_apply_env(c, t.GL, Ek, t.AC, X);
  axpby(c, n, 1.0, X, 1.0, out);
  defect_env(c, t, true, t.GL, gl0, B, t.AL, t.AR, t.AL, t.Ct, t.C, 1, w - 1, Ek);
  hac_apply_env(c, Ek, t.GR, t.AC, X);
  gemm2(c, 'T', 'N', D, d * D, D, 1.0, t.Cinv, D, X, D, 1.0, out, D);
  tangent_project(c, t.AL, out);
  t.hessian_applies++;
}

// acc[W layout (w, d, w, d)] += scale * sum L[b,i,a] bra[b,u,c] ket[a,s,x] R[x,j,c]: the reverse
// H_AC contraction (the derivative of <bra| H_AC[L, R] |ket> with respect to W[i,u,j,s])
void sandwich_acc(imps_ctx* c, const double* L, const double* bra, const double* ket, const double* R, double scale,
                  std::vector<double>& acc) {
  const int D = c->D, w = c->w, d = c->d, wd = w * d;
  double* T1 = ws(c, "rsp_T1", (size_t)D * w * d * D);
  double* T3 = ws(c, "rsp_T2", (size_t)D * w * d * D);
  double* Gp = ws(c, "eg_part", (size_t)wd * wd * D);
  double* Gs = ws(c, "eg_sum", (size_t)wd * wd);
  gemm2(c, 'N', 'N', D * w, d * D, D, 1.0, L, (long)D * w, ket, D, 0.0, T1, (long)D * w);
  gemm2(c, 'N', 'T', D * d, D * w, D, 1.0, bra, (long)D * d, R, (long)D * w, 0.0, T3, (long)D * d);
  HIPCHK(launch_gsand(D, w, d, T1, T3, Gp, c->stream));
  HIPCHK(launch_partials_reduce(wd * wd, D, Gp, Gs, c->stream));
  const double* h = d2h_small(c, Gs, sizeof(double) * wd * wd);
  for (int i = 0; i < w; ++i)
    for (int u = 0; u < d; ++u)
      for (int j = 0; j < w; ++j)
        for (int s2 = 0; s2 < d; ++s2)
          acc[i + (size_t)w * (u + (size_t)d * (j + (size_t)w * s2))] +=
              scale * h[(i + (size_t)w * s2) + (size_t)wd * (u + (size_t)d * j)];
}

// Cinv = pseudo-inverse of the D x D matrix Cm, singular values below rcond * s_max dropped
// (rocSOLVER gesvd, C = U S V^T, Cinv = V S^+ U^T)
void pinv_cutoff(imps_ctx* c, const double* Cm, double rcond, double* Cinv) {
  const int D = c->D;
  const size_t DD = (size_t)D * D;
  double* A = ws(c, "pv_A", DD);
  double* U = ws(c, "pv_U", DD);
  double* Vt = ws(c, "pv_Vt", DD);
  double* S = ws(c, "pv_S", (size_t)2 * D + 2);
  int* info = (int*)ws(c, "pv_info", 1);
  d2d(c, A, Cm, DD);
  RBCHK(rocsolver_dgesvd(c->rb, rocblas_svect_all, rocblas_svect_all, D, D, A, D, S, U, D, Vt, D, S + D,
                         rocblas_outofplace, info));
  std::vector<double> sv(D);
  d2h(c, sv.data(), S, D);
  int inf = 0;
  d2h_bytes(c, &inf, info, sizeof(int));
  require(inf == 0, IMPS_E_SOLVER, "pseudo-inverse of C: gesvd did not converge");
  for (int k = 0; k < D; ++k) {
    const double s = sv[k];
    HIPCHK(launch_scale(D, (s > rcond * sv[0] && s > 0.0) ? 1.0 / s : 0.0, U + (size_t)D * k, c->stream));
  }
  gemm2(c, 'T', 'T', D, D, D, 1.0, Vt, D, U, D, 0.0, Cinv, D);
}

Tangent tangent_setup(imps_ctx* c, double rcond, double tol) {
  const int D = c->D, d = c->d;
  const size_t DD = (size_t)D * D, n = (size_t)D * d * D;
  Tangent t;
  t.AL = c->AL[0];
  t.AR = c->AR[0];
  t.C = c->C[0];
  t.AC = c->AC[0];
  t.GL = c->GL[0];
  t.GR = c->GR[0];
  t.gl_id = c->gl_id[0] != 0;
  t.gr_id = c->gr_id[0] != 0;
  t.tol = tol;
  t.Cinv = ws(c, "tg_Cinv", DD);
  t.Ct = ws(c, "tg_Ct", DD);
  t.CtC = ws(c, "tg_CtC", DD);
  t.CCt = ws(c, "tg_CCt", DD);
  t.Id = ws(c, "tg_Id", DD);
  pinv_cutoff(c, t.C, rcond, t.Cinv);
  HIPCHK(launch_transpose(D, D, t.C, D, t.Ct, D, c->stream));
  gemm2(c, 'T', 'N', D, D, D, 1.0, t.C, D, t.C, D, 0.0, t.CtC, D);
  gemm2(c, 'N', 'T', D, D, D, 1.0, t.C, D, t.C, D, 0.0, t.CCt, D);
  fill(c, DD, 0.0, t.Id);
  HIPCHK(launch_set_identity(D, t.Id, D, c->stream));
  double* y = ws(c, "tg_X", n);
  hac_apply_env(c, t.GL, t.GR, t.AC, y);
  t.lam = dot(c, n, t.AC, y);
  return t;
}

// d/dW <Phi(B)| H[W] |psi> of site 0 in the (w, d, w, d) layout: the defect site, then the adjoint
// environments (bra defect, ket psi) for the MPO tensors right of it (bra AR / ket AR between) and
// left of it (bra AL / ket AL between)
std::vector<double> tangent_w_sandwich(imps_ctx* c, Tangent& t, const double* B) {
  const int D = c->D, w = c->w, d = c->d;
  std::vector<double> S((size_t)w * d * w * d, 0.0);
  double* Ek = ws(c, "tg_env", (size_t)D * w * D);
  const int gl0 = t.gl_id ? 0 : -1, grw = t.gr_id ? w - 1 : -1;
  sandwich_acc(c, t.GL, B, t.AC, t.GR, 1.0, S);
  if (w >= 3) {
    defect_env(c, t, true, t.GL, gl0, B, t.AC, t.AR, t.AR, t.CtC, t.Id, 1, w - 2, Ek);
    sandwich_acc(c, Ek, t.AR, t.AR, t.GR, 1.0, S);
  }
  defect_env(c, t, false, t.GR, grw, B, t.AC, t.AL, t.AL, t.CCt, t.Id, 1, w - 1, Ek);
  sandwich_acc(c, t.GL, t.AL, t.AL, Ek, 1.0, S);
  return S;
}

// MINRES (Paige & Saunders 1975) for a symmetric operator from x = 0, on device vectors with the
// recurrence's scalars on the host; stops when the residual estimate is below max(tol |b|, atol).
// Returns the iteration count; *est receives the estimate relative to |b|.
int minres_dev(imps_ctx* c, const Op& op, size_t n, const double* b, double* x, double tol, int maxiter, double atol,
               double* est) {
  const double beta1 = nrm2(c, n, b);
  fill(c, n, 0.0, x);
  *est = 0.0;
  if (beta1 <= atol) return 0;
  double* r1 = ws(c, "mr_r1", n);
  double* r2 = ws(c, "mr_r2", n);
  double* y = ws(c, "mr_y", n);
  double* v = ws(c, "mr_v", n);
  double* wv = ws(c, "mr_w", n);
  double* w2 = ws(c, "mr_w2", n);
  double* sp = ws(c, "mr_w1", n);
  d2d(c, r1, b, n);
  d2d(c, r2, b, n);
  d2d(c, y, b, n);
  fill(c, n, 0.0, wv);
  fill(c, n, 0.0, w2);
  double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
  int it = 0;
  while (it < maxiter) {
    ++it;
    axpby(c, n, 1.0 / beta, r2, 0.0, v);  // r2 holds the latest Lanczos residual
    op(v, y);
    if (it >= 2) axpby(c, n, -(beta / oldb), r1, 1.0, y);
    const double alfa = dot(c, n, v, y);
    axpby(c, n, -(alfa / beta), r2, 1.0, y);
    std::swap(r1, r2);  // r1 <- r2
    std::swap(r2, y);   // r2 <- y (the old r1 buffer becomes y's, free until the next op)
    oldb = beta;
    beta = nrm2(c, n, r2);
    const double oldeps = epsln;
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    epsln = sn * beta;
    dbar = -cs * beta;
    const double gamma = std::max(std::hypot(gbar, beta), std::numeric_limits<double>::epsilon());
    cs = gbar / gamma;
    sn = beta / gamma;
    const double phi = cs * phibar;
    phibar = sn * phibar;
    // w1, w2 = w2, w; w = (v - oldeps w1 - delta w2) / gamma, built in the spare buffer
    axpby(c, n, 1.0 / gamma, v, 0.0, sp);
    axpby(c, n, -oldeps / gamma, w2, 1.0, sp);
    axpby(c, n, -delta / gamma, wv, 1.0, sp);
    double* old_w1 = w2;
    w2 = wv;
    wv = sp;
    sp = old_w1;
    axpby(c, n, phi, wv, 1.0, x);
    if (!std::isfinite(phibar)) break;
    if (phibar <= std::max(tol * beta1, atol) || beta == 0.0) break;
  }
  *est = phibar / beta1;
  return it;
}

// The context's model and environments, set aside while those of another MPO (an objective) are
// built on the same state; restored on destruction (also when an exception leaves the scope).
struct ModelSwap {
  imps_ctx* c;
  int w, d;
  std::vector<std::vector<double>> Wh;
  std::vector<double*> Wm, GL, GR;
  std::vector<char> gl_id, gr_id;
  bool envs_valid;
  int planes_site, hc_planes_site;
  std::vector<double*> owned;
  ModelSwap(imps_ctx* ctx, int wo, const double* Wobj) : c(ctx) {
    const int N = c->N, D = c->D, dd = c->d;
    w = c->w;
    d = c->d;
    Wh.swap(c->Wh);
    Wm.swap(c->Wm);
    GL.swap(c->GL);
    GR.swap(c->GR);
    gl_id.swap(c->gl_id);
    gr_id.swap(c->gr_id);
    envs_valid = c->envs_valid;
    planes_site = c->planes_site;
    hc_planes_site = c->hc_planes_site;
    c->w = wo;
    c->envs_valid = false;
    c->planes_site = -1;
    c->hc_planes_site = -1;
    const int wd = wo * dd;
    std::vector<double> Wl(Wobj, Wobj + (size_t)wo * dd * wo * dd);
    std::vector<double> Wmh((size_t)wd * wd);
    for (int i = 0; i < wo; ++i)
      for (int u = 0; u < dd; ++u)
        for (int j = 0; j < wo; ++j)
          for (int s = 0; s < dd; ++s)
            Wmh[(u + (size_t)dd * j) + (size_t)wd * (i + (size_t)wo * s)] =
                Wl[i + (size_t)wo * (u + (size_t)dd * (j + (size_t)wo * s))];
    try {
      for (int l = 0; l < N; ++l) {
        c->Wh.push_back(Wl);
        double* dp = dalloc(Wmh.size());
        owned.push_back(dp);
        h2d(c, dp, Wmh.data(), Wmh.size());
        c->Wm.push_back(dp);
        double* gl = dalloc((size_t)D * wo * D);
        owned.push_back(gl);
        double* gr = dalloc((size_t)D * wo * D);
        owned.push_back(gr);
        c->GL.push_back(gl);
        c->GR.push_back(gr);
        c->gl_id.push_back(0);
        c->gr_id.push_back(0);
      }
    } catch (...) {
      restore();
      throw;
    }
  }
  void restore() {
    (void)hipStreamSynchronize(c->stream);
    for (double*& p : owned) dfree(p);
    owned.clear();
    c->w = w;
    c->d = d;
    c->Wh.swap(Wh);
    c->Wm.swap(Wm);
    c->GL.swap(GL);
    c->GR.swap(GR);
    c->gl_id.swap(gl_id);
    c->gr_id.swap(gr_id);
    c->envs_valid = envs_valid;
    c->planes_site = planes_site;
    c->hc_planes_site = hc_planes_site;
  }
  ~ModelSwap() {
    if (c) restore();
  }
};

// ------------------------------------------------------------------ state management

void free_state(imps_ctx* c) {
  clear_graphs(c);
  for (auto* v : {&c->AL, &c->AR, &c->AC, &c->C, &c->pAC, &c->pC, &c->pAL, &c->GL, &c->GR})
    for (double*& p : *v) dfree(p);
  for (auto* v : {&c->AL, &c->AR, &c->AC, &c->C, &c->pAC, &c->pC, &c->pAL, &c->GL, &c->GR}) v->clear();
  c->envs_valid = false;
  c->D = 0;
  c->cplx = false;
  c->chi.clear();
}

void alloc_state(imps_ctx* c, int N, int D, int d) {
  free_state(c);
  c->D = D;
  const size_t nA = (size_t)D * d * D, nC = (size_t)D * D;
  for (int l = 0; l < N; ++l) {
    c->AL.push_back(dalloc(nA));
    c->AR.push_back(dalloc(nA));
    c->AC.push_back(dalloc(nA));
    c->C.push_back(dalloc(nC));
    c->pAC.push_back(dalloc(nA));
    c->pC.push_back(dalloc(nC));
    c->pAL.push_back(dalloc(nA));
  }
}

// the pending ALs of a freshly set state are its ALs, so that gauge_step (imps_gaugefix_right)
// before any regauge re-fixes the gauge of the state itself instead of reading stale buffers
void init_pending(imps_ctx* c) {
  const size_t nA = (size_t)c->D * c->d * c->D;
  for (size_t l = 0; l < c->AL.size(); ++l) d2d(c, c->pAL[l], c->AL[l], nA);
}

void alloc_envs(imps_ctx* c) {
  if (!c->GL.empty()) return;
  const size_t n = (size_t)c->D * c->w * c->D;
  for (int l = 0; l < c->N; ++l) {
    c->GL.push_back(dalloc(n));
    c->GR.push_back(dalloc(n));
  }
  c->gl_id.assign(c->N, 0);
  c->gr_id.assign(c->N, 0);
}

void envs_recalculate(imps_ctx* c, double tol, int maxiter, int krylovdim, EnvStats* st) {
  need_match(c);
  alloc_envs(c);
  envs_build(c, tol, maxiter, krylovdim, st);
}

void gauge_step(imps_ctx* c, double tol, int maxiter, int* sweeps, int eig_miniter = 10) {
  const int N = c->N, D = c->D, d = c->d;
  double* C0 = ws(c, "gs_C0", (size_t)D * D);
  d2d(c, C0, c->C[N - 1], (size_t)D * D);
  const int sw = uniform_rightorth(c, c->pAL, C0, tol, maxiter, std::max(0, eig_miniter));
  for (int l = 0; l < N; ++l) {
    d2d(c, c->AL[l], c->pAL[l], (size_t)D * d * D);
    mul_AL_C(c, c->AL[l], c->C[l], c->AC[l]);
  }
  if (sweeps) *sweeps = sw;
}

void regauge_site(imps_ctx* c, int site) {
  const int D = c->D, d = c->d;
  double* Qac = ws(c, "rg_Qac", (size_t)D * d * D);
  double* Qc = ws(c, "rg_Qc", (size_t)D * D);
  double* R = ws(c, "rg_R", (size_t)D * D);
  d2d(c, Qac, c->pAC[site], (size_t)D * d * D);
  d2d(c, Qc, c->pC[site], (size_t)D * D);
  qrpos(c, D * d, D, Qac, R, nullptr, act_right(c, site));
  qrpos(c, D, D, Qc, R, nullptr, act_right(c, site));
  gemm2(c, 'N', 'T', D * d, D, D, 1.0, Qac, (long)D * d, Qc, D, 0.0, c->pAL[site], (long)D * d);
}

struct EigOut {
  double lambda;
  KInfo info;
};

EigOut eigsolve_ac(imps_ctx* c, int site, double tol, int kdim, int maxiter, bool eager, int prec) {
  const size_t n = (size_t)c->D * c->d * c->D;
  EigOut o;
  site = wrap(c, site);
  // the environments stay fixed during the solve: split their planes once
  struct PlanesScope {
    imps_ctx* c;
    ~PlanesScope() { c->planes_site = -1; }
  } scope{c};
  if (prec == PREC_BF16X3 && hac_planes_ok(c, site)) {
    hac_planes_prepare(c, site);
    c->planes_site = site;
  }
  Op op = [&](const double* x, double* y) { hac_apply(c, site, x, y, prec); };
  o.lambda = lanczos(c, op, n, c->AC[site], c->pAC[site], tol, kdim, maxiter, eager, &o.info);
  return o;
}

// The mixed-precision refinement of eigsolve_ac / eigsolve_c: jd_refine with the FP64 operator for
// the residual and the `prec` operator for the corrections; FP64 Lanczos when it does not converge.
std::atomic<long> g_refine_fallbacks(0);

EigOut eigsolve_refined(imps_ctx* c, int site, bool ac, double tol, int kdim, int maxiter, bool eager, int prec) {
  site = wrap(c, site);
  const size_t n = ac ? (size_t)c->D * c->d * c->D : (size_t)c->D * c->D;
  struct PlanesScope {
    imps_ctx* c;
    ~PlanesScope() {
      c->planes_site = -1;
      c->hc_planes_site = -1;
    }
  } scope{c};
  if (prec == PREC_BF16X3 && ac && hac_planes_ok(c, site)) {
    hac_planes_prepare(c, site);
    c->planes_site = site;
  }
  if (prec == PREC_BF16X3 && !ac && hc_planes_ok(c, site)) {
    hc_planes_prepare(c, site);
    c->hc_planes_site = site;
  }
  Op op64 = ac ? Op([&](const double* x, double* y) { hac_apply(c, site, x, y, PREC_FP64); })
               : Op([&](const double* x, double* y) { hc_apply(c, site, x, y, PREC_FP64); });
  Op oplow = ac ? Op([&](const double* x, double* y) { hac_apply(c, site, x, y, prec); })
                : Op([&](const double* x, double* y) { hc_apply(c, site, x, y, prec); });
  const double* x0 = ac ? c->AC[site] : c->C[site];
  double* xout = ac ? c->pAC[site] : c->pC[site];
  EigOut o;
  if (jd_refine(c, op64, oplow, n, x0, xout, tol, 4, 2000, &o.lambda, &o.info)) return o;
  g_refine_fallbacks++;
  o.lambda = lanczos(c, op64, n, x0, xout, tol, kdim, maxiter, eager, &o.info);
  return o;
}

EigOut eigsolve_c(imps_ctx* c, int site, double tol, int kdim, int maxiter, bool eager, int prec) {
  const size_t n = (size_t)c->D * c->D;
  EigOut o;
  site = wrap(c, site);
  struct PlanesScope {
    imps_ctx* c;
    ~PlanesScope() { c->hc_planes_site = -1; }
  } scope{c};
  if (prec == PREC_BF16X3 && hc_planes_ok(c, site)) {
    hc_planes_prepare(c, site);
    c->hc_planes_site = site;
  }
  Op op = [&](const double* x, double* y) { hc_apply(c, site, x, y, prec); };
  o.lambda = lanczos(c, op, n, c->C[site], c->pC[site], tol, kdim, maxiter, eager, &o.info);
  return o;
}

// ------------------------------------------------------------------ concurrent local update

imps_ctx* worker(imps_ctx* c, int k) {
  while ((int)c->workers.size() <= k) {
    imps_ctx* w = new imps_ctx();
    w->device = c->device;
    c->workers.push_back(w);
    HIPCHK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc((void**)&w->hpart, sizeof(double) * kMaxDotBlocks * kMaxVec, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&w->hscal, sizeof(double) * kSmallPinned, hipHostMallocDefault));
    w->dpart = dalloc((size_t)kMaxDotBlocks * kMaxVec);
    HIPCHK(hipEventCreateWithFlags(&w->ev_coef, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&w->ev_done, hipEventDisableTiming));
  }
  imps_ctx* w = c->workers[k];
  w->N = c->N;
  w->w = c->w;
  w->d = c->d;
  w->D = c->D;
  w->Wm = c->Wm;
  w->AL = c->AL;
  w->AR = c->AR;
  w->AC = c->AC;
  w->C = c->C;
  w->pAC = c->pAC;
  w->pC = c->pC;
  w->pAL = c->pAL;
  w->GL = c->GL;
  w->GR = c->GR;
  w->gl_id = c->gl_id;
  w->gr_id = c->gr_id;
  w->envs_valid = c->envs_valid;
  w->cplx = c->cplx;
  w->chi = c->chi;
  w->pdim = c->pdim;
  return w;
}

// frees what a worker owns (never the borrowed model, state or environments)
void destroy_worker(imps_ctx* w) {
  if (w->stream) (void)hipStreamSynchronize(w->stream);
  clear_graphs(w);
  for (auto& kv : w->ws) dfree(kv.second.p);
  dfree(w->dpart);
  if (w->hpart) (void)hipHostFree(w->hpart);
  if (w->hscal) (void)hipHostFree(w->hscal);
  if (w->hstage) (void)hipHostFree(w->hstage);
  if (w->ev_coef) (void)hipEventDestroy(w->ev_coef);
  if (w->ev_done) (void)hipEventDestroy(w->ev_done);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
}

// localupdate_step! (vumps.jl:7-19): the AC and C eigensolves of all sites depend only on the
// current state and environments, so the 2N solves run concurrently, one host thread and one
// stream each (their device work overlaps; each solve's arithmetic is the same as when run
// alone).  Then regauge! per site on this context's stream.
// The GEMM classes of H_AC and H_C are autotuned on this context's stream, with nothing else
// running, before the concurrent eigensolves would tune them on contended worker streams (one
// application of each operator per site, once per state shape and precision).
void pretune_local_ops(imps_ctx* c, int prec) {
  if (g_gemm_autotune.load(std::memory_order_relaxed) == 0 || (prec != PREC_FP64 && prec != PREC_FP32)) return;
  const long key = (((long)c->D * 64 + c->w) * 64 + c->d) * 4 + prec;
  if (c->pretuned.count(key)) return;
  HIPCHK(hipStreamSynchronize(c->stream));
  const size_t nA = (size_t)c->D * c->d * c->D, nC = (size_t)c->D * c->D;
  double* y = ws(c, "pt_y", nA);
  for (int site = 0; site < c->N; ++site) {
    hac_apply(c, site, c->AC[site], y, prec);
    hc_apply(c, site, c->C[site], y, prec);
  }
  (void)nC;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->hac_applies -= c->N;
  c->hc_applies -= c->N;
  c->pretuned.insert(key);
}

// refine: the eigensolves are eigsolve_refined with `prec` as the correction operator (FP64 residuals)
void local_update(imps_ctx* c, const std::vector<double>& tol_ac, const std::vector<double>& tol_c, int kdim,
                  int maxiter, bool eager, int prec, std::vector<double>& lam_ac, std::vector<double>& lam_c,
                  bool refine = false) {
  const int N = c->N;
  pretune_local_ops(c, refine ? PREC_FP64 : prec);
  if (!c->ev_ready) HIPCHK(hipEventCreateWithFlags(&c->ev_ready, hipEventDisableTiming));
  HIPCHK(hipEventRecord(c->ev_ready, c->stream));
  const int njobs = 2 * N;
  std::vector<imps_ctx*> wk(njobs);
  for (int k = 0; k < njobs; ++k) {
    wk[k] = worker(c, k);
    HIPCHK(hipStreamWaitEvent(wk[k]->stream, c->ev_ready, 0));
  }
  std::vector<int> code(njobs, IMPS_OK);
  std::vector<std::string> msg(njobs);
  std::vector<std::thread> th;
  for (int k = 0; k < njobs; ++k) {
    th.emplace_back([&, k] {
      imps_ctx* w = wk[k];
      try {
        HIPCHK(hipSetDevice(w->device));
        const int site = k / 2;
        if (refine)
          (k % 2 == 0 ? lam_ac : lam_c)[site] =
              eigsolve_refined(w, site, k % 2 == 0, k % 2 == 0 ? tol_ac[site] : tol_c[site], kdim, maxiter, eager, prec)
                  .lambda;
        else if (k % 2 == 0)
          lam_ac[site] = eigsolve_ac(w, site, tol_ac[site], kdim, maxiter, eager, prec).lambda;
        else
          lam_c[site] = eigsolve_c(w, site, tol_c[site], kdim, maxiter, eager, prec).lambda;
        HIPCHK(hipEventRecord(w->ev_done, w->stream));
      } catch (const ImpsError& e) {
        code[k] = e.code;
        msg[k] = e.what();
      } catch (const std::exception& e) {
        code[k] = IMPS_E_ARG;
        msg[k] = e.what();
      }
    });
  }
  for (auto& t : th) t.join();
  for (int k = 0; k < njobs; ++k)
    if (code[k] != IMPS_OK) {
      // let the other solves' enqueued work finish before the state is touched again
      for (imps_ctx* w : wk) (void)hipStreamSynchronize(w->stream);
      throw ImpsError(code[k], msg[k]);
    }
  for (int k = 0; k < njobs; ++k) {
    HIPCHK(hipStreamWaitEvent(c->stream, wk[k]->ev_done, 0));
    c->hac_applies += wk[k]->hac_applies;
    c->hc_applies += wk[k]->hc_applies;
    wk[k]->hac_applies = wk[k]->hc_applies = 0;
  }
  for (int site = 0; site < N; ++site) regauge_site(c, site);
}

int run_guarded(imps_ctx* c, const std::function<void()>& f) {
  try {
    if (c) HIPCHK(hipSetDevice(c->device));
    f();
    return IMPS_OK;
  } catch (const ImpsError& e) {
    if (c) c->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return IMPS_E_ARG;
  }
}

}  // namespace

// =================================================================== C ABI

extern "C" {

int imps_create(int device, imps_ctx** out) {
  if (!out) return IMPS_E_ARG;
  *out = nullptr;
  imps_ctx* c = new imps_ctx();
  c->device = device;
  try {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    RBCHK(rocblas_create_handle(&c->rb));
    RBCHK(rocblas_set_stream(c->rb, c->stream));
    HIPCHK(hipHostMalloc((void**)&c->hpart, sizeof(double) * kMaxDotBlocks * kMaxVec, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&c->hscal, sizeof(double) * kSmallPinned, hipHostMallocDefault));
    c->dpart = dalloc((size_t)kMaxDotBlocks * kMaxVec);
    HIPCHK(hipEventCreateWithFlags(&c->ev_coef, hipEventDisableTiming));
    (void)scratch_ptr(device);  // this device's scratch address, for imps_hac_apply_noalloc
  } catch (const ImpsError& e) {
    g_create_error = e.what();
    if (c->rb) rocblas_destroy_handle(c->rb);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return e.code;
  }
  *out = c;
  return IMPS_OK;
}

int imps_destroy(imps_ctx* c) {
  if (!c) return IMPS_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (imps_ctx* w : c->workers) destroy_worker(w);
  c->workers.clear();
  if (c->ev_ready) (void)hipEventDestroy(c->ev_ready);
  free_state(c);
  for (double*& p : c->Wm) dfree(p);
  for (auto& kv : c->ws) dfree(kv.second.p);
  dfree(c->dpart);
  if (c->hpart) (void)hipHostFree(c->hpart);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->hstage) (void)hipHostFree(c->hstage);
  if (c->ev_coef) (void)hipEventDestroy(c->ev_coef);
  if (c->rb) rocblas_destroy_handle(c->rb);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return IMPS_OK;
}

const char* imps_last_error(const imps_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int imps_set_mpo(imps_ctx* c, int nsites, int w, int d, const double* const* W) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require(nsites > 0 && w >= 2 && d >= 1 && W, IMPS_E_ARG, "imps_set_mpo: bad dimensions");
    require(w * d <= 64, IMPS_E_UNSUPPORTED, "imps_set_mpo: w*d > 64 is not supported");
    require(d <= 8, IMPS_E_UNSUPPORTED, "imps_set_mpo: d > 8 is not supported");
    // a model with another unit cell or physical dimension invalidates the state
    if (!c->AL.empty() && ((int)c->AL.size() != nsites || c->d != d)) free_state(c);
    for (double*& p : c->Wm) dfree(p);
    c->Wm.clear();
    c->Wh.clear();
    c->N = nsites;
    c->w = w;
    c->d = d;
    const int wd = w * d;
    for (int l = 0; l < nsites; ++l) {
      require(W[l] != nullptr, IMPS_E_ARG, "imps_set_mpo: null tensor");
      std::vector<double> Wl(W[l], W[l] + (size_t)w * d * w * d);
      c->Wh.push_back(Wl);
      // unit corners
      for (int u = 0; u < d; ++u)
        for (int s = 0; s < d; ++s) {
          const double e = (u == s) ? 1.0 : 0.0;
          require(Wl[0 + (size_t)w * (u + (size_t)d * (0 + (size_t)w * s))] == e &&
                      Wl[(w - 1) + (size_t)w * (u + (size_t)d * ((w - 1) + (size_t)w * s))] == e,
                  IMPS_E_ARG, "imps_set_mpo: levels 0 and w-1 must be unit levels");
        }
      std::vector<double> Wm((size_t)wd * wd);
      for (int i = 0; i < w; ++i)
        for (int u = 0; u < d; ++u)
          for (int j = 0; j < w; ++j)
            for (int s = 0; s < d; ++s)
              Wm[(u + (size_t)d * j) + (size_t)wd * (i + (size_t)w * s)] =
                  Wl[i + (size_t)w * (u + (size_t)d * (j + (size_t)w * s))];
      double* dp = dalloc(Wm.size());
      h2d(c, dp, Wm.data(), Wm.size());
      c->Wm.push_back(dp);
    }
    if (c->envs_valid) {
      c->envs_valid = false;
    }
    for (double*& p : c->GL) dfree(p);
    for (double*& p : c->GR) dfree(p);
    c->GL.clear();
    c->GR.clear();
    c->gl_id.clear();
    c->gr_id.clear();
    c->pdim.assign(nsites, d);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

// Schur chains with per-site shapes: rectangular Schur tensors (levels that differ from bond to
// bond) and per-site physical dimensions, embedded exactly into the square form that every
// algorithm of the library runs on.  Bond level i < w_b - 1 keeps its index, the closing level
// w_b - 1 goes to w - 1 (w = max w_b), and physical indices s >= dims[l] form an empty sector: the
// unit levels act there as the identity, every other block is zero, so each effective operator
// maps states without weight in the empty sectors to states without weight there, and the QR/LQ
// gauge steps keep their zero rows and columns zero.
int imps_set_mpo_chain(imps_ctx* c, int nsites, const int* dims, const int* levels, const double* const* W) {
  if (!c) return IMPS_E_ARG;
  std::vector<std::vector<double>> sq;
  int wm = 0, dm = 0;
  const int code = run_guarded(c, [&] {
    require(nsites > 0 && dims && levels && W, IMPS_E_ARG, "imps_set_mpo_chain: bad arguments");
    for (int l = 0; l < nsites; ++l) {
      require(dims[l] >= 1 && levels[l] >= 2 && W[l] != nullptr, IMPS_E_ARG,
              "imps_set_mpo_chain: each site needs d >= 1, >= 2 levels on its bond and a tensor");
      wm = std::max(wm, levels[l]);
      dm = std::max(dm, dims[l]);
    }
    require(wm * dm <= 64 && dm <= 8, IMPS_E_UNSUPPORTED, "imps_set_mpo_chain: w*d > 64 or d > 8 is not supported");
    sq.assign(nsites, std::vector<double>((size_t)wm * dm * wm * dm, 0.0));
    for (int l = 0; l < nsites; ++l) {
      const int wl = levels[(l + nsites - 1) % nsites], wr = levels[l], d = dims[l];
      auto lev = [&](int i, int wb) { return i == wb - 1 ? wm - 1 : i; };
      double* P = sq[l].data();
      for (int i = 0; i < wl; ++i)
        for (int u = 0; u < d; ++u)
          for (int j = 0; j < wr; ++j)
            for (int t = 0; t < d; ++t)
              P[lev(i, wl) + (size_t)wm * (u + (size_t)dm * (lev(j, wr) + (size_t)wm * t))] =
                  W[l][i + (size_t)wl * (u + (size_t)d * (j + (size_t)wr * t))];
      for (int u = d; u < dm; ++u) {
        P[0 + (size_t)wm * (u + (size_t)dm * (0 + (size_t)wm * u))] = 1.0;
        P[(wm - 1) + (size_t)wm * (u + (size_t)dm * ((wm - 1) + (size_t)wm * u))] = 1.0;
      }
    }
  });
  if (code != IMPS_OK) return code;
  std::vector<const double*> pp(nsites);
  for (int l = 0; l < nsites; ++l) pp[l] = sq[l].data();
  // the square chain goes through imps_set_mpo's checks (the unit corners included)
  const int rc = imps_set_mpo(c, nsites, wm, dm, pp.data());
  if (rc == IMPS_OK) c->pdim.assign(dims, dims + nsites);
  return rc;
}

// every state tensor (D, d, D) must vanish on the empty physical sectors of its site
static void require_empty_sectors(imps_ctx* c, int D, const double* const* A, const char* what) {
  if (!A) return;
  for (int l = 0; l < c->N; ++l) {
    if (!A[l] || c->pdim.size() != (size_t)c->N || c->pdim[l] >= c->d) continue;
    for (int s = c->pdim[l]; s < c->d; ++s)
      for (int b = 0; b < D; ++b)
        for (int a = 0; a < D; ++a)
          require(A[l][a + (size_t)D * (s + (size_t)c->d * b)] == 0.0, IMPS_E_ARG,
                  std::string(what) + ": a tensor has weight in an empty physical sector of site " +
                      std::to_string(l));
  }
}

int imps_set_state(imps_ctx* c, int nsites, int D, const double* const* AL, const double* const* AR,
                   const double* const* Cm, const double* const* AC) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_mpo(c);
    require(nsites == c->N && D >= 1, IMPS_E_ARG, "imps_set_state: unit cell or D mismatch");
    require(D <= c->d * D, IMPS_E_ARG, "imps_set_state: D exceeds d*D");
    require_empty_sectors(c, D, AL, "imps_set_state");
    require_empty_sectors(c, D, AR, "imps_set_state");
    require_empty_sectors(c, D, AC, "imps_set_state");
    alloc_state(c, nsites, D, c->d);
    const size_t nA = (size_t)D * c->d * D, nC = (size_t)D * D;
    for (int l = 0; l < nsites; ++l) {
      h2d(c, c->AL[l], AL[l], nA);
      h2d(c, c->AR[l], AR[l], nA);
      h2d(c, c->C[l], Cm[l], nC);
      if (AC)
        h2d(c, c->AC[l], AC[l], nA);
      else
        mul_AL_C(c, c->AL[l], c->C[l], c->AC[l]);
    }
    init_pending(c);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int imps_state_dims(imps_ctx* c, int* nsites, int* D, int* d) {
  if (!c || !nsites || !D || !d) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    *nsites = (int)c->AL.size();
    *D = c->cplx ? c->D / 2 : c->D;
    *d = c->d;
  });
}

int imps_get_state(imps_ctx* c, double* const* AL, double* const* AR, double* const* Cm,
                   double* const* AC) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_get_state");
    need_state(c);
    const int N = (int)c->AL.size();
    const size_t nA = (size_t)c->D * c->d * c->D, nC = (size_t)c->D * c->D;
    for (int l = 0; l < N; ++l) {
      if (AL && AL[l]) d2h(c, AL[l], c->AL[l], nA);
      if (AR && AR[l]) d2h(c, AR[l], c->AR[l], nA);
      if (Cm && Cm[l]) d2h(c, Cm[l], c->C[l], nC);
      if (AC && AC[l]) d2h(c, AC[l], c->AC[l], nA);
    }
  });
}

// Active bond dimensions of a state held at the largest one (D): 1 <= chi[l] <= D with some
// chi[l] == D, and each site's tensor can be full rank (chi[l] <= d_l chi[l-1] and
// chi[l-1] <= d_l chi[l]); tensors must vanish outside their active blocks.
static void check_bond_profile(imps_ctx* c, int nsites, int D, const int* chi, const char* what) {
  require(chi != nullptr && nsites == c->N, IMPS_E_ARG, std::string(what) + ": unit cell mismatch");
  int mx = 0;
  for (int l = 0; l < nsites; ++l) {
    require(chi[l] >= 1 && chi[l] <= D, IMPS_E_ARG, std::string(what) + ": bond dimension out of range");
    mx = std::max(mx, chi[l]);
  }
  require(mx == D, IMPS_E_ARG, std::string(what) + ": D must be the largest bond dimension");
  for (int l = 0; l < nsites; ++l) {
    const int dl = c->pdim.size() == (size_t)c->N ? c->pdim[l] : c->d;
    const int cl = chi[(l - 1 + nsites) % nsites], cr = chi[l];
    require(cr <= dl * cl && cl <= dl * cr, IMPS_E_ARG,
            std::string(what) + ": site " + std::to_string(l) + " cannot be full rank (bond dimensions " +
                std::to_string(cl) + ", " + std::to_string(cr) + ")");
  }
}

static void require_outside_zero(imps_ctx* c, int D, const int* chi, const double* const* A, bool mps,
                                 const char* what) {
  if (!A) return;
  const int N = c->N, m = mps ? c->d : 1;
  for (int l = 0; l < N; ++l) {
    if (!A[l]) continue;
    const int ra = mps ? chi[(l - 1 + N) % N] : chi[l], ca = chi[l];
    for (int b = 0; b < D; ++b)
      for (int k = 0; k < m; ++k)
        for (int a = 0; a < D; ++a)
          if (a >= ra || b >= ca)
            require(A[l][a + (size_t)D * (k + (size_t)m * b)] == 0.0, IMPS_E_ARG,
                    std::string(what) + ": site " + std::to_string(l) +
                        " has weight outside its bond dimensions");
  }
}

static void set_state_raw_impl(imps_ctx* c, int nsites, int D, const double* const* A, double tol, int maxiter,
                               const int* chi) {
  need_mpo(c);
  require(nsites == c->N && D >= 1 && A, IMPS_E_ARG, "imps_set_state_raw: unit cell mismatch");
  require_empty_sectors(c, D, A, "imps_set_state_raw");
  alloc_state(c, nsites, D, c->d);
  if (chi) c->chi.assign(chi, chi + nsites);
  const size_t nA = (size_t)D * c->d * D, nC = (size_t)D * D;
  // raw tensors are staged in the pending AC buffers (the gauge sweeps read copies)
  std::vector<double*> raw(nsites);
  for (int l = 0; l < nsites; ++l) {
    raw[l] = c->pAC[l];
    h2d(c, raw[l], A[l], nA);
  }
  double* I = ws(c, "sr_I", nC);
  HIPCHK(launch_set_identity(D, I, D, c->stream));
  if (chi) HIPCHK(launch_mask_block(D, D, chi[nsites - 1], chi[nsites - 1], I, D, c->stream));
  // gaugefix!(:LR): left sweeps from C0 = I, then right sweeps from C[end]
  uniform_leftorth(c, raw, I, tol, maxiter);
  double* C0 = ws(c, "sr_C0", nC);
  d2d(c, C0, c->C[nsites - 1], nC);
  uniform_rightorth(c, c->AL, C0, tol, maxiter);
  for (int l = 0; l < nsites; ++l) mul_AL_C(c, c->AL[l], c->C[l], c->AC[l]);
  init_pending(c);
  c->envs_valid = false;
  HIPCHK(hipStreamSynchronize(c->stream));
}

int imps_set_state_raw(imps_ctx* c, int nsites, int D, const double* const* A, double tol, int maxiter) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] { set_state_raw_impl(c, nsites, D, A, tol, maxiter, nullptr); });
}

int imps_set_state_nonuniform(imps_ctx* c, int nsites, int D, const int* chi, const double* const* AL,
                              const double* const* AR, const double* const* Cm, const double* const* AC) {
  if (!c || !chi || !AL || !AR || !Cm) return IMPS_E_ARG;
  int st = run_guarded(c, [&] {
    need_mpo(c);
    check_bond_profile(c, nsites, D, chi, "imps_set_state_nonuniform");
    require_outside_zero(c, D, chi, AL, true, "imps_set_state_nonuniform");
    require_outside_zero(c, D, chi, AR, true, "imps_set_state_nonuniform");
    require_outside_zero(c, D, chi, AC, true, "imps_set_state_nonuniform");
    require_outside_zero(c, D, chi, Cm, false, "imps_set_state_nonuniform");
  });
  if (st != IMPS_OK) return st;
  st = imps_set_state(c, nsites, D, AL, AR, Cm, AC);
  if (st != IMPS_OK) return st;
  return run_guarded(c, [&] { c->chi.assign(chi, chi + nsites); });
}

int imps_set_state_raw_nonuniform(imps_ctx* c, int nsites, int D, const int* chi, const double* const* A, double tol,
                                  int maxiter) {
  if (!c || !chi || !A) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_mpo(c);
    check_bond_profile(c, nsites, D, chi, "imps_set_state_raw_nonuniform");
    require_outside_zero(c, D, chi, A, true, "imps_set_state_raw_nonuniform");
    set_state_raw_impl(c, nsites, D, A, tol, maxiter, chi);
  });
}

int imps_bond_dims(imps_ctx* c, int* chi) {
  if (!c || !chi) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    const int N = (int)c->AL.size();
    const int f = c->cplx ? 2 : 1;  // a complex state's profile is held on its embedding
    for (int l = 0; l < N; ++l) chi[l] = (c->chi.empty() ? c->D : c->chi[l]) / f;
  });
}

// ---- complex states: the real embedding (2 D, d, 2 D) of the caller's ComplexF64 tensors
static void require_empty_sectors_c(imps_ctx* c, int D, const double* const* A, const char* what) {
  if (!A) return;
  for (int l = 0; l < c->N; ++l) {
    if (!A[l] || c->pdim.size() != (size_t)c->N || c->pdim[l] >= c->d) continue;
    for (int s = c->pdim[l]; s < c->d; ++s)
      for (int b = 0; b < D; ++b)
        for (int a = 0; a < D; ++a) {
          const size_t k = 2 * (a + (size_t)D * (s + (size_t)c->d * b));
          require(A[l][k] == 0.0 && A[l][k + 1] == 0.0, IMPS_E_ARG,
                  std::string(what) + ": a tensor has weight in an empty physical sector of site " +
                      std::to_string(l));
        }
  }
}

// host complex (D1, m, D2) -> embedded device tensor
static void embed_from_host(imps_ctx* c, const double* Z, int D1, int m, int D2, double scale, double* R) {
  const size_t n2 = (size_t)2 * D1 * m * D2;
  double* stagez = ws(c, "cx_stage", n2);
  h2d(c, stagez, Z, n2);
  HIPCHK(launch_cplx_embed(D1, m, D2, stagez, scale, R, c->stream));
}

// embedded device tensor -> host complex (D1, m, D2)
static void extract_to_host(imps_ctx* c, double* R, int D1, int m, int D2, double scale, double* Z) {
  const size_t n2 = (size_t)2 * D1 * m * D2;
  double* stagez = ws(c, "cx_stage", n2);
  HIPCHK(launch_cplx_extract(D1, m, D2, R, scale, stagez, false, c->stream));
  d2h(c, Z, stagez, n2);
}

// projects every state tensor back onto the embedded form (removes rounding drift out of it)
static void project_state_c(imps_ctx* c) {
  const int Dc = c->D / 2, d = c->d;
  for (size_t l = 0; l < c->AL.size(); ++l) {
    HIPCHK(launch_cplx_extract(Dc, d, Dc, c->AL[l], 1.0, nullptr, true, c->stream));
    HIPCHK(launch_cplx_extract(Dc, d, Dc, c->AR[l], 1.0, nullptr, true, c->stream));
    HIPCHK(launch_cplx_extract(Dc, d, Dc, c->AC[l], 1.0, nullptr, true, c->stream));
    HIPCHK(launch_cplx_extract(Dc, 1, Dc, c->C[l], 1.0, nullptr, true, c->stream));
  }
}

int imps_set_state_complex(imps_ctx* c, int nsites, int D, const double* const* AL, const double* const* AR,
                           const double* const* Cm, const double* const* AC) {
  if (!c || !AL || !AR || !Cm) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_mpo(c);
    require(nsites == c->N && D >= 1, IMPS_E_ARG, "imps_set_state_complex: unit cell or D mismatch");
    for (int l = 0; l < nsites; ++l)
      require(AL[l] && AR[l] && Cm[l] && (!AC || AC[l]), IMPS_E_ARG, "imps_set_state_complex: null tensor");
    require_empty_sectors_c(c, D, AL, "imps_set_state_complex");
    require_empty_sectors_c(c, D, AR, "imps_set_state_complex");
    require_empty_sectors_c(c, D, AC, "imps_set_state_complex");
    alloc_state(c, nsites, 2 * D, c->d);
    c->cplx = true;
    const double h = 1.0 / std::sqrt(2.0);
    for (int l = 0; l < nsites; ++l) {
      embed_from_host(c, AL[l], D, c->d, D, 1.0, c->AL[l]);
      embed_from_host(c, AR[l], D, c->d, D, 1.0, c->AR[l]);
      embed_from_host(c, Cm[l], D, 1, D, h, c->C[l]);
      if (AC)
        embed_from_host(c, AC[l], D, c->d, D, h, c->AC[l]);
      else
        mul_AL_C(c, c->AL[l], c->C[l], c->AC[l]);
    }
    init_pending(c);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

// the embedding of a complex bond profile: complex index a occupies rows 2a and 2a + 1, so a
// complex active block of chi indices is the real active block of 2 chi
static void set_profile_c(imps_ctx* c, const int* chi, int nsites) {
  c->chi.resize(nsites);
  for (int l = 0; l < nsites; ++l) c->chi[l] = 2 * chi[l];
}

static void set_state_raw_complex_impl(imps_ctx* c, int nsites, int D, const double* const* A, double tol,
                                       int maxiter, const int* chi) {
  need_mpo(c);
  require(nsites == c->N && D >= 1 && A, IMPS_E_ARG, "imps_set_state_raw_complex: unit cell mismatch");
  for (int l = 0; l < nsites; ++l) require(A[l] != nullptr, IMPS_E_ARG, "imps_set_state_raw_complex: null tensor");
  require_empty_sectors_c(c, D, A, "imps_set_state_raw_complex");
  const int D2 = 2 * D;
  alloc_state(c, nsites, D2, c->d);
  c->cplx = true;
  if (chi) set_profile_c(c, chi, nsites);
  const size_t nC = (size_t)D2 * D2;
  std::vector<double*> raw(nsites);
  for (int l = 0; l < nsites; ++l) {
    raw[l] = c->pAC[l];
    embed_from_host(c, A[l], D, c->d, D, 1.0, raw[l]);
  }
  double* I = ws(c, "sr_I", nC);
  HIPCHK(launch_set_identity(D2, I, D2, c->stream));
  if (chi) HIPCHK(launch_mask_block(D2, D2, c->chi[nsites - 1], c->chi[nsites - 1], I, D2, c->stream));
  uniform_leftorth(c, raw, I, tol, maxiter);
  double* C0 = ws(c, "sr_C0", nC);
  d2d(c, C0, c->C[nsites - 1], nC);
  uniform_rightorth(c, c->AL, C0, tol, maxiter);
  for (int l = 0; l < nsites; ++l) mul_AL_C(c, c->AL[l], c->C[l], c->AC[l]);
  project_state_c(c);
  init_pending(c);
  c->envs_valid = false;
  HIPCHK(hipStreamSynchronize(c->stream));
}

int imps_set_state_raw_complex(imps_ctx* c, int nsites, int D, const double* const* A, double tol, int maxiter) {
  if (!c || !A) return IMPS_E_ARG;
  return run_guarded(c, [&] { set_state_raw_complex_impl(c, nsites, D, A, tol, maxiter, nullptr); });
}

// the complex twin of require_outside_zero (interleaved re, im)
static void require_outside_zero_c(imps_ctx* c, int D, const int* chi, const double* const* A, bool mps,
                                   const char* what) {
  if (!A) return;
  const int N = c->N, m = mps ? c->d : 1;
  for (int l = 0; l < N; ++l) {
    if (!A[l]) continue;
    const int ra = mps ? chi[(l - 1 + N) % N] : chi[l], ca = chi[l];
    for (int b = 0; b < D; ++b)
      for (int k = 0; k < m; ++k)
        for (int a = 0; a < D; ++a)
          if (a >= ra || b >= ca) {
            const size_t e = 2 * (a + (size_t)D * (k + (size_t)m * b));
            require(A[l][e] == 0.0 && A[l][e + 1] == 0.0, IMPS_E_ARG,
                    std::string(what) + ": site " + std::to_string(l) + " has weight outside its bond dimensions");
          }
  }
}

int imps_set_state_complex_nonuniform(imps_ctx* c, int nsites, int D, const int* chi, const double* const* AL,
                                      const double* const* AR, const double* const* Cm, const double* const* AC) {
  if (!c || !chi || !AL || !AR || !Cm) return IMPS_E_ARG;
  int st = run_guarded(c, [&] {
    need_mpo(c);
    check_bond_profile(c, nsites, D, chi, "imps_set_state_complex_nonuniform");
    for (int l = 0; l < nsites; ++l)
      require(AL[l] && AR[l] && Cm[l] && (!AC || AC[l]), IMPS_E_ARG, "imps_set_state_complex_nonuniform: null tensor");
    require_outside_zero_c(c, D, chi, AL, true, "imps_set_state_complex_nonuniform");
    require_outside_zero_c(c, D, chi, AR, true, "imps_set_state_complex_nonuniform");
    require_outside_zero_c(c, D, chi, AC, true, "imps_set_state_complex_nonuniform");
    require_outside_zero_c(c, D, chi, Cm, false, "imps_set_state_complex_nonuniform");
  });
  if (st != IMPS_OK) return st;
  st = imps_set_state_complex(c, nsites, D, AL, AR, Cm, AC);
  if (st != IMPS_OK) return st;
  return run_guarded(c, [&] { set_profile_c(c, chi, nsites); });
}

int imps_set_state_raw_complex_nonuniform(imps_ctx* c, int nsites, int D, const int* chi, const double* const* A,
                                          double tol, int maxiter) {
  if (!c || !chi || !A) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_mpo(c);
    check_bond_profile(c, nsites, D, chi, "imps_set_state_raw_complex_nonuniform");
    for (int l = 0; l < nsites; ++l)
      require(A[l] != nullptr, IMPS_E_ARG, "imps_set_state_raw_complex_nonuniform: null tensor");
    require_outside_zero_c(c, D, chi, A, true, "imps_set_state_raw_complex_nonuniform");
    set_state_raw_complex_impl(c, nsites, D, A, tol, maxiter, chi);
  });
}

int imps_get_state_complex(imps_ctx* c, double* const* AL, double* const* AR, double* const* Cm,
                           double* const* AC) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    require(c->cplx, IMPS_E_STATE, "imps_get_state_complex: the state is real (imps_get_state)");
    const int N = (int)c->AL.size(), Dc = c->D / 2, d = c->d;
    const double s2 = std::sqrt(2.0);
    for (int l = 0; l < N; ++l) {
      if (AL && AL[l]) extract_to_host(c, c->AL[l], Dc, d, Dc, 1.0, AL[l]);
      if (AR && AR[l]) extract_to_host(c, c->AR[l], Dc, d, Dc, 1.0, AR[l]);
      if (Cm && Cm[l]) extract_to_host(c, c->C[l], Dc, 1, Dc, s2, Cm[l]);
      if (AC && AC[l]) extract_to_host(c, c->AC[l], Dc, d, Dc, s2, AC[l]);
    }
  });
}

int imps_state_is_complex(imps_ctx* c, int* flag) {
  if (!c || !flag) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    *flag = c->cplx ? 1 : 0;
  });
}

int imps_hac_apply(imps_ctx* c, int site, const double* x, double* y, int prec) {
  if (!c || !x || !y) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    need_envs(c);
    const size_t n = (size_t)c->D * c->d * c->D;
    double* dx = ws(c, "abi_x", n);
    double* dy = ws(c, "abi_y", n);
    if (c->cplx) {  // ComplexF64 x and y of the complex bond dimension
      const int Dc = c->D / 2;
      embed_from_host(c, x, Dc, c->d, Dc, 1.0, dx);
      hac_apply(c, site, dx, dy, prec);
      extract_to_host(c, dy, Dc, c->d, Dc, 1.0, y);
      return;
    }
    h2d(c, dx, x, n);
    hac_apply(c, site, dx, dy, prec);
    d2h(c, y, dy, n);
  });
}

int imps_hac_apply_dev(imps_ctx* c, int site, const double* d_x, double* d_y, int prec) {
  if (!c || !d_x || !d_y) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_hac_apply_dev");
    need_match(c);
    need_envs(c);
    hac_apply(c, site, d_x, d_y, prec);
  });
}

int imps_hac_apply_noalloc(imps_ctx* c, int site, const double* d_x, double* d_y, int prec) {
  if (!c || !d_x || !d_y) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_hac_apply_noalloc");
    need_match(c);
    need_envs(c);
    require(prec == PREC_FP64 || prec == PREC_FP32, IMPS_E_UNSUPPORTED,
            "imps_hac_apply_noalloc: FP64 and FP32 only (the split-BF16 path needs per-call plane buffers)");
    site = wrap(c, site);
    const int D = c->D, w = c->w, d = c->d;
    const int lv = (w >= 2 && c->gl_id[site] && c->gr_id[site]) ? 1 : 0;
    // T1 and T2 rounded up to 256 bytes, so T2 and the workspace keep the buffers' alignment
    const size_t nT = (((size_t)D * w * d * D) + 31) & ~(size_t)31;
    const size_t nW = std::max(gemm_workspace(D * (w - lv), d * D, D), gemm_workspace(D * d, D, D * (w - lv)));
    const size_t need = 2 * nT + nW;
    require(need <= kScratchDoubles, IMPS_E_ARG,
            "imps_hac_apply_noalloc: needs " + std::to_string(need) + " doubles of scratch (2 D w d D plus the "
            "split-K workspace); the maximum is " + std::to_string(kScratchDoubles) + " (128 MiB)");
    double* S = scratch_ptr(c->device);
    hac_apply_bufs(c, site, d_x, d_y, prec, HacBufs{S, S + nT, nW ? S + 2 * nT : nullptr, true}, nullptr);
  });
}

int imps_hc_apply(imps_ctx* c, int site, const double* x, double* y, int prec) {
  if (!c || !x || !y) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    need_envs(c);
    const size_t n = (size_t)c->D * c->D;
    double* dx = ws(c, "abi_x", n);
    double* dy = ws(c, "abi_y", n);
    if (c->cplx) {
      const int Dc = c->D / 2;
      embed_from_host(c, x, Dc, 1, Dc, 1.0, dx);
      hc_apply(c, site, dx, dy, prec);
      extract_to_host(c, dy, Dc, 1, Dc, 1.0, y);
      return;
    }
    h2d(c, dx, x, n);
    hc_apply(c, site, dx, dy, prec);
    d2h(c, y, dy, n);
  });
}

int imps_hc_apply_dev(imps_ctx* c, int site, const double* d_x, double* d_y, int prec) {
  if (!c || !d_x || !d_y) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_hc_apply_dev");
    need_match(c);
    need_envs(c);
    hc_apply(c, site, d_x, d_y, prec);
  });
}

int imps_transfer_apply(imps_ctx* c, int side, const double* v, double* out) {
  if (!c || !v || !out) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    require(side == IMPS_SIDE_LEFT || side == IMPS_SIDE_RIGHT, IMPS_E_ARG, "imps_transfer_apply: bad side");
    const int N = (int)c->AL.size(), D = c->D;
    const size_t DD = (size_t)D * D;
    double* a = ws(c, "tr_a", DD);
    double* b = ws(c, "tr_b", DD);
    if (c->cplx)
      embed_from_host(c, v, D / 2, 1, D / 2, 1.0, a);
    else
      h2d(c, a, v, DD);
    if (side == IMPS_SIDE_LEFT) {
      for (int l = 0; l < N; ++l) {
        push_left2(c, a, D, c->AL[l], c->AL[l], b, D);
        std::swap(a, b);
      }
    } else {
      for (int l = N - 1; l >= 0; --l) {
        push_right2(c, a, D, c->AR[l], c->AR[l], b, D);
        std::swap(a, b);
      }
    }
    if (c->cplx)
      extract_to_host(c, a, D / 2, 1, D / 2, 1.0, out);
    else
      d2h(c, out, a, DD);
  });
}

int imps_regularize(imps_ctx* c, int n, double* v, const double* lvec, const double* rvec) {
  if (!c || !v || !lvec || !rvec || n <= 0) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    const size_t nn = (size_t)n * n;
    double* dv = ws(c, "rg_v", nn);
    double* dl = ws(c, "rg_l", nn);
    double* dr = ws(c, "rg_r", nn);
    h2d(c, dv, v, nn);
    h2d(c, dl, lvec, nn);
    h2d(c, dr, rvec, nn);
    const double cc = trace_prod(c, n, dl, dv);
    axpby(c, nn, -cc, dr, 1.0, dv);
    d2h(c, v, dv, nn);
  });
}

int imps_envs_recalculate(imps_ctx* c, double tol, int maxiter, int krylovdim) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require(tol > 0 && maxiter > 0 && krylovdim > 0, IMPS_E_ARG, "imps_envs_recalculate: bad parameters");
    envs_recalculate(c, tol, maxiter, krylovdim, nullptr);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int imps_envs_reset(imps_ctx* c) {
  if (!c) return IMPS_E_ARG;
  c->envs_valid = false;
  return IMPS_OK;
}

int imps_get_env(imps_ctx* c, int site, int side, double* out) {
  if (!c || !out) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_envs(c);
    site = wrap(c, site);
    const size_t n = (size_t)c->D * c->w * c->D;
    double* G = side == IMPS_SIDE_LEFT ? c->GL[site] : c->GR[site];
    if (c->cplx) {  // ComplexF64 (D/2, w, D/2): the nearest complex tensor of each 2 x 2 block
      extract_to_host(c, G, c->D / 2, c->w, c->D / 2, 1.0, out);
      return;
    }
    d2h(c, out, G, n);
  });
}

int imps_set_env(imps_ctx* c, int site, int side, const double* in) {
  if (!c || !in) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_set_env");
    need_match(c);
    alloc_envs(c);
    site = wrap(c, site);
    const int D = c->D, w = c->w;
    const size_t n = (size_t)D * w * D;
    h2d(c, side == IMPS_SIDE_LEFT ? c->GL[site] : c->GR[site], in, n);
    // is the structural level (GL level 0, GR level w-1) exactly the identity?
    const int k = side == IMPS_SIDE_LEFT ? 0 : w - 1;
    bool eye = true;
    for (int b = 0; b < D && eye; ++b)
      for (int a = 0; a < D; ++a)
        if (in[a + (size_t)D * (k + (size_t)w * b)] != (a == b ? 1.0 : 0.0)) {
          eye = false;
          break;
        }
    (side == IMPS_SIDE_LEFT ? c->gl_id : c->gr_id)[site] = eye ? 1 : 0;
    c->envs_valid = true;
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

static void fill_info(imps_krylov_info* info, const KInfo& k) {
  if (!info) return;
  info->converged = k.converged;
  info->numiter = k.numiter;
  info->numops = k.numops;
  info->normres = k.normres;
}

int imps_eigsolve_ac(imps_ctx* c, int site, const imps_krylov_params* p, double* lambda, double* vec,
                     imps_krylov_info* info) {
  if (!c || !p) return IMPS_E_ARG;
  return run_guarded(c, [&] {

    need_match(c);
    need_envs(c);
    site = wrap(c, site);
    EigOut o = eigsolve_ac(c, site, p->tol, p->krylovdim, p->maxiter, p->eager != 0, p->prec);
    if (lambda) *lambda = o.lambda;
    fill_info(info, o.info);
    if (vec && c->cplx) {  // unit complex norm (the embedding has twice the squared norm)
      extract_to_host(c, c->pAC[site], c->D / 2, c->d, c->D / 2, std::sqrt(2.0), vec);
    } else if (vec) {
      d2h(c, vec, c->pAC[site], (size_t)c->D * c->d * c->D);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int imps_eigsolve_c(imps_ctx* c, int site, const imps_krylov_params* p, double* lambda, double* vec,
                    imps_krylov_info* info) {
  if (!c || !p) return IMPS_E_ARG;
  return run_guarded(c, [&] {

    need_match(c);
    need_envs(c);
    site = wrap(c, site);
    EigOut o = eigsolve_c(c, site, p->tol, p->krylovdim, p->maxiter, p->eager != 0, p->prec);
    if (lambda) *lambda = o.lambda;
    fill_info(info, o.info);
    if (vec && c->cplx) {
      extract_to_host(c, c->pC[site], c->D / 2, 1, c->D / 2, std::sqrt(2.0), vec);
    } else if (vec) {
      d2h(c, vec, c->pC[site], (size_t)c->D * c->D);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int imps_regauge(imps_ctx* c, int site, double* al) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_regauge");
    need_state(c);
    site = wrap(c, site);
    regauge_site(c, site);
    if (al) d2h(c, al, c->pAL[site], (size_t)c->D * c->d * c->D);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int imps_gaugefix_right(imps_ctx* c, double tol, int maxiter, int* sweeps) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    gauge_step(c, tol, maxiter, sweeps);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int imps_galerkin(imps_ctx* c, double* eps) {
  if (!c || !eps) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    *eps = galerkin(c, PREC_FP64);
  });
}

int imps_energy(imps_ctx* c, double* e) {
  if (!c || !e) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    *e = energy(c);
  });
}

int imps_energy_gradient(imps_ctx* c, double* const* dW) {
  if (!c || !dW) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_energy_gradient");
    need_match(c);
    need_envs(c);
    require(c->w * c->d <= 32, IMPS_E_UNSUPPORTED, "imps_energy_gradient: w*d > 32 is not supported");
    for (int site = 0; site < c->N; ++site) energy_gradient(c, site, dW[site]);
  });
}

int imps_fixedpoint_gradient(imps_ctx* c, int wo, int d, const double* Wobj, double tol, int maxiter, double rcond,
                             double* value, double* dW, imps_response_info* info) {
  if (!c || !Wobj || !dW) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_fixedpoint_gradient");
    need_match(c);
    need_envs(c);
    require(c->N == 1, IMPS_E_UNSUPPORTED, "imps_fixedpoint_gradient: one-site unit cell only");
    require(d == c->d, IMPS_E_ARG, "imps_fixedpoint_gradient: objective's physical dimension differs from the model's");
    require(wo >= 2 && wo * d <= 64, IMPS_E_ARG, "imps_fixedpoint_gradient: objective level count out of range");
    require(c->w * c->d <= 32, IMPS_E_UNSUPPORTED, "imps_fixedpoint_gradient: w*d > 32 is not supported");
    require(tol > 0.0 && maxiter >= 1 && rcond >= 0.0 && rcond < 1.0, IMPS_E_ARG,
            "imps_fixedpoint_gradient: tol > 0, maxiter >= 1 and 0 <= rcond < 1 required");
    for (int u = 0; u < d; ++u)
      for (int s = 0; s < d; ++s) {
        const double e = u == s ? 1.0 : 0.0;
        require(Wobj[(size_t)wo * (u + (size_t)d * ((size_t)wo * s))] == e &&
                    Wobj[(wo - 1) + (size_t)wo * (u + (size_t)d * ((wo - 1) + (size_t)wo * s))] == e,
                IMPS_E_ARG, "imps_fixedpoint_gradient: levels 0 and w-1 of the objective must be unit levels");
      }
    const int D = c->D;
    const size_t n = (size_t)D * d * D;
    double* g = ws(c, "fp_g", n);
    double val = 0.0;
    {
      // the objective's environments on the same state, its value and tangent gradient
      ModelSwap sw(c, wo, Wobj);
      EnvStats est;
      envs_build(c, 1e-14, 300, 30, &est);
      val = energy(c);
      hac_apply(c, 0, c->AC[0], g, PREC_FP64);
      tangent_project(c, c->AL[0], g);
    }
    Tangent t = tangent_setup(c, rcond, std::min(1e-13, 1e-2 * tol));
    double* Lm = ws(c, "fp_L", n);
    // P (H_eff + K) P + (1 - P): the complement of the tangent space (the Hessian's null space,
    // where g has no component) gets the eigenvalue 1, which keeps MINRES's recurrence estimate
    // tied to the true residual (on the singular form rounding components along the null space
    // detach the two)
    double* q = ws(c, "fp_q", (size_t)D * D);
    Op op = [&](const double* v, double* o) {
      tangent_hessian(c, t, v, o);
      gemm2(c, 'T', 'N', D, D, d * D, 1.0, t.AL, (long)d * D, v, (long)d * D, 0.0, q, D);
      gemm2(c, 'N', 'N', d * D, D, D, 1.0, t.AL, (long)d * D, q, D, 1.0, o, (long)d * D);
    };
    double estimate = 0.0;
    const int its = minres_dev(c, op, n, g, Lm, tol, maxiter, 1e-14, &estimate);
    double* res = ws(c, "fp_res", n);
    tangent_hessian(c, t, Lm, res);
    axpby(c, n, -1.0, g, 1.0, res);
    const double ng = nrm2(c, n, g);
    const double relres = nrm2(c, n, res) / std::max(ng, 1e-300);
    std::vector<double> S = tangent_w_sandwich(c, t, Lm);
    for (size_t k = 0; k < S.size(); ++k) dW[k] = -2.0 * S[k];
    if (value) *value = val;
    if (info) {
      info->iterations = its;
      info->converged = (estimate <= tol || ng <= 1e-14) ? 1 : 0;
      info->relres = relres;
      info->estimate = estimate;
      info->hessian_applies = t.hessian_applies;
      info->level_solves = t.solves;
      info->level_ops = t.ops;
      info->unconverged_solves = t.unconverged;
    }
  });
}

int imps_tangent_hessian_apply(imps_ctx* c, double rcond, double tol, const double* B, double* out) {
  if (!c || !B || !out) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_tangent_hessian_apply");
    need_match(c);
    need_envs(c);
    require(c->N == 1, IMPS_E_UNSUPPORTED, "imps_tangent_hessian_apply: one-site unit cell only");
    require(tol > 0.0 && rcond >= 0.0 && rcond < 1.0, IMPS_E_ARG,
            "imps_tangent_hessian_apply: tol > 0 and 0 <= rcond < 1 required");
    const size_t n = (size_t)c->D * c->d * c->D;
    Tangent t = tangent_setup(c, rcond, tol);
    double* x = ws(c, "fp_L", n);
    double* y = ws(c, "fp_res", n);
    h2d(c, x, B, n);
    tangent_hessian(c, t, x, y);
    d2h(c, out, y, n);
  });
}

int imps_local_expectation(imps_ctx* c, int site, const double* O, double* val) {
  if (!c || !O || !val) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    require_symmetric_if_complex(c, O, "imps_local_expectation");
    site = wrap(c, site);
    const size_t n = (size_t)c->D * c->d * c->D;
    double* t = ws(c, "lx_t", n);
    HIPCHK(launch_mix_mid(c->D, c->d, c->D, O, c->AC[site], t, 0.0, c->stream));
    *val = dot(c, n, c->AC[site], t) / dot_psi(c);
  });
}

// <O> for a complex d x d operator O = A + i B (A = Re O, B = Im O, real): <O> = z_A + i z_B with
// z_R = <psi|R|psi>.  On a real state z_R is real.  On a complex state the real trace of the
// embedding gives Re z_R, and Im z_R = Re <psi|R|(-i) psi>, with -i psi the row-pair rotation of AC.
int imps_local_expectation_complex(imps_ctx* c, int site, const double* O, double* val) {
  if (!c || !O || !val) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    site = wrap(c, site);
    const int d = c->d;
    std::vector<double> Ar((size_t)d * d), Bi((size_t)d * d);
    for (size_t k = 0; k < (size_t)d * d; ++k) {
      Ar[k] = O[2 * k];
      Bi[k] = O[2 * k + 1];
    }
    const size_t n = (size_t)c->D * d * c->D;
    double* t = ws(c, "lx_t", n);
    auto expect = [&](const std::vector<double>& R, const double* ket) {
      HIPCHK(launch_mix_mid(c->D, d, c->D, R.data(), ket, t, 0.0, c->stream));
      return dot(c, n, c->AC[site], t);
    };
    const double zAr = expect(Ar, c->AC[site]), zBr = expect(Bi, c->AC[site]);
    double zAi = 0.0, zBi = 0.0;
    if (c->cplx) {
      double* m = ws(c, "lx_mi", n);
      HIPCHK(launch_fill(n, 0.0, m, c->stream));
      HIPCHK(launch_cplx_axpy_i(n, -1.0, c->AC[site], m, c->stream));
      zAi = expect(Ar, m);
      zBi = expect(Bi, m);
    }
    const double np = dot_psi(c);
    val[0] = (zAr - zBi) / np;
    val[1] = (zAi + zBr) / np;
  });
}

int imps_vumps_default_params(imps_vumps_params* p) {
  if (!p) return IMPS_E_ARG;
  p->tol = 1.0e-10;
  p->maxiter = 300;
  p->verbosity = 0;
  p->tol_min = 1.0e-14;
  p->tol_max = 1.0e-4;
  p->eigs_tolfactor = 1.0e-3;
  p->gauge_tolfactor = 1.0e-6;
  p->envs_tolfactor = 1.0e-4;
  p->krylovdim = 30;
  p->eig_maxiter = 300;
  p->gauge_maxiter = 300;
  p->envs_maxiter = 300;
  p->prec = IMPS_PREC_FP64;
  p->refine_below = 0.0;
  p->eig_floor = 0.0;
  p->eig_dynamic = 1;
  p->eig_eager = 1;
  p->eig_tol = 1.0e-10;
  p->gauge_dynamic = 1;
  p->gauge_tol = 1.0e-13;
  p->gauge_tol_min = 1.0e-14;
  p->gauge_tol_max = 1.0e-4;
  p->envs_dynamic = 1;
  p->envs_tol = 1.0e-10;
  p->envs_tol_min = 1.0e-14;
  p->envs_tol_max = 1.0e-4;
  p->envs_krylovdim = 30;
  p->finalize = nullptr;
  p->finalize_user = nullptr;
  p->refine_eig = 1;
  p->gauge_eig_miniter = 10;
  return IMPS_OK;
}

namespace {
// updatetol(alg_*, iter, eps) (utility.jl:150-158) for the three solvers of imps_vumps_params
struct SolverTols {
  const imps_vumps_params* p;
  double eig(int it, double eps) const {
    return p->eig_dynamic ? updatetol(p->eigs_tolfactor, it, eps, p->tol_min, p->tol_max) : p->eig_tol;
  }
  double gauge(int it, double eps) const {
    return p->gauge_dynamic ? updatetol(p->gauge_tolfactor, it, eps, p->gauge_tol_min, p->gauge_tol_max)
                            : p->gauge_tol;
  }
  double envs(int it, double eps) const {
    return p->envs_dynamic ? updatetol(p->envs_tolfactor, it, eps, p->envs_tol_min, p->envs_tol_max)
                           : p->envs_tol;
  }
};

void check_params(const imps_vumps_params* p) {
  require(p->tol >= 0 && p->maxiter >= 0 && p->krylovdim >= 1 && p->eig_maxiter >= 1 && p->gauge_maxiter >= 1 &&
              p->envs_maxiter >= 1 && p->envs_krylovdim >= 1,
          IMPS_E_ARG, "VUMPS parameters: negative tolerance or non-positive solver sizes");
  require(0 <= p->tol_min && p->tol_min <= p->tol_max && 0 <= p->gauge_tol_min &&
              p->gauge_tol_min <= p->gauge_tol_max && 0 <= p->envs_tol_min && p->envs_tol_min <= p->envs_tol_max,
          IMPS_E_ARG, "DynamicTol requires 0 <= tol_min <= tol_max");
  require(p->eig_dynamic || p->eig_tol > 0, IMPS_E_ARG, "VUMPS parameters: fixed eigensolver tolerance must be > 0");
  require(p->gauge_dynamic || p->gauge_tol > 0, IMPS_E_ARG, "VUMPS parameters: fixed gauge tolerance must be > 0");
  require(p->envs_dynamic || p->envs_tol > 0, IMPS_E_ARG, "VUMPS parameters: fixed environment tolerance must be > 0");
}

// finalize (vumps.jl:75, idmrg.jl:60): the callback may replace the state and environments
// through the ABI; environments it invalidated are rebuilt at tol_envs.  Returns true to stop.
bool run_finalize(imps_ctx* c, const imps_vumps_params* p, int it, double tol_envs, EnvStats* est) {
  if (!p->finalize) return false;
  const int N = c->N;
  HIPCHK(hipStreamSynchronize(c->stream));
  const int fr = p->finalize(it, c, p->finalize_user);
  require(fr >= 0, fr, "finalize callback returned " + std::to_string(fr) + (c->err.empty() ? "" : ": " + c->err));
  require(c->N == N && (int)c->AL.size() == N, IMPS_E_ARG, "finalize must not change the unit cell");
  need_match(c);
  if (!c->envs_valid) envs_recalculate(c, tol_envs, p->envs_maxiter, p->envs_krylovdim, est);
  return fr > 0;
}
}  // namespace

int imps_vumps_run(imps_ctx* c, const imps_vumps_params* p, imps_vumps_info* info, double* losses,
                   int nloss, imps_iter_cb cb, void* user) {
  if (!c || !p) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    check_params(p);
    const SolverTols stol{p};
    const int N = c->N;
    const long hac0 = c->hac_applies, hc0 = c->hc_applies;
    auto t0 = std::chrono::steady_clock::now();
    auto secs = [&] {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    const long qf0 = g_qr_fallbacks;
    EnvStats est;
    double tl = 0, tg = 0, te = 0, tk = 0, tf = 0;
    int sweeps_total = 0;
    auto lap = [&](const char* phase, double& acc, auto&& fn) {
      RoctxRange range(phase);
      auto a = std::chrono::steady_clock::now();
      fn();
      HIPCHK(hipStreamSynchronize(c->stream));
      acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    };
    if (!c->envs_valid) lap("vumps_envs", te, [&] { envs_recalculate(c, 1.0e-10, p->envs_maxiter, p->envs_krylovdim, &est); });
    int prec = p->prec;
    bool refining = false;  // refinement phase with JD corrections in p->prec (refine_eig)
    const long rf0 = g_refine_fallbacks;
    int switch_iter = -1;
    std::vector<double> lam_ac(N, 1.0), lam_c(N, 1.0);
    double best_eps = std::numeric_limits<double>::infinity();
    bool stalled = false;
    double eps = 0.0;
    lap("vumps_galerkin", tk, [&] { eps = galerkin(c, PREC_FP64); });
    lap("vumps_envs", te, [&] { envs_recalculate(c, stol.envs(0, eps), p->envs_maxiter, p->envs_krylovdim, &est); });
    int nl = 0;
    if (losses && nl < nloss) losses[nl++] = eps;
    bool converged = false;
    double f = 0.0, t_reach = -1.0;
    int it = 0;
    for (it = 1; it <= p->maxiter; ++it) {
      RoctxRange iter_range("vumps_iteration");
      const int siter = it - 1;
      // FP64 refinement: below refine_below, or as soon as the reduced-precision phase
      // stops improving (an iteration after the third that gains less than 10 % on the best
      // Galerkin error so far: the operator's rounding level has been reached).  With refine_eig
      // the eigensolves from then on are FP64-residual refinements whose corrections keep using
      // the reduced precision (eigsolve_refined); otherwise they become FP64 Lanczos.
      if (prec != PREC_FP64 && !refining && (eps < p->refine_below || (it > 4 && stalled))) {
        switch_iter = it;
        if (p->refine_eig)
          refining = true;
        else
          prec = PREC_FP64;
      }
      double tol_eig = stol.eig(siter, eps);
      // a reduced-precision operator cannot resolve residuals below its own rounding
      // level: the eigensolver tolerance is floored there until the FP64 refinement
      const double floor_rel = (prec == PREC_FP64 || refining) ? 0.0
                               : p->eig_floor > 0.0 ? p->eig_floor
                                                    : (prec == PREC_FP32 ? 1e-6 : 1e-4);
      lap("vumps_localupdate", tl, [&] {
        std::vector<double> ta(N), tc(N);
        for (int site = 0; site < N; ++site) {
          ta[site] = std::max(tol_eig, floor_rel * std::max(1.0, std::fabs(lam_ac[site])));
          tc[site] = std::max(tol_eig, floor_rel * std::max(1.0, std::fabs(lam_c[site])));
        }
        local_update(c, ta, tc, p->krylovdim, p->eig_maxiter, p->eig_eager != 0, prec, lam_ac, lam_c, refining);
      });
      int sw = 0;
      lap("vumps_gauge", tg, [&] { gauge_step(c, stol.gauge(siter, eps), p->gauge_maxiter, &sw, p->gauge_eig_miniter); });
      sweeps_total += sw;
      lap("vumps_envs", te, [&] { envs_recalculate(c, stol.envs(siter, eps), p->envs_maxiter, p->envs_krylovdim, &est); });
      // finalize (vumps.jl:75), between recalculate! and calc_galerkin
      bool fstop = false;
      lap("vumps_finalize", te, [&] { fstop = run_finalize(c, p, it, stol.envs(siter, eps), &est); });
      lap("vumps_galerkin", tk, [&] { eps = galerkin(c, PREC_FP64); });
      stalled = eps > 0.9 * best_eps;
      best_eps = std::min(best_eps, eps);
      if (losses && nl < nloss) losses[nl++] = eps;
      lap("vumps_energy", tf, [&] { f = energy(c); });
      if (p->verbosity > 0) {
        printf("VUMPS iter %4d : eps = %.3e f = %.8g  [%.2fs: local %.2f gauge %.2f (%d sweeps) envs %.2f]\n", it, eps, f,
               secs(), tl, tg, sw, te);
        fflush(stdout);
      }
      if (eps <= p->tol) {
        converged = true;
        t_reach = secs();
      }
      if (cb && cb(it, eps, f, user) != 0) break;
      if (converged || fstop) break;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (info) {
      info->niter = std::min(it, p->maxiter);
      info->converged = converged ? 1 : 0;
      info->itererr = eps;
      info->energy = f;
      info->seconds = secs();
      info->t_galerkin_reached = t_reach;
      info->hac_applies = (int)(c->hac_applies - hac0);
      info->hc_applies = (int)(c->hc_applies - hc0);
      info->prec_switch_iter = switch_iter;
      info->t_local = tl;
      info->t_gauge = tg;
      info->t_envs = te;
      info->t_galerkin = tk;
      info->t_energy = tf;
      info->gauge_sweeps = sweeps_total;
      info->gmres_ops = est.ops;
      info->qr_fallbacks = (int)(g_qr_fallbacks - qf0);
      info->refine_fallbacks = (int)(g_refine_fallbacks - rf0);
    }
  });
}

int imps_idmrg_run(imps_ctx* c, const imps_vumps_params* p, imps_vumps_info* info, double* losses, int nloss,
                   imps_iter_cb cb, void* user) {
  if (!c || !p) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_idmrg_run");
    need_match(c);
    const int N = c->N, D = c->D, d = c->d, w = c->w;
    const long hac0 = c->hac_applies;
    auto t0 = std::chrono::steady_clock::now();
    auto secs = [&] {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    check_params(p);
    const SolverTols stol{p};
    const size_t nA = (size_t)D * d * D, nC = (size_t)D * D, nE = (size_t)D * w * D;
    // envs = DMRGCache(psi0, H) at its default tolerance (idmrg.jl:46)
    if (!c->envs_valid) envs_recalculate(c, 1.0e-10, p->envs_maxiter, p->envs_krylovdim, nullptr);
    double eps = galerkin(c, PREC_FP64);
    int nl = 0;
    if (losses && nl < nloss) losses[nl++] = eps;
    double* Cold = ws(c, "id_Cold", nC);
    double* Env = ws(c, "id_env", nE);
    double E = 0.0;
    bool converged = false;
    int it = 0;
    for (it = 1; it <= p->maxiter; ++it) {
      const double tol_eig = stol.eig(it - 1, eps);
      // _localupdate_sweep_idmrg! (idmrg.jl:12-34)
      d2d(c, Cold, c->C[N - 1], nC);
      for (int pos = 0; pos < N; ++pos) {
        eigsolve_ac(c, pos, tol_eig, p->krylovdim, p->eig_maxiter, p->eig_eager != 0, PREC_FP64);
        d2d(c, c->AL[pos], c->pAC[pos], nA);
        qrpos(c, D * d, D, c->AL[pos], c->C[pos], nullptr, act_right(c, pos));
        d2d(c, c->AC[pos], pos < N - 1 ? c->AL[pos] : c->pAC[pos], nA);
        // transfer_leftenv! (envs.jl:591-596): a pushed environment is no longer pinned
        const int nxt = (pos + 1) % N;
        push_mpo_left(c, pos, c->GL[pos], c->AL[pos], Env);
        d2d(c, c->GL[nxt], Env, nE);
        c->gl_id[nxt] = 0;
      }
      for (int pos = N - 1; pos >= 0; --pos) {
        E = eigsolve_ac(c, pos, tol_eig, p->krylovdim, p->eig_maxiter, p->eig_eager != 0, PREC_FP64).lambda;
        const int prv = (pos - 1 + N) % N;
        lqpos(c, D, d * D, c->pAC[pos], c->C[prv], c->AR[pos], nullptr, act_left(c, pos));
        d2d(c, c->AC[pos], pos > 0 ? c->AR[pos] : c->pAC[pos], nA);
        // transfer_rightenv! (envs.jl:599-604)
        push_mpo_right(c, pos, c->GR[pos], c->AR[pos], Env);
        d2d(c, c->GR[prv], Env, nE);
        c->gr_id[prv] = 0;
      }
      // finalize (idmrg.jl:60), before the error criterion
      const bool fstop = run_finalize(c, p, it, stol.envs(it - 1, eps), nullptr);
      require(c->D == D, IMPS_E_ARG, "IDMRG: finalize must not change the bond dimension");
      // error criterion: drift of the centre matrix on bond 0 (idmrg.jl:63)
      HIPCHK(launch_axpby(nC, 1.0, c->C[N - 1], -1.0, Cold, c->stream));
      eps = nrm2(c, nC, Cold);
      if (losses && nl < nloss) losses[nl++] = eps;
      if (p->verbosity > 0) {
        printf("IDMRG iter %4d : eps = %.3e f = %.10g  [%.2fs]\n", it, eps, E, secs());
        fflush(stdout);
      }
      if (eps <= p->tol) converged = true;
      if (cb && cb(it, eps, E, user) != 0) break;
      if (converged || fstop) break;
    }
    const int niter = std::min(it, p->maxiter);
    // rebuild from AR with the dynamic gauge tolerance, then recalculate! (idmrg.jl:76-84)
    const double tol_g = stol.gauge(niter, eps);
    std::vector<double*> raw(N);
    for (int l = 0; l < N; ++l) {
      raw[l] = c->pAC[l];
      d2d(c, raw[l], c->AR[l], nA);
    }
    double* I = ws(c, "sr_I", nC);
    HIPCHK(launch_set_identity(D, I, D, c->stream));
    uniform_leftorth(c, raw, I, tol_g, p->gauge_maxiter);
    double* C0 = ws(c, "sr_C0", nC);
    d2d(c, C0, c->C[N - 1], nC);
    uniform_rightorth(c, c->AL, C0, tol_g, p->gauge_maxiter);
    for (int l = 0; l < N; ++l) mul_AL_C(c, c->AL[l], c->C[l], c->AC[l]);
    envs_recalculate(c, 1.0e-10, p->envs_maxiter, p->envs_krylovdim, nullptr);
    c->envs_valid = true;
    const double f = energy(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (info) {
      *info = imps_vumps_info{};
      info->niter = niter;
      info->converged = converged ? 1 : 0;
      info->itererr = eps;
      info->energy = f;
      info->seconds = secs();
      info->t_galerkin_reached = converged ? info->seconds : -1.0;
      info->hac_applies = (int)(c->hac_applies - hac0);
      info->prec_switch_iter = -1;
    }
  });
}

// One TDVP step (tdvp.jl:33-52); a complex dt needs a complex state (the embedded form), whose
// tensors are projected back onto the embedded form at the end.
static void tdvp_step(imps_ctx* c, std::complex<double> dt, double tol, int krylovdim, int maxiter,
                      double gauge_tol, int gauge_maxiter) {
  RoctxRange range("tdvp_step");
  need_match(c);
  require(std::isfinite(dt.real()) && std::isfinite(dt.imag()), IMPS_E_ARG, "imps_tdvp_timestep: dt must be finite");
  require(dt.imag() == 0.0 || c->cplx, IMPS_E_STATE,
          "imps_tdvp_timestep: a complex dt (real-time step) needs a complex state (imps_set_state_complex)");
  const int N = c->N, D = c->D, d = c->d;
  // envs = DMRGCache(psi, H) when none exist (tdvp.jl:34)
  if (!c->envs_valid) envs_recalculate(c, 1.0e-10, 300, 30, nullptr);
  const size_t nA = (size_t)D * d * D, nC = (size_t)D * D;
  bool ok = true;
  for (int loc = 0; loc < N; ++loc) {
    Op hac = [&](const double* x, double* y) { hac_apply(c, loc, x, y, PREC_FP64); };
    Op hc = [&](const double* x, double* y) { hc_apply(c, loc, x, y, PREC_FP64); };
    if (dt.imag() == 0.0) {
      ok = lanczos_expm(c, hac, nA, c->AC[loc], c->pAC[loc], dt.real(), tol, krylovdim, maxiter).converged && ok;
      ok = lanczos_expm(c, hc, nC, c->C[loc], c->pC[loc], dt.real(), tol, krylovdim, maxiter).converged && ok;
    } else {
      ok = lanczos_expm_c(c, hac, nA, c->AC[loc], c->pAC[loc], dt, tol, krylovdim, maxiter).converged && ok;
      ok = lanczos_expm_c(c, hc, nC, c->C[loc], c->pC[loc], dt, tol, krylovdim, maxiter).converged && ok;
    }
  }
  if (!ok) fprintf(stderr, "imps: TDVP integrate: exponentiate not converged (dt = %g%+gi)\n", dt.real(), dt.imag());
  // ALs = regauge!(ACs, Cs) (ortho.jl:214-219); CanonicalIMPS(ALs, C[end]) = gaugefix!(:R)
  // (canonicalmps.jl:180-195); recalculate!(envs, psi', H)
  for (int loc = 0; loc < N; ++loc) regauge_site(c, loc);
  gauge_step(c, gauge_tol, gauge_maxiter, nullptr);
  if (c->cplx) project_state_c(c);
  envs_recalculate(c, 1.0e-10, 300, 30, nullptr);
  HIPCHK(hipStreamSynchronize(c->stream));
}

int imps_tdvp_timestep(imps_ctx* c, double dt, double tol, int krylovdim, int maxiter, double gauge_tol,
                       int gauge_maxiter) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] { tdvp_step(c, std::complex<double>(dt, 0.0), tol, krylovdim, maxiter, gauge_tol, gauge_maxiter); });
}

int imps_tdvp_timestep_complex(imps_ctx* c, double dt_re, double dt_im, double tol, int krylovdim, int maxiter,
                               double gauge_tol, int gauge_maxiter) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    tdvp_step(c, std::complex<double>(dt_re, dt_im), tol, krylovdim, maxiter, gauge_tol, gauge_maxiter);
  });
}

// sum over the bra of <AC_i| O1 ... O2_j |ket_i AR ...> for real d x d operators O1, O2 (host),
// each j of js, unnormalised; ket_i is the ket tensor at site i (AC, or a rotation of it)
static std::vector<double> correlator_terms(imps_ctx* c, const double* O1, const double* O2, int i, int nj,
                                            const int* js, const double* ket_i) {
  const int N = (int)c->AL.size(), D = c->D, d = c->d;
  const size_t DD = (size_t)D * D;
  double* V = ws(c, "co_V", DD);
  double* V2 = ws(c, "co_V2", DD);
  double* Z = ws(c, "co_Z", (size_t)D * d * D);
  double* E = ws(c, "co_E", DD);
  double* I = ws(c, "co_I", DD);
  HIPCHK(launch_set_identity(D, I, D, c->stream));
  // V[b, c] = AC[a,u,b] O1[u,s] ket[a,s,c] at site i (correlators.jl:28-30)
  const double* AC = c->AC[wrap(c, i)];
  HIPCHK(launch_mix_mid(D, d, D, O1, ket_i, Z, 0.0, c->stream));
  gemm2(c, 'T', 'N', D, D, d * D, 1.0, AC, (long)d * D, Z, (long)d * D, 0.0, V, D);
  std::vector<double> O2v(O2, O2 + (size_t)d * d);
  auto push = [&](int l) {
    const double* AR = c->AR[((l % N) + N) % N];
    push_left2(c, V, D, AR, AR, V2, D);
    std::swap(V, V2);
  };
  int ctr = i;
  std::vector<double> vals(nj);
  for (int k = 0; k < nj; ++k) {
    const int j = js[k];
    for (int l = ctr + 1; l < j; ++l) push(l);
    // E = V[b, c] AR[b,u,e] O2[u,s] AR[c,s,e] = trace of the slice push of V with O2 (31-38)
    slice_sum_left(c, {SliceTerm{V, (long)D, O2v}}, c->AR[((j % N) + N) % N], E, D);
    vals[k] = trace_prod(c, D, I, E);
    push(j);
    ctr = j;
  }
  return vals;
}

static void require_increasing(int i, int nj, const int* js, const char* what) {
  for (int k = 0; k < nj; ++k)
    require(js[k] > (k == 0 ? i : js[k - 1]), IMPS_E_ARG, std::string(what) + ": need i < j1 < j2 < ...");
}

int imps_correlator(imps_ctx* c, const double* O1, const double* O2, int i, int nj, const int* js, double* out) {
  if (!c || !O1 || !O2 || nj < 0 || (nj > 0 && (!js || !out))) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    require_symmetric_if_complex(c, O1, "imps_correlator");
    require_symmetric_if_complex(c, O2, "imps_correlator");
    require_increasing(i, nj, js, "imps_correlator");
    const std::vector<double> vals = correlator_terms(c, O1, O2, i, nj, js, c->AC[wrap(c, i)]);
    const double nrm = dot_psi(c);
    for (int k = 0; k < nj; ++k) out[k] = vals[k] / nrm;
  });
}

// <O1_i O2_j> for ComplexF64 operators: with O = A + i B (A, B real) and z(R1, R2) the correlator
// of real operators, <O1 O2> = z(A1,A2) - z(B1,B2) + i (z(A1,B2) + z(B1,A2)).  On a complex state
// the embedding's real trace gives Re z, and Im z = Re z evaluated with the ket -i AC at site i.
int imps_correlator_complex(imps_ctx* c, const double* O1, const double* O2, int i, int nj, const int* js,
                            double* out) {
  if (!c || !O1 || !O2 || nj < 0 || (nj > 0 && (!js || !out))) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    require_increasing(i, nj, js, "imps_correlator_complex");
    const int d = c->d;
    const size_t dd = (size_t)d * d;
    std::vector<double> A1(dd), B1(dd), A2(dd), B2(dd);
    for (size_t k = 0; k < dd; ++k) {
      A1[k] = O1[2 * k];
      B1[k] = O1[2 * k + 1];
      A2[k] = O2[2 * k];
      B2[k] = O2[2 * k + 1];
    }
    const double* AC = c->AC[wrap(c, i)];
    auto terms = [&](const std::vector<double>& R1, const std::vector<double>& R2, const double* ket) {
      return correlator_terms(c, R1.data(), R2.data(), i, nj, js, ket);
    };
    const auto aa = terms(A1, A2, AC), bb = terms(B1, B2, AC), ab = terms(A1, B2, AC), ba = terms(B1, A2, AC);
    std::vector<double> aai(nj, 0.0), bbi(nj, 0.0), abi(nj, 0.0), bai(nj, 0.0);
    if (c->cplx) {
      const size_t n = (size_t)c->D * d * c->D;
      double* m = ws(c, "co_mi", n);
      HIPCHK(launch_fill(n, 0.0, m, c->stream));
      HIPCHK(launch_cplx_axpy_i(n, -1.0, AC, m, c->stream));
      aai = terms(A1, A2, m);
      bbi = terms(B1, B2, m);
      abi = terms(A1, B2, m);
      bai = terms(B1, A2, m);
    }
    const double nrm = dot_psi(c);
    for (int k = 0; k < nj; ++k) {
      out[2 * k] = (aa[k] - bb[k] - abi[k] - bai[k]) / nrm;
      out[2 * k + 1] = (aai[k] - bbi[k] + ab[k] + ba[k]) / nrm;
    }
  });
}

int imps_dot_state(imps_ctx* c, int nsites, const double* const* AL2, double* out) {
  if (!c || !AL2 || !out) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_dot_state");
    need_state(c);
    const int N = (int)c->AL.size(), D = c->D, d = c->d;
    require(nsites == N, IMPS_E_ARG, "imps_dot_state: unit cells differ");
    const size_t nA = (size_t)D * d * D, DD = (size_t)D * D;
    std::vector<double*> A2(N);
    for (int l = 0; l < N; ++l) {
      A2[l] = ws(c, "ds_A2_" + std::to_string(l), nA);
      h2d(c, A2[l], AL2[l], nA);
    }
    // dot(psi1, psi2) = dominant eigenvalue of TransferMatrix(psi2.AL, psi1.AL) from v0 = I
    // (canonicalmps.jl:262-269); psi1 is the context's state
    double* v0 = ws(c, "ds_v0", DD);
    double* ev = ws(c, "ds_ev", DD);
    HIPCHK(launch_set_identity(D, v0, D, c->stream));
    Op tm = [&](const double* v, double* o) {
      double* t0 = ws(c, "ds_t0", DD);
      double* t1 = ws(c, "ds_t1", DD);
      const double* cur = v;
      for (int l = 0; l < N; ++l) {
        double* dst = (l == N - 1) ? o : (l % 2 == 0 ? t0 : t1);
        push_left2(c, cur, D, A2[l], c->AL[l], dst, D);
        cur = dst;
      }
    };
    KInfo ki;
    double im = 0.0;
    const double re = arnoldi_dominant(c, tm, DD, v0, ev, 1.0e-10, 30, 300, true, &ki, nullptr, &im);
    require(ki.converged != 0, IMPS_E_SOLVER, "dot: the transfer-matrix eigensolver did not converge");
    out[0] = re;
    out[1] = im;
  });
}

// dot(psi1, psi2) of complex states: psi1 the context's complex state, AL2 ComplexF64 (D, d, D)
// tensors; the same transfer map on the embedding, whose pushes keep the embedded form, so the
// complex Arnoldi finds the complex dominant eigenvalue (re, im)
int imps_dot_state_complex(imps_ctx* c, int nsites, const double* const* AL2, double* out) {
  if (!c || !AL2 || !out) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_state(c);
    require(c->cplx, IMPS_E_STATE, "imps_dot_state_complex: the context holds a real state (imps_dot_state)");
    const int N = (int)c->AL.size(), D = c->D, d = c->d, Dc = D / 2;
    require(nsites == N, IMPS_E_ARG, "imps_dot_state_complex: unit cells differ");
    const size_t nA = (size_t)D * d * D, DD = (size_t)D * D;
    std::vector<double*> A2(N);
    for (int l = 0; l < N; ++l) {
      require(AL2[l] != nullptr, IMPS_E_ARG, "imps_dot_state_complex: null tensor");
      A2[l] = ws(c, "ds_A2_" + std::to_string(l), nA);
      embed_from_host(c, AL2[l], Dc, d, Dc, 1.0, A2[l]);
    }
    double* v0 = ws(c, "ds_v0", DD);
    double* ev = ws(c, "ds_ev", DD);
    HIPCHK(launch_set_identity(D, v0, D, c->stream));
    Op tm = [&](const double* v, double* o) {
      double* t0 = ws(c, "ds_t0", DD);
      double* t1 = ws(c, "ds_t1", DD);
      const double* cur = v;
      for (int l = 0; l < N; ++l) {
        double* dst = (l == N - 1) ? o : (l % 2 == 0 ? t0 : t1);
        push_left2(c, cur, D, A2[l], c->AL[l], dst, D);
        cur = dst;
      }
    };
    KInfo ki;
    double im = 0.0;
    const double re = arnoldi_dominant(c, tm, DD, v0, ev, 1.0e-10, 30, 300, true, &ki, nullptr, &im);
    require(ki.converged != 0, IMPS_E_SOLVER, "dot: the transfer-matrix eigensolver did not converge");
    out[0] = re;
    out[1] = im;
  });
}

// dominant_env (envs.jl:83-99) of the context's state: the unit-cell transfer map of envs.jl:34-68,
// from the left over AL[0..N-1] (side 0) or from the right over AR[N-1..0] (side 1), identity
// channel when Ws is empty, else DenseIMPO tensors Ws[l % nw] (w, d, w, d); Arnoldi (:LM, eager)
// from ones.  out (D w D doubles on the device) receives the eigenvector with unit norm and a
// non-negative entry sum; returns the eigenvalue.
double dominant_env_dev(imps_ctx* c, int side, const std::vector<std::vector<double>>& Ws, int w, double tol,
                        int krylovdim, int maxiter, double* out) {
  const int N = (int)c->AL.size(), D = c->D;
  const bool ident = Ws.empty();
  const int nw = (int)Ws.size();
  const size_t n = (size_t)D * w * D;
  double* v0 = ws(c, "de_v0", n);
  HIPCHK(launch_fill(n, 1.0, v0, c->stream));
  Op tm = [&](const double* v, double* o) {
    double* t0 = ws(c, "de_t0", n);
    double* t1 = ws(c, "de_t1", n);
    const double* cur = v;
    for (int k = 0; k < N; ++k) {
      const int l = side == 0 ? k : N - 1 - k;
      double* dst = (k == N - 1) ? o : (k % 2 == 0 ? t0 : t1);
      if (ident && side == 0)
        push_left2(c, cur, D, c->AL[l], c->AL[l], dst, D);
      else if (ident)
        push_right2(c, cur, D, c->AR[l], c->AR[l], dst, D);
      else if (side == 0)
        push_w_left(c, Ws[l % nw], w, cur, c->AL[l], dst);
      else
        push_w_right(c, Ws[l % nw], w, cur, c->AR[l], dst);
      cur = dst;
    }
  };
  bool cplx = false;
  const double lam = arnoldi_dominant(c, tm, n, v0, out, tol, krylovdim, maxiter, true, nullptr, &cplx);
  require(!cplx, IMPS_E_UNSUPPORTED, "dominant_env: the dominant eigenvalue is complex");
  // L ./= norm(L), sign fixed by sum(L) >= 0
  const double nr = nrm2(c, n, out);
  const double sum = dot(c, n, out, v0);
  require(nr > 0 && std::isfinite(nr), IMPS_E_SOLVER, "dominant_env: zero eigenvector");
  axpby(c, n, (sum < 0 ? -1.0 : 1.0) / nr, out, 0.0, out);
  return lam;
}

// host copies of nw DenseIMPO tensors (w, d, w, d), checked against the context's state
std::vector<std::vector<double>> dense_tensors(imps_ctx* c, int nw, int w, int d, const double* const* W,
                                               const char* who) {
  require(W != nullptr && nw >= 1 && w >= 1 && d == c->d, IMPS_E_ARG,
          std::string(who) + ": operator dimensions do not match the state");
  std::vector<std::vector<double>> Ws;
  const size_t nW = (size_t)w * d * w * d;
  for (int l = 0; l < nw; ++l) {
    require(W[l] != nullptr, IMPS_E_ARG, std::string(who) + ": null tensor");
    Ws.emplace_back(W[l], W[l] + nW);
  }
  return Ws;
}

int imps_dominant_env(imps_ctx* c, int side, int nw, int w, int d, const double* const* W, double tol,
                      int krylovdim, int maxiter, double* lambda, double* env) {
  if (!c || !lambda || !env || (side != 0 && side != 1)) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_dominant_env");
    need_state(c);
    require(tol > 0 && krylovdim >= 1 && maxiter >= 1, IMPS_E_ARG, "imps_dominant_env: bad solver parameters");
    std::vector<std::vector<double>> Ws;
    if (W == nullptr)
      require(w == 1 && nw == 0, IMPS_E_ARG, "imps_dominant_env: the identity channel has w = 1, nw = 0");
    else
      Ws = dense_tensors(c, nw, w, d, W, "imps_dominant_env");
    const size_t n = (size_t)c->D * w * c->D;
    double* ev = ws(c, "de_ev", n);
    *lambda = dominant_env_dev(c, side, Ws, w, tol, krylovdim, maxiter, ev);
    d2h(c, env, ev, n);
  });
}

int imps_dense_expectation(imps_ctx* c, int nw, int w, int d, const double* const* W, double tol, int krylovdim,
                           int maxiter, double* value, double* const* GL_out, double* const* GR_out) {
  if (!c || !value) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require_real_state(c, "imps_dense_expectation");
    need_state(c);
    require(tol > 0 && krylovdim >= 1 && maxiter >= 1, IMPS_E_ARG,
            "imps_dense_expectation: bad solver parameters");
    const std::vector<std::vector<double>> Ws = dense_tensors(c, nw, w, d, W, "imps_dense_expectation");
    const int N = (int)c->AL.size(), D = c->D;
    const size_t n = (size_t)D * w * D, DD = (size_t)D * D;
    // DMRGCache(psi, W::DenseIMPO) (envs.jl:284-304): dominant environments at the cell's ends,
    // MPO-channel pushes in between
    std::vector<double*> GL(N), GR(N);
    for (int l = 0; l < N; ++l) {
      GL[l] = ws(c, "dn_GL_" + std::to_string(l), n);
      GR[l] = ws(c, "dn_GR_" + std::to_string(l), n);
    }
    dominant_env_dev(c, 0, Ws, w, tol, krylovdim, maxiter, GL[0]);
    dominant_env_dev(c, 1, Ws, w, tol, krylovdim, maxiter, GR[N - 1]);
    for (int l = 1; l < N; ++l) push_w_left(c, Ws[(l - 1) % nw], w, GL[l - 1], c->AL[l - 1], GL[l]);
    for (int l = N - 2; l >= 0; --l) push_w_right(c, Ws[(l + 1) % nw], w, GR[l + 1], c->AR[l + 1], GR[l]);
    // normalize_envs! (envs.jl:559-581): unit-norm GR[i], then GL[i+1] /= <C[i], H_C(i) C[i]> with
    // H_C(i) C = sum_w GL[i+1][:, w, :] C GR[i][:, w, :] (effective.jl:37-50)
    double* Y = ws(c, "dn_Y", DD);
    double* M = ws(c, "dn_M", DD);
    for (int i = 0; i < N; ++i) {
      const double nr = nrm2(c, n, GR[i]);
      if (nr > 0) axpby(c, n, 1.0 / nr, GR[i], 0.0, GR[i]);
      const int j = (i + 1) % N;
      double lam = 0.0;
      for (int k = 0; k < w; ++k) {
        // M[b', b] = sum_{a, a'} C[a, b] GL[a, k, a'] C[a', b'] -> lam += sum M[b', b] GR[b', k, b]
        // (elementwise: GR's first index is the ket side)
        gemm2(c, 'N', 'N', D, D, D, 1.0, GL[j] + (size_t)D * k, (long)D * w, c->C[i], D, 0.0, Y, D);
        gemm2(c, 'T', 'N', D, D, D, 1.0, Y, D, c->C[i], D, 0.0, M, D);
        HIPCHK(launch_copy2d(D, D, GR[i] + (size_t)D * k, (long)D * w, Y, D, c->stream));
        lam += dot(c, DD, M, Y);
      }
      if (std::fabs(lam) > 0) axpby(c, n, 1.0 / lam, GL[j], 0.0, GL[j]);
    }
    // expectationvalue(psi, W::DenseIMPO, envs) (expval.jl:60-77): per site
    // sum_{w'} sum_{a, b} Z[a, w', b] GR[b, w', a] with Z the MPO-channel left push of GL through AC
    double* Z = ws(c, "dn_Z", n);
    double E = 0.0;
    for (int l = 0; l < N; ++l) {
      push_w_left(c, Ws[l % nw], w, GL[l], c->AC[l], Z);
      for (int k = 0; k < w; ++k) {
        HIPCHK(launch_copy2d(D, D, Z + (size_t)D * k, (long)D * w, Y, D, c->stream));
        HIPCHK(launch_copy2d(D, D, GR[l] + (size_t)D * k, (long)D * w, M, D, c->stream));
        E += trace_prod(c, D, Y, M);
      }
    }
    *value = E / dot_psi(c);
    for (int l = 0; l < N; ++l) {
      if (GL_out && GL_out[l]) d2h(c, GL_out[l], GL[l], n);
      if (GR_out && GR_out[l]) d2h(c, GR_out[l], GR[l], n);
    }
  });
}

int imps_vumps_batch(int device, int ninst, imps_batch_instance* inst, const imps_vumps_params* p, int nworkers,
                     double gauge_tol) {
  if (ninst < 0 || (ninst > 0 && !inst) || !p || nworkers < 1) return IMPS_E_ARG;
  std::vector<int> order(ninst);
  for (int i = 0; i < ninst; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (double)inst[a].D * inst[a].D * inst[a].D > (double)inst[b].D * inst[b].D * inst[b].D;
  });
  std::atomic<int> next(0);
  std::atomic<int> worst(IMPS_OK);
  auto worker = [&]() {
    imps_ctx* c = nullptr;
    const int st0 = imps_create(device, &c);
    if (st0 != IMPS_OK) {
      worst = st0;
      return;
    }
    while (true) {
      const int k = next.fetch_add(1);
      if (k >= ninst) break;
      imps_batch_instance& in = inst[order[k]];
      auto t0 = std::chrono::steady_clock::now();
      imps_vumps_info info{};
      int st = imps_set_mpo(c, in.nsites, in.w, in.d, in.W);
      if (st == IMPS_OK) st = imps_set_state_raw(c, in.nsites, in.D, in.A, gauge_tol, p->gauge_maxiter);
      if (st == IMPS_OK) st = imps_vumps_run(c, p, &info, nullptr, 0, nullptr, nullptr);
      in.status = st;
      in.niter = info.niter;
      in.converged = info.converged;
      in.itererr = info.itererr;
      in.energy = info.energy;
      in.hac_applies = info.hac_applies;
      in.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (st != IMPS_OK) worst = st;
    }
    imps_destroy(c);
  };
  std::vector<std::thread> th;
  const int nw = std::min(nworkers, std::max(ninst, 1));
  for (int i = 0; i < nw; ++i) th.emplace_back(worker);
  for (auto& t : th) t.join();
  return worst.load();
}

int imps_bench_krylov(imps_ctx* c, long n, int nv, int reps, imps_krylov_timing* t) {
  if (!c || !t || n < 1 || nv < 1 || nv > 32 || reps < 1) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    // three bases of nv vectors plus the vector they act on, used in turn, so that repeated
    // launches read their operands from HBM rather than from the last level of cache
    const int NB = 3;
    const size_t per = (size_t)(nv + 1) * (size_t)n;
    double* B = ws(c, "bk_B", NB * per);
    HIPCHK(launch_fill_random(NB * per, 777, B, c->stream));
    double* coef = ws(c, "bk_coef", 64);
    HIPCHK(launch_fill(64, 1e-3, coef, c->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double ms[3];
    for (int k = 0; k < 3; ++k) {
      auto run = [&](int r) {
        double* X = B + (size_t)(r % NB) * per;
        double* y = X + (size_t)nv * n;
        if (k == 0)
          HIPCHK(launch_multidot((size_t)n, nv, X, (size_t)n, y, c->dpart, c->stream));
        else
          HIPCHK(launch_axpy_reduce((size_t)n, nv, X, (size_t)n, coef, y, c->dpart, k == 1, c->stream));
      };
      run(0);
      HIPCHK(hipEventRecord(e0, c->stream));
      for (int r = 1; r <= reps; ++r) run(r);
      HIPCHK(hipEventRecord(e1, c->stream));
      HIPCHK(hipEventSynchronize(e1));
      float el = 0.f;
      HIPCHK(hipEventElapsedTime(&el, e0, e1));
      ms[k] = el / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    t->ms_multidot = ms[0];
    t->ms_axpy_dots = ms[1];
    t->ms_axpy_norm = ms[2];
    t->bytes_multidot = 8.0 * (double)(nv + 1) * (double)n;
    t->bytes_axpy = 8.0 * (double)(nv + 2) * (double)n;
  });
}

int imps_bench_hac(imps_ctx* c, int site, int reps, int prec, imps_hac_timing* t) {
  if (!c || !t || reps <= 0) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    need_match(c);
    need_envs(c);
    site = wrap(c, site);
    const int D = c->D, w = c->w, d = c->d;
    const size_t n = (size_t)D * d * D;
    double* x = ws(c, "bh_x", n);
    double* y = ws(c, "bh_y", n);
    HIPCHK(launch_fill_random(n, 12345, x, c->stream));
    struct PlanesScope {
      imps_ctx* c;
      ~PlanesScope() { c->planes_site = -1; }
    } scope{c};
    t->ms_prepare = 0.0;
    if (prec == PREC_BF16X3 && hac_planes_ok(c, site)) {
      // the per-eigensolve operand split, timed on its own
      hac_planes_prepare(c, site);  // allocates the planes
      hipEvent_t p0, p1;
      HIPCHK(hipEventCreate(&p0));
      HIPCHK(hipEventCreate(&p1));
      HIPCHK(hipEventRecord(p0, c->stream));
      hac_planes_prepare(c, site);
      HIPCHK(hipEventRecord(p1, c->stream));
      HIPCHK(hipEventSynchronize(p1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, p0, p1));
      t->ms_prepare = ms;
      (void)hipEventDestroy(p0);
      (void)hipEventDestroy(p1);
      c->planes_site = site;
    }
    hac_apply(c, site, x, y, prec);  // warm-up
    hipEvent_t ev[4];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    double s_all = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int r = 0; r < reps; ++r) {
      hac_apply(c, site, x, y, prec, ev);
      HIPCHK(hipEventSynchronize(ev[3]));
      float a, b, cc, tot;
      HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
      HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
      HIPCHK(hipEventElapsedTime(&cc, ev[2], ev[3]));
      HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[3]));
      s1 += a;
      s2 += b;
      s3 += cc;
      s_all += tot;
    }
    // flops of the GEMMs as launched (one level fewer each when the structural levels are
    // identities)
    const int lv = (w >= 2 && c->gl_id[site] && c->gr_id[site]) ? 1 : 0;
    t->flop_gemm1 = 2.0 * (w - lv) * d * (double)D * D * D;
    t->flop_gemm2 = t->flop_gemm1;
    for (auto& e : ev) (void)hipEventDestroy(e);
    t->ms_apply = s_all / reps;
    t->ms_gemm1 = s1 / reps;
    t->ms_wmix = s2 / reps;
    t->ms_gemm2 = s3 / reps;
  });
}

int imps_gemm_dev(imps_ctx* c, char transa, char transb, int m, int n, int k, double alpha, const double* d_a,
                  int lda, const double* d_b, int ldb, double beta, double* d_c, int ldc, int prec) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    require((transa == 'N' || transa == 'T') && (transb == 'N' || transb == 'T'), IMPS_E_ARG, "imps_gemm_dev: trans");
    require(m >= 0 && n >= 0 && k >= 0, IMPS_E_ARG, "imps_gemm_dev: negative size");
    require(ldc >= std::max(1, m), IMPS_E_ARG, "imps_gemm_dev: ldc");
    require(lda >= std::max(1, transa == 'N' ? m : k), IMPS_E_ARG, "imps_gemm_dev: lda");
    require(ldb >= std::max(1, transb == 'N' ? k : n), IMPS_E_ARG, "imps_gemm_dev: ldb");
    gemm2(c, transa, transb, m, n, k, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, prec);
  });
}

int imps_set_graph_capture(int on) {
  if (on != 0 && on != 1) return IMPS_E_ARG;
  g_graphs.store(on);
  return IMPS_OK;
}

int imps_set_gemm_autotune(int on) {
  if (on < 0 || on > 2 + 5) return IMPS_E_ARG;
  g_gemm_autotune.store(on);
  return IMPS_OK;
}

int imps_dev_copy(imps_ctx* c, void* d_dst, const void* d_src, size_t bytes) {
  if (!c || (bytes % sizeof(double)) != 0) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    const size_t n = bytes / sizeof(double);
    const int rows = 1 << 20;
    const size_t cols = (n + rows - 1) / rows;
    for (size_t j = 0; j < cols; ++j) {
      const size_t off = j * (size_t)rows;
      const int r = (int)std::min<size_t>(rows, n - off);
      HIPCHK(launch_copy2d(r, 1, (const double*)d_src + off, r, (double*)d_dst + off, r, c->stream));
    }
  });
}

int imps_dev_alloc(imps_ctx* c, size_t bytes, void** d_ptr) {
  if (!c || !d_ptr) return IMPS_E_ARG;
  return run_guarded(c, [&] { HIPCHK(hipMalloc(d_ptr, std::max<size_t>(bytes, 1))); });
}

int imps_dev_free(imps_ctx* c, void* d_ptr) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    if (d_ptr) HIPCHK(hipFree(d_ptr));
  });
}

int imps_memcpy_h2d(imps_ctx* c, void* d_dst, const void* h_src, size_t bytes) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] { h2d_bytes(c, d_dst, h_src, bytes); });
}

int imps_memcpy_d2h(imps_ctx* c, void* h_dst, const void* d_src, size_t bytes) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] { d2h_bytes(c, h_dst, d_src, bytes); });
}

int imps_synchronize(imps_ctx* c) {
  if (!c) return IMPS_E_ARG;
  return run_guarded(c, [&] { HIPCHK(hipStreamSynchronize(c->stream)); });
}

}  // extern "C"
