// C ABI of the gfx950 ITP engine (declared in include/itp2d.h).
//
// Objects own device memory; all kernels go to the context's one stream.
// Table construction (host, FP64, the reference's formulas) and the sweep
// sequences come from engine.hpp; kernels from kernels.hip.
#include <mutex>
#include <shared_mutex>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <exception>
#include <functional>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/itp2d.h"
#include "engine.hpp"
#include "block_jacobi.hpp"
#include "cluster_eig.hpp"
#include "gpu.hpp"

using namespace itp;

// Device-wide operations (allocation, freeing, device synchronisation,
// synchronous copies) must not run while another thread captures a graph on
// its stream: they would join or invalidate the capture. While any context
// has graphs on, they take this lock exclusively and captures take it
// shared; otherwise (the default) they take no lock.
std::shared_mutex& device_op_mutex() {
  static std::shared_mutex m;
  return m;
}
std::atomic<int>& graph_contexts() {
  static std::atomic<int> n{0};
  return n;
}
template <class F>
hipError_t device_op(F&& f) {
  if (graph_contexts().load() == 0) return f();
  std::unique_lock<std::shared_mutex> lk(device_op_mutex());
  return f();
}
hipError_t dev_malloc(void** p, size_t bytes) { return device_op([&] { return hipMalloc(p, bytes); }); }
hipError_t dev_free(void* p) { return device_op([&] { return hipFree(p); }); }
hipError_t dev_memcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  return device_op([&] { return hipMemcpy(dst, src, bytes, kind); });
}

// HIP graphs of fixed launch sequences on the context's stream, keyed by what
// determines their launches and arguments. The first call with a key runs the
// launches directly (kernel attributes and lazily allocated buffers are set
// up then); the second captures them and instantiates a graph, which later
// calls replay. At most kGraphCap graphs are kept (least recently used out).
struct GraphCache {
  using Key = std::array<uint64_t, 4>;
  struct Entry {
    Key key;
    hipGraphExec_t exec = nullptr;  // null: seen once, not captured yet
    uint64_t last = 0;
  };
  static constexpr size_t kGraphCap = 64;
  std::vector<Entry> entries;
  uint64_t clock = 0;
  void clear() {
    for (auto& e : entries)
      if (e.exec) (void)hipGraphExecDestroy(e.exec);
    entries.clear();
  }
  // Runs fn (launches on st only, nothing that synchronises or copies to the
  // host), directly or as a graph.
  template <class F>
  hipError_t run(hipStream_t st, const Key& key, F&& fn) {
    Entry* hit = nullptr;
    for (auto& e : entries)
      if (e.key == key) hit = &e;
    clock++;
    if (hit && hit->exec) {
      hit->last = clock;
      return hipGraphLaunch(hit->exec, st);
    }
    if (!hit) {
      if (entries.size() >= kGraphCap) {
        auto old = std::min_element(entries.begin(), entries.end(),
                                    [](const Entry& a, const Entry& b) { return a.last < b.last; });
        if (old->exec) (void)hipGraphExecDestroy(old->exec);
        entries.erase(old);
      }
      entries.push_back(Entry{key, nullptr, clock});
      return fn();
    }
    hit->last = clock;
    std::shared_lock<std::shared_mutex> lk(device_op_mutex());
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const hipError_t ef = fn();
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    if (ef != hipSuccess || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return ef != hipSuccess ? ef : e;
    }
    e = hipGraphInstantiate(&hit->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      hit->exec = nullptr;
      return e;
    }
    return hipGraphLaunch(hit->exec, st);
  }
};

// Tuned split-K chunk counts of the overlap per (precision bytes, grid points,
// padded states), shared by every context of the process.
struct GramWisdom {
  std::mutex m;
  std::map<std::array<long long, 3>, int> w;
};
GramWisdom& gram_wisdom() {
  static GramWisdom g;
  return g;
}

// pinned bytes per context for downloads: the largest is the energies of
// 4096 states (E, sigma, <H>, <H^2>: 48 bytes a state)
constexpr size_t kPinBytes = 256 * 1024;

// Pinned download buffers outlive their contexts: a destroyed context's
// buffer goes back to this list for the next context instead of being
// unpinned (hipHostFree) while other threads' contexts keep working; they are
// released when the process ends. At most one buffer per context alive at once;
// portable, so a buffer serves contexts on any device.
struct PinnedBuffers {
  std::mutex m;
  std::vector<void*> free;
};
PinnedBuffers& pinned_buffers() {
  static PinnedBuffers* p = new PinnedBuffers;  // never destroyed: contexts may outlive static destruction
  return *p;
}
void* take_pinned() {
  {
    std::lock_guard<std::mutex> lk(pinned_buffers().m);
    if (!pinned_buffers().free.empty()) {
      void* p = pinned_buffers().free.back();
      pinned_buffers().free.pop_back();
      return p;
    }
  }
  void* p = nullptr;
  if (device_op([&] { return hipHostMalloc(&p, kPinBytes, hipHostMallocPortable); }) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}
void give_pinned(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lk(pinned_buffers().m);
  pinned_buffers().free.push_back(p);
}

uint64_t next_object_id() {
  static std::atomic<uint64_t> id{1};
  return id++;
}

struct itp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // this device's address of the fixed scratch array
  std::string err;
  int num_cus = 0;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  bool timing = false;
  // event pairs recorded around phases without synchronising; resolved when
  // the times are read or the pool is full
  std::vector<hipEvent_t> evpool;
  std::vector<int> pair_phase;
  size_t ev_used = 0;
  double ms[ITP_NUM_PHASES] = {};
  long long cnt[ITP_NUM_PHASES] = {};
  long long eig_calls = 0, eig_sweeps = 0;
  int eig_inner = 1;   // element sweeps per block sub-problem and round (upper bound)
  int eig_grid = 0;    // workgroups of the eigensolver (0: sized by the matrix; the grid barrier cost grows with them)
  int eig_mode = 0;    // 0: Jacobi handing off to the refinement iteration, 1: Jacobi only
  double eig_handoff = 0.5;   // hand-off threshold on |a_ij| / |a_ii - a_jj|
  double eig_cluster = 0.1;   // cluster start: pairs above this |a_ij| / |a_ii - a_jj| are solved exactly (0: off)
  int fp32_rotation = 1;      // FP32 sets' rotation: 0 FP32 MFMA, 1 split-BF16x3 MFMA
  unsigned eig_spin_limit = 1u << 24;  // polls of a block-Jacobi grid barrier before it reports a timeout
  double chol_tol = 0;        // Cholesky pivot threshold relative to the diagonal (0: 1e-5 FP32 sets, 1e-12 FP64)
  // eigenvalue check before the eigendecomposition (tridiagonalization and
  // bisection) from this matrix size on; calls and rejections
  int td_min_n = 1024;
  long long td_checks = 0, td_rejects = 0;
  // fixed launch sequences (propagation, overlap, Cholesky factor) replayed as HIP graphs
  bool graphs = false;
  GraphCache gcache;
  // launch-shape tuning on first use of a shape (the tuned shapes are shared by
  // the process's contexts: gram_wisdom())
  bool autotune = false;
  long long tunings = 0;
  long long oa_done = 0, oa_fallback = 0;
  long long cl_calls = 0, cl_clusters = 0, cl_largest = 0, cl_fallback = 0;
  // pinned host buffer of the small per-call downloads (Downloads)
  char* pin = nullptr;
  size_t pin_cap = 0;
  // Device memory of the context's destroyed objects, kept for the objects
  // created after them (a batch worker runs instance after instance): hipFree
  // waits for the whole device, every other context's queued work included.
  // Blocks by capacity; `sizes` holds every block the context allocated.
  struct Pool {
    std::mutex m;
    std::multimap<size_t, void*> free;
    std::unordered_map<void*, size_t> sizes;
    size_t cached = 0;
  } pool;
  // fills and clears queued on the stream that host-synchronous copies must
  // not overtake (cmemcpy waits for them)
  bool fills_pending = false;
  // the caller's handle plus one per live object created on the context: the
  // last release frees the stream and the memory (objects may outlive the
  // handle's itp_ctx_destroy)
  std::atomic<int> refs{1};
};

namespace {

// The small device-to-host downloads of one call (energies, flags,
// eigenvalues), staged through the context's pinned buffer: the copies queue
// asynchronously on the stream and wait() synchronises the stream once, then
// hands each download to its destination. Copies straight into pageable host
// memory would each wait for the stream on their own. One Downloads at a time
// per context (its adds and wait are not interleaved with other calls).
struct Downloads {
  struct Item {
    void* dst;
    size_t off, bytes;
  };
  static constexpr int kMax = 8;
  itp_ctx* c;
  Item items[kMax];
  int n = 0;
  size_t off = 0;
  explicit Downloads(itp_ctx* ctx) : c(ctx) {}
  hipError_t add(void* dst, const void* src, size_t bytes) {
    const size_t at = (off + 15) / 16 * 16;
    if (c->pin && n < kMax && at + bytes <= c->pin_cap) {
      items[n++] = {dst, at, bytes};
      off = at + bytes;
      return hipMemcpyAsync(c->pin + at, src, bytes, hipMemcpyDeviceToHost, c->stream);
    }
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
  }
  hipError_t wait() {
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess)
      for (int i = 0; i < n; i++) std::memcpy(items[i].dst, c->pin + items[i].off, items[i].bytes);
    n = 0;
    off = 0;
    return e;
  }
};

}  // namespace

struct itp_grid {
  itp_ctx* ctx = nullptr;
  GridInfo g;
  LineGeom lg[2];  // [0] FP64 fields, [1] FP32 fields
  cplx<double>* d_tw = nullptr;
};

struct itp_potential {
  itp_ctx* ctx = nullptr;  // held: destroy needs no parent object
  itp_grid* grid = nullptr;
  bool null = true;
  std::vector<double> V;
  double* dV = nullptr;
};

struct itp_stateset {
  uint64_t id = next_object_id();  // graph keys
  itp_ctx* ctx = nullptr;          // held: destroy needs no parent object
  itp_grid* grid = nullptr;
  size_t N = 0;
  int prec = ITP_FP64;
  size_t esz = 16;
  void* buf[2] = {nullptr, nullptr};
  int cur = 0;
  // orthonormalisation workspace
  int np = 0, nc = 0, nchunk = 1;
  long long kchunk = 0;
  int nchunk32 = 1;             // split-K of the bf16x3 Gram (FP32 sets)
  long long kchunk32 = 0;
  int nchunk_max = 1;           // chunks the partial-product buffer holds
  bool gram_tuned = false;      // chunk count taken from the context's tuning (autotune)
  cplx<float>* C32 = nullptr;   // FP32 copy of the coefficients (FP32 sets)
  std::vector<double> lam_host; // ascending overlap eigenvalues of the last orthonormalization
  cplx<double>* part = nullptr;
  cplx<double>* A[2] = {nullptr, nullptr};
  cplx<double>* V[2] = {nullptr, nullptr};
  cplx<double>* Vinit = nullptr;
  cplx<double>* bjQ = nullptr;     // block Jacobi rotations and diagonalised sub-blocks
  cplx<double>* bjD = nullptr;
  double* bjslots = nullptr;
  int* bjident = nullptr;
  cplx<double>* oa_work = nullptr;  // refinement iteration: Y, S, G, E, second iterate (np x np each)
  double* oa_lam = nullptr;
  double* oa_slots = nullptr;
  int* oa_state = nullptr;
  // cluster start: pair counter and largest ratio, strong pairs, cluster members and offsets
  unsigned long long* cl_hdr = nullptr;
  int* cl_pairs = nullptr;
  int* cl_members = nullptr;
  int* cl_offsets = nullptr;
  cplx<double>* Vsave = nullptr;  // eigenvector estimate before a cluster start, restored if its refinement fails
  cplx<double>* oa_part = nullptr;  // split-K partial products of the refinement (small matrices only)
  cplx<double>* C = nullptr;
  double* lam = nullptr;
  int* perm = nullptr;
  double* info = nullptr;
  int* bad = nullptr;
  // Cholesky orthonormalization: original diagonal, failure flag; field norms
  double* diag0 = nullptr;
  int* chol_fail = nullptr;
  double* chol_mm = nullptr;     // largest and smallest R_ii^2 of the last factorization, then every R_ii^2
  double chol_kappa = 0;         // their ratio (0: no factorization yet or it failed)
  std::vector<double> chol_diag; // R_ii^2 of the last factorization (empty if it failed)
  double* norms = nullptr;
  // energies
  cplx<double>* epart = nullptr;  // [2][N][max(sx,sy)]
  double* dE = nullptr;           // E[N], sigma[N]
  cplx<double>* dmean = nullptr;  // mean[N], msq[N]
  void* td_scratch = nullptr;     // eigenvalue check (tridiagonalization), allocated on first use
};

struct MemberTabs {
  ExpKinTabs host;
  std::vector<double> vh, vf, vp;
  double* dev = nullptr;   // one allocation for all tables of the member
  float* dev32 = nullptr;  // and one for their FP32 copies (FP32 sweep programs)
  MemberDev d;
};

struct itp_split {
  uint64_t id = next_object_id();  // graph keys: id, and gen (bumped when the tables are rebuilt)
  uint64_t gen = 0;
  itp_ctx* ctx = nullptr;  // held: destroy needs no parent object
  itp_potential* pot = nullptr;
  int halforder = 1;
  double B = 0, eps = 0;
  std::vector<double> coefs;
  std::vector<std::unique_ptr<MemberTabs>> members;
};

struct itp_hamiltonian {
  itp_ctx* ctx = nullptr;  // held: destroy needs no parent object
  itp_potential* pot = nullptr;
  double B = 0;
  KinTabs host;
  double* dev = nullptr;
  KinHDev d;
};

// Fixed device scratch of itp_propagate_noalloc: the intermediate fields of
// one propagation for callers that cannot allocate. Each device holds its own
// copy for the life of the process; its address on a device comes from
// hipGetSymbolAddress (the array's name is only a host-side stand-in).
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of the scratch array on `device`, looked up once per device with
// that device current (called from itp_ctx_create).
hipError_t scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

constexpr int kMaxSweeps = 40;
constexpr int kOAMaxIter = 8;
constexpr int kClusterPairCap = 8192;  // strong pairs listed for the cluster start; more: block Jacobi

int set_err(itp_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCK(ctx, expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return set_err((ctx), ITP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class F>
int guarded(itp_ctx* c, F f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return set_err(c, ITP_E_NOMEM, "host allocation failed");
  } catch (const std::exception& ex) {
    return set_err(c, ITP_E_INVALID, ex.what());
  }
}

// Cached bytes per context at most; a cached block serves requests of at
// least half its capacity (or all but 64 KiB of it).
constexpr size_t kPoolCap = 4ull << 30;

int dalloc(itp_ctx* c, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  void* hit = nullptr;
  size_t cap = 0;
  {
    std::lock_guard<std::mutex> lk(c->pool.m);
    auto it = c->pool.free.lower_bound(bytes);
    if (it != c->pool.free.end() && it->first <= std::max(2 * bytes, bytes + (64 << 10))) {
      hit = it->second;
      cap = it->first;
      c->pool.cached -= cap;
      c->pool.free.erase(it);
    }
  }
  if (hit) {
    // a reused block reads as zero, as fresh memory from the driver does (the
    // fill is ordered before host-synchronous copies by cmemcpy)
    const hipError_t e = hipMemsetAsync(hit, 0, cap, c->stream);
    c->fills_pending = true;
    if (e != hipSuccess) {
      std::lock_guard<std::mutex> lk(c->pool.m);
      c->pool.free.emplace(cap, hit);
      c->pool.cached += cap;
      *p = nullptr;
      return set_err(c, ITP_E_HIP, std::string("clearing a reused block: ") + hipGetErrorString(e));
    }
    *p = hit;
    return ITP_OK;
  }
  hipError_t e = dev_malloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_err(c, ITP_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  std::lock_guard<std::mutex> lk(c->pool.m);
  c->pool.sizes[*p] = bytes;
  return ITP_OK;
}

// Host-synchronous copy into or out of a context's memory, after the fills
// and clears queued on its stream.
hipError_t cmemcpy(itp_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (c->fills_pending) {
    const hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    c->fills_pending = false;
  }
  return dev_memcpy(dst, src, bytes, kind);
}

// Clear of a context's memory, queued on its stream.
hipError_t cmemset(itp_ctx* c, void* dst, int v, size_t bytes) {
  c->fills_pending = true;
  return hipMemsetAsync(dst, v, bytes, c->stream);
}

// Returns a block of dalloc to the context's cache (after the context's
// stream has finished every use of it), or frees it when the cache is full.
void dfree(itp_ctx* c, void* p) {
  if (!p) return;
  (void)hipStreamSynchronize(c->stream);
  {
    std::lock_guard<std::mutex> lk(c->pool.m);
    auto it = c->pool.sizes.find(p);
    if (it != c->pool.sizes.end() && c->pool.cached + it->second <= kPoolCap) {
      c->pool.free.emplace(it->second, p);
      c->pool.cached += it->second;
      return;
    }
    if (it != c->pool.sizes.end()) c->pool.sizes.erase(it);
  }
  (void)dev_free(p);
}

// Drops one reference of the context; the last one frees it.
void ctx_release(itp_ctx* c) {
  if (--c->refs > 0) return;
  (void)hipStreamSynchronize(c->stream);
  if (c->ws) (void)dev_free(c->ws);
  for (auto& kv : c->pool.free) (void)dev_free(kv.second);
  for (auto& e : c->evpool)
    if (e) (void)hipEventDestroy(e);
  give_pinned(c->pin);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int ensure_ws(itp_ctx* c, size_t bytes) {
  if (bytes <= c->ws_bytes) return ITP_OK;
  if (c->ws) {
    HIPCK(c, hipStreamSynchronize(c->stream));
    dfree(c, c->ws);
    c->ws = nullptr;
    c->ws_bytes = 0;
  }
  int rc = dalloc(c, &c->ws, bytes);
  if (rc) return rc;
  c->ws_bytes = bytes;
  return ITP_OK;
}

// Accumulate the recorded event pairs into the per-phase totals.
void flush_events(itp_ctx* c) {
  if (c->ev_used == 0) return;
  (void)hipEventSynchronize(c->evpool[c->ev_used - 1]);
  // events 2k and 2k+1 bracket the phase pair_phase[k]
  const size_t pairs = std::min(c->ev_used / 2, c->pair_phase.size());
  for (size_t k = 0; k < pairs; k++) {
    const int ph = c->pair_phase[k];
    float t = 0;
    if (ph >= 0 && ph < ITP_NUM_PHASES &&
        hipEventElapsedTime(&t, c->evpool[2 * k], c->evpool[2 * k + 1]) == hipSuccess) {
      c->ms[ph] += t;
      c->cnt[ph]++;
    }
  }
  c->ev_used = 0;
  c->pair_phase.clear();
}

// HIP-event timing of one phase on the context's stream (only when enabled).
// The stop event is recorded by stop() or, on any other way out of the
// scope (early returns included), by the destructor, so every start event
// has its stop event next to it in the pool.
struct Phase {
  itp_ctx* c;
  int ph;
  bool on;
  Phase(itp_ctx* c_, int ph_) : c(c_), ph(ph_), on(c_->timing) {
    if (!on) return;
    if (c->ev_used + 2 > c->evpool.size()) flush_events(c);
    (void)hipEventRecord(c->evpool[c->ev_used++], c->stream);
  }
  ~Phase() { stop(); }
  Phase(const Phase&) = delete;
  Phase& operator=(const Phase&) = delete;
  void stop() {
    if (!on) return;
    (void)hipEventRecord(c->evpool[c->ev_used++], c->stream);
    c->pair_phase.push_back(ph);
    on = false;
  }
};

LineGeom make_geom(const GridInfo& g, const cplx<double>* d_tw, size_t esz) {
  LineGeom lg;
  make_plan(g.sx, lg.plan_x);
  make_plan(g.sy, lg.plan_y);
  lg.tw_x = d_tw;
  lg.mk_x = d_tw + g.sx;
  lg.tw_y = d_tw + 2 * g.sx;
  lg.mk_y = d_tw + 2 * g.sx + g.sy;
  // lines per workgroup: the most that keep a sweep within 64 KiB of LDS
  // (two workgroups per CU), at least 1, at most 8
  auto pick = [&](int n, int lines) {
    int best = 1;
    for (int L = 8; L >= 1; L /= 2) {
      if (lines % L) continue;
      const size_t bytes = (size_t)3 * L * (n + 1) * esz + (size_t)2 * n * esz + 272 * 16;
      if (bytes <= 64 * 1024) { best = L; break; }
    }
    return best;
  };
  lg.L_row = std::min(4, pick(g.sx, g.sy));
  lg.L_col = std::min(4, pick(g.sy, g.sx));
  return lg;
}

// Host tables in one device allocation, uploaded with one copy; ptrs[i] is
// null for an empty table
int upload_tables(itp_ctx* c, const std::vector<const std::vector<double>*>& tabs, double** dev,
                  std::vector<const double*>& ptrs) {
  size_t total = 0;
  for (auto t : tabs) total += t->size();
  int rc = dalloc(c, (void**)dev, total * sizeof(double));
  if (rc) return rc;
  std::vector<double> h(total);
  size_t off = 0;
  ptrs.clear();
  for (auto t : tabs) {
    ptrs.push_back(t->empty() ? nullptr : *dev + off);
    std::copy(t->begin(), t->end(), h.begin() + off);
    off += t->size();
  }
  if (total) HIPCK(c, cmemcpy(c, *dev, h.data(), total * sizeof(double), hipMemcpyHostToDevice));
  return ITP_OK;
}

// Runs independent host jobs (table loops over the grid) on threads of their
// own when the grid is large enough to pay for them; the first exception
// thrown by a job is rethrown here.
void run_jobs(std::vector<std::function<void()>>& jobs, long long M) {
  if (jobs.size() < 2 || M < 16384) {
    for (auto& j : jobs) j();
    return;
  }
  std::vector<std::exception_ptr> err(jobs.size());
  std::vector<std::thread> th;
  th.reserve(jobs.size() - 1);
  auto run = [&](size_t i) {
    try {
      jobs[i]();
    } catch (...) {
      err[i] = std::current_exception();
    }
  };
  size_t started = 1;
  try {
    for (; started < jobs.size(); started++) th.emplace_back(run, started);
  } catch (const std::system_error&) {
    // no more threads: the jobs not started run here
  }
  run(0);
  for (size_t i = started; i < jobs.size(); i++) run(i);
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

// FP32 copies of host tables in one device allocation; ptrs[i] is null for an empty table
int upload_tables32(itp_ctx* c, std::initializer_list<const std::vector<double>*> tabs, float** dev,
                    std::vector<const float*>& ptrs) {
  size_t total = 0;
  for (auto t : tabs) total += t->size();
  int rc = dalloc(c, (void**)dev, std::max<size_t>(total, 1) * sizeof(float));
  if (rc) return rc;
  std::vector<float> h(total);
  size_t off = 0;
  ptrs.clear();
  for (auto t : tabs) {
    ptrs.push_back(t->empty() ? nullptr : *dev + off);
    for (size_t i = 0; i < t->size(); i++) h[off + i] = float((*t)[i]);
    off += t->size();
  }
  if (total) HIPCK(c, cmemcpy(c, *dev, h.data(), total * sizeof(float), hipMemcpyHostToDevice));
  return ITP_OK;
}

void free_member(itp_ctx* c, MemberTabs* m) {
  if (m && m->dev) { dfree(c, m->dev); m->dev = nullptr; }
  if (m && m->dev32) { dfree(c, m->dev32); m->dev32 = nullptr; }
}

// FP32 copies of a member's tables (for FP32 sweep programs)
int upload_member32(itp_ctx* c, MemberTabs* m) {
  std::vector<const float*> q;
  int rc = upload_tables32(c, {&m->host.mult, &m->host.xm, &m->host.xm2, &m->host.ym, &m->vh, &m->vf, &m->vp},
                           &m->dev32, q);
  if (rc) return rc;
  m->d.kin.mult32 = q[0];
  m->d.kin.xm32 = q[1];
  m->d.kin.xm2_32 = q[2];
  m->d.kin.ym32 = q[3];
  m->d.vhalf32 = q[4];
  m->d.vfull32 = q[5];
  m->d.vpref32 = q[6];
  return ITP_OK;
}

// (Re)build the device tables of every MultiProductSplit member for time step eps
// (multiproductsplit.cpp:23-47, secondordersplit.cpp:21-69).
int build_split_tables(itp_split* sp) {
  itp_ctx* c = sp->pot->grid->ctx;
  const GridInfo& g = sp->pot->grid->g;
  const bool has_v = !sp->pot->null;
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (auto& m : sp->members) free_member(c, m.get());
  sp->members.clear();
  sp->gen = next_object_id();
  // every member's host tables first, the independent loops side by side
  std::vector<std::unique_ptr<MemberTabs>> tabs;
  std::vector<std::function<void()>> jobs;
  const double* V = sp->pot->V.data();
  const double B = sp->B;
  for (int t = 1; t <= sp->halforder; t++) {
    tabs.emplace_back(new MemberTabs);
    MemberTabs* m = tabs.back().get();
    const double h = sp->eps / t;
    const double coef = sp->coefs[t - 1];
    if (has_v) {
      jobs.push_back([&g, B, h, m] { expkin_tables(g, B, h, -1.0, 1.0, m->host); });
      jobs.push_back([&g, V, h, m] { exppot_values(g, V, h, -0.5, 1.0, m->vh); });
      if (t >= 2) jobs.push_back([&g, V, h, m] { exppot_values(g, V, h, -1.0, 1.0, m->vf); });
      jobs.push_back([&g, V, h, coef, m] { exppot_values(g, V, h, -0.5, coef, m->vp); });
    } else {
      jobs.push_back([&g, B, h, t, coef, m] { expkin_tables(g, B, h, -1.0 * t, coef, m->host); });
    }
  }
  run_jobs(jobs, g.M);
  for (int t = 1; t <= sp->halforder; t++) {
    std::unique_ptr<MemberTabs> m = std::move(tabs[t - 1]);
    std::vector<const double*> p;
    int rc = upload_tables(c, {&m->host.mult, &m->host.xm, &m->host.xm2, &m->host.ym, &m->vh, &m->vf, &m->vp},
                           &m->dev, p);
    if (rc) return rc;
    m->d.kin = KinDev{p[0], p[1], p[2], p[3]};
    m->d.vhalf = p[4];
    m->d.vfull = p[5];
    m->d.vpref = p[6];
    rc = upload_member32(c, m.get());
    if (rc) return rc;
    sp->members.push_back(std::move(m));
  }
  return ITP_OK;
}

// One propagation of the whole set. The intermediate fields go to the
// context's workspace (grown with hipMalloc when needed) or, for the
// no-allocation entry point, to the fixed scratch array.
template <class R>
int propagate_impl(itp_split* sp, itp_stateset* s, bool noalloc = false) {
  itp_grid* gr = s->grid;
  itp_ctx* c = gr->ctx;
  const GridInfo& g = gr->g;
  const size_t fbytes = s->N * (size_t)g.M * sizeof(cplx<R>);
  cplx<R>* work = nullptr;
  if (noalloc) {
    if (fbytes > kNoAllocScratchBytes)
      return set_err(c, ITP_E_UNSUPPORTED, "itp_propagate_noalloc: the set needs " + std::to_string(fbytes) +
                                               " bytes of scratch, more than the fixed " +
                                               std::to_string(kNoAllocScratchBytes));
    if (!c->scratch) return set_err(c, ITP_E_HIP, "itp_propagate_noalloc: no scratch address for this device");
    work = (cplx<R>*)c->scratch;
  } else {
    int rc = ensure_ws(c, fbytes);
    if (rc) return rc;
    work = (cplx<R>*)c->ws;
  }
  const Branch br = branch_of(sp->B, g.bt);
  const bool has_v = !sp->pot->null;
  const cplx<R>* in = (const cplx<R>*)s->buf[s->cur];
  cplx<R>* out = (cplx<R>*)s->buf[s->cur ^ 1];
  std::vector<Sweep<R>> prog;
  for (int t = 1; t <= sp->halforder; t++)
    member_program<R>(br, sp->members[t - 1]->d, has_v ? t : 1, has_v, in, work, out, t > 1, prog);
  Phase ph(c, ITP_PHASE_PROPAGATE);
  auto launches = [&]() -> hipError_t {
    GpuBackend be;
    be.st = c->stream;
    run_program<R>(be, g, gr->lg[sizeof(R) == 4], prog, (int)s->N);
    return be.err;
  };
  // (the entry point that allocates nothing launches directly: instantiating a graph allocates)
  const hipError_t err =
      c->graphs && !noalloc
          ? c->gcache.run(c->stream, {sp->id ^ (sp->gen << 20), s->id, (uint64_t)s->cur, (uint64_t)(uintptr_t)work},
                          launches)
          : launches();
  if (err != hipSuccess) return set_err(c, ITP_E_HIP, std::string("propagate: ") + hipGetErrorString(err));
  ph.stop();
  s->cur ^= 1;
  return ITP_OK;
}

template <class R>
int energies_impl(itp_hamiltonian* h, itp_stateset* s, double* E, double* sigma, double* mean, double* msq,
                  bool mean_only = false) {
  itp_grid* gr = s->grid;
  itp_ctx* c = gr->ctx;
  const GridInfo& g = gr->g;
  const size_t fbytes = s->N * (size_t)g.M * sizeof(cplx<R>);
  int rc = ensure_ws(c, 2 * fbytes);
  if (rc) return rc;
  const Branch br = branch_of(h->B, g.bt);
  const double* V = h->pot->null ? nullptr : h->pot->dV;
  const cplx<R>* psi = (const cplx<R>*)s->buf[s->cur];
  cplx<R>* w1 = (cplx<R>*)c->ws;
  cplx<R>* w2 = (cplx<R>*)((char*)c->ws + fbytes);
  const size_t pstride = s->N * (size_t)std::max(g.sx, g.sy);
  cplx<double>* p1 = s->epart;
  cplx<double>* p2 = s->epart + pstride;
  const LineGeom& lg = gr->lg[sizeof(R) == 4];
  std::vector<Sweep<R>> prog;
  Phase ph(c, ITP_PHASE_ENERGIES);
  const int nb = hamiltonian_program<R>(g, lg, br, h->d, V, psi, w1, w2, psi, p1, prog);
  // mean_only: <H> alone (one application; sigma and <H^2> are then meaningless)
  if (!mean_only) hamiltonian_program<R>(g, lg, br, h->d, V, w2, w1, nullptr, psi, p2, prog);
  GpuBackend be;
  be.st = c->stream;
  run_program<R>(be, g, lg, prog, (int)s->N);
  if (be.err != hipSuccess) return set_err(c, ITP_E_HIP, std::string("energies: ") + hipGetErrorString(be.err));
  HIPCK(c, launch_energy_reduce(p1, mean_only ? p1 : p2, nb, (int)s->N, g.dx, s->dE, s->dE + s->N, s->dmean,
                                s->dmean + s->N, c->stream));
  ph.stop();
  Downloads dl(c);
  if (E) HIPCK(c, dl.add(E, s->dE, s->N * sizeof(double)));
  if (sigma) HIPCK(c, dl.add(sigma, s->dE + s->N, s->N * sizeof(double)));
  if (mean) HIPCK(c, dl.add(mean, s->dmean, s->N * 16));
  if (msq) HIPCK(c, dl.add(msq, s->dmean + s->N, s->N * 16));
  HIPCK(c, dl.wait());
  return ITP_OK;
}

// Overlap + eigendecomposition + coefficients; returns after the host has
// checked the eigenvalues (stateset.cpp:145-177).
// The block-Jacobi launch is cooperative: its workgroups spin on each other
// at grid barriers. Contexts on other host threads (batches) could start a
// second one while the first is only partly resident, and each would wait for
// the other's workgroups to leave; one cooperative launch at a time per
// process rules that out. Held from the launch until the stream has drained it.
std::mutex& coop_mutex() {
  static std::mutex m;
  return m;
}

// Connected components of the strong-pair graph on N indices, sizes >= 2,
// as members (by component, ascending indices) and offsets. Returns the
// largest component size (0 if there is no pair).
int strong_clusters(const std::vector<int>& pairs, int N, std::vector<int>& members, std::vector<int>& offsets) {
  std::vector<int> parent(N);
  for (int i = 0; i < N; i++) parent[i] = i;
  auto find = [&](int x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (size_t p = 0; p + 1 < pairs.size(); p += 2) {
    const int a = find(pairs[p]), b = find(pairs[p + 1]);
    if (a != b) parent[a < b ? b : a] = a < b ? a : b;
  }
  std::vector<int> size(N, 0);
  for (int i = 0; i < N; i++) size[find(i)]++;
  std::vector<int> start(N, -1);
  members.clear();
  offsets.assign(1, 0);
  int largest = 0;
  std::vector<std::vector<int>> comp;
  std::vector<int> slot(N, -1);
  for (int i = 0; i < N; i++) {
    const int r = find(i);
    if (size[r] < 2) continue;
    if (slot[r] < 0) {
      slot[r] = (int)comp.size();
      comp.emplace_back();
    }
    comp[slot[r]].push_back(i);
  }
  for (auto& cm : comp) {
    members.insert(members.end(), cm.begin(), cm.end());
    offsets.push_back((int)members.size());
    largest = std::max<int>(largest, (int)cm.size());
  }
  return largest;
}

// Result of one eigensolver attempt, read back after the stream drained.
struct EigRead {
  double info[4] = {0, 0, 0, 0};
  unsigned barrier_err = 0;
  int state[2] = {3, 0};
  int bad = 0;
};

int read_eig(itp_stateset* s, EigRead& r, std::vector<double>& lam, bool oa) {
  itp_ctx* c = s->grid->ctx;
  Downloads dl(c);
  HIPCK(c, dl.add(r.info, s->info, sizeof(r.info)));
  if (oa) HIPCK(c, dl.add(r.state, s->oa_state, sizeof(r.state)));
  HIPCK(c, dl.add(&r.bad, s->bad, sizeof(int)));
  HIPCK(c, dl.add(lam.data(), s->lam, s->N * sizeof(double)));
  HIPCK(c, dl.wait());
  return ITP_OK;
}

// Cluster start for the refinement (cluster_eig.hpp). Acur = V^H S V is the
// current (partly rotated) matrix with eigenvector estimate V = V[0], S = A0
// the original matrix. Strong pairs of Acur on the device, their clusters on
// the host; each cluster's exact rotation is applied to V's cluster columns,
// and the refinement iteration runs on S from there (writing the eigenvalues
// into Acur's diagonal on convergence). Returns 1 on convergence (lam, perm
// and C final), 0 otherwise: too many strong pairs or a cluster too large
// (nothing changed), or no convergence (V restored from Vsave). ident_start:
// V is the identity on entry (no sweep has run).
int cluster_route(itp_stateset* s, cplx<double>* Acur, const cplx<double>* A0, bool ident_start, bool scale,
                  EigRead& rd, std::vector<double>& lam, int& rc) {
  itp_ctx* c = s->grid->ctx;
  const int N = (int)s->N;
  const size_t np2 = (size_t)s->np * s->np * 16;
  rc = ITP_OK;
  auto fail = [&](hipError_t e, const char* what) {
    rc = set_err(c, ITP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
  };
  hipError_t e = launch_cluster_pairs(Acur, s->np, N, c->eig_cluster, kClusterPairCap, s->cl_hdr, s->cl_pairs,
                                      c->stream);
  if (e != hipSuccess) return fail(e, "cluster pairs");
  unsigned long long hdr[2] = {0, 0};
  if ((e = hipMemcpyAsync(hdr, s->cl_hdr, sizeof(hdr), hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess)
    return fail(e, "cluster pairs");
  rd.state[0] = 3;  // refinement not run (yet)
  if (hdr[0] > (unsigned long long)kClusterPairCap) return 0;
  std::vector<int> pairs(2 * hdr[0]);
  if (hdr[0] && ((e = hipMemcpyAsync(pairs.data(), s->cl_pairs, pairs.size() * sizeof(int), hipMemcpyDeviceToHost,
                                     c->stream)) != hipSuccess ||
                 (e = hipStreamSynchronize(c->stream)) != hipSuccess))
    return fail(e, "cluster pairs");
  std::vector<int> members, offsets;
  const int largest = strong_clusters(pairs, N, members, offsets);
  if (largest > kMaxCluster) return 0;
  const int ncl = (int)offsets.size() - 1;
  if (ncl > 0) {
    if ((e = hipMemcpyAsync(s->Vsave, s->V[0], np2, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
      return fail(e, "cluster start");
    // the host vectors outlive the copies: the stream is drained before returning
    if ((e = hipMemcpyAsync(s->cl_members, members.data(), members.size() * sizeof(int), hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(s->cl_offsets, offsets.data(), offsets.size() * sizeof(int), hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess)
      return fail(e, "cluster upload");
    if ((e = launch_cluster_solve(Acur, s->np, s->cl_members, s->cl_offsets, ncl, s->V[0], c->stream)) != hipSuccess)
      return fail(e, "cluster solve");
  }
  // the refinement starts when info[1] == 2 (a hand-off)
  const double start[4] = {0.0, 2.0, 0.0, 0.0};
  if ((e = hipMemcpyAsync(s->info, start, sizeof(start), hipMemcpyHostToDevice, c->stream)) != hipSuccess)
    return fail(e, "cluster start");
  // without clusters, a start that is the identity (before any sweep) needs
  // no products in the refinement's first iteration
  if ((e = launch_oa(A0, s->V[0], s->V[1], s->oa_work, s->oa_lam, s->oa_slots, s->oa_state, s->info, s->np, N,
                     kOAMaxIter, Acur, c->stream, ident_start && ncl == 0 ? s->Vinit : nullptr, s->oa_part)) !=
      hipSuccess)
    return fail(e, "refinement");
  if ((e = launch_eig_finish(Acur, s->V[0], s->np, N, s->lam, s->perm, s->C, s->nc, scale ? 1 : 0, c->stream)) !=
      hipSuccess)
    return fail(e, "eigenvalue sort");
  if ((rc = read_eig(s, rd, lam, true))) return 0;
  if (rd.state[0] != 1) {
    c->cl_fallback++;
    if (ncl > 0 && (e = hipMemcpyAsync(s->V[0], s->Vsave, np2, hipMemcpyDeviceToDevice, c->stream)) != hipSuccess)
      return fail(e, "cluster restore");
    return 0;
  }
  c->cl_calls++;
  c->cl_clusters += ncl;
  c->cl_largest = std::max<long long>(c->cl_largest, largest);
  return 1;
}

// Ascending eigenvalues of the N x N Hermitian matrix A (np x np, device) by
// tridiagonalization and bisection (tridiag.hip), A left unchanged.
int hermitian_eigenvalues(itp_stateset* s, const cplx<double>* A, std::vector<double>& lam) {
  itp_ctx* c = s->grid->ctx;
  const int N = (int)s->N;
  if (!s->td_scratch) {
    int rc = dalloc(c, &s->td_scratch, hermitian_eigenvalues_scratch(s->np, N));
    if (rc) return rc;
  }
  HIPCK(c, launch_hermitian_eigenvalues(A, s->np, N, s->td_scratch, s->lam, c->eig_spin_limit, c->stream));
  lam.resize(N);
  unsigned bar[2] = {0, 0};
  HIPCK(c, hipMemcpyAsync(lam.data(), s->lam, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipMemcpyAsync(bar, (const char*)s->td_scratch + hermitian_eigenvalues_barrier_offset(s->np, N), sizeof(bar),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (bar[1]) return set_err(c, ITP_E_EIGENSOLVER, "tridiagonalization: a grid barrier timed out");
  return ITP_OK;
}

// The reference's checks of the ascending overlap eigenvalues
// (stateset.cpp:164-173): the first that is not positive, or not normal.
int check_eigenvalues(itp_ctx* c, const std::vector<double>& lam, size_t* bad_n, double* bad_lambda) {
  for (size_t n = 0; n < lam.size(); n++) {
    const double ev = lam[n];
    if (ev <= 0) {
      if (bad_n) *bad_n = n;
      if (bad_lambda) *bad_lambda = ev;
      return set_err(c, ITP_E_NONPOSITIVE_EIG, "non-positive overlap eigenvalue " + std::to_string(ev) +
                                                   " at n = " + std::to_string(n));
    }
    if (std::fpclassify(ev) != FP_NORMAL) {
      if (bad_n) *bad_n = n;
      if (bad_lambda) *bad_lambda = ev;
      return set_err(c, ITP_E_NONNORMAL_EIG, "non-normal overlap eigenvalue at n = " + std::to_string(n));
    }
  }
  return ITP_OK;
}

int eigen_stage(itp_stateset* s, size_t* bad_n, double* bad_lambda, bool scale, bool check_positive = true) {
  itp_ctx* c = s->grid->ctx;
  const int N = (int)s->N;
  Phase pe(c, ITP_PHASE_EIGENSOLVE);
  const size_t np2 = (size_t)s->np * s->np * 16;
  const bool hybrid = c->eig_mode == 0;
  const bool clusters = hybrid && c->eig_cluster > 0;
  // Mode 1: Jacobi sweeps to convergence (on A in place, rotations into V).
  // Mode 0 (hybrid): the refinement iteration on the untouched copy A[1]
  // finishes as soon as it can start: from the cluster start (tried before
  // the first sweep and after each sweep, until one attempt's refinement
  // fails), or when the Jacobi sweeps hand off (every unconverged coupling
  // below eig_handoff times its gap). If the refinement fails after a
  // hand-off, the sweeps continue to convergence.
  EigRead rd;
  std::vector<double> lam(N);
  // a non-finite overlap entry (flagged by the Gram kernels) ends here
  {
    Downloads dl(c);
    HIPCK(c, dl.add(&rd.bad, s->bad, sizeof(int)));
    HIPCK(c, hipMemcpyAsync(s->V[0], s->Vinit, np2, hipMemcpyDeviceToDevice, c->stream));
    if (hybrid) HIPCK(c, hipMemcpyAsync(s->A[1], s->A[0], np2, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, dl.wait());
  }
  if (rd.bad) {
    c->eig_calls++;
    return set_err(c, ITP_E_EIGENSOLVER, "overlap matrix has non-finite entries");
  }
  // Large matrices: when the reference's checks would reject the eigenvalues,
  // it raises before any use of the eigenvectors, so those are not computed
  // (the costly case: overlaps of states from noise with a time step too
  // large, which the caller's recovery handles). A Cholesky factorization with
  // every pivot above 1e-12 of its diagonal shows the matrix positive definite
  // far above rounding (every check passes); otherwise the eigenvalues come
  // first, by tridiagonalization and bisection.
  if (check_positive && N >= c->td_min_n) {
    int fail = 0;
    HIPCK(c, launch_cholinv(s->A[0], s->np, N, s->A[1], s->diag0, s->C, s->nc, s->chol_fail, 1e-12, c->stream));
    HIPCK(c, hipMemcpyAsync(&fail, s->chol_fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (hybrid) HIPCK(c, hipMemcpyAsync(s->A[1], s->A[0], np2, hipMemcpyDeviceToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    if (fail) {
      std::vector<double> ev;
      int rc = hermitian_eigenvalues(s, s->A[0], ev);
      if (rc) return rc;
      c->td_checks++;
      if ((rc = check_eigenvalues(c, ev, bad_n, bad_lambda))) {
        c->td_rejects++;
        c->eig_calls++;
        s->lam_host = ev;
        return rc;
      }
    }
  }
  bool done = false, try_clusters = clusters, handoff = hybrid;
  int sweeps = 0, rc = ITP_OK;
  if (try_clusters) {
    done = cluster_route(s, s->A[0], s->A[1], true, scale, rd, lam, rc) != 0;
    if (rc) return rc;
    if (done) c->oa_done++;
    else if (rd.state[0] != 3) try_clusters = false;  // its refinement failed: no more attempts
  }
  // every launch either sweeps, converges, hands off or fails; the launch
  // count bounds the loop as well
  for (int launches = 0; !done && sweeps < kMaxSweeps && launches <= kMaxSweeps; launches++) {
    // one sweep per launch while cluster starts are tried, else up to the budget
    {
      std::unique_lock<std::mutex> coop(coop_mutex());
      const int budget = try_clusters ? 1 : kMaxSweeps - sweeps;
      HIPCK(c, launch_block_jacobi(s->A[0], s->V[0], s->bjQ, s->bjD, s->bjident, s->bjslots, s->np, budget,
                                   c->eig_inner, 1e-14, handoff ? c->eig_handoff : 0.0, s->info, c->eig_grid,
                                   c->eig_spin_limit, c->stream));
      Downloads dl(c);
      HIPCK(c, dl.add(rd.info, s->info, sizeof(rd.info)));
      // the grid barrier's error flag (after the per-check maxima and the arrival counter)
      HIPCK(c, dl.add(&rd.barrier_err, (const char*)(s->bjslots + 2 * (budget + 1)) + sizeof(unsigned),
                      sizeof(unsigned)));
      HIPCK(c, dl.wait());
    }
    if (rd.barrier_err || rd.info[1] < 0) {
      // a grid barrier gave up waiting: A and V are in an unknown state
      c->eig_calls++;
      return set_err(c, ITP_E_EIGENSOLVER, "block Jacobi eigensolver: a grid barrier timed out");
    }
    sweeps += (int)rd.info[0];
    c->eig_sweeps += (long long)rd.info[0];
    if (rd.info[1] == 1.0) {  // converged
      HIPCK(c, launch_eig_finish(s->A[0], s->V[0], s->np, N, s->lam, s->perm, s->C, s->nc, scale ? 1 : 0, c->stream));
      if ((rc = read_eig(s, rd, lam, false))) return rc;
      done = true;
      break;
    }
    if (rd.info[1] == 2.0) {  // handed off: refinement from the Jacobi eigenvectors
      HIPCK(c, launch_oa(s->A[1], s->V[0], s->V[1], s->oa_work, s->oa_lam, s->oa_slots, s->oa_state, s->info, s->np,
                         N, kOAMaxIter, s->A[0], c->stream, nullptr, s->oa_part));
      HIPCK(c, launch_eig_finish(s->A[0], s->V[0], s->np, N, s->lam, s->perm, s->C, s->nc, scale ? 1 : 0, c->stream));
      if ((rc = read_eig(s, rd, lam, true))) return rc;
      if (rd.state[0] == 1) {
        c->oa_done++;
        done = true;
        break;
      }
      c->oa_fallback++;  // the refinement failed: sweeps to convergence
      handoff = false;
      try_clusters = false;
      continue;
    }
    if (try_clusters) {
      done = cluster_route(s, s->A[0], s->A[1], false, scale, rd, lam, rc) != 0;
      if (rc) return rc;
      if (done) c->oa_done++;
      else if (rd.state[0] != 3) try_clusters = false;
    }
  }
  c->eig_calls++;
  pe.stop();
  if (!done)
    return set_err(c, ITP_E_EIGENSOLVER, "Jacobi eigensolver did not converge in " + std::to_string(sweeps) + " sweeps");
  s->lam_host = lam;
  if (!check_positive) {
    for (int n = 0; n < N; n++)
      if (!std::isfinite(lam[n])) return set_err(c, ITP_E_EIGENSOLVER, "non-finite eigenvalue");
    return ITP_OK;
  }
  return check_eigenvalues(c, lam, bad_n, bad_lambda);
}

// Split-K of the overlap products for a requested chunk count (at most the
// buffer's): FP64 chunks of whole 16-element groups, bf16x3 chunks of whole
// 512-element groups.
void set_gram_chunks(itp_stateset* s, long long nchunk) {
  const long long M = s->grid->g.M;
  nchunk = std::max<long long>(1, std::min<long long>(nchunk, s->nchunk_max));
  s->kchunk = ((M + nchunk - 1) / nchunk + 15) / 16 * 16;
  s->nchunk = (int)((M + s->kchunk - 1) / s->kchunk);
  s->kchunk32 = ((M + nchunk - 1) / nchunk + 511) / 512 * 512;
  s->nchunk32 = (int)((M + s->kchunk32 - 1) / s->kchunk32);
}

// The autotuner of the overlap's split-K: on the first overlap of a shape
// (precision, grid points, padded states) it times the formula's chunk count
// (sized for full last waves) and 3/4, 1/2 and 1/4 of it, two launches each,
// and keeps the fastest for every later set of that shape. launches() runs one
// overlap of the set's current buffer; it always runs last with the chosen
// count (also when the shape was tuned before), so the overlap it leaves is
// the one the caller asked for.
template <class F>
int tune_gram(itp_stateset* s, int esz, F&& launches) {
  itp_ctx* c = s->grid->ctx;
  s->gram_tuned = true;
  const std::array<long long, 3> key = {esz, s->grid->g.M, s->nc};
  {
    std::lock_guard<std::mutex> lk(gram_wisdom().m);
    auto it = gram_wisdom().w.find(key);
    if (it != gram_wisdom().w.end()) {
      set_gram_chunks(s, it->second);
      HIPCK(c, launches());
      return ITP_OK;
    }
  }
  const long long n0 = s->nchunk_max;
  std::vector<long long> cand = {n0, (3 * n0 + 3) / 4, (n0 + 1) / 2, (n0 + 3) / 4};
  std::sort(cand.begin(), cand.end());
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  hipEvent_t e0, e1;
  HIPCK(c, hipEventCreate(&e0));
  HIPCK(c, hipEventCreate(&e1));
  long long best = n0;
  float best_ms = 1e30f;
  for (long long n : cand) {
    set_gram_chunks(s, n);
    float ms = 1e30f;
    for (int rep = 0; rep < 2; rep++) {
      HIPCK(c, hipEventRecord(e0, c->stream));
      HIPCK(c, launches());
      HIPCK(c, hipEventRecord(e1, c->stream));
      HIPCK(c, hipEventSynchronize(e1));
      float t = 0;
      HIPCK(c, hipEventElapsedTime(&t, e0, e1));
      ms = std::min(ms, t);
    }
    if (ms < best_ms) { best_ms = ms; best = n; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_gram_chunks(s, best);
  {
    std::lock_guard<std::mutex> lk(gram_wisdom().m);
    gram_wisdom().w[key] = (int)best;
  }
  c->tunings++;
  HIPCK(c, launches());
  return ITP_OK;
}

template <class R>
int gram_impl(itp_stateset* s) {
  itp_ctx* c = s->grid->ctx;
  const GridInfo& g = s->grid->g;
  Phase po(c, ITP_PHASE_OVERLAP);
  auto launches = [&]() -> hipError_t {
    // the non-finite flag is cleared by a kernel (graph-safe)
    hipError_t e = launch_zero_ints(s->bad, 1, c->stream);
    if (e != hipSuccess) return e;
    if (sizeof(R) == 4) {
      const cplx<float>* psi = (const cplx<float>*)s->buf[s->cur];
      return launch_gram_bf16x3(psi, psi, g.M, (int)s->N, s->nc, s->part, s->nchunk32, s->kchunk32, g.dx, s->A[0],
                                s->np, s->bad, c->stream);
    }
    return launch_gram<R>((const cplx<R>*)s->buf[s->cur], (const cplx<R>*)s->buf[s->cur], g.M, (int)s->N, s->nc,
                          s->part, s->nchunk, s->kchunk, g.dx, s->A[0], s->np, s->bad, c->stream);
  };
  if (c->autotune && !s->gram_tuned) {
    // the tuning's last launch is this call's overlap
    int rc = tune_gram(s, (int)sizeof(R), launches);
    po.stop();
    return rc;
  }
  HIPCK(c, c->graphs ? c->gcache.run(c->stream, {s->id, (uint64_t)s->cur, 1, 0}, launches) : launches());
  po.stop();
  return ITP_OK;
}

template <class R>
int rotate_set(itp_stateset* s, const cplx<R>* in, cplx<R>* out, bool upper = false);

template <class R>
int ortho_impl(itp_stateset* s, size_t* bad_n, double* bad_lambda) {
  itp_ctx* c = s->grid->ctx;
  int rc = gram_impl<R>(s);
  if (rc) return rc;
  rc = eigen_stage(s, bad_n, bad_lambda, true);
  if (rc) return rc;
  Phase pl(c, ITP_PHASE_LINCOMB);
  rc = rotate_set<R>(s, (const cplx<R>*)s->buf[s->cur], (cplx<R>*)s->buf[s->cur ^ 1]);
  pl.stop();
  if (rc) return rc;
  s->cur ^= 1;
  return ITP_OK;
}

// Cholesky orthonormalization (the CholeskyQR step of the mixed path; no
// counterpart in the reference): S = R^H R from the same Gram kernels,
// Psi <- Psi R^-1 with the same rotation kernels. When the factorization
// meets a pivot below the precision's threshold relative to its diagonal
// (nearly dependent states), the eigendecomposition route of ortho_impl runs
// on the same overlap matrix instead and *used_eigen is set.
template <class R>
int cholqr_impl(itp_stateset* s, int* used_eigen, size_t* bad_n, double* bad_lambda) {
  itp_ctx* c = s->grid->ctx;
  if (used_eigen) *used_eigen = 0;
  int rc = gram_impl<R>(s);
  if (rc) return rc;
  int fail[2] = {0, 0};
  {
    Phase pe(c, ITP_PHASE_EIGENSOLVE);
    // FP32 fields give an overlap with ~1e-7 relative errors: pivots below
    // 1e-5 of their diagonal would be dominated by them
    const double tol = c->chol_tol > 0 ? c->chol_tol : (sizeof(R) == 4 ? 1e-5 : 1e-12);
    auto launches = [&]() -> hipError_t {
      hipError_t e = launch_cholinv(s->A[0], s->np, (int)s->N, s->A[1], s->diag0, s->C, s->nc, s->chol_fail, tol,
                                    c->stream);
      if (e != hipSuccess) return e;
      return launch_chol_diag_range(s->A[1], s->np, (int)s->N, s->chol_mm, c->stream);
    };
    uint64_t tbits;
    std::memcpy(&tbits, &tol, sizeof(tbits));
    HIPCK(c, c->graphs ? c->gcache.run(c->stream, {s->id, tbits, 2, 0}, launches) : launches());
    std::vector<double> mm(2 + s->N, 0.0);
    Downloads dl(c);
    HIPCK(c, dl.add(&fail[0], s->chol_fail, sizeof(int)));
    HIPCK(c, dl.add(&fail[1], s->bad, sizeof(int)));
    HIPCK(c, dl.add(mm.data(), s->chol_mm, mm.size() * sizeof(double)));
    HIPCK(c, dl.wait());
    const bool ok = !(fail[0] || fail[1]) && mm[1] > 0;
    s->chol_kappa = ok ? mm[0] / mm[1] : 0.0;
    if (ok) s->chol_diag.assign(mm.begin() + 2, mm.end());
    else s->chol_diag.clear();
  }
  if (fail[0] || fail[1]) {
    if (used_eigen) *used_eigen = 1;
    if ((rc = eigen_stage(s, bad_n, bad_lambda, true))) return rc;
  }
  Phase pl(c, ITP_PHASE_LINCOMB);
  // R^-1 is upper triangular (the eigendecomposition's C is not)
  rc = rotate_set<R>(s, (const cplx<R>*)s->buf[s->cur], (cplx<R>*)s->buf[s->cur ^ 1], !(fail[0] || fail[1]));
  pl.stop();
  if (rc) return rc;
  s->cur ^= 1;
  return ITP_OK;
}

// out = in C for N fields of the set's precision: FP64 MFMA for FP64 sets,
// FP32 MFMA (with an FP32 copy of C) for FP32 sets.
template <class R>
int rotate_set(itp_stateset* s, const cplx<R>* in, cplx<R>* out, bool upper) {
  itp_ctx* c = s->grid->ctx;
  const GridInfo& g = s->grid->g;
  if (sizeof(R) == 4) {
    HIPCK(c, launch_convert(s->C, s->C32, (long long)s->nc * s->nc, 1, c->stream));
    if (c->fp32_rotation == 1)
      HIPCK(c, launch_rotate_bf16x3((const cplx<float>*)in, (cplx<float>*)out, s->C32, s->nc, g.M, (int)s->N,
                                    c->stream, upper ? 1 : 0));
    else
      HIPCK(c, launch_rotate_f32((const cplx<float>*)in, (cplx<float>*)out, s->C32, s->nc, g.M, (int)s->N,
                                 c->stream));
    return ITP_OK;
  }
  HIPCK(c, launch_rotate<R>(in, out, s->C, s->nc, g.M, (int)s->N, c->stream));
  return ITP_OK;
}

// Rayleigh-Ritz on the set (the mixed path's final pass; no counterpart in the
// reference): orthonormalize as above, W = H Psi, H-matrix Psi^H W, its
// eigenvectors U and ascending eigenvalues E (the Ritz values), Psi <- Psi U,
// and the residual norms sigma_n = ||H psi_n - E_n psi_n|| (the energy
// deviation: for a normalised psi_n with E_n = <psi_n|H|psi_n> it equals
// sqrt(<H^2> - <H>^2), operators.cpp:40-50).
template <class R>
int rr_impl(itp_hamiltonian* h, itp_stateset* s, double* E, double* sigma) {
  itp_grid* gr = s->grid;
  itp_ctx* c = gr->ctx;
  const GridInfo& g = gr->g;
  const int N = (int)s->N;
  int rc = ortho_impl<R>(s, nullptr, nullptr);
  if (rc) return rc;
  const std::vector<double> overlap_lam = s->lam_host;
  const size_t fbytes = s->N * (size_t)g.M * sizeof(cplx<R>);
  if ((rc = ensure_ws(c, 2 * fbytes))) return rc;
  cplx<R>* W = (cplx<R>*)c->ws;
  cplx<R>* T = (cplx<R>*)((char*)c->ws + fbytes);
  const cplx<R>* psi = (const cplx<R>*)s->buf[s->cur];
  {
    Phase ph(c, ITP_PHASE_ENERGIES);
    const Branch br = branch_of(h->B, g.bt);
    const double* V = h->pot->null ? nullptr : h->pot->dV;
    const LineGeom& lg = gr->lg[sizeof(R) == 4];
    std::vector<Sweep<R>> prog;
    hamiltonian_program<R>(g, lg, br, h->d, V, psi, T, W, nullptr, nullptr, prog);
    GpuBackend be;
    be.st = c->stream;
    run_program<R>(be, g, lg, prog, N);
    if (be.err != hipSuccess) return set_err(c, ITP_E_HIP, std::string("rayleigh-ritz: ") + hipGetErrorString(be.err));
  }
  {
    // the magnetic Dirichlet kinetic operator of the reference is not
    // Hermitian (kinetic.cpp:99-142): use the Hermitian part of the H-matrix,
    // whose expectation values are the energies Re<psi|H psi>
    const int full = branch_of(h->B, g.bt) == BR_BD ? 1 : 0;
    Phase po(c, ITP_PHASE_OVERLAP);
    HIPCK(c, hipMemsetAsync(s->bad, 0, sizeof(int), c->stream));
    if (sizeof(R) == 4)
      HIPCK(c, launch_gram_bf16x3((const cplx<float>*)psi, (const cplx<float>*)W, g.M, N, s->nc, s->part, s->nchunk32,
                                  s->kchunk32, g.dx, s->A[0], s->np, s->bad, c->stream, full));
    else
      HIPCK(c, launch_gram<R>(psi, W, g.M, N, s->nc, s->part, s->nchunk, s->kchunk, g.dx, s->A[0], s->np, s->bad,
                              c->stream, full));
  }
  rc = eigen_stage(s, nullptr, nullptr, false, false);
  s->lam_host = overlap_lam;
  if (rc) return rc;
  {
    Phase pl(c, ITP_PHASE_LINCOMB);
    if ((rc = rotate_set<R>(s, psi, (cplx<R>*)s->buf[s->cur ^ 1]))) return rc;
    if ((rc = rotate_set<R>(s, W, T))) return rc;
  }
  s->cur ^= 1;
  HIPCK(c, launch_residual<R>(T, (const cplx<R>*)s->buf[s->cur], s->lam, g.M, N, g.dx, s->dE + s->N, c->stream));
  Downloads dl(c);
  if (E) HIPCK(c, dl.add(E, s->lam, N * sizeof(double)));
  if (sigma) HIPCK(c, dl.add(sigma, s->dE + s->N, N * sizeof(double)));
  HIPCK(c, dl.wait());
  return ITP_OK;
}

uint64_t splitmix64(uint64_t& st) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// temporary device copy of host fields for the single-operator hooks
struct DevFields {
  itp_ctx* c = nullptr;
  cplx<double>* p = nullptr;
  cplx<double>* w = nullptr;
  ~DevFields() {
    if (c) {
      dfree(c, p);
      dfree(c, w);
    }
  }
};

// a temporary block of dalloc, returned when the scope ends
struct DevHold {
  itp_ctx* c;
  void* p;
  ~DevHold() { dfree(c, p); }
};

int hook_begin(itp_ctx* c, const double* host, size_t nelem, DevFields& df, bool need_work) {
  df.c = c;
  int rc = dalloc(c, (void**)&df.p, nelem * 16);
  if (rc) return rc;
  if (need_work) {
    rc = dalloc(c, (void**)&df.w, nelem * 16);
    if (rc) return rc;
  }
  HIPCK(c, cmemcpy(c, df.p, host, nelem * 16, hipMemcpyHostToDevice));
  return ITP_OK;
}

int hook_end(itp_ctx* c, double* host, size_t nelem, DevFields& df, hipError_t err) {
  if (err != hipSuccess) return set_err(c, ITP_E_HIP, hipGetErrorString(err));
  HIPCK(c, hipStreamSynchronize(c->stream));
  HIPCK(c, cmemcpy(c, host, df.p, nelem * 16, hipMemcpyDeviceToHost));
  return ITP_OK;
}

}  // namespace

extern "C" {

int itp_ctx_create(int device, itp_ctx** out) {
  if (!out) return ITP_E_INVALID;
  *out = nullptr;
  std::unique_ptr<itp_ctx> c(new itp_ctx);
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return ITP_E_HIP;
  // non-blocking: synchronous copies of other threads' contexts (on the null
  // stream) do not join this stream, so a graph capture on it is not broken
  // by them; every host read of this stream's results synchronises it first
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return ITP_E_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return ITP_E_HIP;
  c->num_cus = prop.multiProcessorCount;
  if (scratch_address(device, &c->scratch) != hipSuccess) return ITP_E_HIP;
  c->evpool.resize(2048);
  for (auto& e : c->evpool)
    if (hipEventCreate(&e) != hipSuccess) return ITP_E_HIP;
  // the downloads' pinned buffer (without it, downloads go straight to the caller's memory)
  if (void* pin = take_pinned()) {
    c->pin = static_cast<char*>(pin);
    c->pin_cap = kPinBytes;
  }
  *out = c.release();
  return ITP_OK;
}

void itp_ctx_destroy(itp_ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  c->gcache.clear();
  if (c->graphs) {
    c->graphs = false;
    graph_contexts()--;
  }
  ctx_release(c);
}

const char* itp_last_error(const itp_ctx* c) { return c ? c->err.c_str() : "no context"; }

int itp_ctx_synchronize(itp_ctx* c) {
  HIPCK(c, hipStreamSynchronize(c->stream));
  return ITP_OK;
}

int itp_ctx_num_cus(const itp_ctx* c) { return c ? c->num_cus : 0; }

int itp_ctx_enable_timing(itp_ctx* c, int enable) {
  if (!c) return ITP_E_INVALID;
  c->timing = enable != 0;
  return ITP_OK;
}
int itp_ctx_phase_times(itp_ctx* c, double* ms, long long* counts) {
  if (!c) return ITP_E_INVALID;
  flush_events(c);
  for (int i = 0; i < ITP_NUM_PHASES; i++) {
    if (ms) ms[i] = c->ms[i];
    if (counts) counts[i] = c->cnt[i];
  }
  return ITP_OK;
}
int itp_ctx_set_graphs(itp_ctx* c, int on) {
  if (!c) return ITP_E_INVALID;
  const bool was = c->graphs;
  c->graphs = on != 0;
  if (c->graphs && !was) graph_contexts()++;
  if (!c->graphs && was) {
    (void)hipStreamSynchronize(c->stream);
    c->gcache.clear();
    graph_contexts()--;
  }
  return ITP_OK;
}

int itp_ctx_set_autotune(itp_ctx* c, int on) {
  if (!c) return ITP_E_INVALID;
  c->autotune = on != 0;
  return ITP_OK;
}

int itp_ctx_wisdom_export(const itp_ctx* c, char* out, size_t cap, size_t* len) {
  if (!c || !len) return ITP_E_INVALID;
  std::string t;
  std::lock_guard<std::mutex> lk(gram_wisdom().m);
  for (const auto& kv : gram_wisdom().w)
    t += "gram " + std::to_string(kv.first[0]) + " " + std::to_string(kv.first[1]) + " " +
         std::to_string(kv.first[2]) + " " + std::to_string(kv.second) + "\n";
  *len = t.size();
  if (out && cap > 0) {
    const size_t n = std::min(cap - 1, t.size());
    std::memcpy(out, t.data(), n);
    out[n] = 0;
  }
  return ITP_OK;
}

int itp_ctx_wisdom_import(itp_ctx* c, const char* text) {
  if (!c || !text) return ITP_E_INVALID;
  std::map<std::array<long long, 3>, int> w;
  const char* p = text;
  while (*p) {
    char kind[16] = {0};
    long long a = 0, b = 0, d = 0;
    int n = 0, used = 0;
    if (std::sscanf(p, "%15s %lld %lld %lld %d%n", kind, &a, &b, &d, &n, &used) != 5 || std::strcmp(kind, "gram") ||
        (a != 4 && a != 8) || b < 1 || d < 64 || d % 64 || n < 1)
      return set_err(c, ITP_E_INVALID, "wisdom: malformed line");
    w[{a, b, d}] = n;
    p += used;
    while (*p == '\n' || *p == ' ' || *p == '\r' || *p == '\t') p++;
  }
  std::lock_guard<std::mutex> lk(gram_wisdom().m);
  for (const auto& kv : w) gram_wisdom().w[kv.first] = kv.second;
  return ITP_OK;
}

int itp_ctx_tunings(const itp_ctx* c) { return c ? (int)c->tunings : 0; }

int itp_ctx_set_eig_check_min(itp_ctx* c, int n) {
  if (!c || n < 1) return ITP_E_INVALID;
  c->td_min_n = n;
  return ITP_OK;
}

int itp_ctx_eig_check_stats(itp_ctx* c, long long* checks, long long* rejected) {
  if (!c) return ITP_E_INVALID;
  if (checks) *checks = c->td_checks;
  if (rejected) *rejected = c->td_rejects;
  return ITP_OK;
}

int itp_ctx_eig_stats(itp_ctx* c, long long* calls, long long* sweeps) {
  if (!c) return ITP_E_INVALID;
  if (calls) *calls = c->eig_calls;
  if (sweeps) *sweeps = c->eig_sweeps;
  return ITP_OK;
}

int itp_ctx_set_eig_inner(itp_ctx* c, int inner) {
  if (!c || inner < 1 || inner > 64) return ITP_E_INVALID;
  c->eig_inner = inner;
  return ITP_OK;
}

int itp_ctx_set_eig_mode(itp_ctx* c, int mode) {
  if (!c || mode < 0 || mode > 1) return ITP_E_INVALID;
  c->eig_mode = mode;
  return ITP_OK;
}

int itp_ctx_set_eig_handoff(itp_ctx* c, double ratio) {
  if (!c || !(ratio > 0) || !(ratio < 1e3)) return ITP_E_INVALID;
  c->eig_handoff = ratio;
  return ITP_OK;
}

int itp_ctx_oa_stats(itp_ctx* c, long long* refined, long long* fallbacks) {
  if (!c) return ITP_E_INVALID;
  if (refined) *refined = c->oa_done;
  if (fallbacks) *fallbacks = c->oa_fallback;
  return ITP_OK;
}

int itp_ctx_set_eig_cluster(itp_ctx* c, double tau) {
  if (!c || !(tau >= 0) || !(tau < 0.5)) return ITP_E_INVALID;
  c->eig_cluster = tau;
  return ITP_OK;
}

int itp_ctx_cluster_stats(itp_ctx* c, long long* out) {
  if (!c || !out) return ITP_E_INVALID;
  out[0] = c->cl_calls;
  out[1] = c->cl_clusters;
  out[2] = c->cl_largest;
  out[3] = c->cl_fallback;
  return ITP_OK;
}

int itp_ctx_set_fp32_rotation(itp_ctx* c, int kind) {
  if (!c || kind < 0 || kind > 1) return ITP_E_INVALID;
  c->fp32_rotation = kind;
  return ITP_OK;
}

int itp_ctx_set_eig_spin_limit(itp_ctx* c, unsigned long long polls) {
  if (!c || polls > 0xffffffffull) return ITP_E_INVALID;
  c->eig_spin_limit = (unsigned)polls;
  return ITP_OK;
}

int itp_ctx_set_cholesky_tol(itp_ctx* c, double tol) {
  if (!c || !(tol >= 0) || !(tol < 1)) return ITP_E_INVALID;
  c->chol_tol = tol;
  return ITP_OK;
}

int itp_ctx_set_eig_grid(itp_ctx* c, int workgroups) {
  if (!c || workgroups < 0 || workgroups > 4096) return ITP_E_INVALID;
  c->eig_grid = workgroups;
  return ITP_OK;
}

int itp_ctx_reset_phase_times(itp_ctx* c) {
  if (!c) return ITP_E_INVALID;
  flush_events(c);
  for (int i = 0; i < ITP_NUM_PHASES; i++) { c->ms[i] = 0; c->cnt[i] = 0; }
  c->eig_calls = 0;
  c->eig_sweeps = 0;
  return ITP_OK;
}

int itp_grid_create(itp_ctx* c, size_t sx, size_t sy, double dx, int boundary, itp_grid** out) {
  if (!c || !out || sx < 2 || sy < 2 || !(dx > 0) || (boundary != ITP_PERIODIC && boundary != ITP_DIRICHLET))
    return set_err(c, ITP_E_INVALID, "itp_grid_create: bad arguments");
  if (sx > 4096 || sy > 4096) return set_err(c, ITP_E_UNSUPPORTED, "grid sides above 4096 are not supported");
  return guarded(c, [&]() -> int {
    std::unique_ptr<itp_grid> gr(new itp_grid);
    gr->ctx = c;
    gr->g = make_grid((int)sx, (int)sy, dx, boundary);
    LinePlan px, py;
    if (!make_plan((int)sx, px) || !make_plan((int)sy, py))
      return set_err(c, ITP_E_UNSUPPORTED, "grid sides must be 2^a 3^b 5^c with an even radix-stage count");
    std::vector<cplx<double>> twx, mkx, twy, mky;
    twiddle_tables((int)sx, twx, mkx);
    twiddle_tables((int)sy, twy, mky);
    std::vector<cplx<double>> all;
    all.insert(all.end(), twx.begin(), twx.end());
    all.insert(all.end(), mkx.begin(), mkx.end());
    all.insert(all.end(), twy.begin(), twy.end());
    all.insert(all.end(), mky.begin(), mky.end());
    int rc = dalloc(c, (void**)&gr->d_tw, all.size() * 16);
    if (rc) return rc;
    HIPCK(c, cmemcpy(c, gr->d_tw, all.data(), all.size() * 16, hipMemcpyHostToDevice));
    gr->lg[0] = make_geom(gr->g, gr->d_tw, 16);
    gr->lg[1] = make_geom(gr->g, gr->d_tw, 8);
    c->refs++;
    *out = gr.release();
    return ITP_OK;
  });
}

void itp_grid_destroy(itp_grid* g) {
  if (!g) return;
  itp_ctx* c = g->ctx;
  if (g->d_tw) dfree(c, g->d_tw);
  delete g;
  ctx_release(c);
}

int itp_grid_kvalues(const itp_grid* g, double* fkx, double* fky, double* dkx, double* dky) {
  if (!g) return ITP_E_INVALID;
  if (fkx) std::copy(g->g.fft_kx.begin(), g->g.fft_kx.end(), fkx);
  if (fky) std::copy(g->g.fft_ky.begin(), g->g.fft_ky.end(), fky);
  if (dkx) std::copy(g->g.dsct_kx.begin(), g->g.dsct_kx.end(), dkx);
  if (dky) std::copy(g->g.dsct_ky.begin(), g->g.dsct_ky.end(), dky);
  return ITP_OK;
}

int itp_grid_positions(const itp_grid* g, double* posx, double* posy) {
  if (!g) return ITP_E_INVALID;
  if (posx) std::copy(g->g.posx.begin(), g->g.posx.end(), posx);
  if (posy) std::copy(g->g.posy.begin(), g->g.posy.end(), posy);
  return ITP_OK;
}

double itp_grid_normalization_factor(const itp_grid* g, int kind) {
  if (!g || kind < 0 || kind > ITP_iDCTy) return NAN;
  const GridInfo& gi = g->g;
  const int k = kind / 2;  // pairs (FFT,iFFT), (FFTx,iFFTx), ...
  switch (k) {
    case 0: return gi.nFFT;
    case 1: return gi.nFFTx;
    case 2: return gi.nFFTy;
    case 3: case 6: return gi.nDSCT;
    case 4: case 7: return gi.nDSCTx;
    default: return gi.nDSCTy;
  }
}

int itp_transform_device(itp_grid* g, void* fields, size_t nfields, int kind) {
  if (!g || !fields) return ITP_E_INVALID;
  itp_ctx* c = g->ctx;
  return guarded(c, [&]() -> int {
    std::vector<Sweep<double>> prog;
    if (!transform_program<double>(kind, (cplx<double>*)fields, prog)) return set_err(c, ITP_E_INVALID, "transform kind");
    GpuBackend be;
    be.st = c->stream;
    run_program<double>(be, g->g, g->lg[0], prog, (int)nfields);
    if (be.err != hipSuccess) return set_err(c, ITP_E_HIP, hipGetErrorString(be.err));
    return ITP_OK;
  });
}

int itp_transform_host(itp_grid* g, double* fields, size_t nfields, int kind) {
  if (!g || !fields) return ITP_E_INVALID;
  itp_ctx* c = g->ctx;
  DevFields df;
  const size_t n = nfields * (size_t)g->g.M;
  int rc = hook_begin(c, fields, n, df, false);
  if (rc) return rc;
  rc = itp_transform_device(g, df.p, nfields, kind);
  if (rc) return rc;
  return hook_end(c, fields, n, df, hipSuccess);
}

int itp_potential_create(itp_grid* g, const double* V, itp_potential** out) {
  if (!g || !out) return ITP_E_INVALID;
  itp_ctx* c = g->ctx;
  return guarded(c, [&]() -> int {
    std::unique_ptr<itp_potential> p(new itp_potential);
    p->grid = g;
    p->null = (V == nullptr);
    if (V) {
      p->V.assign(V, V + g->g.M);
      int rc = dalloc(c, (void**)&p->dV, g->g.M * sizeof(double));
      if (rc) return rc;
      HIPCK(c, cmemcpy(c, p->dV, V, g->g.M * sizeof(double), hipMemcpyHostToDevice));
    }
    p->ctx = c;
    c->refs++;
    *out = p.release();
    return ITP_OK;
  });
}

int itp_potential_type_params(const char* name, const double* params, size_t nparams, int* kind, double* resolved) {
  if (!name || (nparams && !params)) return ITP_E_INVALID;
  PotParams q;
  if (pot_resolve(name, params, nparams, q)) return ITP_E_INVALID;
  if (kind) *kind = q.kind;
  if (resolved) std::copy(q.p, q.p + kPotMaxParams, resolved);
  return ITP_OK;
}

int itp_potential_create_type(itp_grid* g, const char* name, const double* params, size_t nparams,
                              itp_potential** out) {
  if (!g || !out || !name || (nparams && !params)) return ITP_E_INVALID;
  itp_ctx* c = g->ctx;
  PotParams q;
  if (const char* why = pot_resolve(name, params, nparams, q)) return set_err(c, ITP_E_INVALID, why);
  return guarded(c, [&]() -> int {
    std::unique_ptr<itp_potential> p(new itp_potential);
    p->grid = g;
    // the zero potential without impurities is the null potential (potential.cpp:28-40)
    p->null = q.kind == kPotZero;
    p->V.assign(g->g.M, 0.0);
    int rc = dalloc(c, (void**)&p->dV, g->g.M * sizeof(double));
    if (rc) return rc;
    HIPCK(c, launch_potential(q, g->g.sx, g->g.sy, g->g.dx, p->dV, c->stream));
    HIPCK(c, hipMemcpyAsync(p->V.data(), p->dV, g->g.M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    p->ctx = c;
    c->refs++;
    *out = p.release();
    return ITP_OK;
  });
}

int itp_potential_add_impurities(itp_potential* p, int kind, const double* data, size_t count,
                                 const double* kind_params) {
  if (!p || kind < 0 || kind >= kImpNumKinds || (count && !data)) return ITP_E_INVALID;
  if ((kind == kImpCoulomb || kind == kImpHemisphere) && !kind_params) return ITP_E_INVALID;
  itp_grid* g = p->grid;
  itp_ctx* c = g->ctx;
  return guarded(c, [&]() -> int {
    const GridInfo& gi = g->g;
    if (!p->dV) {
      int rc = dalloc(c, (void**)&p->dV, gi.M * sizeof(double));
      if (rc) return rc;
      p->V.assign(gi.M, 0.0);
      HIPCK(c, cmemcpy(c, p->dV, p->V.data(), gi.M * sizeof(double), hipMemcpyHostToDevice));
    }
    if (kind == kImpDelta) {
      // one grid point each (DeltaImpurities::add_noise, noise.cpp:80-88): on the host copy
      HIPCK(c, hipMemcpyAsync(p->V.data(), p->dV, gi.M * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      HIPCK(c, hipStreamSynchronize(c->stream));
      // all positions are checked before any value changes, so a rejected
      // call leaves the host and device copies equal
      std::vector<long long> at(count);
      for (size_t n = 0; n < count; n++) {
        const long long x = grid_nearest(data[3 * n], gi.sx, gi.dx), y = grid_nearest(data[3 * n + 1], gi.sy, gi.dx);
        if (x < 0 || y < 0) return set_err(c, ITP_E_INVALID, "delta impurity outside the grid");
        at[n] = y * (long long)gi.sx + x;
      }
      for (size_t n = 0; n < count; n++) p->V[at[n]] += data[3 * n + 2];
      HIPCK(c, cmemcpy(c, p->dV, p->V.data(), gi.M * sizeof(double), hipMemcpyHostToDevice));
      if (count) p->null = false;
      return ITP_OK;
    }
    const int stride = kind == kImpGaussian ? 4 : 2;
    double g0 = 0, g1 = 0;
    if (kind == kImpCoulomb) { g0 = kind_params[0] / 2; g1 = kind_params[1]; }     // exponent e, prefactor alpha
    if (kind == kImpHemisphere) { g0 = kind_params[0]; g1 = kind_params[1] * kind_params[1]; }  // amplitude, radius
    double* dd = nullptr;
    int rc = dalloc(c, (void**)&dd, std::max<size_t>(1, count) * stride * sizeof(double));
    if (rc) return rc;
    hipError_t e = cmemcpy(c, dd, data, count * stride * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_impurities(kind, dd, stride, (long long)count, g0, g1, gi.sx, gi.sy, gi.dx, p->dV,
                                               c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(p->V.data(), p->dV, gi.M * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, dd);
    if (e != hipSuccess) return set_err(c, ITP_E_HIP, std::string("impurities: ") + hipGetErrorString(e));
    if (count) p->null = false;
    return ITP_OK;
  });
}

int itp_potential_values(const itp_potential* p, double* V) {
  if (!p || !V) return ITP_E_INVALID;
  if (p->V.size() != p->grid->g.M) std::fill(V, V + p->grid->g.M, 0.0);
  else std::copy(p->V.begin(), p->V.end(), V);
  return ITP_OK;
}

void itp_potential_destroy(itp_potential* p) {
  if (!p) return;
  itp_ctx* c = p->ctx;  // parents may already be gone
  if (p->dV) dfree(c, p->dV);
  delete p;
  ctx_release(c);
}

}  // extern "C"

namespace {

// Device buffers of a set (every one of them may be null).
void free_stateset_buffers(itp_ctx* c, itp_stateset* s) {
  void* ptrs[] = {s->buf[0], s->buf[1], s->part, s->A[0], s->A[1], s->V[0], s->V[1], s->Vinit, s->bjQ, s->bjD,
                  s->bjslots, s->bjident, s->oa_work, s->oa_lam, s->oa_slots, s->oa_state, s->cl_hdr, s->cl_pairs,
                  s->cl_members, s->cl_offsets, s->Vsave, s->oa_part, s->C, s->C32,
                  s->lam, s->perm, s->info, s->bad, s->epart, s->dE, s->dmean, s->diag0, s->chol_fail, s->chol_mm, s->norms,
                  s->td_scratch};
  for (void* p : ptrs) dfree(c, p);
}

}  // namespace

extern "C" {

int itp_stateset_create(itp_grid* g, size_t N, int precision, itp_stateset** out) {
  if (!g || !out || N < 1 || (precision != ITP_FP64 && precision != ITP_FP32)) return ITP_E_INVALID;
  itp_ctx* c = g->ctx;
  if (N > 4096) return set_err(c, ITP_E_UNSUPPORTED, "more than 4096 states per set is not supported");
  if (precision == ITP_FP32 && (g->g.M % 2))
    return set_err(c, ITP_E_UNSUPPORTED, "FP32 sets need an even number of grid points (sx * sy)");
  if (N > 65535) return set_err(c, ITP_E_UNSUPPORTED, "too many states");
  return guarded(c, [&]() -> int {
    std::unique_ptr<itp_stateset> s(new itp_stateset);
    // a creation that fails part-way returns the buffers it took
    struct Undo {
      itp_ctx* c;
      std::unique_ptr<itp_stateset>& s;
      ~Undo() {
        if (s) free_stateset_buffers(c, s.get());
      }
    } undo{c, s};
    s->grid = g;
    s->N = N;
    s->prec = precision;
    s->esz = precision == ITP_FP64 ? 16 : 8;
    const size_t M = g->g.M;
    int rc;
    for (int b = 0; b < 2; b++) {
      if ((rc = dalloc(c, &s->buf[b], N * M * s->esz))) return rc;
      HIPCK(c, cmemset(c, s->buf[b], 0, N * M * s->esz));
    }
    s->nc = (int)((N + 63) / 64 * 64);
    s->np = s->nc;  // the eigensolver works on the zero-padded nc x nc matrix
    const int nb = s->nc / 64;
    const int npairs = nb * (nb + 1) / 2;
    // split-K chunks: 2048 to 4096 workgroups, the count for which the last
    // wave of resident workgroups (two per CU) is fullest
    long long nchunk = 1;
    {
      const long long Q = 2LL * std::max(1, c->num_cus);
      const long long lo = std::max(1, 2048 / npairs), hi = std::max<long long>(lo, (4096 + npairs - 1) / npairs);
      double best = -1;
      for (long long n = lo; n <= hi; n++) {
        const long long W = n * npairs;
        const double eff = (double)W / (double)(((W + Q - 1) / Q) * Q);
        if (eff > best + 1e-12) { best = eff; nchunk = n; }
      }
    }
    nchunk = std::min<long long>(nchunk, std::max<long long>(1, (long long)M / 256));
    const size_t per_chunk = (size_t)s->nc * s->nc * 16;
    nchunk = std::max<long long>(1, std::min<long long>(nchunk, (long long)((512ull << 20) / per_chunk)));
    s->kchunk = ((long long)M + nchunk - 1) / nchunk;
    s->kchunk = (s->kchunk + 15) / 16 * 16;
    s->nchunk = (int)(((long long)M + s->kchunk - 1) / s->kchunk);
    // bf16x3 Gram: chunks of whole 512-element groups, no more chunks than above
    s->kchunk32 = ((long long)M + nchunk - 1) / nchunk;
    s->kchunk32 = (s->kchunk32 + 511) / 512 * 512;
    s->nchunk32 = (int)(((long long)M + s->kchunk32 - 1) / s->kchunk32);
    s->nchunk_max = s->nchunk;
    const size_t np2 = (size_t)s->np * s->np * 16;
    if ((rc = dalloc(c, (void**)&s->part, (size_t)s->nchunk * per_chunk))) return rc;
    for (int b = 0; b < 2; b++) {
      if ((rc = dalloc(c, (void**)&s->A[b], np2))) return rc;
      if ((rc = dalloc(c, (void**)&s->V[b], np2))) return rc;
      HIPCK(c, cmemset(c, s->A[b], 0, np2));
    }
    if ((rc = dalloc(c, (void**)&s->Vinit, np2))) return rc;
    {
      const size_t qbytes = (size_t)(s->np / JS) * JS2 * 16;
      if ((rc = dalloc(c, (void**)&s->bjQ, qbytes))) return rc;
      if ((rc = dalloc(c, (void**)&s->bjD, qbytes))) return rc;
      if ((rc = dalloc(c, (void**)&s->bjslots, 3 * (kMaxSweeps + 1) * sizeof(double)))) return rc;
      if ((rc = dalloc(c, (void**)&s->bjident, (size_t)(s->np / JS) * sizeof(int)))) return rc;
      if ((rc = dalloc(c, (void**)&s->oa_work, 5 * np2))) return rc;
      if ((rc = dalloc(c, (void**)&s->oa_lam, (size_t)s->np * sizeof(double)))) return rc;
      if ((rc = dalloc(c, (void**)&s->oa_slots, kOAMaxIter * sizeof(double)))) return rc;
      if ((rc = dalloc(c, (void**)&s->oa_state, 2 * sizeof(int)))) return rc;
      if ((rc = dalloc(c, (void**)&s->cl_hdr, 2 * sizeof(unsigned long long)))) return rc;
      if ((rc = dalloc(c, (void**)&s->cl_pairs, 2 * (size_t)kClusterPairCap * sizeof(int)))) return rc;
      if ((rc = dalloc(c, (void**)&s->cl_members, (size_t)s->np * sizeof(int)))) return rc;
      if ((rc = dalloc(c, (void**)&s->cl_offsets, ((size_t)s->np + 1) * sizeof(int)))) return rc;
      if ((rc = dalloc(c, (void**)&s->Vsave, np2))) return rc;
      if (s->np < 1024 && (rc = dalloc(c, (void**)&s->oa_part, 4 * np2))) return rc;
    }
    HIPCK(c, launch_identity(s->Vinit, s->np, c->stream));
    if ((rc = dalloc(c, (void**)&s->C, (size_t)s->nc * s->nc * 16))) return rc;
    if (precision == ITP_FP32 && (rc = dalloc(c, (void**)&s->C32, (size_t)s->nc * s->nc * 8))) return rc;
    HIPCK(c, cmemset(c, s->C, 0, (size_t)s->nc * s->nc * 16));
    if ((rc = dalloc(c, (void**)&s->lam, N * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&s->diag0, (size_t)s->np * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&s->chol_fail, sizeof(int)))) return rc;
    if ((rc = dalloc(c, (void**)&s->chol_mm, (2 + N) * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&s->norms, N * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&s->perm, N * sizeof(int)))) return rc;
    if ((rc = dalloc(c, (void**)&s->info, 8 * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&s->bad, sizeof(int)))) return rc;
    const size_t pstride = N * (size_t)std::max(g->g.sx, g->g.sy);
    if ((rc = dalloc(c, (void**)&s->epart, 2 * pstride * 16))) return rc;
    if ((rc = dalloc(c, (void**)&s->dE, 2 * N * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&s->dmean, 2 * N * 16))) return rc;
    s->ctx = c;
    c->refs++;
    *out = s.release();
    return ITP_OK;
  });
}

void itp_stateset_destroy(itp_stateset* s) {
  if (!s) return;
  itp_ctx* c = s->ctx;  // parents may already be gone
  free_stateset_buffers(c, s);
  delete s;
  ctx_release(c);
}

int itp_stateset_upload(itp_stateset* s, const double* host) {
  if (!s || !host) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  const size_t n = s->N * (size_t)s->grid->g.M;
  if (s->prec == ITP_FP64) {
    HIPCK(c, hipMemcpyAsync(s->buf[s->cur], host, n * 16, hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return ITP_OK;
  }
  // stage through the other buffer's space is too small (FP32); use the workspace
  int rc = ensure_ws(c, n * 16);
  if (rc) return rc;
  HIPCK(c, hipMemcpyAsync(c->ws, host, n * 16, hipMemcpyHostToDevice, c->stream));
  HIPCK(c, launch_convert(c->ws, s->buf[s->cur], (long long)n, 1, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return ITP_OK;
}

int itp_stateset_download(itp_stateset* s, double* host) {
  if (!s || !host) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  const size_t n = s->N * (size_t)s->grid->g.M;
  if (s->prec == ITP_FP64) {
    HIPCK(c, hipMemcpyAsync(host, s->buf[s->cur], n * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return ITP_OK;
  }
  int rc = ensure_ws(c, n * 16);
  if (rc) return rc;
  HIPCK(c, launch_convert(s->buf[s->cur], c->ws, (long long)n, 0, c->stream));
  HIPCK(c, hipMemcpyAsync(host, c->ws, n * 16, hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return ITP_OK;
}

void* itp_stateset_device_ptr(itp_stateset* s) { return s ? s->buf[s->cur] : nullptr; }
size_t itp_stateset_size(const itp_stateset* s) { return s ? s->N : 0; }

int itp_split_create(itp_potential* pot, int halforder, double eps, double B, itp_split** out) {
  if (!pot || !out || halforder < 1 || halforder > 12 || !(eps > 0) || !std::isfinite(B)) return ITP_E_INVALID;
  itp_ctx* c = pot->grid->ctx;
  return guarded(c, [&]() -> int {
    std::unique_ptr<itp_split> sp(new itp_split);
    sp->pot = pot;
    sp->halforder = pot->null ? 1 : halforder;  // multiproductsplit.cpp:24
    sp->B = B;
    sp->eps = eps;
    sp->coefs = mps_coefficients(sp->halforder);
    int rc = build_split_tables(sp.get());
    if (rc) {
      for (auto& m : sp->members) free_member(c, m.get());
      return rc;
    }
    sp->ctx = c;
    c->refs++;
    *out = sp.release();
    return ITP_OK;
  });
}

void itp_split_destroy(itp_split* sp) {
  if (!sp) return;
  itp_ctx* c = sp->ctx;  // parents may already be gone
  for (auto& m : sp->members) free_member(c, m.get());
  delete sp;
  ctx_release(c);
}

int itp_split_set_time_step(itp_split* sp, double eps) {
  if (!sp || !(eps > 0)) return ITP_E_INVALID;
  itp_ctx* c = sp->pot->grid->ctx;
  return guarded(c, [&]() -> int {
    sp->eps = eps;
    return build_split_tables(sp);
  });
}

int itp_split_halforder(const itp_split* sp) { return sp ? sp->halforder : 0; }

int itp_propagate(itp_split* sp, itp_stateset* s) {
  if (!sp || !s || sp->pot->grid != s->grid) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int { return s->prec == ITP_FP64 ? propagate_impl<double>(sp, s) : propagate_impl<float>(sp, s); });
}

int itp_propagate_noalloc(itp_split* sp, itp_stateset* s) {
  if (!sp || !s || sp->pot->grid != s->grid) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    return s->prec == ITP_FP64 ? propagate_impl<double>(sp, s, true) : propagate_impl<float>(sp, s, true);
  });
}

size_t itp_propagate_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

int itp_orthonormalize(itp_stateset* s, size_t* bad_n, double* bad_lambda) {
  if (!s) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    return s->prec == ITP_FP64 ? ortho_impl<double>(s, bad_n, bad_lambda) : ortho_impl<float>(s, bad_n, bad_lambda);
  });
}

int itp_orthonormalize_cholesky(itp_stateset* s, int* used_eigen, size_t* bad_n, double* bad_lambda) {
  if (!s) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    return s->prec == ITP_FP64 ? cholqr_impl<double>(s, used_eigen, bad_n, bad_lambda)
                               : cholqr_impl<float>(s, used_eigen, bad_n, bad_lambda);
  });
}

int itp_stateset_cholesky_condition(const itp_stateset* s, double* kappa) {
  if (!s || !kappa) return ITP_E_INVALID;
  *kappa = s->chol_kappa;
  return ITP_OK;
}

int itp_stateset_cholesky_diagonal(const itp_stateset* s, double* d) {
  if (!s || !d) return ITP_E_INVALID;
  if (s->chol_diag.size() != s->N) return ITP_E_INVALID;
  std::copy(s->chol_diag.begin(), s->chol_diag.end(), d);
  return ITP_OK;
}

int itp_rayleigh_quotients(itp_hamiltonian* h, const itp_stateset* src, itp_stateset* dst, double* E) {
  if (!h || !src || !dst || !E || dst->prec != ITP_FP64 || dst->N != src->N || dst->grid != src->grid ||
      h->pot->grid != src->grid)
    return ITP_E_INVALID;
  itp_ctx* c = src->grid->ctx;
  return guarded(c, [&]() -> int {
    const GridInfo& g = src->grid->g;
    const long long n = (long long)src->N * g.M;
    void* to = dst->buf[dst->cur];
    if (src->prec == ITP_FP64) HIPCK(c, hipMemcpyAsync(to, src->buf[src->cur], n * 16, hipMemcpyDeviceToDevice, c->stream));
    else HIPCK(c, launch_convert(src->buf[src->cur], to, n, 0, c->stream));
    HIPCK(c, launch_field_norms((const cplx<double>*)to, g.M, (int)src->N, dst->norms, c->stream));
    int rc = energies_impl<double>(h, dst, E, nullptr, nullptr, nullptr, true);
    if (rc) return rc;
    std::vector<double> nrm(src->N);
    HIPCK(c, cmemcpy(c, nrm.data(), dst->norms, src->N * sizeof(double), hipMemcpyDeviceToHost));
    const double dx2 = g.dx * g.dx;
    for (size_t k = 0; k < src->N; k++) E[k] /= nrm[k] * dx2;
    return ITP_OK;
  });
}

int itp_overlap(itp_stateset* s, double* out) {
  if (!s || !out) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    int rc = s->prec == ITP_FP64 ? gram_impl<double>(s) : gram_impl<float>(s);
    if (rc) return rc;
    std::vector<cplx<double>> S((size_t)s->np * s->np);
    HIPCK(c, hipMemcpyAsync(S.data(), s->A[0], S.size() * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    const size_t N = s->N;
    cplx<double>* o = (cplx<double>*)out;
    for (size_t i = 0; i < N; i++)
      for (size_t j = 0; j < N; j++) o[j * N + i] = S[i * s->np + j];  // column-major
    return ITP_OK;
  });
}

int itp_rayleigh_ritz(itp_hamiltonian* h, itp_stateset* s, double* E, double* sigma) {
  if (!h || !s || h->pot->grid != s->grid) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    return s->prec == ITP_FP64 ? rr_impl<double>(h, s, E, sigma) : rr_impl<float>(h, s, E, sigma);
  });
}

int itp_stateset_convert(itp_stateset* dst, const itp_stateset* src) {
  if (!dst || !src || dst->grid != src->grid || dst->N != src->N) return ITP_E_INVALID;
  itp_ctx* c = dst->grid->ctx;
  return guarded(c, [&]() -> int {
    const long long n = (long long)src->N * src->grid->g.M;  // complex elements
    if (dst->prec == src->prec)
      HIPCK(c, hipMemcpyAsync(dst->buf[dst->cur], src->buf[src->cur], (size_t)n * src->esz, hipMemcpyDeviceToDevice,
                              c->stream));
    else
      HIPCK(c, launch_convert(src->buf[src->cur], dst->buf[dst->cur], n, dst->prec == ITP_FP32 ? 1 : 0, c->stream));
    return ITP_OK;
  });
}

int itp_stateset_init_noise(itp_stateset* s, uint64_t* rng_state, double dx) {
  if (!s || !rng_state || !(dx > 0)) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    const long long M = s->grid->g.M;
    const size_t need = (size_t)s->N * noise_blocks(M) * sizeof(double);
    const size_t have = 2 * s->N * (size_t)std::max(s->grid->g.sx, s->grid->g.sy) * 16;
    if (need > have) return set_err(c, ITP_E_UNSUPPORTED, "noise partials exceed the energy-partials buffer");
    const uint64_t s0 = *rng_state;
    if (s->prec == ITP_FP64)
      HIPCK(c, launch_noise<double>(s0, M, (int)s->N, dx, (double*)s->epart, (cplx<double>*)s->buf[s->cur],
                                    c->stream));
    else
      HIPCK(c, launch_noise<float>(s0, M, (int)s->N, dx, (double*)s->epart, (cplx<float>*)s->buf[s->cur],
                                   c->stream));
    *rng_state = s0 + 2ull * (unsigned long long)s->N * (unsigned long long)M * 0x9E3779B97F4A7C15ULL;
    return ITP_OK;
  });
}

int itp_stateset_overlap_eigenvalues(const itp_stateset* s, double* lam) {
  if (!s || !lam) return ITP_E_INVALID;
  if (s->lam_host.size() != s->N) return ITP_E_INVALID;
  std::copy(s->lam_host.begin(), s->lam_host.end(), lam);
  return ITP_OK;
}

int itp_hamiltonian_create(itp_potential* pot, double B, itp_hamiltonian** out) {
  if (!pot || !out || !std::isfinite(B)) return ITP_E_INVALID;
  itp_ctx* c = pot->grid->ctx;
  return guarded(c, [&]() -> int {
    std::unique_ptr<itp_hamiltonian> h(new itp_hamiltonian);
    h->pot = pot;
    h->B = B;
    kin_tables(pot->grid->g, B, h->host);
    std::vector<const double*> p;
    int rc = upload_tables(c, {&h->host.mxy, &h->host.mx, &h->host.mx2, &h->host.my}, &h->dev, p);
    if (rc) return rc;
    h->d = KinHDev{p[0], p[1], p[2], p[3]};
    h->ctx = c;
    c->refs++;
    *out = h.release();
    return ITP_OK;
  });
}

void itp_hamiltonian_destroy(itp_hamiltonian* h) {
  if (!h) return;
  itp_ctx* c = h->ctx;  // parents may already be gone
  if (h->dev) dfree(c, h->dev);
  delete h;
  ctx_release(c);
}

int itp_energies(itp_hamiltonian* h, itp_stateset* s, double* E, double* sigma) {
  if (!h || !s || h->pot->grid != s->grid) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    return s->prec == ITP_FP64 ? energies_impl<double>(h, s, E, sigma, nullptr, nullptr)
                               : energies_impl<float>(h, s, E, sigma, nullptr, nullptr);
  });
}

int itp_energies_complex(itp_hamiltonian* h, itp_stateset* s, double* mean, double* meansq) {
  if (!h || !s || h->pot->grid != s->grid) return ITP_E_INVALID;
  itp_ctx* c = s->grid->ctx;
  return guarded(c, [&]() -> int {
    return s->prec == ITP_FP64 ? energies_impl<double>(h, s, nullptr, nullptr, mean, meansq)
                               : energies_impl<float>(h, s, nullptr, nullptr, mean, meansq);
  });
}

int itp_apply_expkinetic_host(itp_grid* g, double B, double e, double coeff, double prefactor, double* fields,
                              size_t n) {
  if (!g || !fields) return ITP_E_INVALID;
  itp_ctx* c = g->ctx;
  return guarded(c, [&]() -> int {
    ExpKinTabs t;
    expkin_tables(g->g, B, e, coeff, prefactor, t);
    double* dev = nullptr;
    std::vector<const double*> p;
    int rc = upload_tables(c, {&t.mult, &t.xm, &t.xm2, &t.ym}, &dev, p);
    if (rc) return rc;
    DevHold hold{c, dev};
    DevFields df;
    const size_t nel = n * (size_t)g->g.M;
    if ((rc = hook_begin(c, fields, nel, df, true))) return rc;
    std::vector<Sweep<double>> prog;
    expkinetic_program<double>(branch_of(B, g->g.bt), KinDev{p[0], p[1], p[2], p[3]}, df.p, df.w, prog);
    GpuBackend be;
    be.st = c->stream;
    run_program<double>(be, g->g, g->lg[0], prog, (int)n);
    return hook_end(c, fields, nel, df, be.err);
  });
}

int itp_apply_exppotential_host(itp_potential* pot, double e, double coeff, double prefactor, double* fields,
                                size_t n) {
  if (!pot || !fields) return ITP_E_INVALID;
  itp_ctx* c = pot->grid->ctx;
  itp_grid* g = pot->grid;
  return guarded(c, [&]() -> int {
    std::vector<double> vals;
    double* dev = nullptr;
    std::vector<const double*> p;
    if (!pot->null) exppot_values(g->g, pot->V.data(), e, coeff, prefactor, vals);
    int rc = upload_tables(c, {&vals}, &dev, p);
    if (rc) return rc;
    DevHold hold{c, dev};
    DevFields df;
    const size_t nel = n * (size_t)g->g.M;
    if ((rc = hook_begin(c, fields, nel, df, false))) return rc;
    std::vector<Sweep<double>> prog;
    exppotential_program<double>(pot->null ? nullptr : p[0], prefactor, df.p, prog);
    GpuBackend be;
    be.st = c->stream;
    run_program<double>(be, g->g, g->lg[0], prog, (int)n);
    return hook_end(c, fields, nel, df, be.err);
  });
}

int itp_apply_secondorder_host(itp_potential* pot, double B, double e, double prefactor, int exponent,
                               double* fields, size_t n) {
  if (!pot || !fields || exponent < 1) return ITP_E_INVALID;
  itp_ctx* c = pot->grid->ctx;
  itp_grid* g = pot->grid;
  return guarded(c, [&]() -> int {
    MemberTabs m;
    const bool has_v = !pot->null;
    if (has_v) {
      expkin_tables(g->g, B, e, -1.0, 1.0, m.host);
      exppot_values(g->g, pot->V.data(), e, -0.5, 1.0, m.vh);
      if (exponent >= 2) exppot_values(g->g, pot->V.data(), e, -1.0, 1.0, m.vf);
      exppot_values(g->g, pot->V.data(), e, -0.5, prefactor, m.vp);
    } else {
      expkin_tables(g->g, B, e, -1.0 * exponent, prefactor, m.host);
    }
    std::vector<const double*> p;
    int rc = upload_tables(c, {&m.host.mult, &m.host.xm, &m.host.xm2, &m.host.ym, &m.vh, &m.vf, &m.vp}, &m.dev, p);
    if (rc) return rc;
    DevHold hold{c, m.dev};
    m.d.kin = KinDev{p[0], p[1], p[2], p[3]};
    m.d.vhalf = p[4];
    m.d.vfull = p[5];
    m.d.vpref = p[6];
    DevFields df;
    const size_t nel = n * (size_t)g->g.M;
    if ((rc = hook_begin(c, fields, nel, df, true))) return rc;
    cplx<double>* out = nullptr;
    if ((rc = dalloc(c, (void**)&out, nel * 16))) return rc;
    DevHold hold2{c, out};
    std::vector<Sweep<double>> prog;
    member_program<double>(branch_of(B, g->g.bt), m.d, has_v ? exponent : 1, has_v, df.p, df.w, out, false, prog);
    GpuBackend be;
    be.st = c->stream;
    run_program<double>(be, g->g, g->lg[0], prog, (int)n);
    if (be.err != hipSuccess) return set_err(c, ITP_E_HIP, hipGetErrorString(be.err));
    HIPCK(c, hipMemcpyAsync(df.p, out, nel * 16, hipMemcpyDeviceToDevice, c->stream));
    return hook_end(c, fields, nel, df, hipSuccess);
  });
}

int itp_apply_hamiltonian_host(itp_hamiltonian* h, double* fields, size_t n) {
  if (!h || !fields) return ITP_E_INVALID;
  itp_grid* g = h->pot->grid;
  itp_ctx* c = g->ctx;
  return guarded(c, [&]() -> int {
    DevFields df;
    const size_t nel = n * (size_t)g->g.M;
    int rc = hook_begin(c, fields, nel, df, true);
    if (rc) return rc;
    cplx<double>* out = nullptr;
    if ((rc = dalloc(c, (void**)&out, nel * 16))) return rc;
    DevHold hold{c, out};
    std::vector<Sweep<double>> prog;
    hamiltonian_program<double>(g->g, g->lg[0], branch_of(h->B, g->g.bt), h->d, h->pot->null ? nullptr : h->pot->dV,
                                df.p, df.w, out, nullptr, nullptr, prog);
    GpuBackend be;
    be.st = c->stream;
    run_program<double>(be, g->g, g->lg[0], prog, (int)n);
    if (be.err != hipSuccess) return set_err(c, ITP_E_HIP, hipGetErrorString(be.err));
    HIPCK(c, hipMemcpyAsync(df.p, out, nel * 16, hipMemcpyDeviceToDevice, c->stream));
    return hook_end(c, fields, nel, df, hipSuccess);
  });
}

int itp_eigh_host(itp_ctx* c, size_t n, double* A, double* w) {
  if (!c || !A || !w || n < 1 || n > 4096) return ITP_E_INVALID;
  return guarded(c, [&]() -> int {
    // a throwaway state set of the right size carries the eigensolver workspace
    itp_grid* g = nullptr;
    int rc = itp_grid_create(c, 8, 8, 1.0, ITP_PERIODIC, &g);
    if (rc) return rc;
    std::unique_ptr<itp_grid, void (*)(itp_grid*)> hg(g, itp_grid_destroy);
    itp_stateset* s = nullptr;
    if ((rc = itp_stateset_create(g, n, ITP_FP64, &s))) return rc;
    std::unique_ptr<itp_stateset, void (*)(itp_stateset*)> hs(s, itp_stateset_destroy);
    const cplx<double>* a = (const cplx<double>*)A;
    std::vector<cplx<double>> full((size_t)s->np * s->np, mk<double>(0.0, 0.0));
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++)
        full[i * s->np + j] = (i <= j) ? a[j * n + i] : conj(a[i * n + j]);  // upper triangle, column-major
    for (size_t i = 0; i < n; i++) full[i * s->np + i].y = 0;
    HIPCK(c, cmemcpy(c, s->A[0], full.data(), full.size() * 16, hipMemcpyHostToDevice));
    HIPCK(c, cmemset(c, s->bad, 0, sizeof(int)));
    rc = eigen_stage(s, nullptr, nullptr, false);
    if (rc && rc != ITP_E_NONPOSITIVE_EIG && rc != ITP_E_NONNORMAL_EIG) return rc;
    std::vector<cplx<double>> C((size_t)s->nc * s->nc);
    HIPCK(c, cmemcpy(c, C.data(), s->C, C.size() * 16, hipMemcpyDeviceToHost));
    HIPCK(c, cmemcpy(c, w, s->lam, n * sizeof(double), hipMemcpyDeviceToHost));
    cplx<double>* o = (cplx<double>*)A;
    for (size_t r = 0; r < n; r++)
      for (size_t cc = 0; cc < n; cc++) o[r * n + cc] = C[cc * s->nc + r];  // column r = eigenvector r
    c->err.clear();
    return ITP_OK;
  });
}

int itp_eigvalsh_host(itp_ctx* c, size_t n, const double* A, double* w) {
  if (!c || !A || !w || n < 1 || n > 4096) return ITP_E_INVALID;
  return guarded(c, [&]() -> int {
    // a throwaway state set of the right size carries the workspace
    itp_grid* g = nullptr;
    int rc = itp_grid_create(c, 8, 8, 1.0, ITP_PERIODIC, &g);
    if (rc) return rc;
    std::unique_ptr<itp_grid, void (*)(itp_grid*)> hg(g, itp_grid_destroy);
    itp_stateset* s = nullptr;
    if ((rc = itp_stateset_create(g, n, ITP_FP64, &s))) return rc;
    std::unique_ptr<itp_stateset, void (*)(itp_stateset*)> hs(s, itp_stateset_destroy);
    const cplx<double>* a = (const cplx<double>*)A;
    std::vector<cplx<double>> full((size_t)s->np * s->np, mk<double>(0.0, 0.0));
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++)
        full[i * s->np + j] = (i <= j) ? a[j * n + i] : conj(a[i * n + j]);  // upper triangle, column-major
    for (size_t i = 0; i < n; i++) full[i * s->np + i].y = 0;
    HIPCK(c, cmemcpy(c, s->A[0], full.data(), full.size() * 16, hipMemcpyHostToDevice));
    std::vector<double> ev;
    if ((rc = hermitian_eigenvalues(s, s->A[0], ev))) return rc;
    std::copy(ev.begin(), ev.end(), w);
    return ITP_OK;
  });
}

int itp_init_states(uint64_t* rng_state, size_t sx, size_t sy, double dx, size_t N, double* out) {
  if (!rng_state || !out) return ITP_E_INVALID;
  const size_t M = sx * sy;
  for (size_t n = 0; n < N; n++) {
    double* s = out + 2 * n * M;
    for (size_t i = 0; i < M; i++) {
      s[2 * i] = 2 * (double(splitmix64(*rng_state) >> 11) * (1.0 / 9007199254740992.0)) - 1;
      s[2 * i + 1] = 2 * (double(splitmix64(*rng_state) >> 11) * (1.0 / 9007199254740992.0)) - 1;
    }
    double acc = 0;
    for (size_t i = 0; i < M; i++) acc += s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1];
    const double nrm = std::sqrt(acc) * dx;
    const double f = 1.0 / nrm;
    for (size_t i = 0; i < 2 * M; i++) s[i] *= f;
  }
  return ITP_OK;
}

}  // extern "C"
