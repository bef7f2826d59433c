// C ABI of liblbswim_amd.so: object lifetimes, host/device mirrors, field
// conversion for the FP32 storage form, kernel launches on the library stream
// and per-kernel event timing.  Declarations and the reference interface each
// entry point replaces are in include/lbswim.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"
#include "lbswim.h"

using namespace lbs;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                                       \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess)                                                                                   \
      return fail(e_ == hipErrorOutOfMemory ? LBSWIM_ENOMEM : LBSWIM_EHIP, "%s failed: %s (%s:%d)", #expr, \
                  hipGetErrorString(e_), __FILE__, __LINE__);                                               \
  } while (0)

#define CHECK_ARG(cond, ...)                     \
  do {                                           \
    if (!(cond)) return fail(LBSWIM_EINVAL, __VA_ARGS__); \
  } while (0)

// One stream orders the device work of every object, like the reference's
// default stream.  Created on first use; default priority.  Inside
// lbswim_system_step the tracer moves run on a second stream (`side`, also
// default priority) beside the swimmer moves and the next step's spreading;
// the main stream waits for them before that call returns, so every other
// entry point sees one ordered stream.
struct Runtime {
  std::atomic<bool> ready{false};
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t u_ready = nullptr;  // main -> side: the u the tracers move in is written
  hipEvent_t joined = nullptr;   // side -> main: every tracer move enqueued so far is done
};
Runtime g_rt;
std::mutex g_rt_mu;  // creation of the runtime's streams and events
// Where this thread's launches go: the library stream, or another stream
// inside a StreamScope (the side stream, an ensemble's capture stream).
thread_local hipStream_t t_cur = nullptr;

int ensure_runtime() {
  if (g_rt.ready) return LBSWIM_OK;
  std::lock_guard<std::mutex> lock(g_rt_mu);
  if (g_rt.ready) return LBSWIM_OK;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) return fail(LBSWIM_ENODEV, "no HIP device available (%s)", hipGetErrorString(e));
  HIP_TRY(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  // the no-allocation step's fixed workspace on this device
  void* scratch = nullptr;
  HIP_TRY(noalloc_scratch(&scratch));
  g_rt.ready = true;
  return LBSWIM_OK;
}

hipStream_t stream() { return t_cur ? t_cur : g_rt.stream; }

int ensure_side_stream() {
  std::lock_guard<std::mutex> lock(g_rt_mu);
  if (g_rt.side) return LBSWIM_OK;
  HIP_TRY(hipEventCreateWithFlags(&g_rt.u_ready, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&g_rt.joined, hipEventDisableTiming));
  HIP_TRY(hipStreamCreateWithFlags(&g_rt.side, hipStreamNonBlocking));
  return LBSWIM_OK;
}

// This thread's launches inside the scope go to stream `s`.
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(t_cur) { t_cur = s; }
  ~StreamScope() { t_cur = prev; }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;
};

int dev_alloc_zero(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return LBSWIM_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(LBSWIM_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  HIP_TRY(hipMemsetAsync(*p, 0, bytes, stream()));
  return LBSWIM_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

// Copy helpers that complete before returning (SharedArray::H2D/D2H are
// synchronous cudaMemcpy calls, Shared.cu:15-26) but stay ordered after the
// work already enqueued on the library stream.
int copy_h2d(void* dst, const void* src, size_t bytes) {
  if (!bytes) return LBSWIM_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
  HIP_TRY(hipStreamSynchronize(stream()));
  return LBSWIM_OK;
}
int copy_d2h(void* dst, const void* src, size_t bytes) {
  if (!bytes) return LBSWIM_OK;
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
  HIP_TRY(hipStreamSynchronize(stream()));
  return LBSWIM_OK;
}

void fill_params(lbswim_params* p, double tau_s, double tau_b) {
  std::memset(p, 0, sizeof *p);
  p->cs2 = val(mk(1, 3));
  for (int q = 0; q < Q; ++q) {
    p->w[q] = TAB.w[q];
    p->norms[q] = TAB.norms[q];
    for (int a = 0; a < D; ++a) {
      p->xi[q][a] = double(CI[q][a]);
      p->ci[q][a] = CI[q][a];
      for (int b = 0; b < D; ++b) p->Q[q][a][b] = val(mk(3 * CI[q][a] * CI[q][b] - (a == b ? 1 : 0), 3));
    }
    for (int c = 0; c < Q; ++c)
      if (CI[c][0] == -CI[q][0] && CI[c][1] == -CI[q][1] && CI[c][2] == -CI[q][2]) p->complement[q] = c;
    for (int m = 0; m < Q; ++m) {
      p->mm[q][m] = TAB.mm[q][m];
      p->mmi[q][m] = TAB.mmi[q][m];
    }
  }
  p->tau_s = tau_s;
  p->tau_b = tau_b;
}

struct TimeRec {
  int kernel;
  hipEvent_t start, stop;
};

}  // namespace

// --------------------------------------------------------------------------
struct lbswim_lattice {
  int storage = LBSWIM_STORE_F64;
  lbswim_params params{};
  lbswim_addressing addr{};
  Geom geom{};
  int time_step = 0;
  size_t n = 0;
  size_t elem = 8;
  void* d_rho = nullptr;
  void* d_u = nullptr;
  void* d_force = nullptr;
  void* d_f[2] = {nullptr, nullptr};
  int cur = 0;  // d_f[cur] holds fOld
  double* h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* d_partial = nullptr;  // 4*nblocks + 4 doubles for totals
  int totals_blocks = 0;
  unsigned timing = 0;  // bit k: record kernel k
  std::vector<TimeRec> recs;
  std::vector<hipEvent_t> pool;
  // workspace of the deterministic spreading in lbswim_system_step
  uint32_t sp_lists = 0;          // one per 64 sites
  uint32_t* sp_counts = nullptr;  // sp_lists
  uint32_t* sp_off = nullptr;     // sp_lists + 1 offsets + scan and sort scratch
  uint32_t* sp_cursor = nullptr;  // sp_lists
  size_t sp_pts_cap = 0;
  void* sp_pts = nullptr;       // SpreadPoint records for sp_pts_cap points
  uint32_t* sp_slot = nullptr;  // point ids in storage order (sp_pts_cap)
  size_t sp_list_cap = 0;
  uint32_t* sp_list = nullptr;
  uint32_t* sp_sorted = nullptr;
  double* sp_edge = nullptr;  // 6 sp_lists: force at both ends of each list's run (fused step)
  size_t sp_rec_cap = 0;      // list entries sp_rec and sp_wz hold (fused step)
  RunEntry* sp_rec = nullptr;
  int* sp_wz = nullptr;
  // Stored force by site (launch_site_force): per-site counts and offsets,
  // and ss_cap point entries (slot, id, site, ordered product).
  uint32_t* ss_counts = nullptr;
  uint32_t* ss_off = nullptr;
  size_t ss_cap = 0;
  uint32_t* ss_local = nullptr;
  uint32_t* ss_ent = nullptr;  // key and site of each entry
  double* ss_sorted = nullptr;
  // Virtual force: after a fused step the force lives in the sorted point
  // lists above and d_force is stale until materialize_force writes it.
  int force_mode = LBSWIM_FORCE_AUTO;
  bool force_virtual = false;
  // Step + CalcHydro fused into one launch in the coupled step (where the
  // lattice qualifies); rows/strips 0 = chosen from the lattice size
  int fuse = LBSWIM_FUSION_AUTO;
  int fuse_rows = 0;
  // u with a site's components together, written by the coupled step's
  // CalcHydro (fused or not) and by ensemble steps (the particle moves gather
  // from it); current only while nothing else has written u since.
  // u_virtual: u was written only there, and d_u is stale until materialize_u
  // writes it (u_virtual implies u_aos_valid).
  // tracers_reading: tracer moves on the side stream may still read d_u_aos;
  // they are done once tracers_done has fired.
  void* d_u_aos = nullptr;
  bool tracers_reading = false;
  hipEvent_t tracers_done = nullptr;
  bool u_aos_valid = false;
  bool u_virtual = false;
  int fuse_strips = 0;
  // Scatter force (LBSWIM_FORCE_SCATTER, FP32 storage, fused step): the
  // swimmer moves add the next step's ZeroForce + AddForces into d_fsc[fsc_cur]
  // with atomics, and the fused launch reads it there while it clears the
  // other buffer for the moves after it.  fsc_valid: d_fsc[fsc_cur] holds that
  // force for the swimmer arrays (and their versions) in fsc_set.  After a
  // scatter step the force is virtual with force_src FORCE_SRC_SCATTER: the
  // force array is copied from d_fsc[fsc_last] on demand.
  void* d_fsc[2] = {nullptr, nullptr};
  int fsc_cur = 0;
  int fsc_last = 0;
  bool fsc_valid = false;
  std::vector<std::pair<const lbswim_swimmers*, uint64_t>> fsc_set;
  int force_src = 0;
};

// What a virtual force is made from.
constexpr int FORCE_SRC_LISTS = 0;    // the sorted point lists
constexpr int FORCE_SRC_SCATTER = 1;  // the scatter buffer d_fsc[fsc_last]

// Device storage of one particle array in slot order (particle_order.hip):
// field f of slot s, component c is at d[f][c*num + s]; perm[s] is the user's
// index of the particle in slot s.  The staging buffer holds 3*num 8-byte
// words (the largest field); every buffer keeps its address until the array
// is destroyed.
struct PStore {
  int num = 0;
  uint64_t* d[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool has[5] = {false, false, false, false, false};
  uint64_t* stage = nullptr;
  int* perm = nullptr;
  int* perm_alt = nullptr;
  uint32_t* keys = nullptr;
  uint32_t* keys_alt = nullptr;
  int* idx = nullptr;
  int* order = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  bool dirty = true;  // positions written from the host since the last ordering
  int moves_since_order = 0;
  void* h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

struct lbswim_swimmers {
  lbswim_common_params common_host{};
  lbswim_common_params common_dev{};  // the copy kernels see (SharedItem::device)
  PStore st;
  unsigned long long* d_tumbles = nullptr;
  // Changes whenever r, n or the parameters may have changed (a value no
  // other array has had): a lattice's scatter force records it.
  uint64_t version = 0;
};

// Fresh swimmer version numbers, unique over the process.
static uint64_t next_version() {
  static std::atomic<uint64_t> counter{0};
  return ++counter;
}

struct lbswim_tracers {
  PStore st;
};

namespace {

SwimParams to_kernel(const lbswim_common_params& c) {
  SwimParams p;
  p.P = c.P;
  p.l = c.l;
  p.alpha = c.alpha;
  p.mobility = c.mobility;
  p.translational_advection_off = c.translational_advection_off;
  p.rotational_advection_off = c.rotational_advection_off;
  p.seed = c.seed;
  return p;
}

hipEvent_t take_event(lbswim_lattice* lat) {
  if (!lat->pool.empty()) {
    hipEvent_t e = lat->pool.back();
    lat->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Records an event pair around one launch when timing is on.
struct KernelTimer {
  lbswim_lattice* lat;
  int kernel;
  hipEvent_t start = nullptr;
  KernelTimer(lbswim_lattice* l, int k) : lat(l), kernel(k) {
    if (lat && ((lat->timing >> k) & 1u)) {
      start = take_event(lat);
      if (start) (void)hipEventRecord(start, stream());
    }
  }
  void done() {
    if (!start) return;
    hipEvent_t stop = take_event(lat);
    if (!stop) return;
    (void)hipEventRecord(stop, stream());
    lat->recs.push_back(TimeRec{kernel, start, stop});
    start = nullptr;
  }
};

size_t field_count(const lbswim_lattice* lat, int field) {
  switch (field) {
    case LBSWIM_RHO: return lat->n;
    case LBSWIM_U:
    case LBSWIM_FORCE: return 3 * lat->n;
    case LBSWIM_FOLD:
    case LBSWIM_FNEW: return 15 * lat->n;
    default: return 0;
  }
}

void* field_device(const lbswim_lattice* lat, int field) {
  switch (field) {
    case LBSWIM_RHO: return lat->d_rho;
    case LBSWIM_U: return lat->d_u;
    case LBSWIM_FORCE: return lat->d_force;
    case LBSWIM_FOLD: return lat->d_f[lat->cur];
    case LBSWIM_FNEW: return lat->d_f[1 - lat->cur];
    default: return nullptr;
  }
}

int lattice_host(lbswim_lattice* lat, int field, double** out) {
  if (!lat->h[field]) {
    const size_t cnt = field_count(lat, field);
    lat->h[field] = static_cast<double*>(std::calloc(cnt ? cnt : 1, sizeof(double)));
    if (!lat->h[field]) return fail(LBSWIM_ENOMEM, "host mirror allocation of %zu doubles failed", cnt);
  }
  *out = lat->h[field];
  return LBSWIM_OK;
}

// Offset added when widening an FP32 deviation field to its FP64 value.
double f32_offset(const lbswim_lattice* lat, int field, size_t idx) {
  if (field == LBSWIM_RHO) return 1.0;
  if (field == LBSWIM_FOLD || field == LBSWIM_FNEW) return TAB.w[idx / lat->n];
  return 0.0;
}

int check_lattice(const lbswim_lattice* lat) {
  if (!lat) return fail(LBSWIM_EINVAL, "lattice handle is NULL");
  return LBSWIM_OK;
}

}  // namespace

namespace {

// The device keeps post-collision distributions g (pull streaming); the API's
// fOld/fNew are post-stream, f_p(x) = g_p(x - c_p).  Calls fn(x, x - c_p) for
// every site x of plane p.
template <typename Fn>
void for_each_upstream(const lbswim_lattice* lat, int p, Fn&& fn) {
  const int nx = lat->addr.size[0], ny = lat->addr.size[1], nz = lat->addr.size[2];
  const int* c = lat->params.ci[p];
  size_t x = 0;
  for (int i = 0; i < nx; ++i) {
    const int ui = (i - c[0] + nx) % nx;
    for (int j = 0; j < ny; ++j) {
      const int uj = (j - c[1] + ny) % ny;
      const size_t row = (size_t(ui) * ny + uj) * nz;
      for (int k = 0; k < nz; ++k, ++x) {
        const int uk = (k - c[2] + nz) % nz;
        fn(x, row + uk);
      }
    }
  }
}

bool is_dist(int field) { return field == LBSWIM_FOLD || field == LBSWIM_FNEW; }

}  // namespace

extern "C" {

const char* lbswim_last_error(void) { return g_err.c_str(); }

int lbswim_synchronize(void) {
  if (!g_rt.ready) return LBSWIM_OK;
  HIP_TRY(hipStreamSynchronize(stream()));
  return LBSWIM_OK;
}

struct lbswim_mark {
  hipEvent_t ev = nullptr;
};

int lbswim_mark_create(lbswim_mark** out) {
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  if (int rc = ensure_runtime()) return rc;
  lbswim_mark* m = new (std::nothrow) lbswim_mark();
  if (!m) return fail(LBSWIM_ENOMEM, "mark allocation failed");
  hipError_t e = hipEventCreate(&m->ev);
  if (e == hipSuccess) e = hipEventRecord(m->ev, stream());
  if (e != hipSuccess) {
    if (m->ev) (void)hipEventDestroy(m->ev);
    delete m;
    return fail(LBSWIM_EHIP, "recording a mark failed: %s", hipGetErrorString(e));
  }
  *out = m;
  return LBSWIM_OK;
}

int lbswim_mark_wait(lbswim_mark* m) {
  CHECK_ARG(m != nullptr, "mark is NULL");
  HIP_TRY(hipEventSynchronize(m->ev));
  return LBSWIM_OK;
}

int lbswim_mark_reached(lbswim_mark* m, int* reached) {
  CHECK_ARG(m != nullptr && reached != nullptr, "NULL argument");
  const hipError_t e = hipEventQuery(m->ev);
  if (e != hipSuccess && e != hipErrorNotReady) return fail(LBSWIM_EHIP, "mark query failed: %s", hipGetErrorString(e));
  *reached = e == hipSuccess ? 1 : 0;
  return LBSWIM_OK;
}

int lbswim_mark_elapsed(lbswim_mark* a, lbswim_mark* b, double* ms) {
  CHECK_ARG(a && b && ms, "NULL argument");
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, a->ev, b->ev));
  *ms = t;
  return LBSWIM_OK;
}

void lbswim_mark_destroy(lbswim_mark* m) {
  if (!m) return;
  (void)hipEventDestroy(m->ev);
  delete m;
}

// ---- lattice ----------------------------------------------------------------

int lbswim_lattice_create(int nx, int ny, int nz, double tau_s, double tau_b, int storage, lbswim_lattice** out) {
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  CHECK_ARG(nx >= 1 && ny >= 1 && nz >= 1, "lattice sizes must be >= 1 (got %d, %d, %d)", nx, ny, nz);
  const unsigned long long n = (unsigned long long)nx * ny * nz;
  CHECK_ARG(n <= (unsigned long long)INT_MAX, "lattice of %llu sites exceeds the int site index", n);
  CHECK_ARG(std::isfinite(tau_s) && tau_s > 0.0 && std::isfinite(tau_b) && tau_b > 0.0,
            "relaxation times must be finite and > 0 (got %g, %g)", tau_s, tau_b);
  CHECK_ARG(storage == LBSWIM_STORE_F64 || storage == LBSWIM_STORE_F32DEV, "unknown storage form %d", storage);
  int rc = ensure_runtime();
  if (rc) return rc;

  lbswim_lattice* lat = new (std::nothrow) lbswim_lattice();
  if (!lat) return fail(LBSWIM_ENOMEM, "lattice allocation failed");
  lat->storage = storage;
  lat->elem = storage == LBSWIM_STORE_F64 ? 8 : 4;
  fill_params(&lat->params, tau_s, tau_b);
  lat->addr.size[0] = nx;
  lat->addr.size[1] = ny;
  lat->addr.size[2] = nz;
  lat->addr.n = int(n);
  lat->addr.strides[0] = ny * nz;
  lat->addr.strides[1] = nz;
  lat->addr.strides[2] = 1;
  lat->n = size_t(n);
  lat->geom = make_geom(nx, ny, nz);
  const size_t e = lat->elem;
  lat->totals_blocks = int(std::min<size_t>(1024, (lat->n + 255) / 256));
  if ((rc = dev_alloc_zero(&lat->d_rho, e * lat->n)) || (rc = dev_alloc_zero(&lat->d_u, 3 * e * lat->n)) ||
      (rc = dev_alloc_zero(&lat->d_force, 3 * e * lat->n)) ||
      (rc = dev_alloc_zero(&lat->d_f[0], 15 * e * lat->n)) ||
      (rc = dev_alloc_zero(&lat->d_f[1], 15 * e * lat->n)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->d_partial),
                           sizeof(double) * (4 * size_t(lat->totals_blocks) + 4)))) {
    std::string msg = g_err;
    lbswim_lattice_destroy(lat);
    g_err = msg;
    return rc;
  }
  hipError_t he = hipStreamSynchronize(stream());
  if (he != hipSuccess) {
    lbswim_lattice_destroy(lat);
    return fail(LBSWIM_EHIP, "lattice initialisation failed: %s", hipGetErrorString(he));
  }
  *out = lat;
  return LBSWIM_OK;
}

void lbswim_lattice_destroy(lbswim_lattice* lat) {
  if (!lat) return;
  if (g_rt.ready) (void)hipStreamSynchronize(stream());
  dev_free(lat->d_rho);
  dev_free(lat->d_u);
  dev_free(lat->d_u_aos);
  if (lat->tracers_done) (void)hipEventDestroy(lat->tracers_done);
  dev_free(lat->d_force);
  dev_free(lat->d_fsc[0]);
  dev_free(lat->d_fsc[1]);
  dev_free(lat->d_f[0]);
  dev_free(lat->d_f[1]);
  dev_free(lat->d_partial);
  dev_free(lat->sp_counts);
  dev_free(lat->sp_off);
  dev_free(lat->sp_cursor);
  dev_free(lat->sp_edge);
  dev_free(lat->sp_rec);
  dev_free(lat->sp_wz);
  dev_free(lat->ss_counts);
  dev_free(lat->ss_off);
  dev_free(lat->ss_local);
  dev_free(lat->ss_ent);
  dev_free(lat->ss_sorted);
  dev_free(lat->sp_pts);
  dev_free(lat->sp_slot);
  dev_free(lat->sp_list);
  dev_free(lat->sp_sorted);
  for (double* h : lat->h) std::free(h);
  for (auto& r : lat->recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (hipEvent_t ev : lat->pool) (void)hipEventDestroy(ev);
  delete lat;
}

// The force of a virtual-force lattice as the reference holds it: written
// from the point lists of the last fused step (ZeroForce + AddForces).
static int materialize_force(lbswim_lattice* lat) {
  if (!lat->force_virtual) return LBSWIM_OK;
  if (lat->force_src == FORCE_SRC_SCATTER) {
    // the force the last scatter step read
    HIP_TRY(hipMemcpyAsync(lat->d_force, lat->d_fsc[lat->fsc_last], 3 * lat->elem * lat->n, hipMemcpyDeviceToDevice,
                           stream()));
    lat->force_virtual = false;
    return LBSWIM_OK;
  }
  KernelTimer t(lat, LBSWIM_K_ADDF);
  const SpreadPoint* pts = static_cast<const SpreadPoint*>(lat->sp_pts);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_spread_force<double>(lat->geom, pts, lat->sp_off, lat->sp_sorted, static_cast<double*>(lat->d_force),
                                    stream());
  else
    e = launch_spread_force<float>(lat->geom, pts, lat->sp_off, lat->sp_sorted, static_cast<float*>(lat->d_force),
                                   stream());
  t.done();
  HIP_TRY(e);
  lat->force_virtual = false;
  return LBSWIM_OK;
}

// u in the reference layout after fused steps, which wrote it only
// component-interleaved.
static int materialize_u(lbswim_lattice* lat) {
  if (!lat->u_virtual) return LBSWIM_OK;
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_u_from_aos<double>(static_cast<const double*>(lat->d_u_aos), static_cast<double*>(lat->d_u), lat->geom,
                                  stream());
  else
    e = launch_u_from_aos<float>(static_cast<const float*>(lat->d_u_aos), static_cast<float*>(lat->d_u), lat->geom,
                                 stream());
  HIP_TRY(e);
  lat->u_virtual = false;
  return LBSWIM_OK;
}

static VForce virtual_force(const lbswim_lattice* lat) {
  return VForce{static_cast<const SpreadPoint*>(lat->sp_pts), lat->sp_off, lat->sp_sorted, lat->sp_edge, lat->sp_rec,
                lat->sp_wz};
}

static int lattice_step_impl(lbswim_lattice* lat) {
  const double ts = lat->params.tau_s, tb = lat->params.tau_b;
  const VForce vf = virtual_force(lat);
  const VForce* pvf = lat->force_virtual ? &vf : nullptr;
  KernelTimer t(lat, LBSWIM_K_COLLIDE);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_collide_stream<double>(static_cast<const double*>(lat->d_f[lat->cur]),
                                      static_cast<const double*>(lat->d_force),
                                      static_cast<double*>(lat->d_f[1 - lat->cur]), lat->geom, ts, tb, pvf, stream());
  else
    e = launch_collide_stream<float>(static_cast<const float*>(lat->d_f[lat->cur]),
                                     static_cast<const float*>(lat->d_force),
                                     static_cast<float*>(lat->d_f[1 - lat->cur]), lat->geom, ts, tb, pvf, stream());
  t.done();
  HIP_TRY(e);
  lat->cur = 1 - lat->cur;  // lb.cu:242-244
  lat->time_step++;         // lb.cu:246
  return LBSWIM_OK;
}

// aos (the coupled step): u is written only component-interleaved, as the
// fused step writes it, for the particle moves to gather from; d_u is then
// stale until materialize_u writes it.  Otherwise the reference layout.
static int lattice_hydro_impl(lbswim_lattice* lat, bool aos = false) {
  if (aos && !lat->d_u_aos)
    if (int rc = dev_alloc_zero(&lat->d_u_aos, 3 * lat->elem * lat->n)) return rc;
  lat->u_aos_valid = aos;
  lat->u_virtual = aos;
  void* u = aos ? lat->d_u_aos : lat->d_u;
  const VForce vf = virtual_force(lat);
  const VForce* pvf = lat->force_virtual ? &vf : nullptr;
  KernelTimer t(lat, LBSWIM_K_HYDRO);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_calc_hydro<double>(static_cast<const double*>(lat->d_f[lat->cur]),
                                  static_cast<const double*>(lat->d_force), static_cast<double*>(lat->d_rho),
                                  static_cast<double*>(u), lat->geom, pvf, stream(), aos);
  else
    e = launch_calc_hydro<float>(static_cast<const float*>(lat->d_f[lat->cur]), static_cast<const float*>(lat->d_force),
                                 static_cast<float*>(lat->d_rho), static_cast<float*>(u), lat->geom, pvf, stream(),
                                 aos);
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

// Tile rows of the fused step, 0 if the coupled step runs the two launches.
// LBSWIM_FUSION_AUTO fuses lattices whose tiles are full (nz % 64 == 0, rows
// dividing ny), where one launch is the faster; LBSWIM_FUSION_ON any lattice
// the launch supports.  The requested rows, else the largest of 8, 4, 2 that
// qualifies (full tiles first).
static int fusion_rows(const lbswim_lattice* lat) {
  if (lat->fuse == LBSWIM_FUSION_OFF) return 0;
  const bool any = lat->fuse == LBSWIM_FUSION_ON;
  auto ok = [&](int ty, bool need_full) {
    return fused_step_supported(lat->geom, ty, lat->elem) && (!need_full || fused_tiles_full(lat->geom, ty));
  };
  if (lat->fuse_rows) return ok(lat->fuse_rows, !any) ? lat->fuse_rows : 0;
  for (int ty : {8, 4, 2})
    if (ok(ty, true)) return ty;
  if (any)
    for (int ty : {8, 4, 2})
      if (ok(ty, false)) return ty;
  return 0;
}

// Step + CalcHydro of the coupled step in one launch (fusion_rows(lat) > 0),
// once the tracer moves still reading the interleaved u (side stream) are done.
// scatter: the force is read from the scatter buffer d_fsc[fsc_cur], and the
// other one is cleared for the moves after the launch.
static int lattice_step_hydro_impl(lbswim_lattice* lat, int rows, bool scatter = false) {
  const double ts = lat->params.tau_s, tb = lat->params.tau_b;
  if (!lat->d_u_aos)
    if (int rc = dev_alloc_zero(&lat->d_u_aos, 3 * lat->elem * lat->n)) return rc;
  if (lat->tracers_reading) {
    HIP_TRY(hipStreamWaitEvent(stream(), lat->tracers_done, 0));
    lat->tracers_reading = false;
  }
  const VForce vf = virtual_force(lat);
  const VForce* pvf = lat->force_virtual && !scatter ? &vf : nullptr;
  void* force = scatter ? lat->d_fsc[lat->fsc_cur] : lat->d_force;
  void* zero = scatter ? lat->d_fsc[lat->fsc_cur ^ 1] : nullptr;
  KernelTimer t(lat, LBSWIM_K_FUSED);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_step_hydro<double>(static_cast<const double*>(lat->d_f[lat->cur]), static_cast<const double*>(force),
                                  static_cast<double*>(lat->d_f[1 - lat->cur]), static_cast<double*>(lat->d_rho),
                                  static_cast<double*>(lat->d_u_aos), lat->geom, ts, tb, pvf, rows, lat->fuse_strips,
                                  stream(), static_cast<double*>(zero));
  else
    e = launch_step_hydro<float>(static_cast<const float*>(lat->d_f[lat->cur]), static_cast<const float*>(force),
                                 static_cast<float*>(lat->d_f[1 - lat->cur]), static_cast<float*>(lat->d_rho),
                                 static_cast<float*>(lat->d_u_aos), lat->geom, ts, tb, pvf, rows, lat->fuse_strips,
                                 stream(), static_cast<float*>(zero));
  t.done();
  HIP_TRY(e);
  lat->u_aos_valid = true;
  lat->u_virtual = true;
  lat->cur = 1 - lat->cur;
  lat->time_step++;
  return LBSWIM_OK;
}

static int lattice_zero_force_impl(lbswim_lattice* lat) {
  lat->force_virtual = false;
  KernelTimer t(lat, LBSWIM_K_ZERO);
  hipError_t e = launch_fill_zero(lat->d_force, 3 * lat->elem * lat->n, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

int lbswim_lattice_step(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  if (int rc = materialize_force(lat)) return rc;
  return lattice_step_impl(lat);
}

int lbswim_lattice_calc_hydro(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  if (int rc = materialize_force(lat)) return rc;
  return lattice_hydro_impl(lat);
}

int lbswim_lattice_set_force_mode(lbswim_lattice* lat, int mode) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(mode == LBSWIM_FORCE_STORED || mode == LBSWIM_FORCE_AUTO || mode == LBSWIM_FORCE_VIRTUAL ||
                mode == LBSWIM_FORCE_SCATTER,
            "unknown force mode %d", mode);
  CHECK_ARG(mode != LBSWIM_FORCE_SCATTER || lat->storage == LBSWIM_STORE_F32DEV,
            "the scatter force is for FP32 storage (its atomics round in a varying order)");
  if (mode == LBSWIM_FORCE_STORED)
    if (int rc = materialize_force(lat)) return rc;
  lat->force_mode = mode;
  return LBSWIM_OK;
}

int lbswim_lattice_get_force_mode(lbswim_lattice* lat, int* mode) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(mode != nullptr, "mode is NULL");
  *mode = lat->force_mode;
  return LBSWIM_OK;
}

int lbswim_lattice_sync_force(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  return materialize_force(lat);
}

int lbswim_lattice_force_is_virtual(lbswim_lattice* lat, int* is_virtual) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(is_virtual != nullptr, "is_virtual is NULL");
  *is_virtual = lat->force_virtual ? 1 : 0;
  return LBSWIM_OK;
}

int lbswim_lattice_sync_u(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  return materialize_u(lat);
}

int lbswim_lattice_u_is_virtual(lbswim_lattice* lat, int* is_virtual) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(is_virtual != nullptr, "is_virtual is NULL");
  *is_virtual = lat->u_virtual ? 1 : 0;
  return LBSWIM_OK;
}

int lbswim_lattice_set_fusion(lbswim_lattice* lat, int enable, int rows, int strips) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(enable == LBSWIM_FUSION_OFF || enable == LBSWIM_FUSION_AUTO || enable == LBSWIM_FUSION_ON,
            "fusion must be LBSWIM_FUSION_OFF, _AUTO or _ON (got %d)", enable);
  CHECK_ARG(rows == 0 || rows == 2 || rows == 4 || rows == 8, "rows must be 0, 2, 4 or 8 (got %d)", rows);
  CHECK_ARG(strips >= 0, "negative strips");
  lat->fuse = enable;
  lat->fuse_rows = rows;
  lat->fuse_strips = strips;
  return LBSWIM_OK;
}

int lbswim_lattice_fusion(lbswim_lattice* lat, int* active, int* rows, int* strips) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(active != nullptr, "active is NULL");
  const int ty = fusion_rows(lat);
  *active = ty > 0 ? 1 : 0;
  if (rows) *rows = ty;
  if (strips) *strips = lat->fuse_strips;
  return LBSWIM_OK;
}

int lbswim_lattice_init_from_hydro(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  if (int rc = materialize_u(lat)) return rc;
  KernelTimer t(lat, LBSWIM_K_INIT);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_init_from_hydro<double>(static_cast<const double*>(lat->d_rho), static_cast<const double*>(lat->d_u),
                                       static_cast<double*>(lat->d_f[lat->cur]), lat->geom, stream());
  else
    e = launch_init_from_hydro<float>(static_cast<const float*>(lat->d_rho), static_cast<const float*>(lat->d_u),
                                      static_cast<float*>(lat->d_f[lat->cur]), lat->geom, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

int lbswim_lattice_zero_force(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  return lattice_zero_force_impl(lat);
}

int lbswim_lattice_params(lbswim_lattice* lat, const lbswim_params** host) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(host != nullptr, "host is NULL");
  *host = &lat->params;
  return LBSWIM_OK;
}

int lbswim_lattice_addressing(lbswim_lattice* lat, const lbswim_addressing** host) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(host != nullptr, "host is NULL");
  *host = &lat->addr;
  return LBSWIM_OK;
}

int lbswim_lattice_get_time_step(lbswim_lattice* lat, int* out) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = lat->time_step;
  return LBSWIM_OK;
}

int lbswim_lattice_set_time_step(lbswim_lattice* lat, int ts) {
  if (int rc = check_lattice(lat)) return rc;
  lat->time_step = ts;
  return LBSWIM_OK;
}

int lbswim_lattice_storage(lbswim_lattice* lat, int* out) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = lat->storage;
  return LBSWIM_OK;
}

int lbswim_lattice_field(lbswim_lattice* lat, int field, double** host, void** device, size_t* size) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(field >= LBSWIM_RHO && field <= LBSWIM_FNEW, "unknown lattice field %d", field);
  if (host) {
    if (int rc = lattice_host(lat, field, host)) return rc;
  }
  if (device) {
    if (field == LBSWIM_FORCE)
      if (int rc = materialize_force(lat)) return rc;
    if (field == LBSWIM_U) {
      if (int rc = materialize_u(lat)) return rc;
      lat->u_aos_valid = false;  // the caller may write u through the pointer
    }
    *device = field_device(lat, field);
  }
  if (size) *size = field_count(lat, field);
  return LBSWIM_OK;
}

int lbswim_lattice_h2d(lbswim_lattice* lat, int field) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(field >= LBSWIM_RHO && field <= LBSWIM_FNEW, "unknown lattice field %d", field);
  double* h = nullptr;
  if (int rc = lattice_host(lat, field, &h)) return rc;
  const size_t cnt = field_count(lat, field);
  void* dev = field_device(lat, field);
  const size_t n = lat->n;
  if (field == LBSWIM_FORCE) lat->force_virtual = false;  // the host array replaces the force
  if (field == LBSWIM_U) lat->u_aos_valid = lat->u_virtual = false;  // the host array replaces u
  if (lat->storage == LBSWIM_STORE_F64) {
    if (!is_dist(field)) return copy_h2d(dev, h, cnt * sizeof(double));
    std::vector<double> stage(cnt);
    for (int p = 0; p < Q; ++p)
      for_each_upstream(lat, p, [&](size_t x, size_t up) { stage[p * n + up] = h[p * n + x]; });
    return copy_h2d(dev, stage.data(), cnt * sizeof(double));
  }
  std::vector<float> stage(cnt);
  if (!is_dist(field)) {
    for (size_t i = 0; i < cnt; ++i) stage[i] = float(h[i] - f32_offset(lat, field, i));
  } else {
    for (int p = 0; p < Q; ++p) {
      const double w = TAB.w[p];
      for_each_upstream(lat, p, [&](size_t x, size_t up) { stage[p * n + up] = float(h[p * n + x] - w); });
    }
  }
  return copy_h2d(dev, stage.data(), cnt * sizeof(float));
}

int lbswim_lattice_d2h(lbswim_lattice* lat, int field) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(field >= LBSWIM_RHO && field <= LBSWIM_FNEW, "unknown lattice field %d", field);
  double* h = nullptr;
  if (int rc = lattice_host(lat, field, &h)) return rc;
  if (field == LBSWIM_FORCE)
    if (int rc = materialize_force(lat)) return rc;
  if (field == LBSWIM_U)
    if (int rc = materialize_u(lat)) return rc;
  const size_t cnt = field_count(lat, field);
  const void* dev = field_device(lat, field);
  const size_t n = lat->n;
  if (lat->storage == LBSWIM_STORE_F64) {
    if (!is_dist(field)) return copy_d2h(h, dev, cnt * sizeof(double));
    std::vector<double> stage(cnt);
    if (int rc = copy_d2h(stage.data(), dev, cnt * sizeof(double))) return rc;
    for (int p = 0; p < Q; ++p)
      for_each_upstream(lat, p, [&](size_t x, size_t up) { h[p * n + x] = stage[p * n + up]; });
    return LBSWIM_OK;
  }
  std::vector<float> stage(cnt);
  if (int rc = copy_d2h(stage.data(), dev, cnt * sizeof(float))) return rc;
  if (!is_dist(field)) {
    for (size_t i = 0; i < cnt; ++i) h[i] = double(stage[i]) + f32_offset(lat, field, i);
  } else {
    for (int p = 0; p < Q; ++p) {
      const double w = TAB.w[p];
      for_each_upstream(lat, p, [&](size_t x, size_t up) { h[p * n + x] = double(stage[p * n + up]) + w; });
    }
  }
  return LBSWIM_OK;
}

// ---- particle storage ----------------------------------------------------------

}  // extern "C"

namespace {

// Moves between re-orderings of a particle array's storage.
constexpr int REORDER_INTERVAL = 64;

void pstore_free(PStore& st) {
  for (auto*& p : st.d) dev_free(p), p = nullptr;
  dev_free(st.stage);
  dev_free(st.perm);
  dev_free(st.perm_alt);
  dev_free(st.keys);
  dev_free(st.keys_alt);
  dev_free(st.idx);
  dev_free(st.order);
  dev_free(st.sort_tmp);
  for (void* h : st.h) std::free(h);
}

int pstore_alloc(PStore& st, int num, std::initializer_list<int> fields) {
  st.num = num;
  const size_t words = 3 * size_t(num);
  int rc;
  for (int f : fields) {
    st.has[f] = true;
    if ((rc = dev_alloc_zero(reinterpret_cast<void**>(&st.d[f]), 8 * words))) return rc;
  }
  if ((rc = dev_alloc_zero(reinterpret_cast<void**>(&st.stage), 8 * words)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&st.perm), 4 * size_t(num))) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&st.perm_alt), 4 * size_t(num))) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&st.keys), 4 * size_t(num))) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&st.keys_alt), 4 * size_t(num))) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&st.idx), 4 * size_t(num))) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&st.order), 4 * size_t(num))))
    return rc;
  if (num > 0) {
    HIP_TRY(sort_scratch_bytes(num, &st.sort_tmp_bytes));
    if ((rc = dev_alloc_zero(&st.sort_tmp, st.sort_tmp_bytes ? st.sort_tmp_bytes : 1))) return rc;
    HIP_TRY(launch_iota(num, st.perm, stream()));
  }
  return LBSWIM_OK;
}

int pstore_ncomp(int field) { return field == LBSWIM_PART_PRNG ? 1 : 3; }

int pstore_host(PStore& st, int field, void** out) {
  if (!st.h[field]) {
    const size_t cnt = size_t(pstore_ncomp(field)) * st.num;
    st.h[field] = std::calloc(cnt ? cnt : 1, 8);
    if (!st.h[field]) return fail(LBSWIM_ENOMEM, "host mirror allocation failed");
  }
  *out = st.h[field];
  return LBSWIM_OK;
}

// Host (user order) -> device (slot order).
int pstore_h2d(PStore& st, int field) {
  void* h = nullptr;
  if (int rc = pstore_host(st, field, &h)) return rc;
  const int nc = pstore_ncomp(field);
  if (int rc = copy_h2d(st.stage, h, 8 * size_t(nc) * st.num)) return rc;
  HIP_TRY(launch_gather<uint64_t>(st.num, nc, st.perm, st.stage, st.d[field], stream()));
  HIP_TRY(hipStreamSynchronize(stream()));
  if (field == LBSWIM_PART_R) st.dirty = true;
  return LBSWIM_OK;
}

// Device (slot order) -> host (user order).
int pstore_d2h(PStore& st, int field) {
  void* h = nullptr;
  if (int rc = pstore_host(st, field, &h)) return rc;
  const int nc = pstore_ncomp(field);
  HIP_TRY(launch_scatter<uint64_t>(st.num, nc, st.perm, st.d[field], st.stage, stream()));
  return copy_d2h(h, st.stage, 8 * size_t(nc) * st.num);
}

// Re-order the storage by the Morton code of each particle's site when the
// positions came from the host or REORDER_INTERVAL moves have passed.
int pstore_maybe_reorder(PStore& st, const Geom& g) {
  if (st.num <= 1 || (!st.dirty && st.moves_since_order < REORDER_INTERVAL)) return LBSWIM_OK;
  const double* r = reinterpret_cast<const double*>(st.d[LBSWIM_PART_R]);
  HIP_TRY(launch_spatial_order(st.num, r, g, st.keys, st.keys_alt, st.idx, st.order, st.sort_tmp,
                               st.sort_tmp_bytes, stream()));
  // Each field is gathered into the staging buffer and copied back, so every
  // field keeps its device buffer for the array's lifetime (as the
  // reference's SharedArray::device does, Shared.h:8): a device pointer
  // obtained once stays valid across re-orderings.
  for (int f = 0; f < 5; ++f) {
    if (!st.has[f]) continue;
    const size_t bytes = 8 * size_t(pstore_ncomp(f)) * st.num;
    HIP_TRY(launch_gather<uint64_t>(st.num, pstore_ncomp(f), st.order, st.d[f], st.stage, stream()));
    HIP_TRY(hipMemcpyAsync(st.d[f], st.stage, bytes, hipMemcpyDeviceToDevice, stream()));
  }
  HIP_TRY(launch_gather<int>(st.num, 1, st.order, st.perm, st.perm_alt, stream()));
  HIP_TRY(hipMemcpyAsync(st.perm, st.perm_alt, 4 * size_t(st.num), hipMemcpyDeviceToDevice, stream()));
  st.dirty = false;
  st.moves_since_order = 0;
  return LBSWIM_OK;
}

double* fld(PStore& st, int f) { return reinterpret_cast<double*>(st.d[f]); }

}  // namespace
extern "C" {

// ---- swimmers ---------------------------------------------------------------

int lbswim_swimmers_create(int n, const lbswim_common_params* p, lbswim_swimmers** out) {
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  CHECK_ARG(n >= 0, "swimmer count must be >= 0 (got %d)", n);
  CHECK_ARG(p != nullptr, "CommonParams is NULL");
  int rc = ensure_runtime();
  if (rc) return rc;
  lbswim_swimmers* sw = new (std::nothrow) lbswim_swimmers();
  if (!sw) return fail(LBSWIM_ENOMEM, "swimmer array allocation failed");
  sw->common_host = *p;
  sw->common_dev = *p;
  sw->version = next_version();
  if ((rc = pstore_alloc(sw->st, n, {LBSWIM_PART_R, LBSWIM_PART_V, LBSWIM_PART_N, LBSWIM_PART_PRNG})) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&sw->d_tumbles), sizeof(unsigned long long)))) {
    std::string msg = g_err;
    lbswim_swimmers_destroy(sw);
    g_err = msg;
    return rc;
  }
  HIP_TRY(hipStreamSynchronize(stream()));
  *out = sw;
  return LBSWIM_OK;
}

void lbswim_swimmers_destroy(lbswim_swimmers* sw) {
  if (!sw) return;
  if (g_rt.ready) (void)hipStreamSynchronize(stream());
  pstore_free(sw->st);
  dev_free(sw->d_tumbles);
  delete sw;
}

int lbswim_swimmers_num(const lbswim_swimmers* sw, int* out) {
  CHECK_ARG(sw && out, "NULL argument");
  *out = sw->st.num;
  return LBSWIM_OK;
}

int lbswim_swimmers_common(lbswim_swimmers* sw, lbswim_common_params** host) {
  CHECK_ARG(sw && host, "NULL argument");
  *host = &sw->common_host;
  return LBSWIM_OK;
}

int lbswim_swimmers_common_h2d(lbswim_swimmers* sw) {
  CHECK_ARG(sw, "NULL argument");
  sw->common_dev = sw->common_host;
  sw->version = next_version();
  return LBSWIM_OK;
}

int lbswim_swimmers_common_d2h(lbswim_swimmers* sw) {
  CHECK_ARG(sw, "NULL argument");
  sw->common_host = sw->common_dev;
  return LBSWIM_OK;
}

int lbswim_swimmers_add_forces(lbswim_swimmers* sw, lbswim_lattice* lat) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = check_lattice(lat)) return rc;
  if (int rc = materialize_force(lat)) return rc;
  KernelTimer t(lat, LBSWIM_K_ADDF);
  const SwimParams p = to_kernel(sw->common_dev);
  PStore& st = sw->st;
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_swimmers_add_forces<double>(st.num, p, fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_N),
                                           static_cast<double*>(lat->d_force), lat->geom, stream());
  else
    e = launch_swimmers_add_forces<float>(st.num, p, fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_N),
                                          static_cast<float*>(lat->d_force), lat->geom, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

// scatter (FP32, after a fused step): the moved swimmers' dipoles are also
// added into this force array.
static int swimmers_move_impl(lbswim_swimmers* sw, lbswim_lattice* lat, void* scatter = nullptr) {
  PStore& st = sw->st;
  if (int rc = pstore_maybe_reorder(st, lat->geom)) return rc;
  KernelTimer t(lat, LBSWIM_K_SMOVE);
  const SwimParams p = to_kernel(sw->common_dev);
  uint64_t* prng = st.d[LBSWIM_PART_PRNG];
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_swimmers_move<double>(st.num, p, fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_V),
                                     fld(st, LBSWIM_PART_N), prng, st.perm,
                                     static_cast<const double*>(lat->u_aos_valid ? lat->d_u_aos : lat->d_u),
                                     lat->u_aos_valid, lat->geom, sw->d_tumbles, stream(),
                                     static_cast<double*>(scatter));
  else
    e = launch_swimmers_move<float>(st.num, p, fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_V),
                                    fld(st, LBSWIM_PART_N), prng, st.perm,
                                    static_cast<const float*>(lat->u_aos_valid ? lat->d_u_aos : lat->d_u),
                                    lat->u_aos_valid, lat->geom, sw->d_tumbles, stream(),
                                    static_cast<float*>(scatter));
  t.done();
  HIP_TRY(e);
  st.moves_since_order++;
  sw->version = next_version();
  return LBSWIM_OK;
}

int lbswim_swimmers_move(lbswim_swimmers* sw, lbswim_lattice* lat) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = check_lattice(lat)) return rc;
  return swimmers_move_impl(sw, lat);
}

static int swimmer_check_field(int field) {
  CHECK_ARG(field == LBSWIM_PART_R || field == LBSWIM_PART_V || field == LBSWIM_PART_N ||
                field == LBSWIM_PART_PRNG,
            "unknown swimmer field %d", field);
  return LBSWIM_OK;
}

int lbswim_swimmers_field(lbswim_swimmers* sw, int field, void** host, void** device, size_t* count,
                          size_t* elem_size) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = swimmer_check_field(field)) return rc;
  if (host) {
    if (int rc = pstore_host(sw->st, field, host)) return rc;
  }
  if (device) {
    *device = sw->st.d[field];
    // the caller may write r or n through it
    if (field == LBSWIM_PART_R || field == LBSWIM_PART_N) sw->version = next_version();
  }
  if (count) *count = size_t(pstore_ncomp(field)) * sw->st.num;
  if (elem_size) *elem_size = 8;
  return LBSWIM_OK;
}

int lbswim_swimmers_h2d(lbswim_swimmers* sw, int field) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = swimmer_check_field(field)) return rc;
  if (field == LBSWIM_PART_R || field == LBSWIM_PART_N) sw->version = next_version();
  return pstore_h2d(sw->st, field);
}

int lbswim_swimmers_d2h(lbswim_swimmers* sw, int field) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = swimmer_check_field(field)) return rc;
  return pstore_d2h(sw->st, field);
}

int lbswim_swimmers_tumbles(lbswim_swimmers* sw, unsigned long long* out) {
  CHECK_ARG(sw && out, "NULL argument");
  return copy_d2h(out, sw->d_tumbles, sizeof(unsigned long long));
}

int lbswim_swimmers_slots(lbswim_swimmers* sw, const int** device) {
  CHECK_ARG(sw && device, "NULL argument");
  *device = sw->st.perm;
  return LBSWIM_OK;
}

// ---- tracers ----------------------------------------------------------------

int lbswim_tracers_create(int n, lbswim_tracers** out) {
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  CHECK_ARG(n >= 0, "tracer count must be >= 0 (got %d)", n);
  int rc = ensure_runtime();
  if (rc) return rc;
  lbswim_tracers* tr = new (std::nothrow) lbswim_tracers();
  if (!tr) return fail(LBSWIM_ENOMEM, "tracer array allocation failed");
  if ((rc = pstore_alloc(tr->st, n, {LBSWIM_PART_R, LBSWIM_PART_S, LBSWIM_PART_V}))) {
    std::string msg = g_err;
    lbswim_tracers_destroy(tr);
    g_err = msg;
    return rc;
  }
  HIP_TRY(hipStreamSynchronize(stream()));
  *out = tr;
  return LBSWIM_OK;
}

void lbswim_tracers_destroy(lbswim_tracers* tr) {
  if (!tr) return;
  if (g_rt.ready) (void)hipStreamSynchronize(stream());
  pstore_free(tr->st);
  delete tr;
}

int lbswim_tracers_num(const lbswim_tracers* tr, int* out) {
  CHECK_ARG(tr && out, "NULL argument");
  *out = tr->st.num;
  return LBSWIM_OK;
}

static int tracers_move_impl(lbswim_tracers* tr, lbswim_lattice* lat) {
  PStore& st = tr->st;
  if (int rc = pstore_maybe_reorder(st, lat->geom)) return rc;
  KernelTimer t(lat, LBSWIM_K_TMOVE);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_tracers_move<double>(st.num, fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_S), fld(st, LBSWIM_PART_V),
                                    static_cast<const double*>(lat->u_aos_valid ? lat->d_u_aos : lat->d_u),
                                    lat->u_aos_valid, lat->geom, stream());
  else
    e = launch_tracers_move<float>(st.num, fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_S), fld(st, LBSWIM_PART_V),
                                   static_cast<const float*>(lat->u_aos_valid ? lat->d_u_aos : lat->d_u),
                                   lat->u_aos_valid, lat->geom, stream());
  t.done();
  HIP_TRY(e);
  st.moves_since_order++;
  return LBSWIM_OK;
}

int lbswim_tracers_move(lbswim_tracers* tr, lbswim_lattice* lat) {
  CHECK_ARG(tr, "tracer handle is NULL");
  if (int rc = check_lattice(lat)) return rc;
  return tracers_move_impl(tr, lat);
}

static int tracer_check_field(int field) {
  CHECK_ARG(field == LBSWIM_PART_R || field == LBSWIM_PART_S || field == LBSWIM_PART_V, "unknown tracer field %d",
            field);
  return LBSWIM_OK;
}

int lbswim_tracers_field(lbswim_tracers* tr, int field, void** host, void** device, size_t* count) {
  CHECK_ARG(tr, "tracer handle is NULL");
  if (int rc = tracer_check_field(field)) return rc;
  if (host) {
    if (int rc = pstore_host(tr->st, field, host)) return rc;
  }
  if (device) *device = tr->st.d[field];
  if (count) *count = 3 * size_t(tr->st.num);
  return LBSWIM_OK;
}

int lbswim_tracers_h2d(lbswim_tracers* tr, int field) {
  CHECK_ARG(tr, "tracer handle is NULL");
  if (int rc = tracer_check_field(field)) return rc;
  return pstore_h2d(tr->st, field);
}

int lbswim_tracers_d2h(lbswim_tracers* tr, int field) {
  CHECK_ARG(tr, "tracer handle is NULL");
  if (int rc = tracer_check_field(field)) return rc;
  return pstore_d2h(tr->st, field);
}

int lbswim_tracers_slots(lbswim_tracers* tr, const int** device) {
  CHECK_ARG(tr && device, "NULL argument");
  *device = tr->st.perm;
  return LBSWIM_OK;
}

// ---- coupled step -------------------------------------------------------------

}  // extern "C"

namespace {

int grow(void** p, size_t* cap, size_t need, size_t elem) {
  if (*cap >= need) return LBSWIM_OK;
  HIP_TRY(hipStreamSynchronize(stream()));
  dev_free(*p);
  *p = nullptr;
  *cap = 0;
  if (int rc = dev_alloc_zero(p, elem * need)) return rc;
  *cap = need;
  return LBSWIM_OK;
}

// The collide and CalcHydro kernels (or the fused step) evaluate the force
// themselves, walking their lists every step, where the lists are short; a
// denser system writes the force array once per step instead
// (launch_site_force and the like), whose cost grows with the lattice.  A
// single system at 32^3-128^3 with 250-4000 swimmers measured the virtual
// force faster only up to about 2 (64^3), 3 (86^3), 4 (96^3) and beyond 4
// (128^3) mean entries per list: the limit taken is 1 + n / 2^18 entries (65
// at 256^3).  In an ensemble the virtual-force members share the batched
// lattice launches (the batched fused step among them) and the stored ones a
// chain beside them; there the limit is 8 entries: on the 64-member bench
// ensemble and the serving mix it measured 4-8 % faster than 32 and than
// 4 or 2, and 10-12 % faster than the single-system limit.
constexpr size_t VF_BATCH_MAX_MEAN_ENTRIES = 8;

bool use_virtual_force(int mode, size_t npts, size_t lists, size_t n, bool batched) {
  // a stencil reaches about 17 lists of 64 sites
  if (mode != LBSWIM_FORCE_AUTO) return mode == LBSWIM_FORCE_VIRTUAL;
  if (batched) return npts * 17 <= VF_BATCH_MAX_MEAN_ENTRIES * lists;
  return npts * 17 * (size_t(1) << 18) <= lists * ((size_t(1) << 18) + n);
}

// How a system's swimmer force is formed: virtual from its lists where
// sparse; otherwise stored, from all its points when they are few enough
// that every run walking them all costs less than building lists that would
// each hold many of them, else from lists.
int stored_force_path(size_t npts, const Geom& g) {
  // each site walks only its own points where every side is >= 4
  if (site_force_ok(g) && npts > SITE_FORCE_MIN_POINTS && npts < SITE_FORCE_MAX_POINTS) return FORCE_SITES_STORED;
  return npts <= RANGE_FORCE_MAX_POINTS ? FORCE_RANGE_STORED : FORCE_LISTS_STORED;
}

int force_path(int mode, size_t npts, const Geom& g, bool batched = false) {
  if (use_virtual_force(mode, npts, spread_lists(g), g.n, batched)) return FORCE_LISTS_VIRTUAL;
  return stored_force_path(npts, g);
}

// ZeroForce + AddForces of every swimmer array, as one deterministic gather.
// With the virtual force the lists are left for the next collide and CalcHydro.
int spread_forces(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw) {
  size_t npts = 0;
  for (int i = 0; i < nsw; ++i) npts += 2 * size_t(sw[i]->st.num);
  if (npts == 0) return lattice_zero_force_impl(lat);
  CHECK_ARG(npts <= size_t(INT_MAX) / 64, "too many swimmers in one system");
  if (!lat->sp_counts) {
    lat->sp_lists = spread_lists(lat->geom);
    const size_t nt = size_t(lat->sp_lists);
    int rc;
    if ((rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->sp_counts), 4 * nt)) ||
        (rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->sp_off), 4 * spread_off_entries(uint32_t(nt)))) ||
        (rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->sp_cursor), 4 * nt)))
      return rc;
  }
  size_t cap_bytes = lat->sp_pts_cap * sizeof(SpreadPoint);
  if (lat->sp_pts_cap < npts) {
    size_t slot_cap = lat->sp_pts_cap;
    if (int rc = grow(&lat->sp_pts, &cap_bytes, npts * sizeof(SpreadPoint), 1)) return rc;
    if (int rc = grow(reinterpret_cast<void**>(&lat->sp_slot), &slot_cap, npts, 4)) return rc;
    lat->sp_pts_cap = npts;
  }
  SpreadPoint* pts = static_cast<SpreadPoint*>(lat->sp_pts);
  int path = force_path(lat->force_mode, npts, lat->geom);
  // the fused launch walks a virtual force only where each 64-site run is one z-row
  if (path == FORCE_LISTS_VIRTUAL && lat->geom.nz % 64 != 0 && fusion_rows(lat) > 0)
    path = stored_force_path(npts, lat->geom);
  if (path == FORCE_SITES_STORED) {
    // dense: per-site segments of the points' stencil entries
    const size_t n = lat->geom.n, nent = 64 * npts;
    int rc;
    if (!lat->ss_counts &&
        ((rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->ss_counts), 4 * n)) ||
         (rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->ss_off), 4 * spread_off_entries(uint32_t(n))))))
      return rc;
    if (lat->ss_cap < nent) {
      size_t c1 = lat->ss_cap, c2 = lat->ss_cap, c4 = lat->ss_cap;
      if ((rc = grow(reinterpret_cast<void**>(&lat->ss_local), &c1, nent, 4)) ||
          (rc = grow(reinterpret_cast<void**>(&lat->ss_ent), &c2, nent, 8)) ||
          (rc = grow(reinterpret_cast<void**>(&lat->ss_sorted), &c4, nent, 3 * sizeof(double))))
        return rc;
      lat->ss_cap = nent;
    }
    KernelTimer t(lat, LBSWIM_K_ADDF);
    hipError_t e = hipSuccess;
    int offset = 0;
    for (int i = 0; i < nsw && e == hipSuccess; ++i) {
      PStore& st = sw[i]->st;
      e = launch_points(st.num, offset, to_kernel(sw[i]->common_dev), fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_N),
                        st.perm, lat->geom, pts, lat->sp_slot, stream());
      offset += st.num;
    }
    if (e == hipSuccess)
      e = lat->storage == LBSWIM_STORE_F64
              ? launch_site_force<double>(int(npts), lat->geom, pts, lat->sp_slot, lat->ss_counts, lat->ss_off,
                                          lat->ss_local, lat->ss_ent, lat->ss_sorted,
                                          static_cast<double*>(lat->d_force), stream())
              : launch_site_force<float>(int(npts), lat->geom, pts, lat->sp_slot, lat->ss_counts, lat->ss_off,
                                         lat->ss_local, lat->ss_ent, lat->ss_sorted,
                                         static_cast<float*>(lat->d_force), stream());
    t.done();
    HIP_TRY(e);
    lat->force_virtual = false;
    return LBSWIM_OK;
  }
  if (path == FORCE_RANGE_STORED) {
    // few points, dense: every run walks all of them, no lists
    KernelTimer t(lat, LBSWIM_K_ADDF);
    hipError_t e = hipSuccess;
    int offset = 0;
    for (int i = 0; i < nsw && e == hipSuccess; ++i) {
      PStore& st = sw[i]->st;
      e = launch_points(st.num, offset, to_kernel(sw[i]->common_dev), fld(st, LBSWIM_PART_R), fld(st, LBSWIM_PART_N),
                        st.perm, lat->geom, pts, lat->sp_slot, stream());
      offset += st.num;
    }
    if (e == hipSuccess)
      e = lat->storage == LBSWIM_STORE_F64
              ? launch_range_force<double>(lat->geom, pts, uint32_t(npts), static_cast<double*>(lat->d_force),
                                           stream())
              : launch_range_force<float>(lat->geom, pts, uint32_t(npts), static_cast<float*>(lat->d_force), stream());
    t.done();
    HIP_TRY(e);
    lat->force_virtual = false;
    return LBSWIM_OK;
  }
  const size_t lcap = spread_list_capacity(npts);
  if (lat->sp_list_cap < lcap) {
    size_t c1 = lat->sp_list_cap, c2 = lat->sp_list_cap;
    if (int rc = grow(reinterpret_cast<void**>(&lat->sp_list), &c1, lcap, 4)) return rc;
    if (int rc = grow(reinterpret_cast<void**>(&lat->sp_sorted), &c2, lcap, 4)) return rc;
    lat->sp_list_cap = lcap;
  }
  const bool virt = path == FORCE_LISTS_VIRTUAL;

  KernelTimer t(lat, LBSWIM_K_ADDF);
  int offset = 0;
  for (int i = 0; i < nsw; ++i) {
    PStore& st = sw[i]->st;
    HIP_TRY(launch_points(st.num, offset, to_kernel(sw[i]->common_dev), fld(st, LBSWIM_PART_R),
                          fld(st, LBSWIM_PART_N), st.perm, lat->geom, pts, lat->sp_slot, stream()));
    offset += st.num;
  }
  // the fused step's z-halo sites read the force at the ends of the runs, its
  // row walks the lists' entries as RunEntry records
  const bool edges = virt && fusion_rows(lat) > 0;
  if (edges && !lat->sp_edge)
    if (int rc = dev_alloc_zero(reinterpret_cast<void**>(&lat->sp_edge), 6 * sizeof(double) * lat->sp_lists))
      return rc;
  if (edges && lat->sp_rec_cap < lcap) {
    size_t c1 = lat->sp_rec_cap, c2 = lat->sp_rec_cap;
    if (int rc = grow(reinterpret_cast<void**>(&lat->sp_rec), &c1, lcap, sizeof(RunEntry))) return rc;
    if (int rc = grow(reinterpret_cast<void**>(&lat->sp_wz), &c2, lcap, sizeof(int))) return rc;
    lat->sp_rec_cap = lcap;
  }
  hipError_t e = launch_spread_lists(int(npts), lat->geom, pts, lat->sp_slot, lat->sp_counts, lat->sp_off,
                                     lat->sp_cursor,
                                     lat->sp_list, lat->sp_sorted, stream());
  if (e == hipSuccess && edges)
    e = launch_edge_force(lat->geom, pts, lat->sp_off, lat->sp_sorted, lat->sp_edge, lat->sp_rec, lat->sp_wz, stream());
  if (e == hipSuccess && !virt) {
    if (lat->storage == LBSWIM_STORE_F64)
      e = launch_spread_force<double>(lat->geom, pts, lat->sp_off, lat->sp_sorted, static_cast<double*>(lat->d_force),
                                      stream());
    else
      e = launch_spread_force<float>(lat->geom, pts, lat->sp_off, lat->sp_sorted, static_cast<float*>(lat->d_force),
                                     stream());
  }
  t.done();
  HIP_TRY(e);
  lat->force_virtual = virt;
  lat->force_src = FORCE_SRC_LISTS;
  return LBSWIM_OK;
}

// The coupled step takes the scatter force: an FP32 lattice set to
// LBSWIM_FORCE_SCATTER, stepped by the fused launch, with swimmers.
bool use_scatter(const lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw) {
  if (lat->force_mode != LBSWIM_FORCE_SCATTER || lat->storage != LBSWIM_STORE_F32DEV || fusion_rows(lat) == 0)
    return false;
  for (int i = 0; i < nsw; ++i)
    if (sw[i]->st.num > 0) return true;
  return false;
}

void scatter_record(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw) {
  lat->fsc_set.clear();
  for (int i = 0; i < nsw; ++i) lat->fsc_set.emplace_back(sw[i], sw[i]->version);
  lat->fsc_valid = true;
}

// Makes d_fsc[fsc_cur] hold ZeroForce + AddForces of the swimmers as they
// are now: kept from the last step's moves if no array has changed since
// (same arrays, same versions), else cleared and scattered here.
int scatter_prepare(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw) {
  const size_t bytes = 3 * lat->elem * lat->n;
  for (int k = 0; k < 2; ++k)
    if (!lat->d_fsc[k]) {
      if (int rc = dev_alloc_zero(&lat->d_fsc[k], bytes)) return rc;
      lat->fsc_valid = false;
    }
  bool same = lat->fsc_valid && lat->fsc_set.size() == size_t(nsw);
  for (int i = 0; same && i < nsw; ++i) same = lat->fsc_set[i].first == sw[i] && lat->fsc_set[i].second == sw[i]->version;
  if (same) return LBSWIM_OK;
  lat->fsc_valid = false;
  KernelTimer t(lat, LBSWIM_K_ADDF);
  float* f = static_cast<float*>(lat->d_fsc[lat->fsc_cur]);
  hipError_t e = launch_fill_zero(f, bytes, stream());
  for (int i = 0; i < nsw && e == hipSuccess; ++i) {
    PStore& st = sw[i]->st;
    e = launch_swimmers_add_forces<float>(st.num, to_kernel(sw[i]->common_dev), fld(st, LBSWIM_PART_R),
                                          fld(st, LBSWIM_PART_N), f, lat->geom, stream());
  }
  t.done();
  HIP_TRY(e);
  scatter_record(lat, sw, nsw);
  return LBSWIM_OK;
}

}  // namespace
extern "C" {

int lbswim_system_step(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr, int ntr,
                       int nsteps) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(nsw >= 0 && ntr >= 0 && nsteps >= 0, "negative count");
  CHECK_ARG(nsw == 0 || sw != nullptr, "swimmer list is NULL");
  CHECK_ARG(ntr == 0 || tr != nullptr, "tracer list is NULL");
  for (int i = 0; i < nsw; ++i) CHECK_ARG(sw[i] != nullptr, "swimmer handle %d is NULL", i);
  for (int i = 0; i < ntr; ++i) CHECK_ARG(tr[i] != nullptr, "tracer handle %d is NULL", i);
  // With the fused step the tracer moves of step s run on the side stream
  // while the main stream goes on with the swimmer moves and the spreading of
  // step s+1; the fused launch of step s+1, which rewrites the interleaved u
  // they read, waits for them.  Nothing else reads the tracers inside the
  // call; the main stream waits for the side stream before returning (also on
  // error), so later calls are ordered.
  const bool overlap = ntr > 0 && nsteps > 0 && fusion_rows(lat) > 0;
  if (overlap)
    if (int rc = ensure_side_stream()) return rc;
  bool side_used = false;
  auto join = [&]() -> int {
    if (!side_used) return LBSWIM_OK;
    side_used = false;
    HIP_TRY(hipEventRecord(g_rt.joined, g_rt.side));
    HIP_TRY(hipStreamWaitEvent(g_rt.stream, g_rt.joined, 0));
    lat->tracers_reading = false;
    return LBSWIM_OK;
  };
  // Scatter force (FP32): each step's force was added by the previous step's
  // swimmer moves into d_fsc[fsc_cur]; the fused launch reads it and clears
  // the other buffer, into which this step's moves add the next force.
  const bool scatter = use_scatter(lat, sw, nsw);
  int rc = LBSWIM_OK;
  for (int s = 0; s < nsteps && !rc; ++s) {
    if (scatter) {
      if ((rc = scatter_prepare(lat, sw, nsw))) break;
      if ((rc = lattice_step_hydro_impl(lat, fusion_rows(lat), true))) break;
      lat->fsc_last = lat->fsc_cur;
      lat->force_virtual = true;
      lat->force_src = FORCE_SRC_SCATTER;
      lat->fsc_valid = false;  // until this step's moves have added the next force
    } else {
      if ((rc = spread_forces(lat, sw, nsw))) break;
      if (const int rows = fusion_rows(lat)) {
        if ((rc = lattice_step_hydro_impl(lat, rows))) break;
      } else {
        if ((rc = lattice_step_impl(lat))) break;
        if ((rc = lattice_hydro_impl(lat, true))) break;
      }
    }
    if (overlap) {
      if (!lat->tracers_done && hipEventCreateWithFlags(&lat->tracers_done, hipEventDisableTiming) != hipSuccess) {
        rc = fail(LBSWIM_EHIP, "hipEventCreateWithFlags failed");
        break;
      }
      if (hipEventRecord(g_rt.u_ready, g_rt.stream) != hipSuccess ||
          hipStreamWaitEvent(g_rt.side, g_rt.u_ready, 0) != hipSuccess) {
        rc = fail(LBSWIM_EHIP, "ordering the tracer moves after the fused step failed");
        break;
      }
      side_used = true;
      {
        StreamScope on_side(g_rt.side);
        for (int i = 0; i < ntr && !rc; ++i) rc = tracers_move_impl(tr[i], lat);
      }
      if (rc) break;
      if (hipEventRecord(lat->tracers_done, g_rt.side) != hipSuccess) {
        rc = fail(LBSWIM_EHIP, "hipEventRecord failed");
        break;
      }
      lat->tracers_reading = true;
    }
    for (int i = 0; i < nsw && !rc; ++i) rc = swimmers_move_impl(sw[i], lat, scatter ? lat->d_fsc[lat->fsc_cur ^ 1] : nullptr);
    if (rc) break;
    if (scatter) {
      lat->fsc_cur ^= 1;
      scatter_record(lat, sw, nsw);
    }
    if (!overlap)
      for (int i = 0; i < ntr && !rc; ++i) rc = tracers_move_impl(tr[i], lat);
  }
  const int jrc = join();
  return rc ? rc : jrc;
}

// ---- autotuning of the coupled step ------------------------------------------

int lbswim_system_autotune(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr,
                           int ntr, int steps_per_trial, int* rows, int* strips) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(steps_per_trial > 0, "steps_per_trial must be > 0");
  int cus = 0, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  // candidates: the two launches (rows 0), then fused tiles of 8, 4, 2 rows
  // with 1, 2 and 4 times the strips that give every CU one tile; for an FP32
  // lattice left to choose (LBSWIM_FORCE_AUTO or _SCATTER) with swimmers,
  // each fused shape also with the scatter force
  struct Choice {
    int rows, strips, force;
  };
  const int mode0 = lat->force_mode == LBSWIM_FORCE_SCATTER ? LBSWIM_FORCE_AUTO : lat->force_mode;
  bool any_swimmers = false;
  for (int i = 0; i < nsw; ++i) any_swimmers = any_swimmers || (sw[i] && sw[i]->st.num > 0);
  const bool try_scatter = lat->storage == LBSWIM_STORE_F32DEV && any_swimmers &&
                           (lat->force_mode == LBSWIM_FORCE_AUTO || lat->force_mode == LBSWIM_FORCE_SCATTER);
  std::vector<Choice> cand{{0, 0, mode0}};
  const Geom& g = lat->geom;
  for (int ty : {8, 4, 2}) {
    if (!fused_step_supported(g, ty, lat->elem)) continue;
    const int tiles = (g.ny + ty - 1) / ty * std::max(1, (g.nz + 63) / 64);
    const int base = std::min(g.nx, std::max(1, (cus + tiles - 1) / tiles));
    int prev = 0;
    for (int k : {1, 2, 4}) {
      const int st = std::min(g.nx, base * k);
      if (st == prev) continue;
      prev = st;
      cand.push_back(Choice{ty, st, mode0});
      if (try_scatter) cand.push_back(Choice{ty, st, LBSWIM_FORCE_SCATTER});
    }
  }
  const int saved_fuse = lat->fuse, saved_rows = lat->fuse_rows, saved_strips = lat->fuse_strips;
  const int saved_mode = lat->force_mode;
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return fail(LBSWIM_EHIP, "hipEventCreate failed");
  }
  int rc = LBSWIM_OK;
  auto trial = [&](const Choice& c, int steps, float* ms) -> int {
    lat->fuse = c.rows ? LBSWIM_FUSION_ON : LBSWIM_FUSION_OFF;
    lat->fuse_rows = c.rows;
    lat->fuse_strips = c.strips;
    lat->force_mode = c.force;
    if (int r = lbswim_system_step(lat, sw, nsw, tr, ntr, 2)) return r;  // warm-up, workspace
    (void)hipEventRecord(a, stream());
    if (int r = lbswim_system_step(lat, sw, nsw, tr, ntr, steps)) return r;
    (void)hipEventRecord(b, stream());
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(ms, a, b) != hipSuccess)
      return fail(LBSWIM_EHIP, "timing the coupled step failed");
    *ms /= float(steps);
    return LBSWIM_OK;
  };
  // Round 1: every candidate once.  Round 2: the three fastest again, three
  // times each in turn with twice the steps, ranked by their minimum: single
  // trials of shapes within a few per cent of each other rank by noise.
  std::vector<float> first(cand.size(), 0.f);
  for (size_t i = 0; i < cand.size() && !rc; ++i) rc = trial(cand[i], steps_per_trial, &first[i]);
  std::vector<size_t> rank(cand.size());
  for (size_t i = 0; i < rank.size(); ++i) rank[i] = i;
  std::stable_sort(rank.begin(), rank.end(), [&](size_t x, size_t y) { return first[x] < first[y]; });
  const size_t finalists = std::min<size_t>(3, rank.size());
  std::vector<float> best_of(finalists, 0.f);
  for (int rep = 0; rep < 3 && !rc && finalists > 1; ++rep)
    for (size_t f = 0; f < finalists && !rc; ++f) {
      float ms = 0.f;
      rc = trial(cand[rank[f]], 2 * steps_per_trial, &ms);
      if (!rc && (rep == 0 || ms < best_of[f])) best_of[f] = ms;
    }
  Choice best = cand[rank[0]];
  if (finalists > 1) {
    size_t bf = 0;
    for (size_t f = 1; f < finalists; ++f)
      if (best_of[f] < best_of[bf]) bf = f;
    best = cand[rank[bf]];
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) {
    lat->fuse = saved_fuse;
    lat->fuse_rows = saved_rows;
    lat->fuse_strips = saved_strips;
    lat->force_mode = saved_mode;
    return rc;
  }
  lat->fuse = best.rows ? LBSWIM_FUSION_ON : LBSWIM_FUSION_OFF;
  lat->fuse_rows = best.rows;
  lat->fuse_strips = best.strips;
  lat->force_mode = best.force;
  if (rows) *rows = best.rows;
  if (strips) *strips = best.strips;
  return LBSWIM_OK;
}

// ---- coupled step without allocation -----------------------------------------

}  // extern "C"

namespace {

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// The spreading workspace of one lattice carved out of the fixed device
// workspace: counts, offsets (+ scan and sort scratch), cursor, point
// records, list and sorted entries.
struct FixedSpread {
  uint32_t* counts = nullptr;
  uint32_t* off = nullptr;
  uint32_t* cursor = nullptr;
  SpreadPoint* pts = nullptr;
  uint32_t* slot = nullptr;
  uint32_t* list = nullptr;
  uint32_t* sorted = nullptr;
  size_t bytes = 0;
};

FixedSpread carve_spread(unsigned char* base, uint32_t nl, size_t npts) {
  FixedSpread w;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base ? base + at : nullptr;
    at += align256(bytes);
    return p;
  };
  w.counts = reinterpret_cast<uint32_t*>(take(4 * size_t(nl)));
  w.off = reinterpret_cast<uint32_t*>(take(4 * spread_off_entries(nl)));
  w.cursor = reinterpret_cast<uint32_t*>(take(4 * size_t(nl)));
  w.pts = reinterpret_cast<SpreadPoint*>(take(sizeof(SpreadPoint) * npts));
  w.slot = reinterpret_cast<uint32_t*>(take(4 * npts));
  w.list = reinterpret_cast<uint32_t*>(take(4 * spread_list_capacity(npts)));
  w.sorted = reinterpret_cast<uint32_t*>(take(4 * spread_list_capacity(npts)));
  w.bytes = at;
  return w;
}

}  // namespace
extern "C" {

int lbswim_system_step_noalloc(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr,
                               int ntr, int nsteps) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(nsw >= 0 && ntr >= 0 && nsteps >= 0, "negative count");
  CHECK_ARG(nsw == 0 || sw != nullptr, "swimmer list is NULL");
  CHECK_ARG(ntr == 0 || tr != nullptr, "tracer list is NULL");
  for (int i = 0; i < nsw; ++i) CHECK_ARG(sw[i] != nullptr, "swimmer handle %d is NULL", i);
  for (int i = 0; i < ntr; ++i) CHECK_ARG(tr[i] != nullptr, "tracer handle %d is NULL", i);
  size_t npts = 0;
  for (int i = 0; i < nsw; ++i) npts += 2 * size_t(sw[i]->st.num);
  const uint32_t nl = spread_lists(lat->geom);
  const size_t need = npts ? carve_spread(nullptr, nl, npts).bytes : 0;
  CHECK_ARG(need <= NOALLOC_SCRATCH_BYTES,
            "the no-allocation step needs %zu bytes of workspace for %zu swimmer ends on %u lists; its fixed "
            "workspace holds %zu bytes (use lbswim_system_step)",
            need, npts, nl, NOALLOC_SCRATCH_BYTES);
  if (nsteps == 0) return LBSWIM_OK;
  void* base = nullptr;
  HIP_TRY(noalloc_scratch(&base));
  const FixedSpread w = carve_spread(static_cast<unsigned char*>(base), nl, npts);
  // The lists' counts must start at zero; the workspace keeps what the last
  // call (possibly of another lattice) left in it.  A fill kernel, not a
  // memset, so that a caller may capture the call into a graph.
  if (npts) HIP_TRY(launch_fill_zero(w.counts, 4 * size_t(nl), stream()));
  const bool fused = lat->d_u_aos != nullptr && fusion_rows(lat) > 0;
  int rc = LBSWIM_OK;
  for (int s = 0; s < nsteps && !rc; ++s) {
    // ZeroForce + AddForces: the force array written from the sorted lists
    if (npts) {
      KernelTimer t(lat, LBSWIM_K_ADDF);
      int base_id = 0;
      hipError_t e = hipSuccess;
      for (int i = 0; i < nsw && e == hipSuccess; ++i) {
        PStore& st = sw[i]->st;
        e = launch_points(st.num, base_id, to_kernel(sw[i]->common_dev), fld(st, LBSWIM_PART_R),
                          fld(st, LBSWIM_PART_N), st.perm, lat->geom, w.pts, w.slot, stream());
        base_id += st.num;
      }
      if (e == hipSuccess)
        e = launch_spread_lists(int(npts), lat->geom, w.pts, w.slot, w.counts, w.off, w.cursor, w.list, w.sorted,
                                stream());
      if (e == hipSuccess)
        e = lat->storage == LBSWIM_STORE_F64
                ? launch_spread_force<double>(lat->geom, w.pts, w.off, w.sorted, static_cast<double*>(lat->d_force),
                                              stream())
                : launch_spread_force<float>(lat->geom, w.pts, w.off, w.sorted, static_cast<float*>(lat->d_force),
                                             stream());
      t.done();
      HIP_TRY(e);
      lat->force_virtual = false;
    } else if ((rc = lattice_zero_force_impl(lat))) {
      break;
    }
    if (fused) {
      if ((rc = lattice_step_hydro_impl(lat, fusion_rows(lat)))) break;
    } else {
      if ((rc = lattice_step_impl(lat))) break;
      // interleaved u only where its buffer exists: this entry point allocates nothing
      if ((rc = lattice_hydro_impl(lat, lat->d_u_aos != nullptr))) break;
    }
    for (int i = 0; i < nsw && !rc; ++i) rc = swimmers_move_impl(sw[i], lat);
    for (int i = 0; i < ntr && !rc; ++i) rc = tracers_move_impl(tr[i], lat);
  }
  return rc;
}

// ---- diagnostics ----------------------------------------------------------------

int lbswim_lattice_totals(lbswim_lattice* lat, double out[4]) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(out != nullptr, "out is NULL");
  double* partial = lat->d_partial;
  double* dres = lat->d_partial + 4 * size_t(lat->totals_blocks);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_totals<double>(static_cast<const double*>(lat->d_f[lat->cur]), lat->geom, partial, lat->totals_blocks,
                              dres, stream());
  else
    e = launch_totals<float>(static_cast<const float*>(lat->d_f[lat->cur]), lat->geom, partial, lat->totals_blocks,
                             dres, stream());
  HIP_TRY(e);
  return copy_d2h(out, dres, 4 * sizeof(double));
}

int lbswim_lattice_set_timing(lbswim_lattice* lat, int on) {
  return lbswim_lattice_set_timing_mask(lat, on ? (1u << LBSWIM_K_COUNT) - 1u : 0u);
}

int lbswim_lattice_set_timing_mask(lbswim_lattice* lat, unsigned mask) {
  if (int rc = check_lattice(lat)) return rc;
  if (mask) {
    HIP_TRY(hipStreamSynchronize(stream()));
    for (auto& r : lat->recs) {
      lat->pool.push_back(r.start);
      lat->pool.push_back(r.stop);
    }
    lat->recs.clear();
  }
  lat->timing = mask & ((1u << LBSWIM_K_COUNT) - 1u);
  return LBSWIM_OK;
}

int lbswim_lattice_kernel_time(lbswim_lattice* lat, int kernel, double* total_ms, long* count) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(kernel >= 0 && kernel < LBSWIM_K_COUNT, "unknown kernel id %d", kernel);
  CHECK_ARG(total_ms && count, "NULL output");
  HIP_TRY(hipStreamSynchronize(stream()));
  double tot = 0.0;
  long cnt = 0;
  for (const auto& r : lat->recs) {
    if (r.kernel != kernel) continue;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.start, r.stop));
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *count = cnt;
  return LBSWIM_OK;
}

int lbswim_device_memory(size_t* free_bytes, size_t* total_bytes) {
  CHECK_ARG(free_bytes && total_bytes, "NULL argument");
  if (int rc = ensure_runtime()) return rc;
  HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
  return LBSWIM_OK;
}

int lbswim_diag_copy(size_t count, int reps, double* ms) {
  CHECK_ARG(count > 0 && reps > 0 && ms, "invalid argument");
  if (int rc = ensure_runtime()) return rc;
  void *a = nullptr, *b = nullptr;
  int rc = dev_alloc_zero(&a, count * sizeof(double));
  if (!rc) rc = dev_alloc_zero(&b, count * sizeof(double));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(LBSWIM_EHIP, "event creation failed");
  if (!rc) {
    (void)launch_copy(static_cast<double*>(a), static_cast<double*>(b), count, stream());
    (void)hipEventRecord(e0, stream());
    for (int r = 0; r < reps; ++r) (void)launch_copy(static_cast<double*>(a), static_cast<double*>(b), count, stream());
    (void)hipEventRecord(e1, stream());
    hipError_t he = hipEventSynchronize(e1);
    float t = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
    if (he != hipSuccess) rc = fail(LBSWIM_EHIP, "copy probe failed: %s", hipGetErrorString(he));
    *ms = double(t) / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(stream());
  dev_free(a);
  dev_free(b);
  return rc;
}

int lbswim_diag_peskin(const double* x, double* out, int count) {
  CHECK_ARG(x && out && count >= 0, "invalid argument");
  if (count == 0) return LBSWIM_OK;
  if (int rc = ensure_runtime()) return rc;
  void *dx = nullptr, *dout = nullptr;
  int rc = dev_alloc_zero(&dx, sizeof(double) * count);
  if (!rc) rc = dev_alloc_zero(&dout, sizeof(double) * count);
  if (!rc) rc = copy_h2d(dx, x, sizeof(double) * count);
  if (!rc) {
    hipError_t e = launch_peskin_probe(static_cast<double*>(dx), static_cast<double*>(dout), count, stream());
    if (e != hipSuccess) rc = fail(LBSWIM_EHIP, "peskin probe launch failed: %s", hipGetErrorString(e));
  }
  if (!rc) rc = copy_d2h(out, dout, sizeof(double) * count);
  dev_free(dx);
  dev_free(dout);
  return rc;
}

int lbswim_diag_interp(int nx, int ny, int nz, const double* u, const double* r, double* v, int count) {
  CHECK_ARG(u && r && v && count >= 0 && nx >= 1 && ny >= 1 && nz >= 1, "invalid argument");
  if (count == 0) return LBSWIM_OK;
  if (int rc = ensure_runtime()) return rc;
  const size_t n = size_t(nx) * ny * nz;
  void *du = nullptr, *dr = nullptr, *dv = nullptr;
  int rc = dev_alloc_zero(&du, 3 * sizeof(double) * n);
  if (!rc) rc = dev_alloc_zero(&dr, 3 * sizeof(double) * count);
  if (!rc) rc = dev_alloc_zero(&dv, 3 * sizeof(double) * count);
  if (!rc) rc = copy_h2d(du, u, 3 * sizeof(double) * n);
  if (!rc) rc = copy_h2d(dr, r, 3 * sizeof(double) * count);
  if (!rc) {
    hipError_t e = launch_interp_probe(static_cast<double*>(du), static_cast<double*>(dr), static_cast<double*>(dv),
                                       count, make_geom(nx, ny, nz), stream());
    if (e != hipSuccess) rc = fail(LBSWIM_EHIP, "interp probe launch failed: %s", hipGetErrorString(e));
  }
  if (!rc) rc = copy_d2h(v, dv, 3 * sizeof(double) * count);
  dev_free(du);
  dev_free(dr);
  dev_free(dv);
  return rc;
}

}  // extern "C"

// ============================================================================
// Ensembles: many independent systems advanced by batched launches.
// ============================================================================

namespace {

struct EnsMember {
  lbswim_lattice* lat = nullptr;  // nullptr: free slot
  std::vector<lbswim_swimmers*> sw;
  std::vector<lbswim_tracers*> tr;
  // lattice buffer role captured in the device tables (a member stepped
  // outside the ensemble may change it; particle buffers never move)
  int cur_at_build = 0;
  bool fused = false;  // advanced by the batched fused step (u written interleaved)
};

// Workgroup orders the autotuner chooses between (results do not depend on it).
enum { ORDER_MEMBER_MAJOR = 0, ORDER_ROUND_ROBIN = 1, ORDER_LARGEST_FIRST = 2, ORDER_COUNT = 3 };

// Capacity for `need` elements when `cap` does not suffice: at least double,
// so that the (device-synchronising) reallocations stay rare as membership
// grows.
size_t grown(size_t cap, size_t need) { return std::max(need, 2 * cap); }


// Reorders per-member blocks: member-major keeps each member contiguous,
// round-robin interleaves members block by block, largest-first is
// member-major with the biggest lattices first.
std::vector<WorkItem> order_work(const std::vector<std::vector<WorkItem>>& per, int order,
                                 const std::vector<size_t>& weight) {
  std::vector<WorkItem> out;
  std::vector<size_t> idx(per.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
  if (order == ORDER_LARGEST_FIRST)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return weight[a] > weight[b]; });
  if (order == ORDER_ROUND_ROBIN) {
    size_t most = 0;
    for (auto& v : per) most = std::max(most, v.size());
    for (size_t k = 0; k < most; ++k)
      for (size_t i : idx)
        if (k < per[i].size()) out.push_back(per[i][k]);
  } else {
    for (size_t i : idx) out.insert(out.end(), per[i].begin(), per[i].end());
  }
  return out;
}

}  // namespace

struct lbswim_ensemble {
  int storage = LBSWIM_STORE_F64;
  std::vector<EnsMember> members;
  bool dirty = true;  // device tables out of date
  int parity = 0;
  int order = ORDER_MEMBER_MAJOR;
  int per = 1;  // 256-site lattice blocks per work item (workgroup) of the batched launches
  bool use_graph = true;
  // device tables
  LatDesc* d_lat = nullptr;
  size_t cap_lat = 0;
  // lattice blocks of the members whose force is stored (written before the
  // collide) and of those whose force is virtual (evaluated in the collide
  // and CalcHydro, written from the lists when a call returns), in the order
  // the autotuner chose
  WorkItem* d_storedwork = nullptr;
  size_t cap_storedwork = 0;
  int n_storedwork = 0;
  // of those, the blocks whose force a separate pass writes (the lists and
  // range members; the collide sums the site members' segments itself)
  WorkItem* d_spreadwork = nullptr;
  size_t cap_spreadwork = 0;
  int n_spreadwork = 0;
  WorkItem* d_vfwork = nullptr;
  size_t cap_vfwork = 0;
  int n_vfwork = 0;
  SwimDesc* d_sw = nullptr;
  size_t cap_sw = 0;
  WorkItem* d_swwork = nullptr;
  size_t cap_swwork = 0;
  int n_swwork = 0;
  WorkItem* d_ptwork = nullptr;
  size_t cap_ptwork = 0;
  int n_ptwork = 0;
  TracerDesc* d_tr = nullptr;
  size_t cap_tr = 0;
  WorkItem* d_trwork = nullptr;
  size_t cap_trwork = 0;
  int n_trwork = 0;
  // Batched fused step (collide/stream + CalcHydro in one launch) of the
  // virtual-force members with full tiles: rows per tile (8 or 4, 0 off) and
  // planes per strip.  Work: tile-strips, 64-list chunks for the RunEntry
  // records, and their lattice blocks (force written when a call returns).
  int fuse_rows = 8;
  int fuse_strip = 32;
  WorkItem* d_fusework = nullptr;
  size_t cap_fusework = 0;
  int n_fusework = 0;
  WorkItem* d_entrywork = nullptr;
  size_t cap_entrywork = 0;
  int n_entrywork = 0;
  WorkItem* d_fuseblk = nullptr;
  size_t cap_fuseblk = 0;
  int n_fuseblk = 0;
  RunEntry* rec = nullptr;
  int* wz = nullptr;
  size_t cap_rec = 0;
  double* edge = nullptr;
  size_t cap_edge = 0;
  size_t sites_fused = 0;
  int nlists = 0;  // spreading lists of all members
  // spreading workspace
  int npts = 0;
  int force_mode = LBSWIM_FORCE_AUTO;
  uint32_t* counts = nullptr;
  size_t cap_counts = 0;
  uint32_t* off = nullptr;
  size_t cap_off = 0;
  uint32_t* cursor = nullptr;
  size_t cap_cursor = 0;
  uint32_t* list = nullptr;
  uint32_t* sorted = nullptr;
  size_t cap_list = 0;
  SpreadPoint* pts = nullptr;
  size_t cap_pts = 0;
  uint32_t* slot_ids = nullptr;  // point ids in storage order
  size_t cap_slot = 0;
  bool dense_lists = false;      // aggregate the list building's atomics
  // site segments of the FORCE_SITES_STORED members (launch_batch_site_lists)
  int nsites = 0;
  uint32_t* ss_counts = nullptr;
  size_t cap_ss_counts = 0;
  uint32_t* ss_off = nullptr;
  size_t cap_ss_off = 0;
  uint32_t* ss_local = nullptr;
  size_t cap_ss_local = 0;
  uint32_t* ss_ent = nullptr;  // key and site of each entry
  size_t cap_ss_ent = 0;
  int ss_nent = 0;  // 64 points of those members
  double* ss_sorted = nullptr;
  size_t cap_ss_sorted = 0;
  // stored-force members' chain forked to the ensemble's side stream (see
  // ens_enqueue_step)
  bool fork_stored = false;
  bool stored_lists = false;  // some stored-force member builds lists
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join = nullptr;
  // graphs are captured on a stream of the ensemble's own, so that other
  // threads' work on the library stream cannot enter or break a capture
  hipStream_t cap = nullptr;
  // two captured steps, one graph for each starting parity (the parity
  // changes when a call steps an odd number of times)
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  hipGraph_t graph[2] = {nullptr, nullptr};
  // per-phase event timing (steps launched eagerly while on)
  unsigned timing = 0;
  std::vector<TimeRec> recs;
  std::vector<hipEvent_t> pool;
  size_t sites_virtual = 0, sites_stored = 0;
  // Lanes: with lanes > 1 the members are split into that many
  // sub-ensembles of about equal sites, each stepped by its own batched
  // launches on a stream of its own, forked from and joined to the caller's
  // stream once per step (a graph capture records them as parallel
  // branches): the latency-bound list building and moves of one lane run
  // beside the lattice sweeps of the others.
  int lanes = 1;
  std::vector<lbswim_ensemble*> sub;
  std::vector<hipStream_t> lane_stream;
  std::vector<hipEvent_t> lane_join;
  hipEvent_t lane_fork = nullptr;
  // A rebuild waits for no device work unless a table outgrows its device
  // buffer (hipFree of the old one waits for the device; capacities grow
  // geometrically): its table copies go through pinned staging on the library
  // stream (ordered after the launches already queued, which still read the
  // old tables), the staging is reused once those copies have run (event
  // `staged`), and the graphs of the old tables are destroyed once their last
  // launch has run (event per retired graph).
  // Two stagings used in turn, so a build waits at most for the copies of
  // the build before the last one.
  unsigned char* h_stage[2] = {nullptr, nullptr};
  size_t cap_stage[2] = {0, 0};
  hipEvent_t staged[2] = {nullptr, nullptr};
  bool staged_pending[2] = {false, false};
  int stage_cur = 0;
  size_t stage_at = 0;
  struct Retired {
    hipGraphExec_t exec;
    hipGraph_t graph;
    hipEvent_t done;
  };
  std::vector<Retired> retired;
};

namespace {

// Records an event pair around one batched phase on stream `s` when timing is on.
struct EnsTimer {
  lbswim_ensemble* e;
  int kernel;
  hipStream_t s;
  hipEvent_t start = nullptr;
  static hipEvent_t take(lbswim_ensemble* e) {
    if (!e->pool.empty()) {
      hipEvent_t ev = e->pool.back();
      e->pool.pop_back();
      return ev;
    }
    hipEvent_t ev = nullptr;
    return hipEventCreate(&ev) == hipSuccess ? ev : nullptr;
  }
  EnsTimer(lbswim_ensemble* en, int k, hipStream_t st) : e(en), kernel(k), s(st) {
    if ((e->timing >> k) & 1u) {
      start = take(e);
      if (start) (void)hipEventRecord(start, s);
    }
  }
  void done() {
    if (!start) return;
    hipEvent_t stop = take(e);
    if (!stop) return;
    (void)hipEventRecord(stop, s);
    e->recs.push_back(TimeRec{kernel, start, stop});
    start = nullptr;
  }
};

// Copies a host table to the device through the ensemble's pinned staging
// (ens_stage_begin reserved room for it), on the library stream.
template <typename T>
int upload(lbswim_ensemble* e, T** dev, size_t* cap, const std::vector<T>& host) {
  if (host.empty()) return LBSWIM_OK;
  if (*cap < host.size()) {
    dev_free(*dev);
    *dev = nullptr;
    const size_t c = grown(*cap, host.size());
    *cap = 0;
    if (int rc = dev_alloc_zero(reinterpret_cast<void**>(dev), sizeof(T) * c)) return rc;
    *cap = c;
  }
  const size_t bytes = sizeof(T) * host.size();
  const int k = e->stage_cur;
  if (e->stage_at + bytes > e->cap_stage[k]) return fail(LBSWIM_EINVAL, "ensemble table staging overflow");
  unsigned char* src = e->h_stage[k] + e->stage_at;
  std::memcpy(src, host.data(), bytes);
  e->stage_at = (e->stage_at + bytes + 255) / 256 * 256;
  HIP_TRY(hipMemcpyAsync(*dev, src, bytes, hipMemcpyHostToDevice, stream()));
  return LBSWIM_OK;
}

// Room for `bytes` of tables (plus alignment) in the pinned staging, once the
// copies of the last build have run.
int ens_stage_begin(lbswim_ensemble* e, size_t bytes) {
  const int k = e->stage_cur ^= 1;
  if (e->staged_pending[k]) {
    HIP_TRY(hipEventSynchronize(e->staged[k]));
    e->staged_pending[k] = false;
  }
  if (e->cap_stage[k] < bytes) {
    if (e->h_stage[k]) (void)hipHostFree(e->h_stage[k]);
    e->h_stage[k] = nullptr;
    e->cap_stage[k] = 0;
    const size_t c = grown(0, bytes);
    if (hipHostMalloc(reinterpret_cast<void**>(&e->h_stage[k]), c, hipHostMallocDefault) != hipSuccess)
      return fail(LBSWIM_ENOMEM, "pinned staging of %zu bytes failed", c);
    e->cap_stage[k] = c;
  }
  if (!e->staged[k]) HIP_TRY(hipEventCreateWithFlags(&e->staged[k], hipEventDisableTiming));
  e->stage_at = 0;
  return LBSWIM_OK;
}

int ens_stage_end(lbswim_ensemble* e) {
  const int k = e->stage_cur;
  HIP_TRY(hipEventRecord(e->staged[k], stream()));
  e->staged_pending[k] = true;
  return LBSWIM_OK;
}

// Destroys the retired graphs whose launches have all run (all of them,
// waiting, with `wait`).
void ens_reap(lbswim_ensemble* e, bool wait) {
  std::vector<lbswim_ensemble::Retired> keep;
  for (auto& r : e->retired) {
    if (!wait && hipEventQuery(r.done) == hipErrorNotReady) {
      keep.push_back(r);
      continue;
    }
    (void)hipEventSynchronize(r.done);
    if (r.exec) (void)hipGraphExecDestroy(r.exec);
    if (r.graph) (void)hipGraphDestroy(r.graph);
    (void)hipEventDestroy(r.done);
  }
  e->retired.swap(keep);
}

// Takes the captured graphs out of use; they are destroyed once every launch
// of them queued so far (on the library stream) has run -- `wait`: before
// returning.
void ens_drop_graph(lbswim_ensemble* e, bool wait = true) {
  for (int p = 0; p < 2; ++p) {
    if (!e->exec[p] && !e->graph[p]) continue;
    hipEvent_t done = nullptr;
    if (hipEventCreateWithFlags(&done, hipEventDisableTiming) == hipSuccess &&
        hipEventRecord(done, stream()) == hipSuccess) {
      e->retired.push_back(lbswim_ensemble::Retired{e->exec[p], e->graph[p], done});
    } else {  // no event: wait here instead
      if (done) (void)hipEventDestroy(done);
      (void)hipStreamSynchronize(stream());
      if (e->exec[p]) (void)hipGraphExecDestroy(e->exec[p]);
      if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
    }
    e->exec[p] = nullptr;
    e->graph[p] = nullptr;
  }
  ens_reap(e, wait);
}

int ens_build_lanes(lbswim_ensemble* e);
void ens_drop_lanes(lbswim_ensemble* e);

int ens_build(lbswim_ensemble* e) {
  ens_drop_graph(e, false);
  if (e->lanes > 1) return ens_build_lanes(e);
  ens_drop_lanes(e);
  std::vector<LatDesc> lat;
  std::vector<char> fused_member;  // lat[i] is advanced by the batched fused step
  // per member: its lattice blocks, in the stored-force (0), virtual-force
  // two-launch (1) or virtual-force fused (2) group, and (3) the stored-force
  // members without site segments again
  std::vector<std::vector<WorkItem>> latwork[4];
  std::vector<size_t> weight[4];
  std::vector<std::vector<WorkItem>> fusetiles;  // per fused member: its tile-strips
  std::vector<WorkItem> entrywork;
  std::vector<SwimDesc> sw;
  std::vector<WorkItem> swwork, ptwork, trwork;
  std::vector<TracerDesc> tr;
  uint32_t tile_base = 0, point_base = 0, site_base = 0;
  size_t site_points = 0;  // points of the FORCE_SITES_STORED members
  for (auto& m : e->members) {
    if (!m.lat) continue;
    lbswim_lattice* L = m.lat;
    const uint32_t mi = uint32_t(lat.size());
    LatDesc d{};
    d.g[0] = L->d_f[0];
    d.g[1] = L->d_f[1];
    d.force = L->d_force;
    d.rho = L->d_rho;
    d.u = L->d_u;
    d.geom = L->geom;
    d.tau_s = L->params.tau_s;
    d.tau_b = L->params.tau_b;
    d.omega_s = 1.0 / (d.tau_s + 0.5);
    d.omega_b = 1.0 / (d.tau_b + 0.5);
    d.cur = L->cur;
    size_t member_pts = 0;
    for (lbswim_swimmers* s : m.sw) member_pts += 2 * size_t(s->st.num);
    // the batch's density rule, member by member
    d.force_path = force_path(e->force_mode, member_pts, L->geom, true);
    d.tile_base = has_lists(d.force_path) ? tile_base : 0u;
    if (has_lists(d.force_path)) tile_base += spread_lists(L->geom);
    d.site_base = d.force_path == FORCE_SITES_STORED ? site_base : 0u;
    if (d.force_path == FORCE_SITES_STORED) {
      site_base += L->geom.n;
      site_points += member_pts;
    }
    d.point_first = point_base;
    d.point_count = uint32_t(member_pts);
    const bool fused = e->fuse_rows > 0 && d.force_path == FORCE_LISTS_VIRTUAL &&
                       batch_fused_member(L->geom, e->fuse_rows, L->elem);
    // every member's CalcHydro (or fused step) writes u only interleaved, the
    // copy the batched moves gather from
    if (!L->d_u_aos)
      if (int rc = dev_alloc_zero(&L->d_u_aos, 3 * L->elem * L->n)) return rc;
    d.u_aos = L->d_u_aos;
    m.fused = fused;
    lat.push_back(d);
    fused_member.push_back(fused);
    std::vector<WorkItem> lw;
    for (uint32_t b = 0; b < lattice_blocks(L->geom); b += uint32_t(e->per)) lw.push_back(WorkItem{mi, b});
    const int group = d.force_path != FORCE_LISTS_VIRTUAL ? 0 : (fused ? 2 : 1);
    latwork[group].push_back(lw);
    weight[group].push_back(L->n);
    if (group == 0 && d.force_path != FORCE_SITES_STORED) {
      latwork[3].push_back(lw);
      weight[3].push_back(L->n);
    }
    if (fused) {
      const uint32_t tiles = uint32_t(L->geom.ny / e->fuse_rows) * uint32_t(L->geom.nz / 64);
      const uint32_t strips = uint32_t((L->geom.nx + e->fuse_strip - 1) / e->fuse_strip);
      std::vector<WorkItem> fw;
      for (uint32_t u = 0; u < tiles * strips; ++u) fw.push_back(WorkItem{mi, u});
      fusetiles.push_back(fw);
      for (uint32_t u = 0; 64 * u < spread_lists(L->geom); ++u) entrywork.push_back(WorkItem{mi, u});
    }
    for (lbswim_swimmers* s : m.sw) {
      if (int rc = pstore_maybe_reorder(s->st, L->geom)) return rc;
      const uint32_t ai = uint32_t(sw.size());
      SwimDesc a{};
      a.r = fld(s->st, LBSWIM_PART_R);
      a.v = fld(s->st, LBSWIM_PART_V);
      a.n = fld(s->st, LBSWIM_PART_N);
      a.prng = s->st.d[LBSWIM_PART_PRNG];
      a.perm = s->st.perm;
      a.tumbles = s->d_tumbles;
      a.p = to_kernel(s->common_dev);
      a.num = s->st.num;
      a.member = int(mi);
      a.point_base = point_base;
      point_base += 2 * uint32_t(a.num);
      sw.push_back(a);
      for (uint32_t g = 0; g < uint32_t((a.num + PARTICLE_GROUPS - 1) / PARTICLE_GROUPS); ++g)
        swwork.push_back(WorkItem{ai, g});
      for (uint32_t b = 0; b < uint32_t((2 * size_t(a.num) + 255) / 256); ++b) ptwork.push_back(WorkItem{ai, b});
    }
    for (lbswim_tracers* t : m.tr) {
      if (int rc = pstore_maybe_reorder(t->st, L->geom)) return rc;
      const uint32_t ai = uint32_t(tr.size());
      TracerDesc a{};
      a.r = fld(t->st, LBSWIM_PART_R);
      a.s = fld(t->st, LBSWIM_PART_S);
      a.v = fld(t->st, LBSWIM_PART_V);
      a.num = t->st.num;
      a.member = int(mi);
      tr.push_back(a);
      for (uint32_t g = 0; g < uint32_t((a.num + TRACER_GROUPS - 1) / TRACER_GROUPS); ++g)
        trwork.push_back(WorkItem{ai, g});
    }
  }
  for (auto& m : e->members) {
    if (!m.lat) continue;
    m.cur_at_build = m.lat->cur;
  }
  const std::vector<WorkItem> storedwork = order_work(latwork[0], e->order, weight[0]);
  const std::vector<WorkItem> vfwork = order_work(latwork[1], e->order, weight[1]);
  const std::vector<WorkItem> fuseblk = order_work(latwork[2], e->order, weight[2]);
  const std::vector<WorkItem> fusework = order_work(fusetiles, e->order, weight[2]);
  const std::vector<WorkItem> spreadwork = order_work(latwork[3], e->order, weight[3]);
  int rc;
  size_t stage_bytes = 0;
  auto add = [&](size_t b) { stage_bytes += (b + 255) / 256 * 256; };
  add(sizeof(LatDesc) * lat.size());
  add(sizeof(SwimDesc) * sw.size());
  add(sizeof(TracerDesc) * tr.size());
  for (const std::vector<WorkItem>* v : std::initializer_list<const std::vector<WorkItem>*>{
           &storedwork, &vfwork, &swwork, &ptwork, &trwork, &fusework, &entrywork, &fuseblk, &spreadwork})
    add(sizeof(WorkItem) * v->size());
  if ((rc = ens_stage_begin(e, stage_bytes))) return rc;
  if ((rc = upload(e, &e->d_lat, &e->cap_lat, lat)) ||
      (rc = upload(e, &e->d_storedwork, &e->cap_storedwork, storedwork)) ||
      (rc = upload(e, &e->d_vfwork, &e->cap_vfwork, vfwork)) || (rc = upload(e, &e->d_sw, &e->cap_sw, sw)) ||
      (rc = upload(e, &e->d_swwork, &e->cap_swwork, swwork)) || (rc = upload(e, &e->d_ptwork, &e->cap_ptwork, ptwork)) ||
      (rc = upload(e, &e->d_tr, &e->cap_tr, tr)) || (rc = upload(e, &e->d_trwork, &e->cap_trwork, trwork)) ||
      (rc = upload(e, &e->d_fusework, &e->cap_fusework, fusework)) ||
      (rc = upload(e, &e->d_entrywork, &e->cap_entrywork, entrywork)) ||
      (rc = upload(e, &e->d_fuseblk, &e->cap_fuseblk, fuseblk)) ||
      (rc = upload(e, &e->d_spreadwork, &e->cap_spreadwork, spreadwork)) || (rc = ens_stage_end(e)))
    return rc;
  e->n_storedwork = int(storedwork.size());
  e->n_spreadwork = int(spreadwork.size());
  e->n_vfwork = int(vfwork.size());
  e->n_fusework = int(fusework.size());
  e->n_entrywork = int(entrywork.size());
  e->n_fuseblk = int(fuseblk.size());
  // aggregated atomics when any member with lists has dense ones
  e->dense_lists = false;
  for (size_t i = 0; i < lat.size(); ++i) {
    if (!has_lists(lat[i].force_path)) continue;
    e->dense_lists = e->dense_lists || dense_lists(lat[i].point_count, spread_lists(lat[i].geom));
  }
  e->sites_virtual = e->sites_stored = e->sites_fused = 0;
  for (const LatDesc& d : lat)
    (d.force_path != FORCE_LISTS_VIRTUAL ? e->sites_stored
                                         : (fused_member[size_t(&d - lat.data())] ? e->sites_fused : e->sites_virtual)) +=
        d.geom.n;
  e->stored_lists = false;
  for (const LatDesc& d : lat) e->stored_lists = e->stored_lists || d.force_path == FORCE_LISTS_STORED;
  e->fork_stored = !storedwork.empty() && !(vfwork.empty() && fusework.empty());
  if (e->fork_stored) {
    if (!e->side) HIP_TRY(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking));
    if (!e->ev_fork) HIP_TRY(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    if (!e->ev_join) HIP_TRY(hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming));
  }
  e->n_swwork = int(swwork.size());
  e->n_ptwork = int(ptwork.size());
  e->n_trwork = int(trwork.size());
  e->nlists = int(tile_base);
  e->npts = int(point_base);
  // spreading workspace
  size_t lcap = 0;  // list entries of the members with lists
  for (const auto& a : sw)
    if (has_lists(lat[a.member].force_path)) lcap += spread_list_capacity(2 * size_t(a.num));
  auto grow_to = [&](auto** p, size_t* cap, size_t need, size_t elem) -> int {
    if (*cap >= need && *p) return LBSWIM_OK;
    dev_free(*p);  // waits for the device: rare, capacities at least double
    *p = nullptr;
    const size_t c = std::max<size_t>(1, grown(*cap, need));
    *cap = 0;
    if (int r = dev_alloc_zero(reinterpret_cast<void**>(p), elem * c)) return r;
    *cap = c;
    return LBSWIM_OK;
  };
  const size_t nt = size_t(e->nlists);
  size_t cap_sorted = e->cap_list;
  if ((rc = grow_to(&e->counts, &e->cap_counts, nt, 4)) ||
      (rc = grow_to(&e->off, &e->cap_off, spread_off_entries(uint32_t(nt)), 4)) ||
      (rc = grow_to(&e->cursor, &e->cap_cursor, nt, 4)) || (rc = grow_to(&e->pts, &e->cap_pts, size_t(e->npts), 176)) ||
      (rc = grow_to(&e->slot_ids, &e->cap_slot, size_t(e->npts), 4)) ||
      (rc = grow_to(&e->sorted, &cap_sorted, lcap, 4)) || (rc = grow_to(&e->list, &e->cap_list, lcap, 4)))
    return rc;
  if (site_base > 0 && point_base >= SITE_FORCE_MAX_POINTS)
    return fail(LBSWIM_EINVAL, "an ensemble with dense members holds at most %zu swimmer ends (it has %u)",
                SITE_FORCE_MAX_POINTS - 1, point_base);
  e->nsites = int(site_base);
  e->ss_nent = int(64 * site_points);
  if (site_base > 0) {
    const size_t ns = site_base;
    if ((rc = grow_to(&e->ss_counts, &e->cap_ss_counts, ns, 4)) ||
        (rc = grow_to(&e->ss_off, &e->cap_ss_off, spread_off_entries(uint32_t(ns)), 4)) ||
        (rc = grow_to(&e->ss_local, &e->cap_ss_local, 64 * size_t(e->npts), 4)) ||
        (rc = grow_to(&e->ss_ent, &e->cap_ss_ent, 2 * 64 * site_points, 4)) ||
        (rc = grow_to(&e->ss_sorted, &e->cap_ss_sorted, 3 * 64 * site_points, 8)))
      return rc;
  }
  if (!fusework.empty()) {  // RunEntry records of every list entry, run-end forces of every list
    size_t cap_wz = e->cap_rec;
    if ((rc = grow_to(&e->rec, &e->cap_rec, lcap, sizeof(RunEntry))) || (rc = grow_to(&e->wz, &cap_wz, lcap, 4)) ||
        (rc = grow_to(&e->edge, &e->cap_edge, 6 * nt, 8)))
      return rc;
  }
  e->parity = 0;
  e->dirty = false;
  return LBSWIM_OK;
}

void ens_drop_lanes(lbswim_ensemble* e) {
  for (lbswim_ensemble* x : e->sub) lbswim_ensemble_destroy(x);
  e->sub.clear();
}

// Splits the members over e->lanes sub-ensembles, largest lattice first to
// the lane with the fewest sites, and builds each.
int ens_build_lanes(lbswim_ensemble* e) {
  std::vector<size_t> live;
  for (size_t i = 0; i < e->members.size(); ++i)
    if (e->members[i].lat) live.push_back(i);
  const size_t nl = std::max<size_t>(1, std::min<size_t>(size_t(e->lanes), live.size()));
  while (e->sub.size() > nl) {
    lbswim_ensemble_destroy(e->sub.back());
    e->sub.pop_back();
  }
  while (e->sub.size() < nl) {
    lbswim_ensemble* x = new (std::nothrow) lbswim_ensemble();
    if (!x) return fail(LBSWIM_ENOMEM, "ensemble lane allocation failed");
    e->sub.push_back(x);
  }
  if (!e->lane_fork) HIP_TRY(hipEventCreateWithFlags(&e->lane_fork, hipEventDisableTiming));
  while (e->lane_stream.size() < nl) {
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
    e->lane_stream.push_back(st);
    hipEvent_t ev = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    e->lane_join.push_back(ev);
  }
  std::stable_sort(live.begin(), live.end(),
                   [&](size_t a, size_t b) { return e->members[a].lat->n > e->members[b].lat->n; });
  std::vector<size_t> load(nl, 0);
  for (lbswim_ensemble* x : e->sub) {
    x->members.clear();
    x->storage = e->storage;
    x->order = e->order;
    x->per = e->per;
    x->fuse_rows = e->fuse_rows;
    x->fuse_strip = e->fuse_strip;
    x->force_mode = e->force_mode;
    x->timing = e->timing;
    x->lanes = 1;
    x->dirty = true;
  }
  std::vector<std::vector<size_t>> owner(nl);
  for (size_t i : live) {
    const size_t k = size_t(std::min_element(load.begin(), load.end()) - load.begin());
    load[k] += e->members[i].lat->n;
    e->sub[k]->members.push_back(e->members[i]);
    owner[k].push_back(i);
  }
  e->sites_virtual = e->sites_stored = e->sites_fused = 0;
  for (size_t k = 0; k < nl; ++k) {
    lbswim_ensemble* x = e->sub[k];
    if (int rc = ens_build(x)) return rc;
    e->sites_virtual += x->sites_virtual;
    e->sites_stored += x->sites_stored;
    e->sites_fused += x->sites_fused;
    for (size_t j = 0; j < owner[k].size(); ++j) {
      e->members[owner[k][j]].cur_at_build = x->members[j].cur_at_build;
      e->members[owner[k][j]].fused = x->members[j].fused;
    }
  }
  e->parity = 0;
  e->dirty = false;
  return LBSWIM_OK;
}

template <typename T>
int ens_enqueue_step(lbswim_ensemble* e, int parity) {
  hipStream_t s = stream();
  EnsTimer t_lists(e, LBSWIM_K_ADDF, s);
  HIP_TRY(launch_batch_points(e->d_lat, e->d_sw, e->d_ptwork, e->n_ptwork, e->pts, e->slot_ids, s));
  // Members with a stored force write it (from their lists or from all their
  // points), then collide and CalcHydro; the others evaluate it from the
  // lists in the collide and CalcHydro.  With both kinds present the stored
  // members' chain runs on the side stream beside the virtual-force members'
  // lattice launches, and, when no stored member needs lists, beside the
  // list building too (a fork and a join by events, which a graph capture
  // records as dependencies).
  const bool fork = e->fork_stored;
  const bool early = fork && !e->stored_lists;
  hipStream_t sb = fork ? e->side : s;
  auto fork_now = [&]() -> int {
    HIP_TRY(hipEventRecord(e->ev_fork, s));
    HIP_TRY(hipStreamWaitEvent(sb, e->ev_fork, 0));
    return LBSWIM_OK;
  };
  if (early)
    if (int rc = fork_now()) return rc;
  auto stored_chain = [&]() -> int {
    EnsTimer t(e, LBSWIM_K_STORED, sb);
    HIP_TRY(launch_batch_site_lists(e->d_lat, e->npts, e->nsites, e->ss_nent, e->pts, e->slot_ids, e->ss_counts,
                                    e->ss_off, e->ss_local, e->ss_ent, e->ss_sorted, sb));
    // the site members' force is summed and stored by the collide; a pass
    // over the other stored members' blocks writes theirs
    HIP_TRY(launch_batch_spread_force<T>(e->d_lat, e->n_spreadwork, e->d_spreadwork, e->per, e->pts, e->off,
                                         e->sorted, nullptr, nullptr, sb));
    HIP_TRY(launch_batch_collide<T>(e->d_lat, e->d_storedwork, e->n_storedwork, e->per, parity, nullptr, sb,
                                    e->ss_off, e->ss_sorted));
    HIP_TRY(launch_batch_hydro<T>(e->d_lat, e->d_storedwork, e->n_storedwork, e->per, parity, nullptr, sb));
    t.done();
    return LBSWIM_OK;
  };
  if (early)
    if (int rc = stored_chain()) return rc;
  HIP_TRY(launch_batch_lists(e->d_lat, e->npts, e->nlists, e->pts, e->slot_ids, e->dense_lists, e->counts, e->off,
                             e->cursor, e->list, e->sorted, s));
  HIP_TRY(launch_batch_run_entries(e->d_lat, e->d_entrywork, e->n_entrywork, e->pts, e->off, e->sorted, e->rec, e->wz,
                                   e->edge, s));
  t_lists.done();
  if (!early) {
    if (fork)
      if (int rc = fork_now()) return rc;
    if (int rc = stored_chain()) return rc;
  }
  const VForce vf{e->pts, e->off, e->sorted};
  if (e->n_fusework) {
    EnsTimer t(e, LBSWIM_K_FUSED, s);
    const VForce fvf{e->pts, e->off, e->sorted, e->edge, e->rec, e->wz};
    HIP_TRY(launch_batch_step_hydro<T>(e->d_lat, e->d_fusework, e->n_fusework, parity, fvf, e->fuse_rows,
                                       e->fuse_strip, s));
    t.done();
  }
  {
    EnsTimer t(e, LBSWIM_K_COLLIDE, s);
    HIP_TRY(launch_batch_collide<T>(e->d_lat, e->d_vfwork, e->n_vfwork, e->per, parity, &vf, s));
    t.done();
  }
  {
    EnsTimer t(e, LBSWIM_K_HYDRO, s);
    HIP_TRY(launch_batch_hydro<T>(e->d_lat, e->d_vfwork, e->n_vfwork, e->per, parity, &vf, s));
    t.done();
  }
  if (fork) {
    HIP_TRY(hipEventRecord(e->ev_join, sb));
    HIP_TRY(hipStreamWaitEvent(s, e->ev_join, 0));
  }
  {
    EnsTimer t(e, LBSWIM_K_SMOVE, s);
    HIP_TRY(launch_batch_swimmers_move<T>(e->d_lat, e->d_sw, e->d_swwork, e->n_swwork, s));
    t.done();
  }
  EnsTimer t(e, LBSWIM_K_TMOVE, s);
  HIP_TRY(launch_batch_tracers_move<T>(e->d_lat, e->d_tr, e->d_trwork, e->n_trwork, s));
  t.done();
  return LBSWIM_OK;
}

int ens_enqueue(lbswim_ensemble* e, int parity) {
  if (e->lanes > 1 && !e->sub.empty()) {
    // fork every lane from the caller's stream, join them all back
    hipStream_t s = stream();
    HIP_TRY(hipEventRecord(e->lane_fork, s));
    for (size_t k = 0; k < e->sub.size(); ++k) {
      HIP_TRY(hipStreamWaitEvent(e->lane_stream[k], e->lane_fork, 0));
      {
        StreamScope on_lane(e->lane_stream[k]);
        if (int rc = ens_enqueue(e->sub[k], parity)) return rc;
      }
      HIP_TRY(hipEventRecord(e->lane_join[k], e->lane_stream[k]));
    }
    for (size_t k = 0; k < e->sub.size(); ++k) HIP_TRY(hipStreamWaitEvent(s, e->lane_join[k], 0));
    return LBSWIM_OK;
  }
  return e->storage == LBSWIM_STORE_F64 ? ens_enqueue_step<double>(e, parity) : ens_enqueue_step<float>(e, parity);
}

// After batched steps with the virtual force: write every member's force
// array from the last step's lists, as the reference leaves it.
int ens_materialize_force(lbswim_ensemble* e) {
  if (e->lanes > 1 && !e->sub.empty()) {
    for (lbswim_ensemble* x : e->sub)
      if (int rc = ens_materialize_force(x)) return rc;
    return LBSWIM_OK;
  }
  for (int k = 0; k < 2; ++k) {
    const int nw = k ? e->n_fuseblk : e->n_vfwork;
    const WorkItem* w = k ? e->d_fuseblk : e->d_vfwork;
    hipError_t r = e->storage == LBSWIM_STORE_F64
                       ? launch_batch_spread_force<double>(e->d_lat, nw, w, e->per, e->pts, e->off, e->sorted,
                                                           e->ss_off, e->ss_sorted, stream())
                       : launch_batch_spread_force<float>(e->d_lat, nw, w, e->per, e->pts, e->off, e->sorted,
                                                          e->ss_off, e->ss_sorted, stream());
    HIP_TRY(r);
  }
  return LBSWIM_OK;
}

// Captures two steps starting with parity p into graph slot p.
int ens_capture(lbswim_ensemble* e, int p) {
  if (!e->cap) HIP_TRY(hipStreamCreateWithFlags(&e->cap, hipStreamNonBlocking));
  StreamScope on_cap(e->cap);
  HIP_TRY(hipStreamBeginCapture(e->cap, hipStreamCaptureModeThreadLocal));
  int rc = ens_enqueue(e, p);
  if (!rc) rc = ens_enqueue(e, p ^ 1);
  hipGraph_t g = nullptr;
  hipError_t he = hipStreamEndCapture(e->cap, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (he != hipSuccess) return fail(LBSWIM_EHIP, "graph capture failed: %s", hipGetErrorString(he));
  hipGraphExec_t x = nullptr;
  he = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (he != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(LBSWIM_EHIP, "graph instantiation failed: %s", hipGetErrorString(he));
  }
  e->graph[p] = g;
  e->exec[p] = x;
  return LBSWIM_OK;
}

// Host bookkeeping after `k` steps: the reference's pointer swap and counter.
void ens_advance(lbswim_ensemble* e, int k) {
  if (e->lanes > 1 && !e->sub.empty()) {  // the lanes own the members' bookkeeping
    for (lbswim_ensemble* x : e->sub) ens_advance(x, k);
    if (k & 1) e->parity ^= 1;
    return;
  }
  for (auto& m : e->members) {
    if (!m.lat) continue;
    if (k & 1) m.lat->cur ^= 1;
    m.lat->time_step += k;
    m.lat->force_virtual = false;  // the batched step spreads into the force array
    // and writes u only interleaved
    m.lat->u_aos_valid = true;
    m.lat->u_virtual = true;
    for (auto* s : m.sw) {
      s->st.moves_since_order += k;
      s->version = next_version();
    }
    for (auto* t : m.tr) t->st.moves_since_order += k;
  }
  if (k & 1) e->parity ^= 1;
}

int ens_check_member(lbswim_ensemble* e, int member) {
  CHECK_ARG(member >= 0 && member < int(e->members.size()) && e->members[member].lat, "no ensemble member %d",
            member);
  return LBSWIM_OK;
}

}  // namespace

extern "C" {

int lbswim_ensemble_create(int storage, lbswim_ensemble** out) {
  CHECK_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  CHECK_ARG(storage == LBSWIM_STORE_F64 || storage == LBSWIM_STORE_F32DEV, "unknown storage form %d", storage);
  if (int rc = ensure_runtime()) return rc;
  lbswim_ensemble* e = new (std::nothrow) lbswim_ensemble();
  if (!e) return fail(LBSWIM_ENOMEM, "ensemble allocation failed");
  e->storage = storage;
  *out = e;
  return LBSWIM_OK;
}

void lbswim_ensemble_destroy(lbswim_ensemble* e) {
  if (!e) return;
  if (g_rt.ready) (void)hipStreamSynchronize(stream());
  ens_drop_graph(e, true);
  for (int k = 0; k < 2; ++k) {
    if (e->staged[k]) (void)hipEventDestroy(e->staged[k]);
    if (e->h_stage[k]) (void)hipHostFree(e->h_stage[k]);
  }
  for (auto& r : e->recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (hipEvent_t ev : e->pool) (void)hipEventDestroy(ev);
  ens_drop_lanes(e);
  for (hipStream_t st : e->lane_stream) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->lane_join) (void)hipEventDestroy(ev);
  if (e->lane_fork) (void)hipEventDestroy(e->lane_fork);
  if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
  if (e->ev_join) (void)hipEventDestroy(e->ev_join);
  if (e->side) (void)hipStreamDestroy(e->side);
  if (e->cap) (void)hipStreamDestroy(e->cap);
  dev_free(e->d_lat);
  dev_free(e->d_storedwork);
  dev_free(e->d_spreadwork);
  dev_free(e->d_vfwork);
  dev_free(e->d_sw);
  dev_free(e->d_swwork);
  dev_free(e->d_ptwork);
  dev_free(e->d_tr);
  dev_free(e->d_trwork);
  dev_free(e->counts);
  dev_free(e->off);
  dev_free(e->cursor);
  dev_free(e->list);
  dev_free(e->sorted);
  dev_free(e->ss_counts);
  dev_free(e->ss_off);
  dev_free(e->ss_local);
  dev_free(e->ss_ent);
  dev_free(e->ss_sorted);
  dev_free(e->pts);
  dev_free(e->slot_ids);
  dev_free(e->d_fusework);
  dev_free(e->d_entrywork);
  dev_free(e->d_fuseblk);
  dev_free(e->rec);
  dev_free(e->wz);
  dev_free(e->edge);
  delete e;
}

int lbswim_ensemble_add(lbswim_ensemble* e, lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw,
                        lbswim_tracers* const* tr, int ntr, int* member) {
  CHECK_ARG(e && lat, "NULL argument");
  CHECK_ARG(lat->storage == e->storage, "lattice storage form differs from the ensemble's");
  CHECK_ARG(nsw >= 0 && ntr >= 0 && (nsw == 0 || sw) && (ntr == 0 || tr), "invalid particle lists");
  for (auto& m : e->members) CHECK_ARG(m.lat != lat, "lattice is already a member");
  EnsMember m;
  m.lat = lat;
  for (int i = 0; i < nsw; ++i) {
    CHECK_ARG(sw[i] != nullptr, "swimmer handle %d is NULL", i);
    m.sw.push_back(sw[i]);
  }
  for (int i = 0; i < ntr; ++i) {
    CHECK_ARG(tr[i] != nullptr, "tracer handle %d is NULL", i);
    m.tr.push_back(tr[i]);
  }
  size_t slot = 0;
  while (slot < e->members.size() && e->members[slot].lat) ++slot;
  if (slot == e->members.size())
    e->members.push_back(m);
  else
    e->members[slot] = m;
  e->dirty = true;
  if (member) *member = int(slot);
  return LBSWIM_OK;
}

int lbswim_ensemble_remove(lbswim_ensemble* e, int member) {
  CHECK_ARG(e, "NULL argument");
  if (int rc = ens_check_member(e, member)) return rc;
  // No wait: steps already enqueued still advance the member, and anything
  // the caller does with it afterwards (copies, destroy, which waits for the
  // stream) is ordered after them on the library stream.
  e->members[member] = EnsMember();
  e->dirty = true;
  return LBSWIM_OK;
}

int lbswim_ensemble_size(lbswim_ensemble* e, int* members) {
  CHECK_ARG(e && members, "NULL argument");
  int k = 0;
  for (auto& m : e->members) k += m.lat ? 1 : 0;
  *members = k;
  return LBSWIM_OK;
}

int lbswim_ensemble_set_graph(lbswim_ensemble* e, int on) {
  CHECK_ARG(e, "NULL argument");
  e->use_graph = on != 0;
  if (!e->use_graph) ens_drop_graph(e);
  return LBSWIM_OK;
}

int lbswim_ensemble_set_order(lbswim_ensemble* e, int order) {
  CHECK_ARG(e && order >= 0 && order < ORDER_COUNT, "invalid argument");
  if (e->order != order) e->dirty = true;
  e->order = order;
  return LBSWIM_OK;
}

int lbswim_ensemble_set_force_mode(lbswim_ensemble* e, int mode) {
  CHECK_ARG(e, "ensemble is NULL");
  CHECK_ARG(mode == LBSWIM_FORCE_STORED || mode == LBSWIM_FORCE_AUTO || mode == LBSWIM_FORCE_VIRTUAL,
            "unknown force mode %d", mode);
  if (e->force_mode != mode) e->dirty = true;
  e->force_mode = mode;
  return LBSWIM_OK;
}

int lbswim_ensemble_step(lbswim_ensemble* e, int nsteps) {
  CHECK_ARG(e && nsteps >= 0, "invalid argument");
  if (nsteps == 0) return LBSWIM_OK;
  for (auto& m : e->members) {
    if (!m.lat || e->dirty) continue;
    if (m.lat->cur != (m.cur_at_build ^ e->parity)) e->dirty = true;
  }
  if (e->dirty)
    if (int rc = ens_build(e)) return rc;
  int done = 0;
  if (e->use_graph && !e->timing && nsteps >= 2) {
    const int p = e->parity;
    if (!e->exec[p])
      if (int rc = ens_capture(e, p)) return rc;
    for (; done + 2 <= nsteps; done += 2) HIP_TRY(hipGraphLaunch(e->exec[p], stream()));
    ens_advance(e, done);
  }
  for (; done < nsteps; ++done) {
    if (int rc = ens_enqueue(e, e->parity)) return rc;
    ens_advance(e, 1);
  }
  return ens_materialize_force(e);
}

int lbswim_ensemble_set_timing(lbswim_ensemble* e, unsigned mask) {
  CHECK_ARG(e, "ensemble is NULL");
  if (mask) {
    HIP_TRY(hipStreamSynchronize(stream()));
    for (auto& r : e->recs) {
      e->pool.push_back(r.start);
      e->pool.push_back(r.stop);
    }
    e->recs.clear();
  }
  e->timing = mask & ((1u << LBSWIM_K_COUNT) - 1u);
  for (lbswim_ensemble* x : e->sub)
    if (int rc = lbswim_ensemble_set_timing(x, mask)) return rc;
  return LBSWIM_OK;
}

int lbswim_ensemble_kernel_time(lbswim_ensemble* e, int kernel, double* total_ms, long* count) {
  CHECK_ARG(e && total_ms && count, "NULL argument");
  CHECK_ARG(kernel >= 0 && kernel < LBSWIM_K_COUNT, "unknown kernel id %d", kernel);
  HIP_TRY(hipStreamSynchronize(stream()));
  double tot = 0.0;
  long cnt = 0;
  for (const auto& r : e->recs) {
    if (r.kernel != kernel) continue;
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.start, r.stop));
    tot += ms;
    ++cnt;
  }
  // lanes: their launches of one step run side by side; the sum over lanes is
  // their total device time
  for (lbswim_ensemble* x : e->sub) {
    double t = 0.0;
    long c = 0;
    if (int rc = lbswim_ensemble_kernel_time(x, kernel, &t, &c)) return rc;
    tot += t;
    cnt += c;
  }
  *total_ms = tot;
  *count = cnt;
  return LBSWIM_OK;
}

int lbswim_ensemble_set_lanes(lbswim_ensemble* e, int lanes) {
  CHECK_ARG(e, "ensemble is NULL");
  CHECK_ARG(lanes >= 1 && lanes <= 8, "lanes must be 1..8 (got %d)", lanes);
  if (e->lanes != lanes) e->dirty = true;
  e->lanes = lanes;
  return LBSWIM_OK;
}

int lbswim_ensemble_lanes(lbswim_ensemble* e, int* lanes) {
  CHECK_ARG(e && lanes, "NULL argument");
  *lanes = e->lanes;
  return LBSWIM_OK;
}

int lbswim_ensemble_sites(lbswim_ensemble* e, size_t* virtual_sites, size_t* stored_sites) {
  CHECK_ARG(e && virtual_sites && stored_sites, "NULL argument");
  if (e->dirty)
    if (int rc = ens_build(e)) return rc;
  *virtual_sites = e->sites_virtual;
  *stored_sites = e->sites_stored;
  return LBSWIM_OK;
}

int lbswim_ensemble_autotune(lbswim_ensemble* e, int steps_per_trial, int* choice) {
  CHECK_ARG(e && steps_per_trial > 0, "invalid argument");
  hipEvent_t a = nullptr, b = nullptr;
  HIP_TRY(hipEventCreate(&a));
  HIP_TRY(hipEventCreate(&b));
  int best_order = e->order, best_per = e->per;
  float best_ms = 0.f;
  bool first = true;
  int rc = LBSWIM_OK;
  // workgroup order x work-item size (1, 2, 4 lattice blocks of 256 sites)
  for (int per : {1, 2, 4}) {
    for (int o = 0; o < ORDER_COUNT && !rc; ++o) {
      e->order = o;
      e->per = per;
      e->dirty = true;
      if ((rc = lbswim_ensemble_step(e, 2))) break;  // build + capture + warm-up
      (void)hipEventRecord(a, stream());
      if ((rc = lbswim_ensemble_step(e, steps_per_trial))) break;
      (void)hipEventRecord(b, stream());
      (void)hipEventSynchronize(b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (first || ms < best_ms) {
        best_ms = ms;
        best_order = o;
        best_per = per;
        first = false;
      }
    }
    if (rc) break;
  }
  // then the fused step's tile shape with that order and work-item size:
  // off, or tiles of 8 rows with strips of 16, 32, 64 planes, or of 4 rows
  // with strips of 32 (only where some member qualifies)
  if (!rc) {
    e->order = best_order;
    e->per = best_per;
    e->dirty = true;  // the tables on the device are the last trial's
    bool any = false;
    for (auto& m : e->members) {
      if (!m.lat) continue;
      size_t npts = 0;
      for (lbswim_swimmers* sw : m.sw) npts += 2 * size_t(sw->st.num);
      any = any || (force_path(e->force_mode, npts, m.lat->geom, true) == FORCE_LISTS_VIRTUAL &&
                    batch_fused_member(m.lat->geom, 4, m.lat->elem));
    }
    if (any) {
      const int cand[5][2] = {{0, 32}, {8, 16}, {8, 32}, {8, 64}, {4, 32}};
      float fbest = 0.f;
      int brows = e->fuse_rows, bstrip = e->fuse_strip;
      for (int i = 0; i < 5 && !rc; ++i) {
        e->fuse_rows = cand[i][0];
        e->fuse_strip = cand[i][1];
        e->dirty = true;
        if ((rc = lbswim_ensemble_step(e, 2))) break;
        (void)hipEventRecord(a, stream());
        if ((rc = lbswim_ensemble_step(e, steps_per_trial))) break;
        (void)hipEventRecord(b, stream());
        (void)hipEventSynchronize(b);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (i == 0 || ms < fbest) {
          fbest = ms;
          brows = cand[i][0];
          bstrip = cand[i][1];
        }
      }
      e->fuse_rows = brows;
      e->fuse_strip = bstrip;
      e->dirty = true;
    }
  }
  // then the lanes, with everything else as chosen
  if (!rc) {
    int members = 0;
    for (auto& m : e->members) members += m.lat ? 1 : 0;
    float lbest = 0.f;
    int blanes = e->lanes;
    bool lfirst = true;
    for (int ln : {1, 2, 4}) {
      if (ln > 1 && ln > members) break;
      e->lanes = ln;
      e->dirty = true;
      if ((rc = lbswim_ensemble_step(e, 2))) break;
      (void)hipEventRecord(a, stream());
      if ((rc = lbswim_ensemble_step(e, steps_per_trial))) break;
      (void)hipEventRecord(b, stream());
      (void)hipEventSynchronize(b);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, a, b);
      if (lfirst || ms < lbest) {
        lbest = ms;
        blanes = ln;
        lfirst = false;
      }
    }
    e->lanes = blanes;
    e->dirty = true;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  e->order = best_order;
  e->per = best_per;
  e->dirty = true;  // built for the last trial
  if (choice) *choice = best_order;
  return LBSWIM_OK;
}

int lbswim_ensemble_set_work(lbswim_ensemble* e, int blocks_per_item) {
  CHECK_ARG(e && blocks_per_item >= 1 && blocks_per_item <= 64, "blocks per work item must be 1..64");
  if (e->per != blocks_per_item) e->dirty = true;
  e->per = blocks_per_item;
  return LBSWIM_OK;
}

int lbswim_ensemble_tuning(lbswim_ensemble* e, int* order, int* blocks_per_item) {
  CHECK_ARG(e, "ensemble is NULL");
  if (order) *order = e->order;
  if (blocks_per_item) *blocks_per_item = e->per;
  return LBSWIM_OK;
}

int lbswim_ensemble_set_fusion(lbswim_ensemble* e, int rows, int strip) {
  CHECK_ARG(e, "ensemble is NULL");
  CHECK_ARG(rows == 0 || rows == 4 || rows == 8, "rows must be 0, 4 or 8 (got %d)", rows);
  CHECK_ARG(strip >= 1, "strip must be >= 1 plane (got %d)", strip);
  if (e->fuse_rows != rows || e->fuse_strip != strip) e->dirty = true;
  e->fuse_rows = rows;
  e->fuse_strip = strip;
  return LBSWIM_OK;
}

int lbswim_ensemble_fusion(lbswim_ensemble* e, int* rows, int* strip) {
  CHECK_ARG(e, "ensemble is NULL");
  if (rows) *rows = e->fuse_rows;
  if (strip) *strip = e->fuse_strip;
  return LBSWIM_OK;
}

int lbswim_ensemble_fused_sites(lbswim_ensemble* e, size_t* fused_sites) {
  CHECK_ARG(e && fused_sites, "NULL argument");
  if (e->dirty)
    if (int rc = ens_build(e)) return rc;
  *fused_sites = e->sites_fused;
  return LBSWIM_OK;
}

}  // extern "C"
