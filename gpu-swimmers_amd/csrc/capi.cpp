// C ABI of liblbswim_amd.so: object lifetimes, host/device mirrors, field
// conversion for the FP32 storage form, kernel launches on the library stream
// and per-kernel event timing.  Declarations and the reference interface each
// entry point replaces are in include/lbswim.h.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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
// default stream.  Created on first use; default priority.
struct Runtime {
  bool ready = false;
  hipStream_t stream = nullptr;
};
Runtime g_rt;

int ensure_runtime() {
  if (g_rt.ready) return LBSWIM_OK;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev < 1) return fail(LBSWIM_ENODEV, "no HIP device available (%s)", hipGetErrorString(e));
  HIP_TRY(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  g_rt.ready = true;
  return LBSWIM_OK;
}

hipStream_t stream() { return g_rt.stream; }

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
  bool timing = false;
  std::vector<TimeRec> recs;
  std::vector<hipEvent_t> pool;
};

struct lbswim_swimmers {
  int num = 0;
  lbswim_common_params common_host{};
  lbswim_common_params common_dev{};  // the copy kernels see (SharedItem::device)
  double* d_r = nullptr;
  double* d_v = nullptr;
  double* d_n = nullptr;
  uint64_t* d_prng = nullptr;
  unsigned long long* d_tumbles = nullptr;
  void* h[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};

struct lbswim_tracers {
  int num = 0;
  double* d_r = nullptr;
  double* d_s = nullptr;
  double* d_v = nullptr;
  void* h[4] = {nullptr, nullptr, nullptr, nullptr};
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
    if (lat && lat->timing) {
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

extern "C" {

const char* lbswim_last_error(void) { return g_err.c_str(); }

int lbswim_synchronize(void) {
  if (!g_rt.ready) return LBSWIM_OK;
  HIP_TRY(hipStreamSynchronize(stream()));
  return LBSWIM_OK;
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
  dev_free(lat->d_force);
  dev_free(lat->d_f[0]);
  dev_free(lat->d_f[1]);
  dev_free(lat->d_partial);
  for (double* h : lat->h) std::free(h);
  for (auto& r : lat->recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (hipEvent_t ev : lat->pool) (void)hipEventDestroy(ev);
  delete lat;
}

static int lattice_step_impl(lbswim_lattice* lat) {
  const double ts = lat->params.tau_s, tb = lat->params.tau_b;
  KernelTimer t(lat, LBSWIM_K_COLLIDE);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_collide_stream<double>(static_cast<const double*>(lat->d_f[lat->cur]),
                                      static_cast<const double*>(lat->d_force),
                                      static_cast<double*>(lat->d_f[1 - lat->cur]), lat->geom, ts, tb, stream());
  else
    e = launch_collide_stream<float>(static_cast<const float*>(lat->d_f[lat->cur]),
                                     static_cast<const float*>(lat->d_force),
                                     static_cast<float*>(lat->d_f[1 - lat->cur]), lat->geom, ts, tb, stream());
  t.done();
  HIP_TRY(e);
  lat->cur = 1 - lat->cur;  // lb.cu:242-244
  lat->time_step++;         // lb.cu:246
  return LBSWIM_OK;
}

static int lattice_hydro_impl(lbswim_lattice* lat, bool zero_force) {
  KernelTimer t(lat, LBSWIM_K_HYDRO);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_calc_hydro<double>(static_cast<const double*>(lat->d_f[lat->cur]), static_cast<double*>(lat->d_force),
                                  static_cast<double*>(lat->d_rho), static_cast<double*>(lat->d_u), lat->geom,
                                  zero_force, stream());
  else
    e = launch_calc_hydro<float>(static_cast<const float*>(lat->d_f[lat->cur]), static_cast<float*>(lat->d_force),
                                 static_cast<float*>(lat->d_rho), static_cast<float*>(lat->d_u), lat->geom,
                                 zero_force, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

static int lattice_zero_force_impl(lbswim_lattice* lat) {
  KernelTimer t(lat, LBSWIM_K_ZERO);
  hipError_t e = launch_fill_zero(lat->d_force, 3 * lat->elem * lat->n, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

int lbswim_lattice_step(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  return lattice_step_impl(lat);
}

int lbswim_lattice_calc_hydro(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
  return lattice_hydro_impl(lat, false);
}

int lbswim_lattice_init_from_hydro(lbswim_lattice* lat) {
  if (int rc = check_lattice(lat)) return rc;
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
  if (device) *device = field_device(lat, field);
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
  if (lat->storage == LBSWIM_STORE_F64) return copy_h2d(dev, h, cnt * sizeof(double));
  std::vector<float> stage(cnt);
  for (size_t i = 0; i < cnt; ++i) stage[i] = float(h[i] - f32_offset(lat, field, i));
  return copy_h2d(dev, stage.data(), cnt * sizeof(float));
}

int lbswim_lattice_d2h(lbswim_lattice* lat, int field) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(field >= LBSWIM_RHO && field <= LBSWIM_FNEW, "unknown lattice field %d", field);
  double* h = nullptr;
  if (int rc = lattice_host(lat, field, &h)) return rc;
  const size_t cnt = field_count(lat, field);
  const void* dev = field_device(lat, field);
  if (lat->storage == LBSWIM_STORE_F64) return copy_d2h(h, dev, cnt * sizeof(double));
  std::vector<float> stage(cnt);
  if (int rc = copy_d2h(stage.data(), dev, cnt * sizeof(float))) return rc;
  for (size_t i = 0; i < cnt; ++i) h[i] = double(stage[i]) + f32_offset(lat, field, i);
  return LBSWIM_OK;
}

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
  sw->num = n;
  sw->common_host = *p;
  sw->common_dev = *p;
  const size_t v3 = 3 * sizeof(double) * size_t(n);
  if ((rc = dev_alloc_zero(reinterpret_cast<void**>(&sw->d_r), v3)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&sw->d_v), v3)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&sw->d_n), v3)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&sw->d_prng), sizeof(uint64_t) * size_t(n))) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&sw->d_tumbles), sizeof(unsigned long long)))) {
    std::string msg = g_err;
    lbswim_swimmers_destroy(sw);
    g_err = msg;
    return rc;
  }
  *out = sw;
  return LBSWIM_OK;
}

void lbswim_swimmers_destroy(lbswim_swimmers* sw) {
  if (!sw) return;
  if (g_rt.ready) (void)hipStreamSynchronize(stream());
  dev_free(sw->d_r);
  dev_free(sw->d_v);
  dev_free(sw->d_n);
  dev_free(sw->d_prng);
  dev_free(sw->d_tumbles);
  for (void* h : sw->h) std::free(h);
  delete sw;
}

int lbswim_swimmers_num(const lbswim_swimmers* sw, int* out) {
  CHECK_ARG(sw && out, "NULL argument");
  *out = sw->num;
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
  KernelTimer t(lat, LBSWIM_K_ADDF);
  const SwimParams p = to_kernel(sw->common_dev);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_swimmers_add_forces<double>(sw->num, p, sw->d_r, sw->d_n, static_cast<double*>(lat->d_force),
                                           lat->geom, stream());
  else
    e = launch_swimmers_add_forces<float>(sw->num, p, sw->d_r, sw->d_n, static_cast<float*>(lat->d_force),
                                          lat->geom, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

int lbswim_swimmers_move(lbswim_swimmers* sw, lbswim_lattice* lat) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = check_lattice(lat)) return rc;
  KernelTimer t(lat, LBSWIM_K_SMOVE);
  const SwimParams p = to_kernel(sw->common_dev);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_swimmers_move<double>(sw->num, p, sw->d_r, sw->d_v, sw->d_n, sw->d_prng,
                                     static_cast<const double*>(lat->d_u), lat->geom, sw->d_tumbles, stream());
  else
    e = launch_swimmers_move<float>(sw->num, p, sw->d_r, sw->d_v, sw->d_n, sw->d_prng,
                                    static_cast<const float*>(lat->d_u), lat->geom, sw->d_tumbles, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

static void* swimmer_dev(lbswim_swimmers* sw, int field) {
  switch (field) {
    case LBSWIM_PART_R: return sw->d_r;
    case LBSWIM_PART_V: return sw->d_v;
    case LBSWIM_PART_N: return sw->d_n;
    case LBSWIM_PART_PRNG: return sw->d_prng;
    default: return nullptr;
  }
}

static int swimmer_check_field(int field) {
  CHECK_ARG(field == LBSWIM_PART_R || field == LBSWIM_PART_V || field == LBSWIM_PART_N ||
                field == LBSWIM_PART_PRNG,
            "unknown swimmer field %d", field);
  return LBSWIM_OK;
}

static size_t swimmer_count(const lbswim_swimmers* sw, int field) {
  return field == LBSWIM_PART_PRNG ? size_t(sw->num) : 3 * size_t(sw->num);
}

static int swimmer_host(lbswim_swimmers* sw, int field, void** out) {
  if (!sw->h[field]) {
    const size_t cnt = swimmer_count(sw, field);
    sw->h[field] = std::calloc(cnt ? cnt : 1, 8);
    if (!sw->h[field]) return fail(LBSWIM_ENOMEM, "host mirror allocation failed");
  }
  *out = sw->h[field];
  return LBSWIM_OK;
}

int lbswim_swimmers_field(lbswim_swimmers* sw, int field, void** host, void** device, size_t* count,
                          size_t* elem_size) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = swimmer_check_field(field)) return rc;
  if (host) {
    if (int rc = swimmer_host(sw, field, host)) return rc;
  }
  if (device) *device = swimmer_dev(sw, field);
  if (count) *count = swimmer_count(sw, field);
  if (elem_size) *elem_size = 8;
  return LBSWIM_OK;
}

int lbswim_swimmers_h2d(lbswim_swimmers* sw, int field) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = swimmer_check_field(field)) return rc;
  void* h = nullptr;
  if (int rc = swimmer_host(sw, field, &h)) return rc;
  return copy_h2d(swimmer_dev(sw, field), h, 8 * swimmer_count(sw, field));
}

int lbswim_swimmers_d2h(lbswim_swimmers* sw, int field) {
  CHECK_ARG(sw, "swimmer handle is NULL");
  if (int rc = swimmer_check_field(field)) return rc;
  void* h = nullptr;
  if (int rc = swimmer_host(sw, field, &h)) return rc;
  return copy_d2h(h, swimmer_dev(sw, field), 8 * swimmer_count(sw, field));
}

int lbswim_swimmers_tumbles(lbswim_swimmers* sw, unsigned long long* out) {
  CHECK_ARG(sw && out, "NULL argument");
  return copy_d2h(out, sw->d_tumbles, sizeof(unsigned long long));
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
  tr->num = n;
  const size_t v3 = 3 * sizeof(double) * size_t(n);
  if ((rc = dev_alloc_zero(reinterpret_cast<void**>(&tr->d_r), v3)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&tr->d_s), v3)) ||
      (rc = dev_alloc_zero(reinterpret_cast<void**>(&tr->d_v), v3))) {
    std::string msg = g_err;
    lbswim_tracers_destroy(tr);
    g_err = msg;
    return rc;
  }
  *out = tr;
  return LBSWIM_OK;
}

void lbswim_tracers_destroy(lbswim_tracers* tr) {
  if (!tr) return;
  if (g_rt.ready) (void)hipStreamSynchronize(stream());
  dev_free(tr->d_r);
  dev_free(tr->d_s);
  dev_free(tr->d_v);
  for (void* h : tr->h) std::free(h);
  delete tr;
}

int lbswim_tracers_num(const lbswim_tracers* tr, int* out) {
  CHECK_ARG(tr && out, "NULL argument");
  *out = tr->num;
  return LBSWIM_OK;
}

int lbswim_tracers_move(lbswim_tracers* tr, lbswim_lattice* lat) {
  CHECK_ARG(tr, "tracer handle is NULL");
  if (int rc = check_lattice(lat)) return rc;
  KernelTimer t(lat, LBSWIM_K_TMOVE);
  hipError_t e;
  if (lat->storage == LBSWIM_STORE_F64)
    e = launch_tracers_move<double>(tr->num, tr->d_r, tr->d_s, tr->d_v, static_cast<const double*>(lat->d_u),
                                    lat->geom, stream());
  else
    e = launch_tracers_move<float>(tr->num, tr->d_r, tr->d_s, tr->d_v, static_cast<const float*>(lat->d_u),
                                   lat->geom, stream());
  t.done();
  HIP_TRY(e);
  return LBSWIM_OK;
}

static double* tracer_dev(lbswim_tracers* tr, int field) {
  switch (field) {
    case LBSWIM_PART_R: return tr->d_r;
    case LBSWIM_PART_S: return tr->d_s;
    case LBSWIM_PART_V: return tr->d_v;
    default: return nullptr;
  }
}

static int tracer_host(lbswim_tracers* tr, int field, void** out) {
  CHECK_ARG(field == LBSWIM_PART_R || field == LBSWIM_PART_S || field == LBSWIM_PART_V, "unknown tracer field %d",
            field);
  if (!tr->h[field]) {
    const size_t cnt = 3 * size_t(tr->num);
    tr->h[field] = std::calloc(cnt ? cnt : 1, sizeof(double));
    if (!tr->h[field]) return fail(LBSWIM_ENOMEM, "host mirror allocation failed");
  }
  *out = tr->h[field];
  return LBSWIM_OK;
}

int lbswim_tracers_field(lbswim_tracers* tr, int field, void** host, void** device, size_t* count) {
  CHECK_ARG(tr, "tracer handle is NULL");
  void* h = nullptr;
  if (int rc = tracer_host(tr, field, &h)) return rc;
  if (host) *host = h;
  if (device) *device = tracer_dev(tr, field);
  if (count) *count = 3 * size_t(tr->num);
  return LBSWIM_OK;
}

int lbswim_tracers_h2d(lbswim_tracers* tr, int field) {
  CHECK_ARG(tr, "tracer handle is NULL");
  void* h = nullptr;
  if (int rc = tracer_host(tr, field, &h)) return rc;
  return copy_h2d(tracer_dev(tr, field), h, 3 * sizeof(double) * size_t(tr->num));
}

int lbswim_tracers_d2h(lbswim_tracers* tr, int field) {
  CHECK_ARG(tr, "tracer handle is NULL");
  void* h = nullptr;
  if (int rc = tracer_host(tr, field, &h)) return rc;
  return copy_d2h(h, tracer_dev(tr, field), 3 * sizeof(double) * size_t(tr->num));
}

// ---- coupled step -------------------------------------------------------------

int lbswim_system_step(lbswim_lattice* lat, lbswim_swimmers* const* sw, int nsw, lbswim_tracers* const* tr, int ntr,
                       int nsteps) {
  if (int rc = check_lattice(lat)) return rc;
  CHECK_ARG(nsw >= 0 && ntr >= 0 && nsteps >= 0, "negative count");
  CHECK_ARG(nsw == 0 || sw != nullptr, "swimmer list is NULL");
  CHECK_ARG(ntr == 0 || tr != nullptr, "tracer list is NULL");
  for (int i = 0; i < nsw; ++i) CHECK_ARG(sw[i] != nullptr, "swimmer handle %d is NULL", i);
  for (int i = 0; i < ntr; ++i) CHECK_ARG(tr[i] != nullptr, "tracer handle %d is NULL", i);
  for (int s = 0; s < nsteps; ++s) {
    int rc;
    if ((rc = lattice_zero_force_impl(lat))) return rc;
    for (int i = 0; i < nsw; ++i)
      if ((rc = lbswim_swimmers_add_forces(sw[i], lat))) return rc;
    if ((rc = lattice_step_impl(lat))) return rc;
    if ((rc = lattice_hydro_impl(lat, false))) return rc;
    for (int i = 0; i < nsw; ++i)
      if ((rc = lbswim_swimmers_move(sw[i], lat))) return rc;
    for (int i = 0; i < ntr; ++i)
      if ((rc = lbswim_tracers_move(tr[i], lat))) return rc;
  }
  return LBSWIM_OK;
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
  if (int rc = check_lattice(lat)) return rc;
  if (on) {
    HIP_TRY(hipStreamSynchronize(stream()));
    for (auto& r : lat->recs) {
      lat->pool.push_back(r.start);
      lat->pool.push_back(r.stop);
    }
    lat->recs.clear();
  }
  lat->timing = on != 0;
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
