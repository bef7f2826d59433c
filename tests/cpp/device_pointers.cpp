// Device pointers of the C++ API (include/lbswim.hpp) read the way the
// reference's C++ users read SharedArray<T>::device (Shared.h:8): typed, and
// fixed for the object's life.
//   device_pointers [steps=70]
// A 16x12x20 FP64 lattice with 40 swimmers and 300 tracers is stepped with the
// individual methods (ZeroForce, AddForces, Step, CalcHydro, Move) for more
// than the 64 moves after which the library re-orders particle storage.  Then:
//   - lat.data->u.device and rho.device, read as double*, equal D2H;
//   - each particle field's device pointer is the one bound at construction,
//     and its contents, mapped through slots(), equal D2H;
//   - an FP32-storage lattice has device == nullptr and a float device_raw
//     holding u.
// Prints "ok" and exits 0 on success.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lbswim.hpp"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what);
    ++failures;
  }
}

template <typename T>
std::vector<T> fetch(const T* dev, size_t count) {
  std::vector<T> out(count);
  if (hipMemcpy(out.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
    std::printf("FAIL: hipMemcpy from a device pointer\n");
    std::exit(1);
  }
  return out;
}

// device[c*num + s] == host[c*num + slots[s]] for every component and slot
template <typename T>
bool matches_through_slots(const lbswim::SharedArray<T>& a, const int* slots_dev, int num, int ncomp) {
  const std::vector<T> dev = fetch(a.device, size_t(ncomp) * num);
  const std::vector<int> slots = fetch(slots_dev, size_t(num));
  std::vector<int> seen(num, 0);
  for (int s = 0; s < num; ++s) {
    if (slots[s] < 0 || slots[s] >= num || seen[slots[s]]++) return false;
    for (int c = 0; c < ncomp; ++c)
      if (dev[size_t(c) * num + s] != a.host[size_t(c) * num + slots[s]]) return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  const int steps = argc > 1 ? std::atoi(argv[1]) : 70;
  try {
    const int nx = 16, ny = 12, nz = 20;
    lbswim::Lattice lat(nx, ny, nz, 0.25, 0.25);
    const int n = lat.addr.host->n;
    unsigned seed = 12345u;
    auto uniform = [&seed]() {
      seed = seed * 1664525u + 1013904223u;
      return double(seed >> 8) / double(1u << 24);
    };
    for (int i = 0; i < n; ++i) lat.data->rho.host[i] = 1.0 + 1e-3 * (uniform() - 0.5);
    for (size_t i = 0; i < lat.data->u.size; ++i) lat.data->u.host[i] = 1e-3 * (uniform() - 0.5);
    lat.data->rho.H2D();
    lat.data->u.H2D();
    lat.ZeroForce();

    lbswim::CommonParams cp{};
    cp.P = 1e-3;
    cp.l = 0.1;
    cp.alpha = 0.0;
    cp.mobility = 5.94;
    cp.seed = 99;
    const int ns = 40, nt = 300;
    lbswim::SwimmerArray sw(ns, &cp);
    lbswim::TracerArray tr(nt);
    const int L[3] = {nx, ny, nz};
    for (int c = 0; c < 3; ++c) {
      for (int i = 0; i < ns; ++i) sw.r.host[c * ns + i] = L[c] * uniform();
      for (int i = 0; i < nt; ++i) tr.r.host[c * nt + i] = L[c] * uniform();
    }
    for (int i = 0; i < ns; ++i) {
      double v[3], nrm = 0.0;
      for (double& x : v) x = uniform() - 0.5, nrm += x * x;
      for (int c = 0; c < 3; ++c) sw.n.host[c * ns + i] = v[c] / std::sqrt(nrm);
    }
    sw.r.H2D();
    sw.n.H2D();
    tr.r.H2D();
    lat.InitFromHydro();

    // typed as in the reference: these lines do not compile against a void* device
    double* const sw_r0 = sw.r.device;
    double* const sw_n0 = sw.n.device;
    uint64_t* const sw_prng0 = sw.prng.device;
    double* const tr_r0 = tr.r.device;
    double* const u0 = lat.data->u.device;

    for (int s = 0; s < steps; ++s) {
      lat.ZeroForce();
      sw.AddForces(&lat);
      lat.Step();
      lat.CalcHydro();
      sw.Move(&lat);
      tr.Move(&lat);
    }
    lbswim::check(lbswim_synchronize());

    // the bound pointers still name the arrays' buffers
    void *dev = nullptr, *host = nullptr;
    size_t cnt = 0, elem = 0;
    lbswim::check(lbswim_swimmers_field(sw.handle(), LBSWIM_PART_R, &host, &dev, &cnt, &elem));
    expect(dev == sw_r0, "swimmer r device pointer changed");
    lbswim::check(lbswim_swimmers_field(sw.handle(), LBSWIM_PART_N, &host, &dev, &cnt, &elem));
    expect(dev == sw_n0, "swimmer n device pointer changed");
    lbswim::check(lbswim_swimmers_field(sw.handle(), LBSWIM_PART_PRNG, &host, &dev, &cnt, &elem));
    expect(dev == sw_prng0, "swimmer prng device pointer changed");
    lbswim::check(lbswim_tracers_field(tr.handle(), LBSWIM_PART_R, &host, &dev, &cnt));
    expect(dev == tr_r0, "tracer r device pointer changed");
    expect(lat.data->u.device == u0, "u device pointer changed");

    // lattice fields through the typed pointers
    lat.data->u.D2H();
    lat.data->rho.D2H();
    const std::vector<double> u = fetch(lat.data->u.device, lat.data->u.size);
    const std::vector<double> rho = fetch(lat.data->rho.device, lat.data->rho.size);
    bool same = true;
    for (size_t i = 0; i < u.size(); ++i) same = same && u[i] == lat.data->u.host[i];
    expect(same, "u read through data->u.device differs from D2H");
    same = true;
    for (size_t i = 0; i < rho.size(); ++i) same = same && rho[i] == lat.data->rho.host[i];
    expect(same, "rho read through data->rho.device differs from D2H");

    // particle fields in storage order, mapped to the user order by slots()
    sw.r.D2H();
    sw.n.D2H();
    sw.prng.D2H();
    tr.r.D2H();
    expect(matches_through_slots(sw.r, sw.slots(), ns, 3), "swimmer r on the device differs from D2H");
    expect(matches_through_slots(sw.n, sw.slots(), ns, 3), "swimmer n on the device differs from D2H");
    expect(matches_through_slots(sw.prng, sw.slots(), ns, 1), "swimmer prng on the device differs from D2H");
    expect(matches_through_slots(tr.r, tr.slots(), nt, 3), "tracer r on the device differs from D2H");
    expect(sw.prng.host[0] == uint64_t(steps), "draw counter is not the step count");

    // the FP32 storage form: no typed pointer, a float buffer behind device_raw
    lbswim::Lattice lat32(8, 6, 5, 0.25, 0.25, LBSWIM_STORE_F32DEV);
    expect(lat32.data->u.device == nullptr, "FP32 lattice exposes a double* device");
    expect(lat32.data->u.device_raw != nullptr, "FP32 lattice has no device buffer");
    for (size_t i = 0; i < lat32.data->u.size; ++i) lat32.data->u.host[i] = 0.25 * double(i % 7);
    lat32.data->u.H2D();
    const std::vector<float> u32 = fetch(static_cast<const float*>(lat32.data->u.device_raw), lat32.data->u.size);
    same = true;
    for (size_t i = 0; i < u32.size(); ++i) same = same && u32[i] == float(0.25 * double(i % 7));
    expect(same, "FP32 u behind device_raw differs from what H2D wrote");
  } catch (const lbswim::Error& e) {
    std::printf("lbswim error %d: %s\n", e.code(), e.what());
    return 1;
  }
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
