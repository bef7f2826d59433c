// C++ classes with the reference's names and members, over the C ABI in lbswim.h.
//
// Mirrors lbswim/liblbswim/Shared.h:4-25 (SharedArray, SharedItem),
// Lattice.h:8-67 (LatticeAddressing, LatticeData, Lattice),
// SwimmerArray.h:8-29 (CommonParams, SwimmerArray) and TracerArray.h:6-14
// (TracerArray), so code written against the reference's C++ API compiles
// against this header.  Differences: failures throw lbswim::Error instead of
// exiting; Lattice takes an optional storage form; swimmer `prng` holds a
// uint64 draw counter per swimmer; device buffers of particle fields hold
// the particles in a library-chosen storage order (H2D/D2H translate to the
// user order; slots() gives the user index of each storage slot); the
// device distributions fOld/fNew hold post-collision values (pull
// streaming, see lbswim_lattice_field).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "lbswim.h"

namespace lbswim {

class Error : public std::runtime_error {
 public:
  Error(int code, const char* msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != LBSWIM_OK) throw Error(rc, lbswim_last_error());
}

using LBParams = lbswim_params;                // LBParams.h:4-30
using LatticeAddressing = lbswim_addressing;  // Lattice.h:8-12
using CommonParams = lbswim_common_params;    // SwimmerArray.h:8-16

// A host mirror and its device copy (Shared.h:4-15).  H2D/D2H forward to the
// owning object's C entry points.  `device` is typed as in the reference
// (Shared.h:8) where the device buffer holds T: every field of the FP64
// lattice form and every particle field.  A lattice in the FP32 storage form
// keeps its fields on the device as float (distributions and rho as
// deviations from the rest state); there `device` is nullptr and
// `device_raw` is the float buffer.  Both stay valid for the owner's life,
// except fOld/fNew of a lattice, which trade buffers at every Step.
template <typename T>
struct SharedArray {
  T* host = nullptr;
  T* device = nullptr;
  void* device_raw = nullptr;
  size_t size = 0;
  void H2D() { check(h2d_(owner_, field_)); }
  void D2H() { check(d2h_(owner_, field_)); }

  using Copy = int (*)(void*, int);
  void bind(void* owner, int field, Copy h2d, Copy d2h) {
    owner_ = owner;
    field_ = field;
    h2d_ = h2d;
    d2h_ = d2h;
  }

 private:
  void* owner_ = nullptr;
  int field_ = 0;
  Copy h2d_ = nullptr;
  Copy d2h_ = nullptr;
};

// A single host struct and its device copy (Shared.h:17-25).  For the
// lattice's params/addr the host copy is authoritative and H2D/D2H do nothing.
template <typename T>
struct SharedItem {
  T* host = nullptr;
  void H2D() {
    if (h2d_) check(h2d_(owner_));
  }
  void D2H() {
    if (d2h_) check(d2h_(owner_));
  }

  using Sync = int (*)(void*);
  void bind(void* owner, Sync h2d, Sync d2h) {
    owner_ = owner;
    h2d_ = h2d;
    d2h_ = d2h;
  }

 private:
  void* owner_ = nullptr;
  Sync h2d_ = nullptr;
  Sync d2h_ = nullptr;
};

namespace detail {
inline int lat_h2d(void* o, int f) { return lbswim_lattice_h2d(static_cast<lbswim_lattice*>(o), f); }
inline int lat_d2h(void* o, int f) { return lbswim_lattice_d2h(static_cast<lbswim_lattice*>(o), f); }
inline int sw_h2d(void* o, int f) { return lbswim_swimmers_h2d(static_cast<lbswim_swimmers*>(o), f); }
inline int sw_d2h(void* o, int f) { return lbswim_swimmers_d2h(static_cast<lbswim_swimmers*>(o), f); }
inline int tr_h2d(void* o, int f) { return lbswim_tracers_h2d(static_cast<lbswim_tracers*>(o), f); }
inline int tr_d2h(void* o, int f) { return lbswim_tracers_d2h(static_cast<lbswim_tracers*>(o), f); }
inline int cp_h2d(void* o) { return lbswim_swimmers_common_h2d(static_cast<lbswim_swimmers*>(o)); }
inline int cp_d2h(void* o) { return lbswim_swimmers_common_d2h(static_cast<lbswim_swimmers*>(o)); }
}  // namespace detail

// Lattice.h:38-50
struct LatticeData {
  SharedArray<double> rho, u, force, fOld, fNew;
};

// Lattice.h:53-67; lb.cu:40-428
class Lattice {
 public:
  Lattice(int nx, int ny, int nz, double tau_s, double tau_b, int storage = LBSWIM_STORE_F64) {
    check(lbswim_lattice_create(nx, ny, nz, tau_s, tau_b, storage, &impl_));
    const lbswim_params* p = nullptr;
    const lbswim_addressing* a = nullptr;
    check(lbswim_lattice_params(impl_, &p));
    check(lbswim_lattice_addressing(impl_, &a));
    params.host = const_cast<LBParams*>(p);
    addr.host = const_cast<LatticeAddressing*>(a);
    data = &data_;
    bind(data_.rho, LBSWIM_RHO);
    bind(data_.u, LBSWIM_U);
    bind(data_.force, LBSWIM_FORCE);
    bind(data_.fOld, LBSWIM_FOLD);
    bind(data_.fNew, LBSWIM_FNEW);
  }
  ~Lattice() { lbswim_lattice_destroy(impl_); }
  Lattice(const Lattice&) = delete;
  Lattice& operator=(const Lattice&) = delete;

  void Step() {
    check(lbswim_lattice_step(impl_));
    refresh_devices();  // fOld and fNew swap device buffers (lb.cu:242-244)
    ++time_step;
  }
  void CalcHydro() { check(lbswim_lattice_calc_hydro(impl_)); }
  void InitFromHydro() { check(lbswim_lattice_init_from_hydro(impl_)); }
  void ZeroForce() { check(lbswim_lattice_zero_force(impl_)); }

  SharedItem<LBParams> params;
  SharedItem<LatticeAddressing> addr;
  LatticeData* data = nullptr;
  int time_step = 0;

  lbswim_lattice* handle() const { return impl_; }
  void refresh_devices() {
    set_device(data_.fOld, LBSWIM_FOLD);
    set_device(data_.fNew, LBSWIM_FNEW);
  }
  // After coupled steps (lbswim_system_step) u and the force may be held in
  // another form on the device; these write the reference-layout arrays that
  // data->u.device and data->force.device point at.
  void sync_u() { check(lbswim_lattice_sync_u(impl_)); }
  void sync_force() { check(lbswim_lattice_sync_force(impl_)); }

 private:
  void set_device(SharedArray<double>& a, int field) {
    void* dev = nullptr;
    check(lbswim_lattice_field(impl_, field, nullptr, &dev, nullptr));
    int storage = LBSWIM_STORE_F64;
    check(lbswim_lattice_storage(impl_, &storage));
    a.device_raw = dev;
    a.device = storage == LBSWIM_STORE_F64 ? static_cast<double*>(dev) : nullptr;
  }
  void bind(SharedArray<double>& a, int field) {
    check(lbswim_lattice_field(impl_, field, &a.host, nullptr, &a.size));
    set_device(a, field);
    a.bind(impl_, field, detail::lat_h2d, detail::lat_d2h);
  }
  lbswim_lattice* impl_ = nullptr;
  LatticeData data_;
};

// SwimmerArray.h:18-29; SwimmerArray.cu:16-220
class SwimmerArray {
 public:
  static const int BlockSize = 512;  // kept for source compatibility (SwimmerArray.h:19)

  SwimmerArray(int n, const CommonParams* p) : num(n) {
    check(lbswim_swimmers_create(n, p, &impl_));
    CommonParams* host = nullptr;
    check(lbswim_swimmers_common(impl_, &host));
    common.host = host;
    common.bind(impl_, detail::cp_h2d, detail::cp_d2h);
    bind(r, LBSWIM_PART_R);
    bind(v, LBSWIM_PART_V);
    bind(this->n, LBSWIM_PART_N);
    bind(prng, LBSWIM_PART_PRNG);
  }
  ~SwimmerArray() { lbswim_swimmers_destroy(impl_); }
  SwimmerArray(const SwimmerArray&) = delete;
  SwimmerArray& operator=(const SwimmerArray&) = delete;

  void AddForces(Lattice* lat) const { check(lbswim_swimmers_add_forces(impl_, lat->handle())); }
  void Move(Lattice* lat) { check(lbswim_swimmers_move(impl_, lat->handle())); }

  int num;
  SharedItem<CommonParams> common;
  SharedArray<double> r;
  SharedArray<double> v;
  SharedArray<double> n;
  SharedArray<uint64_t> prng;

  lbswim_swimmers* handle() const { return impl_; }
  // Device array of num ints: the user index of the swimmer in each storage
  // slot of the device fields (no reference counterpart).
  const int* slots() const {
    const int* p = nullptr;
    check(lbswim_swimmers_slots(impl_, &p));
    return p;
  }

 private:
  template <typename T>
  void bind(SharedArray<T>& a, int field) {
    void* host = nullptr;
    void* dev = nullptr;
    size_t elem = 0;
    check(lbswim_swimmers_field(impl_, field, &host, &dev, &a.size, &elem));
    a.host = static_cast<T*>(host);
    a.device = static_cast<T*>(dev);
    a.device_raw = dev;
    a.bind(impl_, field, detail::sw_h2d, detail::sw_d2h);
  }
  lbswim_swimmers* impl_ = nullptr;
};

// TracerArray.h:6-14; TracerArray.cu:5-52
class TracerArray {
 public:
  static const int BlockSize = 512;  // kept for source compatibility (TracerArray.h:11)

  explicit TracerArray(int n) : num(n) {
    check(lbswim_tracers_create(n, &impl_));
    bind(r, LBSWIM_PART_R);
    bind(s, LBSWIM_PART_S);
    bind(v, LBSWIM_PART_V);
  }
  ~TracerArray() { lbswim_tracers_destroy(impl_); }
  TracerArray(const TracerArray&) = delete;
  TracerArray& operator=(const TracerArray&) = delete;

  void Move(Lattice* lat) { check(lbswim_tracers_move(impl_, lat->handle())); }

  int num;
  SharedArray<double> r;
  SharedArray<double> s;
  SharedArray<double> v;

  lbswim_tracers* handle() const { return impl_; }
  // Device array of num ints: the user index of the tracer in each storage slot.
  const int* slots() const {
    const int* p = nullptr;
    check(lbswim_tracers_slots(impl_, &p));
    return p;
  }

 private:
  void bind(SharedArray<double>& a, int field) {
    void* host = nullptr;
    void* dev = nullptr;
    check(lbswim_tracers_field(impl_, field, &host, &dev, &a.size));
    a.host = static_cast<double*>(host);
    a.device = static_cast<double*>(dev);
    a.device_raw = dev;
    a.bind(impl_, field, detail::tr_h2d, detail::tr_d2h);
  }
  lbswim_tracers* impl_ = nullptr;
};

}  // namespace lbswim
