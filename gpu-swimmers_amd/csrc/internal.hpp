// Internal launcher interface between the C ABI (capi.cpp) and the kernels.
// Every launcher only enqueues work on `stream`; none synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "d3q15.hpp"

namespace lbs {

// Swimmer parameters as the kernels receive them (by value).  Same members as
// the reference's CommonParams (lbswim/liblbswim/SwimmerArray.h:8-16).
struct SwimParams {
  double P;
  double l;
  double alpha;
  double mobility;
  unsigned char translational_advection_off;
  unsigned char rotational_advection_off;
  unsigned long long seed;
};

Geom make_geom(int nx, int ny, int nz);

// ---- lattice (storage T = double: reference arithmetic; T = float: deviation form)
template <typename T>
hipError_t launch_collide_stream(const T* fOld, const T* force, T* fNew, const Geom& g, double tau_s,
                                 double tau_b, hipStream_t s);
template <typename T>
hipError_t launch_calc_hydro(const T* f, T* force, T* rho, T* u, const Geom& g, bool zero_force,
                             hipStream_t s);
template <typename T>
hipError_t launch_init_from_hydro(const T* rho, const T* u, T* f, const Geom& g, hipStream_t s);

hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s);

// FP64 mass and momentum totals of the current distributions: out[0] = sum rho,
// out[1..3] = sum of momentum; `partial` holds 4*nblocks doubles.
template <typename T>
hipError_t launch_totals(const T* f, const Geom& g, double* partial, int nblocks, double* out,
                         hipStream_t s);

// ---- particles
template <typename T>
hipError_t launch_swimmers_add_forces(int ns, const SwimParams& p, const double* r, const double* n,
                                      T* force, const Geom& g, hipStream_t s);
template <typename T>
hipError_t launch_swimmers_move(int ns, const SwimParams& p, double* r, double* v, double* n,
                                uint64_t* prng, const T* u, const Geom& g,
                                unsigned long long* tumbles, hipStream_t s);
template <typename T>
hipError_t launch_tracers_move(int nt, double* r, double* sdisp, double* v, const T* u, const Geom& g,
                               hipStream_t s);

// ---- diagnostics
hipError_t launch_copy(const double* src, double* dst, size_t count, hipStream_t s);
hipError_t launch_peskin_probe(const double* x, double* out, int count, hipStream_t s);
hipError_t launch_interp_probe(const double* u, const double* r, double* v, int count, const Geom& g,
                               hipStream_t s);

}  // namespace lbs
