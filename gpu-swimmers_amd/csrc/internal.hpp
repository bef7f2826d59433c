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
                                uint64_t* prng, const int* perm, const T* u, const Geom& g,
                                unsigned long long* tumbles, hipStream_t s);
template <typename T>
hipError_t launch_tracers_move(int nt, double* r, double* sdisp, double* v, const T* u, const Geom& g,
                               hipStream_t s);

// ---- deterministic spreading (ZeroForce + AddForces of all swimmer arrays)
// Lattice tiles of tx*ty*tz <= 512 sites; each stencil touches at most
// min(4, tiles) tiles per dimension.
struct SpreadGeom {
  int tx, ty, tz;
  int ntx, nty, ntz;
  int ntiles;
};
// Stencil record of one spreading point (tail or head of a swimmer), stored
// by point id = 2*(array base + user index) + (0 tail, 1 head): Peskin weights
// del[d*4+k] and wrapped site coordinates wc[d*4+k] of the four stencil
// positions k along each axis d, and the point force F.
struct SpreadPoint {
  double del[12];
  double F[3];
  int wc[12];
  int pad[2];
};
static_assert(sizeof(SpreadPoint) == 176, "SpreadPoint layout");
SpreadGeom make_spread_geom(const Geom& g);
size_t spread_list_capacity(const SpreadGeom& sg, size_t npts);
// counts[ntiles] must be zero before the first launch_points of a step.
hipError_t launch_points(int ns, int base, const SwimParams& p, const double* r, const double* n, const int* perm,
                         const Geom& g, const SpreadGeom& sg, SpreadPoint* pts, uint32_t* counts, hipStream_t s);
// off[ntiles + 1 + ceil(ntiles/1024)] (offsets, then scan scratch), cursor[ntiles],
// list/sorted[spread_list_capacity]
template <typename T>
hipError_t launch_spread(int npts, const Geom& g, const SpreadGeom& sg, const SpreadPoint* pts, uint32_t* counts,
                         uint32_t* off, uint32_t* cursor, uint32_t* list, uint32_t* sorted, T* force, hipStream_t s);

// ---- batched launches over the members of an ensemble
// One lattice of a batch.  g[cur ^ parity] holds the state at the start of a
// step launched with `parity`.
struct LatDesc {
  void* g[2];
  void* force;
  void* rho;
  void* u;
  Geom geom;
  double tau_s, tau_b, omega_s, omega_b;
  int cur;
  uint32_t tile_base;  // first spreading tile of this member in the batch
  SpreadGeom sg;
};
// One workgroup of a batched launch: member index and the block (or tile,
// or particle group) of that member it processes.
struct WorkItem {
  uint32_t member, unit;
};
// One swimmer array of a batch.
struct SwimDesc {
  double* r;
  double* v;
  double* n;
  uint64_t* prng;
  const int* perm;
  unsigned long long* tumbles;
  SwimParams p;
  int num;
  int member;
  uint32_t point_base;  // first point id of this array in the batch
};
// One tracer array of a batch.
struct TracerDesc {
  double* r;
  double* s;
  double* v;
  int num;
  int member;
};
uint32_t lattice_blocks(const Geom& g);  // workgroups of a one-lattice launch
constexpr int PARTICLE_GROUPS = 16;      // particles per workgroup of the move kernels
template <typename T>
hipError_t launch_batch_collide(const LatDesc* desc, const WorkItem* work, int nwork, int parity, hipStream_t s);
template <typename T>
hipError_t launch_batch_hydro(const LatDesc* desc, const WorkItem* work, int nwork, int parity, hipStream_t s);
// Force spreading of a batch: zero the tile counts first; points of array a go
// to ids point_base_a + 2*user + end; work over swimmer blocks of 256.
hipError_t launch_batch_points(const LatDesc* ldesc, const SwimDesc* sdesc, const WorkItem* work, int nwork,
                               SpreadPoint* pts, uint32_t* counts, hipStream_t s);
template <typename T>
hipError_t launch_batch_spread(const LatDesc* ldesc, int npts, int ntiles, const WorkItem* tile_work,
                               const SpreadPoint* pts, uint32_t* counts, uint32_t* off, uint32_t* cursor,
                               uint32_t* list, uint32_t* sorted, hipStream_t s);
template <typename T>
hipError_t launch_batch_swimmers_move(const LatDesc* ldesc, const SwimDesc* sdesc, const WorkItem* work, int nwork,
                                      hipStream_t s);
template <typename T>
hipError_t launch_batch_tracers_move(const LatDesc* ldesc, const TracerDesc* tdesc, const WorkItem* work, int nwork,
                                     hipStream_t s);

// ---- spatial ordering of particle storage (particle_order.hip)
hipError_t launch_iota(int np, int* p, hipStream_t s);
hipError_t sort_scratch_bytes(int np, size_t* bytes);
// order_out[t] = slot that moves to position t (stable Morton order of floor(r))
hipError_t launch_spatial_order(int np, const double* r, const Geom& g, uint32_t* keys, uint32_t* keys_alt, int* idx,
                                int* order_out, void* scratch, size_t scratch_bytes, hipStream_t s);
// dst[c*np + t] = src[c*np + order[t]]
template <typename T>
hipError_t launch_gather(int np, int ncomp, const int* order, const T* src, T* dst, hipStream_t s);
// dst[c*np + order[t]] = src[c*np + t]
template <typename T>
hipError_t launch_scatter(int np, int ncomp, const int* order, const T* src, T* dst, hipStream_t s);

// ---- diagnostics
hipError_t launch_copy(const double* src, double* dst, size_t count, hipStream_t s);
hipError_t launch_peskin_probe(const double* x, double* out, int count, hipStream_t s);
hipError_t launch_interp_probe(const double* u, const double* r, double* v, int count, const Geom& g,
                               hipStream_t s);

}  // namespace lbs
