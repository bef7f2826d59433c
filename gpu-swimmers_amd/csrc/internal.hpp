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
struct SpreadPoint;
// Entry k of the sorted lists for the fused step's row walk (lattices with nz
// % 64 == 0, where a list's 64-site run is one row x, y): the point's row
// weight dx_i dy_j at that run, its z weights and its force.  The z of its
// first stencil site is kept beside it (wz[k]), -1 for an entry the walk
// skips (a repeated id, or a point whose stencil misses the row).
struct RunEntry {
  double dxy;
  double zw[4];
  double F[3];
};
static_assert(sizeof(RunEntry) == 64, "RunEntry layout");
// Sorted point lists of the lattice's 64-site runs (see launch_spread_lists).
// Passed to the collide and CalcHydro launchers instead of a stored force
// field, they evaluate every site's force from its block's list.
struct VForce {
  const SpreadPoint* pts;
  const uint32_t* off;
  const uint32_t* sorted;
  const double* edge;    // launch_edge_force output (fused step only)
  const RunEntry* rec;   // launch_edge_force entries (fused step only)
  const int* wz;
};
// `vf` null: the force is read from `force`.
template <typename T>
hipError_t launch_collide_stream(const T* fOld, const T* force, T* fNew, const Geom& g, double tau_s,
                                 double tau_b, const VForce* vf, hipStream_t s);
// aos: u written component-interleaved (u[3 ijk + a], 3 n values).
template <typename T>
hipError_t launch_calc_hydro(const T* f, const T* force, T* rho, T* u, const Geom& g, const VForce* vf,
                             hipStream_t s, bool aos = false);
// Collide+stream followed by CalcHydro in one launch, with the results of
// launch_collide_stream + launch_calc_hydro bit for bit, for lattices where
// fused_step_supported(g, ty, sizeof(T)): ny >= ty, ty in {2, 4, 8} (rows
// per tile), and 15 n sizeof(T) < 4 GiB (32-bit buffer offsets); a virtual
// force `vf` also needs nz % 64 == 0.  `strips` splits the x sweep (0: one
// workgroup per CU).
bool fused_step_supported(const Geom& g, int ty, size_t elem);
// nz % 64 == 0 and ny % ty == 0: no tile overlaps another and no lane idles
// (the fused launch is then the faster one; the virtual force's row walk
// needs nz % 64 == 0).
bool fused_tiles_full(const Geom& g, int ty);
// u is written only component-interleaved, u_aos[3 ijk + a] (3 n values),
// which the particle moves gather from; launch_u_from_aos writes the
// reference layout from it.  zero_next (stored force only, not `force`):
// a force array the launch clears, for the scatter of the next step's force.
template <typename T>
hipError_t launch_step_hydro(const T* fOld, const T* force, T* fNew, T* rho, T* u_aos, const Geom& g, double tau_s,
                             double tau_b, const VForce* vf, int ty, int strips, hipStream_t s,
                             T* zero_next = nullptr);
template <typename T>
hipError_t launch_u_from_aos(const T* aos, T* u, const Geom& g, hipStream_t s);
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
// aos: u holds the three components of a site together (u[3 ijk + a]),
// the copy the fused step writes for the moves; else u[a n + ijk].
// scatter (FP32 storage with aos only): also adds each moved swimmer's dipole
// -- the next step's ZeroForce + AddForces -- into this force array with
// atomics (the array must hold zeros, or other arrays' dipoles, before).
template <typename T>
hipError_t launch_swimmers_move(int ns, const SwimParams& p, double* r, double* v, double* n,
                                uint64_t* prng, const int* perm, const T* u, bool aos, const Geom& g,
                                unsigned long long* tumbles, hipStream_t s, T* scatter = nullptr);
template <typename T>
hipError_t launch_tracers_move(int nt, double* r, double* sdisp, double* v, const T* u, bool aos, const Geom& g,
                               hipStream_t s);

// ---- deterministic spreading (ZeroForce + AddForces of all swimmer arrays)
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
// One list per run of 64 consecutive sites -- one wavefront of the lattice
// kernels' 256-site workgroups: spread_lists(g) = 4 ceil(n / 256) of them.
uint32_t spread_lists(const Geom& g);
size_t spread_list_capacity(size_t npts);
// Entries of the `off` array of launch_spread_lists for nlists lists.
size_t spread_off_entries(uint32_t nlists);
// Point records of the swimmers of one array, ids from 2*base;
// slot_ids[2 base + t] receives the id of the t-th point in storage order.
hipError_t launch_points(int ns, int base, const SwimParams& p, const double* r, const double* n, const int* perm,
                         const Geom& g, SpreadPoint* pts, uint32_t* slot_ids, hipStream_t s);
// Lists averaging more than 8 entries: the list building aggregates its
// atomics per wavefront.
inline bool dense_lists(size_t npts, uint32_t nlists) { return npts * 17 > size_t(8) * nlists; }
// counts[nl] (zero on entry: zeroed at allocation, left zero by every build),
// off[spread_off_entries(nl)] (offsets, then scan and sort scratch),
// cursor[nl], list/sorted[spread_list_capacity]; afterwards
// sorted[off[l]..off[l+1]) holds the ids of the points reaching sites
// 64 l .. 64 l + 63 in increasing order.
hipError_t launch_spread_lists(int npts, const Geom& g, const SpreadPoint* pts, const uint32_t* slot_ids,
                               uint32_t* counts, uint32_t* off, uint32_t* cursor, uint32_t* list, uint32_t* sorted,
                               hipStream_t s);
// Writes the force of every site from all npts point records in id order,
// without lists: the sum the sorted lists give (dense small systems).
template <typename T>
hipError_t launch_range_force(const Geom& g, const SpreadPoint* pts, uint32_t npts, T* force, hipStream_t s);
// Writes the force of every site from the lists (ZeroForce + AddForces).
template <typename T>
hipError_t launch_spread_force(const Geom& g, const SpreadPoint* pts, const uint32_t* off, const uint32_t* sorted,
                               T* force, hipStream_t s);
// Writes the force of every site from per-site segments of the points' 64
// stencil entries, for lattices with every side >= 4 (site_force_ok): the
// sum the sorted lists give.  counts[g.n] (zero on entry, left zero),
// off[spread_off_entries(g.n)]; scratch local[64 npts] (16-byte aligned),
// entries[2 * 64 npts] (key and site of each entry, 8-byte aligned) and
// sorted[3 * 64 npts].
inline bool site_force_ok(const Geom& g) { return g.nx >= 4 && g.ny >= 4 && g.nz >= 4; }
template <typename T>
hipError_t launch_site_force(int npts, const Geom& g, const SpreadPoint* pts, const uint32_t* slot_ids,
                             uint32_t* counts, uint32_t* off, uint32_t* local, uint32_t* entries, double* sorted,
                             T* force, hipStream_t s);

// Force at the first and last site of every list's 64-site run:
// edge[6 l + 3 e + a] (e = 0 first, 1 last), 6 spread_lists(g) doubles; and,
// for g.nz % 64 == 0, the RunEntry rec[k] and wz[k] of every list entry k
// (capacity spread_list_capacity).
hipError_t launch_edge_force(const Geom& g, const SpreadPoint* pts, const uint32_t* off, const uint32_t* sorted,
                             double* edge, RunEntry* rec, int* wz, hipStream_t s);

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
  uint32_t tile_base;    // first spreading list of this member in the batch
  uint32_t point_first;  // first point id of this member (its arrays' points are consecutive)
  uint32_t point_count;
  uint32_t site_base;    // first site of this member in the batch's site segments (FORCE_SITES_STORED)
  // how the member's swimmer force reaches its collide and CalcHydro:
  // FORCE_LISTS_VIRTUAL (evaluated from its lists there), FORCE_LISTS_STORED
  // (written from its lists to the force array first), FORCE_RANGE_STORED
  // (written from all its points in id order, no lists), FORCE_SITES_STORED
  // (written from per-site segments of its points' stencil entries, no lists)
  int force_path;
  // the interleaved u of a member the batched fused step advances (it then
  // writes u only there, and the batched moves gather from it); null for the
  // others
  void* u_aos;
};
constexpr int FORCE_LISTS_VIRTUAL = 0;
constexpr int FORCE_LISTS_STORED = 1;
constexpr int FORCE_RANGE_STORED = 2;
constexpr int FORCE_SITES_STORED = 3;
__host__ __device__ inline bool has_lists(int force_path) {
  return force_path == FORCE_LISTS_VIRTUAL || force_path == FORCE_LISTS_STORED;
}
// Points up to which a dense (stored-force) system takes FORCE_RANGE_STORED:
// every 64-site run walks all of them instead of building lists that would
// hold a large share of them anyway.
constexpr size_t RANGE_FORCE_MAX_POINTS = 4096;
// On lattices with every side >= 4, dense systems with more points than this
// take FORCE_SITES_STORED instead.
constexpr size_t SITE_FORCE_MIN_POINTS = 1024;
// Site segments key their entries by point id * 64 + stencil index: point
// ids (numbered across an ensemble) below 2^26.
constexpr size_t SITE_FORCE_MAX_POINTS = size_t(1) << 26;
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
constexpr int PARTICLE_GROUPS = 16;      // swimmers per workgroup of the swimmer move kernel
constexpr int TRACER_GROUPS = 64;        // tracers per workgroup of the tracer move kernel
// vf: the work covers members with FORCE_LISTS_VIRTUAL, whose force of block
// b comes from lists tile_base_m + 4 b ...; null: members that read their
// force array.
// Each work item covers `per` consecutive 256-site blocks of its member.
// Without vf, the members with FORCE_SITES_STORED take their force from the
// site segments soff/ssorted (the sum launch_batch_spread_force forms) and
// store it in their force arrays for the CalcHydro.
template <typename T>
hipError_t launch_batch_collide(const LatDesc* desc, const WorkItem* work, int nwork, int per, int parity,
                                const VForce* vf, hipStream_t s, const uint32_t* soff = nullptr,
                                const double* ssorted = nullptr);
template <typename T>
hipError_t launch_batch_hydro(const LatDesc* desc, const WorkItem* work, int nwork, int per, int parity,
                              const VForce* vf, hipStream_t s);
// Members the batched fused step can advance with `rows`-row tiles (8 or 4):
// full tiles (nz % 64 == 0, rows dividing ny), nx and ny >= 4 (the row walk
// over RunEntry records), offsets below 4 GiB.
bool batch_fused_member(const Geom& g, int rows, size_t elem);
// Collide/stream + CalcHydro of those members in one launch, bit for bit the
// two batched launches; each work item is one tile-strip (k_step_hydro's
// index within its member) with strips of `strip` planes.  vf holds the
// batch's point records, list bounds and sorted ids and the RunEntry records
// and run-end forces of launch_batch_run_entries.
template <typename T>
hipError_t launch_batch_step_hydro(const LatDesc* desc, const WorkItem* work, int nwork, int parity, const VForce& vf,
                                   int rows, int strip, hipStream_t s);
// RunEntry records, wz and run-end forces (launch_edge_force's outputs) of
// the lists of the members in `work`: each work item is 64 consecutive lists
// (unit 64 l ...) of its member.
hipError_t launch_batch_run_entries(const LatDesc* ldesc, const WorkItem* work, int nwork, const SpreadPoint* pts,
                                    const uint32_t* off, const uint32_t* sorted, RunEntry* rec, int* wz, double* edge,
                                    hipStream_t s);
// Point records of a batch: points of array a get ids point_base_a + 2*user +
// end; work items are blocks of 128 swimmers (256 points) of an array.
hipError_t launch_batch_points(const LatDesc* ldesc, const SwimDesc* sdesc, const WorkItem* work, int nwork,
                               SpreadPoint* pts, uint32_t* slot_ids, hipStream_t s);
// Sorted lists of all nlists lists of the batch (see launch_spread_lists);
// dense: aggregate the atomics per wavefront.
hipError_t launch_batch_lists(const LatDesc* ldesc, int npts, int nlists, const SpreadPoint* pts,
                              const uint32_t* slot_ids, bool dense, uint32_t* counts, uint32_t* off, uint32_t* cursor,
                              uint32_t* list, uint32_t* sorted, hipStream_t s);
// Per-site segments of the FORCE_SITES_STORED members' points (see
// launch_site_force), nsites = their sites (site_base numbering), npts all
// the batch's points, nent = 64 points of those members: counts[nsites]
// (zero on entry, left zero), off[spread_off_entries(nsites)],
// local[64 npts] (16-byte aligned), entries[2 nent] (8-byte aligned),
// sorted[3 nent].
hipError_t launch_batch_site_lists(const LatDesc* ldesc, int npts, int nsites, int nent, const SpreadPoint* pts,
                                   const uint32_t* slot_ids, uint32_t* counts, uint32_t* off, uint32_t* local,
                                   uint32_t* entries, double* sorted, hipStream_t s);
// Writes the force arrays of the lattice blocks in block_work: from the lists
// (FORCE_LISTS_STORED), from all the member's points (FORCE_RANGE_STORED) or
// from its site segments soff/ssorted (FORCE_SITES_STORED; none where soff is
// null, for launch_batch_collide to write from the segments).
template <typename T>
hipError_t launch_batch_spread_force(const LatDesc* ldesc, int nwork, const WorkItem* block_work, int per,
                                     const SpreadPoint* pts, const uint32_t* off, const uint32_t* sorted,
                                     const uint32_t* soff, const double* ssorted, hipStream_t s);
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

// ---- fixed workspace of the no-allocation coupled step (noalloc.hip)
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
// Device address of this device's workspace (looked up once per device).
hipError_t noalloc_scratch(void** out);

// ---- diagnostics
hipError_t launch_copy(const double* src, double* dst, size_t count, hipStream_t s);
hipError_t launch_peskin_probe(const double* x, double* out, int count, hipStream_t s);
hipError_t launch_interp_probe(const double* u, const double* r, double* v, int count, const Geom& g,
                               hipStream_t s);

}  // namespace lbs
