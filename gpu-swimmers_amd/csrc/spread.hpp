// Force at a lattice site from the sorted point list of its 64-site run.
//
// The coupled step's ZeroForce + AddForces (SwimmerArray.cu:42-112) is
// evaluated as a gather: every point (the tail or head of a swimmer) is listed
// under each run of 64 consecutive sites (one wavefront of the lattice
// kernels) its 4x4x4 stencil reaches, the lists are sorted by point id =
// 2*(array base + user index) + end, and a site adds the contributions of its
// list's points in list order -- the swimmer order of a sequential
// restatement -- starting from zero.  A point may be listed twice in one list;
// the second entry is skipped.  Shared by the spreading kernel (which writes
// the force field) and the collide/CalcHydro kernels (which use the force
// without storing it).
#pragma once

#include <hip/hip_runtime.h>

#include "d3q15.hpp"
#include "internal.hpp"

namespace lbs {

constexpr int SITE_BLOCK = 256;  // sites per workgroup of the lattice kernels
constexpr int WAVE_SITES = 64;   // sites per list: one wavefront's sites
constexpr int LISTS_PER_BLOCK = SITE_BLOCK / WAVE_SITES;

struct SiteXYZ {
  int x, y, z;
};

// Coordinates of linear site s = (x*ny + y)*nz + z.
__device__ __forceinline__ SiteXYZ site_xyz(uint32_t s, const Geom& g) {
  const uint32_t t = fastdiv(s, g.div_nz);
  const uint32_t x = fastdiv(t, g.div_ny);
  return SiteXYZ{int(x), int(t - x * uint32_t(g.ny)), int(s - t * uint32_t(g.nz))};
}

// Calls f(list) for the list (site / 64) of each of the 4 stencil sites of
// row `row` = 4 i + j of a point, skipping a list equal to the previous one.
template <typename F>
__device__ __forceinline__ void for_each_row_list(const SpreadPoint& p, int row, const Geom& g, F&& f) {
  const uint32_t base =
      (uint32_t(p.wc[row >> 2]) * uint32_t(g.ny) + uint32_t(p.wc[4 + (row & 3)])) * uint32_t(g.nz);
  uint32_t last = 0xFFFFFFFFu;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t b = (base + uint32_t(p.wc[8 + k])) / WAVE_SITES;
    if (b != last) f(b);
    last = b;
  }
}

__device__ __forceinline__ double pick4(double a0, double a1, double a2, double a3, int i) {
  return i == 0 ? a0 : (i == 1 ? a1 : (i == 2 ? a2 : a3));
}

// Adds one point's contributions to site (X, Y, Z) in the (i, j, k) loop order
// of AccumulateDeltaForce.  `simple`: every lattice side is >= 4, so each axis
// has at most one matching stencil position, found by a wrapped difference.
__device__ __forceinline__ void add_point(const SpreadPoint* __restrict__ p, int X, int Y, int Z, const Geom& g,
                                          bool simple, double (&acc)[3]) {
  if (simple) {
    int di = X - p->wc[0];
    di += di < 0 ? g.nx : 0;
    int dj = Y - p->wc[4];
    dj += dj < 0 ? g.ny : 0;
    int dk = Z - p->wc[8];
    dk += dk < 0 ? g.nz : 0;
    if ((di | dj | dk) >= 4) return;
    // p is wavefront-uniform: the weights come through the scalar cache and
    // each lane selects its own, so no vector memory access is made here
    const double* d = p->del;
    const double delta3d =
        pick4(d[0], d[1], d[2], d[3], di) * pick4(d[4], d[5], d[6], d[7], dj) * pick4(d[8], d[9], d[10], d[11], dk);
    acc[0] += delta3d * p->F[0];
    acc[1] += delta3d * p->F[1];
    acc[2] += delta3d * p->F[2];
    return;
  }
  for (int i = 0; i < 4; ++i) {
    if (p->wc[i] != X) continue;
    for (int j = 0; j < 4; ++j) {
      if (p->wc[4 + j] != Y) continue;
      const double dxy = p->del[i] * p->del[4 + j];
      for (int k = 0; k < 4; ++k) {
        if (p->wc[8 + k] != Z) continue;
        const double delta3d = dxy * p->del[8 + k];
        acc[0] += delta3d * p->F[0];
        acc[1] += delta3d * p->F[1];
        acc[2] += delta3d * p->F[2];
      }
    }
  }
}

// add_point for a record held in LDS: in the simple case each lane reads the
// three weights it selects by index (one LDS read per axis) instead of all
// twelve -- the same factors in the same order, so the same bits.
__device__ __forceinline__ void add_point_lds(const SpreadPoint* __restrict__ p, int X, int Y, int Z, const Geom& g,
                                              bool simple, double (&acc)[3]) {
  if (!simple) {
    add_point(p, X, Y, Z, g, simple, acc);
    return;
  }
  int di = X - p->wc[0];
  di += di < 0 ? g.nx : 0;
  int dj = Y - p->wc[4];
  dj += dj < 0 ? g.ny : 0;
  int dk = Z - p->wc[8];
  dk += dk < 0 ? g.nz : 0;
  if ((di | dj | dk) >= 4) return;
  const double delta3d = p->del[di] * p->del[4 + dj] * p->del[8 + dk];
  acc[0] += delta3d * p->F[0];
  acc[1] += delta3d * p->F[1];
  acc[2] += delta3d * p->F[2];
}

// Force of the site whose segment is gs in per-site segments of ordered
// products (off: segment bounds, sorted: 3 doubles per entry): the products
// added in segment order from zero (loaded four entries ahead of the adds).
__device__ __forceinline__ void segment_force(const uint32_t* __restrict__ off, const double* __restrict__ sorted,
                                              uint32_t gs, double (&acc)[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  uint32_t e = off[gs];
  const uint32_t end = off[gs + 1];
  for (; e + 4 <= end; e += 4) {
    double v[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) v[q][a] = sorted[3 * size_t(e + q) + a];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) acc[a] += v[q][a];
  }
  for (; e < end; ++e) {
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] += sorted[3 * size_t(e) + a];
  }
}

// Force at this lane's site (X, Y, Z) from list `wl`, the list of the
// wavefront's 64 sites (ok = false for lanes past the lattice end).  The list
// index is wavefront-uniform, so ids and point records are read through the
// scalar cache; no LDS and no workgroup barrier.
//
// When nz is a multiple of 64 every list lies inside one z-row, so X and Y are
// the same on all lanes: each listed point then has one stencil row (i, j) on
// this list, found and weighted (dx_i * dy_j) once for the wavefront, and a
// lane only selects its z weight -- the same products in the same order as
// add_point, with a quarter of the vector instructions.
__device__ __forceinline__ void wave_force(const SpreadPoint* __restrict__ pts, const uint32_t* __restrict__ off,
                                           const uint32_t* __restrict__ sorted, uint32_t wl, bool ok, int X, int Y,
                                           int Z, const Geom& g, double (&acc)[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  wl = __builtin_amdgcn_readfirstlane(wl);
  const uint32_t o = off[wl], end = off[wl + 1];
  const bool simple = g.nx >= 4 && g.ny >= 4 && g.nz >= 4;
  uint32_t prev = 0xFFFFFFFFu;
  if (simple && g.nz % WAVE_SITES == 0) {
    // all 64 sites valid and in one row: X, Y uniform
    const int Xu = __builtin_amdgcn_readfirstlane(X), Yu = __builtin_amdgcn_readfirstlane(Y);
    for (uint32_t e = o; e < end; ++e) {
      const uint32_t id = sorted[e];
      if (id == prev) continue;
      prev = id;
      const SpreadPoint* __restrict__ p = pts + id;
      int di = Xu - p->wc[0];
      di += di < 0 ? g.nx : 0;
      int dj = Yu - p->wc[4];
      dj += dj < 0 ? g.ny : 0;
      if ((di | dj) >= 4) continue;  // not on this row (cannot happen for a listed point)
      const double dxy = p->del[di] * p->del[4 + dj];
      const double F0 = p->F[0], F1 = p->F[1], F2 = p->F[2];
      int dk = Z - p->wc[8];
      dk += dk < 0 ? g.nz : 0;
      // per-lane select of the z weight from the four uniform ones, with bit
      // masks so that the compiler keeps it branch-free
      uint64_t zbits = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        zbits |= __builtin_bit_cast(uint64_t, p->del[8 + q]) & (uint64_t(0) - uint64_t(dk == q));
      const double delta3d = dxy * __builtin_bit_cast(double, zbits);
      if (dk < 4) {
        acc[0] += delta3d * F0;
        acc[1] += delta3d * F1;
        acc[2] += delta3d * F2;
      }
    }
    return;
  }
  for (uint32_t e = o; e < end; ++e) {
    const uint32_t id = sorted[e];
    if (id == prev) continue;
    prev = id;
    if (ok) add_point(pts + id, X, Y, Z, g, simple, acc);
  }
}

// Force at this lane's site from every point first .. first+count-1, added
// in id order: exactly the sum of the site's sorted list (a point not in the
// list reaches no site of its run and adds nothing; a listed point is added
// once), without the list -- for dense small systems, whose lists would each
// hold a large share of all points.  For a whole workgroup (it holds
// barriers): chunks of 256 consecutive point records are copied to `st` with
// coalesced loads (one round trip per chunk); each lane then tests one record
// of every 64 against the x range (and, when the wavefront's sites lie in one
// x plane, the y range) of its wavefront's sites, and the wavefront adds, in
// id order, only the points that pass -- one that fails reaches none of its
// sites.
constexpr int RANGE_STAGE = 256;
struct RangeStage {
  SpreadPoint pt[RANGE_STAGE];
};

__device__ __forceinline__ void range_force_staged(const SpreadPoint* __restrict__ pts, uint32_t first,
                                                   uint32_t count, bool ok, int X, int Y, int Z, const Geom& g,
                                                   RangeStage* __restrict__ st, double (&acc)[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  first = __builtin_amdgcn_readfirstlane(first);
  count = __builtin_amdgcn_readfirstlane(count);
  const bool simple = g.nx >= 4 && g.ny >= 4 && g.nz >= 4;
  const int lane = int(threadIdx.x % WAVE_SITES);
  // the wavefront's sites are consecutive, so its valid lanes are a prefix
  // and their x (and, within a plane, y) do not decrease
  const uint64_t okm = __ballot(ok);
  int Xlo = 0, Xhi = -1, Ylo = 0, Yhi = -1;
  if (okm) {
    const int last = 63 - __builtin_clzll(okm);
    Xlo = __builtin_amdgcn_readfirstlane(X);
    Xhi = __builtin_amdgcn_readlane(X, last);
    Ylo = __builtin_amdgcn_readfirstlane(Y);
    Yhi = __builtin_amdgcn_readlane(Y, last);
  }
  const bool one_plane = Xlo == Xhi;
  for (uint32_t base = first; base < first + count; base += RANGE_STAGE) {
    const uint32_t m = min(uint32_t(RANGE_STAGE), first + count - base);
    __syncthreads();  // the previous chunk is consumed
    uint64_t* __restrict__ dst = reinterpret_cast<uint64_t*>(st->pt);
    const uint64_t* __restrict__ src = reinterpret_cast<const uint64_t*>(pts + base);
    for (uint32_t w = threadIdx.x; w < m * uint32_t(sizeof(SpreadPoint) / 8); w += blockDim.x) dst[w] = src[w];
    __syncthreads();
    for (uint32_t sub = 0; sub < m; sub += WAVE_SITES) {
      bool cand = false;
      if (sub + uint32_t(lane) < m && okm) {
        const SpreadPoint& p = st->pt[sub + lane];
        bool hx = false, hy = !one_plane;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          hx = hx || (p.wc[k] >= Xlo && p.wc[k] <= Xhi);
          if (one_plane) hy = hy || (p.wc[4 + k] >= Ylo && p.wc[4 + k] <= Yhi);
        }
        cand = hx && hy;
      }
      uint64_t mask = __ballot(cand);
      while (mask) {
        const int k = __builtin_ctzll(mask);
        mask &= mask - 1;
        if (ok) add_point_lds(&st->pt[sub + k], X, Y, Z, g, simple, acc);
      }
    }
  }
}

// wave_force for a lane whose list differs from the other lanes' (`wl` per
// lane, vector loads): the sum of add_point over the list, the same products
// in the same order as wave_force gives for that site.
__device__ __forceinline__ void lane_force(const SpreadPoint* __restrict__ pts, const uint32_t* __restrict__ off,
                                           const uint32_t* __restrict__ sorted, uint32_t wl, int X, int Y, int Z,
                                           const Geom& g, double (&acc)[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  const uint32_t o = off[wl], end = off[wl + 1];
  const bool simple = g.nx >= 4 && g.ny >= 4 && g.nz >= 4;
  uint32_t prev = 0xFFFFFFFFu;
  for (uint32_t e = o; e < end; ++e) {
    const uint32_t id = sorted[e];
    if (id == prev) continue;
    prev = id;
    add_point(pts + id, X, Y, Z, g, simple, acc);
  }
}

// One list entry spread over four lanes: lane 4e + q holds entry e (of the
// first 16) with its z weight q, from the entry's RunEntry.
struct QuadEntry {
  double dxy;  // row weight dx_i dy_j at this run
  double zq;   // z weight q
  double F[3];
  int wz;      // z of the first stencil site, -1: skipped entry
};

__device__ __forceinline__ void load_quad_entry(const RunEntry* __restrict__ rec, const int* __restrict__ wz,
                                                uint32_t k, QuadEntry& r) {
  const RunEntry* __restrict__ p = rec + k;
  r.dxy = p->dxy;
  r.zq = p->zw[threadIdx.x % 4];
#pragma unroll
  for (int a = 0; a < 3; ++a) r.F[a] = p->F[a];
  r.wz = wz[k];
}

// The row walk of wave_force (simple lattice, nz % 64 == 0: X and Y uniform)
// over list entries o .. o+m-1 whose first min(m, 16) entries are held four
// lanes each (QuadEntry): lane 4e + q forms entry e's contribution to the
// site at z offset q once (delta3d = dx_i dy_j dz_q, times F), and every site
// then takes, entry by entry in list order, the contribution of the lane
// whose q is its own offset.  Entries past the 16th are read from memory.
// Same products, same order, same sum as wave_force.
__device__ __forceinline__ void quad_row_force(const QuadEntry& r, uint32_t o, uint32_t m,
                                               const RunEntry* __restrict__ rec, const int* __restrict__ wz, int Z,
                                               const Geom& g, double (&acc)[3]) {
  constexpr uint32_t QUADS = WAVE_SITES / 4;
  acc[0] = acc[1] = acc[2] = 0.0;
  const int lane = int(threadIdx.x % WAVE_SITES), e = lane >> 2, q = lane & 3;
  const bool use = uint32_t(e) < m && r.wz >= 0;
  const double delta3d = r.dxy * r.zq;
  const double c0 = delta3d * r.F[0], c1 = delta3d * r.F[1], c2 = delta3d * r.F[2];
  uint64_t mask = __ballot(use && q == 0);
  while (mask) {
    const int L = __builtin_ctzll(mask);  // lane 4e of the next entry
    mask &= mask - 1;
    int dk = Z - __builtin_amdgcn_readlane(r.wz, L);
    dk += dk < 0 ? g.nz : 0;
    const int src = L + (dk & 3);
    const double t0 = __shfl(c0, src, WAVE_SITES), t1 = __shfl(c1, src, WAVE_SITES),
                 t2 = __shfl(c2, src, WAVE_SITES);
    if (dk < 4) {
      acc[0] += t0;
      acc[1] += t1;
      acc[2] += t2;
    }
  }
  if (m > QUADS) {  // rare: the rest of a long list, one entry at a time
    for (uint32_t k = o + QUADS; k < o + m; ++k) {
      const int w = wz[k];
      if (w < 0) continue;
      const RunEntry* __restrict__ p = rec + k;
      int dk = Z - w;
      dk += dk < 0 ? g.nz : 0;
      uint64_t zbits = 0;
#pragma unroll
      for (int qq = 0; qq < 4; ++qq)
        zbits |= __builtin_bit_cast(uint64_t, p->zw[qq]) & (uint64_t(0) - uint64_t(dk == qq));
      const double d3 = p->dxy * __builtin_bit_cast(double, zbits);
      if (dk < 4) {
        acc[0] += d3 * p->F[0];
        acc[1] += d3 * p->F[1];
        acc[2] += d3 * p->F[2];
      }
    }
  }
}

// LDS staging of one chunk of a list, one per wavefront.
constexpr int STAGE_POINTS = 32;
constexpr int POINT_WORDS = int(sizeof(SpreadPoint) / 8);
struct WaveStage {
  SpreadPoint pt[STAGE_POINTS];
  uint32_t id[STAGE_POINTS];
};

// wave_force for long lists: the wavefront copies chunks of 32 point records
// into its LDS slice with all lanes, then every lane walks the chunk from LDS,
// so an entry costs a few LDS reads instead of a dependent round trip to the
// caches.  Same order, same result.
__device__ __forceinline__ void wave_force_staged(const SpreadPoint* __restrict__ pts,
                                                  const uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ sorted, uint32_t wl, bool ok, int X,
                                                  int Y, int Z, const Geom& g, WaveStage* __restrict__ st,
                                                  double (&acc)[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
  wl = __builtin_amdgcn_readfirstlane(wl);
  const uint32_t o = off[wl], end = off[wl + 1];
  const bool simple = g.nx >= 4 && g.ny >= 4 && g.nz >= 4;
  const int lane = int(threadIdx.x % WAVE_SITES);
  uint32_t prev = 0xFFFFFFFFu;
  for (uint32_t base = o; base < end; base += STAGE_POINTS) {
    const int m = int(min(uint32_t(STAGE_POINTS), end - base));
    if (lane < m) st->id[lane] = sorted[base + lane];
    // the wavefront's LDS accesses complete in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ids from LDS: the record loads of a lane are independent of each other
    uint64_t* dst = reinterpret_cast<uint64_t*>(st->pt);
    for (int w = lane; w < m * POINT_WORDS; w += WAVE_SITES) {
      const int e = w / POINT_WORDS, c = w - e * POINT_WORDS;
      dst[w] = reinterpret_cast<const uint64_t*>(pts + st->id[e])[c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int e = 0; e < m; ++e) {
      const uint32_t id = st->id[e];
      if (id == prev) continue;
      prev = id;
      if (ok) add_point_lds(&st->pt[e], X, Y, Z, g, simple, acc);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

}  // namespace lbs
