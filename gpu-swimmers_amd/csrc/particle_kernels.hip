// Particle kernels for gfx950: Peskin 4-point delta, velocity interpolation,
// singular-force spreading, swimmer and tracer moves.
//
// Particle state is FP64 in both storage forms; u and force are read/written
// in the lattice's storage type T.  The arithmetic of the reference is kept
// operation for operation (delta.cu, interp.cu, SwimmerArray.cu,
// TracerArray.cu), so with FP64 storage a swimmer or tracer moves exactly as a
// CPU restatement does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "d3q15.hpp"
#include "internal.hpp"
#include "spread.hpp"
#include "xcd.hpp"

namespace lbs {
namespace {

constexpr int PBLOCK = 256;

// Peskin (2002) four-point regularised delta, delta.cu:6-25.
__device__ __forceinline__ double peskin(double x) {
  const double abs_x = fabs(x);
  double root = -4. * x * x;
  double phi = -2. * abs_x;
  if (abs_x >= 2.0) return 0.;
  if (abs_x >= 1.0) {
    root += 12. * abs_x - 7.;
    phi += 5.;
    phi -= sqrt(root);
  } else {
    root += 4. * abs_x + 1;
    phi += 3.;
    phi += sqrt(root);
  }
  return 0.125 * phi;
}

// ((int)x + n) % n of interp.cu:26 for x >= -n, extended to a true periodic
// wrap (and NaN mapped to 0) so that no input can index outside the lattice.
// Coordinates already inside [0, n) skip the division.
__device__ __forceinline__ uint32_t wrap_index(double x, int n) {
  if (!(x > -2147483648.0 && x < 2147483647.0)) x = 0.0;
  const int v0 = int(x);
  if (uint32_t(v0) < uint32_t(n)) return uint32_t(v0);
  const int v = v0 % n;
  return uint32_t(v < 0 ? v + n : v);
}

// fmod(x, y) of SwimmerArray.cu:161 / TracerArray.cu:33 for y > 0: when
// 0 <= x < 2y the result is x or x - y, both exact (Sterbenz), so the
// division is needed only outside that range.
__device__ __forceinline__ double wrap_position(double x, double y) {
  if (x >= 0.0 && x < y) return x;
  if (x >= y && x < 2.0 * y) return x - y;
  return fmod(x, y);
}

// InterpVelocity, interp.cu:8-47: v_d = sum_{i,j,k} dx_i dy_j dz_k u_d(site),
// summed in i, j, k order.
template <typename T>
__device__ __forceinline__ void interp(const T* __restrict__ u, const Geom& g, const double (&r)[3],
                                       double (&v)[3]) {
  const int size[3] = {g.nx, g.ny, g.nz};
  uint32_t idx[3][4];
  double del[3][4];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double x0 = ceil(r[d] - 2.);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double x = x0 + i;
      idx[d][i] = wrap_index(x, size[d]);
      del[d][i] = peskin(r[d] - x);
    }
  }
  const size_t n = g.n;
  v[0] = v[1] = v[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double dxy = del[0][i] * del[1][j];
      const uint32_t row = idx[0][i] * g.nyz + idx[1][j] * uint32_t(g.nz);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double delta3d = dxy * del[2][k];
        const uint32_t ijk = row + idx[2][k];
        v[0] += delta3d * double(u[ijk]);
        v[1] += delta3d * double(u[n + ijk]);
        v[2] += delta3d * double(u[2 * n + ijk]);
      }
    }
  }
}

// ---- stencil gathers by lane groups -----------------------------------------
// A group of 16 lanes serves one particle (group_interp below): lane c = 4j + k
// owns stencil column (j, k), so in each load instruction the group reads four
// runs of four consecutive z sites instead of 16 scattered words.
constexpr int GROUP = 16;
constexpr int GROUPS = PBLOCK / GROUP;

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----
struct U4 {
  uint32_t x[4];
};

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x[0], hi0 = __umulhi(0xD2511F53u, c.x[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c.x[2], hi1 = __umulhi(0xCD9E8D57u, c.x[2]);
    U4 o;
    o.x[0] = hi1 ^ c.x[1] ^ k0;
    o.x[1] = lo1;
    o.x[2] = hi0 ^ c.x[3] ^ k1;
    o.x[3] = lo0;
    c = o;
  }
  return c;
}

// Block `tag` of swimmer `is` at draw counter `ctr`, keyed by the seed.
__device__ __forceinline__ U4 draw(unsigned long long seed, int is, uint64_t ctr, uint32_t tag) {
  U4 c;
  c.x[0] = uint32_t(is);
  c.x[1] = uint32_t(ctr);
  c.x[2] = uint32_t(ctr >> 32);
  c.x[3] = tag;
  return philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
}

// float in (0,1] from the top 24 bits (the range of curand_uniform).
__device__ __forceinline__ float u01f(uint32_t x) { return float((x >> 8) + 1u) * 0x1p-24f; }
// double in (0,1) from 53 bits.
__device__ __forceinline__ double u01d(uint32_t a, uint32_t b) {
  uint64_t m = (uint64_t(a) << 21) ^ uint64_t(b >> 11);
  m &= (uint64_t(1) << 53) - 1;
  return (double(m) + 0.5) * 0x1p-53;
}

// Box-Muller needs ln(u) and cos/sin(2 pi u).  Library versions of these
// differ between the device and a host C library in the last bit, so the
// tumble orientations would differ between a CPU restatement and this code;
// these are written out from series with +, -, *, / and sqrt only (all
// correctly rounded on both), so they give the same bits everywhere.  Accuracy
// about 2e-16 absolute, checked against numpy by the oracle's tests.

// ln(x) for a normal x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)), and
// ln(m) = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.172, by its Taylor series to
// s^23 (the next term is below 1e-19).
__device__ __forceinline__ double bm_log(double x) {
  const uint64_t bits = __builtin_bit_cast(uint64_t, x);
  int e = int((bits >> 52) & 0x7FF) - 1022;
  double m = __builtin_bit_cast(double, (bits & 0x000FFFFFFFFFFFFFull) | (uint64_t(1022) << 52));  // [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  return double(e) * 0.69314718055994530942 + 2.0 * s * p;
}

// cos and sin of 2 pi u for u in [0, 1): 4u = q + f (both exact), the angle
// q pi/2 + f pi/2 is reduced to theta = f pi/2 or (1-f) pi/2 in [0, pi/4],
// whose sine and cosine come from their Taylor series to theta^19 and
// theta^18, then rotated by the quadrant q.
__device__ __forceinline__ void bm_cossin_2pi(double u, double& c, double& s) {
  const double t = 4.0 * u;
  const double q = floor(t);
  double f = t - q;
  const bool swap = f > 0.5;
  if (swap) f = 1.0 - f;
  const double th = f * 1.57079632679489661923;
  const double t2 = th * th;
  double ps = -1.0 / 121645100408832000.0;  // -1/19!
  ps = ps * t2 + 1.0 / 355687428096000.0;    // 1/17!
  ps = ps * t2 - 1.0 / 1307674368000.0;      // 1/15!
  ps = ps * t2 + 1.0 / 6227020800.0;         // 1/13!
  ps = ps * t2 - 1.0 / 39916800.0;           // 1/11!
  ps = ps * t2 + 1.0 / 362880.0;             // 1/9!
  ps = ps * t2 - 1.0 / 5040.0;               // 1/7!
  ps = ps * t2 + 1.0 / 120.0;                // 1/5!
  ps = ps * t2 - 1.0 / 6.0;                  // 1/3!
  ps = ps * t2 + 1.0;
  double pc = -1.0 / 6402373705728000.0;     // -1/18!
  pc = pc * t2 + 1.0 / 20922789888000.0;     // 1/16!
  pc = pc * t2 - 1.0 / 87178291200.0;        // 1/14!
  pc = pc * t2 + 1.0 / 479001600.0;          // 1/12!
  pc = pc * t2 - 1.0 / 3628800.0;            // 1/10!
  pc = pc * t2 + 1.0 / 40320.0;              // 1/8!
  pc = pc * t2 - 1.0 / 720.0;                // 1/6!
  pc = pc * t2 + 1.0 / 24.0;                 // 1/4!
  pc = pc * t2 - 1.0 / 2.0;                  // 1/2!
  pc = pc * t2 + 1.0;
  double sv = th * ps, cv = pc;
  if (swap) {
    const double tmp = sv;
    sv = cv;
    cv = tmp;
  }
  const int quad = int(q);
  c = quad == 0 ? cv : (quad == 1 ? -sv : (quad == 2 ? -cv : sv));
  s = quad == 0 ? sv : (quad == 1 ? cv : (quad == 2 ? -sv : -cv));
}

// Three N(0,1) normals by Box-Muller from draw blocks 1 and 2.
__device__ __forceinline__ void normals3(unsigned long long seed, int is, uint64_t ctr, double (&g)[3]) {
  const U4 b1 = draw(seed, is, ctr, 1u), b2 = draw(seed, is, ctr, 2u);
  const double u1 = u01d(b1.x[0], b1.x[1]), u2 = u01d(b1.x[2], b1.x[3]);
  const double u3 = u01d(b2.x[0], b2.x[1]), u4 = u01d(b2.x[2], b2.x[3]);
  double c, sn;
  double rad = sqrt(-2.0 * bm_log(u1));
  bm_cossin_2pi(u2, c, sn);
  g[0] = rad * c;
  g[1] = rad * sn;
  rad = sqrt(-2.0 * bm_log(u3));
  bm_cossin_2pi(u4, c, sn);
  g[2] = rad * c;
}

// ---- force spreading -------------------------------------------------------
// DoSwimmerArrayAddForces + AccumulateDeltaForce (SwimmerArray.cu:42-112).
// One wavefront per swimmer.  The tail stencil (at r - l n, force -P n) and
// the head stencil (at r, force +P n) usually share their 4x4x4 box; each
// lane takes one site of the union box and adds tail + head with one atomic
// per component, so a site seen only by this swimmer gets 0 + t + h exactly
// as the reference's two atomics give.  Boxes wider than 5 fall back to two
// separate 4x4x4 passes.
template <typename T>
__device__ __forceinline__ void spread_atomic(T* addr, double val) {
  unsafeAtomicAdd(addr, T(val));
}

template <typename T>
__global__ __launch_bounds__(PBLOCK) void k_swimmers_add_forces(int ns, double P, double l,
                                                               const double* __restrict__ r,
                                                               const double* __restrict__ nv,
                                                               T* __restrict__ force, Geom g) {
  const int lane = threadIdx.x & 63;
  const int is = int((blockIdx.x * PBLOCK + threadIdx.x) >> 6);
  if (is >= ns) return;
  const size_t n = g.n;
  const int size[3] = {g.nx, g.ny, g.nz};
  double rt[3], rh[3], Ft[3], Fh[3], x0t[3], x0h[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double ra = r[size_t(ns) * a + is], na = nv[size_t(ns) * a + is];
    rt[a] = ra - na * l;
    Ft[a] = -P * na;
    rh[a] = ra;
    Fh[a] = Ft[a] * -1.0;
    x0t[a] = ceil(rt[a] - 2.);
    x0h[a] = ceil(rh[a] - 2.);
  }
  double lo[3];
  int ext[3];
  bool fused = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = fmin(x0t[a], x0h[a]);
    const double span = fmax(x0t[a], x0h[a]) - lo[a];
    fused = fused && (span <= 1.0);
    ext[a] = span <= 1.0 ? int(span) + 4 : 4;
  }
  if (fused) {
    const int vol = ext[0] * ext[1] * ext[2];
    for (int c = lane; c < vol; c += 64) {
      const int cz = c % ext[2], t = c / ext[2];
      const int cy = t % ext[1], cx = t / ext[1];
      const double X = lo[0] + cx, Y = lo[1] + cy, Z = lo[2] + cz;
      const double wt = peskin(rt[0] - X) * peskin(rt[1] - Y) * peskin(rt[2] - Z);
      const double wh = peskin(rh[0] - X) * peskin(rh[1] - Y) * peskin(rh[2] - Z);
      if (wt == 0.0 && wh == 0.0) continue;
      const uint32_t ijk = wrap_index(X, size[0]) * g.nyz + wrap_index(Y, size[1]) * uint32_t(g.nz) +
                           wrap_index(Z, size[2]);
#pragma unroll
      for (int d = 0; d < 3; ++d) spread_atomic(&force[size_t(d) * n + ijk], wt * Ft[d] + wh * Fh[d]);
    }
  } else {
    const int ci = lane >> 4, cj = (lane >> 2) & 3, ck = lane & 3;
    for (int end = 0; end < 2; ++end) {
      const double* rp = end ? rh : rt;
      const double* x0 = end ? x0h : x0t;
      const double* Fp = end ? Fh : Ft;
      const double X = x0[0] + ci, Y = x0[1] + cj, Z = x0[2] + ck;
      const double w = peskin(rp[0] - X) * peskin(rp[1] - Y) * peskin(rp[2] - Z);
      const uint32_t ijk = wrap_index(X, size[0]) * g.nyz + wrap_index(Y, size[1]) * uint32_t(g.nz) +
                           wrap_index(Z, size[2]);
#pragma unroll
      for (int d = 0; d < 3; ++d) spread_atomic(&force[size_t(d) * n + ijk], w * Fp[d]);
    }
  }
}

// InterpVelocity (interp.cu:8-47) by a group of 16 lanes, the stencil
// streamed plane by plane: lane q < 12 forms stencil coordinate q (axis q / 4,
// offset q % 4) once and the lanes read the ones they need from each other;
// lane c = 4j + k gathers column (j, k) of every x plane i; for each plane the
// 16 lanes write their products to the group's LDS slice, SUM_PLANES stencil
// planes i at a time, component-major (slice[(3p + a) 16 + c]), and lanes
// 0..2 add the 16 products of their component of each plane in t = 16i + 4j
// + k order -- the i, j, k order of the reference -- reading them 16 bytes at
// a time; then the next planes reuse the slice.  A group's lanes are in one
// wavefront and LDS operations of a wavefront complete in order, so no
// workgroup barrier is needed.  Returns component c on lanes c < 3 (0 on the
// others).  Every lane of the group must call it.
constexpr int SUM_PLANES = 1;
// products of SUM_PLANES planes + 2: 16-byte aligned, groups start on different banks
constexpr int PLANE_STRIDE = 48 * SUM_PLANES + 2;

// One lane's share of a stencil: its column's weights and the gathered u.
struct GroupGather {
  double dx[4], dy, dz;
  double uv[4][3];
};

// AOS: u holds the three components of a site together (u[3 ijk + a], the
// copy the fused step writes for the moves) instead of u[a n + ijk].
template <typename T, bool AOS = false>
__device__ __forceinline__ GroupGather group_gather(const T* __restrict__ u, const Geom& g, const double (&r)[3],
                                                    int c) {
  GroupGather G;
  const int q = c < 12 ? c : 0;
  const int axis = q >> 2;
  const double ra = axis == 0 ? r[0] : (axis == 1 ? r[1] : r[2]);
  const int na = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  const double x = ceil(ra - 2.) + (q & 3);
  const int mine_idx = int(wrap_index(x, na));
  const double mine_del = peskin(ra - x);
  const int j = c >> 2, k = c & 3;
  uint32_t ix[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ix[i] = uint32_t(__shfl(mine_idx, i, GROUP));
    G.dx[i] = __shfl(mine_del, i, GROUP);
  }
  const uint32_t iyz =
      uint32_t(__shfl(mine_idx, 4 + j, GROUP)) * uint32_t(g.nz) + uint32_t(__shfl(mine_idx, 8 + k, GROUP));
  G.dy = __shfl(mine_del, 4 + j, GROUP);
  G.dz = __shfl(mine_del, 8 + k, GROUP);
  const size_t n = g.n;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t ijk = ix[i] * g.nyz + iyz;
    if constexpr (AOS) {
      const T* __restrict__ us = u + 3 * size_t(ijk);
      G.uv[i][0] = double(us[0]);
      G.uv[i][1] = double(us[1]);
      G.uv[i][2] = double(us[2]);
    } else {
      G.uv[i][0] = double(u[ijk]);
      G.uv[i][1] = double(u[n + ijk]);
      G.uv[i][2] = double(u[2 * n + ijk]);
    }
  }
  return G;
}

__device__ __forceinline__ double group_sum(const GroupGather& G, int c, double* __restrict__ terms) {
  const int d = c < 3 ? c : 0;
  double vel = 0.0;
#pragma unroll
  for (int i0 = 0; i0 < 4; i0 += SUM_PLANES) {
#pragma unroll
    for (int p = 0; p < SUM_PLANES; ++p) {
      const double delta3d = G.dx[i0 + p] * G.dy * G.dz;
#pragma unroll
      for (int a = 0; a < 3; ++a) terms[(3 * p + a) * GROUP + c] = delta3d * G.uv[i0 + p][a];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (c < 3) {
#pragma unroll
      for (int p = 0; p < SUM_PLANES; ++p) {
        const double2* __restrict__ row = reinterpret_cast<const double2*>(terms + (3 * p + d) * GROUP);
#pragma unroll
        for (int h = 0; h < GROUP / 2; ++h) {
          const double2 two = row[h];
          vel += two.x;
          vel += two.y;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return vel;
}

template <typename T, bool AOS = false>
__device__ __forceinline__ double group_interp(const T* __restrict__ u, const Geom& g, const double (&r)[3], int c,
                                               double* __restrict__ terms) {
  return group_sum(group_gather<T, AOS>(u, g, r, c), c, terms);
}

// ZeroForce + AddForces of the swimmer at R with orientation N, added into
// `force` with hardware atomics (the reference's own scatter,
// SwimmerArray.cu:42-112: tail at R - l N with force -P N, head at R with
// +P N; weights peskin(x) peskin(y) peskin(z)), by the GROUP lanes of the
// swimmer: lane q < 12 forms stencil coordinate q of each end, lane c =
// 4j + k adds column (j, k) of the four x planes.  Where both ends share
// their stencil box (|l n| below one site: the usual case) one atomic per
// site and component adds tail + head.  Atomics make the order of the
// additions, and so the rounding of the sums, vary from run to run.
template <typename T>
__device__ __forceinline__ void scatter_dipole(const double (&R)[3], const double (&N)[3], double P, double l,
                                               const Geom& g, int c, T* __restrict__ force) {
  double rp[2][3], F[2][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rp[0][a] = R[a] - N[a] * l;
    F[0][a] = -P * N[a];
    rp[1][a] = R[a];
    F[1][a] = F[0][a] * -1.0;
  }
  const int q = c < 12 ? c : 0;
  const int axis = q >> 2;
  const int na = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
  int idx[2];
  double del[2];
  bool same = true;
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const double ra = axis == 0 ? rp[e][0] : (axis == 1 ? rp[e][1] : rp[e][2]);
    const double b = ceil(ra - 2.);
    const double x = b + (q & 3);
    idx[e] = int(wrap_index(x, na));
    del[e] = peskin(ra - x);
    if (e == 1) {
      const double b0 = ceil((axis == 0 ? rp[0][0] : (axis == 1 ? rp[0][1] : rp[0][2])) - 2.);
      same = b0 == b;
    }
  }
  // uniform over the group: every axis of the two ends has the same base
  same = __shfl(same ? 1 : 0, 0, GROUP) && __shfl(same ? 1 : 0, 4, GROUP) && __shfl(same ? 1 : 0, 8, GROUP);
  const int j = c >> 2, k = c & 3;
  const size_t n = g.n;
  double w[2][4];
  uint32_t ijk[2][4];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const double dy = __shfl(del[e], 4 + j, GROUP), dz = __shfl(del[e], 8 + k, GROUP);
    const uint32_t iyz =
        uint32_t(__shfl(idx[e], 4 + j, GROUP)) * uint32_t(g.nz) + uint32_t(__shfl(idx[e], 8 + k, GROUP));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      w[e][i] = __shfl(del[e], i, GROUP) * dy * dz;
      ijk[e][i] = uint32_t(__shfl(idx[e], i, GROUP)) * g.nyz + iyz;
    }
  }
  if (same) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (w[0][i] == 0.0 && w[1][i] == 0.0) continue;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        spread_atomic(&force[size_t(d) * n + ijk[0][i]], w[0][i] * F[0][d] + w[1][i] * F[1][d]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < 2; ++e)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (w[e][i] == 0.0) continue;
#pragma unroll
        for (int d = 0; d < 3; ++d) spread_atomic(&force[size_t(d) * n + ijk[e][i]], w[e][i] * F[e][d]);
      }
  }
}

// DoSwimmerArrayMove, SwimmerArray.cu:123-210 (norm starts at 0; with
// rotation off and no tumble, n is kept).  A lane group per swimmer
// interpolates at r and at r - l n; lane 0 of the group finishes the move.
// SCATTER: the group then adds the moved swimmer's dipole into `fnext` (the
// next step's ZeroForce + AddForces, scatter_dipole).
template <typename T, bool AOS = false, bool SCATTER = false>
__device__ __forceinline__ void swimmers_move_body(int ns, const SwimParams& cp, double* __restrict__ r,
                                                   double* __restrict__ v, double* __restrict__ nv,
                                                   uint64_t* __restrict__ prng, const int* __restrict__ perm,
                                                   const T* __restrict__ u, const Geom& g,
                                                   unsigned long long* __restrict__ tumbles, uint32_t blk,
                                                   double* __restrict__ lds, T* __restrict__ fnext = nullptr) {
  const int grp = int(threadIdx.x) / GROUP, c = int(threadIdx.x) % GROUP;
  const int is = int(blk) * GROUPS + grp;
  if (is >= ns) return;  // whole group: the lanes of a group share `is`
  double* terms = lds + grp * PLANE_STRIDE;
  double R[3], N[3], rMinus[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    R[d] = r[size_t(ns) * d + is];
    N[d] = nv[size_t(ns) * d + is];
    rMinus[d] = R[d] - N[d] * cp.l;
  }
  // both stencils' gathers in flight together, then the two ordered sums
  const GroupGather GA = group_gather<T, AOS>(u, g, R, c);
  const GroupGather GB = group_gather<T, AOS>(u, g, rMinus, c);
  const double a = group_sum(GA, c, terms);
  const double b = group_sum(GB, c, terms);
  double vel[3], vMinus[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    vel[d] = __shfl(a, d, GROUP);
    vMinus[d] = __shfl(b, d, GROUP);
  }
  double Rout[3] = {0.0, 0.0, 0.0}, Nout[3] = {0.0, 0.0, 0.0};  // lane 0: the moved swimmer's r and n
  if (c == 0) {
    const int size[3] = {g.nx, g.ny, g.nz};
    const int user = perm[is];  // the random stream belongs to the swimmer, not its storage slot
    double rDot[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rDot[d] = cp.mobility * cp.P * N[d];
      if (!cp.translational_advection_off) rDot[d] += vel[d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[size_t(ns) * d + is] = rDot[d];
      const double Rn = R[d] + rDot[d];
      Rout[d] = wrap_position(Rn + size[d], double(size[d]));
      r[size_t(ns) * d + is] = Rout[d];
    }
    const uint64_t ctr = prng[is];
    const float U = u01f(draw(cp.seed, user, ctr, 0u).x[0]);
    double newn[3];
    double norm = 0.0;
    bool turned = true;
    if (double(U) < cp.alpha) {
      normals3(cp.seed, user, ctr, newn);
#pragma unroll
      for (int d = 0; d < 3; ++d) norm += newn[d] * newn[d];
      atomicAdd(tumbles, 1ull);
    } else if (!cp.rotational_advection_off) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double nDot = vel[d] - vMinus[d];
        nDot /= cp.l;
        newn[d] = N[d] + nDot;
        norm += newn[d] * newn[d];
      }
    } else {
      turned = false;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) Nout[d] = N[d];
    if (turned) {
      norm = sqrt(norm);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        Nout[d] = newn[d] / norm;
        nv[size_t(ns) * d + is] = Nout[d];
      }
    }
    prng[is] = ctr + 1;
  }
  if constexpr (SCATTER) {
    double Rb[3], Nb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      Rb[d] = __shfl(Rout[d], 0, GROUP);
      Nb[d] = __shfl(Nout[d], 0, GROUP);
    }
    scatter_dipole(Rb, Nb, cp.P, cp.l, g, c, fnext);
  }
}

// ---- stencil gathers by pairs of columns (tracer moves) ----------------------
// A group of 8 lanes serves one tracer: lane c = 2j + h owns the stencil
// columns (j, 2h) and (j, 2h + 1), whose products are consecutive in the
// reference's i, j, k order, so each lane writes them to the plane's LDS
// slice together, and a wavefront sums 8 tracers' stencils at once.
constexpr int PAIR = 8;

struct PairGather {
  double dx[4], dy, dz[2];
  double uv[4][2][3];
};

// The weights and wrapped coordinates are those of group_gather: lane c
// computes the x (c < 4) or y weight of offset c & 3, and every lane the z
// weight of offset c & 3, with the same expressions.
template <typename T, bool AOS>
__device__ __forceinline__ PairGather pair_gather(const T* __restrict__ u, const Geom& g, const double (&r)[3],
                                                  int c) {
  PairGather G;
  const int axis = c >> 2;
  const double ra = axis == 0 ? r[0] : r[1];
  const int na = axis == 0 ? g.nx : g.ny;
  const double x = ceil(ra - 2.) + (c & 3);
  const int mine_idx = int(wrap_index(x, na));
  const double mine_del = peskin(ra - x);
  const double zq = ceil(r[2] - 2.) + (c & 3);
  const int z_idx = int(wrap_index(zq, g.nz));
  const double z_del = peskin(r[2] - zq);
  const int j = c >> 1, h = c & 1;
  uint32_t ix[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ix[i] = uint32_t(__shfl(mine_idx, i, PAIR));
    G.dx[i] = __shfl(mine_del, i, PAIR);
  }
  const uint32_t iy = uint32_t(__shfl(mine_idx, 4 + j, PAIR)) * uint32_t(g.nz);
  G.dy = __shfl(mine_del, 4 + j, PAIR);
  uint32_t iz[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    iz[e] = uint32_t(__shfl(z_idx, 2 * h + e, PAIR));
    G.dz[e] = __shfl(z_del, 2 * h + e, PAIR);
  }
  const size_t n = g.n;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const uint32_t ijk = ix[i] * g.nyz + iy + iz[e];
      if constexpr (AOS) {
        const T* __restrict__ us = u + 3 * size_t(ijk);
#pragma unroll
        for (int a = 0; a < 3; ++a) G.uv[i][e][a] = double(us[a]);
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) G.uv[i][e][a] = double(u[a * n + ijk]);
      }
    }
  }
  return G;
}

// group_sum for a pair gather: per plane i the lane's two products of each
// component go to slice[a 16 + 2c + e] in one 16-byte write, and lanes 0..2
// add the 16 products of their component in column order.  Same products,
// same order.
__device__ __forceinline__ double pair_sum(const PairGather& G, int c, double* __restrict__ terms) {
  const int d = c < 3 ? c : 0;
  double vel = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double delta0 = G.dx[i] * G.dy * G.dz[0];
    const double delta1 = G.dx[i] * G.dy * G.dz[1];
#pragma unroll
    for (int a = 0; a < 3; ++a)
      *reinterpret_cast<double2*>(terms + a * GROUP + 2 * c) = double2{delta0 * G.uv[i][0][a], delta1 * G.uv[i][1][a]};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (c < 3) {
      const double2* __restrict__ row = reinterpret_cast<const double2*>(terms + d * GROUP);
#pragma unroll
      for (int hh = 0; hh < GROUP / 2; ++hh) {
        const double2 two = row[hh];
        vel += two.x;
        vel += two.y;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  return vel;
}

// DoTracerArrayMove, TracerArray.cu:12-43 (s updated per particle), a lane
// group of 8 per tracer (pair_gather, pair_sum).
constexpr int TGROUPS = TRACER_GROUPS;  // tracers per workgroup: 512 threads, contiguous in storage order
constexpr int TBLOCK = TGROUPS * PAIR;
// Tracers per workgroup of the single-array move: 256 threads, one wavefront
// per SIMD, so that at 86 VGPRs five workgroups fill a CU's 5 wavefronts per
// SIMD (512-thread workgroups leave one slot per SIMD empty).  At 256^3 with
// 1e6 tracers 2 % faster than 64 tracers; 16 and 8 are slower beside the
// list building.
constexpr int TMOVE_GROUPS = 32;
static_assert(SUM_PLANES == 1, "pair_sum stages one plane at a time");

template <typename T, bool AOS = false, int TG = TGROUPS>
__device__ __forceinline__ void tracers_move_body(int nt, double* __restrict__ r, double* __restrict__ sd,
                                                  double* __restrict__ v, const T* __restrict__ u, const Geom& g,
                                                  uint32_t blk, double* __restrict__ lds) {
  const int grp = int(threadIdx.x) / PAIR, c = int(threadIdx.x) % PAIR;
  const int it = int(blk) * TG + grp;
  if (it >= nt) return;  // whole group: the lanes of a group share `it`
  double* terms = lds + grp * PLANE_STRIDE;
  const double R[3] = {r[it], r[size_t(nt) + it], r[2 * size_t(nt) + it]};
  const int d = c < 3 ? c : 0;
  const double s_old = c < 3 ? sd[size_t(nt) * d + it] : 0.0;
  const double vel = pair_sum(pair_gather<T, AOS>(u, g, R, c), c, terms);
  if (c >= 3) return;
  const int size = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
  const double Rd = d == 0 ? R[0] : (d == 1 ? R[1] : R[2]);
  v[size_t(nt) * d + it] = vel;
  const double Rn = Rd + vel;
  r[size_t(nt) * d + it] = wrap_position(Rn + size, double(size));
  sd[size_t(nt) * d + it] = s_old + vel;
}

// Particles are stored in spatial order, so the move kernels give each XCD one
// contiguous range of particle blocks (xcd_block): neighbouring particles
// share their stencils' lines in one L2.

template <typename T, bool AOS, bool SCATTER = false>
__global__ __launch_bounds__(PBLOCK) void k_swimmers_move(int ns, SwimParams cp, double* __restrict__ r,
                                                          double* __restrict__ v, double* __restrict__ nv,
                                                          uint64_t* __restrict__ prng, const int* __restrict__ perm,
                                                          const T* __restrict__ u, Geom g,
                                                          unsigned long long* __restrict__ tumbles, uint32_t nb,
                                                          T* __restrict__ fnext) {
  __shared__ __attribute__((aligned(16))) double lds[GROUPS * PLANE_STRIDE];
  const uint32_t blk = xcd_block(round_xcd(nb));
  if (blk >= nb) return;
  swimmers_move_body<T, AOS, SCATTER>(ns, cp, r, v, nv, prng, perm, u, g, tumbles, blk, lds, fnext);
}

template <typename T, bool AOS>
__global__ __launch_bounds__(TMOVE_GROUPS * PAIR) void k_tracers_move(int nt, double* __restrict__ r,
                                                                     double* __restrict__ sd, double* __restrict__ v,
                                                                     const T* __restrict__ u, Geom g, uint32_t nb) {
  __shared__ __attribute__((aligned(16))) double lds[TMOVE_GROUPS * PLANE_STRIDE];
  const uint32_t blk = xcd_block(round_xcd(nb));
  if (blk >= nb) return;
  tracers_move_body<T, AOS, TMOVE_GROUPS>(nt, r, sd, v, u, g, blk, lds);
}

// Batched moves: WorkItem.member indexes the particle-array table, unit is
// the group block of that array; u and the geometry come from its lattice.
// The work list is split into XCD ranges the same way.
template <typename T>
__global__ __launch_bounds__(PBLOCK) void kb_swimmers_move(const LatDesc* __restrict__ ldesc,
                                                           const SwimDesc* __restrict__ sdesc,
                                                           const WorkItem* __restrict__ work, uint32_t nwork) {
  __shared__ __attribute__((aligned(16))) double lds[GROUPS * PLANE_STRIDE];
  const uint32_t wi = xcd_block(round_xcd(nwork));
  if (wi >= nwork) return;
  const WorkItem w = work[wi];
  const SwimDesc& a = sdesc[w.member];
  const LatDesc& m = ldesc[a.member];
  if (m.u_aos)  // the batched fused step wrote u only component-interleaved
    swimmers_move_body<T, true>(a.num, a.p, a.r, a.v, a.n, a.prng, a.perm, static_cast<const T*>(m.u_aos), m.geom,
                                a.tumbles, w.unit, lds);
  else
    swimmers_move_body<T, false>(a.num, a.p, a.r, a.v, a.n, a.prng, a.perm, static_cast<const T*>(m.u), m.geom,
                                 a.tumbles, w.unit, lds);
}

template <typename T>
__global__ __launch_bounds__(TBLOCK) void kb_tracers_move(const LatDesc* __restrict__ ldesc,
                                                          const TracerDesc* __restrict__ tdesc,
                                                          const WorkItem* __restrict__ work, uint32_t nwork) {
  __shared__ __attribute__((aligned(16))) double lds[TGROUPS * PLANE_STRIDE];
  const uint32_t wi = xcd_block(round_xcd(nwork));
  if (wi >= nwork) return;
  const WorkItem w = work[wi];
  const TracerDesc& a = tdesc[w.member];
  const LatDesc& m = ldesc[a.member];
  if (m.u_aos)
    tracers_move_body<T, true>(a.num, a.r, a.s, a.v, static_cast<const T*>(m.u_aos), m.geom, w.unit, lds);
  else
    tracers_move_body<T, false>(a.num, a.r, a.s, a.v, static_cast<const T*>(m.u), m.geom, w.unit, lds);
}

// ---- deterministic force spreading -----------------------------------------
// ZeroForce followed by AddForces of every swimmer array, computed as a gather
// per run of 64 consecutive lattice sites (one list per wavefront of the
// lattice kernels) so that each site adds its contributions in swimmer order,
// tail before head -- the order a sequential restatement of
// SwimmerArray.cu:42-112 uses -- without atomics on the force.
//   1. k_points: the tail and head stencils of every swimmer (wrapped site
//      coordinates and Peskin weights, exactly as AccumulateDeltaForce forms
//      them) are stored by point id = 2*(array base + user index) + end;
//   2. k_count_rows: every list a stencil row touches is counted;
//   3. k_scan_sums / k_scan_apply: exclusive scan of the counts;
//   4. k_fill_rows: point ids appended to their lists;
//   5. k_sort_lists: each list ordered by point id (stable rank sort);
//   6. either k_spread_blocks writes the force of every site (zeros where no
//      point reaches), or the collide and CalcHydro kernels evaluate the force
//      of their own sites from the same lists (wave_force).

// One thread per point: thread t is end t & 1 of the swimmer in slot t >> 1.
// slot_ids[point_base + t] = the id of the point thread t forms: points in
// storage (spatial) order, which the list building walks.
__device__ __forceinline__ void points_body(int ns, uint32_t point_base, double P, double l,
                                            const double* __restrict__ r, const double* __restrict__ nv,
                                            const int* __restrict__ perm, const Geom& g, int member,
                                            SpreadPoint* __restrict__ pts, uint32_t* __restrict__ slot_ids,
                                            uint32_t blk) {
  const int t = int(blk * PBLOCK + threadIdx.x);
  if (t >= 2 * ns) return;
  const int slot = t >> 1, end = t & 1;
  const int size[3] = {g.nx, g.ny, g.nz};
  SpreadPoint sp;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double ra = r[size_t(ns) * d + slot], na = nv[size_t(ns) * d + slot];
    const double Ft = -P * na;
    const double rp = end ? ra : ra - na * l;
    const double x0 = ceil(rp - 2.);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double x = x0 + k;
      sp.wc[d * 4 + k] = int(wrap_index(x, size[d]));
      sp.del[d * 4 + k] = peskin(rp - x);
    }
    sp.F[d] = end ? Ft * -1.0 : Ft;
  }
  sp.pad[0] = member;
  sp.pad[1] = 0;
  const uint32_t id = point_base + 2 * uint32_t(perm[slot]) + uint32_t(end);
  pts[id] = sp;
  slot_ids[point_base + uint32_t(t)] = id;
}

__global__ __launch_bounds__(PBLOCK) void k_points(int ns, int base, double P, double l, const double* __restrict__ r,
                                                  const double* __restrict__ nv, const int* __restrict__ perm,
                                                  Geom g, SpreadPoint* __restrict__ pts,
                                                  uint32_t* __restrict__ slot_ids) {
  points_body(ns, 2 * uint32_t(base), P, l, r, nv, perm, g, 0, pts, slot_ids, blockIdx.x);
}

// Batched: work items are blocks of 128 swimmers (256 points) of an array.
__global__ __launch_bounds__(PBLOCK) void kb_points(const LatDesc* __restrict__ ldesc,
                                                   const SwimDesc* __restrict__ sdesc,
                                                   const WorkItem* __restrict__ work, SpreadPoint* __restrict__ pts,
                                                   uint32_t* __restrict__ slot_ids) {
  const WorkItem w = work[blockIdx.x];
  const SwimDesc& a = sdesc[w.member];
  const LatDesc& m = ldesc[a.member];
  points_body(a.num, a.point_base, a.p.P, a.p.l, a.r, a.n, a.perm, m.geom, a.member, pts, slot_ids, w.unit);
}

// Up to 4 lists (each once) that stencil row `row` of point p reaches.
struct RowLists {
  uint32_t l[4];
  int count;
};
__device__ __forceinline__ RowLists row_lists(const SpreadPoint& p, int row, const Geom& g, uint32_t base) {
  RowLists r{{0u, 0u, 0u, 0u}, 0};
  for_each_row_list(p, row, g, [&](uint32_t b) { r.l[r.count++] = base + b; });
  return r;
}

// Dense systems: the lanes of a wavefront (points in storage order, so
// spatial neighbours) often hit the same list; one lane per distinct list
// adds the wavefront's total with one atomic and hands each lane its slot.
// The distinct lists are found by ballots first (no memory traffic), then
// every leader lane issues its atomic at once, so the wavefront waits for one
// round trip instead of one per distinct list.  Returns the value the
// atomic returned plus the lane's rank among the lanes with the same list.
// Every lane of the wavefront calls it; lanes with `key` ~0u take part
// without a list.
struct WaveGroup {
  int leader;      // lowest lane with this lane's key
  uint32_t rank;   // lanes below with the same key
  uint32_t total;  // lanes with this lane's key (valid on the leader)
};
__device__ __forceinline__ WaveGroup wave_group(uint32_t key) {
  const int lane = int(threadIdx.x % WAVE_SITES);
  const uint64_t below = (uint64_t(1) << lane) - 1;
  WaveGroup w{lane, 0u, 0u};
  uint64_t pending = __ballot(key != 0xFFFFFFFFu);
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    const uint32_t lk = uint32_t(__builtin_amdgcn_readlane(int(key), leader));
    const uint64_t same = __ballot(key == lk);
    if (key == lk) {
      w.leader = leader;
      w.rank = uint32_t(__builtin_popcountll(same & below));
      w.total = uint32_t(__builtin_popcountll(same));
    }
    pending &= ~same;
  }
  return w;
}
__device__ __forceinline__ uint32_t wave_aggregated_add(uint32_t* __restrict__ ctr, uint32_t key) {
  const WaveGroup w = wave_group(key);
  const int lane = int(threadIdx.x % WAVE_SITES);
  uint32_t base = 0;
  if (key != 0xFFFFFFFFu && w.leader == lane) base = atomicAdd(&ctr[key], w.total);
  return uint32_t(__shfl(int(base), w.leader)) + w.rank;
}
// The same counts without returning slots (no wait for the atomics).
__device__ __forceinline__ void wave_aggregated_count(uint32_t* __restrict__ ctr, uint32_t key) {
  const WaveGroup w = wave_group(key);
  if (key != 0xFFFFFFFFu && w.leader == int(threadIdx.x % WAVE_SITES)) atomicAdd(&ctr[key], w.total);
}

// One thread per (point, stencil row), points in storage order (slot_ids):
// counts the lists the row reaches.  `Batched`: geometry and first list from
// the point's member.  AGG: wavefront-aggregated atomics (dense systems).
template <bool Batched, bool AGG>
__global__ __launch_bounds__(PBLOCK) void k_count_rows(uint32_t npts, Geom g, const LatDesc* __restrict__ ldesc,
                                                      const SpreadPoint* __restrict__ pts,
                                                      const uint32_t* __restrict__ slot_ids,
                                                      uint32_t* __restrict__ counts) {
  const uint32_t t = blockIdx.x * PBLOCK + threadIdx.x;
  RowLists rl{{0u, 0u, 0u, 0u}, 0};
  if (t < 16 * npts) {
    const SpreadPoint& p = pts[slot_ids[t >> 4]];
    uint32_t base = 0;
    bool has = true;
    if constexpr (Batched) {
      const LatDesc& m = ldesc[p.pad[0]];
      has = has_lists(m.force_path);  // a member without lists
      g = m.geom;
      base = m.tile_base;
    }
    if (has) rl = row_lists(p, int(t & 15), g, base);
  }
  if constexpr (AGG) {
#pragma unroll
    for (int k = 0; k < 4; ++k) wave_aggregated_count(counts, k < rl.count ? rl.l[k] : 0xFFFFFFFFu);
  } else {
    for (int k = 0; k < rl.count; ++k) atomicAdd(&counts[rl.l[k]], 1u);
  }
}

// Exclusive scan of counts[0..nt) into off[0..nt] in three passes: the
// total of every tile of SCAN_TILE counts, an exclusive scan of those totals
// by one workgroup, and the scan within each tile from its offset.  Each
// thread holds SCAN_ITEMS consecutive counts (two 16-byte accesses; the
// arrays are allocation bases).  The last pass also resets counts to zero for
// the next build (its last reader) and clears cursor where one is given.
// Small workgroups: the scans run beside the tracer moves, and a workgroup
// has to wait for a CU with room for all of its wavefronts.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;
inline unsigned scan_blocks(size_t nt) { return unsigned((nt + SCAN_TILE - 1) / SCAN_TILE); }

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* part, uint32_t* total) {
  const int t = int(threadIdx.x);
  part[t] = v;
  __syncthreads();
  for (int w = 1; w < SCAN_BLOCK; w <<= 1) {
    const uint32_t add = t >= w ? part[t - w] : 0u;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  const uint32_t incl = part[t];
  *total = part[SCAN_BLOCK - 1];
  __syncthreads();
  return incl - v;
}

__device__ __forceinline__ void scan_load(const uint32_t* __restrict__ a, int i0, int nt, uint32_t (&v)[SCAN_ITEMS]) {
  if (i0 + SCAN_ITEMS <= nt) {
    const uint4 x = *reinterpret_cast<const uint4*>(a + i0), y = *reinterpret_cast<const uint4*>(a + i0 + 4);
    v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = i0 + k < nt ? a[i0 + k] : 0u;
  }
}

__device__ __forceinline__ void scan_store(uint32_t* __restrict__ a, int i0, int nt, const uint32_t (&v)[SCAN_ITEMS]) {
  if (i0 + SCAN_ITEMS <= nt) {
    *reinterpret_cast<uint4*>(a + i0) = make_uint4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<uint4*>(a + i0 + 4) = make_uint4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
      if (i0 + k < nt) a[i0 + k] = v[k];
  }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int nt, const uint32_t* __restrict__ counts,
                                                          uint32_t* __restrict__ sums) {
  __shared__ uint32_t part[SCAN_BLOCK];
  const int i0 = (int(blockIdx.x) * SCAN_BLOCK + int(threadIdx.x)) * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  scan_load(counts, i0, nt, v);
  uint32_t mine = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) mine += v[k];
  uint32_t total = 0;
  block_exclusive_scan(mine, part, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// The nsb tile totals replaced by their exclusive prefix sums, by one
// workgroup (SCAN_PER consecutive totals per thread, in chunks).
constexpr int SCAN_PER = 16;
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_block_sums(int nsb, uint32_t* __restrict__ sums) {
  __shared__ uint32_t part[SCAN_BLOCK];
  uint32_t carry = 0;
  for (int base = 0; base < nsb; base += SCAN_BLOCK * SCAN_PER) {
    const int i0 = base + int(threadIdx.x) * SCAN_PER;
    uint32_t v[SCAN_PER];
    uint32_t mine = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      v[k] = i0 + k < nsb ? sums[i0 + k] : 0u;
      mine += v[k];
    }
    uint32_t total = 0;
    uint32_t run = carry + block_exclusive_scan(mine, part, &total);
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      if (i0 + k < nsb) sums[i0 + k] = run;
      run += v[k];
    }
    carry += total;
  }
}

// `sums`: the tile totals after k_scan_block_sums (exclusive prefixes).
// `cursor` may be null.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(int nt, uint32_t* __restrict__ counts,
                                                           const uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ off, uint32_t* __restrict__ cursor) {
  __shared__ uint32_t part[SCAN_BLOCK];
  const uint32_t base = sums[blockIdx.x];  // offset of this tile
  const int i0 = (int(blockIdx.x) * SCAN_BLOCK + int(threadIdx.x)) * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  scan_load(counts, i0, nt, v);
  uint32_t mine = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) mine += v[k];
  uint32_t total = 0;
  uint32_t run = base + block_exclusive_scan(mine, part, &total);
  uint32_t o[SCAN_ITEMS], zero[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    o[k] = run;
    run += v[k];
    zero[k] = 0u;
  }
  scan_store(off, i0, nt, o);
  scan_store(counts, i0, nt, zero);
  if (cursor) scan_store(cursor, i0, nt, zero);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    off[nt] = base + total;
    off[nt + 1 + gridDim.x] = 0u;  // queue counter of the sort, after the scan's partial sums
  }
}

// One thread per (point, stencil row), points in storage order: appends the
// point id to the lists the row reaches (in any order; k_sort_lists orders
// them).  AGG as for k_count_rows.
template <bool Batched, bool AGG>
__global__ __launch_bounds__(PBLOCK) void k_fill_rows(uint32_t npts, Geom g, const LatDesc* __restrict__ ldesc,
                                                     const SpreadPoint* __restrict__ pts,
                                                     const uint32_t* __restrict__ slot_ids,
                                                     const uint32_t* __restrict__ off, uint32_t* __restrict__ cursor,
                                                     uint32_t* __restrict__ list) {
  const uint32_t t = blockIdx.x * PBLOCK + threadIdx.x;
  RowLists rl{{0u, 0u, 0u, 0u}, 0};
  uint32_t id = 0;
  if (t < 16 * npts) {
    id = slot_ids[t >> 4];
    const SpreadPoint& p = pts[id];
    uint32_t base = 0;
    bool has = true;
    if constexpr (Batched) {
      const LatDesc& m = ldesc[p.pad[0]];
      has = has_lists(m.force_path);
      g = m.geom;
      base = m.tile_base;
    }
    if (has) rl = row_lists(p, int(t & 15), g, base);
  }
  if constexpr (AGG) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t key = k < rl.count ? rl.l[k] : 0xFFFFFFFFu;
      const uint32_t pos = wave_aggregated_add(cursor, key);
      if (key != 0xFFFFFFFFu) list[off[key] + pos] = id;
    }
  } else {
    for (int k = 0; k < rl.count; ++k) list[off[rl.l[k]] + atomicAdd(&cursor[rl.l[k]], 1u)] = id;
  }
}

// Sort of every list by point id, in two passes.  k_sort_lists: one lane per
// list sorts lists of up to 16 entries in registers with Batcher's odd-even
// merge networks (19 comparators for 8 inputs, 63 for 16; equal ids -- a point
// listed twice -- are interchangeable) and queues the longer ones.
// k_sort_long: one wavefront per queued list ranks its entries (stable: equal
// ids keep their order, so every position of `sorted` is written once), from
// a copy in LDS when it fits.
constexpr uint32_t SHORT_LIST = 16;
constexpr uint32_t LONG_LDS = 2048;  // entries per wavefront staged in LDS
constexpr int LONG_BLOCK = 256;

__device__ __forceinline__ void cmp_swap(uint32_t& a, uint32_t& b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}

constexpr uint8_t NET8[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                 {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
constexpr uint8_t NET16[63][2] = {
    {0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}, {4, 5}, {6, 7}, {4, 6}, {5, 7}, {5, 6}, {0, 4}, {2, 6},
    {2, 4}, {1, 5}, {3, 7}, {3, 5}, {1, 2}, {3, 4}, {5, 6}, {8, 9}, {10, 11}, {8, 10}, {9, 11}, {9, 10},
    {12, 13}, {14, 15}, {12, 14}, {13, 15}, {13, 14}, {8, 12}, {10, 14}, {10, 12}, {9, 13}, {11, 15}, {11, 13}, {9, 10},
    {11, 12}, {13, 14}, {0, 8}, {4, 12}, {4, 8}, {2, 10}, {6, 14}, {6, 10}, {2, 4}, {6, 8}, {10, 12}, {1, 9},
    {5, 13}, {5, 9}, {3, 11}, {7, 15}, {7, 11}, {3, 5}, {7, 9}, {11, 13}, {1, 2}, {3, 4}, {5, 6}, {7, 8},
    {9, 10}, {11, 12}, {13, 14}};

template <int N, int C>
__device__ __forceinline__ void sort_net(const uint32_t* __restrict__ src, uint32_t cnt, uint32_t* __restrict__ dst,
                                         const uint8_t (&net)[C][2]) {
  uint32_t v[N];
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = uint32_t(k) < cnt ? src[k] : 0xFFFFFFFFu;  // ids stay below 2^31
#pragma unroll
  for (int c = 0; c < C; ++c) cmp_swap(v[net[c][0]], v[net[c][1]]);
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (uint32_t(k) < cnt) dst[k] = v[k];
}

// `queue` (nlists entries) and `nqueued` receive the lists longer than 16;
// nqueued is zeroed by the scan.
__global__ __launch_bounds__(PBLOCK) void k_sort_lists(uint32_t nlists, const uint32_t* __restrict__ off,
                                                      const uint32_t* __restrict__ list,
                                                      uint32_t* __restrict__ sorted, uint32_t* __restrict__ queue,
                                                      uint32_t* __restrict__ nqueued) {
  const uint32_t wl = blockIdx.x * PBLOCK + threadIdx.x;
  if (wl >= nlists) return;
  const uint32_t o = off[wl], cnt = off[wl + 1] - o;
  if (cnt <= 8)
    sort_net<8>(list + o, cnt, sorted + o, NET8);
  else if (cnt <= SHORT_LIST)
    sort_net<16>(list + o, cnt, sorted + o, NET16);
  else
    queue[atomicAdd(nqueued, 1u)] = wl;
}

// Stable ranks of list src[0..cnt) by the lanes of one wavefront.
__device__ __forceinline__ void rank_list(const uint32_t* __restrict__ src, uint32_t cnt, uint32_t lane,
                                          uint32_t* __restrict__ dst) {
  for (uint32_t e = lane; e < cnt; e += WAVE_SITES) {
    const uint32_t id = src[e];
    uint32_t rank = 0;
    for (uint32_t k = 0; k < cnt; ++k) {
      const uint32_t v = src[k];
      rank += (v < id || (v == id && k < e)) ? 1u : 0u;
    }
    dst[rank] = id;
  }
}

// Sorts v[0..cnt) in place by one wavefront with a bitonic network over the
// next power of two (padded with the largest id); the lanes take disjoint
// pairs of each stage, and a wavefront's LDS accesses complete in order.
__device__ __forceinline__ void bitonic_lds(uint32_t* __restrict__ v, uint32_t cnt, uint32_t lane) {
  uint32_t P = 1;
  while (P < cnt) P <<= 1;
  for (uint32_t e = cnt + lane; e < P; e += WAVE_SITES) v[e] = 0xFFFFFFFFu;
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (uint32_t i = lane; i < P; i += WAVE_SITES) {
        const uint32_t x = i ^ j;
        if (x > i) {
          const uint32_t a = v[i], b = v[x];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            v[i] = b;
            v[x] = a;
          }
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Fixed grid; wavefront w takes queued lists w, w + waves, ...  Lists that
// fit in the wavefront's LDS slice are sorted there (bitonic, equal ids are
// interchangeable), longer ones ranked from memory.
__global__ __launch_bounds__(LONG_BLOCK) void k_sort_long(const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ list,
                                                         uint32_t* __restrict__ sorted,
                                                         const uint32_t* __restrict__ queue,
                                                         const uint32_t* __restrict__ nqueued) {
  __shared__ uint32_t stage[LONG_BLOCK / WAVE_SITES][LONG_LDS];
  const uint32_t wave = threadIdx.x / WAVE_SITES, lane = threadIdx.x % WAVE_SITES;
  const uint32_t waves = gridDim.x * (LONG_BLOCK / WAVE_SITES);
  const uint32_t n = *nqueued;
  uint32_t* mine = stage[wave];
  for (uint32_t q = blockIdx.x * (LONG_BLOCK / WAVE_SITES) + wave; q < n; q += waves) {
    const uint32_t wl = queue[q];
    const uint32_t o = off[wl], cnt = off[wl + 1] - o;
    if (cnt <= LONG_LDS) {
      for (uint32_t e = lane; e < cnt; e += WAVE_SITES) mine[e] = list[o + e];
      bitonic_lds(mine, cnt, lane);
      for (uint32_t e = lane; e < cnt; e += WAVE_SITES) sorted[o + e] = mine[e];
      // the next list reuses the slice
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
      rank_list(list + o, cnt, lane, sorted + o);
    }
  }
}

// Force of the sites of lattice block `blk` (lists first_list.. of its
// wavefronts) written to the force array.
template <typename T>
__device__ __forceinline__ void spread_block_body(const Geom& g, const SpreadPoint* __restrict__ pts,
                                                  const uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ sorted, T* __restrict__ force,
                                                  uint32_t blk, uint32_t first_list, WaveStage* __restrict__ stage) {
  const uint32_t s = blk * SITE_BLOCK + threadIdx.x;
  const bool ok = s < g.n;
  const SiteXYZ c = site_xyz(ok ? s : 0u, g);
  double acc[3];
  wave_force_staged(pts, off, sorted, first_list + threadIdx.x / WAVE_SITES, ok, c.x, c.y, c.z, g,
                    stage + threadIdx.x / WAVE_SITES, acc);
  if (!ok) return;
  const size_t n = g.n;
  force[s] = T(acc[0]);
  force[n + s] = T(acc[1]);
  force[2 * n + s] = T(acc[2]);
}

// Force of the sites of lattice block `blk` from all points first ..
// first+count-1 (range_force), written to the force array.
// The whole workgroup calls it (range_force_staged holds barriers).
template <typename T>
__device__ __forceinline__ void range_block_body(const Geom& g, const SpreadPoint* __restrict__ pts, uint32_t first,
                                                 uint32_t count, T* __restrict__ force, uint32_t blk,
                                                 RangeStage* __restrict__ st) {
  const uint32_t s = blk * SITE_BLOCK + threadIdx.x;
  const bool ok = s < g.n;
  const SiteXYZ c = site_xyz(ok ? s : 0u, g);
  double acc[3];
  range_force_staged(pts, first, count, ok, c.x, c.y, c.z, g, st, acc);
  if (!ok) return;
  const size_t n = g.n;
  force[s] = T(acc[0]);
  force[n + s] = T(acc[1]);
  force[2 * n + s] = T(acc[2]);
}

template <typename T>
__global__ __launch_bounds__(SITE_BLOCK) void k_range_blocks(Geom g, const SpreadPoint* __restrict__ pts,
                                                            uint32_t npts, T* __restrict__ force) {
  __shared__ RangeStage st;
  range_block_body(g, pts, 0u, npts, force, blockIdx.x, &st);
}

template <typename T>
__global__ __launch_bounds__(SITE_BLOCK) void k_spread_blocks(Geom g, const SpreadPoint* __restrict__ pts,
                                                             const uint32_t* __restrict__ off,
                                                             const uint32_t* __restrict__ sorted,
                                                             T* __restrict__ force) {
  __shared__ WaveStage stage[LISTS_PER_BLOCK];
  spread_block_body(g, pts, off, sorted, force, blockIdx.x, blockIdx.x * LISTS_PER_BLOCK, stage);
}

// ---- stored force by site (dense systems, every lattice side >= 4) ----------
// With every side >= 4 a point's 4x4x4 stencil covers 64 distinct sites, and
// add_point gives site (wc[i], wc[4+j], wc[8+k]) the product
// (del[i] * del[4+j]) * del[8+k] * F.  The points' entries are counted per
// site (each entry keeps the slot its atomic returned), the counts scanned
// into segments, the point ids written to their slots, every product placed
// at the rank of its point id within its segment, and a thread per site adds
// its segment in order from zero: the products and the order of the sorted
// 64-site lists (wave_force), which follow the sequential AccumulateDeltaForce
// order (SwimmerArray.cu:42-79) -- but each site walks only the points that
// reach it instead of every point of its 64-site run.  Entry e = 16 id + row
// holds the 4 sites of stencil row `row` = 4 i + j.

// Linear site of row `row` = 4 i + j of point p's stencil at z offset 0.
__device__ __forceinline__ uint32_t stencil_row_site(const SpreadPoint& p, int row, const Geom& g) {
  return (uint32_t(p.wc[row >> 2]) * uint32_t(g.ny) + uint32_t(p.wc[4 + (row & 3)])) * uint32_t(g.nz);
}

// The point's member (Batched: geometry and first site from its LatDesc);
// false for a batched point whose member takes another force path.
template <bool Batched>
__device__ __forceinline__ bool site_member(const SpreadPoint& p, const LatDesc* __restrict__ ldesc, Geom& g,
                                            uint32_t& base) {
  base = 0;
  if constexpr (Batched) {
    const LatDesc& m = ldesc[p.pad[0]];
    if (m.force_path != FORCE_SITES_STORED) return false;
    g = m.geom;
    base = m.site_base;
  }
  return true;
}

// One thread per (point, stencil row), points in storage order (slot_ids:
// spatial neighbours): counts the row's 4 sites and keeps each site's slot
// within its segment.  A workgroup's 16 points share most of their sites (a
// swimmer's tail and head lie a fraction of a site apart), so the workgroup
// counts its entries per site in an LDS hash table first (linear probing;
// at most 1024 distinct sites in SITE_HASH slots, so the table stays at most
// half full and every probe ends), then reserves each distinct site's slots
// with one global atomic, and every entry takes its site's reserved base
// plus its rank from the LDS count.
constexpr int SITE_HASH = 2048;
template <bool Batched>
__global__ __launch_bounds__(PBLOCK) void k_site_count(uint32_t npts, Geom g, const LatDesc* __restrict__ ldesc,
                                                      const SpreadPoint* __restrict__ pts,
                                                      const uint32_t* __restrict__ slot_ids,
                                                      uint32_t* __restrict__ counts, uint4* __restrict__ local) {
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  __shared__ uint32_t key[SITE_HASH];
  __shared__ uint32_t num[SITE_HASH];
  for (int i = int(threadIdx.x); i < SITE_HASH; i += PBLOCK) {
    key[i] = NONE;
    num[i] = 0u;
  }
  const uint32_t t = blockIdx.x * PBLOCK + threadIdx.x;
  uint32_t site[4] = {NONE, NONE, NONE, NONE};
  if (t < 16 * npts) {
    const SpreadPoint& p = pts[slot_ids[t >> 4]];
    uint32_t base;
    if (site_member<Batched>(p, ldesc, g, base)) {
      const uint32_t row = base + stencil_row_site(p, int(t & 15), g);
#pragma unroll
      for (int k = 0; k < 4; ++k) site[k] = row + uint32_t(p.wc[8 + k]);
    }
  }
  __syncthreads();
  uint32_t h[4] = {0u, 0u, 0u, 0u}, rank[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (site[k] == NONE) continue;
    uint32_t x = (site[k] * 2654435761u) >> 21;  // 11 bits: SITE_HASH slots
    for (;;) {
      const uint32_t old = atomicCAS(&key[x], NONE, site[k]);
      if (old == NONE || old == site[k]) break;
      x = (x + 1u) & uint32_t(SITE_HASH - 1);
    }
    h[k] = x;
    rank[k] = atomicAdd(&num[x], 1u);
  }
  __syncthreads();
  for (int i = int(threadIdx.x); i < SITE_HASH; i += PBLOCK) {
    const uint32_t sk = key[i];
    if (sk != NONE) num[i] = atomicAdd(&counts[sk], num[i]);  // the workgroup's first slot of site sk
  }
  __syncthreads();
  if (site[0] != NONE)
    local[t] = make_uint4(num[h[0]] + rank[0], num[h[1]] + rank[1], num[h[2]] + rank[2], num[h[3]] + rank[3]);
}

// One thread per (point, stencil row): the entry key -- point id * 64 +
// stencil index 16 i + 4 j + k (ids below 2^26) -- and the site into the
// row's 4 slots, one 8-byte store per entry.
template <bool Batched>
__global__ __launch_bounds__(PBLOCK) void k_site_fill(uint32_t npts, Geom g, const LatDesc* __restrict__ ldesc,
                                                     const SpreadPoint* __restrict__ pts,
                                                     const uint32_t* __restrict__ slot_ids,
                                                     const uint32_t* __restrict__ off,
                                                     const uint4* __restrict__ local, uint2* __restrict__ ent) {
  const uint32_t t = blockIdx.x * PBLOCK + threadIdx.x;
  if (t >= 16 * npts) return;
  const uint32_t id = slot_ids[t >> 4];
  const SpreadPoint& p = pts[id];
  uint32_t base;
  if (!site_member<Batched>(p, ldesc, g, base)) return;
  const uint32_t row = base + stencil_row_site(p, int(t & 15), g);
  const uint4 q = local[t];
  const uint32_t slot[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t site = row + uint32_t(p.wc[8 + k]);
    const uint32_t e = off[site] + slot[k];
    ent[e] = make_uint2((id << 6) | (uint32_t(t & 15) << 2) | uint32_t(k), site);
  }
}

// One thread per slot, in slot order (so a wavefront's lanes mostly scan
// the same segment): the entry's product, formed from its point and stencil
// index as add_point forms it, goes to the rank of its point id within the
// segment (the points of a segment are distinct, so keys rank as ids do).
// nent = 64 points of the members concerned.
__global__ __launch_bounds__(PBLOCK) void k_site_place(uint32_t nent, const SpreadPoint* __restrict__ pts,
                                                      const uint32_t* __restrict__ off,
                                                      const uint2* __restrict__ ent, double* __restrict__ sorted) {
  const uint32_t e = blockIdx.x * PBLOCK + threadIdx.x;
  if (e >= nent) return;
  const uint2 me = ent[e];
  const uint32_t key = me.x, site = me.y;
  const SpreadPoint& p = pts[key >> 6];
  const int di = int(key >> 4) & 3, dj = int(key >> 2) & 3, dk = int(key) & 3;
  const uint32_t o = off[site], end = off[site + 1];
  uint32_t rank = 0;
  for (uint32_t q = o; q < end; ++q) rank += ent[q].x < key ? 1u : 0u;
  const double d3 = p.del[di] * p.del[4 + dj] * p.del[8 + dk];
  double* __restrict__ dst = sorted + 3 * size_t(o + rank);
#pragma unroll
  for (int a = 0; a < 3; ++a) dst[a] = d3 * p.F[a];
}

// Force at site s (segment gs) of a lattice of n sites: its segment's
// products added in order (segment_force).
template <typename T>
__device__ __forceinline__ void site_sum_body(uint32_t n, uint32_t s, uint32_t gs, const uint32_t* __restrict__ off,
                                              const double* __restrict__ sorted, T* __restrict__ force) {
  double acc[3];
  segment_force(off, sorted, gs, acc);
#pragma unroll
  for (int a = 0; a < 3; ++a) force[size_t(a) * n + s] = T(acc[a]);
}

// One thread per site of a single lattice.
template <typename T>
__global__ __launch_bounds__(PBLOCK) void k_site_sum(uint32_t n, const uint32_t* __restrict__ off,
                                                    const double* __restrict__ sorted, T* __restrict__ force) {
  const uint32_t s = blockIdx.x * PBLOCK + threadIdx.x;
  if (s < n) site_sum_body(n, s, s, off, sorted, force);
}

// Force at the first and the last site of every 64-site run (edge[6 l + 3 e
// + a]: list l, e = 0 first, 1 last), the sum wave_force gives there.  The
// fused step's z-halo sites are such sites of neighbouring runs; they load
// it instead of walking one list per lane.
__global__ __launch_bounds__(SITE_BLOCK) void k_edge_force(Geom g, uint32_t nlists,
                                                          const SpreadPoint* __restrict__ pts,
                                                          const uint32_t* __restrict__ off,
                                                          const uint32_t* __restrict__ sorted,
                                                          double* __restrict__ edge) {
  const uint32_t i = blockIdx.x * SITE_BLOCK + threadIdx.x;
  if (i >= 2 * nlists) return;
  const uint32_t l = i >> 1;
  const uint32_t s = l * WAVE_SITES + ((i & 1) ? WAVE_SITES - 1 : 0);
  double acc[3] = {0.0, 0.0, 0.0};
  if (s < g.n) {
    const SiteXYZ c = site_xyz(s, g);
    lane_force(pts, off, sorted, l, c.x, c.y, c.z, g, acc);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) edge[3 * size_t(i) + a] = acc[a];
}

// RunEntry rec[k] and wz[k] of every list entry k (g.nz % 64 == 0: the run of
// list l is one row x, y): the row weight dx_i dy_j at that row as the row
// walk of wave_force forms it, and wz = -1 where that walk skips the entry
// (the id repeats the previous entry's, or the stencil misses the row).  A
// wavefront takes ENTRY_LISTS consecutive lists, whose entries are
// contiguous, one lane per entry, and finds an entry's list among the
// lists' bounds held by lanes 0..ENTRY_LISTS.
//
// With `edge` (simple lattice: nx, ny >= 4) it also writes k_edge_force's
// output: the entries reaching the first or last site of their run are taken
// lane by lane in list order and added, with add_point's products, by the
// lane that holds that list (lane j: list l0 + j).
constexpr uint32_t ENTRY_LISTS = 16;
// One wavefront's ENTRY_LISTS lists from l0 (< nlists).
__device__ __forceinline__ void run_entries_wave(const Geom& g, uint32_t nlists, uint32_t l0,
                                                 const SpreadPoint* __restrict__ pts,
                                                 const uint32_t* __restrict__ off,
                                                 const uint32_t* __restrict__ sorted, RunEntry* __restrict__ rec,
                                                 int* __restrict__ wz, double* __restrict__ edge) {
  const int lane = int(threadIdx.x % WAVE_SITES);
  // lane j <= ENTRY_LISTS: first entry of list l0 + j (the end for j = ENTRY_LISTS)
  const uint32_t oj = off[min(l0 + min(uint32_t(lane), ENTRY_LISTS), nlists)];
  const uint32_t k0 = uint32_t(__builtin_amdgcn_readlane(int(oj), 0));
  const uint32_t k1 = uint32_t(__builtin_amdgcn_readlane(int(oj), int(ENTRY_LISTS)));
  double ef[3] = {0.0, 0.0, 0.0}, el[3] = {0.0, 0.0, 0.0};  // lane j: first and last site of list l0 + j
  for (uint32_t base = k0; base < k1; base += WAVE_SITES) {
    const uint32_t k = base + uint32_t(lane);
    // list of entry k: the last j with off[l0 + j] <= k (every lane searches,
    // so every lane the shuffles read from is active)
    int j = 0;
#pragma unroll
    for (int step = int(ENTRY_LISTS) / 2; step; step >>= 1)
      if (uint32_t(__shfl(int(oj), j + step)) <= k) j += step;
    const uint32_t first = uint32_t(__shfl(int(oj), j));
    const bool valid = k < k1;
    RunEntry r{};
    int w = -1;
    if (valid) {
      const uint32_t l = l0 + uint32_t(j);
      const SiteXYZ c = site_xyz(l * WAVE_SITES, g);
      const uint32_t id = sorted[k];
      const uint32_t prev = k > first ? sorted[k - 1] : 0xFFFFFFFFu;
      const SpreadPoint* __restrict__ p = pts + id;
      int di = c.x - p->wc[0];
      di += di < 0 ? g.nx : 0;
      int dj = c.y - p->wc[4];
      dj += dj < 0 ? g.ny : 0;
      const bool skip = id == prev || (di | dj) >= 4;
      r.dxy = pick4(p->del[0], p->del[1], p->del[2], p->del[3], di & 3) *
              pick4(p->del[4], p->del[5], p->del[6], p->del[7], dj & 3);
#pragma unroll
      for (int q = 0; q < 4; ++q) r.zw[q] = p->del[8 + q];
#pragma unroll
      for (int a = 0; a < 3; ++a) r.F[a] = p->F[a];
      rec[k] = r;
      w = skip ? -1 : p->wc[8];
      wz[k] = w;
    }
    if (!edge) continue;
    // z of the run's first site: l0 + j's runs start at multiples of 64 in z
    const int z0 = int((l0 + uint32_t(j)) * WAVE_SITES % uint32_t(g.nz));
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      int dk = z0 + (e ? WAVE_SITES - 1 : 0) - w;
      dk += dk < 0 ? g.nz : 0;
      const bool reach = w >= 0 && dk < 4;
      const double d3 = r.dxy * pick4(r.zw[0], r.zw[1], r.zw[2], r.zw[3], dk & 3);
      const double t[3] = {d3 * r.F[0], d3 * r.F[1], d3 * r.F[2]};
      uint64_t mask = __ballot(reach);
      while (mask) {
        const int L = __builtin_ctzll(mask);
        mask &= mask - 1;
        const int jl = __builtin_amdgcn_readlane(j, L);
        const double v0 = __shfl(t[0], L), v1 = __shfl(t[1], L), v2 = __shfl(t[2], L);
        if (lane == jl) {
          double* acc = e ? el : ef;
          acc[0] += v0;
          acc[1] += v1;
          acc[2] += v2;
        }
      }
    }
  }
  if (edge && uint32_t(lane) < ENTRY_LISTS && l0 + uint32_t(lane) < nlists) {
    double* out = edge + 6 * size_t(l0 + uint32_t(lane));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      out[a] = ef[a];
      out[3 + a] = el[a];
    }
  }
}

__global__ __launch_bounds__(SITE_BLOCK) void k_run_entries(Geom g, uint32_t nlists,
                                                           const SpreadPoint* __restrict__ pts,
                                                           const uint32_t* __restrict__ off,
                                                           const uint32_t* __restrict__ sorted,
                                                           RunEntry* __restrict__ rec, int* __restrict__ wz,
                                                           double* __restrict__ edge) {
  const uint32_t l0 = (blockIdx.x * SITE_BLOCK + threadIdx.x) / WAVE_SITES * ENTRY_LISTS;
  if (l0 >= nlists) return;  // whole wavefront
  run_entries_wave(g, nlists, l0, pts, off, sorted, rec, wz, edge);
}

// Batched: work item (member, unit) takes the member's lists 64 unit ...
// (ENTRY_LISTS per wavefront); list bounds and run-end forces are offset by
// the member's tile_base, entries and point ids are numbered across the batch.
constexpr uint32_t BATCH_ENTRY_LISTS = SITE_BLOCK / WAVE_SITES * ENTRY_LISTS;
static_assert(BATCH_ENTRY_LISTS == 64, "lists per work item of kb_run_entries");
__global__ __launch_bounds__(SITE_BLOCK) void kb_run_entries(const LatDesc* __restrict__ ldesc,
                                                            const WorkItem* __restrict__ work,
                                                            const SpreadPoint* __restrict__ pts,
                                                            const uint32_t* __restrict__ off,
                                                            const uint32_t* __restrict__ sorted,
                                                            RunEntry* __restrict__ rec, int* __restrict__ wz,
                                                            double* __restrict__ edge) {
  const WorkItem w = work[blockIdx.x];
  const LatDesc& m = ldesc[w.member];
  const uint32_t nlists = 4 * ((m.geom.n + SITE_BLOCK - 1) / SITE_BLOCK);
  const uint32_t l0 = w.unit * BATCH_ENTRY_LISTS + threadIdx.x / WAVE_SITES * ENTRY_LISTS;
  if (l0 >= nlists) return;  // whole wavefront
  run_entries_wave(m.geom, nlists, l0, pts, off + m.tile_base, sorted, rec, wz, edge + 6 * size_t(m.tile_base));
}

template <typename T>
__global__ __launch_bounds__(SITE_BLOCK) void kb_spread_blocks(const LatDesc* __restrict__ ldesc,
                                                              const WorkItem* __restrict__ block_work,
                                                              const SpreadPoint* __restrict__ pts,
                                                              const uint32_t* __restrict__ off,
                                                              const uint32_t* __restrict__ sorted, int per,
                                                              const uint32_t* __restrict__ soff,
                                                              const double* __restrict__ ssorted) {
  __shared__ union {
    WaveStage list[LISTS_PER_BLOCK];
    RangeStage range;
  } stage;
  const WorkItem w = block_work[blockIdx.x];
  const LatDesc& m = ldesc[w.member];
  const uint32_t nb = (m.geom.n + SITE_BLOCK - 1) / SITE_BLOCK;
  for (int b = 0; b < per; ++b) {  // the work item's `per` consecutive blocks
    const uint32_t blk = w.unit + uint32_t(b);
    if (blk >= nb) break;
    if (m.force_path == FORCE_SITES_STORED) {
      // soff null: the batched collide sums these members' segments itself
      if (soff) {
        const uint32_t site = blk * SITE_BLOCK + threadIdx.x;
        if (site < m.geom.n)
          site_sum_body(m.geom.n, site, m.site_base + site, soff, ssorted, static_cast<T*>(m.force));
      }
    } else if (m.force_path == FORCE_RANGE_STORED) {
      range_block_body(m.geom, pts, m.point_first, m.point_count, static_cast<T*>(m.force), blk, &stage.range);
    } else
      spread_block_body(m.geom, pts, off, sorted, static_cast<T*>(m.force), blk, m.tile_base + blk * LISTS_PER_BLOCK,
                        stage.list);
  }
}

// ---- device-math probes used by the parity tests ----
__global__ void k_peskin_probe(const double* __restrict__ x, double* __restrict__ out, int count) {
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < count) out[i] = peskin(x[i]);
}

__global__ void k_interp_probe(const double* __restrict__ u, const double* __restrict__ r,
                               double* __restrict__ v, int count, Geom g) {
  const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= count) return;
  const double R[3] = {r[i], r[size_t(count) + i], r[2 * size_t(count) + i]};
  double out[3];
  interp(u, g, R, out);
  v[i] = out[0];
  v[size_t(count) + i] = out[1];
  v[2 * size_t(count) + i] = out[2];
}

inline unsigned blocks_for(size_t threads) { return unsigned((threads + PBLOCK - 1) / PBLOCK); }

}  // namespace

template <typename T>
hipError_t launch_swimmers_add_forces(int ns, const SwimParams& p, const double* r, const double* n, T* force,
                                      const Geom& g, hipStream_t s) {
  if (ns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_swimmers_add_forces<T>, dim3(blocks_for(size_t(ns) * 64)), dim3(PBLOCK), 0, s, ns, p.P,
                     p.l, r, n, force, g);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_swimmers_move(int ns, const SwimParams& p, double* r, double* v, double* n, uint64_t* prng,
                                const int* perm, const T* u, bool aos, const Geom& g, unsigned long long* tumbles,
                                hipStream_t s, T* scatter) {
  if (ns <= 0) return hipSuccess;
  const uint32_t nb = uint32_t((ns + GROUPS - 1) / GROUPS);
  if (scatter) {
    // the scatter form is built for FP32 storage after a fused step only
    if constexpr (sizeof(T) == 4) {
      if (!aos) return hipErrorInvalidValue;
      hipLaunchKernelGGL((k_swimmers_move<T, true, true>), dim3(round_xcd(nb)), dim3(PBLOCK), 0, s, ns, p, r, v, n,
                         prng, perm, u, g, tumbles, nb, scatter);
    } else {
      return hipErrorInvalidValue;
    }
  } else if (aos) {
    hipLaunchKernelGGL((k_swimmers_move<T, true>), dim3(round_xcd(nb)), dim3(PBLOCK), 0, s, ns, p, r, v, n, prng, perm,
                       u, g, tumbles, nb, nullptr);
  } else {
    hipLaunchKernelGGL((k_swimmers_move<T, false>), dim3(round_xcd(nb)), dim3(PBLOCK), 0, s, ns, p, r, v, n, prng,
                       perm, u, g, tumbles, nb, nullptr);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tracers_move(int nt, double* r, double* sd, double* v, const T* u, bool aos, const Geom& g,
                               hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  const uint32_t nb = uint32_t((nt + TMOVE_GROUPS - 1) / TMOVE_GROUPS);
  const dim3 block(TMOVE_GROUPS * PAIR);
  if (aos)
    hipLaunchKernelGGL((k_tracers_move<T, true>), dim3(round_xcd(nb)), block, 0, s, nt, r, sd, v, u, g, nb);
  else
    hipLaunchKernelGGL((k_tracers_move<T, false>), dim3(round_xcd(nb)), block, 0, s, nt, r, sd, v, u, g, nb);
  return hipGetLastError();
}

template hipError_t launch_swimmers_add_forces<double>(int, const SwimParams&, const double*, const double*,
                                                       double*, const Geom&, hipStream_t);
template hipError_t launch_swimmers_add_forces<float>(int, const SwimParams&, const double*, const double*,
                                                      float*, const Geom&, hipStream_t);
template hipError_t launch_swimmers_move<double>(int, const SwimParams&, double*, double*, double*, uint64_t*,
                                                 const int*, const double*, bool, const Geom&, unsigned long long*,
                                                 hipStream_t, double*);
template hipError_t launch_swimmers_move<float>(int, const SwimParams&, double*, double*, double*, uint64_t*,
                                                const int*, const float*, bool, const Geom&, unsigned long long*,
                                                hipStream_t, float*);
template hipError_t launch_tracers_move<double>(int, double*, double*, double*, const double*, bool, const Geom&,
                                                hipStream_t);
template hipError_t launch_tracers_move<float>(int, double*, double*, double*, const float*, bool, const Geom&,
                                               hipStream_t);


size_t spread_off_entries(uint32_t nlists) {
  // nlists + 1 offsets, the scan's per-workgroup sums, the sort's queue counter
  return size_t(nlists) + 2 + scan_blocks(nlists);
}

size_t spread_list_capacity(size_t npts) {
  // A stencil has 16 (x, y) rows of 4 sites; the z-run of a row (one wrap at
  // most) lies in at most 4 lists, so a point is listed at most 64 times.
  return npts * 64;
}

uint32_t spread_lists(const Geom& g) { return (g.n + SITE_BLOCK - 1) / SITE_BLOCK * LISTS_PER_BLOCK; }

hipError_t launch_points(int ns, int base, const SwimParams& p, const double* r, const double* n, const int* perm,
                         const Geom& g, SpreadPoint* pts, uint32_t* slot_ids, hipStream_t s) {
  if (ns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_points, dim3(blocks_for(2 * size_t(ns))), dim3(PBLOCK), 0, s, ns, base, p.P, p.l, r, n, perm, g,
                     pts, slot_ids);
  return hipGetLastError();
}

// Counts, offsets, lists and sorted lists of nt lists from the npts point
// records.  counts[0..nt) must be zero on entry: they are zeroed when
// allocated and reset by k_scan_apply at the end of every build.
template <bool Batched>
static hipError_t build_lists(int npts, int nt, const Geom& g, const LatDesc* ldesc, const SpreadPoint* pts,
                              const uint32_t* slot_ids, bool dense, uint32_t* counts, uint32_t* off, uint32_t* cursor,
                              uint32_t* list, uint32_t* sorted, hipStream_t s) {
  const unsigned rows = blocks_for(16 * size_t(npts));
  if (npts > 0) {
    if (dense)
      hipLaunchKernelGGL((k_count_rows<Batched, true>), dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), g, ldesc, pts,
                         slot_ids, counts);
    else
      hipLaunchKernelGGL((k_count_rows<Batched, false>), dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), g, ldesc,
                         pts, slot_ids, counts);
  }
  // the per-workgroup totals of the scan live after the nt+1 offsets
  const unsigned nsb = scan_blocks(size_t(nt));
  uint32_t* sums = off + nt + 1;
  hipLaunchKernelGGL(k_scan_sums, dim3(nsb), dim3(SCAN_BLOCK), 0, s, nt, counts, sums);
  hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, int(nsb), sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(SCAN_BLOCK), 0, s, nt, counts, sums, off, cursor);
  if (npts > 0) {
    if (dense)
      hipLaunchKernelGGL((k_fill_rows<Batched, true>), dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), g, ldesc, pts,
                         slot_ids, off, cursor, list);
    else
      hipLaunchKernelGGL((k_fill_rows<Batched, false>), dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), g, ldesc, pts,
                         slot_ids, off, cursor, list);
  }
  // lists longer than 16 are queued in `cursor` (free once the lists are filled)
  uint32_t* nqueued = off + nt + 1 + nsb;
  hipLaunchKernelGGL(k_sort_lists, dim3(blocks_for(size_t(nt))), dim3(PBLOCK), 0, s, uint32_t(nt), off, list, sorted,
                     cursor, nqueued);
  // one wavefront per list, at most 512 workgroups (dense ensemble members
  // queue many long lists)
  const unsigned long_blocks = unsigned(std::min<size_t>(512, (size_t(nt) + 3) / 4));
  hipLaunchKernelGGL(k_sort_long, dim3(long_blocks), dim3(LONG_BLOCK), 0, s, off, list, sorted, cursor, nqueued);
  return hipGetLastError();
}

hipError_t launch_spread_lists(int npts, const Geom& g, const SpreadPoint* pts, const uint32_t* slot_ids,
                               uint32_t* counts, uint32_t* off, uint32_t* cursor, uint32_t* list, uint32_t* sorted,
                               hipStream_t s) {
  const uint32_t nt = spread_lists(g);
  return build_lists<false>(npts, int(nt), g, nullptr, pts, slot_ids, dense_lists(size_t(npts), nt), counts, off,
                            cursor, list, sorted, s);
}

template <typename T>
hipError_t launch_range_force(const Geom& g, const SpreadPoint* pts, uint32_t npts, T* force, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_range_blocks<T>, dim3((g.n + SITE_BLOCK - 1) / SITE_BLOCK), dim3(SITE_BLOCK), 0, s, g, pts, npts,
                     force);
  return hipGetLastError();
}
template hipError_t launch_range_force<double>(const Geom&, const SpreadPoint*, uint32_t, double*, hipStream_t);
template hipError_t launch_range_force<float>(const Geom&, const SpreadPoint*, uint32_t, float*, hipStream_t);

template <typename T>
hipError_t launch_site_force(int npts, const Geom& g, const SpreadPoint* pts, const uint32_t* slot_ids,
                             uint32_t* counts, uint32_t* off, uint32_t* local, uint32_t* entries, double* sorted,
                             T* force, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  const uint32_t nt = g.n;
  const unsigned rows = blocks_for(16 * size_t(npts));
  uint4* loc = reinterpret_cast<uint4*>(local);
  if (npts > 0)
    hipLaunchKernelGGL(k_site_count<false>, dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), g, nullptr, pts, slot_ids,
                       counts, loc);
  // per-workgroup totals of the scan after the nt + 1 offsets
  const unsigned nsb = scan_blocks(size_t(nt));
  uint32_t* sums = off + nt + 1;
  hipLaunchKernelGGL(k_scan_sums, dim3(nsb), dim3(SCAN_BLOCK), 0, s, int(nt), counts, sums);
  hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, int(nsb), sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(SCAN_BLOCK), 0, s, int(nt), counts, sums, off, nullptr);
  if (npts > 0) {
    uint2* ent = reinterpret_cast<uint2*>(entries);
    hipLaunchKernelGGL(k_site_fill<false>, dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), g, nullptr, pts, slot_ids,
                       off, loc, ent);
    const uint32_t nent = 64u * uint32_t(npts);
    hipLaunchKernelGGL(k_site_place, dim3(blocks_for(nent)), dim3(PBLOCK), 0, s, nent, pts, off, ent, sorted);
  }
  hipLaunchKernelGGL(k_site_sum<T>, dim3(blocks_for(nt)), dim3(PBLOCK), 0, s, nt, off, sorted, force);
  return hipGetLastError();
}

hipError_t launch_batch_site_lists(const LatDesc* ldesc, int npts, int nsites, int nent, const SpreadPoint* pts,
                                   const uint32_t* slot_ids, uint32_t* counts, uint32_t* off, uint32_t* local,
                                   uint32_t* entries, double* sorted, hipStream_t s) {
  if (nsites <= 0) return hipSuccess;
  const uint32_t nt = uint32_t(nsites);
  const unsigned rows = blocks_for(16 * size_t(npts));
  uint4* loc = reinterpret_cast<uint4*>(local);
  const Geom none{};
  if (npts > 0)
    hipLaunchKernelGGL(k_site_count<true>, dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), none, ldesc, pts, slot_ids,
                       counts, loc);
  const unsigned nsb = scan_blocks(size_t(nt));
  uint32_t* sums = off + nt + 1;
  hipLaunchKernelGGL(k_scan_sums, dim3(nsb), dim3(SCAN_BLOCK), 0, s, int(nt), counts, sums);
  hipLaunchKernelGGL(k_scan_block_sums, dim3(1), dim3(SCAN_BLOCK), 0, s, int(nsb), sums);
  hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(SCAN_BLOCK), 0, s, int(nt), counts, sums, off, nullptr);
  if (npts > 0) {
    uint2* ent = reinterpret_cast<uint2*>(entries);
    hipLaunchKernelGGL(k_site_fill<true>, dim3(rows), dim3(PBLOCK), 0, s, uint32_t(npts), none, ldesc, pts, slot_ids,
                       off, loc, ent);
    if (nent > 0)
      hipLaunchKernelGGL(k_site_place, dim3(blocks_for(size_t(nent))), dim3(PBLOCK), 0, s, uint32_t(nent), pts, off,
                         ent, sorted);
  }
  return hipGetLastError();
}
template hipError_t launch_site_force<double>(int, const Geom&, const SpreadPoint*, const uint32_t*, uint32_t*,
                                              uint32_t*, uint32_t*, uint32_t*, double*, double*, hipStream_t);
template hipError_t launch_site_force<float>(int, const Geom&, const SpreadPoint*, const uint32_t*, uint32_t*,
                                             uint32_t*, uint32_t*, uint32_t*, double*, float*, hipStream_t);

template <typename T>
hipError_t launch_spread_force(const Geom& g, const SpreadPoint* pts, const uint32_t* off, const uint32_t* sorted,
                               T* force, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_spread_blocks<T>, dim3((g.n + SITE_BLOCK - 1) / SITE_BLOCK), dim3(SITE_BLOCK), 0, s, g, pts, off,
                     sorted, force);
  return hipGetLastError();
}
template hipError_t launch_spread_force<double>(const Geom&, const SpreadPoint*, const uint32_t*, const uint32_t*,
                                                double*, hipStream_t);
template hipError_t launch_spread_force<float>(const Geom&, const SpreadPoint*, const uint32_t*, const uint32_t*,
                                               float*, hipStream_t);

hipError_t launch_edge_force(const Geom& g, const SpreadPoint* pts, const uint32_t* off, const uint32_t* sorted,
                             double* edge, RunEntry* rec, int* wz, hipStream_t s) {
  const uint32_t nl = spread_lists(g);
  if (nl == 0) return hipSuccess;
  const bool entries = rec && wz && g.nz % WAVE_SITES == 0;
  // k_run_entries also writes the run-end forces where every axis has a
  // single matching stencil position (add_point's simple case)
  const bool merged = entries && g.nx >= 4 && g.ny >= 4;
  if (!merged)
    hipLaunchKernelGGL(k_edge_force, dim3((2 * nl + SITE_BLOCK - 1) / SITE_BLOCK), dim3(SITE_BLOCK), 0, s, g, nl,
                       pts, off, sorted, edge);
  if (entries) {
    const uint32_t waves = (nl + ENTRY_LISTS - 1) / ENTRY_LISTS;
    const uint32_t per_block = SITE_BLOCK / WAVE_SITES;
    hipLaunchKernelGGL(k_run_entries, dim3((waves + per_block - 1) / per_block), dim3(SITE_BLOCK), 0, s, g, nl, pts,
                       off, sorted, rec, wz, merged ? edge : nullptr);
  }
  return hipGetLastError();
}

hipError_t launch_batch_run_entries(const LatDesc* ldesc, const WorkItem* work, int nwork, const SpreadPoint* pts,
                                    const uint32_t* off, const uint32_t* sorted, RunEntry* rec, int* wz, double* edge,
                                    hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_run_entries, dim3(nwork), dim3(SITE_BLOCK), 0, s, ldesc, work, pts, off, sorted, rec, wz,
                     edge);
  return hipGetLastError();
}

hipError_t launch_batch_points(const LatDesc* ldesc, const SwimDesc* sdesc, const WorkItem* work, int nwork,
                               SpreadPoint* pts, uint32_t* slot_ids, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_points, dim3(nwork), dim3(PBLOCK), 0, s, ldesc, sdesc, work, pts, slot_ids);
  return hipGetLastError();
}

// Lists of all members are numbered consecutively (member m from its
// tile_base); block_work lists (member, lattice block) for every block.
hipError_t launch_batch_lists(const LatDesc* ldesc, int npts, int nlists, const SpreadPoint* pts,
                              const uint32_t* slot_ids, bool dense, uint32_t* counts, uint32_t* off, uint32_t* cursor,
                              uint32_t* list, uint32_t* sorted, hipStream_t s) {
  if (nlists <= 0) return hipSuccess;
  return build_lists<true>(npts, nlists, Geom{}, ldesc, pts, slot_ids, dense, counts, off, cursor, list, sorted, s);
}

template <typename T>
hipError_t launch_batch_spread_force(const LatDesc* ldesc, int nwork, const WorkItem* block_work, int per,
                                     const SpreadPoint* pts, const uint32_t* off, const uint32_t* sorted,
                                     const uint32_t* soff, const double* ssorted, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_spread_blocks<T>, dim3(unsigned(nwork)), dim3(SITE_BLOCK), 0, s, ldesc, block_work, pts, off,
                     sorted, per, soff, ssorted);
  return hipGetLastError();
}
template hipError_t launch_batch_spread_force<double>(const LatDesc*, int, const WorkItem*, int, const SpreadPoint*,
                                                      const uint32_t*, const uint32_t*, const uint32_t*, const double*,
                                                      hipStream_t);
template hipError_t launch_batch_spread_force<float>(const LatDesc*, int, const WorkItem*, int, const SpreadPoint*,
                                                     const uint32_t*, const uint32_t*, const uint32_t*, const double*,
                                                     hipStream_t);

template <typename T>
hipError_t launch_batch_swimmers_move(const LatDesc* ldesc, const SwimDesc* sdesc, const WorkItem* work, int nwork,
                                      hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_swimmers_move<T>, dim3(round_xcd(uint32_t(nwork))), dim3(PBLOCK), 0, s, ldesc, sdesc, work,
                     uint32_t(nwork));
  return hipGetLastError();
}
template <typename T>
hipError_t launch_batch_tracers_move(const LatDesc* ldesc, const TracerDesc* tdesc, const WorkItem* work, int nwork,
                                     hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  hipLaunchKernelGGL(kb_tracers_move<T>, dim3(round_xcd(uint32_t(nwork))), dim3(TBLOCK), 0, s, ldesc, tdesc, work,
                     uint32_t(nwork));
  return hipGetLastError();
}
template hipError_t launch_batch_swimmers_move<double>(const LatDesc*, const SwimDesc*, const WorkItem*, int,
                                                       hipStream_t);
template hipError_t launch_batch_swimmers_move<float>(const LatDesc*, const SwimDesc*, const WorkItem*, int,
                                                      hipStream_t);
template hipError_t launch_batch_tracers_move<double>(const LatDesc*, const TracerDesc*, const WorkItem*, int,
                                                      hipStream_t);
template hipError_t launch_batch_tracers_move<float>(const LatDesc*, const TracerDesc*, const WorkItem*, int,
                                                     hipStream_t);

hipError_t launch_peskin_probe(const double* x, double* out, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_peskin_probe, dim3(blocks_for(size_t(count))), dim3(PBLOCK), 0, s, x, out, count);
  return hipGetLastError();
}

hipError_t launch_interp_probe(const double* u, const double* r, double* v, int count, const Geom& g,
                               hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_interp_probe, dim3(blocks_for(size_t(count))), dim3(PBLOCK), 0, s, u, r, v, count, g);
  return hipGetLastError();
}

}  // namespace lbs
