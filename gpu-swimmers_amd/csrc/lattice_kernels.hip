// Lattice kernels for gfx950: collide+stream, CalcHydro, InitFromHydro,
// ZeroForce and FP64 conservation totals.
//
// Layout (unchanged from the reference, lb.cu:49-57,117): site
// ijk = i*ny*nz + j*nz + k, distributions f[p*n + ijk], vectors u[a*n + ijk].
// One lane per site over the linear site index, so the 64 lanes of a
// wavefront read and write 64 consecutive sites of each plane p: fully
// coalesced for every lattice shape (the reference maps threadIdx.x to x, the
// slowest axis, lb.cu:75-88).  Everything is bandwidth-bound, so the D3Q15
// transforms are unrolled against compile-time tables and stay in registers.
//
// Streaming is done by pulling: the device buffers hold post-collision
// distributions g, and the post-stream value the reference stores
// (lb.cu:206-216) is f_p(x) = g_p(x - c_p).  The collide kernel reads g at
// x - c_p (shifted reads hit lines already in L2) and writes its own site
// aligned with non-temporal stores; on MI355X this runs ~19 % faster than
// pushing to x + c_p, whose misaligned stores split cache lines.  CalcHydro
// pulls the same way, InitFromHydro writes its equilibrium to x - c_p, and
// the host side shifts on H2D/D2H so the API sees post-stream fields.
//
// Two storage forms:
//   double -- the reference's arithmetic, operation for operation, so that
//             results match a CPU restatement bit for bit.  The only changes
//             are dropped terms whose coefficient is exactly zero and the
//             ghost modes that the reference zeroes before use.  Build with
//             -ffp-contract=off.
//   float  -- distributions stored as deviations g_p = f_p - w_p from the
//             rest state (rho=1, u=0), density as rho-1, collided in FP32
//             with an algebraically simplified transform.
#include <hip/hip_runtime.h>

#include <utility>

#include "d3q15.hpp"
#include "internal.hpp"
#include "spread.hpp"
#include "xcd.hpp"

namespace lbs {

Geom make_geom(int nx, int ny, int nz) {
  Geom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.n = uint32_t(nx) * uint32_t(ny) * uint32_t(nz);
  g.nyz = uint32_t(ny) * uint32_t(nz);
  g.div_nz = make_fastdiv(uint32_t(nz));
  g.div_ny = make_fastdiv(uint32_t(ny));
  return g;
}

namespace {

constexpr int BLOCK = 256;
static_assert(BLOCK == SITE_BLOCK, "force lists are numbered by lattice workgroup");
__host__ __device__ inline uint32_t grid_blocks(uint32_t n) { return (n + BLOCK - 1) / BLOCK; }

struct Site {
  uint32_t i, j, k;
};

__device__ __forceinline__ Site decompose(uint32_t s, const Geom& g) {
  const uint32_t t = fastdiv(s, g.div_nz);
  const uint32_t k = s - t * uint32_t(g.nz);
  const uint32_t i = fastdiv(t, g.div_ny);
  const uint32_t j = t - i * uint32_t(g.ny);
  return Site{i, j, k};
}

// Wrapped offsets of the x, y, z neighbours of a site for c = -1, 0, 1
// (periodic boundaries, lb.cu:210-214).
struct Nbr {
  uint32_t ox[3], oy[3], oz[3];
};

__device__ __forceinline__ Nbr neighbours(Site s, const Geom& g) {
  Nbr d;
  const uint32_t im = s.i == 0 ? uint32_t(g.nx) - 1 : s.i - 1;
  const uint32_t ip = s.i + 1 == uint32_t(g.nx) ? 0 : s.i + 1;
  const uint32_t jm = s.j == 0 ? uint32_t(g.ny) - 1 : s.j - 1;
  const uint32_t jp = s.j + 1 == uint32_t(g.ny) ? 0 : s.j + 1;
  const uint32_t km = s.k == 0 ? uint32_t(g.nz) - 1 : s.k - 1;
  const uint32_t kp = s.k + 1 == uint32_t(g.nz) ? 0 : s.k + 1;
  d.ox[0] = im * g.nyz;
  d.ox[1] = s.i * g.nyz;
  d.ox[2] = ip * g.nyz;
  d.oy[0] = jm * uint32_t(g.nz);
  d.oy[1] = s.j * uint32_t(g.nz);
  d.oy[2] = jp * uint32_t(g.nz);
  d.oz[0] = km;
  d.oz[1] = s.k;
  d.oz[2] = kp;
  return d;
}

// The upstream site x - c_P.
template <int P>
__device__ __forceinline__ uint32_t upstream(const Nbr& d) {
  return d.ox[1 - CI[P][0]] + d.oy[1 - CI[P][1]] + d.oz[1 - CI[P][2]];
}

template <typename T>
__device__ __forceinline__ void store_stream(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// ------------------------------------------------------------------------
// FP64: the reference's arithmetic.
// ------------------------------------------------------------------------

// mode[M] = sum_p f[p]*mm[M][p] in p order, zero terms skipped (lb.cu:114-119).
template <int M, int P>
__device__ __forceinline__ void fwd_term(double& acc, const double (&f)[Q]) {
  constexpr double c = TAB.mm[M][P];
  if constexpr (c != 0.0) {
    if constexpr (P == first_nz_mm(M))
      acc = f[P] * c;
    else
      acc = acc + f[P] * c;
  }
}
template <int M, int... P>
__device__ __forceinline__ double fwd_row(const double (&f)[Q], std::integer_sequence<int, P...>) {
  double acc = 0.0;
  (fwd_term<M, P>(acc, f), ...);
  return acc;
}
template <int M>
__device__ __forceinline__ double fwd(const double (&f)[Q]) {
  return fwd_row<M>(f, std::make_integer_sequence<int, Q>{});
}

// f[P] = sum_m mode[m]*mmi[P][m] in m order over the kept modes (lb.cu:198-203).
template <int P, int M>
__device__ __forceinline__ void inv_term(double& acc, const double (&mode)[NKEEP]) {
  constexpr double c = TAB.mmi[P][M];
  if constexpr (c != 0.0) {
    if constexpr (M == first_nz_mmi(P))
      acc = mode[M] * c;
    else
      acc = acc + mode[M] * c;
  }
}
template <int P, int... M>
__device__ __forceinline__ double inv_row(const double (&mode)[NKEEP], std::integer_sequence<int, M...>) {
  double acc = 0.0;
  (inv_term<P, M>(acc, mode), ...);
  return acc;
}
template <int P>
__device__ __forceinline__ double inv(const double (&mode)[NKEEP]) {
  return inv_row<P>(mode, std::make_integer_sequence<int, NKEEP>{});
}

template <int... M>
__device__ __forceinline__ void fwd_all(const double (&f)[Q], double (&mode)[NKEEP], std::integer_sequence<int, M...>) {
  ((mode[M] = fwd<M>(f)), ...);
}

// MRT relaxation with forcing, lb.cu:127-195, in the same operation order.
// Terms multiplied by an off-diagonal DQ_delta (exactly 0) are dropped:
// x - usq*0.0 and x - 2*uF*0.0 equal x.
__device__ __forceinline__ void relax_f64(double (&mode)[NKEEP], const double (&F)[3], double tau_s,
                                          double tau_b, double omega_s, double omega_b) {
  double u[3];
  double usq = 0., uF = 0.;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    u[a] = (mode[1 + a] + 0.5 * F[a]) / mode[M_RHO];
    mode[1 + a] += F[a];
    usq += u[a] * u[a];
    uF += u[a] * F[a];
  }
  double Sxx = mode[M_SXX], Sxy = mode[M_SXY], Sxz = mode[M_SXZ];
  double Syy = mode[M_SYY], Syz = mode[M_SYZ], Szz = mode[M_SZZ];
  double TrS = 0.;
  TrS += Sxx;
  TrS += Syy;
  TrS += Szz;
  Sxx -= TrS / 3;
  Syy -= TrS / 3;
  Szz -= TrS / 3;
  TrS -= omega_b * (TrS - mode[M_RHO] * usq);
  TrS += 2. * omega_b * tau_b * uF;
  const double rho = mode[M_RHO];
  const double cf = 2. * omega_s * tau_s;
  // a = x
  Sxx -= omega_s * (Sxx - rho * (u[0] * u[0] - usq));
  Sxx += cf * (u[0] * F[0] + F[0] * u[0] - 2. * uF);
  Sxy -= omega_s * (Sxy - rho * (u[0] * u[1]));
  Sxy += cf * (u[0] * F[1] + F[0] * u[1]);
  Sxz -= omega_s * (Sxz - rho * (u[0] * u[2]));
  Sxz += cf * (u[0] * F[2] + F[0] * u[2]);
  Sxx += TrS / 3;
  // a = y (the yx element is recomputed by the reference but never used)
  Syy -= omega_s * (Syy - rho * (u[1] * u[1] - usq));
  Syy += cf * (u[1] * F[1] + F[1] * u[1] - 2. * uF);
  Syz -= omega_s * (Syz - rho * (u[1] * u[2]));
  Syz += cf * (u[1] * F[2] + F[1] * u[2]);
  Syy += TrS / 3;
  // a = z
  Szz -= omega_s * (Szz - rho * (u[2] * u[2] - usq));
  Szz += cf * (u[2] * F[2] + F[2] * u[2] - 2. * uF);
  Szz += TrS / 3;
  mode[M_SXX] = Sxx;
  mode[M_SXY] = Sxy;
  mode[M_SXZ] = Sxz;
  mode[M_SYY] = Syy;
  mode[M_SYZ] = Syz;
  mode[M_SZZ] = Szz;
}

// Post-stream distributions at a site: f_p = g_p(x - c_p).
template <typename T, int... P>
__device__ __forceinline__ void pull_all(const T* __restrict__ g, uint32_t n, const Nbr& d, T (&out)[Q],
                                         std::integer_sequence<int, P...>) {
  ((out[P] = g[size_t(P) * n + upstream<P>(d)]), ...);
}

template <int... P>
__device__ __forceinline__ void store_all(double* __restrict__ gNew, uint32_t n, uint32_t s,
                                          const double (&mode)[NKEEP], std::integer_sequence<int, P...>) {
  (store_stream(&gNew[size_t(P) * n + s], inv<P>(mode)), ...);
}

// Stored force from per-site segments of ordered products (segment_force):
// the collide sums its site's segment (site sbase + s) and stores the force
// where the CalcHydro reads it.  off null: the force array is read.
struct SegForce {
  const uint32_t* off = nullptr;
  const double* sorted = nullptr;
  uint32_t sbase = 0;
  void* out = nullptr;  // the member's force array
};

// DoStep (lb.cu:72-218): pull-stream, then collide into the site's own slots.
// VF: the force comes from the wavefront's point list (wave_force; lists
// vlist.. of the block) instead of the force array.  No barrier is involved,
// so lanes past the lattice end leave at once.
template <bool VF>
__device__ __forceinline__ void collide_f64_body(const double* __restrict__ gOld, const double* __restrict__ force,
                                                 double* __restrict__ gNew, Geom g, double tau_s, double tau_b,
                                                 double omega_s, double omega_b, uint32_t blk, const VForce& vf,
                                                 uint32_t vlist, const SegForce& sf = SegForce{}) {
  const uint32_t s = blk * BLOCK + threadIdx.x;
  if (s >= g.n) return;
  const uint32_t n = g.n;
  const Site c = decompose(s, g);
  const Nbr d = neighbours(c, g);
  double f[Q];
  pull_all(gOld, n, d, f, std::make_integer_sequence<int, Q>{});
  double F[3];
  // the list walk uses scalar loads only, so it overlaps the loads of f
  if constexpr (VF)
    wave_force(vf.pts, vf.off, vf.sorted, vlist + threadIdx.x / WAVE_SITES, true, int(c.i), int(c.j), int(c.k), g, F);
  else if (sf.off) {
    segment_force(sf.off, sf.sorted, sf.sbase + s, F);
#pragma unroll
    for (int a = 0; a < 3; ++a) static_cast<double*>(sf.out)[size_t(a) * n + s] = F[a];
  } else {
    F[0] = force[s];
    F[1] = force[size_t(n) + s];
    F[2] = force[2 * size_t(n) + s];
  }
  double mode[NKEEP];
  fwd_all(f, mode, std::make_integer_sequence<int, NKEEP>{});
  relax_f64(mode, F, tau_s, tau_b, omega_s, omega_b);
  store_all(gNew, n, s, mode, std::make_integer_sequence<int, Q>{});
}

// lb.cu:250-295: only the density and momentum modes are needed.  aos: u
// goes component-interleaved, u[3 s + a] (the copy the particle moves gather
// from), instead of u[a n + s].
template <bool VF>
__device__ __forceinline__ void hydro_f64_body(const double* __restrict__ gCur, const double* __restrict__ force,
                                               double* __restrict__ rho, double* __restrict__ u, Geom g, uint32_t blk,
                                               const VForce& vf, uint32_t vlist, bool aos) {
  const uint32_t s = blk * BLOCK + threadIdx.x;
  if (s >= g.n) return;
  const uint32_t n = g.n;
  const Site c = decompose(s, g);
  double fl[Q];
  pull_all(gCur, n, neighbours(c, g), fl, std::make_integer_sequence<int, Q>{});
  double F[3];
  if constexpr (VF)
    wave_force(vf.pts, vf.off, vf.sorted, vlist + threadIdx.x / WAVE_SITES, true, int(c.i), int(c.j), int(c.k), g, F);
  const double m0 = fwd<0>(fl);
  const double m1 = fwd<1>(fl);
  const double m2 = fwd<2>(fl);
  const double m3 = fwd<3>(fl);
  if constexpr (!VF) {
    F[0] = force[s];
    F[1] = force[size_t(n) + s];
    F[2] = force[2 * size_t(n) + s];
  }
  const double uv[3] = {(m1 + 0.5 * F[0]) / m0, (m2 + 0.5 * F[1]) / m0, (m3 + 0.5 * F[2]) / m0};
  // streaming stores: the fields are far larger than the caches (the
  // interleaved u through the caches, where L2 merges the strided stores)
  store_stream(&rho[s], m0);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (aos)
      u[3 * size_t(s) + a] = uv[a];
    else
      store_stream(&u[size_t(a) * n + s], uv[a]);
  }
}

// lb.cu:314-371, including the reference's SXZ = rho*ux*uy (lb.cu:356).
__global__ __launch_bounds__(BLOCK) void k_init_from_hydro_f64(const double* __restrict__ rho,
                                                               const double* __restrict__ u,
                                                               double* __restrict__ f, Geom g) {
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= g.n) return;
  const uint32_t n = g.n;
  const double r = rho[s];
  const double ux = u[s], uy = u[size_t(n) + s], uz = u[2 * size_t(n) + s];
  double mode[NKEEP];
  mode[0] = r;
  mode[1] = r * ux;
  mode[2] = r * uy;
  mode[3] = r * uz;
  mode[M_SXX] = r * ux * ux;
  mode[M_SXY] = r * ux * uy;
  mode[M_SXZ] = r * ux * uy;
  mode[M_SYY] = r * uy * uy;
  mode[M_SYZ] = r * uy * uz;
  mode[M_SZZ] = r * uz * uz;
  const Nbr d = neighbours(decompose(s, g), g);
  [&]<int... P>(std::integer_sequence<int, P...>) {
    ((f[size_t(P) * n + upstream<P>(d)] = inv<P>(mode)), ...);
  }(std::make_integer_sequence<int, Q>{});
}

// ------------------------------------------------------------------------
// FP32 deviation form.
// ------------------------------------------------------------------------

struct Moments32 {
  float drho, j[3], S[6];  // S: xx, xy, xz, yy, yz, zz
};

__device__ __forceinline__ Moments32 moments_dev32(const float (&g)[Q]) {
  Moments32 m;
  const float cpx = (g[7] + g[8]) + (g[9] + g[10]);     // cube velocities with c_x = +1
  const float cmx = (g[11] + g[12]) + (g[13] + g[14]);  // c_x = -1
  const float cube = cpx + cmx;
  m.drho = ((g[0] + (g[1] + g[2])) + ((g[3] + g[4]) + (g[5] + g[6]))) + cube;
  m.j[0] = (g[1] - g[2]) + (cpx - cmx);
  m.j[1] = (g[3] - g[4]) + (((g[7] + g[8]) + (g[11] + g[12])) - ((g[9] + g[10]) + (g[13] + g[14])));
  m.j[2] = (g[5] - g[6]) + (((g[7] + g[9]) + (g[11] + g[13])) - ((g[8] + g[10]) + (g[12] + g[14])));
  const float third = m.drho * (1.0f / 3.0f);
  m.S[0] = ((g[1] + g[2]) + cube) - third;
  m.S[3] = ((g[3] + g[4]) + cube) - third;
  m.S[5] = ((g[5] + g[6]) + cube) - third;
  m.S[1] = ((g[7] + g[8]) + (g[13] + g[14])) - ((g[9] + g[10]) + (g[11] + g[12]));
  m.S[2] = ((g[7] + g[9]) + (g[12] + g[14])) - ((g[8] + g[10]) + (g[11] + g[13]));
  m.S[4] = ((g[7] + g[10]) + (g[11] + g[14])) - ((g[8] + g[9]) + (g[12] + g[13]));
  return m;
}

// Same relaxation as relax_f64 (lb.cu:127-195, including its usq and 2*uF
// diagonal terms), on rho = 1 + drho, with fused multiply-adds.
__device__ __forceinline__ void relax_dev32(Moments32& m, const float (&F)[3], float tau_s, float tau_b,
                                            float omega_s, float omega_b) {
  const float rho = 1.0f + m.drho;
  const float inv_rho = 1.0f / rho;
  float u[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    u[a] = __builtin_fmaf(0.5f, F[a], m.j[a]) * inv_rho;
    m.j[a] += F[a];
  }
  const float usq = __builtin_fmaf(u[0], u[0], __builtin_fmaf(u[1], u[1], u[2] * u[2]));
  const float uF = __builtin_fmaf(u[0], F[0], __builtin_fmaf(u[1], F[1], u[2] * F[2]));
  float* S = m.S;
  float TrS = (S[0] + S[3]) + S[5];
  const float t3 = TrS * (1.0f / 3.0f);
  S[0] -= t3;
  S[3] -= t3;
  S[5] -= t3;
  TrS = __builtin_fmaf(-omega_b, TrS - rho * usq, TrS);
  TrS = __builtin_fmaf(2.0f * omega_b * tau_b, uF, TrS);
  const float cf = 2.0f * omega_s * tau_s;
  const float nt3 = TrS * (1.0f / 3.0f);
  constexpr int DI[3] = {0, 3, 5};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float& s = S[DI[a]];
    s = __builtin_fmaf(-omega_s, s - rho * __builtin_fmaf(u[a], u[a], -usq), s);
    s = __builtin_fmaf(cf, 2.0f * __builtin_fmaf(u[a], F[a], -uF), s);
    s += nt3;
  }
  constexpr int OFF[3][3] = {{1, 0, 1}, {2, 0, 2}, {4, 1, 2}};  // slot, a, b
#pragma unroll
  for (int o = 0; o < 3; ++o) {
    float& s = S[OFF[o][0]];
    const int a = OFF[o][1], b = OFF[o][2];
    s = __builtin_fmaf(-omega_s, s - rho * (u[a] * u[b]), s);
    s = __builtin_fmaf(cf, __builtin_fmaf(u[a], F[b], F[a] * u[b]), s);
  }
}

// g_p = w_p (drho - 3/2 T + 3 c.j + 9/2 (sum_a c_a^2 S_aa + 2 sum_{a<b} c_a c_b S_ab)),
// the kept-mode part of mmi (d3q15.c:171-185) with mode 0 replaced by drho.
__device__ __forceinline__ void inverse_dev32(const Moments32& m, float (&g)[Q]) {
  constexpr float W0 = 2.0f / 9.0f, W1 = 1.0f / 9.0f, W2 = 1.0f / 72.0f;
  const float* S = m.S;
  const float T = (S[0] + S[3]) + S[5];
  const float base = __builtin_fmaf(-1.5f, T, m.drho);
  g[0] = W0 * base;
  const float bx = __builtin_fmaf(4.5f, S[0], base), by = __builtin_fmaf(4.5f, S[3], base),
              bz = __builtin_fmaf(4.5f, S[5], base);
  const float jx3 = 3.0f * m.j[0], jy3 = 3.0f * m.j[1], jz3 = 3.0f * m.j[2];
  g[1] = W1 * (bx + jx3);
  g[2] = W1 * (bx - jx3);
  g[3] = W1 * (by + jy3);
  g[4] = W1 * (by - jy3);
  g[5] = W1 * (bz + jz3);
  g[6] = W1 * (bz - jz3);
  const float A = __builtin_fmaf(3.0f, T, m.drho);
  const float sxy = 9.0f * S[1], sxz = 9.0f * S[2], syz = 9.0f * S[4];
  // pairs (p, complement p): 7/14 (+,+,+), 8/13 (+,+,-), 9/12 (+,-,+), 10/11 (+,-,-)
  const float s7 = A + (sxy + sxz + syz), t7 = jx3 + jy3 + jz3;
  const float s8 = A + (sxy - sxz - syz), t8 = jx3 + jy3 - jz3;
  const float s9 = A + (-sxy + sxz - syz), t9 = jx3 - jy3 + jz3;
  const float s10 = A + (-sxy - sxz + syz), t10 = jx3 - jy3 - jz3;
  g[7] = W2 * (s7 + t7);
  g[14] = W2 * (s7 - t7);
  g[8] = W2 * (s8 + t8);
  g[13] = W2 * (s8 - t8);
  g[9] = W2 * (s9 + t9);
  g[12] = W2 * (s9 - t9);
  g[10] = W2 * (s10 + t10);
  g[11] = W2 * (s10 - t10);
}

// The force is rounded to FP32 as the stored force array holds it.
template <bool VF>
__device__ __forceinline__ void collide_f32_body(const float* __restrict__ fOld, const float* __restrict__ force,
                                                 float* __restrict__ fNew, Geom g, float tau_s, float tau_b,
                                                 float omega_s, float omega_b, uint32_t blk, const VForce& vf,
                                                 uint32_t vlist, const SegForce& sf = SegForce{}) {
  const uint32_t s = blk * BLOCK + threadIdx.x;
  if (s >= g.n) return;
  const uint32_t n = g.n;
  const Site c = decompose(s, g);
  float gl[Q];
  pull_all(fOld, n, neighbours(c, g), gl, std::make_integer_sequence<int, Q>{});
  float F[3];
  if constexpr (VF) {
    double acc[3];
    wave_force(vf.pts, vf.off, vf.sorted, vlist + threadIdx.x / WAVE_SITES, true, int(c.i), int(c.j), int(c.k), g,
               acc);
#pragma unroll
    for (int a = 0; a < 3; ++a) F[a] = float(acc[a]);
  } else if (sf.off) {
    double acc[3];
    segment_force(sf.off, sf.sorted, sf.sbase + s, acc);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      F[a] = float(acc[a]);
      static_cast<float*>(sf.out)[size_t(a) * n + s] = F[a];
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) F[a] = force[size_t(a) * n + s];
  }
  Moments32 m = moments_dev32(gl);
  relax_dev32(m, F, tau_s, tau_b, omega_s, omega_b);
  float out[Q];
  inverse_dev32(m, out);
#pragma unroll
  for (int p = 0; p < Q; ++p) store_stream(&fNew[size_t(p) * n + s], out[p]);
}

template <bool VF>
__device__ __forceinline__ void hydro_f32_body(const float* __restrict__ f, const float* __restrict__ force,
                                               float* __restrict__ drho_out, float* __restrict__ u, Geom g,
                                               uint32_t blk, const VForce& vf, uint32_t vlist, bool aos) {
  const uint32_t s = blk * BLOCK + threadIdx.x;
  if (s >= g.n) return;
  const uint32_t n = g.n;
  const Site c = decompose(s, g);
  float gl[Q];
  pull_all(f, n, neighbours(c, g), gl, std::make_integer_sequence<int, Q>{});
  float F[3];
  if constexpr (VF) {
    double acc[3];
    wave_force(vf.pts, vf.off, vf.sorted, vlist + threadIdx.x / WAVE_SITES, true, int(c.i), int(c.j), int(c.k), g,
               acc);
#pragma unroll
    for (int a = 0; a < 3; ++a) F[a] = float(acc[a]);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) F[a] = force[size_t(a) * n + s];
  }
  const Moments32 m = moments_dev32(gl);
  const float inv_rho = 1.0f / (1.0f + m.drho);
  store_stream(&drho_out[s], m.drho);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float ua = __builtin_fmaf(0.5f, F[a], m.j[a]) * inv_rho;
    if (aos)
      u[3 * size_t(s) + a] = ua;
    else
      store_stream(&u[size_t(a) * n + s], ua);
  }
}

__global__ __launch_bounds__(BLOCK) void k_init_from_hydro_f32(const float* __restrict__ drho,
                                                               const float* __restrict__ u,
                                                               float* __restrict__ f, Geom g) {
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= g.n) return;
  const uint32_t n = g.n;
  Moments32 m;
  m.drho = drho[s];
  const float r = 1.0f + m.drho;
  const float ux = u[s], uy = u[size_t(n) + s], uz = u[2 * size_t(n) + s];
  m.j[0] = r * ux;
  m.j[1] = r * uy;
  m.j[2] = r * uz;
  // The stress modes as the reference sets them, SXZ = rho*ux*uy included
  // (lb.cu:354-361); the rest state has zero stress, so no offset applies.
  const float sxx = r * ux * ux, syy = r * uy * uy, szz = r * uz * uz;
  m.S[0] = sxx;
  m.S[1] = r * ux * uy;
  m.S[2] = r * ux * uy;
  m.S[3] = syy;
  m.S[4] = r * uy * uz;
  m.S[5] = szz;
  float out[Q];
  inverse_dev32(m, out);
  const Nbr d = neighbours(decompose(s, g), g);
  [&]<int... P>(std::integer_sequence<int, P...>) {
    ((f[size_t(P) * n + upstream<P>(d)] = out[P]), ...);
  }(std::make_integer_sequence<int, Q>{});
}

// ------------------------------------------------------------------------
// Fused step: collide+stream (lb.cu:72-218) and CalcHydro (lb.cu:250-295) in
// one pass, for lattices with nz % 64 == 0 and ny % TY == 0.
//
// CalcHydro at site x pulls g'_p(x - c_p), the values the collision has just
// written around x.  A workgroup owns a tile of TY rows (y) by 64 sites (z)
// and sweeps it along x over a strip of planes [xa, xb).  Row waves 1..TY
// collide the tile's own sites; row waves 0 and TY+1 (the rows above and
// below) and the first 2 (TY+2) lanes of one more wave (the sites just before
// and after each row in z) collide the ring of neighbour sites; planes xa-1
// and xb are collided as well.  Each collided value that a tile site pulls
// goes to LDS at that target site's index, in ring slots by c_x, and
// CalcHydro of plane x-1 runs once plane x is collided: it reads the c_x = -1
// values of plane x, the c_x = 0 values of x-1 and the c_x = +1 values of
// x-2.  Neighbour sites are computed by the same code from the same inputs as
// their owners compute them, so the distributions and rho/u equal the
// two-kernel results bit for bit, and the distributions are read from HBM
// once per step instead of twice.
// ------------------------------------------------------------------------

constexpr int FUSED_Z = 64;  // sites of a tile along z: one wavefront

// Directions grouped by c_x (5 each): group 0 c_x = -1, 1 c_x = 0, 2 c_x = +1,
// and the index of p within its group.
__host__ __device__ constexpr int dir_group(int p) { return CI[p][0] + 1; }
__host__ __device__ constexpr int dir_index(int p) {
  int k = 0;
  for (int q = 0; q < p; ++q) k += CI[q][0] == CI[p][0] ? 1 : 0;
  return k;
}
// LDS ring slots.  Six slot-sets (FP64): group 0 slot 0 (read in the
// iteration that writes it), group 1 slots 1-2 (read one iteration later),
// group 2 slots 3-5 (two later); a second barrier per plane keeps the next
// plane's writes off the slots CalcHydro is still reading.  Nine (FP32, whose
// LDS has room): groups 0, 1, 2 rotate over 2, 3, 4 slots (0-1, 2-4, 5-8), so
// the next plane writes other slots and one barrier per plane suffices (a
// wavefront cannot pass the next plane's barrier before every wavefront has
// read this plane's slots).  The slot-set after the ring takes the unused
// writes of the z-halo wave.
template <typename T>
constexpr int fused_slots() { return sizeof(T) == 4 ? 9 : 6; }

// Tiles of TY rows by `width` z sites (64, or nz below 64).  Where TY does
// not divide ny, or 64 does not divide nz (>= 64), the last tile along that
// axis is moved back to end at the lattice edge and overlaps its neighbour:
// the sites both compute get the same values from the same inputs, so both
// stores write the same bits.
struct FusedTiles {
  uint32_t nyt, nzt;  // tiles along y and z
  uint32_t strip;     // planes per strip along x
  uint32_t nblk;      // tiles x strips
  int width;          // z sites of a tile
};

template <typename T>
__device__ __forceinline__ void collide_site(const T (&f)[Q], const T (&F)[3], T tau_s, T tau_b, T omega_s,
                                             T omega_b, T (&out)[Q]) {
  if constexpr (sizeof(T) == 8) {
    double mode[NKEEP];
    fwd_all(f, mode, std::make_integer_sequence<int, NKEEP>{});
    relax_f64(mode, F, tau_s, tau_b, omega_s, omega_b);
    [&]<int... P>(std::integer_sequence<int, P...>) {
      ((out[P] = inv<P>(mode)), ...);
    }(std::make_integer_sequence<int, Q>{});
  } else {
    Moments32 m = moments_dev32(f);
    relax_dev32(m, F, tau_s, tau_b, omega_s, omega_b);
    inverse_dev32(m, out);
  }
}

// rho (FP32: rho - 1) and u of pulled distributions fl, as hydro_*_body.
template <typename T>
__device__ __forceinline__ void hydro_site(const T (&fl)[Q], const T (&F)[3], T& rho, T (&u)[3]) {
  if constexpr (sizeof(T) == 8) {
    const double m0 = fwd<0>(fl);
    const double m1 = fwd<1>(fl);
    const double m2 = fwd<2>(fl);
    const double m3 = fwd<3>(fl);
    rho = m0;
    u[0] = (m1 + 0.5 * F[0]) / m0;
    u[1] = (m2 + 0.5 * F[1]) / m0;
    u[2] = (m3 + 0.5 * F[2]) / m0;
  } else {
    const Moments32 m = moments_dev32(fl);
    const float inv_rho = 1.0f / (1.0f + m.drho);
    rho = m.drho;
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = __builtin_fmaf(0.5f, F[a], m.j[a]) * inv_rho;
  }
}

// Barrier for LDS only: pending global loads (the next plane's pulls) and
// stores stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Wavefront roles in a fused tile.  Each role runs its own copy of the sweep,
// so that every load and store inside the loop is unconditional (the wait
// counts then let the next plane's pulls stay in flight) and the collision
// computes only the outputs the role stores.
enum FusedRole { ROLE_TILE, ROLE_ABOVE, ROLE_BELOW, ROLE_ZHALO };

template <typename T>
struct FusedArgs {
  const T* gOld;
  const T* force;
  VForce vf;
  T* gNew;
  T* rho;
  T* u_aos;  // u with a site's components together, u_aos[3 ijk + a]
  // stored force only: where non-null, the tile's sites of this force array
  // are cleared (for the next step's scatter into it)
  T* zero_next;
  Geom g;
  T tau_s, tau_b, omega_s, omega_b;
};

// LDS rows hold tile z = -1..64 (66 words), so a row wavefront writes every
// z-shifted value without a lane test; the two end words are never read.
constexpr int FUSED_ROW = FUSED_Z + 2;

// Buffer loads and stores for the fused sweep: one resource per array, a
// wavefront-uniform byte offset (SGPR) and a 32-bit per-lane byte offset, so
// the 15 pull addresses cost 3 VGPRs.  Needs every offset below 4 GiB
// (fused_step_supported checks Q n sizeof(T) < 2^32).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, 0xFFFFFFFFu, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// Store through the caches (L2 merges the partial lines of strided stores).
template <typename T>
__device__ __forceinline__ void buffer_store(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
}
// Non-temporal store (cache policy bit 1: nt).
template <typename T>
__device__ __forceinline__ void buffer_store_nt(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, voff, soff, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 2);
}

// One wavefront's sweep of the tile.  r: tile row (0..TY+1; uniform except in
// the z-halo wave), zz: tile z (-1..W), (y, z) the lattice row and z.
// PARTIAL (nz < 64, tiles W = nz sites wide): lanes with `active` false take
// a valid site (zz = W - 1) and write nothing.
template <typename T, bool VF, int TY, int ROLE, bool PARTIAL>
__device__ __forceinline__ void fused_sweep(const FusedArgs<T>& A, T* __restrict__ lds, int xa, int last, int r,
                                            int zz, bool active, int y, int z, int W) {
  constexpr int PLANE = TY * FUSED_ROW;  // LDS words of one direction slot
  constexpr int SLOTS = fused_slots<T>(), TRASH = SLOTS * 5;
  constexpr bool TILE = ROLE == ROLE_TILE;
  const Geom& g = A.g;
  const uint32_t n = g.n, nz = uint32_t(g.nz), nyz = g.nyz;
  const int lane = int(threadIdx.x) % 64;
  constexpr uint32_t B = sizeof(T);
  const __amdgpu_buffer_rsrc_t rOld = buffer_rsrc(A.gOld), rNew = buffer_rsrc(A.gNew);
  // row offsets: uniform in row wavefronts (part of the SGPR offset), per lane
  // in the z-halo wavefront (part of the lane offset)
  uint32_t oy[3], oz[3];
  oy[0] = uint32_t(y == 0 ? g.ny - 1 : y - 1) * nz;
  oy[1] = uint32_t(y) * nz;
  oy[2] = uint32_t(y + 1 == g.ny ? 0 : y + 1) * nz;
  oz[0] = uint32_t(z == 0 ? g.nz - 1 : z - 1);
  oz[1] = uint32_t(z);
  oz[2] = uint32_t(z + 1 == g.nz ? 0 : z + 1);
  constexpr bool ZH = ROLE == ROLE_ZHALO;
  // Row wavefronts: y and the tile's first z are wavefront-uniform, and the
  // buffer instructions need them in SGPRs (a VGPR offset would be looped
  // over lane by lane).
  if constexpr (!ZH) {
#pragma unroll
    for (int i = 0; i < 3; ++i) oy[i] = uint32_t(__builtin_amdgcn_readfirstlane(int(oy[i])));
  }
  // site = srow (x, y and the tile's first z: uniform) + zl (lane part)
  const uint32_t zrow = ZH ? 0u : uint32_t(__builtin_amdgcn_readfirstlane(z - zz));
  const uint32_t zl = ZH ? oy[1] + oz[1] : uint32_t(zz);
  // lanes that store (every lane of a row wavefront unless PARTIAL)
  const bool put = ZH || !PARTIAL || active;
  const uint32_t yrow = ZH ? 0u : oy[1];
  auto plane = [&](int it) {
    int x = (xa - 1 + it) % g.nx;  // the list pipeline looks up to two planes past the strip
    return x < 0 ? x + g.nx : x;
  };
  // Virtual force of a row wavefront: the row's list for plane it is walked
  // four lanes per entry (quad_row_force) from entries loaded ahead -- list
  // bounds (offv) two planes before, the entries' RunEntry records (rec)
  // during the previous plane -- so each dependent load has at least half a
  // plane's work to complete behind.
  constexpr bool VFROW = VF && !ZH;
  const bool simple = g.nx >= 4 && g.ny >= 4;
  // offv[p & 1]: bounds of plane p (lane 0 first entry, lane 1 end), loaded
  // two iterations before take_list reads them
  uint32_t offv[2] = {0, 0}, kv = 0, o0 = 0, m0 = 0, o1 = 0, m1 = 0;
  QuadEntry rec{};
  auto list_of = [&](int it) { return (uint32_t(plane(it)) * nyz + yrow + zrow) / WAVE_SITES; };
  auto load_off = [&](int it) {
    const uint32_t v = A.vf.off[list_of(it) + uint32_t(lane & 1)];
    if (it & 1)
      offv[1] = v;
    else
      offv[0] = v;
  };
  auto take_list = [&](int it, uint32_t& o, uint32_t& m) {
    const uint32_t v = (it & 1) ? offv[1] : offv[0];
    o = uint32_t(__builtin_amdgcn_readlane(int(v), 0));
    m = uint32_t(__builtin_amdgcn_readlane(int(v), 1)) - o;
    kv = uint32_t(lane >> 2) < m ? o + uint32_t(lane >> 2) : 0u;  // four lanes per entry
  };
  // pulls of plane it (and its stored or run-end force)
  auto pull = [&](int it, T(&f)[Q], T(&sf)[3]) {
    const int x = plane(it);
    const uint32_t ox[3] = {uint32_t(x == 0 ? g.nx - 1 : x - 1) * nyz, uint32_t(x) * nyz,
                            uint32_t(x + 1 == g.nx ? 0 : x + 1) * nyz};
    [&]<int... P>(std::integer_sequence<int, P...>) {
      ((f[P] = ZH ? buffer_load<T>(rOld, (oy[1 - CI[P][1]] + oz[1 - CI[P][2]]) * B,
                                    (uint32_t(P) * n + ox[1 - CI[P][0]]) * B)
                  : buffer_load<T>(rOld, oz[1 - CI[P][2]] * B,
                                    (uint32_t(P) * n + ox[1 - CI[P][0]] + oy[1 - CI[P][1]]) * B)),
       ...);
    }(std::make_integer_sequence<int, Q>{});
    if constexpr (!VF) {
      const __amdgpu_buffer_rsrc_t rF = buffer_rsrc(A.force);
#pragma unroll
      for (int a = 0; a < 3; ++a) sf[a] = buffer_load<T>(rF, zl * B, (uint32_t(a) * n + ox[1] + yrow + zrow) * B);
    } else if constexpr (ZH) {
      // z-halo sites are the last (z % 64 == 63) or first site of a run
      const uint32_t site = ox[1] + zl;
      const uint32_t e = 2 * (site / WAVE_SITES) + (site % WAVE_SITES == WAVE_SITES - 1 ? 1 : 0);
      const __amdgpu_buffer_rsrc_t rE = buffer_rsrc(A.vf.edge);
#pragma unroll
      for (int a = 0; a < 3; ++a) sf[a] = T(buffer_load<double>(rE, (3 * e + uint32_t(a)) * 8u, 0));
    }
  };
  // issued before plane it-1 is collided: list bounds of plane it+2 and the
  // pulls of plane it (the list records of plane it follow the collision,
  // when its registers are free)
  auto load = [&](int it, T(&f)[Q], T(&sf)[3]) {
    if constexpr (VFROW) load_off(it + 2);
    pull(it, f, sf);
  };
  // force at this lane's site on plane it
  auto force = [&](int it, const T(&sf)[3], T(&F)[3]) {
    if constexpr (VFROW) {
      const int x = plane(it);
      double acc[3];
      if (simple)
        quad_row_force(rec, o0, m0, A.vf.rec, A.vf.wz, z, g, acc);
      else
        wave_force(A.vf.pts, A.vf.off, A.vf.sorted, list_of(it), true, x, y, z, g, acc);
#pragma unroll
      for (int a = 0; a < 3; ++a) F[a] = T(acc[a]);
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) F[a] = sf[a];
    }
  };
  T Fh[3] = {T(0), T(0), T(0)};  // force at this site on the previous plane
  // STORE: write the distributions (tile planes); HYDRO: CalcHydro of plane it-1
  auto process = [&]<bool STORE, bool HYDRO>(int it, const T(&f)[Q], const T(&F)[3]) {
    const int x = plane(it);
    const uint32_t srow = uint32_t(x) * nyz + yrow + zrow;
    T out[Q];
    collide_site(f, F, A.tau_s, A.tau_b, A.omega_s, A.omega_b, out);
    uint32_t o2 = 0, m2 = 0;
    if constexpr (VFROW) {
      load_quad_entry(A.vf.rec, A.vf.wz, kv, rec);  // entries of plane it+1
      take_list(it + 2, o2, m2);                    // plane it+2's entries, from bounds loaded with plane it
    }
    if constexpr (STORE) {
      if (put) {
#pragma unroll
        for (int p = 0; p < Q; ++p) buffer_store_nt(out[p], rNew, zl * B, (uint32_t(p) * n + srow) * B);
      }
      if constexpr (!VF) {
        if (A.zero_next && put) {
          const __amdgpu_buffer_rsrc_t rZ = buffer_rsrc(A.zero_next);
#pragma unroll
          for (int a = 0; a < 3; ++a) buffer_store_nt(T(0), rZ, zl * B, (uint32_t(a) * n + srow) * B);
        }
      }
    }
    const int w0 = SLOTS == 6 ? 0 : it % 2;
    const int w1 = SLOTS == 6 ? 1 + it % 2 : 2 + it % 3;
    const int w2 = SLOTS == 6 ? 3 + it % 3 : 5 + it % 4;
    [&]<int... P>(std::integer_sequence<int, P...>) {
      (
          [&] {
            constexpr int cy = CI[P][1], cz = CI[P][2];
            if constexpr (ROLE == ROLE_ZHALO && cz == 0) return;
            if constexpr (ROLE == ROLE_ABOVE && cy != 1) return;
            if constexpr (ROLE == ROLE_BELOW && cy != -1) return;
            constexpr int grp = dir_group(P), k = dir_index(P);
            const int slot = grp == 0 ? w0 : (grp == 1 ? w1 : w2);
            const int tr = r + cy;  // target tile row
            const int at = (slot * 5 + k) * PLANE + (tr - 1) * FUSED_ROW + zz + cz + 1;
            if constexpr (ROLE == ROLE_ZHALO) {
              const bool ok = active && tr >= 1 && tr <= TY && zz + cz >= 0 && zz + cz < W;
              lds[ok ? at : TRASH * PLANE + lane] = out[P];
            } else if constexpr (ROLE == ROLE_TILE) {
              if (tr >= 1 && tr <= TY && put) lds[at] = out[P];  // uniform test unless PARTIAL
            } else {
              if (put) lds[at] = out[P];
            }
          }(),
          ...);
    }(std::make_integer_sequence<int, Q>{});
    lds_barrier();
    // CalcHydro of plane it-1: the LDS reads come before the second barrier
    // (six slot-sets; it only has to keep the next plane's writes off these
    // slots), the arithmetic and stores after it, overlapping the other
    // wavefronts' next plane.  Slots read: group 0 written this iteration,
    // group 1 the one before, group 2 two before.
    T fl[Q];
    if constexpr (HYDRO) {
      const int r0 = SLOTS == 6 ? 0 : it % 2;
      const int r1 = SLOTS == 6 ? 1 + (it + 1) % 2 : 2 + (it + 2) % 3;
      const int r2 = SLOTS == 6 ? 3 + (it + 1) % 3 : 5 + (it + 2) % 4;
      [&]<int... P>(std::integer_sequence<int, P...>) {
        ((fl[P] = lds[((dir_group(P) == 0 ? r0 : (dir_group(P) == 1 ? r1 : r2)) * 5 + dir_index(P)) * PLANE +
                      (r - 1) * FUSED_ROW + zz + 1]),
         ...);
      }(std::make_integer_sequence<int, Q>{});
    }
    if constexpr (SLOTS == 6) lds_barrier();
    if constexpr (HYDRO) {
      T rh, uh[3];
      hydro_site(fl, Fh, rh, uh);
      const uint32_t sp = uint32_t(plane(it - 1)) * nyz + yrow + zrow;
      if (put) {
        buffer_store_nt(rh, buffer_rsrc(A.rho), zl * B, sp * B);
        // u goes only to the component-interleaved array; the reference
        // layout is written from it when asked for (launch_u_from_aos)
#pragma unroll
        for (int a = 0; a < 3; ++a)
          buffer_store(uh[a], buffer_rsrc(A.u_aos), (3 * zl + uint32_t(a)) * B, 3 * sp * B);
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) Fh[a] = F[a];
    o0 = o1;
    m0 = m1;
    o1 = o2;
    m1 = m2;
  };
  // Iterations: 0 (plane xa-1, halo), 1 (xa: stored, no CalcHydro yet),
  // 2..last-1 (stored + CalcHydro of the plane before), last (plane xb, halo:
  // CalcHydro of xb-1 only).  last >= 2.  Two register sets: the next plane's
  // pulls are in flight while this plane is collided (a third set, for FP32,
  // measured slower).
  T fa[Q], fb[Q], sa[3] = {}, sb[3] = {}, F[3];
  if constexpr (VFROW) {  // prime the list pipeline: plane 0 entries, plane 1 indices, plane 2 bounds
    load_off(0);
    take_list(0, o0, m0);
    load_quad_entry(A.vf.rec, A.vf.wz, kv, rec);
    load_off(1);
    take_list(1, o1, m1);
    load_off(2);
  }
  pull(0, fa, sa);
  force(0, sa, F);
  load(1, fb, sb);
  process.template operator()<false, false>(0, fa, F);
  force(1, sb, F);
  load(2, fa, sa);
  process.template operator()<TILE, false>(1, fb, F);
  int it = 2;
  for (; it + 1 < last; it += 2) {
    force(it, sa, F);
    load(it + 1, fb, sb);
    process.template operator()<TILE, TILE>(it, fa, F);
    force(it + 1, sb, F);
    load(it + 2, fa, sa);
    process.template operator()<TILE, TILE>(it + 1, fb, F);
  }
  if (it + 1 == last) {
    force(it, sa, F);
    load(last, fb, sb);
    process.template operator()<TILE, TILE>(it, fa, F);
    force(last, sb, F);
    process.template operator()<false, TILE>(last, fb, F);
  } else {
    force(last, sa, F);
    process.template operator()<false, TILE>(last, fa, F);
  }
}

template <typename T, int TY>
constexpr int fused_lds_words() { return fused_slots<T>() * 5 * TY * FUSED_ROW + 64; }

// Tile-strip L of the sweep (z tiles fastest, then y tiles, then strips).
template <typename T, bool VF, int TY, bool PARTIAL>
__device__ __forceinline__ void step_hydro_tile(const FusedArgs<T>& A, const FusedTiles& t, uint32_t L,
                                                T* __restrict__ lds) {
  constexpr int ROWS = TY + 2;
  const Geom& g = A.g;
  // Tiles along z, then along y, are adjacent in L, so one XCD sweeps whole
  // z-rings of neighbouring tiles together: the lines a 64-site run's z-shifted
  // pulls share with the next run (a third of a FP32 run's lines) and most row
  // halos come from its own L2.
  const uint32_t zt = L % t.nzt, rest = L / t.nzt;
  const uint32_t yt = rest % t.nyt, sx = rest / t.nyt;
  const int xa = int(sx * t.strip), xb = min(xa + int(t.strip), g.nx);
  const int last = xb - xa + 1;  // iterations 0..last collide planes xa-1..xb
  const int W = t.width;
  const int z0 = min(int(zt) * FUSED_Z, g.nz - W), y0 = min(int(yt) * TY, g.ny - TY);
  const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64), lane = int(threadIdx.x) % 64;
  // this lane's position: tile row r (0 and ROWS-1 are halo rows) and tile z
  // index zz (-1 and W are halo sites); spare lanes of the z-halo wave take
  // a valid position and write nothing
  auto wrap = [](int v, int size) { return v < 0 ? v + size : (v >= size ? v - size : v); };
  if (wv == ROWS) {
    const bool active = lane < 2 * ROWS;
    const int r = active ? lane >> 1 : 0;
    const int zz = (active && (lane & 1)) ? W : -1;
    const int y = wrap(y0 + r - 1, g.ny), z = wrap(z0 + zz, g.nz);
    fused_sweep<T, VF, TY, ROLE_ZHALO, PARTIAL>(A, lds, xa, last, r, zz, active, y, z, W);
    return;
  }
  const bool active = lane < W;
  const int zz = active ? lane : W - 1;
  const int y = wrap(y0 + wv - 1, g.ny), z = z0 + zz;
  if (wv == 0)
    fused_sweep<T, VF, TY, ROLE_ABOVE, PARTIAL>(A, lds, xa, last, wv, zz, active, y, z, W);
  else if (wv == ROWS - 1)
    fused_sweep<T, VF, TY, ROLE_BELOW, PARTIAL>(A, lds, xa, last, wv, zz, active, y, z, W);
  else
    fused_sweep<T, VF, TY, ROLE_TILE, PARTIAL>(A, lds, xa, last, wv, zz, active, y, z, W);
}

template <typename T, bool VF, int TY, bool PARTIAL>
__global__ __launch_bounds__((TY + 3) * 64) void k_step_hydro(FusedArgs<T> A, FusedTiles t) {
  __shared__ T lds[fused_lds_words<T, TY>()];
  const uint32_t L = xcd_block(round_xcd(t.nblk));
  if (L >= t.nblk) return;
  step_hydro_tile<T, VF, TY, PARTIAL>(A, t, L, lds);
}

// Batched fused step over members with full tiles (nz % 64 == 0, TY dividing
// ny, nx and ny >= 4) and a virtual force: WorkItem.unit is k_step_hydro's
// tile-strip index within member `member`, whose x sweep is split into
// strips of `strip` planes.  The member's lists start at its tile_base (the
// list bounds and run-end forces are offset by it; the entries, point
// records and point ids are numbered across the batch).  State in g[cur ^
// parity], as kb_collide.
template <typename T, int TY>
__global__ __launch_bounds__((TY + 3) * 64) void kb_step_hydro(const LatDesc* __restrict__ desc,
                                                               const WorkItem* __restrict__ work, uint32_t nwork,
                                                               int parity, VForce vf, int strip) {
  __shared__ T lds[fused_lds_words<T, TY>()];
  const uint32_t wi = xcd_block(round_xcd(nwork));
  if (wi >= nwork) return;
  const WorkItem w = work[wi];
  const LatDesc& m = desc[w.member];
  const int c = m.cur ^ parity;
  FusedArgs<T> A;
  A.gOld = static_cast<const T*>(m.g[c]);
  A.force = nullptr;
  A.vf = VForce{vf.pts, vf.off + m.tile_base, vf.sorted, vf.edge + 6 * size_t(m.tile_base), vf.rec, vf.wz};
  A.gNew = static_cast<T*>(m.g[c ^ 1]);
  A.rho = static_cast<T*>(m.rho);
  A.u_aos = static_cast<T*>(m.u_aos);
  A.zero_next = nullptr;
  A.g = m.geom;
  A.tau_s = T(m.tau_s);
  A.tau_b = T(m.tau_b);
  A.omega_s = T(m.omega_s);
  A.omega_b = T(m.omega_b);
  FusedTiles t;
  t.nyt = uint32_t(m.geom.ny / TY);
  t.nzt = uint32_t(m.geom.nz / FUSED_Z);
  t.strip = uint32_t(strip);
  t.width = FUSED_Z;
  t.nblk = 0;  // not read by the tile sweep
  step_hydro_tile<T, true, TY, false>(A, t, w.unit, lds);
}

// ------------------------------------------------------------------------
// Launch wrappers: one lattice, or a batch of lattices described by LatDesc
// entries with one WorkItem (member, block of that member) per workgroup.
// ------------------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_collide_stream_f64(const double* __restrict__ gOld,
                                                              const double* __restrict__ force,
                                                              double* __restrict__ gNew, Geom g, double tau_s,
                                                              double tau_b, double omega_s, double omega_b) {
  collide_f64_body<false>(gOld, force, gNew, g, tau_s, tau_b, omega_s, omega_b, blockIdx.x, VForce{}, 0);
}
__global__ __launch_bounds__(BLOCK) void k_collide_stream_vf_f64(const double* __restrict__ gOld, VForce vf,
                                                                 double* __restrict__ gNew, Geom g, double tau_s,
                                                                 double tau_b, double omega_s, double omega_b) {
  collide_f64_body<true>(gOld, nullptr, gNew, g, tau_s, tau_b, omega_s, omega_b, blockIdx.x, vf, blockIdx.x * LISTS_PER_BLOCK);
}
__global__ __launch_bounds__(BLOCK) void k_collide_stream_f32(const float* __restrict__ gOld,
                                                              const float* __restrict__ force,
                                                              float* __restrict__ gNew, Geom g, float tau_s,
                                                              float tau_b, float omega_s, float omega_b) {
  collide_f32_body<false>(gOld, force, gNew, g, tau_s, tau_b, omega_s, omega_b, blockIdx.x, VForce{}, 0);
}
__global__ __launch_bounds__(BLOCK) void k_collide_stream_vf_f32(const float* __restrict__ gOld, VForce vf,
                                                                 float* __restrict__ gNew, Geom g, float tau_s,
                                                                 float tau_b, float omega_s, float omega_b) {
  collide_f32_body<true>(gOld, nullptr, gNew, g, tau_s, tau_b, omega_s, omega_b, blockIdx.x, vf, blockIdx.x * LISTS_PER_BLOCK);
}
__global__ __launch_bounds__(BLOCK) void k_calc_hydro_f64(const double* __restrict__ gCur,
                                                          const double* __restrict__ force, double* __restrict__ rho,
                                                          double* __restrict__ u, Geom g, bool aos) {
  hydro_f64_body<false>(gCur, force, rho, u, g, blockIdx.x, VForce{}, 0, aos);
}
__global__ __launch_bounds__(BLOCK) void k_calc_hydro_vf_f64(const double* __restrict__ gCur, VForce vf,
                                                             double* __restrict__ rho, double* __restrict__ u,
                                                             Geom g, bool aos) {
  hydro_f64_body<true>(gCur, nullptr, rho, u, g, blockIdx.x, vf, blockIdx.x * LISTS_PER_BLOCK, aos);
}
__global__ __launch_bounds__(BLOCK) void k_calc_hydro_f32(const float* __restrict__ gCur,
                                                          const float* __restrict__ force, float* __restrict__ rho,
                                                          float* __restrict__ u, Geom g, bool aos) {
  hydro_f32_body<false>(gCur, force, rho, u, g, blockIdx.x, VForce{}, 0, aos);
}
__global__ __launch_bounds__(BLOCK) void k_calc_hydro_vf_f32(const float* __restrict__ gCur, VForce vf,
                                                             float* __restrict__ rho, float* __restrict__ u, Geom g,
                                                             bool aos) {
  hydro_f32_body<true>(gCur, nullptr, rho, u, g, blockIdx.x, vf, blockIdx.x * LISTS_PER_BLOCK, aos);
}

// At the start of a batched step with `parity` p the state of member m is in
// g[cur_m ^ p]; the collision writes g[cur_m ^ p ^ 1], which CalcHydro reads.
// VF: the launch covers members with FORCE_LISTS_VIRTUAL, whose force of
// block `unit` comes from lists tile_base_m + 4 unit ...; otherwise members
// whose force array holds it.  (One kernel for both would need more registers
// than either.)
// A work item covers `per` consecutive 256-site blocks of its member (the
// work-unit size the ensemble autotuner chooses).
template <typename T, bool VF>
__global__ __launch_bounds__(BLOCK) void kb_collide(const LatDesc* __restrict__ desc, const WorkItem* __restrict__ work,
                                                    int parity, VForce vf, int per, const uint32_t* __restrict__ soff,
                                                    const double* __restrict__ ssorted) {
  const WorkItem w = work[blockIdx.x];
  const LatDesc& m = desc[w.member];
  const int c = m.cur ^ parity;
  const uint32_t nb = grid_blocks(m.geom.n);
  SegForce sf;
  if (!VF && soff && m.force_path == FORCE_SITES_STORED) sf = SegForce{soff, ssorted, m.site_base, m.force};
  for (int b = 0; b < per; ++b) {
    const uint32_t blk = w.unit + uint32_t(b);
    if (blk >= nb) break;
    const uint32_t vlist = m.tile_base + blk * LISTS_PER_BLOCK;
    if constexpr (sizeof(T) == 8)
      collide_f64_body<VF>(static_cast<const double*>(m.g[c]), static_cast<const double*>(m.force),
                           static_cast<double*>(m.g[c ^ 1]), m.geom, m.tau_s, m.tau_b, m.omega_s, m.omega_b, blk, vf,
                           vlist, sf);
    else
      collide_f32_body<VF>(static_cast<const float*>(m.g[c]), static_cast<const float*>(m.force),
                           static_cast<float*>(m.g[c ^ 1]), m.geom, float(m.tau_s), float(m.tau_b), float(m.omega_s),
                           float(m.omega_b), blk, vf, vlist, sf);
  }
}

template <typename T, bool VF>
__global__ __launch_bounds__(BLOCK) void kb_hydro(const LatDesc* __restrict__ desc, const WorkItem* __restrict__ work,
                                                  int parity, VForce vf, int per) {
  const WorkItem w = work[blockIdx.x];
  const LatDesc& m = desc[w.member];
  const int c = m.cur ^ parity ^ 1;
  const uint32_t nb = grid_blocks(m.geom.n);
  for (int b = 0; b < per; ++b) {
    const uint32_t blk = w.unit + uint32_t(b);
    if (blk >= nb) break;
    const uint32_t vlist = m.tile_base + blk * LISTS_PER_BLOCK;
    // u interleaved where the member has that copy (the batched moves gather from it)
    const bool aos = m.u_aos != nullptr;
    if constexpr (sizeof(T) == 8)
      hydro_f64_body<VF>(static_cast<const double*>(m.g[c]), static_cast<const double*>(m.force),
                         static_cast<double*>(m.rho), static_cast<double*>(aos ? m.u_aos : m.u), m.geom, blk, vf,
                         vlist, aos);
    else
      hydro_f32_body<VF>(static_cast<const float*>(m.g[c]), static_cast<const float*>(m.force),
                         static_cast<float*>(m.rho), static_cast<float*>(aos ? m.u_aos : m.u), m.geom, blk, vf,
                         vlist, aos);
  }
}

// ------------------------------------------------------------------------
// Utilities
// ------------------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_fill_zero(uint4* __restrict__ p, size_t n16) {
  const size_t stride = size_t(gridDim.x) * BLOCK;
  for (size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}
__global__ void k_fill_zero_tail(unsigned char* __restrict__ p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename T>
__device__ __forceinline__ double rest_weight(int p) {
  if constexpr (sizeof(T) == 4) return TAB.w[p];
  return 0.0;
}

// Block-partial FP64 sums of mass and momentum (deterministic order).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_totals_partial(const T* __restrict__ f, Geom g,
                                                          double* __restrict__ partial) {
  __shared__ double red[4][BLOCK];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const uint32_t n = g.n;
  for (uint32_t s = blockIdx.x * BLOCK + threadIdx.x; s < n; s += gridDim.x * BLOCK) {
#pragma unroll
    for (int p = 0; p < Q; ++p) {
      const double fp = double(f[size_t(p) * n + s]) + rest_weight<T>(p);
      acc[0] += fp;
      acc[1] += CI[p][0] * fp;
      acc[2] += CI[p][1] * fp;
      acc[3] += CI[p][2] * fp;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
#pragma unroll
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_totals_final(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
  if (threadIdx.x < 4) {
    double a = 0.0;
    for (int b = 0; b < nblocks; ++b) a += partial[b * 4 + threadIdx.x];
    out[threadIdx.x] = a;
  }
}

__global__ __launch_bounds__(BLOCK) void k_copy(const double* __restrict__ src, double* __restrict__ dst, size_t n) {
  const size_t i = size_t(blockIdx.x) * BLOCK + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

inline unsigned grid_for(uint32_t n) { return (n + BLOCK - 1) / BLOCK; }

}  // namespace

// ---- launchers -------------------------------------------------------------

template <>
hipError_t launch_collide_stream<double>(const double* fOld, const double* force, double* fNew, const Geom& g,
                                         double tau_s, double tau_b, const VForce* vf, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  // omega = 1/(tau + 1/2) as lb.cu:107-108 computes it, once per launch.
  const double omega_s = 1.0 / (tau_s + 0.5), omega_b = 1.0 / (tau_b + 0.5);
  if (vf)
    hipLaunchKernelGGL(k_collide_stream_vf_f64, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, fOld, *vf, fNew, g, tau_s,
                       tau_b, omega_s, omega_b);
  else
    hipLaunchKernelGGL(k_collide_stream_f64, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, fOld, force, fNew, g, tau_s,
                       tau_b, omega_s, omega_b);
  return hipGetLastError();
}

template <>
hipError_t launch_collide_stream<float>(const float* fOld, const float* force, float* fNew, const Geom& g,
                                        double tau_s, double tau_b, const VForce* vf, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  const float os = float(1.0 / (tau_s + 0.5)), ob = float(1.0 / (tau_b + 0.5));
  if (vf)
    hipLaunchKernelGGL(k_collide_stream_vf_f32, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, fOld, *vf, fNew, g,
                       float(tau_s), float(tau_b), os, ob);
  else
    hipLaunchKernelGGL(k_collide_stream_f32, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, fOld, force, fNew, g,
                       float(tau_s), float(tau_b), os, ob);
  return hipGetLastError();
}

bool fused_step_supported(const Geom& g, int ty, size_t elem) {
  return g.n > 0 && (ty == 2 || ty == 4 || ty == 8) && g.ny >= ty && uint64_t(Q) * g.n * elem < (uint64_t(1) << 32);
}

bool fused_tiles_full(const Geom& g, int ty) { return g.nz % FUSED_Z == 0 && g.ny % ty == 0; }

template <typename T>
hipError_t launch_step_hydro(const T* fOld, const T* force, T* fNew, T* rho, T* u_aos, const Geom& g, double tau_s,
                             double tau_b, const VForce* vf, int ty, int strips, hipStream_t s, T* zero_next) {
  if (!fused_step_supported(g, ty, sizeof(T)) || strips < 0) return hipErrorInvalidValue;
  if (zero_next && (vf || zero_next == force)) return hipErrorInvalidValue;
  // the virtual force's row walk needs each 64-site run to be one z-row
  if (vf && g.nz % FUSED_Z != 0) return hipErrorInvalidValue;
  FusedTiles t;
  t.nyt = uint32_t((g.ny + ty - 1) / ty);
  t.nzt = uint32_t((g.nz + FUSED_Z - 1) / FUSED_Z);
  t.width = std::min(g.nz, FUSED_Z);
  const bool partial = g.nz < FUSED_Z;
  const uint32_t tiles = t.nyt * t.nzt;
  if (strips == 0) {
    // one workgroup per CU (the tile's LDS and VGPRs allow one): enough
    // strips for every CU to have a tile, each as long as possible
    static int cus = 0;
    if (!cus) {
      int dev = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    }
    strips = int(std::min<uint32_t>(uint32_t(g.nx), std::max<uint32_t>(1, (uint32_t(cus) + tiles - 1) / tiles)));
  }
  strips = std::min(strips, g.nx);
  t.strip = uint32_t((g.nx + strips - 1) / strips);
  t.nblk = tiles * uint32_t((g.nx + int(t.strip) - 1) / int(t.strip));
  FusedArgs<T> A;
  A.gOld = fOld;
  A.force = force;
  A.vf = vf ? *vf : VForce{};
  if (vf && !(vf->edge && vf->rec && vf->wz)) return hipErrorInvalidValue;
  A.gNew = fNew;
  A.rho = rho;
  A.u_aos = u_aos;
  A.zero_next = zero_next;
  A.g = g;
  A.tau_s = T(tau_s);
  A.tau_b = T(tau_b);
  A.omega_s = T(1.0 / (tau_s + 0.5));
  A.omega_b = T(1.0 / (tau_b + 0.5));
  const dim3 grid(round_xcd(t.nblk));
#define LBS_FUSED(TYV)                                                                                       \
  if (ty == TYV) {                                                                                           \
    if (vf)                                                                                                  \
      hipLaunchKernelGGL((k_step_hydro<T, true, TYV, false>), grid, dim3((TYV + 3) * 64), 0, s, A, t);       \
    else if (partial)                                                                                        \
      hipLaunchKernelGGL((k_step_hydro<T, false, TYV, true>), grid, dim3((TYV + 3) * 64), 0, s, A, t);       \
    else                                                                                                     \
      hipLaunchKernelGGL((k_step_hydro<T, false, TYV, false>), grid, dim3((TYV + 3) * 64), 0, s, A, t);      \
  }
  LBS_FUSED(2)
  LBS_FUSED(4)
  LBS_FUSED(8)
#undef LBS_FUSED
  return hipGetLastError();
}
template hipError_t launch_step_hydro<double>(const double*, const double*, double*, double*, double*, const Geom&,
                                              double, double, const VForce*, int, int, hipStream_t, double*);
template hipError_t launch_step_hydro<float>(const float*, const float*, float*, float*, float*, const Geom&, double,
                                             double, const VForce*, int, int, hipStream_t, float*);

// u[a n + s] = u_aos[3 s + a]: the reference layout from the fused step's
// component-interleaved u.  Thread per site; the three loads of a site are
// contiguous and each component's stores coalesce.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_u_from_aos(const T* __restrict__ aos, T* __restrict__ u, uint32_t n) {
  const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
  if (s >= n) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) store_stream(&u[size_t(a) * n + s], aos[3 * size_t(s) + a]);
}

template <typename T>
hipError_t launch_u_from_aos(const T* aos, T* u, const Geom& g, hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_u_from_aos<T>, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, aos, u, g.n);
  return hipGetLastError();
}
template hipError_t launch_u_from_aos<double>(const double*, double*, const Geom&, hipStream_t);
template hipError_t launch_u_from_aos<float>(const float*, float*, const Geom&, hipStream_t);

template <>
hipError_t launch_calc_hydro<double>(const double* f, const double* force, double* rho, double* u, const Geom& g,
                                     const VForce* vf, hipStream_t s, bool aos) {
  if (g.n == 0) return hipSuccess;
  if (vf)
    hipLaunchKernelGGL(k_calc_hydro_vf_f64, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, f, *vf, rho, u, g, aos);
  else
    hipLaunchKernelGGL(k_calc_hydro_f64, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, f, force, rho, u, g, aos);
  return hipGetLastError();
}

template <>
hipError_t launch_calc_hydro<float>(const float* f, const float* force, float* rho, float* u, const Geom& g,
                                    const VForce* vf, hipStream_t s, bool aos) {
  if (g.n == 0) return hipSuccess;
  if (vf)
    hipLaunchKernelGGL(k_calc_hydro_vf_f32, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, f, *vf, rho, u, g, aos);
  else
    hipLaunchKernelGGL(k_calc_hydro_f32, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, f, force, rho, u, g, aos);
  return hipGetLastError();
}

template <>
hipError_t launch_init_from_hydro<double>(const double* rho, const double* u, double* f, const Geom& g,
                                          hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_init_from_hydro_f64, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, rho, u, f, g);
  return hipGetLastError();
}

template <>
hipError_t launch_init_from_hydro<float>(const float* rho, const float* u, float* f, const Geom& g,
                                         hipStream_t s) {
  if (g.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_init_from_hydro_f32, dim3(grid_for(g.n)), dim3(BLOCK), 0, s, rho, u, f, g);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_batch_collide(const LatDesc* desc, const WorkItem* work, int nwork, int per, int parity,
                                const VForce* vf, hipStream_t s, const uint32_t* soff, const double* ssorted) {
  if (nwork <= 0) return hipSuccess;
  if (vf)
    hipLaunchKernelGGL((kb_collide<T, true>), dim3(nwork), dim3(BLOCK), 0, s, desc, work, parity, *vf, per, nullptr,
                       nullptr);
  else
    hipLaunchKernelGGL((kb_collide<T, false>), dim3(nwork), dim3(BLOCK), 0, s, desc, work, parity, VForce{}, per,
                       soff, ssorted);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_batch_hydro(const LatDesc* desc, const WorkItem* work, int nwork, int per, int parity,
                              const VForce* vf, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  if (vf)
    hipLaunchKernelGGL((kb_hydro<T, true>), dim3(nwork), dim3(BLOCK), 0, s, desc, work, parity, *vf, per);
  else
    hipLaunchKernelGGL((kb_hydro<T, false>), dim3(nwork), dim3(BLOCK), 0, s, desc, work, parity, VForce{}, per);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_batch_step_hydro(const LatDesc* desc, const WorkItem* work, int nwork, int parity, const VForce& vf,
                                   int rows, int strip, hipStream_t s) {
  if (nwork <= 0) return hipSuccess;
  if (strip < 1 || !(vf.edge && vf.rec && vf.wz)) return hipErrorInvalidValue;
  const dim3 grid(round_xcd(uint32_t(nwork)));
  if (rows == 8)
    hipLaunchKernelGGL((kb_step_hydro<T, 8>), grid, dim3((8 + 3) * 64), 0, s, desc, work, uint32_t(nwork), parity,
                       vf, strip);
  else if (rows == 4)
    hipLaunchKernelGGL((kb_step_hydro<T, 4>), grid, dim3((4 + 3) * 64), 0, s, desc, work, uint32_t(nwork), parity,
                       vf, strip);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
template hipError_t launch_batch_step_hydro<double>(const LatDesc*, const WorkItem*, int, int, const VForce&, int,
                                                    int, hipStream_t);
template hipError_t launch_batch_step_hydro<float>(const LatDesc*, const WorkItem*, int, int, const VForce&, int,
                                                   int, hipStream_t);

bool batch_fused_member(const Geom& g, int rows, size_t elem) {
  return (rows == 8 || rows == 4) && g.nx >= 4 && g.ny >= 4 && g.nz % FUSED_Z == 0 && g.ny % rows == 0 &&
         fused_step_supported(g, rows, elem);
}

template hipError_t launch_batch_collide<double>(const LatDesc*, const WorkItem*, int, int, int, const VForce*,
                                                 hipStream_t, const uint32_t*, const double*);
template hipError_t launch_batch_collide<float>(const LatDesc*, const WorkItem*, int, int, int, const VForce*,
                                                hipStream_t, const uint32_t*, const double*);
template hipError_t launch_batch_hydro<double>(const LatDesc*, const WorkItem*, int, int, int, const VForce*,
                                               hipStream_t);
template hipError_t launch_batch_hydro<float>(const LatDesc*, const WorkItem*, int, int, int, const VForce*,
                                              hipStream_t);

uint32_t lattice_blocks(const Geom& g) { return grid_for(g.n); }

hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  const size_t n16 = bytes / 16;
  if (n16) {
    const size_t blocks = (n16 + BLOCK - 1) / BLOCK;
    const unsigned grid = unsigned(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(BLOCK), 0, s, static_cast<uint4*>(p), n16);
  }
  const size_t tail = bytes - n16 * 16;
  if (tail) {
    hipLaunchKernelGGL(k_fill_zero_tail, dim3(1), dim3(64), 0, s, static_cast<unsigned char*>(p) + n16 * 16,
                       tail);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_totals(const T* f, const Geom& g, double* partial, int nblocks, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_totals_partial<T>, dim3(nblocks), dim3(BLOCK), 0, s, f, g, partial);
  hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(64), 0, s, partial, nblocks, out);
  return hipGetLastError();
}
template hipError_t launch_totals<double>(const double*, const Geom&, double*, int, double*, hipStream_t);
template hipError_t launch_totals<float>(const float*, const Geom&, double*, int, double*, hipStream_t);

hipError_t launch_copy(const double* src, double* dst, size_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3(unsigned((count + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, src, dst, count);
  return hipGetLastError();
}

}  // namespace lbs
