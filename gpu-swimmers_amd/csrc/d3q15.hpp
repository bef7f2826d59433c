// D3Q15 MRT model constants as compile-time tables, shared by host and device.
//
// The reference fills these at run time into an LBParams struct and its
// kernels read them from global memory (lbswim/liblbswim/d3q15.c:2-186,
// LBParams.h:4-30, lb.cu:117,201).  Here every entry is built from its closed
// form as an exact rational rounded once to double -- the value the
// reference's literal expressions (e.g. 2./3.) fold to -- so kernels can
// unroll over the tables and drop the zero entries at compile time.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lbs {

constexpr int Q = 15;
constexpr int D = 3;

// Velocity set in the reference's order (d3q15.c:24-38).
constexpr int CI[Q][D] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                          {0, 0, 1},   {0, 0, -1},  {1, 1, 1},  {1, 1, -1},  {1, -1, 1},
                          {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

struct Rat {
  long long n, d;
};

constexpr long long gcd_ll(long long a, long long b) {
  a = a < 0 ? -a : a;
  b = b < 0 ? -b : b;
  while (b) {
    long long t = a % b;
    a = b;
    b = t;
  }
  return a ? a : 1;
}
constexpr Rat mk(long long n, long long d) {
  if (d < 0) {
    n = -n;
    d = -d;
  }
  long long g = gcd_ll(n, d);
  return Rat{n / g, d / g};
}
constexpr Rat rmul(Rat a, Rat b) { return mk(a.n * b.n, a.d * b.d); }
constexpr Rat radd(Rat a, Rat b) { return mk(a.n * b.d + b.n * a.d, a.d * b.d); }
constexpr double val(Rat a) { return double(a.n) / double(a.d); }

constexpr int speed2(int p) { return CI[p][0] * CI[p][0] + CI[p][1] * CI[p][1] + CI[p][2] * CI[p][2]; }

// Quadrature weights 2/9, 1/9, 1/72 by |c|^2 = 0, 1, 3.
constexpr Rat weight_r(int p) {
  return speed2(p) == 0 ? mk(2, 9) : (speed2(p) == 1 ? mk(1, 9) : mk(1, 72));
}

// Mode m on velocity p: 1; c_a; c_a c_b - delta_ab/3 (xx,xy,xz,yy,yz,zz);
// chi1 = -2 + 9/2 s - 3/2 s^2 (s = |c|^2); chi1 c_a; c_x c_y c_z.
constexpr Rat mode_r(int m, int p) {
  constexpr int AB[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  const int s = speed2(p);
  const long long chi1x2 = -4 + 9 * s - 3 * s * s;
  if (m == 0) return mk(1, 1);
  if (m <= 3) return mk(CI[p][m - 1], 1);
  if (m <= 9) {
    const int a = AB[m - 4][0], b = AB[m - 4][1];
    return mk(3 * CI[p][a] * CI[p][b] - (a == b ? 1 : 0), 3);
  }
  if (m == 10) return mk(chi1x2, 2);
  if (m <= 13) return mk(chi1x2 * CI[p][m - 11], 2);
  return mk(CI[p][0] * CI[p][1] * CI[p][2], 1);
}

constexpr Rat norm_r(int m) {
  Rat acc = mk(0, 1);
  for (int p = 0; p < Q; ++p) acc = radd(acc, rmul(weight_r(p), rmul(mode_r(m, p), mode_r(m, p))));
  return mk(acc.d, acc.n);
}

struct Tables {
  double w[Q];
  double mm[Q][Q];   // rows are modes
  double mmi[Q][Q];  // mmi[p][m] = w_p mm[m][p] norms[m]
  double norms[Q];
};

constexpr Tables make_tables() {
  Tables t{};
  for (int p = 0; p < Q; ++p) t.w[p] = val(weight_r(p));
  for (int m = 0; m < Q; ++m) {
    t.norms[m] = val(norm_r(m));
    for (int p = 0; p < Q; ++p) {
      t.mm[m][p] = val(mode_r(m, p));
      t.mmi[p][m] = val(rmul(weight_r(p), rmul(mode_r(m, p), norm_r(m))));
    }
  }
  return t;
}

constexpr Tables TAB = make_tables();

// Mode indices (d3q15.h:4-28).
constexpr int M_RHO = 0, M_SXX = 4, M_SXY = 5, M_SXZ = 6, M_SYY = 7, M_SYZ = 8, M_SZZ = 9;
// Only modes 0..9 survive a collision: the ghosts are set to zero (lb.cu:191-195).
constexpr int NKEEP = 10;

// First column p with a nonzero entry in mode row m, and first mode m with a
// nonzero entry in mmi row p among the kept modes.
constexpr int first_nz_mm(int m) {
  for (int p = 0; p < Q; ++p)
    if (TAB.mm[m][p] != 0.0) return p;
  return Q;
}
constexpr int first_nz_mmi(int p) {
  for (int m = 0; m < NKEEP; ++m)
    if (TAB.mmi[p][m] != 0.0) return m;
  return NKEEP;
}

// Unsigned division by a run-time constant (Granlund & Montgomery, PLDI'94,
// "round-up" variant with a 33-bit intermediate): q = (mulhi(x, mul) + x) >> shift
// for every 32-bit x.
struct FastDiv {
  uint32_t d, mul, shift;
};

inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t s = 0;
  while ((uint64_t(1) << s) < d) ++s;
  const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1;
  return FastDiv{d, uint32_t(m), s};
}

__host__ __device__ inline uint32_t fastdiv(uint32_t x, FastDiv f) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t hi = __umulhi(x, f.mul);
#else
  const uint32_t hi = uint32_t((uint64_t(x) * f.mul) >> 32);
#endif
  return uint32_t((uint64_t(hi) + x) >> f.shift);
}

// Lattice geometry as the kernels see it: sizes, site count and divisors.
struct Geom {
  int nx, ny, nz;
  uint32_t n;     // nx*ny*nz
  uint32_t nyz;   // ny*nz  (stride of x; y stride is nz, z stride 1, lb.cu:55-57)
  FastDiv div_nz, div_ny;
};

}  // namespace lbs
