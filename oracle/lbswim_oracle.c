/*
 * lbswim CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference gpu-swimmers coupled step,
 * written to follow the reference's arithmetic order operation by operation
 * so that a kernel which keeps that order can be compared bit for bit.
 * It is used only as the checker (tests/, __graft_entry__.smoke()) and as
 * the timed CPU baseline in bench.py.  No product code path may link or call
 * it.  Build with -ffp-contract=off and without -ffast-math (oracle/Makefile):
 * the restatement relies on C's "as written" evaluation.
 *
 * Reference files followed (paths relative to the reference checkout):
 *   D3Q15 tables ........ lbswim/liblbswim/d3q15.c:2-186, LBParams.h:4-30
 *   DoInitFromHydro ..... lbswim/liblbswim/lb.cu:314-371
 *   DoLatticeZeroForce .. lbswim/liblbswim/lb.cu:390-411
 *   DoStep .............. lbswim/liblbswim/lb.cu:72-218
 *   DoCalcHydro ......... lbswim/liblbswim/lb.cu:250-295
 *   peskin_delta ........ lbswim/liblbswim/delta.cu:6-25
 *   InterpVelocity ...... lbswim/liblbswim/interp.cu:8-47
 *   AccumulateDeltaForce  lbswim/liblbswim/SwimmerArray.cu:42-79
 *   DoSwimmerArrayAddForces lbswim/liblbswim/SwimmerArray.cu:82-112
 *   DoSwimmerArrayMove .. lbswim/liblbswim/SwimmerArray.cu:123-210
 *   DoTracerArrayMove ... lbswim/liblbswim/TracerArray.cu:12-43
 *   System.Step order ... lbswim/System.py:9-16
 *
 * Deliberate departures (all documented, none touches alpha=0 arithmetic):
 *   - the swimmer random stream is Philox4x32-10 (Salmon et al., SC'11)
 *     keyed by (seed, swimmer, draw counter) instead of cuRAND XORWOW
 *     (SwimmerArray.cu:13,173,179); the uniform is a float in (0,1] and the
 *     normals come from Box-Muller;
 *   - the orientation norm starts at 0 (SwimmerArray.cu:171 leaves it
 *     uninitialised), and with rotation off and no tumble n is left as is;
 *   - tracer s is updated per particle, s[d*N+i] += v_d
 *     (TracerArray.cu:41 races on s[0..2]);
 *   - stencil indices ((int)x + n) % n (interp.cu:26) are wrapped once more
 *     when negative, which only happens for x < -n where the reference reads
 *     out of bounds.
 * Force spreading is done in swimmer order, tail then head, which is one of
 * the orders the reference's atomics may produce (SwimmerArray.cu:71).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_Q 15
#define ORC_D 3

/* Same member order and types as the reference's LBParams (LBParams.h:4-30). */
typedef struct {
  double cs2;
  double w[ORC_Q];
  double xi[ORC_Q][ORC_D];
  int ci[ORC_Q][ORC_D];
  double Q[ORC_Q][ORC_D][ORC_D];
  int complement[ORC_Q];
  double norms[ORC_Q];
  double mm[ORC_Q][ORC_Q];
  double mmi[ORC_Q][ORC_Q];
  double tau_s;
  double tau_b;
} orc_params;

/* ------------------------------------------------------------------ */
/* D3Q15 tables from closed forms (d3q15.c:2-186).                    */
/* Every entry is an exact rational rounded once to double, which is  */
/* what the reference's literal expressions such as 2./3. evaluate to.*/
/* ------------------------------------------------------------------ */

/* Velocity order of the reference (d3q15.c:24-38). */
static const int ORC_C[ORC_Q][ORC_D] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
    {0, 0, 1},  {0, 0, -1},  {1, 1, 1},  {1, 1, -1}, {1, -1, 1},
    {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

/* Rational number with small integer parts. */
typedef struct { long num, den; } orc_rat;

static long orc_gcd(long a, long b) {
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) { long t = a % b; a = b; b = t; }
  return a ? a : 1;
}
static orc_rat orc_mk(long n, long d) {
  orc_rat r;
  long g = orc_gcd(n, d);
  if (d < 0) { n = -n; d = -d; }
  r.num = n / g; r.den = d / g;
  return r;
}
static orc_rat orc_mul(orc_rat a, orc_rat b) { return orc_mk(a.num * b.num, a.den * b.den); }
static double orc_val(orc_rat a) { return (double)a.num / (double)a.den; }

/* Mode m evaluated on velocity p as a rational (rows of mm, d3q15.c:154-168):
 *   0: 1;  1..3: xi_a;  4..9: xi_a xi_b - delta_ab/3 in the order
 *   xx, xy, xz, yy, yz, zz;  10: chi1 = -2 + 9/2 s - 3/2 s^2 with s = |xi|^2;
 *   11..13: chi1 xi_a;  14: xi_x xi_y xi_z. */
static orc_rat orc_mode_on(int m, int p) {
  const int *c = ORC_C[p];
  static const int AB[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  int s = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
  /* 2*chi1 = -4 + 9 s - 3 s^2 */
  long chi1x2 = -4 + 9 * s - 3 * s * s;
  if (m == 0) return orc_mk(1, 1);
  if (m <= 3) return orc_mk(c[m - 1], 1);
  if (m <= 9) {
    int a = AB[m - 4][0], b = AB[m - 4][1];
    return orc_mk(3 * c[a] * c[b] - (a == b ? 1 : 0), 3);
  }
  if (m == 10) return orc_mk(chi1x2, 2);
  if (m <= 13) return orc_mk(chi1x2 * c[m - 11], 2);
  return orc_mk(c[0] * c[1] * c[2], 1);
}

static orc_rat orc_weight(int p) {
  int s = ORC_C[p][0] * ORC_C[p][0] + ORC_C[p][1] * ORC_C[p][1] + ORC_C[p][2] * ORC_C[p][2];
  if (s == 0) return orc_mk(2, 9);
  if (s == 1) return orc_mk(1, 9);
  return orc_mk(1, 72);
}

/* norms[m] = 1 / sum_p w_p mm[m][p]^2 */
static orc_rat orc_norm(int m) {
  long num = 0, den = 1;
  int p;
  for (p = 0; p < ORC_Q; p++) {
    orc_rat t = orc_mul(orc_weight(p), orc_mul(orc_mode_on(m, p), orc_mode_on(m, p)));
    /* num/den += t */
    num = num * t.den + t.num * den;
    den = den * t.den;
    {
      orc_rat r = orc_mk(num, den);
      num = r.num; den = r.den;
    }
  }
  return orc_mk(den, num);
}

void orc_eigenset(orc_params *lat, double tau_s, double tau_b) {
  int p, m, a, b;
  memset(lat, 0, sizeof(*lat));
  lat->cs2 = orc_val(orc_mk(1, 3));
  for (p = 0; p < ORC_Q; p++) {
    lat->w[p] = orc_val(orc_weight(p));
    for (a = 0; a < ORC_D; a++) {
      lat->xi[p][a] = (double)ORC_C[p][a];
      lat->ci[p][a] = ORC_C[p][a];
    }
    for (a = 0; a < ORC_D; a++)
      for (b = 0; b < ORC_D; b++)
        lat->Q[p][a][b] = orc_val(orc_mk(3 * ORC_C[p][a] * ORC_C[p][b] - (a == b ? 1 : 0), 3));
    /* complement: the velocity -xi_p */
    for (m = 0; m < ORC_Q; m++)
      if (ORC_C[m][0] == -ORC_C[p][0] && ORC_C[m][1] == -ORC_C[p][1] && ORC_C[m][2] == -ORC_C[p][2])
        lat->complement[p] = m;
  }
  for (m = 0; m < ORC_Q; m++) {
    orc_rat nm = orc_norm(m);
    lat->norms[m] = orc_val(nm);
    for (p = 0; p < ORC_Q; p++) {
      lat->mm[m][p] = orc_val(orc_mode_on(m, p));
      /* mmi[p][m] = w_p mm[m][p] norms[m] */
      lat->mmi[p][m] = orc_val(orc_mul(orc_weight(p), orc_mul(orc_mode_on(m, p), nm)));
    }
  }
  lat->tau_s = tau_s;
  lat->tau_b = tau_b;
}

size_t orc_params_size(void) { return sizeof(orc_params); }

/* ------------------------------------------------------------------ */
/* Lattice kernels.  Sites are ijk = i*ny*nz + j*nz + k (lb.cu:49-57) */
/* and fields are element-slowest: f[p*n + ijk], u[a*n + ijk].        */
/* ------------------------------------------------------------------ */

/* Shared table instance; orc_set_threads() controls OpenMP width. */
static orc_params g_tab;
static int g_tab_ready = 0;
static int g_threads = 1;

static const orc_params *orc_tab(void) {
  if (!g_tab_ready) {
    orc_eigenset(&g_tab, 0.0, 0.0);
    g_tab_ready = 1;
  }
  return &g_tab;
}

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int orc_get_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* lb.cu:314-371 (note SXZ = rho*ux*uy as in lb.cu:356) */
void orc_init_from_hydro(int nx, int ny, int nz, const double *rho, const double *u, double *f) {
  const orc_params *P = orc_tab();
  const long n = (long)nx * ny * nz;
  long ijk;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (ijk = 0; ijk < n; ijk++) {
    double mode[ORC_Q];
    int m, p, a;
    for (m = 0; m < ORC_Q; m++) mode[m] = 0.0;
    mode[0] = rho[ijk];
    for (a = 0; a < ORC_D; a++) mode[1 + a] = rho[ijk] * u[a * n + ijk];
    mode[4] = rho[ijk] * u[0 * n + ijk] * u[0 * n + ijk];
    mode[5] = rho[ijk] * u[0 * n + ijk] * u[1 * n + ijk];
    mode[6] = rho[ijk] * u[0 * n + ijk] * u[1 * n + ijk];
    mode[7] = rho[ijk] * u[1 * n + ijk] * u[1 * n + ijk];
    mode[8] = rho[ijk] * u[1 * n + ijk] * u[2 * n + ijk];
    mode[9] = rho[ijk] * u[2 * n + ijk] * u[2 * n + ijk];
    for (p = 0; p < ORC_Q; p++) {
      f[p * n + ijk] = 0.;
      for (m = 0; m < ORC_Q; m++) f[p * n + ijk] += mode[m] * P->mmi[p][m];
    }
  }
}

/* lb.cu:390-411 */
void orc_zero_force(long n, double *force) {
  long i;
  for (i = 0; i < 3 * n; i++) force[i] = 0.0;
}

/* lb.cu:72-218: MRT collide with forcing, then push-stream with PBC into fNew. */
void orc_step(int nx, int ny, int nz, double tau_s, double tau_b, const double *fOld,
              const double *force, double *fNew) {
  const orc_params *P = orc_tab();
  const long n = (long)nx * ny * nz;
  const int size[3] = {nx, ny, nz};
  const long strides[3] = {(long)ny * nz, nz, 1};
  const double omega_s = 1.0 / (tau_s + 0.5);
  const double omega_b = 1.0 / (tau_b + 0.5);
  static const double DELTA[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  long ijk;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (ijk = 0; ijk < n; ijk++) {
    const int site[3] = {(int)(ijk / strides[0]), (int)((ijk / strides[1]) % ny), (int)(ijk % nz)};
    double mode[ORC_Q], S[3][3], u[3], fpc[ORC_Q];
    double usq, TrS, uF;
    int m, p, a, b;
    for (m = 0; m < ORC_Q; m++) {
      mode[m] = 0.;
      for (p = 0; p < ORC_Q; p++) mode[m] += fOld[p * n + ijk] * P->mm[m][p];
    }
    usq = 0.;
    uF = 0.;
    for (a = 0; a < 3; a++) {
      u[a] = (mode[1 + a] + 0.5 * force[a * n + ijk]) / mode[0];
      mode[1 + a] += force[a * n + ijk];
      usq += u[a] * u[a];
      uF += u[a] * force[a * n + ijk];
    }
    S[0][0] = mode[4]; S[0][1] = mode[5]; S[0][2] = mode[6];
    S[1][0] = mode[5]; S[1][1] = mode[7]; S[1][2] = mode[8];
    S[2][0] = mode[6]; S[2][1] = mode[8]; S[2][2] = mode[9];
    TrS = 0.;
    for (a = 0; a < 3; a++) TrS += S[a][a];
    for (a = 0; a < 3; a++) S[a][a] -= TrS / 3;
    TrS -= omega_b * (TrS - mode[0] * usq);
    TrS += 2. * omega_b * tau_b * uF;
    for (a = 0; a < 3; a++) {
      for (b = 0; b < 3; b++) {
        S[a][b] -= omega_s * (S[a][b] - mode[0] * (u[a] * u[b] - usq * DELTA[a][b]));
        S[a][b] += 2. * omega_s * tau_s *
                   (u[a] * force[b * n + ijk] + force[a * n + ijk] * u[b] - 2. * uF * DELTA[a][b]);
      }
      S[a][a] += TrS / 3;
    }
    mode[4] = S[0][0]; mode[5] = S[0][1]; mode[6] = S[0][2];
    mode[7] = S[1][1]; mode[8] = S[1][2]; mode[9] = S[2][2];
    mode[10] = 0.; mode[11] = 0.; mode[12] = 0.; mode[13] = 0.; mode[14] = 0.;
    for (p = 0; p < ORC_Q; p++) {
      fpc[p] = 0.;
      for (m = 0; m < ORC_Q; m++) fpc[p] += mode[m] * P->mmi[p][m];
    }
    for (p = 0; p < ORC_Q; p++) {
      long dest = 0;
      for (a = 0; a < 3; a++) {
        int di = (site[a] + P->ci[p][a] + size[a]) % size[a];
        dest += di * strides[a];
      }
      fNew[p * n + dest] = fpc[p];
    }
  }
}

/* lb.cu:250-295 */
void orc_calc_hydro(int nx, int ny, int nz, const double *f, const double *force, double *rho,
                    double *u) {
  const orc_params *P = orc_tab();
  const long n = (long)nx * ny * nz;
  long ijk;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (ijk = 0; ijk < n; ijk++) {
    double mode[ORC_Q];
    int m, p, a;
    for (m = 0; m < ORC_Q; m++) {
      mode[m] = 0.;
      for (p = 0; p < ORC_Q; p++) mode[m] += f[p * n + ijk] * P->mm[m][p];
    }
    rho[ijk] = mode[0];
    for (a = 0; a < 3; a++) u[a * n + ijk] = (mode[1 + a] + 0.5 * force[a * n + ijk]) / mode[0];
  }
}

/* ------------------------------------------------------------------ */
/* Particles                                                           */
/* ------------------------------------------------------------------ */

/* delta.cu:6-25 (Peskin 2002 four-point function) */
double orc_peskin_delta(double x) {
  double abs_x = fabs(x);
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

/* Stencil set-up shared by interp.cu:23-30 and SwimmerArray.cu:50-57. */
static void orc_stencil(const int size[3], const double r[3], int idx[3][4], double del[3][4]) {
  int d, i;
  for (d = 0; d < 3; d++) {
    double x0 = ceil(r[d] - 2.);
    for (i = 0; i < 4; i++) {
      double x = x0 + i;
      int v = ((int)x + size[d]) % size[d];
      /* the reference indexes out of bounds for x < -n; wrap instead */
      idx[d][i] = v < 0 ? v + size[d] : v;
      del[d][i] = orc_peskin_delta(r[d] - x);
    }
  }
}

/* interp.cu:8-47 */
void orc_interp(int nx, int ny, int nz, const double *u, const double *r, double *v) {
  const int size[3] = {nx, ny, nz};
  const long n = (long)nx * ny * nz;
  int idx[3][4];
  double del[3][4];
  int i, j, k, d;
  v[0] = v[1] = v[2] = 0.0;
  orc_stencil(size, r, idx, del);
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 4; ++j)
      for (k = 0; k < 4; ++k) {
        double delta3d = del[0][i] * del[1][j] * del[2][k];
        long ijk = (long)ny * nz * idx[0][i] + (long)nz * idx[1][j] + idx[2][k];
        for (d = 0; d < 3; ++d) v[d] += delta3d * u[d * n + ijk];
      }
}

/* SwimmerArray.cu:42-79, with the atomics applied in program order. */
static void orc_accumulate(int nx, int ny, int nz, double *force, const double *r, const double *F) {
  const int size[3] = {nx, ny, nz};
  const long n = (long)nx * ny * nz;
  int idx[3][4];
  double del[3][4];
  int i, j, k, d;
  orc_stencil(size, r, idx, del);
  for (i = 0; i < 4; ++i)
    for (j = 0; j < 4; ++j)
      for (k = 0; k < 4; ++k) {
        long ijk = (long)ny * nz * idx[0][i] + (long)nz * idx[1][j] + idx[2][k];
        double delta3d = del[0][i] * del[1][j] * del[2][k];
        for (d = 0; d < 3; ++d) force[d * n + ijk] += delta3d * F[d];
      }
}

/* Same member order and types as the reference's CommonParams (SwimmerArray.h:8-16). */
typedef struct {
  double P;
  double l;
  double alpha;
  double mobility;
  unsigned char translational_advection_off;
  unsigned char rotational_advection_off;
  unsigned long long seed;
} orc_common;

/* SwimmerArray.cu:82-112: tail (r - l n, force -P n) then head (r, force +P n). */
void orc_swimmers_add_forces(int nx, int ny, int nz, int ns, const orc_common *cp,
                             const double *sr, const double *sn, double *force) {
  int is, a;
  for (is = 0; is < ns; is++) {
    double r[3], F[3];
    for (a = 0; a < 3; a++) {
      r[a] = sr[ns * a + is] - sn[ns * a + is] * cp->l;
      F[a] = -cp->P * sn[ns * a + is];
    }
    orc_accumulate(nx, ny, nz, force, r, F);
    for (a = 0; a < 3; a++) {
      r[a] = sr[ns * a + is];
      F[a] *= -1.0;
    }
    orc_accumulate(nx, ny, nz, force, r, F);
  }
}

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
static void orc_philox_round(uint32_t ctr[4], const uint32_t key[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  uint32_t out[4];
  out[0] = hi1 ^ ctr[1] ^ key[0];
  out[1] = lo1;
  out[2] = hi0 ^ ctr[3] ^ key[1];
  out[3] = lo0;
  memcpy(ctr, out, sizeof(out));
}

void orc_philox4x32_10(const uint32_t in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c[4], k[2];
  int r;
  memcpy(c, in, sizeof(c));
  memcpy(k, key_in, sizeof(k));
  for (r = 0; r < 10; r++) {
    if (r) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    orc_philox_round(c, k);
  }
  memcpy(out, c, sizeof(c));
}

/* Draw block `tag` of swimmer `is` at draw counter `ctr` under `seed`. */
static void orc_draw(unsigned long long seed, int is, uint64_t ctr, uint32_t tag, uint32_t out[4]) {
  uint32_t in[4], key[2];
  in[0] = (uint32_t)is;
  in[1] = (uint32_t)ctr;
  in[2] = (uint32_t)(ctr >> 32);
  in[3] = tag;
  key[0] = (uint32_t)seed;
  key[1] = (uint32_t)(seed >> 32);
  orc_philox4x32_10(in, key, out);
}

/* float uniform in (0,1] from the top 24 bits */
static float orc_u01f(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }
/* double uniform in (0,1) from 53 bits */
static double orc_u01d(uint32_t a, uint32_t b) {
  uint64_t m = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
  m &= ((uint64_t)1 << 53) - 1;
  return ((double)m + 0.5) * 0x1p-53;
}

/* ln(x) and cos/sin(2 pi u) for Box-Muller, from series with +, -, *, /
 * only, so that GPU and CPU give the same bits (library log/sin/cos differ in
 * the last bit).  ln: x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s),
 * s = (m-1)/(m+1), Taylor series to s^23. */
double orc_bm_log(double x) {
  uint64_t bits, mb;
  int e;
  double m, s, s2, p;
  memcpy(&bits, &x, 8);
  e = (int)((bits >> 52) & 0x7FF) - 1022;
  mb = (bits & 0x000FFFFFFFFFFFFFull) | ((uint64_t)1022 << 52);
  memcpy(&m, &mb, 8); /* [0.5, 1) */
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  s = (m - 1.0) / (m + 1.0);
  s2 = s * s;
  p = 1.0 / 23.0;
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
  return (double)e * 0.69314718055994530942 + 2.0 * s * p;
}

/* cos and sin of 2 pi u, u in [0, 1): 4u = q + f, theta = f pi/2 (or
 * (1-f) pi/2 with sine and cosine swapped) in [0, pi/4], Taylor series to
 * theta^19 / theta^18, rotated by the quadrant q. */
void orc_bm_cossin_2pi(double u, double *c, double *s) {
  const double t = 4.0 * u;
  const double q = floor(t);
  double f = t - q, th, t2, ps, pc, sv, cv;
  const int swap = f > 0.5;
  int quad;
  if (swap) f = 1.0 - f;
  th = f * 1.57079632679489661923;
  t2 = th * th;
  ps = -1.0 / 121645100408832000.0;
  ps = ps * t2 + 1.0 / 355687428096000.0;
  ps = ps * t2 - 1.0 / 1307674368000.0;
  ps = ps * t2 + 1.0 / 6227020800.0;
  ps = ps * t2 - 1.0 / 39916800.0;
  ps = ps * t2 + 1.0 / 362880.0;
  ps = ps * t2 - 1.0 / 5040.0;
  ps = ps * t2 + 1.0 / 120.0;
  ps = ps * t2 - 1.0 / 6.0;
  ps = ps * t2 + 1.0;
  pc = -1.0 / 6402373705728000.0;
  pc = pc * t2 + 1.0 / 20922789888000.0;
  pc = pc * t2 - 1.0 / 87178291200.0;
  pc = pc * t2 + 1.0 / 479001600.0;
  pc = pc * t2 - 1.0 / 3628800.0;
  pc = pc * t2 + 1.0 / 40320.0;
  pc = pc * t2 - 1.0 / 720.0;
  pc = pc * t2 + 1.0 / 24.0;
  pc = pc * t2 - 1.0 / 2.0;
  pc = pc * t2 + 1.0;
  sv = th * ps;
  cv = pc;
  if (swap) {
    const double tmp = sv;
    sv = cv;
    cv = tmp;
  }
  quad = (int)q;
  *c = quad == 0 ? cv : (quad == 1 ? -sv : (quad == 2 ? -cv : sv));
  *s = quad == 0 ? sv : (quad == 1 ? cv : (quad == 2 ? -sv : -cv));
}

/* Three N(0,1) normals by Box-Muller from draw blocks 1 and 2. */
static void orc_normals3(unsigned long long seed, int is, uint64_t ctr, double g[3]) {
  uint32_t b1[4], b2[4];
  double u1, u2, u3, u4, rad, c, s;
  orc_draw(seed, is, ctr, 1u, b1);
  orc_draw(seed, is, ctr, 2u, b2);
  u1 = orc_u01d(b1[0], b1[1]);
  u2 = orc_u01d(b1[2], b1[3]);
  u3 = orc_u01d(b2[0], b2[1]);
  u4 = orc_u01d(b2[2], b2[3]);
  rad = sqrt(-2.0 * orc_bm_log(u1));
  orc_bm_cossin_2pi(u2, &c, &s);
  g[0] = rad * c;
  g[1] = rad * s;
  rad = sqrt(-2.0 * orc_bm_log(u3));
  orc_bm_cossin_2pi(u4, &c, &s);
  g[2] = rad * c;
}

/* Uniform and (if it tumbles) normals for one swimmer step; exported for tests. */
float orc_swimmer_uniform(unsigned long long seed, int is, uint64_t ctr) {
  uint32_t b0[4];
  orc_draw(seed, is, ctr, 0u, b0);
  return orc_u01f(b0[0]);
}
void orc_swimmer_normals(unsigned long long seed, int is, uint64_t ctr, double *g) {
  orc_normals3(seed, is, ctr, g);
}

/* SwimmerArray.cu:123-210.  prng[is] is the per-swimmer draw counter;
 * returns the number of tumbles in this call. */
long orc_swimmers_move(int nx, int ny, int nz, int ns, const orc_common *cp, double *sr,
                       double *sv, double *sn, uint64_t *prng, const double *u) {
  const int size[3] = {nx, ny, nz};
  long tumbles = 0;
  int is;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(+ : tumbles)
  for (is = 0; is < ns; is++) {
    double r[3] = {sr[ns * 0 + is], sr[ns * 1 + is], sr[ns * 2 + is]};
    double v[3], rDot[3], rMinus[3], newn[3];
    double norm = 0.0;
    int d, turned = 1;
    float U;
    orc_interp(nx, ny, nz, u, r, v);
    for (d = 0; d < 3; d++) {
      rDot[d] = cp->mobility * cp->P * sn[ns * d + is];
      if (!cp->translational_advection_off) rDot[d] += v[d];
      rMinus[d] = r[d] - sn[ns * d + is] * cp->l;
    }
    for (d = 0; d < 3; d++) {
      sv[ns * d + is] = rDot[d];
      r[d] += rDot[d];
      sr[ns * d + is] = fmod(r[d] + size[d], (double)size[d]);
    }
    U = orc_swimmer_uniform(cp->seed, is, prng[is]);
    if (U < cp->alpha) {
      orc_normals3(cp->seed, is, prng[is], newn);
      for (d = 0; d < 3; d++) norm += newn[d] * newn[d];
      tumbles++;
    } else if (!cp->rotational_advection_off) {
      double vMinus[3], nDot[3];
      orc_interp(nx, ny, nz, u, rMinus, vMinus);
      for (d = 0; d < 3; d++) {
        nDot[d] = v[d] - vMinus[d];
        nDot[d] /= cp->l;
      }
      for (d = 0; d < 3; d++) {
        newn[d] = sn[ns * d + is] + nDot[d];
        norm += newn[d] * newn[d];
      }
    } else {
      turned = 0;
    }
    if (turned) {
      norm = sqrt(norm);
      for (d = 0; d < 3; d++) sn[ns * d + is] = newn[d] / norm;
    }
    prng[is] += 1;
  }
  return tumbles;
}

/* TracerArray.cu:12-43 */
void orc_tracers_move(int nx, int ny, int nz, int nt, double *tr, double *ts, double *tv,
                      const double *u) {
  const int size[3] = {nx, ny, nz};
  int it;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (it = 0; it < nt; it++) {
    double r[3] = {tr[nt * 0 + it], tr[nt * 1 + it], tr[nt * 2 + it]};
    double v[3];
    int d;
    orc_interp(nx, ny, nz, u, r, v);
    for (d = 0; d < 3; d++) {
      tv[nt * d + it] = v[d];
      r[d] += v[d];
      tr[nt * d + it] = fmod(r[d] + size[d], (double)size[d]);
      ts[nt * d + it] += v[d];
    }
  }
}

/* ------------------------------------------------------------------ */
/* One coupled step in System.Step order (System.py:9-16) for one      */
/* lattice, one swimmer array (may be empty) and one tracer array.     */
/* fOld/fNew are swapped by pointer by the caller after every call:    */
/* on return the new distributions are in fNew.                        */
/* ------------------------------------------------------------------ */
long orc_system_step(int nx, int ny, int nz, double tau_s, double tau_b, double *fOld,
                     double *fNew, double *rho, double *u, double *force, int ns,
                     const orc_common *cp, double *sr, double *sv, double *sn, uint64_t *prng,
                     int nt, double *tr, double *ts, double *tv) {
  const long n = (long)nx * ny * nz;
  long tumbles = 0;
  orc_zero_force(n, force);
  if (ns > 0) orc_swimmers_add_forces(nx, ny, nz, ns, cp, sr, sn, force);
  orc_step(nx, ny, nz, tau_s, tau_b, fOld, force, fNew);
  orc_calc_hydro(nx, ny, nz, fNew, force, rho, u);
  if (ns > 0) tumbles = orc_swimmers_move(nx, ny, nz, ns, cp, sr, sv, sn, prng, u);
  if (nt > 0) orc_tracers_move(nx, ny, nz, nt, tr, ts, tv, u);
  return tumbles;
}

/* Run `steps` coupled steps, swapping f internally; the final distributions
 * end up in f0 (copied back if the step count is odd). */
long orc_run(int nx, int ny, int nz, double tau_s, double tau_b, int steps, double *f0,
             double *f1, double *rho, double *u, double *force, int ns, const orc_common *cp,
             double *sr, double *sv, double *sn, uint64_t *prng, int nt, double *tr, double *ts,
             double *tv) {
  const long n = (long)nx * ny * nz;
  double *a = f0, *b = f1;
  long tumbles = 0;
  int s;
  for (s = 0; s < steps; s++) {
    double *t;
    tumbles += orc_system_step(nx, ny, nz, tau_s, tau_b, a, b, rho, u, force, ns, cp, sr, sv,
                               sn, prng, nt, tr, ts, tv);
    t = a; a = b; b = t;
  }
  if (a != f0) memcpy(f0, a, sizeof(double) * 15 * n);
  return tumbles;
}
