// The README usage example (README.md:83-129) written against the C++ API
// (include/lbswim.hpp): 16^3 lattice, tau 0.25, one swimmer at (8,8,8) facing
// +z, 4096 tracers on the grid points, the System.Step order by hand.
//   readme_coupled OUT.bin [steps=1000] [alpha=0.01]
// Prints the swimmer position every 100 steps and writes fOld, rho, u,
// swimmer r and n and tracer r (doubles, reference layout) to OUT.bin.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "lbswim.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT.bin [steps] [alpha]\n", argv[0]);
    return 2;
  }
  const int steps = argc > 2 ? std::atoi(argv[2]) : 1000;
  const double alpha = argc > 3 ? std::atof(argv[3]) : 0.01;
  try {
    const double tau = 0.25, eta = tau / 3.0;
    lbswim::Lattice lat(16, 16, 16, tau, tau);
    const int n = lat.addr.host->n;
    for (int i = 0; i < n; ++i) lat.data->rho.host[i] = 1.0;
    lat.data->rho.H2D();
    for (size_t i = 0; i < lat.data->u.size; ++i) lat.data->u.host[i] = 0.0;
    lat.data->u.H2D();
    lat.ZeroForce();

    lbswim::CommonParams cp{};
    cp.P = 1e-3;
    cp.l = 0.1;
    cp.alpha = alpha;
    const double a = 0.1, hydro = 1.5;
    cp.mobility = (1.0 / a - 1.0 / hydro) / (6.0 * M_PI * eta);
    cp.translational_advection_off = 0;
    cp.rotational_advection_off = 0;
    cp.seed = 1234;
    lbswim::SwimmerArray sw(1, &cp);
    sw.r.host[0] = sw.r.host[1] = sw.r.host[2] = 8.0;
    sw.n.host[0] = 0.0;
    sw.n.host[1] = 0.0;
    sw.n.host[2] = 1.0;
    sw.r.H2D();
    sw.n.H2D();

    lbswim::TracerArray tr(16 * 16 * 16);
    for (int i = 0; i < 16; ++i)
      for (int j = 0; j < 16; ++j)
        for (int k = 0; k < 16; ++k) {
          const int t = (i * 16 + j) * 16 + k;
          tr.r.host[t] = i;
          tr.r.host[tr.num + t] = j;
          tr.r.host[2 * tr.num + t] = k;
        }
    tr.r.H2D();
    lat.InitFromHydro();

    while (lat.time_step < steps) {
      if (lat.time_step % 100 == 0) {
        sw.r.D2H();
        std::printf("step %4d  swimmer r = (%.12f, %.12f, %.12f)\n", lat.time_step, sw.r.host[0], sw.r.host[1],
                    sw.r.host[2]);
      }
      lat.ZeroForce();
      sw.AddForces(&lat);
      lat.Step();
      lat.CalcHydro();
      sw.Move(&lat);
      tr.Move(&lat);
    }
    lat.data->fOld.D2H();
    lat.data->rho.D2H();
    lat.data->u.D2H();
    sw.r.D2H();
    sw.n.D2H();
    tr.r.D2H();
    std::FILE* f = std::fopen(argv[1], "wb");
    if (!f) {
      std::perror(argv[1]);
      return 1;
    }
    std::fwrite(lat.data->fOld.host, sizeof(double), lat.data->fOld.size, f);
    std::fwrite(lat.data->rho.host, sizeof(double), lat.data->rho.size, f);
    std::fwrite(lat.data->u.host, sizeof(double), lat.data->u.size, f);
    std::fwrite(sw.r.host, sizeof(double), sw.r.size, f);
    std::fwrite(sw.n.host, sizeof(double), sw.n.size, f);
    std::fwrite(tr.r.host, sizeof(double), tr.r.size, f);
    std::fclose(f);
    std::printf("done %d steps\n", lat.time_step);
  } catch (const lbswim::Error& e) {
    std::fprintf(stderr, "lbswim error %d: %s\n", e.code(), e.what());
    return 1;
  }
  return 0;
}
