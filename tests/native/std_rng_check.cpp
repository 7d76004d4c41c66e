// Reference values for the composition workload's initial state, from the C++ standard
// library the reference is built with: std::mt19937(seed) driving
// std::uniform_real_distribution<double>(-0.5, 0.5) (kh.cpp:38-39).
//   std_rng_check draws <seed> <n>        -> n values, one per line (%a)
//   std_rng_check khcomp <gid> <nx> <ny> <x0> <y0> <dx> <dy> <nlset>
//        -> the conserved fields of one block's interior cells in the pgen's loop order
//           (kh.cpp:274-310, iprob 6, the deck inputs/kh-comp-advection.in), one cell
//           per line: 5 + 7 + 2 nlset values (%a)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

int main(int argc, char **argv) {
  if (argc >= 4 && std::strcmp(argv[1], "draws") == 0) {
    std::mt19937 gen((unsigned)std::strtoul(argv[2], nullptr, 10));
    std::uniform_real_distribution<double> ran(-0.5, 0.5);
    const long n = std::strtol(argv[3], nullptr, 10);
    for (long i = 0; i < n; ++i) std::printf("%a\n", ran(gen));
    return 0;
  }
  if (argc >= 10 && std::strcmp(argv[1], "khcomp") == 0) {
    const unsigned gid = (unsigned)std::strtoul(argv[2], nullptr, 10);
    const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]);
    const double x0 = std::strtod(argv[5], nullptr), y0 = std::strtod(argv[6], nullptr);
    const double dx = std::strtod(argv[7], nullptr), dy = std::strtod(argv[8], nullptr);
    const int nlset = std::atoi(argv[9]);
    (void)x0;
    (void)dx;
    const double vflow = 1.0, drat = 0.1, amp = 0.01, ofrac = 0.65, ye = 0.4983, pamp = 2.5;
    const double gam = 1.666666666666667, gm1 = gam - 1.0, rho = 1.0;
    std::mt19937 gen(gid);
    std::uniform_real_distribution<double> ran(-0.5, 0.5);
    const int nv = 5 + 7 + 2 * nlset;
    std::vector<double> u(nv);
    for (int j = 0; j < ny; ++j) {
      const double y = y0 + (j + 0.5) * dy;
      for (int i = 0; i < nx; ++i) {
        for (double &v : u) v = 0.0;
        u[0] = rho;
        u[1] = vflow + amp * ran(gen);
        u[2] = amp * ran(gen);
        u[3] = 0.0;
        u[5 + 2] = ofrac * rho;
        u[5 + 3] = (1.0 - ofrac) * rho;
        u[5 + 6] = ye * rho;
        if (std::abs(y) < 0.25) {
          u[0] = drat * rho;
          u[1] = -drat * (vflow + amp * ran(gen)) * rho;
          u[2] = drat * amp * ran(gen) * rho;
          u[5 + 2] = 1.0 * drat * rho;
          u[5 + 3] = 0.0 * drat * rho;
          u[5 + 6] = 0.5 * drat * rho;
        }
        if (nlset > 0) u[5 + 7] = (0.25 - std::abs(y)) * u[0];
        u[4] = pamp / gm1 + 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0];
        for (int v = 0; v < nv; ++v) std::printf("%a%c", u[v], v + 1 < nv ? ' ' : '\n');
      }
    }
    return 0;
  }
  std::fprintf(stderr, "usage: see the header comment\n");
  return 2;
}
