// Host build of the device implicit solve (koral_amd/csrc/dev_implicit.hpp,
// the per-lane Newton step the GPU passes run, here in the sequential
// fallback chain) over a batch of independent cells, for comparison with the
// oracle on the CPU: separates differences of the algorithm from differences
// of the GPU's arithmetic.
//
// usage: implicit_check BATCH OUT   (BATCH: params, n, pp[n][nv], uu[n][nv],
// g20[n][20], G20[n][20], dt[n]; OUT: pp[n][nv], info[n][6] as int32)
#define KD_HOST_CHECK 1
#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include "dev_implicit.hpp"

using namespace kd;

static void rd(FILE *f, void *p, size_t n)
{
  if (fread(p, 1, n, f) != n) {
    fprintf(stderr, "implicit_check: short read\n");
    exit(2);
  }
}

template <int MAG>
static void solve_all(const koral_params &P, long n, std::vector<double> &pp, std::vector<double> &uu,
                      const std::vector<double> &g20, const std::vector<double> &G20, const std::vector<double> &dt,
                      std::vector<int> &info)
{
  constexpr int NV = Lay<MAG, 1>::NV;
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < n; i++) {
    Geom g;
    load_geom(g, [&](int c) { return g20[20 * i + c]; }, [&](int c) { return G20[20 * i + c]; });
    ImpInfo inf{};
    solve_implicit_lab<MAG, 1>(P, &uu[NV * i], &pp[NV * i], g, dt[i], 0., inf);
    const int v[6] = {inf.iters, inf.attempt, inf.whichprim, inf.whicheq, inf.whichframe, inf.total};
    for (int k = 0; k < 6; k++) info[6 * i + k] = v[k];
  }
}

int main(int argc, char **argv)
{
  if (argc < 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  koral_params P;
  long n = 0;
  rd(f, &P, sizeof P);
  rd(f, &n, sizeof n);
  const int nv = P.nv;
  std::vector<double> pp(nv * n), uu(nv * n), g20(20 * n), G20(20 * n), dt(n);
  rd(f, pp.data(), 8 * pp.size());
  rd(f, uu.data(), 8 * uu.size());
  rd(f, g20.data(), 8 * g20.size());
  rd(f, G20.data(), 8 * G20.size());
  rd(f, dt.data(), 8 * dt.size());
  fclose(f);
  std::vector<int> info(6 * n);
  if (P.mag)
    solve_all<1>(P, n, pp, uu, g20, G20, dt, info);
  else
    solve_all<0>(P, n, pp, uu, g20, G20, dt, info);
  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(pp.data(), 8, pp.size(), o);
  fwrite(info.data(), 4, info.size(), o);
  fclose(o);
  return 0;
}
