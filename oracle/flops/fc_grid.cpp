// Operation counts per zone-update of each operator of one RK2IMEX grid step
// (the oracle's o_grid.c compiled with the counting double), on a grid set-up
// read from a file in the host check's input format.
#include <stdio.h>
#include <algorithm>
#include <initializer_list>
#include <vector>
#include "cd.hpp"
#define O_FLOPCOUNT 1
#define double cd
#include "../o_grid.c"
#undef double

static double fc_total(const FlopCount &a) { return (double)(a.add + a.mul + a.div + a.sqrt + a.trans); }

struct FcAcc {
  FlopCount c = {};
  template <class F>
  void run(F fn)
  {
    FlopCount a = g_flops;
    fn();
    c.add += g_flops.add - a.add;
    c.mul += g_flops.mul - a.mul;
    c.div += g_flops.div - a.div;
    c.sqrt += g_flops.sqrt - a.sqrt;
    c.trans += g_flops.trans - a.trans;
  }
};

// op_explicit (o_op_explicit_parts) split into its operators
static void fc_explicit(const oracle_params *P, const o_grid *G, cd dt, cd *u, cd *p, FcAcc &bc, FcAcc &ws_acc,
                        FcAcc sw[3], FcAcc &ct, FcAcc &upd)
{
  const int nv = P->nv;
  bc.run([&] { o_set_bc(P, G, u, p); });
  const long ncell = (long)G->sx * G->sy * G->sz;
  cd *ws = (cd *)calloc(ncell * NWS, sizeof(cd));
  ws_acc.run([&] { calc_wavespeeds(P, G, p, ws, NULL); });
  cd *flb[3] = {NULL, NULL, NULL};
  const long nf[3] = {(long)(G->sx + 1) * G->sy * G->sz, (long)G->sx * (G->sy + 1) * G->sz,
                      (long)G->sx * G->sy * (G->sz + 1)};
  const int n[3] = {G->nx, G->ny, G->nz};
  for (int d = 0; d < 3; d++) {
    if (n[d] == 1) continue;
    flb[d] = (cd *)calloc(nf[d] * nv, sizeof(cd));
    sw[d].run([&] { sweep(P, G, p, ws, d, flb[d]); });
  }
  if (P->mag) ct.run([&] { flux_ct(P, G, flb[0], flb[1], flb[2]); });
  upd.run([&] {
    for (int ix = 0; ix < G->nx; ix++)
      for (int iy = 0; iy < G->ny; iy++)
        for (int iz = 0; iz < G->nz; iz++) {
          long c = cidx(G, ix, iy, iz);
          cd ms[O_NVMAX];
          metric_source(P, G, p + (long)nv * c, ix, iy, ms);
          cd dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
          for (int iv = 0; iv < nv; iv++) {
            cd du = -(flb[0][fxidx(G, ix + 1, iy, iz) * nv + iv] - flb[0][fxidx(G, ix, iy, iz) * nv + iv]) * dt / dx;
            if (G->ny > 1)
              du = du - (flb[1][fyidx(G, ix, iy + 1, iz) * nv + iv] - flb[1][fyidx(G, ix, iy, iz) * nv + iv]) * dt / dy;
            if (G->nz > 1)
              du = du - (flb[2][fzidx(G, ix, iy, iz + 1) * nv + iv] - flb[2][fzidx(G, ix, iy, iz) * nv + iv]) * dt / dz;
            u[(long)nv * c + iv] = u[(long)nv * c + iv] + du + ms[iv] * dt;
          }
        }
  });
  for (int d = 0; d < 3; d++) free(flb[d]);
  free(ws);
}

int fc_grid_counts(FILE *f)
{
  int32_t hdr[12];
  if (fread(hdr, 4, 12, f) != 12 || hdr[0] != 0x4b4f5241 || hdr[11] != (int)sizeof(oracle_params)) return 2;
  oracle_params P;
  if (fread(&P, sizeof P, 1, f) != 1) return 2;
  o_grid G;
  G.nx = hdr[1]; G.ny = hdr[2]; G.nz = hdr[3]; G.ng = hdr[4];
  for (int k = 0; k < 6; k++) G.bc[k] = hdr[5 + k];
  G.ngx = G.ng; G.ngy = G.ny > 1 ? G.ng : 0; G.ngz = G.nz > 1 ? G.ng : 0;
  G.sx = G.nx + 2 * G.ngx; G.sy = G.ny + 2 * G.ngy; G.sz = G.nz + 2 * G.ngz;
  const int nd[3] = {G.nx, G.ny, G.nz};
  for (int d = 0; d < 3; d++) {
    long n = nd[d] + 2 * G.ng + 1;
    cd *y = (cd *)malloc(sizeof(cd) * n);
    if (fread(y, sizeof(cd), n, f) != (size_t)n) return 2;
    G.xb[d] = y;
  }
  const long npl = (long)G.sy * G.sx, nplx = (long)G.sy * (G.sx + 1), nply = (long)(G.sy + 1) * G.sx;
  const cd **tabs[7] = {&G.gc, &G.Gc, &G.kr, &G.gfx, &G.Gfx, &G.gfy, &G.Gfy};
  const long tn[7] = {20 * npl, 20 * npl, 64 * npl, 20 * nplx, 20 * nplx, 20 * nply, 20 * nply};
  // the file holds the device layout [c][point]; the oracle's tables are
  // [point][c]
  const long tc[7] = {20, 20, 64, 20, 20, 20, 20};
  for (int k = 0; k < 7; k++) {
    std::vector<double> raw(tn[k]);
    if (fread(raw.data(), sizeof(double), tn[k], f) != (size_t)tn[k]) return 2;
    const long np = tn[k] / tc[k];
    cd *y = (cd *)malloc(sizeof(cd) * tn[k]);
    for (long q = 0; q < np; q++)
      for (long c = 0; c < tc[k]; c++) y[q * tc[k] + c] = raw[c * np + q];
    *tabs[k] = y;
  }
  const int nv = P.nv;
  const long nc = (long)G.sx * G.sy * G.sz;
  cd *p = (cd *)malloc(sizeof(cd) * nv * nc), *u = (cd *)calloc(nv * nc, sizeof(cd));
  if (fread(p, sizeof(cd), nv * nc, f) != (size_t)(nv * nc)) return 2;
  int *flags = (int *)calloc(O_NFLAGS * nc, sizeof(int));
  // initial state as the step tests take it: u = p2u(p) everywhere, ghosts
  for (int iz = -G.ngz; iz < G.nz + G.ngz; iz++)
    for (int iy = -G.ngy; iy < G.ny + G.ngy; iy++)
      for (int ix = -G.ngx; ix < G.nx + G.ngx; ix++) {
        o_geom g;
        o_geom_cell(&G, ix, iy, &g);
        const long c = cidx(&G, ix, iy, iz);
        o_p2u(&P, p + (long)nv * c, u + (long)nv * c, &g);
      }
  o_set_bc(&P, &G, u, p);
  const cd dt = cd(1.) / o_tstepden(&P, &G, p);
  const cd gamma = cd(1.) - cd(1.) / sqrt(cd(2.));
  // the operators of one step in the order of o_step_rk2imex (the stage
  // algebra is counted from its statements below)
  if (getenv("FC_U2P_HIST")) {
    // per-cell operation counts of one u2p call on the initial state after
    // one explicit update (the cells the deferred pass takes stand out)
    FcAcc b2, w2, s2[3], c2, up2;
    fc_explicit(&P, &G, dt, u, p, b2, w2, s2, c2, up2);
    std::vector<double> cnt;
    const int nvv = P.nv;
    for (int ix = 0; ix < G.nx; ix++)
      for (int iy = 0; iy < G.ny; iy++)
        for (int iz = 0; iz < G.nz; iz++) {
          const long c = cidx(&G, ix, iy, iz);
          o_geom g;
          o_geom_cell(&G, ix, iy, &g);
          int corr[3] = {0, 0, 0}, fix[2] = {0, 0};
          FlopCount a = g_flops;
          const int r = o_u2p(&P, u + (long)nvv * c, p + (long)nvv * c, &g, corr, fix);
          cnt.push_back(fc_total(g_flops) - fc_total(a));
          (void)r;
        }
    std::sort(cnt.begin(), cnt.end());
    double tot = 0.;
    for (double v : cnt) tot += v;
    const size_t m = cnt.size();
    printf("{\"cells\": %zu, \"mean\": %.1f, \"median\": %.1f, \"p99\": %.1f, \"p999\": %.1f, \"max\": %.1f, "
           "\"top_0.1pct_share\": %.3f}\n", m, tot / m, cnt[m / 2], cnt[m * 99 / 100], cnt[m * 999 / 1000], cnt[m - 1],
           [&] { double s = 0.; for (size_t k = m - m / 1000; k < m; k++) s += cnt[k]; return s / tot; }());
    return 0;
  }
  FcAcc bc, ws, sw[3], ct, upd, u2p, imp, pol;
  int counts[6] = {0};
  for (int stage = 0; stage < 2; stage++) {
    imp.run([&] { o_op_implicit(&P, &G, 0., dt * gamma, u, p, flags, counts); });
    u2p.run([&] { o_calc_u2p(&P, &G, u, p, flags, counts + 3); });
    pol.run([&] { o_do_correct(&P, &G, u, p); });
    fc_explicit(&P, &G, dt, u, p, bc, ws, sw, ct, upd);
    u2p.run([&] { o_calc_u2p(&P, &G, u, p, flags, counts + 3); });
    pol.run([&] { o_do_correct(&P, &G, u, p); });
  }
  const double zones = (double)G.nx * G.ny * G.nz;
  // 11 linear combinations per (cell, variable) of ko.c:444-570: 25 FP64
  // operations in all
  const double comb = 25. * nv;
  double step = comb;
  for (FcAcc *a : {&bc, &ws, &sw[0], &sw[1], &sw[2], &ct, &upd, &u2p, &imp, &pol}) step += fc_total(a->c) / zones;
  printf("{\"zones\": %.0f, \"sweep_x\": %.1f, \"sweep_y\": %.1f, \"sweep_z\": %.1f, \"wavespeeds\": %.1f, "
         "\"set_bc\": %.1f, \"flux_ct\": %.1f, \"update\": %.1f, \"u2p\": %.1f, \"implicit\": %.1f, "
         "\"polar\": %.1f, \"comb\": %.1f, \"step\": %.1f, \"u2p_trans\": %.2f, \"sweep_trans\": %.2f}\n",
         zones, fc_total(sw[0].c) / zones, fc_total(sw[1].c) / zones, fc_total(sw[2].c) / zones,
         fc_total(ws.c) / zones, fc_total(bc.c) / zones, fc_total(ct.c) / zones, fc_total(upd.c) / zones,
         fc_total(u2p.c) / zones, fc_total(imp.c) / zones, fc_total(pol.c) / zones, comb, step,
         u2p.c.trans / zones, (sw[0].c.trans + sw[1].c.trans + sw[2].c.trans) / zones);
  return 0;
}
