// Host executor of the grid operators (koral_amd/csrc/grid_ops.hpp compiled as
// plain C++ with KD_HOST_CHECK) with bounds-checked buffers.
//
// Every device buffer of a grid (grid_buffers, the sizes koral_hip_grid_create
// allocates) and every metric table is placed between two 64 GiB PROT_NONE
// guard regions.  Mode 0 puts each buffer's first byte at a page start (any
// access below it faults), mode 1 its last byte at a page end (any access
// above it faults).  The index functions of dev_grid.hpp stop the run when a
// coordinate leaves the box they serve, which catches accesses that stay
// inside a buffer but land on another variable, face or plane.  Buffers that
// the library leaves uninitialised start as NaN (lists as 0xffffffff), so a
// read of an unwritten value shows in the results or trips an index check.
//
// The operators run item by item in the step order of grid_ops.hpp (items of
// one launch in parallel with OpenMP, as they are independent on the
// device).  The implicit sweep runs OpImplicit per cell, the solve the
// device's per-attempt passes reproduce.  With a sample stride s > 1 a launch
// runs only its first and last 4096 items and every s-th item in between
// (for index checks at full size, where the results are not meaningful).
//
// Usage: grid_check IN OUT MODE STRIDE CMD [ARG]
//   IN:  int32 magic 0x4b4f5241, nx, ny, nz, ng, bc[6], sizeof(koral_params);
//        koral_params; xb[3] (n_d + 2 ng + 1 doubles each); gc, Gc [20][SY][SX],
//        kr [64][SY][SX], gfx, Gfx [20][SY][SX+1], gfy, Gfy [20][SY+1][SX];
//        p [SZ][SY][SX][NV]; with CMD explicit also u [SZ][SY][SX][NV]
//   CMD step N:        u = p2u(p), set_bc, dt = 1/tstepden, then N RK2IMEX steps
//                      (each at the previous step's dt, as ko.c's loop)
//   CMD explicit DT:   op_explicit at t = 0, dt = DT on the given p, u
//   CMD legacy_bc:     the ghost fill in the launch geometry of the earlier
//                      build (one thread per edge position running the edge
//                      recipe in sequence)
//   CMD legacy_explicit DT: op_explicit as the earlier build launched it (its
//                      ghost fill, one fused reconstruction + full floors +
//                      flux kernel per direction), on the given p, u
//   OUT: p, u [SZ][SY][SX][NV] doubles, flags [SZ][SY][SX][7] int32, then per
//        step tstepdenmax (double) and counters[6] (int32)
#define KD_HOST_CHECK 1
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <unistd.h>
#include <type_traits>
#include <vector>
#include "grid_ops.hpp"

using namespace kd;

namespace {

int g_mode = 0;

struct Region {
  const char *name;
  char *lo, *hi;  // [lo, hi) the buffer's bytes
  char *map;
  size_t maplen;
};
std::vector<Region> g_regions;

const size_t GUARD = (size_t)1 << 36;

void *guarded(size_t bytes, const char *name)
{
  if (bytes == 0) bytes = 1;
  const size_t pg = (size_t)sysconf(_SC_PAGESIZE);
  const size_t span = (bytes + pg - 1) / pg * pg;
  const size_t len = GUARD + span + GUARD;
  char *m = (char *)mmap(nullptr, len, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (m == MAP_FAILED) {
    fprintf(stderr, "mmap failed for %s (%zu bytes)\n", name, bytes);
    exit(3);
  }
  if (mprotect(m + GUARD, span, PROT_READ | PROT_WRITE) != 0) {
    fprintf(stderr, "mprotect failed for %s\n", name);
    exit(3);
  }
  char *user = g_mode == 0 ? m + GUARD : m + GUARD + span - bytes;
  g_regions.push_back(Region{name, user, user + bytes, m, len});
  return user;
}

// a fault inside a guard region names the buffer and the offset
void on_segv(int sig, siginfo_t *si, void *)
{
  char *a = (char *)si->si_addr;
  char msg[256];
  int n = 0;
  for (const Region &r : g_regions)
    if (a >= r.map && a < r.map + r.maplen) {
      long off = a < r.lo ? -(long)(r.lo - a) : (long)(a - r.lo);
      n = snprintf(msg, sizeof msg, "GUARD FAULT: %s at byte offset %ld (buffer %ld bytes, mode %d)\n", r.name, off,
                   (long)(r.hi - r.lo), g_mode);
      break;
    }
  if (!n) n = snprintf(msg, sizeof msg, "SEGV at %p outside every guarded buffer\n", (void *)a);
  ssize_t w = write(2, msg, (size_t)n);
  (void)w;
  signal(sig, SIG_DFL);
  raise(sig);
}

template <class T>
void fill(T *p, long n, T v)
{
  for (long i = 0; i < n; i++) p[i] = v;
}

// ---------------------------------------------------------------------------
// executor
// ---------------------------------------------------------------------------
struct HostExec {
  long stride = 1;
  static constexpr long EDGE = 4096;
  bool want(long i, long n) const { return stride <= 1 || i < EDGE || i >= n - EDGE || i % stride == 0; }
  template <class Op>
  void each(const Op &op)
  {
    const long n = op.size();
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; i++)
      if (want(i, n)) op(i);
  }
  template <class Op>
  void each_on(int, const Op &op)
  {
    each(op);
  }
  void fork(int) {}
  void join(int) {}
  template <class Op>
  void max(const Op &op, double *out)
  {
    const long n = op.size();
    double m = 0.;
#pragma omp parallel for schedule(dynamic, 256) reduction(max : m)
    for (long i = 0; i < n; i++)
      if (want(i, n)) {
        const double v = op(i);
        if (v > m) m = v;
      }
    if (m > *out) *out = m;
  }
  template <class Op>
  void count(const Op &op, int *out)
  {
    const long n = op.size();
    long c0 = 0, c1 = 0, c2 = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : c0, c1, c2)
    for (long i = 0; i < n; i++)
      if (want(i, n)) {
        const Cnt3 c = op(i);
        c0 += c.v[0];
        c1 += c.v[1];
        c2 += c.v[2];
      }
    out[0] += (int)c0;
    out[1] += (int)c1;
    out[2] += (int)c2;
  }
  template <int MAG>
  void implicit_cells(const OpImplicit<MAG> &op, int *out)
  {
    count(op, out);
  }
};

// The ghost fill of edges in the earlier build's launch geometry: one thread
// per (corner of the a-b edge, position k along the third dimension) doing
// the copies of the far ghosts, then the two diagonal averages, in sequence.
template <int MAG, int RAD>
struct LegacyEdge {
  koral_params P;
  DGrid G;
  double *p, *u;
  int a, b;
  int c() const { return 3 - a - b; }
  int ngc() const { return c() == 0 ? G.ngx : (c() == 1 ? G.ngy : G.ngz); }
  int nk() const { return dimn(G, c()) + 2 * ngc(); }
  long size() const { return 4L * nk(); }
  long cell(int va, int vb, int k, int *xy) const
  {
    int q[3];
    q[a] = va;
    q[b] = vb;
    q[c()] = k;
    if (xy) {
      xy[0] = q[0];
      xy[1] = q[1];
    }
    return cidx(G, q[0], q[1], q[2]);
  }
  void set(long d, int xy[2], const double *pv) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    store_p<MAG, RAD>(p, G.ncell, d, pv);
    Geom g;
    geom_cell(G, xy[0], xy[1], g);
    double uu[NV];
    kd::p2u<MAG, RAD>(P, pv, uu, g);
    store_p<MAG, RAD>(u, G.ncell, d, uu);
  }
  void operator()(long t) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const int NG = G.ng, na = dimn(G, a), nb = dimn(G, b);
    const int corner = (int)(t / nk()), k = -ngc() + (int)(t % nk());
    const int ahi = corner >> 1, bhi = corner & 1;
    const int ea = ahi ? na : -1, ia = ahi ? na - 1 : 0, ea2 = ahi ? na + 1 : -2;
    const int eb = bhi ? nb : -1, ib = bhi ? nb - 1 : 0, eb2 = bhi ? nb + 1 : -2;
    double pv[NV];
    int xy[2];
    for (int i = 0; i < NG - 1; i++) {
      const int ga = ahi ? na + i + 1 : -NG + i, gb = bhi ? nb + i + 1 : -NG + i;
      load_p<MAG, RAD>(p, G.ncell, cell(ga, ib, k, nullptr), pv);
      set(cell(ga, eb, k, xy), xy, pv);
      load_p<MAG, RAD>(p, G.ncell, cell(ia, gb, k, nullptr), pv);
      set(cell(ea, gb, k, xy), xy, pv);
    }
    const long q1 = cell(ea, ib, k, nullptr), q2 = cell(ia, eb, k, nullptr);
    for (int iv = 0; iv < NV; iv++) pv[iv] = .5 * (p[iv * G.ncell + q1] + p[iv * G.ncell + q2]);
    set(cell(ea, eb, k, xy), xy, pv);
    const long r1 = cell(ea2, eb, k, nullptr), r2 = cell(ea, eb2, k, nullptr);
    for (int iv = 0; iv < NV; iv++) pv[iv] = .5 * (p[iv * G.ncell + r1] + p[iv * G.ncell + r2]);
    set(cell(ea2, eb2, k, xy), xy, pv);
  }
};

// the earlier build's face fill: one item per cell of the whole
// ghost-inclusive box, writing the cells with exactly one coordinate outside
// the domain
template <int MAG, int RAD>
struct LegacyFaces {
  koral_params P;
  DGrid G;
  double *p, *u;
  long size() const { return G.ncell; }
  void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const Box bx{-G.ngx, -G.ngy, -G.ngz, G.sx, G.sy, G.sz};
    int q[3];
    bx.at(i, q[0], q[1], q[2]);
    int d = -1, hi = 0, out = 0;
    for (int k = 0; k < 3; k++) {
      const int n = dimn(G, k);
      if (n > 1 && q[k] < 0) { d = k; hi = 0; out++; }
      if (n > 1 && q[k] >= n) { d = k; hi = 1; out++; }
    }
    if (out != 1) return;
    int src[3] = {q[0], q[1], q[2]};
    const int n = dimn(G, d), type = G.bc[2 * d + hi];
    bool flip = false;
    if (type == BC_PERIODIC) {
      if (src[d] < 0) src[d] += n;
      if (src[d] > n - 1) src[d] -= n;
    } else if (type == BC_REFLECT) {
      src[d] = hi ? 2 * n - src[d] - 1 : -src[d] - 1;
      flip = true;
    } else {
      src[d] = hi ? n - 1 : 0;
    }
    double pv[NV], uu[NV];
    load_p<MAG, RAD>(p, G.ncell, cidx(G, src[0], src[1], src[2]), pv);
    if (flip) {
      pv[2 + d] = -pv[2 + d];
      if (MAG) pv[6 + d] = -pv[6 + d];
      if (RAD) pv[Lay<MAG, RAD>::EE + 1 + d] = -pv[Lay<MAG, RAD>::EE + 1 + d];
    }
    const long c = cidx(G, q[0], q[1], q[2]);
    store_p<MAG, RAD>(p, G.ncell, c, pv);
    Geom g;
    geom_cell(G, q[0], q[1], g);
    kd::p2u<MAG, RAD>(P, pv, uu, g);
    store_p<MAG, RAD>(u, G.ncell, c, uu);
  }
};

template <class Ex>
void legacy_bc(Ex &ex, const GridBufs &s)
{
  const DGrid &G = s.G;
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    constexpr int M = T::MAG, R = T::RAD;
    ex.each(LegacyFaces<M, R>{s.P, G, s.p, s.u});
    if (G.ny > 1) {
      ex.each(LegacyEdge<M, R>{s.P, G, s.p, s.u, 0, 1});
      if (G.nz > 1) {
        ex.each(LegacyEdge<M, R>{s.P, G, s.p, s.u, 0, 2});
        ex.each(LegacyEdge<M, R>{s.P, G, s.p, s.u, 1, 2});
        ex.each(OpBcTotal<M, R>{s.P, G, s.p, s.u});
      }
    }
  });
}

// The earlier build's fused face kernel: reconstruction of both sides from
// the six cells around the face, the full face floors (no deferral), fluxes
// and HLL/LF, one item per face of box b
template <int MAG, int RAD, int ORDER, int D>
struct LegacyFlux {
  koral_params P;
  DGrid G;
  const double *p, *ws;
  double *flb;
  Box b;
  long size() const { return b.size(); }
  void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    long c[6];
    double dx[6];
    for (int o = 0; o < 6; o++) {
      int q[3] = {ix, iy, iz};
      q[D] += o - 3;
      const bool need = ORDER > 1 || (o >= 1 && o <= 4);
      c[o] = need ? cidx(G, q[0], q[1], q[2]) : 0;
      dx[o] = (ORDER == 2 && need) ? size_x(G, q[D], D) : 0.;
    }
    const long n = G.ncell;
    double pL[NV], pR[NV];
    for (int iv = 0; iv < NV; iv++) {
      double v[6];
      for (int o = 0; o < 6; o++) v[o] = p[iv * n + c[o]];
      if (ORDER == 4) {
        pL[iv] = mp5(v[0], v[1], v[2], v[3], v[4]);
        pR[iv] = mp5(v[5], v[4], v[3], v[2], v[1]);
      } else if (ORDER == 2) {
        double ul, ur;
        ppm(v, dx, ul, ur);
        pL[iv] = ur;
        ppm(v + 1, dx + 1, ul, ur);
        pR[iv] = ul;
      } else {
        const double th = P.minmod_theta;
        pL[iv] = v[2] + .5 * minmod3(th * (v[3] - v[2]), .5 * (v[3] - v[1]), th * (v[2] - v[1]));
        pR[iv] = v[3] - .5 * minmod3(th * (v[4] - v[3]), .5 * (v[4] - v[2]), th * (v[3] - v[2]));
      }
    }
    Geom g;
    geom_face(G, D, ix, iy, g);
    floors_mhd<MAG, RAD>(P, pL, g);
    floors_mhd<MAG, RAD>(P, pR, g);
    face_flux<MAG, RAD, D>(P, G, ws, flb, ix, iy, iz, pL, pR, g);
  }
};

// op_explicit as the earlier build launched it: its ghost fill, wave speeds,
// the fused face kernel per direction in the order x1, x2, x3, flux-CT and
// the update
template <class Ex>
void legacy_explicit(Ex &ex, const GridBufs &s)
{
  const DGrid &G = s.G;
  const koral_params &P = s.P;
  legacy_bc(ex, s);
  tstepden(ex, s);
  const int tt = P.mag ? 1 : 0;
  const int n[3] = {G.nx, G.ny, G.nz};
  for (int d = 0; d < 3; d++) {
    if (n[d] == 1) continue;
    int lo[3], cnt[3];
    for (int k = 0; k < 3; k++) {
      if (k == d) { lo[k] = 0; cnt[k] = n[k] + 1; }
      else if (n[k] == 1) { lo[k] = 0; cnt[k] = 1; }
      else { lo[k] = -tt; cnt[k] = n[k] + 2 * tt; }
    }
    const Box fb{lo[0], lo[1], lo[2], cnt[0], cnt[1], cnt[2]};
    dispatch(P, [&](auto t) {
      using T = decltype(t);
      constexpr int M = T::MAG, R = T::RAD;
      auto sweep = [&](auto dd) {
        constexpr int DD = decltype(dd)::value;
        if (P.int_order == 4)
          ex.each(LegacyFlux<M, R, 4, DD>{P, G, s.p, s.ws, s.flb[DD], fb});
        else if (P.int_order == 2)
          ex.each(LegacyFlux<M, R, 2, DD>{P, G, s.p, s.ws, s.flb[DD], fb});
        else
          ex.each(LegacyFlux<M, R, 1, DD>{P, G, s.p, s.ws, s.flb[DD], fb});
      };
      if (d == 0) sweep(std::integral_constant<int, 0>{});
      else if (d == 1) sweep(std::integral_constant<int, 1>{});
      else sweep(std::integral_constant<int, 2>{});
    });
  }
  Flx F{s.flb[0], s.flb[1], s.flb[2], G.nfx, G.nfy, G.nfz};
  if (P.mag) {
    ex.each(OpEmf{G, F, s.emf});
    ex.each(OpCtFaces{G, F, s.emf});
  }
  dispatch(P, [&](auto t) {
    using T = decltype(t);
    ex.each(OpUpdate<T::MAG, T::RAD>{P, G, s.p, s.u, F, s.clock});
  });
}

void read_all(FILE *f, void *dst, size_t bytes, const char *what)
{
  if (fread(dst, 1, bytes, f) != bytes) {
    fprintf(stderr, "input truncated at %s\n", what);
    exit(2);
  }
}

void aos_to_soa(const double *aos, double *soa, long nc, int nv)
{
  for (long c = 0; c < nc; c++)
    for (int iv = 0; iv < nv; iv++) soa[iv * nc + c] = aos[c * nv + iv];
}

}  // namespace

int main(int argc, char **argv)
{
  if (argc < 6) {
    fprintf(stderr, "usage: grid_check IN OUT MODE STRIDE CMD [ARG]\n");
    return 2;
  }
  g_mode = atoi(argv[3]);
  HostExec ex;
  ex.stride = atol(argv[4]);
  const char *cmd = argv[5];
  struct sigaction sa;
  memset(&sa, 0, sizeof sa);
  sa.sa_sigaction = on_segv;
  sa.sa_flags = SA_SIGINFO;
  sigaction(SIGSEGV, &sa, nullptr);
  sigaction(SIGBUS, &sa, nullptr);

  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[12];
  read_all(f, hdr, sizeof hdr, "header");
  if (hdr[0] != 0x4b4f5241 || hdr[11] != (int)sizeof(koral_params)) {
    fprintf(stderr, "bad header\n");
    return 2;
  }
  koral_params P;
  read_all(f, &P, sizeof P, "params");
  koral_grid_desc D{};
  D.nx = hdr[1];
  D.ny = hdr[2];
  D.nz = hdr[3];
  D.ng = hdr[4];
  for (int k = 0; k < 6; k++) D.bc[k] = hdr[5 + k];
  GridBufs s{};
  s.tune = default_tune();
  // KORAL_HC_TUNE selects launch choices that change which functors run:
  // "fused" (reconstruction inside the flux functor for every direction),
  // "lane" (one lane per face or cell in the deferred passes), "pre" (the
  // update loading u and p with the fluxes), "combrt" (the stage algebra
  // with a run-time variable count), "budget" (a u2p budget of 3 residual
  // evaluations, so most cells take the deferred pass), or several
  if (const char *tn = getenv("KORAL_HC_TUNE")) {
    if (strstr(tn, "fused"))
      for (int d = 0; d < 3; d++) s.tune.fused[d] = 1;
    if (strstr(tn, "lane")) s.tune.floor_wave = 0;
    if (strstr(tn, "pre")) s.tune.update_pre = 1;
    if (strstr(tn, "combrt")) s.tune.comb_nv = 0;
    if (strstr(tn, "budget")) s.tune.u2p_budget = 3;
  }
  s.P = P;
  DGrid &G = s.G;
  grid_layout(&D, G);
  const int nv = P.nv;
  const int nd3[3] = {D.nx, D.ny, D.nz};
  const double **xbs[3] = {&G.xb0, &G.xb1, &G.xb2};
  for (int d = 0; d < 3; d++) {
    const long n = nd3[d] + 2 * D.ng + 1;
    double *x = (double *)guarded(n * sizeof(double), d == 0 ? "xb0" : d == 1 ? "xb1" : "xb2");
    read_all(f, x, n * sizeof(double), "xb");
    *xbs[d] = x;
  }
  struct Tab {
    const double **dst;
    long n;
    const char *name;
  } tabs[7] = {{&G.gc, 20 * G.npl, "gc"},      {&G.Gc, 20 * G.npl, "Gc"},      {&G.kr, 64 * G.npl, "kr"},
               {&G.gfx, 20 * G.nplx, "gfx"},   {&G.Gfx, 20 * G.nplx, "Gfx"},   {&G.gfy, 20 * G.nply, "gfy"},
               {&G.Gfy, 20 * G.nply, "Gfy"}};
  for (const Tab &t : tabs) {
    double *x = (double *)guarded(t.n * sizeof(double), t.name);
    read_all(f, x, t.n * sizeof(double), t.name);
    *t.dst = x;
  }
  grid_buffers(s, [&](auto *&ptr, long count, const char *name) {
    using T = typename std::remove_reference<decltype(*ptr)>::type;
    ptr = (T *)guarded((size_t)count * sizeof(T), name);
  });
  const long nc = G.ncell;
  std::vector<double> aos((size_t)nc * nv);
  read_all(f, aos.data(), aos.size() * sizeof(double), "p");
  // the library zeroes p, u, flags, fluxes and wave speeds at creation and
  // leaves the rest unset: here the rest starts as NaN / all ones
  const double qnan = __builtin_nan("");
  fill(s.u, nv * nc, 0.);
  fill(s.fl, (long)NFLAGS * nc, (signed char)0);
  fill(s.ws, (long)NWS * nc, 0.);
  for (int d = 0; d < 3; d++)
    if (s.flb[d]) fill(s.flb[d], nv * (d == 0 ? G.nfx : d == 1 ? G.nfy : G.nfz), 0.);
  for (int d = 0; d < 3; d++)
    if (s.fs[d]) {
      fill(s.fs[d], 2 * nv * face_states_cap(G), qnan);
      fill(s.flist[d], face_states_cap(G), 0xffffffffu);
    }
  const long nd = (long)G.nx * G.ny * G.nz;
  fill(s.ulist, nd, 0xffffffffu);
  fill(s.uwhy, nd, (unsigned char)0xff);
  fill(s.fcnt, 4L, 0x7fffffffu);
  fill(s.emf, 3L * (G.nx + 1) * (G.ny + 1) * (G.nz + 1), qnan);
  double *work[9] = {s.pn, s.un, s.ut0, s.ut1, s.uforget, s.dut1, s.dut2, s.drt1, s.drt2};
  for (double *w : work) fill(w, nv * nc, qnan);
  fill(s.mark, nc, -1);
  aos_to_soa(aos.data(), s.p, nc, nv);

  std::vector<double> tsds;
  std::vector<int> ctrs;
  if (!strcmp(cmd, "step")) {
    const int nsteps = argc > 6 ? atoi(argv[6]) : 1;
    grid_p2u(ex, s);
    set_bc(ex, s);
    s.tsd[0] = 0.;
    tstepden(ex, s);
    double t = 0., dt = 1. / s.tsd[0];
    for (int k = 0; k < nsteps; k++) {
      s.clock[0] = t;
      s.clock[1] = dt;
      s.tsd[0] = 0.;
      for (int q = 0; q < 8; q++) s.counters[q] = 0;
      step_rk2imex(ex, s);
      tsds.push_back(s.tsd[0]);
      for (int q = 0; q < 6; q++) ctrs.push_back(s.counters[q]);
      t += dt;
      dt = 1. / s.tsd[0];
      fprintf(stderr, "step %d done: tstepdenmax %.17e\n", k, s.tsd[0]);
    }
  } else if (!strcmp(cmd, "explicit")) {
    const double dt = argc > 6 ? atof(argv[6]) : 0.;
    read_all(f, aos.data(), aos.size() * sizeof(double), "u");
    aos_to_soa(aos.data(), s.u, nc, nv);
    s.clock[0] = 0.;
    s.clock[1] = dt;
    s.tsd[0] = 0.;
    explicit_op(ex, s);
    tsds.push_back(s.tsd[0]);
    for (int q = 0; q < 6; q++) ctrs.push_back(0);
  } else if (!strcmp(cmd, "legacy_explicit")) {
    const double dt = argc > 6 ? atof(argv[6]) : 0.;
    read_all(f, aos.data(), aos.size() * sizeof(double), "u");
    aos_to_soa(aos.data(), s.u, nc, nv);
    s.clock[0] = 0.;
    s.clock[1] = dt;
    s.tsd[0] = 0.;
    legacy_explicit(ex, s);
    tsds.push_back(s.tsd[0]);
    for (int q = 0; q < 6; q++) ctrs.push_back(0);
  } else if (!strcmp(cmd, "legacy_bc")) {
    grid_p2u(ex, s);
    legacy_bc(ex, s);
  } else {
    fprintf(stderr, "unknown command %s\n", cmd);
    return 2;
  }
  fclose(f);

  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  for (const double *x : {(const double *)s.p, (const double *)s.u}) {
    for (long c = 0; c < nc; c++)
      for (int iv = 0; iv < nv; iv++) aos[c * nv + iv] = x[iv * nc + c];
    fwrite(aos.data(), sizeof(double), aos.size(), o);
  }
  std::vector<int32_t> fl((size_t)nc * NFLAGS);
  for (long c = 0; c < nc; c++)
    for (int q = 0; q < NFLAGS; q++) fl[c * NFLAGS + q] = s.fl[q * nc + c];
  fwrite(fl.data(), sizeof(int32_t), fl.size(), o);
  for (size_t k = 0; k < tsds.size(); k++) {
    fwrite(&tsds[k], sizeof(double), 1, o);
    fwrite(&ctrs[6 * k], sizeof(int), 6, o);
  }
  fclose(o);
  fprintf(stderr, "grid_check: %s done, mode %d, stride %ld, %zu guarded buffers\n", cmd, g_mode, ex.stride,
          g_regions.size());
  return 0;
}
