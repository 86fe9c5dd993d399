// Grid operators of one KORAL timestep on device-resident SoA state (the
// per-thread work and the step order are in grid_ops.hpp): the HIP executor
// that launches each operator as one kernel on the grid's stream, and the
// grid C-ABI of include/koral_hip.h.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>
#include <string>
#include <vector>
#include "koral_hip.h"
#include "grid_ops.hpp"
#include "imp_persist.hpp"
#include "capi_util.hpp"

using namespace kd;

static_assert(IMP_NCTR == IMP_LIST_COUNTERS, "implicit pass counters");
static_assert(sizeof(koral_grid_tune) == sizeof(GridTune) && sizeof(((koral_grid_tune *)0)->block) ==
                  sizeof(((GridTune *)0)->block), "koral_grid_tune mirrors GridTune");

namespace {

constexpr int BS = 256;  // largest block of the grid kernels (launch bounds)

inline dim3 blocks_for(long n) { return dim3((unsigned)((n + BS - 1) / BS)); }
inline dim3 blocks_for(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// Op::waves_per_simd (default 1) asks the compiler to fit that many waves
// per SIMD in the register file
template <class Op, class = void>
struct WavesPerSimd {
  static constexpr int value = 1;
};
template <class Op>
struct WavesPerSimd<Op, decltype((void)Op::waves_per_simd, void())> {
  static constexpr int value = Op::waves_per_simd;
};

// the grid kernels run with 64, 128 or 256 threads per block (the tuned
// launch choice of each operator kind)
template <class Op>
__global__ void __launch_bounds__(BS, WavesPerSimd<Op>::value) k_each(Op op)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < op.size()) op(i);
}

__device__ __forceinline__ void atomic_max_pos(double *addr, double v)
{
  // non-negative doubles order as their bit patterns
  atomicMax((unsigned long long *)addr, (unsigned long long)__double_as_longlong(v));
}

// max over threads: wave shuffle, block via LDS, one atomic per block
template <class Op>
__global__ void __launch_bounds__(BS, WavesPerSimd<Op>::value) k_max(Op op, double *out)
{
  __shared__ double red[BS / 64];
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.;
  if (i < op.size()) v = op(i);
  for (int off = 32; off > 0; off >>= 1) v = mx(v, __shfl_xor(v, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    const int nwav = (int)(blockDim.x >> 6);
    for (int k = 1; k < nwav; k++) m = mx(m, red[k]);
    if (m > 0.) atomic_max_pos(out, m);
  }
}

// three counters: wave sums, one atomic per wave and counter
template <class Op>
__global__ void __launch_bounds__(BS, WavesPerSimd<Op>::value) k_count(Op op, int *out)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  Cnt3 c{{0, 0, 0}};
  if (i < op.size()) c = op(i);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int v = c.v[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&out[k], v);
  }
}

// Optional per-launch timing: an event pair around every launch, summed per
// operator kind when read back.
struct Timing {
  bool on = false;
  unsigned mask = ~0u;  // operator kinds that get events while on
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Rec {
    int kind;
    long units;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  hipEvent_t next()
  {
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[used++];
  }
  ~Timing()
  {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

// cell source of the implicit passes: interior cells in (x, y, z) order of
// the index
template <int MAG>
struct GridCells {
  DGrid G;
  double *u, *p;
  signed char *fl;
  const double *clk;  // device clock {t, dt}
  double fac;         // the solve's time step is dt * fac
  __device__ __forceinline__ double time() const { return clk[0]; }
  __device__ __forceinline__ long cell(long i, int &ix, int &iy) const
  {
    Box b{0, 0, 0, G.nx, G.ny, G.nz};
    int iz;
    b.at(i, ix, iy, iz);
    return cidx(G, ix, iy, iz);
  }
  __device__ __forceinline__ void load(long i, double *uu00, double *pp00, Geom &g, double &dtc) const
  {
    int ix, iy;
    const long c = cell(i, ix, iy);
    geom_cell(G, ix, iy, g);
    load_p<MAG, 1>(p, G.ncell, c, pp00);
    load_p<MAG, 1>(u, G.ncell, c, uu00);
    dtc = clk[1] * fac;
  }
  __device__ __forceinline__ void init(long) const {}
  __device__ __forceinline__ void add_iters(long, int) const {}
  __device__ __forceinline__ void done_ok(long i, const double *uu, const double *pp, int, int, int, int,
                                          int) const
  {
    int ix, iy;
    const long c = cell(i, ix, iy), n = G.ncell;
    store_p<MAG, 1>(p, n, c, pp);
    store_p<MAG, 1>(u, n, c, uu);
    fl[F_RADSOURCETYPE * n + c] = 3;
    fl[F_RADFIXUP * n + c] = 0;
    fl[F_RADIMPFIXUP * n + c] = 0;
  }
  __device__ __forceinline__ void done_fail(long i) const
  {
    int ix, iy;
    const long c = cell(i, ix, iy), n = G.ncell;
    fl[F_RADSOURCETYPE * n + c] = 3;
    fl[F_RADFIXUP * n + c] = -1;
    fl[F_RADIMPFIXUP * n + c] = -1;
  }
};

// Learned Newton starts of the grid's implicit solve (no reference
// counterpart): the networks, the gate threshold, the learned budget, and
// work arrays over the n interior cells: primitives and dt gathered in
// batch layout [NV][n] and [n], features [11][n], outputs [4][n], the
// starts [4][n] and the second try's lists.
struct GridLearned {
  koral_hip_mlp *gas = nullptr, *rad = nullptr, *gate = nullptr;
  float thr = 0.f;
  int maxiter = 0;
  long n = 0;
  double *pb = nullptr, *dtb = nullptr, *gxg = nullptr, *gxr = nullptr;
  float *feat = nullptr, *y = nullptr;
  unsigned *second = nullptr;
};

// interior primitives and the solve's time step in batch layout
__global__ void k_gather_interior(DGrid G, int nv, const double *p, const double *clk, double fac, double *pb,
                                  double *dtb)
{
  const long n = (long)G.nx * G.ny * G.nz;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Box b{0, 0, 0, G.nx, G.ny, G.nz};
  int ix, iy, iz;
  b.at(i, ix, iy, iz);
  const long c = cidx(G, ix, iy, iz);
  for (int iv = 0; iv < nv; iv++) pb[iv * n + i] = p[iv * G.ncell + c];
  dtb[i] = clk[1] * fac;
}

struct HipExec {
  hipStream_t s;
  Timing *tm;
  const GridLearned *lrn = nullptr;  // learned starts of the implicit solve, when set
  const GridTune *tune = nullptr;  // threads per block of each operator kind
  int rc = 0;  // first launch-configuration error
  template <class Op>
  int bs() const
  {
    const int b = tune ? tune->block[Op::kind] : BS;
    return (b == 64 || b == 128 || b == 256) ? b : BS;
  }
  // side streams and their fork/join events (work that may overlap the
  // main stream, e.g. the few faces of a sweep left to the floor pass)
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork[3] = {nullptr, nullptr, nullptr}, ev_join[3] = {nullptr, nullptr, nullptr};
  template <class Op, class L>
  void timed(const Op &op, L launch, hipStream_t st)
  {
    long n = op.size();
    if (n <= 0) return;
    hipEvent_t a = nullptr, b = nullptr;
    if (tm && tm->on && ((tm->mask >> Op::kind) & 1u)) {
      a = tm->next();
      b = tm->next();
    }
    if (a && b) (void)hipEventRecord(a, st);
    launch(n);
    if (a && b) {
      (void)hipEventRecord(b, st);
      tm->recs.push_back(Timing::Rec{Op::kind, n, a, b});
    }
  }
  template <class Op, class L>
  void timed(const Op &op, L launch)
  {
    timed(op, launch, s);
  }
  void note(hipError_t e, const char *what)
  {
    if (e != hipSuccess && !rc) rc = check_hip(e, what);
  }
  // side stream k starts after the work launched so far on the main stream
  void fork(int k)
  {
    note(hipEventRecord(ev_fork[k], s), "fork record");
    note(hipStreamWaitEvent(side[k], ev_fork[k], 0), "fork wait");
  }
  // the main stream continues after the work launched so far on side stream k
  void join(int k)
  {
    note(hipEventRecord(ev_join[k], side[k]), "join record");
    note(hipStreamWaitEvent(s, ev_join[k], 0), "join wait");
  }
  template <class Op>
  void each_on(int k, const Op &op)
  {
    hipStream_t st = side[k];
    const int b = bs<Op>();
    timed(op, [&](long n) { k_each<Op><<<blocks_for(n, b), b, 0, st>>>(op); }, st);
  }
  template <class Op>
  void each(const Op &op)
  {
    const int b = bs<Op>();
    timed(op, [&](long n) { k_each<Op><<<blocks_for(n, b), b, 0, s>>>(op); });
  }
  template <class Op>
  void max(const Op &op, double *out)
  {
    const int b = bs<Op>();
    timed(op, [&](long n) { k_max<Op><<<blocks_for(n, b), b, 0, s>>>(op, out); });
  }
  template <class Op>
  void count(const Op &op, int *out)
  {
    const int b = bs<Op>();
    timed(op, [&](long n) { k_count<Op><<<blocks_for(n, b), b, 0, s>>>(op, out); });
  }
  // the implicit sweep as six per-attempt persistent passes
  ImpBufs imp{};
  template <int MAG>
  void implicit_cells(const OpImplicit<MAG> &op, int *out)
  {
    timed(op, [&](long n) {
      GridCells<MAG> src{op.G, op.u, op.p, op.fl, op.clk, op.fac};
      int r = 0;
      if (lrn && lrn->gas) {
        // the starts from the networks on the old state, then the passes
        // with the learned tries first (as the batch learned solve)
        const GridLearned &L = *lrn;
        const int nv = op.P.nv;
        k_gather_interior<<<blocks_for(n), BS, 0, s>>>(op.G, nv, op.p, op.clk, op.fac, L.pb, L.dtb);
        r = koral_hip_mlp_features(&op.P, L.gas, n, L.pb, L.dtb, L.feat, s);
        if (!r) r = koral_hip_mlp_infer(L.gas, L.feat, n, L.y, s);
        if (!r) r = koral_hip_mlp_guess(L.gas, n, nv, L.pb, L.y, L.gxg, s);
        if (!r && L.rad) r = koral_hip_mlp_features(&op.P, L.rad, n, L.pb, L.dtb, L.feat, s);
        if (!r && L.rad) r = koral_hip_mlp_infer(L.rad, L.feat, n, L.y, s);
        if (!r && L.rad) r = koral_hip_mlp_guess(L.rad, n, nv, L.pb, L.y, L.gxr, s);
        if (!r && L.gate) r = koral_hip_mlp_features(&op.P, L.gate, n, L.pb, L.dtb, L.feat, s);
        if (!r && L.gate) r = koral_hip_mlp_infer(L.gate, L.feat, n, L.y, s);
        if (!r && L.gate) r = koral_hip_mlp_gate(L.gate, n, L.y, L.thr, L.gxg, s);
        ImpBufs b2 = imp;
        b2.second[0] = L.second;
        b2.second[1] = L.second + n;
        if (!r) r = implicit_passes<MAG>(op.P, src, n, b2, out, s, L.gxg, L.rad ? L.gxr : nullptr, L.maxiter);
      } else {
        r = implicit_passes<MAG>(op.P, src, n, imp, out, s);
      }
      if (r && !rc) rc = r;
    });
  }
};

// AoS (KORAL host layout, [cell][nv]) <-> SoA
__global__ void k_aos2soa(long ncell, int nv, const double *aos, double *soa)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * nv) return;
  long c = i / nv;
  int iv = (int)(i % nv);
  soa[iv * ncell + c] = aos[i];
}
__global__ void k_soa2aos(long ncell, int nv, const double *soa, double *aos)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * nv) return;
  long c = i / nv;
  int iv = (int)(i % nv);
  aos[i] = soa[iv * ncell + c];
}
__global__ void k_flags_i2c(long ncell, const int *aos, signed char *soa)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * NFLAGS) return;
  long c = i / NFLAGS;
  int f = (int)(i % NFLAGS);
  soa[f * ncell + c] = (signed char)aos[i];
}
__global__ void k_flags_c2i(long ncell, const signed char *soa, int *aos)
{
  long i = (long)blockIdx.x * BS + threadIdx.x;
  if (i >= ncell * NFLAGS) return;
  long c = i / NFLAGS;
  int f = (int)(i % NFLAGS);
  aos[i] = (int)soa[f * ncell + c];
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side grid object
// ---------------------------------------------------------------------------
struct koral_hip_grid {
  GridBufs b{};
  hipStream_t s = nullptr;
  hipStream_t side[3] = {nullptr, nullptr, nullptr};  // default priority, non-blocking
  hipEvent_t ev_fork[3] = {nullptr, nullptr, nullptr}, ev_join[3] = {nullptr, nullptr, nullptr};
  Timing timing;
  GridLearned lrn;
  std::vector<void *> allocs;
  double *stage = nullptr;  // AoS staging for upload/download
  // one RK2IMEX step plus the clock advance, captured once and replayed by
  // koral_hip_run_rk2imex when graphs are enabled
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  void drop_graph()
  {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  template <class T>
  int alloc(T **ptr, size_t count)
  {
    void *q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return check_hip(e, "hipMalloc");
    allocs.push_back(q);
    *ptr = (T *)q;
    return 0;
  }
  ~koral_hip_grid()
  {
    drop_graph();
    for (int k = 0; k < 3; k++) {
      if (side[k]) (void)hipStreamSynchronize(side[k]);
      if (side[k]) (void)hipStreamDestroy(side[k]);
      if (ev_fork[k]) (void)hipEventDestroy(ev_fork[k]);
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
    }
    for (void *q : allocs) (void)hipFree(q);
  }
  int make_side()
  {
    for (int k = 0; k < 3; k++) {
      int rc = check_hip(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking), "side stream");
      if (!rc) rc = check_hip(hipEventCreateWithFlags(&ev_fork[k], hipEventDisableTiming), "fork event");
      if (!rc) rc = check_hip(hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming), "join event");
      if (rc) return rc;
    }
    return 0;
  }
  HipExec ex()
  {
    HipExec e{s, &timing};
    e.tune = &b.tune;
    e.lrn = &lrn;
    for (int k = 0; k < 3; k++) {
      e.side[k] = side[k];
      e.ev_fork[k] = ev_fork[k];
      e.ev_join[k] = ev_join[k];
    }
    e.imp.list[0][0] = b.imp_list[0][0];
    e.imp.list[0][1] = b.imp_list[0][1];
    e.imp.list[1][0] = b.imp_list[1][0];
    e.imp.list[1][1] = b.imp_list[1][1];
    e.imp.ctr = b.imp_ctr;
    return e;
  }
};

namespace {

// The explicit operator with both magnetic field and radiation faulted the
// device once (illegal address) in an earlier build of these kernels and the
// cause has not been identified; until it is, that combination is refused
// rather than launched.  The same kernels without radiation (GRMHD) or
// without field (radiation hydro) and the per-cell implicit solver are
// enabled.
int refuse_mag_rad(const koral_hip_grid *g, const char *what)
{
  if (g->b.P.mag && g->b.P.rad) {
    char msg[256];
    snprintf(msg, sizeof msg, "%s: field + radiation explicit operator is disabled on the device "
                              "(unresolved device fault)", what);
    return set_error(KORAL_E_STATE, msg);
  }
  return 0;
}

// device clock and scalars, set and advanced by kernels so that a sequence
// of steps needs no host round trip
__global__ void k_clock_set(double *clock, double t, double dt)
{
  clock[0] = t;
  clock[1] = dt;
}
__global__ void k_zero_scalars(double *tsd, int *counters, int ncounters)
{
  if (threadIdx.x == 0) tsd[0] = 0.;
  if ((int)threadIdx.x < ncounters) counters[threadIdx.x] = 0;
}
// after a step: t += dt, dt = 1 / tstepdenmax (ko.c:399), tstepdenmax = 0
__global__ void k_clock_advance(double *clock, double *tsd)
{
  clock[0] = clock[0] + clock[1];
  clock[1] = 1. / tsd[0];
  tsd[0] = 0.;
}

int set_clock(koral_hip_grid *g, double t, double dt)
{
  k_clock_set<<<1, 1, 0, g->s>>>(g->b.clock, t, dt);
  return check_launch("clock");
}

int zero_tsd(koral_hip_grid *g)
{
  k_zero_scalars<<<1, 64, 0, g->s>>>(g->b.tsd, g->b.counters, 0);
  return check_launch("zero tsd");
}

int read_tsd(koral_hip_grid *g, double *tsd)
{
  int rc = check_hip(hipMemcpyAsync(tsd, g->b.tsd, sizeof(double), hipMemcpyDeviceToHost, g->s), "tsd");
  return rc ? rc : check_hip(hipStreamSynchronize(g->s), "sync");
}

}  // namespace

extern "C" {

int koral_hip_grid_create(const koral_params *P, const koral_grid_desc *D, koral_hip_grid **out)
{
  if (check_params(P) || !D || !out) return set_error(KORAL_E_ARG, "grid_create: bad arguments");
  if (D->nx < 1 || D->ny < 1 || D->nz < 1 || D->ng < 2 || D->ng > 4)
    return set_error(KORAL_E_ARG, "grid_create: bad sizes");
  if ((P->int_order == 4 && D->ng < 4) || (P->int_order == 2 && D->ng < 3))
    return set_error(KORAL_E_ARG, "grid_create: NG too small for INT_ORDER");
  if (D->ny == 1 && D->nz > 1) return set_error(KORAL_E_ARG, "grid_create: NY=1 with NZ>1 not supported");
  {
    // the kernels address a variable's cells and faces with 32-bit byte offsets
    const long sx = D->nx + 2L * D->ng, sy = D->ny > 1 ? D->ny + 2L * D->ng : 1, sz = D->nz > 1 ? D->nz + 2L * D->ng : 1;
    if ((sx + 1) * (sy + 1) * (sz + 1) >= (1L << 29))
      return set_error(KORAL_E_ARG, "grid_create: more than 2^29 cells or faces per variable (ghosts included)");
  }
  if (P->polar_nc < 0 || (P->polar_nc > 0 && D->ny > 1 && D->ny < 2 * P->polar_nc + 2))
    return set_error(KORAL_E_ARG, "grid_create: polar-axis correction needs NY >= 2 NCCORRECTPOLAR + 2");
  for (int k = 0; k < 3; k++)
    if (!D->xb[k]) return set_error(KORAL_E_ARG, "grid_create: missing face coordinates");
  if (!D->gc || !D->Gc || !D->kr || !D->gfx || !D->Gfx || !D->gfy || !D->Gfy)
    return set_error(KORAL_E_ARG, "grid_create: missing metric tables");
  auto *g = new koral_hip_grid();
  GridBufs &s = g->b;
  s.tune = default_tune();
  s.P = *P;
  DGrid &G = s.G;
  grid_layout(D, G);
  const int nv = P->nv;
  int rc = 0;
  auto up = [&](const double *h, size_t n, const double **dst) {
    if (rc) return;
    double *d = nullptr;
    rc = g->alloc(&d, n);
    if (!rc) rc = check_hip(hipMemcpy(d, h, n * sizeof(double), hipMemcpyHostToDevice), "upload table");
    *dst = d;
  };
  up(D->xb[0], D->nx + 2 * D->ng + 1, &G.xb0);
  up(D->xb[1], D->ny + 2 * D->ng + 1, &G.xb1);
  up(D->xb[2], D->nz + 2 * D->ng + 1, &G.xb2);
  up(D->gc, 20 * G.npl, &G.gc);
  up(D->Gc, 20 * G.npl, &G.Gc);
  up(D->kr, 64 * G.npl, &G.kr);
  up(D->gfx, 20 * G.nplx, &G.gfx);
  up(D->Gfx, 20 * G.nplx, &G.Gfx);
  up(D->gfy, 20 * G.nply, &G.gfy);
  up(D->Gfy, 20 * G.nply, &G.Gfy);
  const long nc = G.ncell;
  grid_buffers(s, [&](auto *&ptr, long count, const char *) {
    if (!rc) rc = g->alloc(&ptr, (size_t)count);
  });
  if (!rc) rc = g->make_side();
  if (!rc) rc = g->alloc(&g->stage, nv * nc);
  if (!rc) rc = check_hip(hipMemset(s.p, 0, nv * nc * sizeof(double)), "memset");
  if (!rc) rc = check_hip(hipMemset(s.u, 0, nv * nc * sizeof(double)), "memset");
  if (!rc) rc = check_hip(hipMemset(s.fl, 0, NFLAGS * nc), "memset");
  for (int d = 0; d < 3 && !rc; d++)
    if (s.flb[d]) rc = check_hip(hipMemset(s.flb[d], 0, nv * (d == 0 ? G.nfx : d == 1 ? G.nfy : G.nfz) * sizeof(double)), "memset");
  if (!rc) rc = check_hip(hipMemset(s.ws, 0, NWS * nc * sizeof(double)), "memset");
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return 0;
}

int koral_hip_grid_destroy(koral_hip_grid *g)
{
  if (!g) return 0;
  (void)hipStreamSynchronize(g->s);
  delete g;
  return 0;
}

int koral_hip_grid_set_stream(koral_hip_grid *g, void *stream)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  g->s = (hipStream_t)stream;
  return 0;
}

int koral_hip_grid_upload(koral_hip_grid *g, const double *p_aos, const double *u_aos, const int *flags_aos)
{
  if (!g || !p_aos) return set_error(KORAL_E_ARG, "grid_upload: bad arguments");
  const GridBufs &s = g->b;
  const long nc = s.G.ncell;
  const int nv = s.P.nv;
  int rc = check_hip(hipMemcpyAsync(g->stage, p_aos, nv * nc * sizeof(double), hipMemcpyHostToDevice, g->s), "upload p");
  if (rc) return rc;
  k_aos2soa<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, g->stage, s.p);
  if (u_aos) {
    rc = check_hip(hipStreamSynchronize(g->s), "sync");
    if (!rc) rc = check_hip(hipMemcpyAsync(g->stage, u_aos, nv * nc * sizeof(double), hipMemcpyHostToDevice, g->s), "upload u");
    if (rc) return rc;
    k_aos2soa<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, g->stage, s.u);
  }
  if (flags_aos) {
    rc = check_hip(hipStreamSynchronize(g->s), "sync");
    int *tmp = (int *)s.pn;  // nv*nc doubles hold 7*nc ints
    if (!rc) rc = check_hip(hipMemcpyAsync(tmp, flags_aos, NFLAGS * nc * sizeof(int), hipMemcpyHostToDevice, g->s), "upload flags");
    if (rc) return rc;
    k_flags_i2c<<<blocks_for(nc * NFLAGS), BS, 0, g->s>>>(nc, tmp, s.fl);
  }
  rc = check_hip(hipStreamSynchronize(g->s), "upload sync");
  return rc ? rc : check_launch("grid_upload");
}

int koral_hip_grid_download(koral_hip_grid *g, double *p_aos, double *u_aos, int *flags_aos)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  const GridBufs &s = g->b;
  const long nc = s.G.ncell;
  const int nv = s.P.nv;
  int rc = 0;
  if (p_aos) {
    k_soa2aos<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, s.p, g->stage);
    rc = check_hip(hipMemcpyAsync(p_aos, g->stage, nv * nc * sizeof(double), hipMemcpyDeviceToHost, g->s), "download p");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  if (!rc && u_aos) {
    k_soa2aos<<<blocks_for(nc * nv), BS, 0, g->s>>>(nc, nv, s.u, g->stage);
    rc = check_hip(hipMemcpyAsync(u_aos, g->stage, nv * nc * sizeof(double), hipMemcpyDeviceToHost, g->s), "download u");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  if (!rc && flags_aos) {
    int *tmp = (int *)g->stage;
    k_flags_c2i<<<blocks_for(nc * NFLAGS), BS, 0, g->s>>>(nc, s.fl, tmp);
    rc = check_hip(hipMemcpyAsync(flags_aos, tmp, NFLAGS * nc * sizeof(int), hipMemcpyDeviceToHost, g->s), "download flags");
    if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  }
  return rc;
}

int koral_hip_grid_device_state(koral_hip_grid *g, double **p, double **u, signed char **flags)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  if (p) *p = g->b.p;
  if (u) *u = g->b.u;
  if (flags) *flags = g->b.fl;
  return 0;
}

int koral_hip_grid_download_work(koral_hip_grid *g, int which, double *out)
{
  if (!g || !out || which < 0 || which > 3) return set_error(KORAL_E_ARG, "grid_download_work: bad arguments");
  const GridBufs &s = g->b;
  const DGrid &G = s.G;
  const double *src = which == 0 ? s.ws : s.flb[which - 1];
  if (!src) return set_error(KORAL_E_ARG, "grid_download_work: collapsed direction has no fluxes");
  long n = which == 0 ? (long)NWS * G.ncell : (long)s.P.nv * (which == 1 ? G.nfx : which == 2 ? G.nfy : G.nfz);
  int rc = check_hip(hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToHost, g->s), "download work");
  return rc ? rc : check_hip(hipStreamSynchronize(g->s), "sync");
}

int koral_hip_grid_timing(koral_hip_grid *g, int enable, double *ms, long *units, int *launches, int nkinds)
{
  if (!g || nkinds < 0) return set_error(KORAL_E_ARG, "grid_timing: bad arguments");
  Timing &T = g->timing;
  int rc = 0;
  if (ms || units || launches) {
    rc = check_hip(hipStreamSynchronize(g->s), "timing sync");
    for (int k = 0; k < nkinds; k++) {
      if (ms) ms[k] = 0.;
      if (units) units[k] = 0;
      if (launches) launches[k] = 0;
    }
    for (const Timing::Rec &r : T.recs) {
      if (rc || r.kind >= nkinds) continue;
      float e = 0.f;
      rc = check_hip(hipEventElapsedTime(&e, r.a, r.b), "event elapsed");
      if (ms) ms[r.kind] += e;
      if (units) units[r.kind] += r.units;
      if (launches) launches[r.kind] += 1;
    }
  }
  T.recs.clear();
  T.used = 0;
  T.on = enable != 0;
  return rc;
}

int koral_hip_grid_timing_mask(koral_hip_grid *g, unsigned mask)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  g->timing.mask = mask;
  return 0;
}

const char *koral_hip_grid_op_name(int kind)
{
  static const char *names[K_NKINDS] = {"p2u", "bc_faces", "bc_edge", "bc_total", "wavespeeds", "flux",
                                        "emf", "ct_faces", "update", "u2p", "fixup", "fixup_apply",
                                        "implicit", "comb", "polar", "recon", "flux_floor", "u2p_floor", "zero"};
  return (kind >= 0 && kind < K_NKINDS) ? names[kind] : "";
}

int koral_hip_grid_p2u(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  HipExec ex = g->ex();
  grid_p2u(ex, g->b);
  return check_launch("grid_p2u");
}

int koral_hip_set_bc(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  HipExec ex = g->ex();
  set_bc(ex, g->b);
  return check_launch("set_bc");
}

int koral_hip_tstepden(koral_hip_grid *g, double *tstepdenmax)
{
  if (!g || !tstepdenmax) return set_error(KORAL_E_ARG, "tstepden: bad arguments");
  int rc = zero_tsd(g);
  if (rc) return rc;
  HipExec ex = g->ex();
  tstepden(ex, g->b);
  rc = check_launch("tstepden");
  return rc ? rc : read_tsd(g, tstepdenmax);
}

int koral_hip_op_explicit(koral_hip_grid *g, double t, double dt, double *tstepdenmax)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  int rc = refuse_mag_rad(g, "op_explicit");
  if (!rc) rc = zero_tsd(g);
  if (!rc) rc = set_clock(g, t, dt);
  if (rc) return rc;
  HipExec ex = g->ex();
  explicit_op(ex, g->b);
  rc = check_launch("op_explicit");
  if (!rc && tstepdenmax) rc = read_tsd(g, tstepdenmax);
  return rc;
}

int koral_hip_calc_u2p(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  HipExec ex = g->ex();
  u2p(ex, g->b);
  return check_launch("calc_u2p");
}

int koral_hip_do_correct(koral_hip_grid *g)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  HipExec ex = g->ex();
  do_correct(ex, g->b);
  return check_launch("do_correct");
}

int koral_hip_op_implicit(koral_hip_grid *g, double t, double dt)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  int rc = set_clock(g, t, dt);
  if (rc) return rc;
  HipExec ex = g->ex();
  implicit_op(ex, g->b, 1.);
  if (ex.rc) return ex.rc;
  return check_launch("op_implicit");
}

int koral_hip_step_rk2imex(koral_hip_grid *g, double t, double dt, double *tstepdenmax, int *counters)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  int rc = refuse_mag_rad(g, "step_rk2imex");
  if (rc) return rc;
  k_zero_scalars<<<1, 64, 0, g->s>>>(g->b.tsd, g->b.counters, 8);
  rc = set_clock(g, t, dt);
  if (rc) return rc;
  HipExec ex = g->ex();
  step_rk2imex(ex, g->b);
  rc = ex.rc ? ex.rc : check_launch("step_rk2imex");
  if (!rc && tstepdenmax) rc = check_hip(hipMemcpyAsync(tstepdenmax, g->b.tsd, sizeof(double), hipMemcpyDeviceToHost, g->s), "tsd");
  if (!rc && counters) rc = check_hip(hipMemcpyAsync(counters, g->b.counters, 6 * sizeof(int), hipMemcpyDeviceToHost, g->s), "counters");
  if (!rc && (tstepdenmax || counters)) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  return rc;
}

// ---------------------------------------------------------------------------
// restart files (fileop.c:911-982 writer, fileop.c:1009-1108 reader): a text
// header "## nout1 nout2 t PROBLEM NX NY NZ" and a binary body of NX*NY*NZ
// int index triples (ix outer, iz inner) followed by NV doubles of primitives
// per cell in the same order
// ---------------------------------------------------------------------------
int koral_hip_grid_write_restart(koral_hip_grid *g, const char *folder, int nout1, int nout2, double t,
                                 int problem)
{
  if (!g || !folder || nout1 < 0) return set_error(KORAL_E_ARG, "write_restart: bad arguments");
  const DGrid &G = g->b.G;
  const int nv = g->b.P.nv;
  std::vector<double> p((size_t)nv * G.ncell);
  int rc = koral_hip_grid_download(g, p.data(), nullptr, nullptr);
  if (rc) return rc;
  char name[1024], link[1024], target[64];
  snprintf(name, sizeof name, "%s/res%04d.head", folder, nout1);
  FILE *f = fopen(name, "w");
  if (!f) return set_error(KORAL_E_STATE, "write_restart: cannot create the header file");
  int bad = fprintf(f, "## %5d %5d %10.6e %5d %5d %5d %5d\n", nout1, nout2, t, problem, G.nx, G.ny, G.nz) < 0;
  bad |= fclose(f) != 0;
  snprintf(name, sizeof name, "%s/res%04d.dat", folder, nout1);
  f = bad ? nullptr : fopen(name, "wb");
  if (!f) return set_error(KORAL_E_STATE, "write_restart: cannot write the restart files");
  for (int ix = 0; ix < G.nx && !bad; ix++)
    for (int iy = 0; iy < G.ny && !bad; iy++)
      for (int iz = 0; iz < G.nz && !bad; iz++) {
        const int idx[3] = {ix, iy, iz};
        bad |= fwrite(idx, sizeof(int), 3, f) != 3;
      }
  for (int ix = 0; ix < G.nx && !bad; ix++)
    for (int iy = 0; iy < G.ny && !bad; iy++)
      for (int iz = 0; iz < G.nz && !bad; iz++)
        bad |= fwrite(&p[(size_t)nv * cidx(G, ix, iy, iz)], sizeof(double), nv, f) != (size_t)nv;
  bad |= fclose(f) != 0;
  if (bad) return set_error(KORAL_E_STATE, "write_restart: write error");
  // reslast.{head,dat} point at the newest files, as the reference's links do
  const char *ext[2] = {"head", "dat"};
  for (int k = 0; k < 2; k++) {
    snprintf(link, sizeof link, "%s/reslast.%s", folder, ext[k]);
    snprintf(target, sizeof target, "res%04d.%s", nout1, ext[k]);
    (void)unlink(link);
    if (symlink(target, link) != 0) return set_error(KORAL_E_STATE, "write_restart: cannot link reslast");
  }
  return 0;
}

int koral_hip_grid_read_restart(koral_hip_grid *g, const char *folder, int nout1, double *t, int *nfout1,
                                int *nfout2)
{
  if (!g || !folder || !t) return set_error(KORAL_E_ARG, "read_restart: bad arguments");
  const DGrid &G = g->b.G;
  const int nv = g->b.P.nv;
  char name[1024];
  if (nout1 >= 0)
    snprintf(name, sizeof name, "%s/res%04d.head", folder, nout1);
  else
    snprintf(name, sizeof name, "%s/reslast.head", folder);
  FILE *f = fopen(name, "r");
  if (!f) return 1;  // no restart file: start from scratch
  int ip[6];
  double tt = 0.;
  const int nr = fscanf(f, "## %d %d %lf %d %d %d %d\n", &ip[0], &ip[1], &tt, &ip[2], &ip[3], &ip[4], &ip[5]);
  fclose(f);
  if (nr != 7) return set_error(KORAL_E_STATE, "read_restart: malformed header");
  if (ip[3] != G.nx || ip[4] != G.ny || ip[5] != G.nz)
    return set_error(KORAL_E_STATE, "read_restart: grid size of the file differs from the grid");
  if (nout1 >= 0)
    snprintf(name, sizeof name, "%s/res%04d.dat", folder, nout1);
  else
    snprintf(name, sizeof name, "%s/reslast.dat", folder);
  f = fopen(name, "rb");
  if (!f) return set_error(KORAL_E_STATE, "read_restart: cannot open the body file");
  const long ncells = (long)G.nx * G.ny * G.nz;
  std::vector<int> idx((size_t)3 * ncells);
  int bad = fread(idx.data(), sizeof(int), idx.size(), f) != idx.size();
  for (long c = 0; c < ncells && !bad; c++)
    bad |= idx[3 * c] < 0 || idx[3 * c] >= G.nx || idx[3 * c + 1] < 0 || idx[3 * c + 1] >= G.ny ||
           idx[3 * c + 2] < 0 || idx[3 * c + 2] >= G.nz;
  std::vector<double> p((size_t)nv * G.ncell);
  if (!bad) bad = koral_hip_grid_download(g, p.data(), nullptr, nullptr) != 0;  // ghosts kept
  for (long c = 0; c < ncells && !bad; c++) {
    double *dst = &p[(size_t)nv * cidx(G, idx[3 * c], idx[3 * c + 1], idx[3 * c + 2])];
    bad |= fread(dst, sizeof(double), nv, f) != (size_t)nv;
  }
  fclose(f);
  if (bad) return set_error(KORAL_E_STATE, "read_restart: truncated or inconsistent body file");
  int rc = koral_hip_grid_upload(g, p.data(), nullptr, nullptr);
  if (rc) return rc;
  HipExec ex = g->ex();
  grid_p2u(ex, g->b, true);
  rc = check_hip(hipStreamSynchronize(g->s), "read_restart sync");
  if (rc) return rc;
  *t = tt;
  if (nfout1) *nfout1 = ip[0] + 1;
  if (nfout2) *nfout2 = ip[1];
  return check_launch("read_restart");
}

int koral_hip_grid_use_graph(koral_hip_grid *g, int enable)
{
  if (!g) return set_error(KORAL_E_ARG, "null grid");
  if (enable && !g->s) return set_error(KORAL_E_ARG, "grid_use_graph: set a stream first (graphs are captured on it)");
  g->use_graph = enable != 0;
  if (!g->use_graph) g->drop_graph();
  return 0;
}

int koral_hip_run_rk2imex(koral_hip_grid *g, int nsteps, double *t, double *dt, int *counters)
{
  if (!g || !t || !dt || nsteps < 0) return set_error(KORAL_E_ARG, "run_rk2imex: bad arguments");
  int rc = refuse_mag_rad(g, "run_rk2imex");
  if (rc) return rc;
  k_zero_scalars<<<1, 64, 0, g->s>>>(g->b.tsd, g->b.counters, 8);
  rc = set_clock(g, *t, *dt);
  const bool graph = g->use_graph && g->s && !g->timing.on;
  if (!rc && graph && !g->gexec) {
    // capture one step; every value that changes between steps lives in
    // device memory (clock, counters, work lists), so the graph is replayed
    // unchanged
    rc = check_hip(hipStreamBeginCapture(g->s, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!rc) {
      HipExec ex = g->ex();
      step_rk2imex(ex, g->b);
      k_clock_advance<<<1, 1, 0, g->s>>>(g->b.clock, g->b.tsd);
      int rc1 = ex.rc ? ex.rc : check_launch("capture step");
      int rc2 = check_hip(hipStreamEndCapture(g->s, &g->graph), "end capture");
      rc = rc1 ? rc1 : rc2;
    }
    if (!rc) rc = check_hip(hipGraphInstantiate(&g->gexec, g->graph, nullptr, nullptr, 0), "graph instantiate");
    if (rc) g->drop_graph();
  }
  if (graph) {
    for (int k = 0; k < nsteps && !rc; k++) rc = check_hip(hipGraphLaunch(g->gexec, g->s), "graph launch");
  } else {
    HipExec ex = g->ex();
    for (int k = 0; k < nsteps && !rc; k++) {
      step_rk2imex(ex, g->b);
      k_clock_advance<<<1, 1, 0, g->s>>>(g->b.clock, g->b.tsd);
      rc = ex.rc ? ex.rc : check_launch("run_rk2imex");
    }
  }
  double clk[2];
  if (!rc) rc = check_hip(hipMemcpyAsync(clk, g->b.clock, 2 * sizeof(double), hipMemcpyDeviceToHost, g->s), "clock");
  if (!rc && counters) rc = check_hip(hipMemcpyAsync(counters, g->b.counters, 6 * sizeof(int), hipMemcpyDeviceToHost, g->s), "counters");
  if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  if (!rc) {
    *t = clk[0];
    *dt = clk[1];
  }
  return rc;
}

}  // extern "C"

namespace {

void tune_to_c(const GridTune &a, koral_grid_tune *t)
{
  for (int k = 0; k < TUNE_KINDS; k++) t->block[k] = a.block[k];
  for (int d = 0; d < 3; d++) t->fused[d] = a.fused[d];
  t->floor_wave = a.floor_wave;
  t->update_pre = a.update_pre;
  t->comb_nv = a.comb_nv;
  t->u2p_budget = a.u2p_budget;
  t->u2p_group = a.u2p_group;
  t->floor_nw = a.floor_nw;
  t->u2p_nw = a.u2p_nw;
}

int tune_from_c(const koral_grid_tune *t, GridTune &a)
{
  for (int k = 0; k < TUNE_KINDS; k++)
    if (t->block[k] != 64 && t->block[k] != 128 && t->block[k] != 256)
      return set_error(KORAL_E_ARG, "set_tune: block sizes must be 64, 128 or 256");
  if (t->u2p_budget < 0) return set_error(KORAL_E_ARG, "set_tune: u2p_budget must be >= 0");
  if (t->u2p_group != 32 && t->u2p_group != 64) return set_error(KORAL_E_ARG, "set_tune: u2p_group must be 32 or 64");
  if (t->floor_nw < 1 || t->floor_nw > (1L << 20) || t->u2p_nw < 1 || t->u2p_nw > (1L << 20))
    return set_error(KORAL_E_ARG, "set_tune: waves in flight must be in [1, 2^20]");
  for (int k = 0; k < TUNE_KINDS; k++) a.block[k] = t->block[k];
  for (int d = 0; d < 3; d++) a.fused[d] = t->fused[d] ? 1 : 0;
  a.floor_wave = t->floor_wave ? 1 : 0;
  a.update_pre = t->update_pre ? 1 : 0;
  a.comb_nv = t->comb_nv ? 1 : 0;
  a.u2p_budget = t->u2p_budget;
  a.u2p_group = t->u2p_group;
  a.floor_nw = t->floor_nw;
  a.u2p_nw = t->u2p_nw;
  return 0;
}

// state the autotuner's trial steps change, saved once and restored before
// every trial
struct TuneSnap {
  double *p = nullptr, *u = nullptr, *clock = nullptr;
  signed char *fl = nullptr;
  ~TuneSnap()
  {
    if (p) (void)hipFree(p);
    if (u) (void)hipFree(u);
    if (clock) (void)hipFree(clock);
    if (fl) (void)hipFree(fl);
  }
  int save(const GridBufs &b, hipStream_t s)
  {
    const size_t nvc = (size_t)b.P.nv * b.G.ncell;
    int rc = check_hip(hipMalloc(&p, nvc * sizeof(double)), "autotune snapshot");
    if (!rc) rc = check_hip(hipMalloc(&u, nvc * sizeof(double)), "autotune snapshot");
    if (!rc) rc = check_hip(hipMalloc(&clock, 4 * sizeof(double)), "autotune snapshot");
    if (!rc) rc = check_hip(hipMalloc(&fl, (size_t)NFLAGS * b.G.ncell), "autotune snapshot");
    if (!rc) rc = copy(b.p, b.u, b.clock, b.fl, p, u, clock, fl, b, s);
    return rc;
  }
  int restore(const GridBufs &b, hipStream_t s) { return copy(p, u, clock, fl, b.p, b.u, b.clock, b.fl, b, s); }
  static int copy(const double *p0, const double *u0, const double *c0, const signed char *f0, double *p1,
                  double *u1, double *c1, signed char *f1, const GridBufs &b, hipStream_t s)
  {
    const size_t nvc = (size_t)b.P.nv * b.G.ncell;
    int rc = check_hip(hipMemcpyAsync(p1, p0, nvc * sizeof(double), hipMemcpyDeviceToDevice, s), "snapshot copy");
    if (!rc) rc = check_hip(hipMemcpyAsync(u1, u0, nvc * sizeof(double), hipMemcpyDeviceToDevice, s), "snapshot copy");
    if (!rc) rc = check_hip(hipMemcpyAsync(c1, c0, 4 * sizeof(double), hipMemcpyDeviceToDevice, s), "snapshot copy");
    if (!rc)
      rc = check_hip(hipMemcpyAsync(f1, f0, (size_t)NFLAGS * b.G.ncell, hipMemcpyDeviceToDevice, s), "snapshot copy");
    return rc;
  }
};

// ms per step of 1 + steps steps from the snapshot, the last `steps` timed
int tune_trial(koral_hip_grid *g, TuneSnap &snap, int steps, hipEvent_t e0, hipEvent_t e1, float *ms)
{
  int rc = snap.restore(g->b, g->s);
  HipExec ex = g->ex();
  for (int k = 0; k < 1 + steps && !rc; k++) {
    if (k == 1) rc = check_hip(hipEventRecord(e0, g->s), "event");
    k_zero_scalars<<<1, 64, 0, g->s>>>(g->b.tsd, g->b.counters, 8);
    step_rk2imex(ex, g->b);
    k_clock_advance<<<1, 1, 0, g->s>>>(g->b.clock, g->b.tsd);
    rc = ex.rc ? ex.rc : check_launch("autotune step");
  }
  if (!rc) rc = check_hip(hipEventRecord(e1, g->s), "event");
  if (!rc) rc = check_hip(hipEventSynchronize(e1), "event sync");
  if (!rc) rc = check_hip(hipEventElapsedTime(ms, e0, e1), "event time");
  if (!rc) *ms /= (float)(steps > 0 ? steps : 1);
  return rc;
}

}  // namespace

extern "C" {

int koral_hip_grid_set_learned(koral_hip_grid *g, koral_hip_mlp *gas, koral_hip_mlp *rad, koral_hip_mlp *gate,
                               float thr, int maxiter)
{
  if (!g) return set_error(KORAL_E_ARG, "grid_set_learned: null grid");
  GridLearned &L = g->lrn;
  g->drop_graph();  // a captured step holds the other implicit path
  if (!gas) {
    L.gas = L.rad = L.gate = nullptr;
    return 0;
  }
  if (!g->b.P.rad) return set_error(KORAL_E_ARG, "grid_set_learned: the grid has no radiation");
  if (koral_hip_mlp_kind(gas) != 0 || (rad && koral_hip_mlp_kind(rad) != 1) || (gate && koral_hip_mlp_kind(gate) != 3))
    return set_error(KORAL_E_ARG, "grid_set_learned: networks of the wrong kind");
  if (maxiter < 0) return set_error(KORAL_E_ARG, "grid_set_learned: maxiter must be >= 0");
  const DGrid &G = g->b.G;
  const long n = (long)G.nx * G.ny * G.nz;
  if (L.n != n) {
    int rc = g->alloc(&L.pb, (size_t)g->b.P.nv * n);
    if (!rc) rc = g->alloc(&L.dtb, (size_t)n);
    if (!rc) rc = g->alloc(&L.gxg, (size_t)4 * n);
    if (!rc) rc = g->alloc(&L.gxr, (size_t)4 * n);
    if (!rc) rc = g->alloc(&L.feat, (size_t)11 * n);
    if (!rc) rc = g->alloc(&L.y, (size_t)4 * n);
    if (!rc) rc = g->alloc(&L.second, (size_t)2 * n);
    if (rc) return rc;
    L.n = n;
  }
  L.gas = gas;
  L.rad = rad;
  L.gate = gate;
  L.thr = thr;
  L.maxiter = maxiter;
  return 0;
}

int koral_hip_grid_download_guess(koral_hip_grid *g, double *gxg, double *gxr)
{
  if (!g || !gxg) return set_error(KORAL_E_ARG, "grid_download_guess: bad arguments");
  const GridLearned &L = g->lrn;
  if (!L.gas || !L.n) return set_error(KORAL_E_STATE, "grid_download_guess: no learned starts set");
  const size_t bytes = (size_t)4 * L.n * sizeof(double);
  int rc = check_hip(hipMemcpyAsync(gxg, L.gxg, bytes, hipMemcpyDeviceToHost, g->s), "download gas starts");
  if (!rc && gxr && L.rad) rc = check_hip(hipMemcpyAsync(gxr, L.gxr, bytes, hipMemcpyDeviceToHost, g->s), "download rad starts");
  if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "download starts sync");
  return rc;
}

int koral_hip_grid_deferred_counts(koral_hip_grid *g, unsigned counts[4])
{
  if (!g || !counts) return set_error(KORAL_E_ARG, "deferred_counts: bad arguments");
  int rc = check_hip(hipMemcpyAsync(counts, g->b.fcnt, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, g->s),
                     "deferred_counts copy");
  if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "deferred_counts sync");
  return rc;
}

int koral_hip_grid_get_tune(koral_hip_grid *g, koral_grid_tune *t)
{
  if (!g || !t) return set_error(KORAL_E_ARG, "get_tune: bad arguments");
  tune_to_c(g->b.tune, t);
  return 0;
}

int koral_hip_grid_set_tune(koral_hip_grid *g, const koral_grid_tune *t)
{
  if (!g || !t) return set_error(KORAL_E_ARG, "set_tune: bad arguments");
  GridTune a = g->b.tune;
  int rc = tune_from_c(t, a);
  if (rc) return rc;
  g->b.tune = a;
  g->drop_graph();  // a captured step holds the old launch shapes
  return 0;
}

int koral_hip_grid_autotune(koral_hip_grid *g, double t, double dt, int steps, char *report, int report_len)
{
  if (!g || steps < 1 || steps > 100 || !(dt > 0.) || (report && report_len < 1))
    return set_error(KORAL_E_ARG, "autotune: bad arguments");
  int rc = refuse_mag_rad(g, "autotune");
  if (!rc) rc = set_clock(g, t, dt);
  if (rc) return rc;
  std::string rep;
  TuneSnap snap;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = snap.save(g->b, g->s);
  if (!rc) rc = check_hip(hipEventCreate(&e0), "event");
  if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
  const bool timing_was = g->timing.on;
  g->timing.on = false;
  g->drop_graph();
  GridTune &T = g->b.tune;
  float best = 0.f;
  char line[160];
  if (!rc) rc = tune_trial(g, snap, steps, e0, e1, &best);
  if (!rc) {
    snprintf(line, sizeof line, "start %.4f\n", best);
    rep += line;
  }
  // one coordinate: try each value, keep the fastest if it beats the current
  // one by at least 1 %
  auto coord = [&](const char *name, auto get, auto set, std::initializer_list<long> vals) {
    if (rc) return;
    const long cur = get();
    long pick = cur;
    float pick_ms = best;
    for (long v : vals) {
      if (v == cur || rc) continue;
      set(v);
      float ms = 0.f;
      rc = tune_trial(g, snap, steps, e0, e1, &ms);
      snprintf(line, sizeof line, "%s %ld %.4f\n", name, v, ms);
      rep += line;
      if (!rc && ms < pick_ms * 0.99f) {
        pick = v;
        pick_ms = ms;
      }
    }
    set(pick);
    best = pick_ms;
    snprintf(line, sizeof line, "chose %s %ld %.4f\n", name, pick, best);
    rep += line;
  };
  const DGrid &G = g->b.G;
  const int n3[3] = {G.nx, G.ny, G.nz};
  const char *fname[3] = {"fused_x1", "fused_x2", "fused_x3"};
  for (int d = 0; d < 3; d++)
    if (n3[d] > 1)
      coord(fname[d], [&] { return (long)T.fused[d]; }, [&](long v) { T.fused[d] = (int)v; }, {0, 1});
  if (g->b.P.mag) {
    coord("floor_wave", [&] { return (long)T.floor_wave; }, [&](long v) { T.floor_wave = (int)v; }, {1, 0});
    coord("floor_nw", [&] { return T.floor_nw; }, [&](long v) { T.floor_nw = v; }, {1024, 4096, 16384});
  }
  coord("u2p_nw", [&] { return T.u2p_nw; }, [&](long v) { T.u2p_nw = v; }, {1024, 4096, 16384});
  coord("update_pre", [&] { return (long)T.update_pre; }, [&](long v) { T.update_pre = (int)v; }, {0, 1});
  coord("comb_nv", [&] { return (long)T.comb_nv; }, [&](long v) { T.comb_nv = (int)v; }, {1, 0});
  coord("u2p_budget", [&] { return (long)T.u2p_budget; }, [&](long v) { T.u2p_budget = (int)v; }, {32, 0, 16, 64});
  coord("u2p_group", [&] { return (long)T.u2p_group; }, [&](long v) { T.u2p_group = (int)v; }, {32, 64});
  // block sizes of the operator kinds a step launches
  const int kinds[] = {K_BC_FACES, K_WAVESPEEDS, K_RECON, K_FLUX, K_FLUX_FLOOR, K_EMF, K_CT,
                       K_UPDATE, K_U2P, K_U2P_FLOOR, K_FIXUP, K_FIXUP_APPLY, K_COMB, K_POLAR};
  for (int k : kinds) {
    snprintf(line, sizeof line, "block_%s", koral_hip_grid_op_name(k));
    const std::string nm = line;
    coord(nm.c_str(), [&] { return (long)T.block[k]; }, [&](long v) { T.block[k] = (int)v; }, {64, 128, 256});
  }
  if (!rc) rc = snap.restore(g->b, g->s);
  if (!rc) rc = check_hip(hipStreamSynchronize(g->s), "sync");
  g->timing.on = timing_was;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (report) {
    snprintf(report, (size_t)report_len, "%s", rep.c_str());
  }
  return rc;
}

}  // extern "C"
