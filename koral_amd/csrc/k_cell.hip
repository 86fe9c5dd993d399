// Batched per-point kernels and their C-ABI entry points (include/koral_hip.h).
// One lane per point; SoA inputs (variable-major) so each load instruction of
// a wavefront reads 64 consecutive doubles (512 B, fully coalesced).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include "koral_hip.h"
#include <mutex>
#include <vector>
#include "imp_persist.hpp"
#include "capi_util.hpp"

using namespace kd;

namespace {

template <int MAG, int RAD>
__device__ __forceinline__ void load_point(const double *g20, const double *G20, long n, long i, Geom &g)
{
  load_geom(g, [&](int c) { return g20[c * n + i]; }, [&](int c) { return G20[c * n + i]; });
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(256) k_p2u(koral_params P, long n, const double *__restrict__ pp,
                                             double *__restrict__ uu, const double *__restrict__ g20,
                                             const double *__restrict__ G20)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  load_point<MAG, RAD>(g20, G20, n, i, g);
  double p[NV], u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) p[v] = pp[v * n + i];
  p2u<MAG, RAD>(P, p, u, g);
#pragma unroll
  for (int v = 0; v < NV; v++) uu[v * n + i] = u[v];
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(256) k_u2p(koral_params P, long n, double *__restrict__ uu,
                                             double *__restrict__ pp, const double *__restrict__ g20,
                                             const double *__restrict__ G20, int *__restrict__ ret,
                                             int *__restrict__ flags)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  load_point<MAG, RAD>(g20, G20, n, i, g);
  double p[NV], u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    p[v] = pp[v * n + i];
    u[v] = uu[v * n + i];
  }
  int fe, fe2, fh, fr;
  int r = calc_primitives<MAG, RAD>(P, u, p, g, fe, fe2, fh, fr);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    pp[v * n + i] = p[v];
    uu[v * n + i] = u[v];
  }
  ret[i] = r;
  flags[F_ENTROPY * n + i] = fe;
  flags[F_ENTROPY2 * n + i] = fe2;
  flags[F_HDFIXUP * n + i] = fh;
  flags[F_RADFIXUP * n + i] = fr;
}

// the W inversion from a learned W start (W0[i] > 0), retried from the
// plain start when that fails; total Newton iterations and whether the
// learned start solved the point
template <int MAG, int RAD>
__global__ void __launch_bounds__(256) k_u2p_solver_from(koral_params P, long n, const double *__restrict__ uu,
                                                         double *__restrict__ pp, const double *__restrict__ g20,
                                                         const double *__restrict__ G20, int etype,
                                                         const double *__restrict__ W0, int *__restrict__ ret,
                                                         int *__restrict__ iters, int *__restrict__ guessed)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  load_point<MAG, RAD>(g20, G20, n, i, g);
  double p[NV], u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    p[v] = pp[v * n + i];
    u[v] = uu[v * n + i];
  }
  int it = 0, it2 = 0, r = -1;
  bool gd = false;
  const double w0 = W0 ? W0[i] : 0.;
  if (w0 > 0.) {
    r = u2p_solver_W<MAG, RAD>(P, u, p, g, etype, w0, &it);
    gd = r >= 0;
  }
  if (!gd) {
    r = u2p_solver_W<MAG, RAD>(P, u, p, g, etype, 0., &it2);
    it += it2;
  }
  ret[i] = r;
  if (iters) iters[i] = it;
  if (guessed) guessed[i] = gd ? 1 : 0;
#pragma unroll
  for (int v = 0; v < NV; v++) pp[v * n + i] = p[v];
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(256) k_u2p_solver(koral_params P, long n, const double *__restrict__ uu,
                                                    double *__restrict__ pp, const double *__restrict__ g20,
                                                    const double *__restrict__ G20, int etype,
                                                    int *__restrict__ ret)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  load_point<MAG, RAD>(g20, G20, n, i, g);
  double p[NV], u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    p[v] = pp[v * n + i];
    u[v] = uu[v * n + i];
  }
  ret[i] = u2p_solver_W<MAG, RAD>(P, u, p, g, etype);
#pragma unroll
  for (int v = 0; v < NV; v++) pp[v * n + i] = p[v];
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(256) k_u2p_rad(koral_params P, long n, const double *__restrict__ uu,
                                                 double *__restrict__ pp, const double *__restrict__ g20,
                                                 const double *__restrict__ G20, int *__restrict__ ret,
                                                 int *__restrict__ corr)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  load_point<MAG, RAD>(g20, G20, n, i, g);
  double p[NV], u[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    p[v] = pp[v * n + i];
    u[v] = uu[v * n + i];
  }
  int c = 0;
  ret[i] = u2p_rad<MAG, RAD>(P, u, p, g, c);
  corr[i] = c;
#pragma unroll
  for (int v = 0; v < NV; v++) pp[v * n + i] = p[v];
}

template <int MAG, int RAD>
__global__ void __launch_bounds__(256) k_Gi(koral_params P, long n, const double *__restrict__ pp,
                                            const double *__restrict__ g20, const double *__restrict__ G20,
                                            int labframe, double *__restrict__ Gi)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Geom g;
  load_point<MAG, RAD>(g20, G20, n, i, g);
  double p[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) p[v] = pp[v * n + i];
  double G4[4];
  calc_Gi<MAG, RAD>(P, p, g, G4, labframe);
#pragma unroll
  for (int k = 0; k < 4; k++) Gi[k * n + i] = G4[k];
}

// cell source of the batched implicit solve: SoA arrays with stride n;
// info[5] accumulates the iterations of all attempts
template <int MAG, int RAD>
struct BatchCells {
  long n;
  double *uu, *pp;
  const double *g20, *G20, *dt;
  int *flags, *info;
  double gt;
  __device__ __forceinline__ double time() const { return gt; }
  __device__ __forceinline__ void load(long c, double *uu00, double *pp00, Geom &g, double &dtc) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    load_point<MAG, RAD>(g20, G20, n, c, g);
#pragma unroll
    for (int v = 0; v < NV; v++) {
      pp00[v] = pp[v * n + c];
      uu00[v] = uu[v * n + c];
    }
    dtc = dt[c];
  }
  __device__ __forceinline__ void init(long c) const { info[5 * n + c] = 0; }
  __device__ __forceinline__ void add_iters(long c, int it) const { info[5 * n + c] += it; }
  __device__ __forceinline__ void done_ok(long c, const double *u, const double *p, int iters, int attempt, int wp,
                                          int we, int wf) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
#pragma unroll
    for (int v = 0; v < NV; v++) {
      pp[v * n + c] = p[v];
      uu[v * n + c] = u[v];
    }
    flags[F_RADSOURCETYPE * n + c] = 3;
    flags[F_RADFIXUP * n + c] = 0;
    flags[F_RADIMPFIXUP * n + c] = 0;
    info[0 * n + c] = iters;
    info[1 * n + c] = attempt;
    info[2 * n + c] = wp;
    info[3 * n + c] = we;
    info[4 * n + c] = wf;
  }
  __device__ __forceinline__ void done_fail(long c) const
  {
    flags[F_RADSOURCETYPE * n + c] = 3;
    flags[F_RADFIXUP * n + c] = -1;
    flags[F_RADIMPFIXUP * n + c] = -1;
    info[0 * n + c] = 0;
    info[1 * n + c] = -1;
    info[2 * n + c] = 0;
    info[3 * n + c] = 0;
    info[4 * n + c] = 0;
  }
};

// Work lists and counters of the implicit passes, per device.  A set is
// reused by a call on the stream of its last call (stream order separates
// them) or, on another stream, once the work of its last call has finished
// (an event recorded after the launches), so overlapping calls never share
// one; the sets grow to the largest batch seen and live until the process
// ends.
struct ImpSet {
  int dev = 0;
  long cap = 0;
  bool busy = false;
  hipStream_t last = nullptr;
  ImpBufs b{};
  void *mem = nullptr;
  void *lmem = nullptr;  // learned-guess buffers (features, outputs, guesses), made on first use
  float *feat = nullptr, *y = nullptr;
  double *gxg = nullptr, *gxr = nullptr;
  hipEvent_t done = nullptr;
};
std::mutex pool_mu;
std::vector<ImpSet *> pool;

int acquire_set(long n, hipStream_t st, ImpSet **out)
{
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(pool_mu);
  for (ImpSet *s : pool)
    if (!s->busy && s->dev == dev && s->cap >= n && (s->last == st || hipEventQuery(s->done) == hipSuccess)) {
      s->busy = true;
      *out = s;
      return 0;
    }
  long cap = n < (1L << 16) ? (1L << 16) : n;
  auto *s = new ImpSet();
  s->dev = dev;
  s->cap = cap;
  size_t bytes = (4 * (size_t)cap + IMP_NCTR) * sizeof(unsigned);
  rc = check_hip(hipMalloc(&s->mem, bytes), "hipMalloc implicit lists");
  if (!rc) rc = check_hip(hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "event");
  if (rc) {
    if (s->mem) (void)hipFree(s->mem);
    delete s;
    return rc;
  }
  unsigned *m = (unsigned *)s->mem;
  s->b.list[0][0] = m;
  s->b.list[0][1] = m + cap;
  s->b.list[1][0] = m + 2 * cap;
  s->b.list[1][1] = m + 3 * cap;
  s->b.ctr = m + 4 * cap;
  s->busy = true;
  pool.push_back(s);
  *out = s;
  return 0;
}

int learned_buffers(ImpSet *s)
{
  if (s->lmem) return 0;
  const size_t cap = (size_t)s->cap;
  int rc = check_hip(
      hipMalloc(&s->lmem, cap * (11 * sizeof(float) + 4 * sizeof(float) + 8 * sizeof(double) + 2 * sizeof(unsigned))),
      "hipMalloc learned buffers");
  if (rc) return rc;
  s->gxg = (double *)s->lmem;
  s->gxr = s->gxg + 4 * cap;
  s->feat = (float *)(s->gxr + 4 * cap);
  s->y = s->feat + 11 * cap;
  // work lists of the second learned try
  s->b.second[0] = (unsigned *)(s->y + 4 * cap);
  s->b.second[1] = s->b.second[0] + cap;
  return 0;
}

int release_set(ImpSet *s, hipStream_t st)
{
  int rc = check_hip(hipEventRecord(s->done, st), "event record");
  std::lock_guard<std::mutex> lk(pool_mu);
  s->last = st;
  s->busy = false;
  return rc;
}

inline dim3 grid_for(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// Work lists and counters of koral_hip_implicit_batch_noalloc: a file-scope
// device array (each GPU holds its own copy from module load until the
// process ends), laid out as an ImpSet of KORAL_HIP_NOALLOC_MAX_CELLS cells.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): it uses the address hipGetSymbolAddress
// returns on each device, kept in scratch_dev.
constexpr long NOALLOC_LISTS = 4L * KORAL_HIP_NOALLOC_MAX_CELLS;
static __device__ __align__(256) unsigned scratch[NOALLOC_LISTS + IMP_NCTR];
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex scratch_mu;
unsigned *scratch_dev[NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on the current device, looked up once per
// device with that device current
int scratch_on_device(unsigned **out)
{
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return set_error(KORAL_E_STATE, "implicit_batch_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void *ptr = nullptr;
    rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    scratch_dev[dev] = (unsigned *)ptr;
  }
  *out = scratch_dev[dev];
  return 0;
}

}  // namespace

#define KORAL_DISPATCH(P, KERNEL, ...)                                          \
  do {                                                                          \
    if ((P)->mag && (P)->rad) KERNEL<1, 1> __VA_ARGS__;                         \
    else if ((P)->mag) KERNEL<1, 0> __VA_ARGS__;                                \
    else if ((P)->rad) KERNEL<0, 1> __VA_ARGS__;                                \
    else KERNEL<0, 0> __VA_ARGS__;                                              \
  } while (0)

extern "C" {

int koral_hip_p2u_batch(const koral_params *P, long n, const double *pp, double *uu, const double *g20,
                        const double *G20, void *stream)
{
  if (check_params(P) || n < 0) return KORAL_E_ARG;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  KORAL_DISPATCH(P, k_p2u, <<<grid_for(n, 256), 256, 0, s>>>(*P, n, pp, uu, g20, G20));
  return check_launch("koral_hip_p2u_batch");
}

int koral_hip_u2p_batch(const koral_params *P, long n, double *uu, double *pp, const double *g20,
                        const double *G20, int *ret, int *flags, void *stream)
{
  if (check_params(P) || n < 0) return KORAL_E_ARG;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  KORAL_DISPATCH(P, k_u2p, <<<grid_for(n, 256), 256, 0, s>>>(*P, n, uu, pp, g20, G20, ret, flags));
  return check_launch("koral_hip_u2p_batch");
}

int koral_hip_u2p_solver_batch(const koral_params *P, long n, const double *uu, double *pp, const double *g20,
                               const double *G20, int etype, int *ret, void *stream)
{
  if (check_params(P) || n < 0 || (etype != 0 && etype != 1)) return set_error(KORAL_E_ARG, "bad etype");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  KORAL_DISPATCH(P, k_u2p_solver, <<<grid_for(n, 256), 256, 0, s>>>(*P, n, uu, pp, g20, G20, etype, ret));
  return check_launch("koral_hip_u2p_solver_batch");
}

int koral_hip_u2p_solver_batch_from(const koral_params *P, long n, const double *uu, double *pp,
                                    const double *g20, const double *G20, int etype, const double *W0, int *ret,
                                    int *iters, int *guessed, void *stream)
{
  if (check_params(P) || n < 0 || !ret || (etype != 0 && etype != 1))
    return set_error(KORAL_E_ARG, "u2p_solver_batch_from: bad arguments");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  KORAL_DISPATCH(P, k_u2p_solver_from,
                 <<<grid_for(n, 256), 256, 0, s>>>(*P, n, uu, pp, g20, G20, etype, W0, ret, iters, guessed));
  return check_launch("koral_hip_u2p_solver_batch_from");
}

int koral_hip_u2p_solver_batch_learned(const koral_params *P, koral_hip_mlp *m, long n, const double *uu,
                                       double *pp, const double *g20, const double *G20, int etype, int *ret,
                                       int *iters, int *guessed, void *stream)
{
  if (check_params(P) || n < 0 || !m || koral_hip_mlp_kind(m) != 2)
    return set_error(KORAL_E_ARG, "u2p_solver_batch_learned: bad arguments");
  if (n == 0) return 0;
  if (n >= (1L << 31)) return set_error(KORAL_E_ARG, "u2p batch larger than 2^31 points");
  hipStream_t s = (hipStream_t)stream;
  ImpSet *set = nullptr;
  int rc = acquire_set(n, s, &set);
  if (rc) return rc;
  rc = learned_buffers(set);
  // the W guess lands in the first n doubles of the gas-guess buffer
  if (!rc) rc = koral_hip_mlp_u2p_features(P, m, n, uu, pp, g20, G20, set->feat, stream);
  if (!rc) rc = koral_hip_mlp_infer(m, set->feat, n, set->y, stream);
  if (!rc) rc = koral_hip_mlp_u2p_guess(P, m, n, pp, g20, G20, set->y, set->gxg, stream);
  if (!rc) rc = koral_hip_u2p_solver_batch_from(P, n, uu, pp, g20, G20, etype, set->gxg, ret, iters, guessed, stream);
  int rc2 = release_set(set, s);
  if (rc) return rc;
  if (rc2) return rc2;
  return check_launch("koral_hip_u2p_solver_batch_learned");
}

int koral_hip_u2p_rad_batch(const koral_params *P, long n, const double *uu, double *pp, const double *g20,
                            const double *G20, int *ret, int *corr, void *stream)
{
  if (check_params(P) || n < 0) return KORAL_E_ARG;
  if (!P->rad) return set_error(KORAL_E_ARG, "u2p_rad needs a radiation layout");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (P->mag)
    k_u2p_rad<1, 1><<<grid_for(n, 256), 256, 0, s>>>(*P, n, uu, pp, g20, G20, ret, corr);
  else
    k_u2p_rad<0, 1><<<grid_for(n, 256), 256, 0, s>>>(*P, n, uu, pp, g20, G20, ret, corr);
  return check_launch("koral_hip_u2p_rad_batch");
}

int koral_hip_Gi_batch(const koral_params *P, long n, const double *pp, const double *g20, const double *G20,
                       int labframe, double *Gi, void *stream)
{
  if (check_params(P) || n < 0) return KORAL_E_ARG;
  if (!P->rad) return set_error(KORAL_E_ARG, "Gi needs a radiation layout");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (P->mag)
    k_Gi<1, 1><<<grid_for(n, 256), 256, 0, s>>>(*P, n, pp, g20, G20, labframe, Gi);
  else
    k_Gi<0, 1><<<grid_for(n, 256), 256, 0, s>>>(*P, n, pp, g20, G20, labframe, Gi);
  return check_launch("koral_hip_Gi_batch");
}

int koral_hip_implicit_batch(const koral_params *P, long n, double *uu, double *pp, const double *g20,
                             const double *G20, const double *dt, double global_time, int *flags, int *info,
                             void *stream)
{
  if (check_params(P) || n < 0) return KORAL_E_ARG;
  if (!P->rad) return set_error(KORAL_E_ARG, "implicit solve needs a radiation layout");
  if (n == 0) return 0;
  if (n >= (1L << 31)) return set_error(KORAL_E_ARG, "implicit batch larger than 2^31 cells");
  hipStream_t s = (hipStream_t)stream;
  ImpSet *set = nullptr;
  int rc = acquire_set(n, s, &set);
  if (rc) return rc;
  if (P->mag) {
    BatchCells<1, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
    rc = implicit_passes<1>(*P, src, n, set->b, nullptr, s);
  } else {
    BatchCells<0, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
    rc = implicit_passes<0>(*P, src, n, set->b, nullptr, s);
  }
  int rc2 = release_set(set, s);
  if (rc) return rc;
  if (rc2) return rc2;
  return check_launch("koral_hip_implicit_batch");
}

int koral_hip_implicit_batch_noalloc(const koral_params *P, long n, double *uu, double *pp, const double *g20,
                                     const double *G20, const double *dt, double global_time, int *flags, int *info,
                                     void *stream)
{
  if (check_params(P) || n < 0) return KORAL_E_ARG;
  if (!P->rad) return set_error(KORAL_E_ARG, "implicit solve needs a radiation layout");
  if (n > KORAL_HIP_NOALLOC_MAX_CELLS)
    return set_error(KORAL_E_ARG, "implicit_batch_noalloc: batch larger than KORAL_HIP_NOALLOC_MAX_CELLS "
                                  "(use koral_hip_implicit_batch)");
  unsigned *ptr = nullptr;
  int rc = scratch_on_device(&ptr);
  if (rc) return rc;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  // the passes clear the counters with a kernel before they use them; list
  // entries are written before they are read
  ImpBufs b{};
  b.list[0][0] = ptr;
  b.list[0][1] = ptr + KORAL_HIP_NOALLOC_MAX_CELLS;
  b.list[1][0] = ptr + 2L * KORAL_HIP_NOALLOC_MAX_CELLS;
  b.list[1][1] = ptr + 3L * KORAL_HIP_NOALLOC_MAX_CELLS;
  b.ctr = ptr + NOALLOC_LISTS;
  if (P->mag) {
    BatchCells<1, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
    rc = implicit_passes<1>(*P, src, n, b, nullptr, s);
  } else {
    BatchCells<0, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
    rc = implicit_passes<0>(*P, src, n, b, nullptr, s);
  }
  if (rc) return rc;
  return check_launch("koral_hip_implicit_batch_noalloc");
}

int koral_hip_implicit_batch_from(const koral_params *P, long n, double *uu, double *pp, const double *g20,
                                  const double *G20, const double *dt, double global_time, const double *gxg,
                                  const double *gxr, int maxiter, int *flags, int *info, void *stream)
{
  if (check_params(P) || n < 0 || !gxg) return set_error(KORAL_E_ARG, "implicit_batch_from: bad arguments");
  if (!P->rad) return set_error(KORAL_E_ARG, "implicit solve needs a radiation layout");
  if (n == 0) return 0;
  if (n >= (1L << 31)) return set_error(KORAL_E_ARG, "implicit batch larger than 2^31 cells");
  hipStream_t s = (hipStream_t)stream;
  ImpSet *set = nullptr;
  int rc = acquire_set(n, s, &set);
  if (rc) return rc;
  // the second try on the other primitives needs the lists made with the learned buffers
  if (gxr) rc = learned_buffers(set);
  if (rc) {
    (void)release_set(set, s);
    return rc;
  }
  if (P->mag) {
    BatchCells<1, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
    rc = implicit_passes<1>(*P, src, n, set->b, nullptr, s, gxg, gxr, maxiter);
  } else {
    BatchCells<0, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
    rc = implicit_passes<0>(*P, src, n, set->b, nullptr, s, gxg, gxr, maxiter);
  }
  int rc2 = release_set(set, s);
  if (rc) return rc;
  if (rc2) return rc2;
  return check_launch("koral_hip_implicit_batch_from");
}

namespace {

// the learned solve with an optional gate network
int learned_solve(const koral_params *P, koral_hip_mlp *gas, koral_hip_mlp *rad, koral_hip_mlp *gate, float thr,
                  long n, double *uu, double *pp, const double *g20, const double *G20, const double *dt,
                  double global_time, int maxiter, int *flags, int *info, void *stream, const char *what)
{
  if (check_params(P) || n < 0 || !gas) return set_error(KORAL_E_ARG, "implicit_batch_learned: bad arguments");
  if (koral_hip_mlp_kind(gas) != 0 || (rad && koral_hip_mlp_kind(rad) != 1) || (gate && koral_hip_mlp_kind(gate) != 3))
    return set_error(KORAL_E_ARG, "implicit_batch_learned: networks of the wrong kind");
  if (!P->rad) return set_error(KORAL_E_ARG, "implicit solve needs a radiation layout");
  if (n == 0) return 0;
  if (n >= (1L << 31)) return set_error(KORAL_E_ARG, "implicit batch larger than 2^31 cells");
  hipStream_t s = (hipStream_t)stream;
  ImpSet *set = nullptr;
  int rc = acquire_set(n, s, &set);
  if (rc) return rc;
  rc = learned_buffers(set);
  if (!rc) rc = koral_hip_mlp_features(P, gas, n, pp, dt, set->feat, stream);
  if (!rc) rc = koral_hip_mlp_infer(gas, set->feat, n, set->y, stream);
  if (!rc) rc = koral_hip_mlp_guess(gas, n, P->nv, pp, set->y, set->gxg, stream);
  if (!rc && rad) rc = koral_hip_mlp_features(P, rad, n, pp, dt, set->feat, stream);
  if (!rc && rad) rc = koral_hip_mlp_infer(rad, set->feat, n, set->y, stream);
  if (!rc && rad) rc = koral_hip_mlp_guess(rad, n, P->nv, pp, set->y, set->gxr, stream);
  if (!rc && gate) rc = koral_hip_mlp_features(P, gate, n, pp, dt, set->feat, stream);
  if (!rc && gate) rc = koral_hip_mlp_infer(gate, set->feat, n, set->y, stream);
  if (!rc && gate) rc = koral_hip_mlp_gate(gate, n, set->y, thr, set->gxg, stream);
  const double *gxr = rad ? set->gxr : nullptr;
  if (!rc) {
    if (P->mag) {
      BatchCells<1, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
      rc = implicit_passes<1>(*P, src, n, set->b, nullptr, s, set->gxg, gxr, maxiter);
    } else {
      BatchCells<0, 1> src{n, uu, pp, g20, G20, dt, flags, info, global_time};
      rc = implicit_passes<0>(*P, src, n, set->b, nullptr, s, set->gxg, gxr, maxiter);
    }
  }
  int rc2 = release_set(set, s);
  if (rc) return rc;
  if (rc2) return rc2;
  return check_launch(what);
}

}  // namespace

int koral_hip_implicit_batch_learned(const koral_params *P, koral_hip_mlp *gas, koral_hip_mlp *rad, long n,
                                     double *uu, double *pp, const double *g20, const double *G20,
                                     const double *dt, double global_time, int maxiter, int *flags, int *info,
                                     void *stream)
{
  return learned_solve(P, gas, rad, nullptr, 0.f, n, uu, pp, g20, G20, dt, global_time, maxiter, flags, info, stream,
                       "koral_hip_implicit_batch_learned");
}

int koral_hip_implicit_batch_learned_gated(const koral_params *P, koral_hip_mlp *gas, koral_hip_mlp *rad,
                                           koral_hip_mlp *gate, float thr, long n, double *uu, double *pp,
                                           const double *g20, const double *G20, const double *dt,
                                           double global_time, int maxiter, int *flags, int *info,
                                           void *stream)
{
  if (!gate) return set_error(KORAL_E_ARG, "implicit_batch_learned_gated: no gate network");
  return learned_solve(P, gas, rad, gate, thr, n, uu, pp, g20, G20, dt, global_time, maxiter, flags, info, stream,
                       "koral_hip_implicit_batch_learned_gated");
}

}  // extern "C"
