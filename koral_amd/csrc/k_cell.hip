// Batched per-point kernels and their C-ABI entry points (include/koral_hip.h).
// One lane per point; SoA inputs (variable-major) so each load instruction of
// a wavefront reads 64 consecutive doubles (512 B, fully coalesced).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include "koral_hip.h"
#include <mutex>
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

// cell source of the batched implicit solve: SoA arrays with stride n
template <int MAG, int RAD>
struct BatchCells {
  long n;
  double *uu, *pp;
  const double *g20, *G20, *dt;
  int *flags, *info;
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
  __device__ __forceinline__ void done_ok(long c, const double *u, const double *p, int iters, int attempt, int wp,
                                          int we, int wf, int total) const
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
    info[5 * n + c] = total;
  }
  __device__ __forceinline__ void done_fail(long c, int total) const
  {
    flags[F_RADSOURCETYPE * n + c] = 3;
    flags[F_RADFIXUP * n + c] = -1;
    flags[F_RADIMPFIXUP * n + c] = -1;
    info[0 * n + c] = 0;
    info[1 * n + c] = -1;
    info[2 * n + c] = 0;
    info[3 * n + c] = 0;
    info[4 * n + c] = 0;
    info[5 * n + c] = total;
  }
};

constexpr int IMP_BS = 256;

template <int MAG, int RAD>
__global__ void __launch_bounds__(IMP_BS) k_implicit(koral_params P, BatchCells<MAG, RAD> src,
                                                     unsigned int *queue, double global_time)
{
  implicit_persistent<MAG, RAD>(P, src, src.n, queue, global_time, nullptr);
}

// Work counters of the persistent implicit launch: one allocation per device
// (made on the first call), a ring of slots so calls on different streams
// never share a counter; each call clears its slot with a kernel on its
// stream before the launch.
constexpr int QRING = 1024;
struct QueueRing {
  unsigned int *slots = nullptr;
  unsigned int next = 0;
};
std::mutex qmu;
QueueRing qrings[64];

int queue_slot(unsigned int **out)
{
  int dev = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= 64) return set_error(KORAL_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(qmu);
  QueueRing &r = qrings[dev];
  if (!r.slots) {
    rc = check_hip(hipMalloc(&r.slots, QRING * sizeof(unsigned int)), "hipMalloc queue");
    if (rc) return rc;
  }
  *out = r.slots + (r.next++ % QRING);
  return 0;
}


inline dim3 grid_for(long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

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
  unsigned int *q = nullptr;
  int rc = queue_slot(&q);
  if (rc) return rc;
  k_zero_u32<<<1, 1, 0, s>>>(q);
  unsigned blocks = 0;
  if (P->mag) {
    rc = resident_blocks(k_implicit<1, 1>, IMP_BS, n, &blocks);
    if (rc) return rc;
    BatchCells<1, 1> src{n, uu, pp, g20, G20, dt, flags, info};
    k_implicit<1, 1><<<blocks, IMP_BS, 0, s>>>(*P, src, q, global_time);
  } else {
    rc = resident_blocks(k_implicit<0, 1>, IMP_BS, n, &blocks);
    if (rc) return rc;
    BatchCells<0, 1> src{n, uu, pp, g20, G20, dt, flags, info};
    k_implicit<0, 1><<<blocks, IMP_BS, 0, s>>>(*P, src, q, global_time);
  }
  return check_launch("koral_hip_implicit_batch");
}

}  // extern "C"
