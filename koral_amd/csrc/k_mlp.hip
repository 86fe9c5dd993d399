// Learned initial guesses for the implicit radiation solve and the u2p W
// inversion: a small MLP
// (11 -> 128 -> 128 -> 128 -> 4, GELU in tanh form) that maps the
// dimensionless local state of a cell to corrections of the plain Newton
// start.  A gas network corrects (u, u~^1..3), a radiation network
// (E, F~^1..3); the solve picks one per cell.  A W network (kind 2)
// predicts the start of the W Newton iteration of u2p.  Training is FP32 (forward,
// backward and Adam kernels below, deterministic reductions); inference runs
// the weights in BF16 on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32
// accumulation).
//
// Parameters: one FP32 vector W1[128][11], b1, W2[128][128], b2, W3[128][128],
// b3, W4[4][128], b4 (row-major W[out][in]).  Normalisation: feature mean and
// std (11 + 11), target mean and std (4 + 4).
//
// Features of a cell: log10 rho, log10(u/rho), log10(E/u), gas u~^i (3),
// radiation-frame u~^i (3), log10(dt kappa_abs), log10(dt kappa_es).
// Targets between the converged energy-equation lab-frame state (x*) and the
// initial state (x0) of the network's four primitives: log(x*_0 / x0_0) and
// asinh((x*_i - x0_i) / VSCALE) for the three velocity components.
#include <hip/hip_runtime.h>
#include <math.h>
#include <mutex>
#include <vector>
#include "koral_hip.h"
#include "dev_cell.hpp"
#include "capi_util.hpp"

using namespace kd;

namespace {

constexpr int NIN = 11, NH = 128, NOUT = 4;
constexpr int OW1 = 0, OB1 = OW1 + NH * NIN, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NOUT * NH, NPARAM = OB4 + NOUT;
constexpr int NNORM = 2 * NIN + 2 * NOUT;  // feature mean, std, target mean, std
constexpr int BMAX = 4096;                 // largest training batch
constexpr float C_GELU = 0.7978845608028654f;
// velocity-like targets are asinh(delta / VSCALE): linear below VSCALE,
// logarithmic above, so changes from 1e-10 to 1 carry comparable weight
constexpr double VSCALE = 1e-6;

__device__ __forceinline__ float gelu(float z)
{
  return 0.5f * z * (1.f + tanhf(C_GELU * (z + 0.044715f * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z)
{
  float t = tanhf(C_GELU * (z + 0.044715f * z * z * z));
  return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * C_GELU * (1.f + 3.f * 0.044715f * z * z);
}

// ---------------------------------------------------------------------------
// features and targets
// ---------------------------------------------------------------------------
template <int MAG>
__global__ void __launch_bounds__(256) k_features(koral_params P, long n, const double *__restrict__ pp,
                                                  const double *__restrict__ dt, const float *__restrict__ norm,
                                                  float *__restrict__ feat)
{
  constexpr int EE = Lay<MAG, 1>::EE;
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const double rho = pp[0 * n + c], u = pp[1 * n + c], E = pp[EE * n + c], d = dt[c];
  double f[NIN];
  f[0] = log10(rho);
  f[1] = log10(u / rho);
  f[2] = log10(E / u);
  f[3] = pp[2 * n + c];
  f[4] = pp[3 * n + c];
  f[5] = pp[4 * n + c];
  f[6] = pp[(EE + 1) * n + c];
  f[7] = pp[(EE + 2) * n + c];
  f[8] = pp[(EE + 3) * n + c];
  f[9] = log10(d * kappa_abs(P, rho, u));
  f[10] = log10(d * kappa_es(P, rho));
#pragma unroll
  for (int k = 0; k < NIN; k++) {
    float v = (float)f[k];
    if (norm) v = (v - norm[k]) / norm[NIN + k];
    feat[k * n + c] = v;
  }
}

__global__ void __launch_bounds__(256) k_targets(long n, int base, const double *__restrict__ p0,
                                                 const double *__restrict__ p1, float *__restrict__ t)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  t[c] = (float)log(p1[base * n + c] / p0[base * n + c]);
#pragma unroll
  for (int i = 1; i < NOUT; i++)
    t[i * n + c] = (float)asinh((p1[(base + i) * n + c] - p0[(base + i) * n + c]) / VSCALE);
}

// The W network (kind 2): features of a point of the u2p W inversion from
// its conserved state uu and the guess pp (the primitives the inversion
// starts from), and the target log(W1 / W0) between the solution pp1 and
// the guess pp0.  W = (rho + Gamma u) gamma^2, the start u2p.c:2094-2112
// computes from a primitive state.
template <int MAG>
__device__ __forceinline__ double plain_W(const koral_params &P, const double *p, const Geom &g)
{
  return (p[0] + P.gamma * p[1]) * (1. + qsq_of(p[2], p[3], p[4], g));
}

template <int MAG>
__device__ __forceinline__ void load_mhd(const double *__restrict__ a, long n, long c, double *x)
{
  constexpr int NVM = 6 + 3 * MAG;
#pragma unroll
  for (int k = 0; k < 9; k++) x[k] = k < NVM ? a[k * n + c] : 0.;
}

template <int MAG>
__global__ void __launch_bounds__(256) k_u2p_features(koral_params P, long n, const double *__restrict__ uu,
                                                      const double *__restrict__ pp, const double *__restrict__ g20,
                                                      const double *__restrict__ G20, const float *__restrict__ norm,
                                                      float *__restrict__ feat)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Geom g;
  load_geom(g, [&](int k) { return g20[k * n + c]; }, [&](int k) { return G20[k * n + c]; });
  double p[9], U[9], U0[9];
  load_mhd<MAG>(pp, n, c, p);
  load_mhd<MAG>(uu, n, c, U);
  p2u_mhd<MAG, 0>(P, p, U0, g);
  const double rho = p[0], u = p[1];
  double bsq = 0.;
  if (MAG) {
    double ucon[4], ucov[4], bcon[4], bcov[4];
    relr_to_4vel(p[2], p[3], p[4], ucon, ucov, g);
    bcon4(p[6], p[7], p[8], ucon, ucov, bcon);
    lower(bcon, bcov, g);
    bsq = dot4(bcon, bcov);
  }
  double f[NIN];
  f[0] = log10(u / rho);
  f[1] = log10(bsq / rho + 1e-30);
#pragma unroll
  for (int i = 0; i < 3; i++) f[2 + i] = p[2 + i] * sqrt(g.g[1 + i][1 + i]);
  f[5] = log(U[0] / U0[0]);
  f[6] = asinh((U[1] - U0[1]) / fabs(U0[1]));
  const double s2 = fabs(U0[0]) + fabs(U0[1]);
#pragma unroll
  for (int i = 0; i < 3; i++) f[7 + i] = asinh((U[2 + i] - U0[2 + i]) / s2);
  f[10] = log10(plain_W<MAG>(P, p, g) / rho);
#pragma unroll
  for (int k = 0; k < NIN; k++) {
    float v = (float)f[k];
    if (norm) v = (v - norm[k]) / norm[NIN + k];
    feat[k * n + c] = v;
  }
}

template <int MAG>
__global__ void __launch_bounds__(256) k_u2p_targets(koral_params P, long n, const double *__restrict__ p0,
                                                     const double *__restrict__ p1, const double *__restrict__ g20,
                                                     const double *__restrict__ G20, float *__restrict__ t)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Geom g;
  load_geom(g, [&](int k) { return g20[k * n + c]; }, [&](int k) { return G20[k * n + c]; });
  double a[9], b[9];
  load_mhd<MAG>(p0, n, c, a);
  load_mhd<MAG>(p1, n, c, b);
  t[c] = (float)log(plain_W<MAG>(P, b, g) / plain_W<MAG>(P, a, g));
#pragma unroll
  for (int i = 1; i < NOUT; i++) t[i * n + c] = 0.f;
}

template <int MAG>
__global__ void __launch_bounds__(256) k_u2p_guess(koral_params P, long n, const double *__restrict__ pp,
                                                   const double *__restrict__ g20, const double *__restrict__ G20,
                                                   const float *__restrict__ y, const float *__restrict__ norm,
                                                   double *__restrict__ W0)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Geom g;
  load_geom(g, [&](int k) { return g20[k * n + c]; }, [&](int k) { return G20[k * n + c]; });
  double p[9];
  load_mhd<MAG>(pp, n, c, p);
  const double t = (double)y[c] * norm[2 * NIN + NOUT] + norm[2 * NIN];
  const double w = plain_W<MAG>(P, p, g) * exp(t);
  W0[c] = (isfinite(w) && w > 0.) ? w : 0.;
}

// ---------------------------------------------------------------------------
// FP32 training kernels; activations are [feature][sample] with stride B
// ---------------------------------------------------------------------------
// Z[o][b] = bias[o] + sum_i W[o][i] X[i][b];  A = gelu(Z) when act
__global__ void __launch_bounds__(256) k_dense_fwd(const float *__restrict__ W, const float *__restrict__ bias,
                                                   const float *__restrict__ X, int I, int O, int B,
                                                   float *__restrict__ Z, float *__restrict__ A, int act)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x, o = blockIdx.y;
  if (b >= B) return;
  float z = bias[o];
  for (int i = 0; i < I; i++) z += W[o * I + i] * X[i * B + b];
  Z[o * B + b] = z;
  if (A) A[o * B + b] = act ? gelu(z) : z;
}

// loss = sum (Y - T)^2 / (4 B), dY = 2 (Y - T) / (4 B); one block
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ Y, const float *__restrict__ T, int B,
                                             float *__restrict__ dY, float *__restrict__ loss)
{
  __shared__ float red[256];
  float s = 0.f;
  const float scale = 1.f / (float)(NOUT * B);
  for (int k = threadIdx.x; k < NOUT * B; k += blockDim.x) {
    float r = Y[k] - T[k];
    s += r * r;
    dY[k] = 2.f * r * scale;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] * scale;
}

// dZ = dA * gelu'(Z)
__global__ void __launch_bounds__(256) k_gelu_bwd(const float *__restrict__ dA, const float *__restrict__ Z, long n,
                                                  float *__restrict__ dZ)
{
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dZ[k] = dA[k] * gelu_grad(Z[k]);
}

// dX[i][b] = sum_o W[o][i] dZ[o][b]
__global__ void __launch_bounds__(256) k_dense_bwd_x(const float *__restrict__ W, const float *__restrict__ dZ,
                                                     int I, int O, int B, float *__restrict__ dX)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (b >= B) return;
  float s = 0.f;
  for (int o = 0; o < O; o++) s += W[o * I + i] * dZ[o * B + b];
  dX[i * B + b] = s;
}

// gW[o][i] = sum_b dZ[o][b] X[i][b] for i < I, gb[o] = sum_b dZ[o][b]
// (column i = I); block (o, 16 columns), fixed-order reduction
__global__ void __launch_bounds__(256) k_dense_bwd_w(const float *__restrict__ dZ, const float *__restrict__ X,
                                                     int I, int B, float *__restrict__ gW, float *__restrict__ gb)
{
  __shared__ float part[16][257];
  const int o = blockIdx.x, i0 = blockIdx.y * 16, t = threadIdx.x;
  float acc[16];
#pragma unroll
  for (int j = 0; j < 16; j++) acc[j] = 0.f;
  for (int b = t; b < B; b += 256) {
    const float d = dZ[o * B + b];
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int i = i0 + j;
      const float x = (i < I) ? X[i * B + b] : (i == I ? 1.f : 0.f);
      acc[j] += d * x;
    }
  }
#pragma unroll
  for (int j = 0; j < 16; j++) part[j][t] = acc[j];
  __syncthreads();
  if (t < 16) {
    float s = 0.f;
    for (int k = 0; k < 256; k++) s += part[t][k];
    const int i = i0 + t;
    if (i < I) gW[o * I + i] = s;
    else if (i == I) gb[o] = s;
  }
}

__global__ void __launch_bounds__(256) k_adam(float *__restrict__ th, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int n, float lr,
                                              float b1, float b2, float c1, float c2, float eps)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float gk = g[k];
  float mk = b1 * m[k] + (1.f - b1) * gk;
  float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  th[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// ---------------------------------------------------------------------------
// BF16 matrix-core inference
// ---------------------------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float x)
{
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}

// weights in BF16 as the matrix cores read them: W1 padded to 16 inputs,
// W4 padded to 16 outputs (zero rows/columns)
constexpr int PW1 = 0, PW2 = PW1 + NH * 16, PW3 = PW2 + NH * NH, PW4 = PW3 + NH * NH, NPACK = PW4 + 16 * NH;

__global__ void __launch_bounds__(256) k_pack_bf16(const float *__restrict__ th, unsigned short *__restrict__ wp)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NPACK) return;
  float v = 0.f;
  if (k < PW2) {
    int o = (k - PW1) / 16, i = (k - PW1) % 16;
    v = i < NIN ? th[OW1 + o * NIN + i] : 0.f;
  } else if (k < PW3) {
    v = th[OW2 + (k - PW2)];
  } else if (k < PW4) {
    v = th[OW3 + (k - PW3)];
  } else {
    int o = (k - PW4) / NH, i = (k - PW4) % NH;
    v = o < NOUT ? th[OW4 + o * NH + i] : 0.f;
  }
  wp[k] = f2bf(v);
}

// One wave per 16 cells.  For D = A B on the 16x16x16 BF16 matrix core,
// lane l holds A[l%16][4(l/16)..+3] and B[4(l/16)..+3][l%16] and receives
// D[4(l/16)..+3][l%16].  A is the activation tile [cell][k] and B the
// transposed weights, B[k][o] = W[o][k], so both are 4 consecutive k.
__global__ void __launch_bounds__(256) k_mlp_infer(const unsigned short *__restrict__ wp,
                                                   const float *__restrict__ th, const float *__restrict__ feat,
                                                   long n, float *__restrict__ out)
{
  __shared__ unsigned short sw[NPACK];
  __shared__ float sb[3 * NH + 16];
  __shared__ unsigned short sh[4][2][16][NH + 4];  // per wave: two activation tiles, padded rows
  for (int k = threadIdx.x; k < NPACK; k += blockDim.x) sw[k] = wp[k];
  for (int k = threadIdx.x; k < 3 * NH + 16; k += blockDim.x) {
    float v = 0.f;
    if (k < NH) v = th[OB1 + k];
    else if (k < 2 * NH) v = th[OB2 + k - NH];
    else if (k < 3 * NH) v = th[OB3 + k - 2 * NH];
    else if (k - 3 * NH < NOUT) v = th[OB4 + k - 3 * NH];
    sb[k] = v;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, kq = 4 * (lane >> 4);
  for (long base = (long)blockIdx.x * 64; base < n; base += (long)gridDim.x * 64) {
    const long c0 = base + 16 * wave;
    // layer 1: A from the features (k < 11), zero padded
    bf16x4 a;
    {
      const long c = c0 + row;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int k = kq + q;
        const float v = (c < n && k < NIN) ? feat[k * n + c] : 0.f;
        a[q] = (short)f2bf(v);
      }
    }
    unsigned short(*h0)[NH + 4] = sh[wave][0];
    unsigned short(*h1)[NH + 4] = sh[wave][1];
#pragma unroll
    for (int ot = 0; ot < NH / 16; ot++) {
      const int o = 16 * ot + row;
      bf16x4 b = *(const bf16x4 *)&sw[PW1 + o * 16 + kq];
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; r++) h0[kq + r][o] = f2bf(gelu(acc[r] + sb[o]));
    }
    __syncthreads();
    // layers 2 and 3: 128 -> 128
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
      const int pw = layer == 0 ? PW2 : PW3;
      unsigned short(*hin)[NH + 4] = layer == 0 ? h0 : h1;
      unsigned short(*hout)[NH + 4] = layer == 0 ? h1 : h0;
      const float *bias = sb + (layer + 1) * NH;
#pragma unroll
      for (int ot = 0; ot < NH / 16; ot++) {
        const int o = 16 * ot + row;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < NH / 16; kc++) {
          bf16x4 av = *(const bf16x4 *)&hin[row][16 * kc + kq];
          bf16x4 bv = *(const bf16x4 *)&sw[pw + o * NH + 16 * kc + kq];
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) hout[kq + r][o] = f2bf(gelu(acc[r] + bias[o]));
      }
      __syncthreads();
    }
    // output layer: 128 -> 4 (padded to 16)
    {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kc = 0; kc < NH / 16; kc++) {
        bf16x4 av = *(const bf16x4 *)&h0[row][16 * kc + kq];
        bf16x4 bv = *(const bf16x4 *)&sw[PW4 + row * NH + 16 * kc + kq];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
      }
      const int o = row;
      if (o < NOUT) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const long c = c0 + kq + r;
          if (c < n) out[o * n + c] = acc[r] + sb[3 * NH + o];
        }
      }
    }
    __syncthreads();
  }
}

// guess (u, u~^i) = plain start corrected by the de-normalised outputs;
// non-finite or non-positive guesses fall back to the plain start
__global__ void __launch_bounds__(256) k_guess(long n, int base, const double *__restrict__ pp,
                                               const float *__restrict__ y, const float *__restrict__ norm,
                                               double *__restrict__ gx)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double t[NOUT];
#pragma unroll
  for (int j = 0; j < NOUT; j++) t[j] = (double)y[j * n + c] * norm[2 * NIN + NOUT + j] + norm[2 * NIN + j];
  const double x0 = pp[base * n + c], x1 = pp[(base + 1) * n + c], x2 = pp[(base + 2) * n + c],
               x3 = pp[(base + 3) * n + c];
  const double g0 = x0 * exp(t[0]);
  const double g1 = x1 + VSCALE * sinh(t[1]), g2 = x2 + VSCALE * sinh(t[2]), g3 = x3 + VSCALE * sinh(t[3]);
  const bool ok = isfinite(g0) && g0 > 0. && isfinite(g1) && isfinite(g2) && isfinite(g3);
  gx[0 * n + c] = ok ? g0 : x0;
  gx[1 * n + c] = ok ? g1 : x1;
  gx[2 * n + c] = ok ? g2 : x2;
  gx[3 * n + c] = ok ? g3 : x3;
}

// the gate network (kind 3) scores whether a learned try will solve a cell;
// a cell scored below thr gets no learned try: its gas start becomes NaN,
// which the learned solve reads as "start the chain at once"
__global__ void __launch_bounds__(256) k_gate(long n, const float *__restrict__ y, const float *__restrict__ norm,
                                              float thr, double *__restrict__ gxg)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const float score = y[c] * norm[2 * NIN + NOUT] + norm[2 * NIN];
  if (!(score >= thr)) gxg[c] = __builtin_nan("");
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

// ---------------------------------------------------------------------------
// the network object
// ---------------------------------------------------------------------------
struct koral_hip_mlp {
  int kind = 0;  // 0: gas primitives (u, u~^i), 1: radiation primitives (E, F~^i)
  float *th = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *norm = nullptr;
  unsigned short *wp = nullptr;
  bool packed = false;
  int step = 0;
  float *Z[3] = {}, *A[3] = {}, *Y = nullptr, *dY = nullptr, *d0 = nullptr, *d1 = nullptr, *loss = nullptr;
  std::vector<void *> allocs;
  template <class T>
  int alloc(T **p, size_t count)
  {
    void *q = nullptr;
    int rc = check_hip(hipMalloc(&q, count * sizeof(T)), "hipMalloc mlp");
    if (rc) return rc;
    allocs.push_back(q);
    *p = (T *)q;
    return 0;
  }
  ~koral_hip_mlp()
  {
    for (void *q : allocs) (void)hipFree(q);
  }
};

namespace {

int pack(koral_hip_mlp *m, hipStream_t s)
{
  if (!m->packed) {
    k_pack_bf16<<<grid1(NPACK), 256, 0, s>>>(m->th, m->wp);
    m->packed = true;
  }
  return check_launch("mlp pack");
}

// forward pass of a batch X [11][B] into m->Z/A and m->Y [4][B]
void forward(koral_hip_mlp *m, const float *X, int B, hipStream_t s)
{
  dim3 gb((B + 255) / 256, NH);
  k_dense_fwd<<<gb, 256, 0, s>>>(m->th + OW1, m->th + OB1, X, NIN, NH, B, m->Z[0], m->A[0], 1);
  k_dense_fwd<<<gb, 256, 0, s>>>(m->th + OW2, m->th + OB2, m->A[0], NH, NH, B, m->Z[1], m->A[1], 1);
  k_dense_fwd<<<gb, 256, 0, s>>>(m->th + OW3, m->th + OB3, m->A[1], NH, NH, B, m->Z[2], m->A[2], 1);
  dim3 go((B + 255) / 256, NOUT);
  k_dense_fwd<<<go, 256, 0, s>>>(m->th + OW4, m->th + OB4, m->A[2], NH, NOUT, B, m->Y, nullptr, 0);
}

// loss and gradient of the mean squared error into m->g
void backward(koral_hip_mlp *m, const float *X, const float *T, int B, hipStream_t s)
{
  k_mse<<<1, 256, 0, s>>>(m->Y, T, B, m->dY, m->loss);
  const long nh = (long)NH * B;
  // output layer
  k_dense_bwd_w<<<dim3(NOUT, (NH + 1 + 15) / 16), 256, 0, s>>>(m->dY, m->A[2], NH, B, m->g + OW4, m->g + OB4);
  k_dense_bwd_x<<<dim3((B + 255) / 256, NH), 256, 0, s>>>(m->th + OW4, m->dY, NH, NOUT, B, m->d0);
  k_gelu_bwd<<<grid1(nh), 256, 0, s>>>(m->d0, m->Z[2], nh, m->d1);
  // layer 3
  k_dense_bwd_w<<<dim3(NH, (NH + 1 + 15) / 16), 256, 0, s>>>(m->d1, m->A[1], NH, B, m->g + OW3, m->g + OB3);
  k_dense_bwd_x<<<dim3((B + 255) / 256, NH), 256, 0, s>>>(m->th + OW3, m->d1, NH, NH, B, m->d0);
  k_gelu_bwd<<<grid1(nh), 256, 0, s>>>(m->d0, m->Z[1], nh, m->d1);
  // layer 2
  k_dense_bwd_w<<<dim3(NH, (NH + 1 + 15) / 16), 256, 0, s>>>(m->d1, m->A[0], NH, B, m->g + OW2, m->g + OB2);
  k_dense_bwd_x<<<dim3((B + 255) / 256, NH), 256, 0, s>>>(m->th + OW2, m->d1, NH, NH, B, m->d0);
  k_gelu_bwd<<<grid1(nh), 256, 0, s>>>(m->d0, m->Z[0], nh, m->d1);
  // layer 1
  k_dense_bwd_w<<<dim3(NH, (NIN + 1 + 15) / 16), 256, 0, s>>>(m->d1, X, NIN, B, m->g + OW1, m->g + OB1);
}

int read_loss(koral_hip_mlp *m, float *loss, hipStream_t s)
{
  if (!loss) return 0;
  int rc = check_hip(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, s), "loss");
  return rc ? rc : check_hip(hipStreamSynchronize(s), "sync");
}

}  // namespace

extern "C" {

int koral_hip_mlp_nparam(void) { return NPARAM; }

int koral_hip_mlp_create(int kind, koral_hip_mlp **out)
{
  if (!out || kind < 0 || kind > 3) return set_error(KORAL_E_ARG, "mlp_create: bad arguments");
  auto *m = new koral_hip_mlp();
  m->kind = kind;
  int rc = 0;
  float **vecs[4] = {&m->th, &m->m, &m->v, &m->g};
  for (int k = 0; k < 4 && !rc; k++) rc = m->alloc(vecs[k], NPARAM);
  if (!rc) rc = m->alloc(&m->norm, NNORM);
  if (!rc) rc = m->alloc(&m->wp, NPACK);
  for (int k = 0; k < 3 && !rc; k++) rc = m->alloc(&m->Z[k], (size_t)NH * BMAX);
  for (int k = 0; k < 3 && !rc; k++) rc = m->alloc(&m->A[k], (size_t)NH * BMAX);
  if (!rc) rc = m->alloc(&m->Y, (size_t)NOUT * BMAX);
  if (!rc) rc = m->alloc(&m->dY, (size_t)NOUT * BMAX);
  if (!rc) rc = m->alloc(&m->d0, (size_t)NH * BMAX);
  if (!rc) rc = m->alloc(&m->d1, (size_t)NH * BMAX);
  if (!rc) rc = m->alloc(&m->loss, 1);
  if (!rc) rc = check_hip(hipMemset(m->m, 0, NPARAM * sizeof(float)), "memset");
  if (!rc) rc = check_hip(hipMemset(m->v, 0, NPARAM * sizeof(float)), "memset");
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return 0;
}

int koral_hip_mlp_destroy(koral_hip_mlp *m)
{
  if (m) {
    (void)hipDeviceSynchronize();
    delete m;
  }
  return 0;
}

int koral_hip_mlp_set(koral_hip_mlp *m, const float *theta, const float *norm)
{
  if (!m || !theta || !norm) return set_error(KORAL_E_ARG, "mlp_set: bad arguments");
  int rc = check_hip(hipMemcpy(m->th, theta, NPARAM * sizeof(float), hipMemcpyHostToDevice), "mlp_set");
  if (!rc) rc = check_hip(hipMemcpy(m->norm, norm, NNORM * sizeof(float), hipMemcpyHostToDevice), "mlp_set");
  if (!rc) rc = check_hip(hipMemset(m->m, 0, NPARAM * sizeof(float)), "memset");
  if (!rc) rc = check_hip(hipMemset(m->v, 0, NPARAM * sizeof(float)), "memset");
  m->step = 0;
  m->packed = false;
  return rc;
}

int koral_hip_mlp_get(koral_hip_mlp *m, float *theta, float *norm)
{
  if (!m) return set_error(KORAL_E_ARG, "mlp_get: null");
  int rc = check_hip(hipDeviceSynchronize(), "sync");
  if (!rc && theta) rc = check_hip(hipMemcpy(theta, m->th, NPARAM * sizeof(float), hipMemcpyDeviceToHost), "mlp_get");
  if (!rc && norm) rc = check_hip(hipMemcpy(norm, m->norm, NNORM * sizeof(float), hipMemcpyDeviceToHost), "mlp_get");
  return rc;
}

int koral_hip_mlp_features(const koral_params *P, const koral_hip_mlp *m, long n, const double *pp,
                           const double *dt, float *feat, void *stream)
{
  if (check_params(P) || !P->rad || n < 0 || !pp || !dt || !feat) return set_error(KORAL_E_ARG, "mlp_features");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const float *norm = m ? m->norm : nullptr;
  if (P->mag)
    k_features<1><<<grid1(n), 256, 0, s>>>(*P, n, pp, dt, norm, feat);
  else
    k_features<0><<<grid1(n), 256, 0, s>>>(*P, n, pp, dt, norm, feat);
  return check_launch("mlp_features");
}

int koral_hip_mlp_kind(const koral_hip_mlp *m) { return m ? m->kind : -1; }

int koral_hip_mlp_targets(int kind, long n, int nv, const double *pp0, const double *pp1, float *targ,
                          void *stream)
{
  if (n < 0 || !pp0 || !pp1 || !targ || kind < 0 || kind > 1 || (kind == 1 && nv < 10))
    return set_error(KORAL_E_ARG, "mlp_targets: bad arguments");
  if (n == 0) return 0;
  k_targets<<<grid1(n), 256, 0, (hipStream_t)stream>>>(n, kind == 0 ? 1 : nv - 4, pp0, pp1, targ);
  return check_launch("mlp_targets");
}

int koral_hip_mlp_u2p_features(const koral_params *P, const koral_hip_mlp *m, long n, const double *uu,
                               const double *pp, const double *g20, const double *G20, float *feat, void *stream)
{
  if (check_params(P) || n < 0 || !uu || !pp || !g20 || !G20 || !feat || (m && m->kind != 2))
    return set_error(KORAL_E_ARG, "mlp_u2p_features: bad arguments");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const float *norm = m ? m->norm : nullptr;
  if (P->mag)
    k_u2p_features<1><<<grid1(n), 256, 0, s>>>(*P, n, uu, pp, g20, G20, norm, feat);
  else
    k_u2p_features<0><<<grid1(n), 256, 0, s>>>(*P, n, uu, pp, g20, G20, norm, feat);
  return check_launch("mlp_u2p_features");
}

int koral_hip_mlp_u2p_targets(const koral_params *P, long n, const double *pp0, const double *pp1,
                              const double *g20, const double *G20, float *targ, void *stream)
{
  if (check_params(P) || n < 0 || !pp0 || !pp1 || !g20 || !G20 || !targ)
    return set_error(KORAL_E_ARG, "mlp_u2p_targets: bad arguments");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (P->mag)
    k_u2p_targets<1><<<grid1(n), 256, 0, s>>>(*P, n, pp0, pp1, g20, G20, targ);
  else
    k_u2p_targets<0><<<grid1(n), 256, 0, s>>>(*P, n, pp0, pp1, g20, G20, targ);
  return check_launch("mlp_u2p_targets");
}

int koral_hip_mlp_u2p_guess(const koral_params *P, const koral_hip_mlp *m, long n, const double *pp,
                            const double *g20, const double *G20, const float *Y, double *W0, void *stream)
{
  if (check_params(P) || n < 0 || !m || m->kind != 2 || !pp || !g20 || !G20 || !Y || !W0)
    return set_error(KORAL_E_ARG, "mlp_u2p_guess: bad arguments");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (P->mag)
    k_u2p_guess<1><<<grid1(n), 256, 0, s>>>(*P, n, pp, g20, G20, Y, m->norm, W0);
  else
    k_u2p_guess<0><<<grid1(n), 256, 0, s>>>(*P, n, pp, g20, G20, Y, m->norm, W0);
  return check_launch("mlp_u2p_guess");
}

int koral_hip_mlp_forward(koral_hip_mlp *m, const float *X, int B, float *Y, void *stream)
{
  if (!m || !X || !Y || B < 1 || B > BMAX) return set_error(KORAL_E_ARG, "mlp_forward: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  forward(m, X, B, s);
  int rc = check_hip(hipMemcpyAsync(Y, m->Y, (size_t)NOUT * B * sizeof(float), hipMemcpyDeviceToDevice, s), "copy");
  return rc ? rc : check_launch("mlp_forward");
}

int koral_hip_mlp_grad(koral_hip_mlp *m, const float *X, const float *T, int B, float *grad, float *loss,
                       void *stream)
{
  if (!m || !X || !T || B < 1 || B > BMAX) return set_error(KORAL_E_ARG, "mlp_grad: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  forward(m, X, B, s);
  backward(m, X, T, B, s);
  int rc = grad ? check_hip(hipMemcpyAsync(grad, m->g, NPARAM * sizeof(float), hipMemcpyDeviceToDevice, s), "copy")
                : 0;
  if (!rc) rc = check_launch("mlp_grad");
  return rc ? rc : read_loss(m, loss, s);
}

int koral_hip_mlp_train_step(koral_hip_mlp *m, const float *X, const float *T, int B, float lr, float *loss,
                             void *stream)
{
  if (!m || !X || !T || B < 1 || B > BMAX) return set_error(KORAL_E_ARG, "mlp_train_step: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  forward(m, X, B, s);
  backward(m, X, T, B, s);
  m->step++;
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1. - pow(b1, m->step)), c2 = (float)(1. - pow(b2, m->step));
  k_adam<<<grid1(NPARAM), 256, 0, s>>>(m->th, m->g, m->m, m->v, NPARAM, lr, (float)b1, (float)b2, c1, c2, 1e-8f);
  m->packed = false;
  int rc = check_launch("mlp_train_step");
  return rc ? rc : read_loss(m, loss, s);
}

int koral_hip_mlp_infer(koral_hip_mlp *m, const float *feat, long n, float *Y, void *stream)
{
  if (!m || !feat || !Y || n < 0) return set_error(KORAL_E_ARG, "mlp_infer: bad arguments");
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  int rc = pack(m, s);
  if (rc) return rc;
  long blocks = (n + 63) / 64;
  if (blocks > 2048) blocks = 2048;
  k_mlp_infer<<<(unsigned)blocks, 256, 0, s>>>(m->wp, m->th, feat, n, Y);
  return check_launch("mlp_infer");
}

int koral_hip_mlp_gate(koral_hip_mlp *m, long n, const float *Y, float thr, double *gxg, void *stream)
{
  if (!m || m->kind != 3 || n < 0 || !Y || !gxg) return set_error(KORAL_E_ARG, "mlp_gate: bad arguments");
  if (n == 0) return 0;
  k_gate<<<grid1(n), 256, 0, (hipStream_t)stream>>>(n, Y, m->norm, thr, gxg);
  return check_launch("mlp_gate");
}

int koral_hip_mlp_guess(koral_hip_mlp *m, long n, int nv, const double *pp, const float *Y, double *gx,
                        void *stream)
{
  if (!m || n < 0 || !pp || !Y || !gx || (m->kind == 1 && nv < 10))
    return set_error(KORAL_E_ARG, "mlp_guess: bad arguments");
  if (n == 0) return 0;
  k_guess<<<grid1(n), 256, 0, (hipStream_t)stream>>>(n, m->kind == 0 ? 1 : nv - 4, pp, Y, m->norm, gx);
  return check_launch("mlp_guess");
}

}  // extern "C"
