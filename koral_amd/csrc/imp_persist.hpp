// The implicit-source solve of many cells as six passes, one per attempt of
// solve_implicit_lab's fallback chain (rad.c:1389-1464: energy LAB, energy
// FF, swapped LAB, swapped FF, entropy LAB, entropy swapped).
//
// Pass a works on the cells still unsolved after a - 1 attempts, held in two
// lists by the primitives attempt a iterates on (MHD or RAD).  Each list is
// solved by its own kernel, specialised at compile time for the attempt's
// (primitives, equation, frame), so the lanes of a wave run the same branch
// of the residual.  The kernels are persistent: a lane owns one cell at a
// time and advances it one damped-Newton iteration per pass of its loop
// (imp_newton_step); a lane whose cell is done takes the next list entry
// from a work counter (the wave ballots its idle lanes, one lane claims that
// many entries with one atomic add, each idle lane takes its rank).  A cell
// that fails the attempt is appended, by the same ballot/prefix-count
// scheme, to the list of the next attempt.  Cells need 1 to 51 iterations
// per attempt, so no lane waits for the slowest cell of its wave.
//
// With a learned initial guess a first pass runs before the chain: the
// energy-equation lab-frame solve from a start given per cell.  With a
// radiation guess the learned classification takes the primitives of the
// chain's attempt 0 (the old state's E^ < RADIMPLICITTHRESHOLD u choice,
// rad.c:1381): the radiation primitives from the radiation network's guess
// (the gas part from the gas guess), or the gas primitives from the gas
// guess; without a radiation guess every cell takes the gas primitives.
// Iterations on the gas primitives are much cheaper (no inner gas inversion
// per residual), so a cell whose old state favours them tries them first
// even where it will need the radiation primitives.  With both guesses a
// cell the first learned try does not solve within its iteration budget
// tries the other primitive set from its guess (as the plain chain falls
// back between primitive sets).  A cell the learned tries do not solve enters
// attempt 0 of the plain chain unchanged, so the guess can change the
// iteration count but not what the chain would do without it.  A cell
// solved by a learned try reports attempt 6.
//
// Src provides, for a cell index c:
//   time()                         global time of the solve (device side)
//   load(c, uu00, pp00, g, dt)     inputs of the solve
//   init(c)                        before the first attempt
//   add_iters(c, it)               after every attempt
//   done_ok(c, uu, pp, it, attempt, whichprim, whicheq, whichframe)
//   done_fail(c)                   after the sixth failed attempt
#pragma once
#include "dev_implicit.hpp"
#include "capi_util.hpp"

namespace kd {

// Device buffers of the passes: two list pairs (ping-pong, each n entries per
// primitive kind) and IMP_NCTR counters.  Counter 4 a + k: k = 0, 1 the list
// lengths for attempt a (MHD, RAD); k = 2, 3 the work counters of its two
// kernels.  The first learned try reads list pair 1 (free until attempt 0
// writes its failures there) with lengths IMP_LEARN_LEN + k and work
// counters IMP_LEARN_Q + k; the second try reads the lists second[k] with
// lengths IMP_SECOND_LEN + k and work counters IMP_SECOND_Q + k (counters
// 24-27 are otherwise unused: attempt 5 appends nothing).
constexpr int IMP_NCTR = 32;
constexpr int IMP_SECOND_LEN = 24;
constexpr int IMP_SECOND_Q = 26;
constexpr int IMP_LEARN_LEN = 28;
constexpr int IMP_LEARN_Q = 30;
constexpr int IMP_LEARNED_ATTEMPT = 6;
struct ImpBufs {
  unsigned *list[2][2];
  unsigned *second[2];  // lists of the second learned try (learned solves only)
  unsigned *ctr;
};

// pass modes of k_imp_pass
enum { IMP_CHAIN = 0, IMP_LEARN1 = 1, IMP_LEARN2 = 2 };

KD_INL int imp_other(int wp) { return wp == P_RAD ? P_MHD : P_RAD; }

// primitives of attempt a + 1 for a cell iterating on wp in attempt a
KD_INL int imp_next_wp(int a, int wp)
{
  const bool swapped = (a == 2 || a == 3 || a == 5);
  const int startwith = swapped ? imp_other(wp) : wp;
  const bool swapped_next = (a + 1 == 2 || a + 1 == 3 || a + 1 == 5);
  return swapped_next ? imp_other(startwith) : startwith;
}

// sorts the n cells into the two attempt-0 lists by the primitives the
// chain starts with
template <int MAG, int RAD, class Src>
__global__ void __launch_bounds__(256) k_imp_classify(koral_params P, Src src, long n, ImpBufs b)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int wp = 0;
  if (c < n) {
    src.init(c);
    double uu00[NV], pp00[NV], dt;
    Geom g;
    src.load(c, uu00, pp00, g, dt);
    wp = imp_startwith<MAG, RAD>(P, pp00, g);
  }
  append_index(wp == P_MHD, (unsigned)c, b.list[0][0], &b.ctr[0]);
  append_index(wp == P_RAD, (unsigned)c, b.list[0][1], &b.ctr[1]);
}

// learned classification: the primitive kind of every cell's first learned
// try (see above), or, for a cell whose gas start gxg is NaN, of its first
// chain attempt; gxr may be null
template <int MAG, class Src>
__global__ void __launch_bounds__(256) k_imp_learned_classify(koral_params P, Src src, long n, ImpBufs b,
                                                              const double *__restrict__ gxg,
                                                              const double *__restrict__ gxr)
{
  constexpr int NV = Lay<MAG, 1>::NV;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int wp = 0;
  bool skip = false;  // no learned try: a NaN gas start (gated out)
  if (c < n) {
    src.init(c);
    wp = P_MHD;
    skip = isnan(gxg[c]);
    if (gxr || skip) {
      double uu00[NV], pp00[NV], dt;
      Geom g;
      src.load(c, uu00, pp00, g, dt);
      wp = imp_startwith<MAG, 1>(P, pp00, g);
    }
  }
  // a skipped cell joins attempt 0 of the chain as a failed learned try does
  append_index(skip && wp == P_MHD, (unsigned)c, b.list[0][0], &b.ctr[0]);
  append_index(skip && wp == P_RAD, (unsigned)c, b.list[0][1], &b.ctr[1]);
  append_index(!skip && wp == P_MHD, (unsigned)c, b.list[1][0], &b.ctr[IMP_LEARN_LEN]);
  append_index(!skip && wp == P_RAD, (unsigned)c, b.list[1][1], &b.ctr[IMP_LEARN_LEN + 1]);
}

// attempt a on the cells of one list, all iterating on primitives WP with
// equation WE in frame WF
template <int MAG, int RAD, int WP, int WE, int WF, int MODE, class Src>
__global__ void __launch_bounds__(256) k_imp_pass(koral_params P, Src src, int a, ImpBufs b, int *counts,
                                                  const double *gxg, const double *gxr, long n_all)
{
  const double global_time = src.time();
  constexpr int NV = Lay<MAG, RAD>::NV;
  const int lane = threadIdx.x & 63;
  const int kind = (WP == P_MHD) ? 0 : 1;
  // the learned pass takes every cell in index order and hands failures to
  // attempt 0; attempt a takes its list and hands failures to attempt a + 1
  constexpr bool LEARN = MODE != IMP_CHAIN;
  // the first learned try hands failures to the second try on the other
  // primitives when there is one (b.second set), else to attempt 0; the
  // second try hands them to attempt 0; attempt a hands them to a + 1
  const bool to_second = MODE == IMP_LEARN1 && b.second[0] != nullptr;
  const unsigned *list = MODE == IMP_LEARN1 ? b.list[1][kind] : (MODE == IMP_LEARN2 ? b.second[kind] : b.list[a & 1][kind]);
  unsigned *out0 = to_second ? b.second[0] : (LEARN ? b.list[0][0] : b.list[(a + 1) & 1][0]);
  unsigned *out1 = to_second ? b.second[1] : (LEARN ? b.list[0][1] : b.list[(a + 1) & 1][1]);
  const long n = MODE == IMP_LEARN1 ? (long)b.ctr[IMP_LEARN_LEN + kind]
                                    : (MODE == IMP_LEARN2 ? (long)b.ctr[IMP_SECOND_LEN + kind] : (long)b.ctr[4 * a + kind]);
  unsigned *queue = MODE == IMP_LEARN1 ? &b.ctr[IMP_LEARN_Q + kind]
                                       : (MODE == IMP_LEARN2 ? &b.ctr[IMP_SECOND_Q + kind] : &b.ctr[4 * a + 2 + kind]);
  unsigned *out_len = to_second ? &b.ctr[IMP_SECOND_LEN] : (LEARN ? &b.ctr[0] : &b.ctr[4 * (a + 1)]);
  int next_kind = to_second ? 1 - kind : (LEARN ? 0 : ((imp_next_wp(a, WP) == P_MHD) ? 0 : 1));
  bool active = false, fresh = false, exhausted = n == 0;
  unsigned c = 0, c_push = 0;
  bool push = false;
  int iter = 0;
  int n_ok = 0, n_it = 0, n_bad = 0;
  double pp[NV];
  // The inputs of a lane's current cell, copied once when the lane takes the
  // cell: the lanes of a wave hold unrelated cells, so reloading them from
  // the cell arrays every iteration would touch one cache line per lane and
  // value.  Column threadIdx.x belongs to this lane alone (no barrier).
  constexpr int NG = 43;  // Geom: g[4][5], G[4][5], gdet, alpha, gttpert
  static_assert(sizeof(Geom) == NG * sizeof(double), "Geom layout");
  __shared__ double keep[2 * NV + NG + 1][256];
  const int me = threadIdx.x;
  for (;;) {
    // failures of the last iteration go to the next attempt's list
    append_index(push && next_kind == 0, c_push, out0, &out_len[0]);
    append_index(push && next_kind == 1, c_push, out1, &out_len[1]);
    push = false;
    if (!exhausted) {
      const unsigned long long need = __ballot(!active);
      if (need) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((unsigned long long)need) - 1;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(queue, (unsigned)cnt);
        base = __shfl(base, leader);
        if ((long)base + cnt >= n) exhausted = true;
        if (!active) {
          const long k = (long)base + __popcll(need & ((1ull << lane) - 1ull));
          if (k < n) {
            c = list[k];
            active = true;
            fresh = true;
          }
        }
      }
    }
    if (__ballot(active) == 0) break;
    if (active) {
      double uu00[NV], pp00[NV], dt;
      Geom g;
      if (fresh) {
        src.load(c, uu00, pp00, g, dt);
#pragma unroll
        for (int i = 0; i < NV; i++) {
          keep[i][me] = uu00[i];
          keep[NV + i][me] = pp00[i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 5; j++) {
            keep[2 * NV + 5 * i + j][me] = g.g[i][j];
            keep[2 * NV + 20 + 5 * i + j][me] = g.G[i][j];
          }
        keep[2 * NV + 40][me] = g.gdet;
        keep[2 * NV + 41][me] = g.alpha;
        keep[2 * NV + 42][me] = g.gttpert;
        keep[2 * NV + NG][me] = dt;
#pragma unroll
        for (int i = 0; i < NV; i++) pp[i] = pp00[i];
        if (LEARN) {
          // the gas guess also starts the gas inversion of a radiation-primitive attempt
#pragma unroll
          for (int i = 0; i < 4; i++) pp[1 + i] = gxg[i * n_all + c];
          if (WP == P_RAD) {
            constexpr int EE = Lay<MAG, RAD>::EE;
#pragma unroll
            for (int i = 0; i < 4; i++) pp[EE + i] = gxr[i * n_all + c];
          }
        }
        iter = 0;
        fresh = false;
      } else {
#pragma unroll
        for (int i = 0; i < NV; i++) {
          uu00[i] = keep[i][me];
          pp00[i] = keep[NV + i][me];
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 5; j++) {
            g.g[i][j] = keep[2 * NV + 5 * i + j][me];
            g.G[i][j] = keep[2 * NV + 20 + 5 * i + j][me];
          }
        g.gdet = keep[2 * NV + 40][me];
        g.alpha = keep[2 * NV + 41][me];
        g.gttpert = keep[2 * NV + 42][me];
        dt = keep[2 * NV + NG][me];
      }
      const int st = imp_newton_step<MAG, RAD, WP, WE, WF>(P, uu00, pp00, g, dt, pp, global_time, iter);
      if (st != IMP_CONT) {
        src.add_iters(c, iter);
        if (st == IMP_OK) {
          double uu[NV];
          p2u<MAG, RAD>(P, pp, uu, g);
          src.done_ok(c, uu, pp, iter, LEARN ? IMP_LEARNED_ATTEMPT : a, WP, WE, WF);
          n_ok++;
          n_it += iter;
        } else if (!LEARN && a == 5) {
          src.done_fail(c);
          n_bad++;
        } else {
          if (LEARN && !to_second) next_kind = (imp_startwith<MAG, RAD>(P, pp00, g) == P_MHD) ? 0 : 1;
          push = true;
          c_push = c;
        }
        active = false;
      }
    }
  }
  if (counts) {
    for (int off = 32; off > 0; off >>= 1) {
      n_ok += __shfl_xor(n_ok, off);
      n_it += __shfl_xor(n_it, off);
      n_bad += __shfl_xor(n_bad, off);
    }
    if (lane == 0) {
      if (n_ok) atomicAdd(&counts[0], n_ok);
      if (n_it) atomicAdd(&counts[1], n_it);
      if (n_bad) atomicAdd(&counts[2], n_bad);
    }
  }
}

// clears n (<= 64) counters on the launching stream (a kernel, so the clear
// can sit inside a captured graph)
static __global__ void k_zero_u32(unsigned *q, int n)
{
  if ((int)threadIdx.x < n) q[threadIdx.x] = 0u;
}

// Blocks of bs threads that fill every CU at the kernel's occupancy (never
// more than the n work items need).
template <class K>
int resident_blocks(K kernel, int bs, long n, unsigned *blocks)
{
  int dev = 0, cus = 0, per = 0;
  int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (!rc) rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
  if (!rc) rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, bs, 0), "occupancy");
  if (rc) return rc;
  if (per < 1) per = 1;
  long want = (long)cus * per;
  long need = (n + bs - 1) / bs;
  if (need < 1) need = 1;
  *blocks = (unsigned)(want < need ? want : need);
  return 0;
}

template <int MAG, int WP, int WE, int WF, int MODE = IMP_CHAIN, class Src>
int imp_launch_pass(const koral_params &P, const Src &src, long n, int a, const ImpBufs &b, int *counts,
                    hipStream_t s, const double *gxg = nullptr, const double *gxr = nullptr)
{
  unsigned blocks = 0;
  int rc = resident_blocks(k_imp_pass<MAG, 1, WP, WE, WF, MODE, Src>, 256, n, &blocks);
  if (rc) return rc;
  k_imp_pass<MAG, 1, WP, WE, WF, MODE, Src><<<blocks, 256, 0, s>>>(P, src, a, b, counts, gxg, gxr, n);
  return 0;
}

// all passes for n cells (asynchronous on s); counts may be null.  With a
// learned guess, gxg [4][n] is the gas start (u, u~^1..3) and gxr [4][n] (may
// be null) the radiation start (E, F~^1..3); the learned pass stops a cell
// after learned_maxiter iterations when that is positive and below
// RADIMPMAXITER.
template <int MAG, class Src>
int implicit_passes(const koral_params &P, const Src &src, long n, const ImpBufs &b, int *counts, hipStream_t s, const double *gxg = nullptr, const double *gxr = nullptr, int learned_maxiter = 0)
{
  const unsigned nb = (unsigned)((n + 255) / 256);
  k_zero_u32<<<1, 64, 0, s>>>(b.ctr, IMP_NCTR);
  int rc = 0;
  if (gxg) {
    k_imp_learned_classify<MAG, Src><<<nb, 256, 0, s>>>(P, src, n, b, gxg, gxr);
    koral_params Pl = P;
    if (learned_maxiter > 0 && learned_maxiter < Pl.radimpmaxiter) Pl.radimpmaxiter = learned_maxiter;
    // without a radiation guess there is no second try: b.second must then be null
    ImpBufs bl = b;
    if (!gxr) bl.second[0] = bl.second[1] = nullptr;
    rc = imp_launch_pass<MAG, P_MHD, ENERGYEQ, LAB, IMP_LEARN1>(Pl, src, n, 0, bl, counts, s, gxg, gxr);
    if (!rc && gxr) rc = imp_launch_pass<MAG, P_RAD, ENERGYEQ, LAB, IMP_LEARN1>(Pl, src, n, 0, bl, counts, s, gxg, gxr);
    if (!rc && gxr && bl.second[0]) {
      rc = imp_launch_pass<MAG, P_MHD, ENERGYEQ, LAB, IMP_LEARN2>(Pl, src, n, 0, bl, counts, s, gxg, gxr);
      if (!rc) rc = imp_launch_pass<MAG, P_RAD, ENERGYEQ, LAB, IMP_LEARN2>(Pl, src, n, 0, bl, counts, s, gxg, gxr);
    }
  } else {
    k_imp_classify<MAG, 1, Src><<<nb, 256, 0, s>>>(P, src, n, b);
  }
  for (int a = 0; a < 6 && !rc; a++) {
    if (a >= 4) {
      rc = imp_launch_pass<MAG, P_MHD, ENTROPYEQ, LAB>(P, src, n, a, b, counts, s);
      if (!rc) rc = imp_launch_pass<MAG, P_RAD, ENTROPYEQ, LAB>(P, src, n, a, b, counts, s);
    } else if (a == 1 || a == 3) {
      rc = imp_launch_pass<MAG, P_MHD, ENERGYEQ, FF>(P, src, n, a, b, counts, s);
      if (!rc) rc = imp_launch_pass<MAG, P_RAD, ENERGYEQ, FF>(P, src, n, a, b, counts, s);
    } else {
      rc = imp_launch_pass<MAG, P_MHD, ENERGYEQ, LAB>(P, src, n, a, b, counts, s);
      if (!rc) rc = imp_launch_pass<MAG, P_RAD, ENERGYEQ, LAB>(P, src, n, a, b, counts, s);
    }
  }
  return rc;
}

}  // namespace kd
