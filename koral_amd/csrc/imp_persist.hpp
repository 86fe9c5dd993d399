// Persistent implicit-source solve: every lane of a resident wave owns one
// cell at a time and advances it one damped-Newton iteration per pass of the
// loop (imp_newton_step), walking the six-attempt fallback chain of
// solve_implicit_lab (rad.c:1225-1570).  A lane whose cell has finished
// (solved or out of attempts) takes the next cell from a global work
// counter: the wave ballots its idle lanes, one lane claims that many cells
// with a single atomic add, and each idle lane takes its rank among them
// (prefix count of the ballot).  Cells need 1 to 306 iterations, so a lane
// never waits for the slowest cell of its wave; the wave leaves when the
// counter is exhausted and none of its lanes holds a cell.
//
// Per-iteration state kept in registers: the iterate (NV doubles), the cell
// index and the attempt/iteration counters.  The cell's inputs (U, P, the
// metric, dt) are re-read from memory every iteration.
//
// Src provides, for a work index c in [0, n):
//   load(c, uu00, pp00, g, dt)
//   done_ok(c, uu, pp, iters, attempt, whichprim, whicheq, whichframe, total)
//   done_fail(c, total)
#pragma once
#include "dev_implicit.hpp"
#include "capi_util.hpp"

namespace kd {

template <int MAG, int RAD, class Src>
__device__ __forceinline__ void implicit_persistent(const koral_params &P, const Src &src, long n,
                                                    unsigned int *queue, double global_time, int *counts)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  const int lane = threadIdx.x & 63;
  bool active = false, fresh = false, exhausted = false;
  long c = 0;
  int a = 0, iter = 0, total = 0, startwith = P_MHD;
  int n_ok = 0, n_it = 0, n_bad = 0;
  double pp[NV];
  for (;;) {
    if (!exhausted) {
      const unsigned long long need = __ballot(!active);
      if (need) {
        const int cnt = __popcll(need);
        const int leader = __ffsll((unsigned long long)need) - 1;
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(queue, (unsigned int)cnt);
        base = __shfl(base, leader);
        if ((long)base + cnt >= n) exhausted = true;
        if (!active) {
          const long cc = (long)base + __popcll(need & ((1ull << lane) - 1ull));
          if (cc < n) {
            c = cc;
            active = true;
            fresh = true;
            a = 0;
            iter = 0;
            total = 0;
          }
        }
      }
    }
    if (__ballot(active) == 0) break;
    if (active) {
      double uu00[NV], pp00[NV], dt;
      Geom g;
      src.load(c, uu00, pp00, g, dt);
      if (fresh) {
        startwith = imp_startwith<MAG, RAD>(P, pp00, g);
#pragma unroll
        for (int i = 0; i < NV; i++) pp[i] = pp00[i];
        fresh = false;
      }
      int wp, we, wf;
      imp_attempt(a, startwith, wp, we, wf);
      const int st = imp_newton_step<MAG, RAD>(P, uu00, pp00, g, dt, wp, we, wf, pp, global_time, iter);
      if (st != IMP_CONT) {
        total += iter;
        if (st == IMP_OK) {
          double uu[NV];
          p2u<MAG, RAD>(P, pp, uu, g);
          src.done_ok(c, uu, pp, iter, a, wp, we, wf, total);
          n_ok++;
          n_it += iter;
          active = false;
        } else if (++a == 6) {
          src.done_fail(c, total);
          n_bad++;
          active = false;
        } else {
          iter = 0;
#pragma unroll
          for (int i = 0; i < NV; i++) pp[i] = pp00[i];
        }
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

// clears a work counter on the launching stream (a kernel, so the clear can
// sit inside a captured graph)
static __global__ void k_zero_u32(unsigned int *q) { *q = 0u; }

}  // namespace kd
