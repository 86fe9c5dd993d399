// Grid-level operators of one KORAL timestep as per-thread functors, and the
// order in which a step applies them.
//
// Each Op* functor does the work of one thread (one cell, face or edge) in
// operator()(i) for i in [0, size()).  The step order (set_bc, op_explicit,
// calc_u2p with cell_fixup, op_implicit, RK2IMEX combinations) is written
// once against an executor Ex that provides
//   ex.each(op)          run op(i) for every i
//   ex.fork(k), ex.each_on(k, op), ex.join(k)   the same on side stream k,
//                        ordered after the main stream's work at fork and
//                        before its later work at join
//   ex.max(op, out)      *out = max(*out, max_i op(i))     (op(i) >= 0)
//   ex.count(op, out3)   out3[k] += sum_i op(i).v[k], k < 3
//   ex.implicit_cells(op, out3)   the implicit sweep OpImplicit with its
//                        counts (the HIP executor runs it as six persistent
//                        per-attempt passes over compacted work lists)
// The library's executor launches one HIP kernel per call on the grid's
// stream (k_grid.hip).
#pragma once
#include <initializer_list>
#include <type_traits>
#include "dev_grid.hpp"

namespace kd {

// Operator kinds (per-kind launch timing and names in the C-ABI)
enum OpKind {
  K_P2U, K_BC_FACES, K_BC_EDGE, K_BC_TOTAL, K_WAVESPEEDS, K_FLUX, K_EMF, K_CT, K_UPDATE,
  K_U2P, K_FIXUP, K_FIXUP_APPLY, K_IMPLICIT, K_COMB, K_POLAR, K_RECON, K_FLUX_FLOOR, K_U2P_FLOOR, K_ZERO, K_NKINDS
};

struct Box {
  int x0, y0, z0, nx, ny, nz;
  KD_HD long size() const { return (long)nx * ny * nz; }
  KD_HD void at(long i, int &x, int &y, int &z) const
  {
    x = x0 + (int)(i % nx);
    long r = i / nx;
    y = y0 + (int)(r % ny);
    z = z0 + (int)(r / ny);
  }
};

struct Cnt3 {
  int v[3];
};

template <int MAG, int RAD>
KD_INL void load_p(const double *p, long n, long c, double *v)
{
  const unsigned cb = soa_off(c);
#pragma unroll
  for (int iv = 0; iv < Lay<MAG, RAD>::NV; iv++) v[iv] = soa_ld(p, iv * n, cb);
}
template <int MAG, int RAD>
KD_INL void store_p(double *p, long n, long c, const double *v)
{
  const unsigned cb = soa_off(c);
#pragma unroll
  for (int iv = 0; iv < Lay<MAG, RAD>::NV; iv++) soa_st(p, iv * n, cb, v[iv]);
}

template <int MAG, int RAD>
KD_INL void p2u_at(const koral_params &P, const DGrid &G, const double *p, double *u, int ix, int iy, int iz)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  long c = cidx(G, ix, iy, iz);
  Geom g;
  geom_cell(G, ix, iy, g);
  double pp[NV], uu[NV];
  load_p<MAG, RAD>(p, G.ncell, c, pp);
  kd::p2u<MAG, RAD>(P, pp, uu, g);
  store_p<MAG, RAD>(u, G.ncell, c, uu);
}

// u = p2u(p) over a box (set_initial_profile leaves u consistent with p)
template <int MAG, int RAD>
struct OpP2U {
  static constexpr int kind = K_P2U;
  koral_params P;
  DGrid G;
  const double *p;
  double *u;
  Box b;
  KD_HD long size() const { return b.size(); }
  KD_INL void operator()(long i) const
  {
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    p2u_at<MAG, RAD>(P, G, p, u, ix, iy, iz);
  }
};

// ---------------------------------------------------------------------------
// ghost fill (set_bc finite.c:2752-3700)
// ---------------------------------------------------------------------------
// Primitives of the face-ghost cell (x, y, z), index d outside the domain and
// the other two inside, as the face fill writes them: an exact copy of the
// cell that the boundary type of that side maps it to, with the d components
// of the vectors negated on a reflecting side.  Every coordinate and
// component goes through a select over all three directions (a loop that
// touches only index d is folded into a run-time index, and the array then
// lives in scratch memory).
template <int MAG, int RAD>
KD_INL void face_ghost_p(const DGrid &G, const double *p, int d, int x, int y, int z, double *pv)
{
  const int gk = d == 0 ? x : (d == 1 ? y : z);
  const int n = dimn(G, d), hi = gk >= n;
  const int type = d == 0 ? (hi ? G.bc[1] : G.bc[0]) : (d == 1 ? (hi ? G.bc[3] : G.bc[2]) : (hi ? G.bc[5] : G.bc[4]));
  int sk = gk;
  bool flip = false;
  if (type == BC_PERIODIC) {
    if (sk < 0) sk += n;
    if (sk > n - 1) sk -= n;
  } else if (type == BC_REFLECT) {
    sk = hi ? 2 * n - sk - 1 : -sk - 1;
    flip = true;
  } else {
    sk = hi ? n - 1 : 0;
  }
  load_p<MAG, RAD>(p, G.ncell, cidx(G, d == 0 ? sk : x, d == 1 ? sk : y, d == 2 ? sk : z), pv);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const bool f = flip && k == d;
    pv[2 + k] = f ? -pv[2 + k] : pv[2 + k];
    if (MAG) pv[6 + k] = f ? -pv[6 + k] : pv[6 + k];
    if (RAD) pv[Lay<MAG, RAD>::EE + 1 + k] = f ? -pv[Lay<MAG, RAD>::EE + 1 + k] : pv[Lay<MAG, RAD>::EE + 1 + k];
  }
}

// ghost cells outside the domain in exactly one dimension
template <int MAG, int RAD>
struct OpBcFaces {
  static constexpr int kind = K_BC_FACES;
  koral_params P;
  DGrid G;
  double *p, *u;
  // the items are the face-ghost cells only: for each direction d that is
  // not collapsed, its 2 NG_d ghost layers over the interior of the other
  // two directions (the cells with exactly one index outside the domain)
  KD_HD int ngd(int d) const { return d == 0 ? G.ngx : (d == 1 ? G.ngy : G.ngz); }
  KD_HD long slab(int d) const
  {
    return dimn(G, d) > 1 ? 2L * ngd(d) * (d == 0 ? 1L : G.nx) * (d == 1 ? 1L : G.ny) * (d == 2 ? 1L : G.nz) : 0;
  }
  KD_HD long size() const { return slab(0) + slab(1) + slab(2); }
  KD_INL void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    int d = 0;
    if (i >= slab(0)) {
      i -= slab(0);
      d = 1;
      if (i >= slab(1)) {
        i -= slab(1);
        d = 2;
      }
    }
    const int n = dimn(G, d), ng = ngd(d);
    const Box b{0, 0, 0, d == 0 ? 2 * ng : G.nx, d == 1 ? 2 * ng : G.ny, d == 2 ? 2 * ng : G.nz};
    int ii[3];
    b.at(i, ii[0], ii[1], ii[2]);
#pragma unroll
    for (int k = 0; k < 3; k++) ii[k] = k != d ? ii[k] : (ii[k] >= ng ? n + ii[k] - ng : ii[k] - ng);
    double pv[NV];
    face_ghost_p<MAG, RAD>(G, p, d, ii[0], ii[1], ii[2], pv);
    const long c = cidx(G, ii[0], ii[1], ii[2]);
    store_p<MAG, RAD>(p, G.ncell, c, pv);
    Geom g;
    geom_cell(G, ii[0], ii[1], g);
    double uu[NV];
    kd::p2u<MAG, RAD>(P, pv, uu, g);
    store_p<MAG, RAD>(u, G.ncell, c, uu);
  }
};

// Ghosts along the four a-b edges of the domain, at the positions k inside
// the domain along the third dimension c: finite.c:3336-3600 recipe.  The
// copies of the far face ghosts (NG-1 deep on each side of the edge) and the
// average next to the corner, (ea, eb) = <(ea, ib), (ia, eb)>, read face
// ghosts.  The reference then averages two of those copies on the diagonal,
// (ea2, eb2) = <(ea2, eb), (ea, eb2)>; the copies are exact, so the same item
// reads their sources (ea2, ib) and (ia, eb2) instead, with the same operand
// order.  Each written cell gets u = p2u(p); NG >= 2.
template <int MAG, int RAD>
struct EdgeGeo {
  DGrid G;
  int a, b;
  KD_HD int c() const { return 3 - a - b; }
  KD_HD int ngc() const { return c() == 0 ? G.ngx : (c() == 1 ? G.ngy : G.ngz); }  // 0 if collapsed
  KD_HD int nk() const { return (c() == 0 ? G.nx : (c() == 1 ? G.ny : G.nz)) + 2 * ngc(); }
  // index d of the cell at (va, vb) in the a-b plane and k along c, by
  // selects (an array indexed by a and b would live in scratch memory)
  KD_INL int coord(int d, int va, int vb, int k) const { return a == d ? va : (b == d ? vb : k); }
  KD_INL long cell(int va, int vb, int k) const
  {
    return cidx(G, coord(0, va, vb, k), coord(1, va, vb, k), coord(2, va, vb, k));
  }
  KD_INL void pos(int va, int vb, int k, int &x, int &y) const
  {
    x = coord(0, va, vb, k);
    y = coord(1, va, vb, k);
  }
};

// the edge cell (da, db) at k gets the face ghost (s1a, s1b) (index d1
// outside the domain) or, averaging, <(s1a, s1b), (s2a, s2b)> with the
// second an E.b face ghost; face ghosts are read through the face fill's map
// from the cells inside, so this needs no earlier face fill
template <int MAG, int RAD>
KD_INL void edge_set(const koral_params &P, const EdgeGeo<MAG, RAD> &E, double *p, double *u, int k, int da, int db,
                     int d1, int s1a, int s1b, int s2a, int s2b, bool average)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  const long n = E.G.ncell;
  double pv[NV];
  face_ghost_p<MAG, RAD>(E.G, p, d1, E.coord(0, s1a, s1b, k), E.coord(1, s1a, s1b, k), E.coord(2, s1a, s1b, k), pv);
  if (average) {
    double pw[NV];
    face_ghost_p<MAG, RAD>(E.G, p, E.b, E.coord(0, s2a, s2b, k), E.coord(1, s2a, s2b, k), E.coord(2, s2a, s2b, k),
                           pw);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) pv[iv] = .5 * (pv[iv] + pw[iv]);
  }
  const long d = E.cell(da, db, k);
  store_p<MAG, RAD>(p, n, d, pv);
  int x, y;
  E.pos(da, db, k, x, y);
  Geom g;
  geom_cell(E.G, x, y, g);
  double uu[NV];
  kd::p2u<MAG, RAD>(P, pv, uu, g);
  store_p<MAG, RAD>(u, n, d, uu);
}

template <int MAG, int RAD>
struct OpBcEdgeA {
  static constexpr int kind = K_BC_EDGE;
  koral_params P;
  EdgeGeo<MAG, RAD> E;
  double *p, *u;
  KD_HD int per() const { return 2 * (E.G.ng - 1) + 2; }  // items per (corner, k)
  KD_HD int nkk() const { return E.nk() - 2 * E.ngc(); }  // positions inside along c
  KD_HD long size() const { return 4L * nkk() * per(); }
  KD_INL void operator()(long t) const
  {
    const int NG = E.G.ng;
    const int m = (int)(t % per());
    const long r = t / per();
    const int k = (int)(r % nkk());
    const int corner = (int)(r / nkk());
    const int ahi = corner >> 1, bhi = corner & 1;
    const int na = dimn(E.G, E.a), nb = dimn(E.G, E.b);
    const int ea = ahi ? na : -1, ia = ahi ? na - 1 : 0;
    const int eb = bhi ? nb : -1, ib = bhi ? nb - 1 : 0;
    if (m < NG - 1) {
      const int ga = ahi ? na + m + 1 : -NG + m;
      edge_set<MAG, RAD>(P, E, p, u, k, ga, eb, E.a, ga, ib, 0, 0, false);
    } else if (m < 2 * (NG - 1)) {
      const int i = m - (NG - 1);
      const int gb = bhi ? nb + i + 1 : -NG + i;
      edge_set<MAG, RAD>(P, E, p, u, k, ea, gb, E.b, ia, gb, 0, 0, false);
    } else if (m == 2 * (NG - 1)) {
      edge_set<MAG, RAD>(P, E, p, u, k, ea, eb, E.a, ea, ib, ia, eb, true);
    } else {
      const int ea2 = ahi ? na + 1 : -2, eb2 = bhi ? nb + 1 : -2;
      edge_set<MAG, RAD>(P, E, p, u, k, ea2, eb2, E.a, ea2, ib, ia, eb2, true);
    }
  }
};

// the eight NG^3 blocks at the domain corners: copies of the corner cell
template <int MAG, int RAD>
struct OpBcTotal {
  static constexpr int kind = K_BC_TOTAL;
  koral_params P;
  DGrid G;
  double *p, *u;
  KD_HD long size() const { return 8L * G.ng * G.ng * G.ng; }
  KD_INL void operator()(long t) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const int NG = G.ng;
    const long per = (long)NG * NG * NG;
    int blk = (int)(t / per);
    int r = (int)(t % per);
    int hx = blk & 1, hy = (blk >> 1) & 1, hz = (blk >> 2) & 1;
    int ix = (hx ? G.nx : -NG) + r % NG;
    int iy = (hy ? G.ny : -NG) + (r / NG) % NG;
    int iz = (hz ? G.nz : -NG) + r / (NG * NG);
    long s = cidx(G, hx ? G.nx - 1 : 0, hy ? G.ny - 1 : 0, hz ? G.nz - 1 : 0);
    long d = cidx(G, ix, iy, iz);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) p[iv * G.ncell + d] = p[iv * G.ncell + s];
    p2u_at<MAG, RAD>(P, G, p, u, ix, iy, iz);
  }
};

// The whole ghost fill in one launch: the face ghosts, the edge ghosts at the
// positions inside the domain along their third dimension, and on a 3-D grid
// the corner blocks.  The edge items read face ghosts through the face map
// (from cells inside the domain) and OpBcTotal reads cells inside only, so
// the launch has no internal dependence.  The reference's order (faces, the
// x-y, x-z and y-z edges at every position, then the corners) also writes the
// edge items at ghost positions; each of those cells has all three indices
// outside the domain and OpBcTotal overwrites it (p and u), so the final
// values are the same.
template <int MAG, int RAD>
struct OpBcAll {
  static constexpr int kind = K_BC_FACES;
  // the face fill is the one member (the kernel argument itself): it selects
  // grid fields by a run-time direction, which a local copy would keep in
  // scratch memory; the edge and corner sub-operators index them by
  // constants only and are built per item from its fields
  OpBcFaces<MAG, RAD> f;
  KD_HD OpBcEdgeA<MAG, RAD> edge(int a, int b) const { return {f.P, EdgeGeo<MAG, RAD>{f.G, a, b}, f.p, f.u}; }
  KD_HD OpBcTotal<MAG, RAD> corners() const { return {f.P, f.G, f.p, f.u}; }
  KD_HD bool has2() const { return f.G.ny > 1; }
  KD_HD bool has3() const { return f.G.ny > 1 && f.G.nz > 1; }
  KD_HD long size() const
  {
    return f.size() + (has2() ? edge(0, 1).size() : 0) +
           (has3() ? edge(0, 2).size() + edge(1, 2).size() + corners().size() : 0);
  }
  KD_INL void operator()(long i) const
  {
    const long nf = f.size();
    if (i < nf) {
      f(i);
      return;
    }
    i -= nf;
    const long n01 = edge(0, 1).size();
    if (i < n01) {
      edge(0, 1)(i);
      return;
    }
    i -= n01;
    const long n02 = edge(0, 2).size();
    if (i < n02) {
      edge(0, 2)(i);
      return;
    }
    i -= n02;
    const long n12 = edge(1, 2).size();
    if (i < n12)
      edge(1, 2)(i);
    else
      corners()(i - n12);
  }
};

// ---------------------------------------------------------------------------
// wave speeds on loop_1 (domain +- lim) and the time-step density
// (calc_wavespeeds finite.c:568-599, save_wavespeeds finite.c:296-398)
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
struct OpWavespeeds {
  static constexpr int kind = K_WAVESPEEDS;
  koral_params P;
  DGrid G;
  const double *p;
  double *ws;
  Box b;
  KD_HD long size() const { return b.size(); }
  KD_INL double operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    int out = (ix < 0 || ix >= G.nx) + (iy < 0 || iy >= G.ny) + (iz < 0 || iz >= G.nz);
    if (!(MAG || out <= 1)) return 0.;
    long c = cidx(G, ix, iy, iz);
    double pp[NV], a[NWS];
    load_p<MAG, RAD>(p, G.ncell, c, pp);
    Geom g;
    geom_cell(G, ix, iy, g);
    wavespeeds_cell<MAG, RAD>(P, G, pp, ix, iy, iz, g, a);
#pragma unroll
    for (int k = 0; k < NWS; k++)
      // without radiation only the gas slots (0..5, 12..14) are read
      if (RAD || k < 6 || (k >= 12 && k < 15)) ws[k * G.ncell + c] = a[k];
    if (out != 0) return 0.;
    double w[3];
#pragma unroll
    for (int d = 0; d < 3; d++) w[d] = RAD ? mx(a[12 + d], a[15 + d]) : a[12 + d];
    double dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
    double t;
    if (G.nz > 1 && G.ny > 1)
      t = w[0] / dx + w[1] / dy + w[2] / dz;
    else if (G.nz == 1 && G.ny > 1)
      t = w[0] / dx + w[1] / dy;
    else if (G.ny == 1 && G.nz > 1)
      t = w[0] / dx + w[2] / dz;
    else
      t = w[0] / dx;
    return t / P.tsteplim;
  }
};

// ---------------------------------------------------------------------------
// per face: reconstruct both sides, face floors, fluxes, Riemann solver
// (finite.c:694-1014 and f_calc_fluxes_at_faces finite.c:1297-1415)
// ---------------------------------------------------------------------------
// stride between neighbouring cells along the swept direction D (never a
// collapsed one) and the coordinate index along it
template <int D>
KD_INL long dir_stride(const DGrid &G)
{
  return D == 0 ? 1L : (D == 1 ? (long)G.sx : (long)G.sx * G.sy);
}
template <int D>
KD_INL int dir_index(int ix, int iy, int iz)
{
  return D == 0 ? ix : (D == 1 ? iy : iz);
}

// Reconstruction of one variable at a face of direction D from the cells
// -3..+2 around it (finite.c:694-1014, avg2point finite.c:90-289): q points
// at the variable in cell 0 (the right cell of the face), st is the cell
// stride of direction D, dx the cell widths (PPM only).  l is the state on
// the left side of the face (right edge of cell -1), r on the right side.
template <int ORDER>
KD_INL void recon_var(const koral_params &P, const double *__restrict__ q, long st, const double *dx, double &l,
                      double &r)
{
  if (ORDER == 1) {
    const double th = P.minmod_theta;
    double um1 = q[-2 * st], u0 = q[-st], up1 = q[0];
    l = u0 + .5 * minmod3(th * (up1 - u0), .5 * (up1 - um1), th * (u0 - um1));
    double vm1 = u0, v0 = up1, vp1 = q[st];
    r = v0 - .5 * minmod3(th * (vp1 - v0), .5 * (vp1 - vm1), th * (v0 - vm1));
    return;
  }
  double v[6];
#pragma unroll
  for (int o = 0; o < 6; o++) v[o] = q[(o - 3) * st];
  if (ORDER == 4) {
    l = mp5(v[0], v[1], v[2], v[3], v[4]);
    r = mp5(v[5], v[4], v[3], v[2], v[1]);
  } else {
    double ul, ur;
    ppm(v, dx, ul, ur);
    l = ur;
    ppm(v + 1, dx + 1, ul, ur);
    r = ul;
  }
}

// Face states of one sweep, KORAL's pbL/pbR arrays: both sides of every face
// of box b, stored compactly by box index, fs[iv][i] (left) and fs[NV + iv][i]
// (right).  One variable at a time, so the kernel holds few registers and
// runs at full occupancy; the branchy limiters (MP5) stay out of the
// register-heavy flux kernel.
template <int MAG, int RAD, int ORDER, int D>
struct OpRecon {
  static constexpr int kind = K_RECON;
  koral_params P;
  DGrid G;
  const double *p;
  double *fs;
  Box b;
  KD_HD long size() const { return b.size(); }
  KD_INL void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    const long n = G.ncell, nb = b.size();
    const long c0 = cidx(G, ix, iy, iz), st = dir_stride<D>(G);
    double dx[6];
    const int id = dir_index<D>(ix, iy, iz);
#pragma unroll
    for (int o = 0; o < 6; o++) dx[o] = (ORDER == 2) ? size_x(G, id + o - 3, D) : 0.;
#pragma unroll 1
    for (int iv = 0; iv < NV; iv++) {
      double l, r;
      recon_var<ORDER>(P, p + iv * n + c0, st, dx, l, r);
      fs[iv * nb + i] = l;
      fs[(NV + iv) * nb + i] = r;
    }
  }
};

// OpRecon for MP5 across x2 or x3 (D = 1, 2), a thread marching over a run
// of RUN faces along D: the six-cell window of each variable slides by one
// cell per face, so each cell is loaded once per run instead of six times
// (the stencil cells of these sweeps lie rows or planes apart and were
// re-read from HBM).  The same values reach mp5 in the same order.
template <int MAG, int RAD, int D>
struct OpReconRun {
  static constexpr int kind = K_RECON;
  static constexpr int RUN = 4;
  koral_params P;
  DGrid G;
  const double *p;
  double *fs;
  Box b;
  KD_HD int len() const { return D == 1 ? b.ny : b.nz; }
  KD_HD int runs() const { return (len() + RUN - 1) / RUN; }
  // items: x fastest, then the other transverse index, then the run
  KD_HD long size() const { return (long)b.nx * (D == 1 ? b.nz : b.ny) * runs(); }
  KD_INL void operator()(long t) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const int ix = b.x0 + (int)(t % b.nx);
    const long r1 = t / b.nx;
    const int nt = D == 1 ? b.nz : b.ny;
    const int it = (int)(r1 % nt);
    const int run = (int)(r1 / nt);
    const int j0 = run * RUN;
    const int iy0 = D == 1 ? b.y0 + j0 : b.y0 + it;
    const int iz0 = D == 1 ? b.z0 + it : b.z0 + j0;
    const long n = G.ncell, nb = b.size();
    const long st = dir_stride<D>(G);
    const long c0 = cidx(G, ix, iy0, iz0);
    const int m = len() - j0 < RUN ? len() - j0 : RUN;
    // box index of the run's first face and its step along D
    const long i0 = (ix - b.x0) + (long)b.nx * ((iy0 - b.y0) + (long)b.ny * (iz0 - b.z0));
    const long di = D == 1 ? (long)b.nx : (long)b.nx * b.ny;
#pragma unroll 1
    for (int iv = 0; iv < NV; iv++) {
      const double *q = p + iv * n + c0;
      double v[6];
#pragma unroll
      for (int o = 0; o < 6; o++) v[o] = q[(o - 3) * st];
#pragma unroll
      for (int j = 0; j < RUN; j++) {
        if (j < m) {
          const long i = i0 + j * di;
          fs[iv * nb + i] = mp5(v[0], v[1], v[2], v[3], v[4]);
          fs[(NV + iv) * nb + i] = mp5(v[5], v[4], v[3], v[2], v[1]);
          if (j + 1 < m) {
#pragma unroll
            for (int o = 0; o < 5; o++) v[o] = v[o + 1];
            v[5] = q[(j + 3) * st];
          }
        }
      }
    }
  }
};

// Fluxes at the face (ix, iy, iz) of direction D from its floored face
// states: physical fluxes, conserved states, HLL or LF with the wave speeds
// of the two cells (f_calc_fluxes_at_faces finite.c:1297-1415)
template <int MAG, int RAD, int D>
KD_INL void face_flux(const koral_params &P, const DGrid &G, const double *__restrict__ ws, double *flb, int ix,
                      int iy, int iz, const double *pL, const double *pR, const Geom &g, bool store = true)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  constexpr int NVMHD = 6 + 3 * MAG;
  double fL[NV], fR[NV], uL[NV], uR[NV];
  flux_prime<MAG, RAD>(P, pL, D + 1, g, fL);
  flux_prime<MAG, RAD>(P, pR, D + 1, g, fR);
  kd::p2u<MAG, RAD>(P, pL, uL, g);
  kd::p2u<MAG, RAD>(P, pR, uR, g);
  const long n = G.ncell;
  const long cR = cidx(G, ix, iy, iz), cL = cR - dir_stride<D>(G);
  const long f = fidx(G, D, ix, iy, iz);
  const long nf = D == 0 ? G.nfx : (D == 1 ? G.nfy : G.nfz);
#pragma unroll
  for (int iv = 0; iv < NV; iv++) {
    const int r = (RAD && iv >= NVMHD) ? 1 : 0;
    double ag = mx(ws[(12 + 3 * r + D) * n + cR], ws[(12 + 3 * r + D) * n + cL]);
    double al = mn(ws[(6 * r + 2 * D) * n + cR], ws[(6 * r + 2 * D) * n + cL]);
    double ar = mx(ws[(6 * r + 2 * D + 1) * n + cR], ws[(6 * r + 2 * D + 1) * n + cL]);
    double F;
    if (P.fluxmethod == 0) {
      F = .5 * (fR[iv] + fL[iv] - ag * (uR[iv] - uL[iv]));
    } else {
      if (al > 0.)
        F = fL[iv];
      else if (ar < 0.)
        F = fL[iv];
      else
        F = (-al * fR[iv] + ar * fL[iv] + al * ar * (uR[iv] - uL[iv])) / (ar - al);
    }
    if (store) flb[iv * nf + f] = F;
  }
}

// face states of face i of a sweep box, stored by OpRecon
template <int NV>
KD_INL void load_face_states(const double *__restrict__ fs, long nb, long i, double *pL, double *pR)
{
#pragma unroll
  for (int iv = 0; iv < NV; iv++) {
    pL[iv] = fs[iv * nb + i];
    pR[iv] = fs[(NV + iv) * nb + i];
  }
}

// Fluxes at the faces of box b from the face states of OpRecon, with the
// face floors (check_floors_mhd, finite.c:801-993) on both states.  A face
// whose states need the b^2/rho floor (mass injection and a Newton
// re-inversion, rare and long) is appended to list (length *cnt, cleared
// before the sweep) and left to OpFluxFloor, so the common path carries no
// inversion and those few faces run in a handful of waves.
template <int MAG, int RAD, int D>
struct OpFlux {
  static constexpr int kind = K_FLUX;
  // 3 waves per SIMD for MHD (168 VGPRs, 10-18 spilled): 2.5 % faster steps
  // than 2 waves with no spill
  static constexpr int waves_per_simd = RAD ? 2 : 3;
  koral_params P;
  DGrid G;
  const double *fs, *ws;
  double *flb;
  unsigned *list, *cnt;
  Box b;
  KD_HD long size() const { return b.size(); }
  KD_INL void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    double pL[NV], pR[NV];
    load_face_states<NV>(fs, b.size(), i, pL, pR);
    Geom g;
    geom_face(G, D, ix, iy, g);
    const int dl = floors_mhd<MAG, RAD, true>(P, pL, g);
    const int dr = floors_mhd<MAG, RAD, true>(P, pR, g);
    const bool defer = dl == FLOORS_DEFERRED || dr == FLOORS_DEFERRED;
    append_index(defer, (unsigned)i, list, cnt);
    if (defer) return;
    face_flux<MAG, RAD, D>(P, G, ws, flb, ix, iy, iz, pL, pR, g);
  }
};

// The face floors of both states of a face.  WAVE: the two halves of the
// wave floor the left and the right state at once (each half a 32-lane
// group of the wave-parallel W inversion), then every lane takes both
// results by shuffle; the two states are floored independently, so the
// results are the sequential ones.
template <int MAG, int RAD, bool WAVE>
KD_INL void floors_pair(const koral_params &P, double *pL, double *pR, const Geom &g)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
#if KD_WAVE_SOLVER
  if constexpr (WAVE) {
    const int lane = threadIdx.x & 63;
    const bool right = lane >= 32;
    double ps[NV];
#pragma unroll
    for (int iv = 0; iv < NV; iv++) ps[iv] = right ? pR[iv] : pL[iv];
    floors_mhd<MAG, RAD, false, true, 32>(P, ps, g);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) {
      pL[iv] = __shfl(ps[iv], lane & 31);
      pR[iv] = __shfl(ps[iv], 32 + (lane & 31));
    }
    return;
  }
#endif
  floors_mhd<MAG, RAD, false, WAVE>(P, pL, g);
  floors_mhd<MAG, RAD, false, WAVE>(P, pR, g);
}

// The faces OpFlux left: the same face states, the full face floors, the
// same fluxes.  WAVE: a group of WAVE_GROUP lanes per face (the floors' W
// inversions evaluate their trials one per lane; the group's first lane
// stores); otherwise one lane per face.  nw groups (WAVE) or lanes stride
// over the list, whose length is known only on the device.
template <int MAG, int RAD, int D, bool WAVE = true>
struct OpFluxFloor {
  static constexpr int kind = K_FLUX_FLOOR;
  static constexpr int waves_per_simd = 2;
  koral_params P;
  DGrid G;
  const double *fs, *ws;
  double *flb;
  const unsigned *list, *cnt;
  Box b;
  long nw;
  KD_HD long size() const { return WAVE ? (long)WAVE_GROUP * nw : nw; }
  KD_INL void operator()(long t) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const long m = (long)*cnt;
    const bool lead = WAVE ? t % WAVE_GROUP == 0 : true;
    for (long j = WAVE ? t / WAVE_GROUP : t; j < m; j += nw) {
      const long i = list[j];
      int ix, iy, iz;
      b.at(i, ix, iy, iz);
      double pL[NV], pR[NV];
      load_face_states<NV>(fs, b.size(), i, pL, pR);
      Geom g;
      geom_face(G, D, ix, iy, g);
      floors_pair<MAG, RAD, WAVE>(P, pL, pR, g);
      face_flux<MAG, RAD, D>(P, G, ws, flb, ix, iy, iz, pL, pR, g, lead);
    }
  }
};

// Both face states of the face (ix, iy, iz) of direction D reconstructed from
// the cell values, variable by variable: the arithmetic of OpRecon, so the
// states are the same bit for bit.
template <int MAG, int RAD, int ORDER, int D>
KD_INL void face_states(const koral_params &P, const DGrid &G, const double *__restrict__ p, int ix, int iy, int iz,
                        double *pL, double *pR)
{
  constexpr int NV = Lay<MAG, RAD>::NV;
  const long n = G.ncell, c0 = cidx(G, ix, iy, iz), st = dir_stride<D>(G);
  double dx[6];
  const int id = dir_index<D>(ix, iy, iz);
#pragma unroll
  for (int o = 0; o < 6; o++) dx[o] = (ORDER == 2) ? size_x(G, id + o - 3, D) : 0.;
#pragma unroll
  for (int iv = 0; iv < NV; iv++) {
    recon_var<ORDER>(P, p + iv * n + c0, st, dx, pL[iv], pR[iv]);
    // one variable's stencil in flight at a time
    compiler_fence();
  }
}

// One face sweep of direction D in a single kernel: reconstruction from the
// cells (finite.c:694-1014), face floors (check_floors_mhd, finite.c:801-993),
// fluxes and HLL/LF (f_calc_fluxes_at_faces finite.c:1297-1415).  The face
// states never leave the registers.  A face whose states need the b^2/rho
// floor is appended to list (cleared before the sweep) and left to
// OpSweepFloor, which reconstructs it again.  Same results as OpRecon +
// OpFlux bit for bit.
template <int MAG, int RAD, int ORDER, int D>
struct OpSweep {
  static constexpr int kind = K_FLUX;
  static constexpr int waves_per_simd = RAD ? 2 : 3;  // as OpFlux
  koral_params P;
  DGrid G;
  const double *p, *ws;
  double *flb;
  unsigned *list, *cnt;
  Box b;
  KD_HD long size() const { return b.size(); }
  KD_INL void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    double pL[NV], pR[NV];
    face_states<MAG, RAD, ORDER, D>(P, G, p, ix, iy, iz, pL, pR);
    Geom g;
    geom_face(G, D, ix, iy, g);
    const int dl = floors_mhd<MAG, RAD, true>(P, pL, g);
    const int dr = floors_mhd<MAG, RAD, true>(P, pR, g);
    const bool defer = dl == FLOORS_DEFERRED || dr == FLOORS_DEFERRED;
    append_index(defer, (unsigned)i, list, cnt);
    if (defer) return;
    face_flux<MAG, RAD, D>(P, G, ws, flb, ix, iy, iz, pL, pR, g);
  }
};

// The faces OpSweep left, reconstructed again from the cells, with the full
// face floors; WAVE and nw as in OpFluxFloor.
template <int MAG, int RAD, int ORDER, int D, bool WAVE = true>
struct OpSweepFloor {
  static constexpr int kind = K_FLUX_FLOOR;
  static constexpr int waves_per_simd = 2;
  koral_params P;
  DGrid G;
  const double *p, *ws;
  double *flb;
  const unsigned *list, *cnt;
  Box b;
  long nw;
  KD_HD long size() const { return WAVE ? (long)WAVE_GROUP * nw : nw; }
  KD_INL void operator()(long t) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const long m = (long)*cnt;
    const bool lead = WAVE ? t % WAVE_GROUP == 0 : true;
    for (long j = WAVE ? t / WAVE_GROUP : t; j < m; j += nw) {
      const long i = list[j];
      int ix, iy, iz;
      b.at(i, ix, iy, iz);
      double pL[NV], pR[NV];
      face_states<MAG, RAD, ORDER, D>(P, G, p, ix, iy, iz, pL, pR);
      Geom g;
      geom_face(G, D, ix, iy, g);
      floors_pair<MAG, RAD, WAVE>(P, pL, pR, g);
      face_flux<MAG, RAD, D>(P, G, ws, flb, ix, iy, iz, pL, pR, g, lead);
    }
  }
};

// n (<= 64) counters set to zero by a kernel (safe inside a captured graph)
struct OpZeroU32 {
  static constexpr int kind = K_ZERO;
  unsigned *q;
  int n;
  KD_HD long size() const { return n; }
  KD_INL void operator()(long i) const { q[i] = 0u; }
};

// ---------------------------------------------------------------------------
// flux-CT (magn.c:199-339): corner EMFs, then the B fluxes of every face
// ---------------------------------------------------------------------------
struct Flx {
  double *x, *y, *z;
  long nfx, nfy, nfz;
};

struct OpEmf {
  static constexpr int kind = K_EMF;
  DGrid G;
  Flx F;
  double *emf;
  KD_HD long size() const { return (long)(G.nx + 1) * (G.ny + 1) * (G.nz + 1); }
  KD_INL void operator()(long t) const
  {
    const int NX = G.nx, NY = G.ny, NZ = G.nz;
    const long ne = size();
    int ix = (int)(t % (NX + 1));
    int iy = (int)((t / (NX + 1)) % (NY + 1));
    int iz = (int)(t / ((long)(NX + 1) * (NY + 1)));
    const double c1 = (NY > 1 && NZ > 1) ? 0.25 : 0.5;
    const double c2 = (NX > 1 && NZ > 1) ? 0.25 : 0.5;
    const double c3 = (NX > 1 && NY > 1) ? 0.25 : 0.5;
    auto FX = [&](int iv, int a, int b, int c) { return F.x[iv * F.nfx + fxidx(G, a, NY > 1 ? b : 0, NZ > 1 ? c : 0)]; };
    auto FY = [&](int iv, int a, int b, int c) { return F.y[iv * F.nfy + fyidx(G, NX > 1 ? a : 0, b, NZ > 1 ? c : 0)]; };
    auto FZ = [&](int iv, int a, int b, int c) { return F.z[iv * F.nfz + fzidx(G, NX > 1 ? a : 0, NY > 1 ? b : 0, c)]; };
    double e1 = 0., e2 = 0., e3 = 0.;
    if (NY > 1 || NZ > 1) {
      double s = 0.;
      if (NY > 1) s = s + FY(8, ix, iy, iz) + FY(8, ix, iy, iz - 1);
      if (NZ > 1) s = s - FZ(7, ix, iy, iz) - FZ(7, ix, iy - 1, iz);
      e1 = c1 * (s);
    }
    if (NX > 1 || NZ > 1) {
      double s = 0.;
      if (NZ > 1) s = s + FZ(6, ix, iy, iz) + FZ(6, ix - 1, iy, iz);
      if (NX > 1) s = s - FX(8, ix, iy, iz) - FX(8, ix, iy, iz - 1);
      e2 = c2 * (s);
    }
    if (NX > 1 || NY > 1) {
      double s = 0.;
      if (NX > 1) s = s + FX(7, ix, iy, iz) + FX(7, ix, iy - 1, iz);
      if (NY > 1) s = s - FY(6, ix, iy, iz) - FY(6, ix - 1, iy, iz);
      e3 = c3 * (s);
    }
    emf[0 * ne + t] = e1;
    emf[1 * ne + t] = e2;
    emf[2 * ne + t] = e3;
  }
};

struct OpCtFaces {
  static constexpr int kind = K_CT;
  DGrid G;
  Flx F;
  const double *emf;
  KD_HD long size() const { return (long)(G.nx + 1) * (G.ny + 1) * (G.nz + 1); }
  KD_INL void operator()(long t) const
  {
    const int NX = G.nx, NY = G.ny, NZ = G.nz;
    const long ne = size();
    int ix = (int)(t % (NX + 1));
    int iy = (int)((t / (NX + 1)) % (NY + 1));
    int iz = (int)(t / ((long)(NX + 1) * (NY + 1)));
    auto E = [&](int k, int a, int b, int c) {
      return emf[(k - 1) * ne + a + (long)b * (NX + 1) + (long)c * (NY + 1) * (NX + 1)];
    };
    if (NX > 1 && iy < NY && iz < NZ) {
      long f = fxidx(G, ix, iy, iz);
      F.x[6 * F.nfx + f] = 0.;
      F.x[7 * F.nfx + f] = 0.5 * (E(3, ix, iy, iz) + E(3, ix, iy + 1, iz));
      F.x[8 * F.nfx + f] = -0.5 * (E(2, ix, iy, iz) + E(2, ix, iy, iz + 1));
    }
    if (NY > 1 && ix < NX && iz < NZ) {
      long f = fyidx(G, ix, iy, iz);
      F.y[6 * F.nfy + f] = -0.5 * (E(3, ix, iy, iz) + E(3, ix + 1, iy, iz));
      F.y[7 * F.nfy + f] = 0.;
      F.y[8 * F.nfy + f] = 0.5 * (E(1, ix, iy, iz) + E(1, ix, iy, iz + 1));
    }
    if (NZ > 1 && ix < NX && iy < NY) {
      long f = fzidx(G, ix, iy, iz);
      F.z[6 * F.nfz + f] = 0.5 * (E(2, ix, iy, iz) + E(2, ix + 1, iy, iz));
      F.z[7 * F.nfz + f] = -0.5 * (E(1, ix, iy, iz) + E(1, ix, iy + 1, iz));
      F.z[8 * F.nfz + f] = 0.;
    }
  }
};

// A coefficient of the step algebra as the expression of ko.c:444-570 it
// stands for, evaluated on the device from the clock's dt with the same
// operations the host would use: c, c / (dt * gamma), c / dt, dt * c, dt / c.
struct Coef {
  enum Kind { CONST, C_OVER_DTG, C_OVER_DT, DT_TIMES_C, DT_OVER_C } kind;
  double c;
  KD_HD double eval(double dt, double gamma) const
  {
    switch (kind) {
    case C_OVER_DTG: return c / (dt * gamma);
    case C_OVER_DT: return c / dt;
    case DT_TIMES_C: return dt * c;
    case DT_OVER_C: return dt / c;
    default: return c;
    }
  }
};

// RK2IMEX stage algebra (ko.c:444-570): up to four statements out = a x
// [+ b y [+ c z]], evaluated left to right, per variable of one interior
// cell, each term an array or the result of an earlier statement of the
// same launch.  A statement's arithmetic does not depend on what it is
// fused with, so fusing gives the same bits as separate launches.
constexpr int COMB_MAXST = 4;
struct CombTerm {
  Coef c;
  const double *x;  // the array, or null for the result of statement prev
  int prev;
};
struct CombStmt {
  double *out;  // null: the result only feeds later statements
  int nt;
  CombTerm t[3];
};
struct OpCombSeq {
  static constexpr int kind = K_COMB;
  DGrid G;
  int nv;
  int ns;
  CombStmt st[COMB_MAXST];
  const double *clk;
  double gamma;
  unsigned *zero = nullptr;  // counters cleared by the first nzero items
  int nzero = 0;
  // one interior cell per item, all its variables
  KD_HD long size() const { return (long)G.nx * G.ny * G.nz; }
  KD_INL void coefs(double coef[COMB_MAXST][3]) const
  {
    const double dt = clk[1];
#pragma unroll
    for (int q = 0; q < COMB_MAXST; q++)
#pragma unroll
      for (int j = 0; j < 3; j++) coef[q][j] = st[q].t[j].c.eval(dt, gamma);
  }
  // the statements for variable entry k
  KD_INL void var(long k, const double coef[COMB_MAXST][3]) const
  {
    double r[COMB_MAXST] = {0., 0., 0., 0.};
#pragma unroll
    for (int q = 0; q < COMB_MAXST; q++) {
      if (q >= ns) break;
      const CombStmt &S = st[q];
      double v[3] = {0., 0., 0.};
#pragma unroll
      for (int j = 0; j < 3; j++)
        if (j < S.nt) {
          const int pq = S.t[j].prev;
          const double *x = S.t[j].x;
          double rp = r[0];
#pragma unroll
          for (int m = 1; m < COMB_MAXST - 1; m++)
            if (pq == m) rp = r[m];
          v[j] = x ? x[k] : rp;
        }
      const double a = coef[q][0], b = coef[q][1], cc = coef[q][2];
      double w;
      if (S.nt == 3)
        w = a * v[0] + b * v[1] + cc * v[2];
      else if (S.nt == 2)
        w = a * v[0] + b * v[1];
      else
        w = a * v[0];
      r[q] = w;
      if (S.out) S.out[k] = w;
    }
  }
  // the statements for the NV variables of interior cell c, statement by
  // statement: every load of a statement is issued before its stores, so a
  // statement costs one memory round trip however many variables and terms
  // it has (the stores may alias later loads, so per variable they would
  // not overlap).  The sum a x + b y + c z is accumulated left to right,
  // the same operations in the same order as var().  No statement refers
  // to the result of the last, so COMB_MAXST - 1 results are kept.
  template <int NV>
  KD_INL void cell_vars(long c, const double coef[COMB_MAXST][3]) const
  {
    const unsigned cb = soa_off(c);
    double r[COMB_MAXST - 1][NV];
#pragma unroll
    for (int q = 0; q < COMB_MAXST; q++) {
      if (q >= ns) break;
      const CombStmt &S = st[q];
      double w[NV];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        if (j >= S.nt) break;
        const double *x = S.t[j].x;
        const int pq = S.t[j].prev;
        const double a = coef[q][j];
        double v[NV];
        if (x) {
#pragma unroll
          for (int iv = 0; iv < NV; iv++) v[iv] = soa_ld(x, iv * G.ncell, cb);
        } else {
#pragma unroll
          for (int iv = 0; iv < NV; iv++) {
            v[iv] = r[0][iv];
#pragma unroll
            for (int m = 1; m < COMB_MAXST - 1; m++)
              if (pq == m) v[iv] = r[m][iv];
          }
        }
#pragma unroll
        for (int iv = 0; iv < NV; iv++) w[iv] = j == 0 ? a * v[iv] : w[iv] + a * v[iv];
      }
      if (S.out)
#pragma unroll
        for (int iv = 0; iv < NV; iv++) soa_st(S.out, iv * G.ncell, cb, w[iv]);
      if (q < COMB_MAXST - 1)
#pragma unroll
        for (int iv = 0; iv < NV; iv++) r[q][iv] = w[iv];
    }
  }
  KD_INL void operator()(long i) const
  {
    if (i < nzero) zero[i] = 0u;
    const int ix = (int)(i % G.nx), iy = (int)((i / G.nx) % G.ny), iz = (int)(i / ((long)G.nx * G.ny));
    const long c = cidx(G, ix, iy, iz);
    double coef[COMB_MAXST][3];
    coefs(coef);
    for (int iv = 0; iv < nv; iv++) var((long)iv * G.ncell + c, coef);
  }
};

// OpCombSeq with the variable count fixed at compile time: each statement's
// loads for all NV variables are in flight together (cell_vars)
template <int NV>
struct OpCombSeqN : OpCombSeq {
  KD_INL void operator()(long i) const
  {
    if (i < nzero) zero[i] = 0u;
    const int ix = (int)(i % G.nx), iy = (int)((i / G.nx) % G.ny), iz = (int)(i / ((long)G.nx * G.ny));
    double coef[COMB_MAXST][3];
    coefs(coef);
    this->template cell_vars<NV>(cidx(G, ix, iy, iz), coef);
  }
};

// ---------------------------------------------------------------------------
// conservative update with metric sources (finite.c:1062-1160)
// ---------------------------------------------------------------------------
// PRE: the cell's own u and p are loaded together with the fluxes, so their
// loads are in flight at once instead of in three dependent rounds (more
// registers: 128 VGPRs for MHD, so still 4 waves per SIMD; 3 with
// radiation); otherwise p is loaded for the metric source and u at the end.
// The arithmetic of every term is the same either way.
template <int MAG, int RAD, bool PRE = false>
struct OpUpdate {
  static constexpr int kind = K_UPDATE;
  static constexpr int waves_per_simd = RAD ? 3 : 4;
  koral_params P;
  DGrid G;
  const double *p;
  double *u;
  Flx F;
  const double *clk;  // device clock {t, dt}
  KD_HD long size() const { return (long)G.nx * G.ny * G.nz; }
  KD_INL void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    const double dt = clk[1];
    Box b{0, 0, 0, G.nx, G.ny, G.nz};
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    long c = cidx(G, ix, iy, iz);
    // flux divergence first, metric source second: the compiler fence keeps
    // the 64 Christoffel and 40 metric loads from being hoisted beside the
    // 6 NV flux loads (253 VGPRs, 2 waves per SIMD otherwise)
    double du[NV], u0[NV], pp[NV];
    if (PRE) {
      load_p<MAG, RAD>(u, G.ncell, c, u0);
      load_p<MAG, RAD>(p, G.ncell, c, pp);
    }
    double dx = size_x(G, ix, 0), dy = size_x(G, iy, 1), dz = size_x(G, iz, 2);
    long fxl = fxidx(G, ix, iy, iz), fxr = fxidx(G, ix + 1, iy, iz);
    long fyl = 0, fyr = 0, fzl = 0, fzr = 0;
    if (G.ny > 1) { fyl = fyidx(G, ix, iy, iz); fyr = fyidx(G, ix, iy + 1, iz); }
    if (G.nz > 1) { fzl = fzidx(G, ix, iy, iz); fzr = fzidx(G, ix, iy, iz + 1); }
    const unsigned bxl = soa_off(fxl), bxr = soa_off(fxr), byl = soa_off(fyl), byr = soa_off(fyr),
                   bzl = soa_off(fzl), bzr = soa_off(fzr);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) {
      du[iv] = -(soa_ld(F.x, iv * F.nfx, bxr) - soa_ld(F.x, iv * F.nfx, bxl)) * dt / dx;
      if (G.ny > 1) du[iv] = du[iv] - (soa_ld(F.y, iv * F.nfy, byr) - soa_ld(F.y, iv * F.nfy, byl)) * dt / dy;
      if (G.nz > 1) du[iv] = du[iv] - (soa_ld(F.z, iv * F.nfz, bzr) - soa_ld(F.z, iv * F.nfz, bzl)) * dt / dz;
    }
    compiler_fence();
    if (!PRE) load_p<MAG, RAD>(p, G.ncell, c, pp);
    double ms[NV];
    Geom g;
    geom_cell(G, ix, iy, g);
    metric_source<MAG, RAD>(P, G, pp, ix, iy, g, ms);
    {
      const unsigned cb = soa_off(c);
#pragma unroll
      for (int iv = 0; iv < NV; iv++) {
        const double uo = PRE ? u0[iv] : soa_ld(u, iv * G.ncell, cb);
        soa_st(u, iv * G.ncell, cb, uo + du[iv] + ms[iv] * dt);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// u2p sweep (calc_u2p finite.c:502-561 / calc_primitives u2p.c:13-133);
// counts {hd fixup requests, rad fixup requests, entropy inversions}
// ---------------------------------------------------------------------------
// calc_primitives per interior cell.  A cell whose hot inversion fails
// (entropy fallback) or whose state needs the b^2/rho floor is appended to
// list (length *cnt, cleared before) and left unwritten for OpU2PFloor; the
// common path then carries one Newton inversion instead of four.
template <int MAG, int RAD>
struct OpU2P {
  static constexpr int kind = K_U2P;
  static constexpr int waves_per_simd = 2;
  koral_params P;
  DGrid G;
  double *u, *p;
  signed char *fl;
  unsigned *list, *cnt;
  KD_HD long size() const { return (long)G.nx * G.ny * G.nz; }
  unsigned char *why;  // per list entry: 1 when the redo may skip the failed hot inversion
  // residual evaluations the hot inversion of the common pass may spend
  // before the cell is left to the deferred pass (0: no bound); a cell with
  // a long run of W bumps or step halvings then runs them lane-parallel
  int budget = 0;
  KD_INL Cnt3 operator()(long i) const { return cell<true, false, false>(i, list, cnt, true); }
  // the deferring calc_primitives, or the full one (optionally from the
  // entropy inversion on) on cell i; WAVE: the calling wave's lanes all run
  // cell i together and only the lane with store writes it
  template <bool DEFER, bool SKIP_HOT, bool WAVE, int WG = WAVE_GROUP>
  KD_INL Cnt3 cell(long i, unsigned *lst, unsigned *len, bool store) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    Box b{0, 0, 0, G.nx, G.ny, G.nz};
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    long c = cidx(G, ix, iy, iz);
    double pp[NV], uu[NV];
    load_p<MAG, RAD>(p, G.ncell, c, pp);
    load_p<MAG, RAD>(u, G.ncell, c, uu);
    Geom g;
    geom_cell(G, ix, iy, g);
    int fe, fe2, fh, fr;
    const int r = calc_primitives<MAG, RAD, DEFER, SKIP_HOT, WAVE, WG>(P, uu, pp, g, fe, fe2, fh, fr, DEFER ? budget : 0);
    if (DEFER) {
      const bool defer = r == CP_DEFERRED || r == CP_DEFERRED_HOT;
      const long slot = append_index(defer, (unsigned)i, lst, len);
      if (defer) {
        why[slot] = r == CP_DEFERRED_HOT ? 1 : 0;
        return Cnt3{{0, 0, 0}};
      }
    }
    if (!store) return Cnt3{{0, 0, 0}};
    store_p<MAG, RAD>(p, G.ncell, c, pp);
    store_p<MAG, RAD>(u, G.ncell, c, uu);
    const long n = G.ncell;
    fl[F_ENTROPY * n + c] = (signed char)fe;
    fl[F_ENTROPY2 * n + c] = (signed char)fe2;
    fl[F_HDFIXUP * n + c] = (signed char)fh;
    fl[F_RADFIXUP * n + c] = (signed char)fr;
    return Cnt3{{fh != 0, fr != 0, fe != 0}};
  }
};

// the cells OpU2P left, with the full calc_primitives.  WAVE: a group of
// WG lanes per cell (64: one cell per wave, 32: two), the W inversions
// evaluating their trials one per lane; otherwise one lane per cell.  nw
// groups (or lanes) stride over the list, whose length is known only on
// the device
template <int MAG, int RAD, bool WAVE = true, int WG = WAVE_GROUP>
struct OpU2PFloor {
  static constexpr int kind = K_U2P_FLOOR;
  static constexpr int waves_per_simd = 2;
  OpU2P<MAG, RAD> op;
  long nw;  // lane groups (WAVE: one cell per group) or lanes (one cell per lane) in flight
  KD_HD long size() const { return WAVE ? (long)WG * nw : nw; }
  KD_INL Cnt3 operator()(long t) const
  {
    Cnt3 sum{{0, 0, 0}};
    const long m = (long)*op.cnt;
    const bool lead = WAVE ? t % WG == 0 : true;
    for (long j = WAVE ? t / WG : t; j < m; j += nw) {
      const long i = (long)op.list[j];
      const Cnt3 c = op.why[j] ? op.template cell<false, true, WAVE, WG>(i, nullptr, nullptr, lead)
                                : op.template cell<false, false, WAVE, WG>(i, nullptr, nullptr, lead);
#pragma unroll
      for (int k = 0; k < 3; k++) sum.v[k] += c.v[k];
    }
    return sum;
  }
};

// ---------------------------------------------------------------------------
// cell_fixup (finite.c:4468-4640): the reference processes flagged cells in
// (ix, iy, iz) order and clears each flag when processed, so a lower
// neighbour always counts as good and an upper neighbour only if it was not
// flagged.  Results go to scratch (pn, un, mark) and are applied by
// OpFixupApply, so every cell reads its neighbours' values from before the
// sweep.
// ---------------------------------------------------------------------------
template <int MAG, int RAD>
struct OpFixup {
  static constexpr int kind = K_FIXUP;
  koral_params P;
  DGrid G;
  const double *p;
  const signed char *fl;
  int type;
  double *pn, *un;
  int *mark;
  KD_HD long size() const { return (long)G.nx * G.ny * G.nz; }
  KD_INL void operator()(long i) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    constexpr int EE = Lay<MAG, RAD>::EE;
    Box b{0, 0, 0, G.nx, G.ny, G.nz};
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    const long n = G.ncell;
    long c = cidx(G, ix, iy, iz);
    mark[c] = 0;
    if (fl[type * n + c] == 0) return;
    const int NX = G.nx, NY = G.ny, NZ = G.nz;
    const int nb[6][3] = {{ix - 1, iy, iz}, {ix + 1, iy, iz}, {ix, iy - 1, iz},
                          {ix, iy + 1, iz}, {ix, iy, iz - 1}, {ix, iy, iz + 1}};
    const bool inb[6] = {ix - 1 >= 0, ix + 1 < NX, iy - 1 >= 0, iy + 1 < NY, iz - 1 >= 0, iz + 1 < NZ};
    long cn[6];
    int in = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      if (!inb[k]) continue;
      long q = cidx(G, nb[k][0], nb[k][1], nb[k][2]);
      bool lower = (k & 1) == 0;
      if (lower || fl[type * n + q] == 0) cn[in++] = q;
    }
    if ((NZ == 1 && NY == 1 && in >= 1) || (NZ == 1 && in >= 2) || (NY == 1 && in >= 2) || in >= 3) {
      double pp[NV], uu[NV];
#pragma unroll
      for (int iv = 0; iv < NV; iv++) {
        bool fix = false;
        if (type == F_HDFIXUP && iv != 0 && iv < 6) fix = true;
        if (type == F_RADFIXUP && iv != 0 && RAD && iv >= EE && iv <= EE + 3) fix = true;
        if (type == F_RADIMPFIXUP && iv != 0 && (iv < 6 || (RAD && iv >= EE && iv <= EE + 3))) fix = true;
        if (fix) {
          double s = 0.;
          for (int q = 0; q < in; q++) s += p[iv * n + cn[q]];
          pp[iv] = s / (double)in;
        } else {
          pp[iv] = p[iv * n + c];
        }
      }
      Geom g;
      geom_cell(G, ix, iy, g);
      kd::p2u<MAG, RAD>(P, pp, uu, g);
#pragma unroll
      for (int iv = 0; iv < 6 + 3 * MAG; iv++) {
        pn[iv * n + c] = pp[iv];
        un[iv * n + c] = uu[iv];
      }
      mark[c] = 1;
    }
  }
};

template <int MAG>
struct OpFixupApply {
  static constexpr int kind = K_FIXUP_APPLY;
  DGrid G;
  double *p, *u;
  signed char *fl;
  int type;
  const double *pn, *un;
  const int *mark;
  KD_HD long size() const { return (long)G.nx * G.ny * G.nz; }
  KD_INL void operator()(long i) const
  {
    Box b{0, 0, 0, G.nx, G.ny, G.nz};
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    const long n = G.ncell;
    long c = cidx(G, ix, iy, iz);
    fl[type * n + c] = 0;
    if (!mark[c]) return;
#pragma unroll
    for (int iv = 0; iv < 6 + 3 * MAG; iv++) {
      p[iv * n + c] = pn[iv * n + c];
      u[iv * n + c] = un[iv * n + c];
    }
  }
};

// ---------------------------------------------------------------------------
// implicit sweep (op_implicit finite.c:1242-1292);
// counts {successes, Newton iterations of the successful attempts, failures}
// ---------------------------------------------------------------------------
template <int MAG>
struct OpImplicit {
  static constexpr int kind = K_IMPLICIT;
  koral_params P;
  DGrid G;
  double *u, *p;
  signed char *fl;
  const double *clk;  // device clock {t, dt}; the solve's time step is dt * fac
  double fac;
  KD_HD long size() const { return (long)G.nx * G.ny * G.nz; }
  KD_INL Cnt3 operator()(long i) const
  {
    constexpr int NV = Lay<MAG, 1>::NV;
    const double t = clk[0], dt = clk[1] * fac;
    Box b{0, 0, 0, G.nx, G.ny, G.nz};
    int ix, iy, iz;
    b.at(i, ix, iy, iz);
    long c = cidx(G, ix, iy, iz);
    double pp[NV], uu[NV];
    load_p<MAG, 1>(p, G.ncell, c, pp);
    load_p<MAG, 1>(u, G.ncell, c, uu);
    Geom g;
    geom_cell(G, ix, iy, g);
    ImpInfo inf;
    int r = solve_implicit_lab<MAG, 1>(P, uu, pp, g, dt, t, inf);
    if (r == 0) {
      store_p<MAG, 1>(p, G.ncell, c, pp);
      store_p<MAG, 1>(u, G.ncell, c, uu);
    }
    const long n = G.ncell;
    fl[F_RADSOURCETYPE * n + c] = 3;
    fl[F_RADFIXUP * n + c] = r == 0 ? 0 : -1;
    fl[F_RADIMPFIXUP * n + c] = r == 0 ? 0 : -1;
    return Cnt3{{r == 0, inf.iters, r != 0}};
  }
};

// ---------------------------------------------------------------------------
// step order, written once for any executor
// ---------------------------------------------------------------------------
// Device-resident state of a grid: SoA primitives/conserved over the
// ghost-inclusive box, int8 flags, and the work arrays of a step.
// Launch choices of the grid operators, the space the autotuner searches
// (koral_hip_grid_autotune).  Every choice computes the same results bit for
// bit; they differ only in speed.
constexpr int TUNE_KINDS = 32;  // koral_grid_tune.block entries (>= K_NKINDS)
static_assert(K_NKINDS <= TUNE_KINDS, "operator kinds");
struct GridTune {
  int block[TUNE_KINDS];  // threads per block of each operator kind: 64, 128 or 256 (0: 256)
  int fused[3];         // sweep of direction d: 1 reconstruction inside the flux kernel
                        // (OpSweep), 0 face states through HBM (OpRecon + OpFlux)
  int floor_wave;       // deferred passes: 1 a group of WAVE_GROUP lanes per face or cell, 0 one lane
  int update_pre;       // update: 1 the cell's u and p loaded with the fluxes
  int comb_nv;          // 1: stage algebra with the variable count a template parameter
  int u2p_budget;       // OpU2P::budget
  int u2p_group;        // lanes per cell of the wave-mapped deferred u2p pass (64 or 32)
  long floor_nw;        // waves (or lanes) in flight in the deferred face passes
  long u2p_nw;          // waves (or lanes) in flight in the deferred u2p pass
};
inline GridTune default_tune()
{
  GridTune t{};
  for (int k = 0; k < TUNE_KINDS; k++) t.block[k] = 256;
  for (int d = 0; d < 3; d++) t.fused[d] = 0;
  t.floor_wave = 1;
  t.update_pre = 0;
  t.comb_nv = 1;
  t.u2p_budget = 32;
  t.u2p_group = 32;
  t.floor_nw = 4096;
  t.u2p_nw = 4096;
  return t;
}

struct GridBufs {
  GridTune tune;
  koral_params P;
  DGrid G;
  double *p, *u;
  signed char *fl;
  double *ws, *flb[3], *emf;
  double *fs[3];      // face states of the sweep of each direction, 2 NV x face_states_cap(G)
  unsigned *flist[3];  // faces of each sweep left to the b^2/rho floor pass, face_states_cap(G)
  unsigned *fcnt;      // their counts [3], then the count of ulist
  unsigned *ulist;     // cells left to the full calc_primitives pass, nx ny nz
  unsigned char *uwhy;  // per ulist entry: the redo may skip the failed hot inversion
  double *pn, *un;
  int *mark;
  double *ut0, *ut1, *uforget, *dut1, *dut2, *drt1, *drt2;
  double *tsd;    // time-step density max (one double)
  double *clock;  // {t, dt} of the operators (device scalars)
  int *counters;  // step counters [6]
  unsigned *imp_list[2][2];  // work lists of the implicit passes (device executor)
  unsigned *imp_ctr;         // their counters
};

// Faces per sweep box bound: a sweep of direction d covers n_d + 1 faces
// along d and at most n_k + 2 across (flux-CT needs one ghost layer)
inline long face_states_cap(const DGrid &G)
{
  auto w = [](int n) { return n == 1 ? 1L : (long)n + 2; };
  return w(G.nx) * w(G.ny) * w(G.nz);
}

// Every device buffer of a grid state with its element count: f(ptr, count,
// name) for each, ptr a reference to the GridBufs member.  The library
// allocates exactly these (koral_hip_grid_create); the host check of the
// operators allocates the same sizes behind guard pages.  Directions that are
// collapsed get no flux, face-state or floor-list buffers.
constexpr int IMP_LIST_COUNTERS = 32;
template <class F>
void grid_buffers(GridBufs &s, F &&f)
{
  const DGrid &G = s.G;
  const long nv = s.P.nv, nc = G.ncell, nd = (long)G.nx * G.ny * G.nz;
  const long nf[3] = {G.nfx, G.nfy, G.nfz};
  f(s.p, nv * nc, "p");
  f(s.u, nv * nc, "u");
  f(s.fl, (long)NFLAGS * nc, "flags");
  f(s.ws, (long)NWS * nc, "wavespeeds");
  for (int d = 0; d < 3; d++)
    if (dimn(G, d) > 1) {
      f(s.flb[d], nv * nf[d], d == 0 ? "flux x" : d == 1 ? "flux y" : "flux z");
      f(s.fs[d], 2 * nv * face_states_cap(G), d == 0 ? "face states x" : d == 1 ? "face states y" : "face states z");
      f(s.flist[d], face_states_cap(G), d == 0 ? "floor list x" : d == 1 ? "floor list y" : "floor list z");
    }
  f(s.fcnt, 4L, "list counts");
  f(s.ulist, nd, "u2p list");
  f(s.uwhy, nd, "u2p list kinds");
  f(s.emf, 3L * (G.nx + 1) * (G.ny + 1) * (G.nz + 1), "emf");
  f(s.pn, nv * nc, "fixup p");
  f(s.un, nv * nc, "fixup u");
  f(s.mark, nc, "fixup marks");
  f(s.ut0, nv * nc, "ut0");
  f(s.ut1, nv * nc, "ut1");
  f(s.uforget, nv * nc, "uforget");
  f(s.dut1, nv * nc, "dut1");
  f(s.dut2, nv * nc, "dut2");
  f(s.drt1, nv * nc, "drt1");
  f(s.drt2, nv * nc, "drt2");
  f(s.tsd, 1L, "tstepden");
  f(s.clock, 4L, "clock");
  f(s.counters, 8L, "counters");
  for (int i = 0; i < 2; i++)
    for (int k = 0; k < 2; k++) f(s.imp_list[i][k], nd, "implicit list");
  f(s.imp_ctr, (long)IMP_LIST_COUNTERS, "implicit counters");
}

// Ghost-inclusive layout of a grid description (a collapsed dimension,
// n = 1, has no ghosts).  Table pointers are left for the caller.
inline void grid_layout(const koral_grid_desc *D, DGrid &G)
{
  G.nx = D->nx; G.ny = D->ny; G.nz = D->nz; G.ng = D->ng;
  G.ngx = D->ng; G.ngy = D->ny > 1 ? D->ng : 0; G.ngz = D->nz > 1 ? D->ng : 0;
  G.sx = G.nx + 2 * G.ngx; G.sy = G.ny + 2 * G.ngy; G.sz = G.nz + 2 * G.ngz;
  G.ncell = (long)G.sx * G.sy * G.sz;
  G.nfx = (long)(G.sx + 1) * G.sy * G.sz;
  G.nfy = (long)G.sx * (G.sy + 1) * G.sz;
  G.nfz = (long)G.sx * G.sy * (G.sz + 1);
  G.npl = (long)G.sy * G.sx;
  G.nplx = (long)G.sy * (G.sx + 1);
  G.nply = (long)(G.sy + 1) * G.sx;
  for (int i = 0; i < 6; i++) G.bc[i] = D->bc[i];
}

template <int M, int R>
struct Tag {
  static constexpr int MAG = M, RAD = R;
};

template <class F>
void dispatch(const koral_params &P, F f)
{
  if (P.mag && P.rad)
    f(Tag<1, 1>{});
  else if (P.mag)
    f(Tag<1, 0>{});
  else if (P.rad)
    f(Tag<0, 1>{});
  else
    f(Tag<0, 0>{});
}

inline Box wavespeed_box(const koral_params &P, const DGrid &G)
{
  int lim = (P.int_order == 4) ? 2 : 1;
  int lx = lim, ly = G.ny > 1 ? lim : 0, lz = G.nz > 1 ? lim : 0;
  return Box{-lx, -ly, -lz, G.nx + 2 * lx, G.ny + 2 * ly, G.nz + 2 * lz};
}

template <class Ex>
void grid_p2u(Ex &ex, const GridBufs &s, bool interior_only = false)
{
  const DGrid &G = s.G;
  Box b{-G.ngx, -G.ngy, -G.ngz, G.sx, G.sy, G.sz};
  if (interior_only) b = Box{0, 0, 0, G.nx, G.ny, G.nz};
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    ex.each(OpP2U<T::MAG, T::RAD>{s.P, G, s.p, s.u, b});
  });
}

template <class Ex>
void set_bc(Ex &ex, const GridBufs &s)
{
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    ex.each(OpBcAll<T::MAG, T::RAD>{OpBcFaces<T::MAG, T::RAD>{s.P, s.G, s.p, s.u}});
  });
}

template <class Ex>
void tstepden(Ex &ex, const GridBufs &s)
{
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    ex.max(OpWavespeeds<T::MAG, T::RAD>{s.P, s.G, s.p, s.ws, wavespeed_box(s.P, s.G)}, s.tsd);
  });
}

// op_explicit: ghosts, wave speeds (tsd accumulates the max), face fluxes,
// flux-CT, update.  ghosts_current: the ghost cells already hold set_bc of
// the current p and u (the ghost fill is a function of the interior alone,
// so a second fill would write the same values) and are not refilled.
template <class Ex>
void explicit_op(Ex &ex, const GridBufs &s, bool ghosts_current = false, bool counts_cleared = false)
{
  const DGrid &G = s.G;
  const koral_params &P = s.P;
  if (!ghosts_current) set_bc(ex, s);
  tstepden(ex, s);
  const int tt = P.mag ? 1 : 0;
  const int n[3] = {G.nx, G.ny, G.nz};
  if (!counts_cleared) ex.each(OpZeroU32{s.fcnt, 3});
  // counts_cleared: the caller's last launch cleared the face-list counts
  // (the stage algebra's first items do); the host check build verifies it
  KD_CHECK(!counts_cleared || (s.fcnt[0] == 0u && s.fcnt[1] == 0u && s.fcnt[2] == 0u),
           "explicit_op: face-list counts not cleared", s.fcnt[0], s.fcnt[1], s.fcnt[2]);
  // x2 first: its floor pass (faces next to the magnetised torus surface,
  // the longest) then overlaps the x1 and x3 sweeps.  The sweeps write
  // separate flux arrays, so their order does not change any result.
  const int order[3] = {1, 0, 2};
  for (int od = 0; od < 3; od++) {
    const int d = order[od];
    if (n[d] == 1) continue;
    int lo[3], cnt[3];
    for (int k = 0; k < 3; k++) {
      if (k == d) { lo[k] = 0; cnt[k] = n[k] + 1; }
      else if (n[k] == 1) { lo[k] = 0; cnt[k] = 1; }
      else { lo[k] = -tt; cnt[k] = n[k] + 2 * tt; }
    }
    Box fb{lo[0], lo[1], lo[2], cnt[0], cnt[1], cnt[2]};
    dispatch(P, [&](auto t) {
      using T = decltype(t);
      constexpr int M = T::MAG, R = T::RAD;
      const GridTune &tn = s.tune;
      // the faces left to the floor pass (a few, each a long Newton
      // inversion) run on a side stream while the next sweeps proceed
      auto floor_pass = [&](int D, auto op_wave, auto op_lane) {
        if (!M) return;
        ex.fork(D);
        if (tn.floor_wave)
          ex.each_on(D, op_wave);
        else
          ex.each_on(D, op_lane);
      };
      auto fused = [&](auto dd, auto oo) {
        constexpr int D = decltype(dd)::value, O = decltype(oo)::value;
        ex.each(OpSweep<M, R, O, D>{P, G, s.p, s.ws, s.flb[D], s.flist[D], s.fcnt + D, fb});
        floor_pass(D, OpSweepFloor<M, R, O, D, true>{P, G, s.p, s.ws, s.flb[D], s.flist[D], s.fcnt + D, fb, tn.floor_nw},
                   OpSweepFloor<M, R, O, D, false>{P, G, s.p, s.ws, s.flb[D], s.flist[D], s.fcnt + D, fb, tn.floor_nw});
      };
      auto sweep = [&](auto dd) {
        constexpr int D = decltype(dd)::value;
        if (tn.fused[D]) {
          if (P.int_order == 4)
            fused(dd, std::integral_constant<int, 4>{});
          else if (P.int_order == 2)
            fused(dd, std::integral_constant<int, 2>{});
          else
            fused(dd, std::integral_constant<int, 1>{});
          return;
        }
        if (P.int_order == 4 && D > 0)
          ex.each(OpReconRun<M, R, (D > 0 ? D : 1)>{P, G, s.p, s.fs[D], fb});
        else if (P.int_order == 4)
          ex.each(OpRecon<M, R, 4, D>{P, G, s.p, s.fs[D], fb});
        else if (P.int_order == 2)
          ex.each(OpRecon<M, R, 2, D>{P, G, s.p, s.fs[D], fb});
        else
          ex.each(OpRecon<M, R, 1, D>{P, G, s.p, s.fs[D], fb});
        ex.each(OpFlux<M, R, D>{P, G, s.fs[D], s.ws, s.flb[D], s.flist[D], s.fcnt + D, fb});
        floor_pass(D, OpFluxFloor<M, R, D, true>{P, G, s.fs[D], s.ws, s.flb[D], s.flist[D], s.fcnt + D, fb, tn.floor_nw},
                   OpFluxFloor<M, R, D, false>{P, G, s.fs[D], s.ws, s.flb[D], s.flist[D], s.fcnt + D, fb, tn.floor_nw});
      };
      if (d == 0)
        sweep(std::integral_constant<int, 0>{});
      else if (d == 1)
        sweep(std::integral_constant<int, 1>{});
      else
        sweep(std::integral_constant<int, 2>{});
    });
  }
  if (P.mag)
    for (int d = 0; d < 3; d++)
      if (n[d] > 1) ex.join(d);
  Flx F{s.flb[0], s.flb[1], s.flb[2], G.nfx, G.nfy, G.nfz};
  if (P.mag) {
    ex.each(OpEmf{G, F, s.emf});
    ex.each(OpCtFaces{G, F, s.emf});
  }
  dispatch(P, [&](auto t) {
    using T = decltype(t);
    if (s.tune.update_pre)
      ex.each(OpUpdate<T::MAG, T::RAD, true>{P, G, s.p, s.u, F, s.clock});
    else
      ex.each(OpUpdate<T::MAG, T::RAD, false>{P, G, s.p, s.u, F, s.clock});
  });
}

template <class Ex>
void fixup(Ex &ex, const GridBufs &s, int type)
{
  const koral_params &P = s.P;
  if (!P.dofixups) return;
  if (type == F_HDFIXUP && !P.dou2pmhdfixups) return;
  if (type == F_RADFIXUP && !P.dou2pradfixups) return;
  if (type == F_RADIMPFIXUP && !P.doradimpfixups) return;
  dispatch(P, [&](auto t) {
    using T = decltype(t);
    ex.each(OpFixup<T::MAG, T::RAD>{P, s.G, s.p, s.fl, type, s.pn, s.un, s.mark});
    ex.each(OpFixupApply<T::MAG>{s.G, s.p, s.u, s.fl, type, s.pn, s.un, s.mark});
  });
}

// calc_u2p with its cell_fixup passes and the ghost refresh
template <class Ex>
void u2p(Ex &ex, const GridBufs &s, bool count_cleared = false)
{
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    if (!count_cleared) ex.each(OpZeroU32{s.fcnt + 3, 1});
    KD_CHECK(!count_cleared || s.fcnt[3] == 0u, "u2p: cell-list count not cleared", s.fcnt[3], 0, 0);
    const OpU2P<T::MAG, T::RAD> op{s.P, s.G, s.u, s.p, s.fl, s.ulist, s.fcnt + 3, s.uwhy, s.tune.u2p_budget};
    ex.count(op, s.counters + 3);
    if (s.tune.floor_wave)
      if (s.tune.u2p_group == 32)
        ex.count(OpU2PFloor<T::MAG, T::RAD, true, 32>{op, s.tune.u2p_nw}, s.counters + 3);
      else
        ex.count(OpU2PFloor<T::MAG, T::RAD, true, 64>{op, s.tune.u2p_nw}, s.counters + 3);
    else
      ex.count(OpU2PFloor<T::MAG, T::RAD, false>{op, s.tune.u2p_nw}, s.counters + 3);
  });
  fixup(ex, s, F_HDFIXUP);
  if (s.P.rad) fixup(ex, s, F_RADFIXUP);
  set_bc(ex, s);
}

// whether do_correct changes anything (ko.c calls it after every calc_u2p)
inline bool polar_correction_active(const GridBufs &s) { return s.P.polar_nc > 0 && s.G.ny >= 2; }

// do_correct (finite.c:605-636) with CORRECT_POLARAXIS: correct_polaraxis
// for spherical-like coordinates (finite.c:4924-5060).  One item per
// (ix, iz, face, ic): the cell ic away from the polar face takes the gas and
// radiation densities and the x1/x3 velocities of the cell nc away, the x2
// velocities scaled by |x2 - x2_axis| / |x2_src - x2_axis| (cell-centre code
// coordinates, set_grid finite.c:1630-1660); p and u = p2u(p) are written for
// every variable but the field.  The items are independent when
// NY >= 2 nc + 2 (checked at grid creation).
template <int MAG, int RAD>
struct OpPolar {
  static constexpr int kind = K_POLAR;
  koral_params P;
  DGrid G;
  double *p, *u;
  KD_HD long size() const { return 2L * P.polar_nc * G.nx * G.nz; }
  KD_INL void operator()(long t) const
  {
    constexpr int NV = Lay<MAG, RAD>::NV;
    constexpr int EE = Lay<MAG, RAD>::EE;
    const int nc = P.polar_nc;
    const int ix = (int)(t % G.nx);
    long r = t / G.nx;
    const int iz = (int)(r % G.nz);
    r /= G.nz;
    const int side = (int)(r % 2), ic = (int)(r / 2);
    const int iy = side == 0 ? ic : G.ny - 1 - ic;
    const int iysrc = side == 0 ? nc : G.ny - 1 - nc;
    const double *xb = G.xb1 + G.ng;
    const double thaxis = side == 0 ? xb[0] : xb[G.ny];
    const double th = .5 * (xb[iy + 1] + xb[iy]), thsrc = .5 * (xb[iysrc + 1] + xb[iysrc]);
    const double fac = fabs((th - thaxis) / (thsrc - thaxis));
    const long n = G.ncell, c = cidx(G, ix, iy, iz), cs = cidx(G, ix, iysrc, iz);
    double pp[NV], uu[NV];
    load_p<MAG, RAD>(p, n, c, pp);
    pp[0] = p[0 * n + cs];
    pp[1] = p[1 * n + cs];
    pp[5] = p[5 * n + cs];
    pp[2] = p[2 * n + cs];
    pp[4] = p[4 * n + cs];
    pp[3] = fac * p[3 * n + cs];
    if (RAD) {
      pp[EE] = p[EE * n + cs];
      pp[EE + 1] = p[(EE + 1) * n + cs];
      pp[EE + 3] = p[(EE + 3) * n + cs];
      pp[EE + 2] = fac * p[(EE + 2) * n + cs];
    }
    Geom g;
    geom_cell(G, ix, iy, g);
    kd::p2u<MAG, RAD>(P, pp, uu, g);
#pragma unroll
    for (int iv = 0; iv < NV; iv++) {
      if (MAG && iv >= 6 && iv <= 8) continue;
      p[iv * n + c] = pp[iv];
      u[iv * n + c] = uu[iv];
    }
  }
};

template <class Ex>
void do_correct(Ex &ex, const GridBufs &s)
{
  if (!polar_correction_active(s)) return;
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    ex.each(OpPolar<T::MAG, T::RAD>{s.P, s.G, s.p, s.u});
  });
}

template <class Ex>
void implicit_op(Ex &ex, const GridBufs &s, double fac)
{
  if (!s.P.rad) return;
  if (s.P.mag)
    ex.implicit_cells(OpImplicit<1>{s.P, s.G, s.u, s.p, s.fl, s.clock, fac}, s.counters);
  else
    ex.implicit_cells(OpImplicit<0>{s.P, s.G, s.u, s.p, s.fl, s.clock, fac}, s.counters);
  fixup(ex, s, F_RADIMPFIXUP);
}

// launches a stage-algebra op (compile-time variable count when tune.comb_nv)
template <class Ex>
void each_comb(Ex &ex, const GridBufs &s, const OpCombSeq &op)
{
  if (!s.tune.comb_nv) {
    ex.each(op);
    return;
  }
  dispatch(s.P, [&](auto t) {
    using T = decltype(t);
    ex.each(OpCombSeqN<Lay<T::MAG, T::RAD>::NV>{op});
  });
}

template <class Ex>
void comb(Ex &ex, const GridBufs &s, double gamma, double *out, Coef a, const double *x,
          Coef b = {Coef::CONST, 0.}, const double *y = nullptr, Coef c = {Coef::CONST, 0.},
          const double *z = nullptr)
{
  const int nt = z ? 3 : (y ? 2 : 1);
  OpCombSeq op{s.G, s.P.nv, 1, {}, s.clock, gamma};
  op.st[0] = CombStmt{out, nt, {CombTerm{a, x, 0}, CombTerm{b, y, 0}, CombTerm{c, z, 0}}};
  each_comb(ex, s, op);
}

// body of one RK2IMEX step, ko.c:444-570, at the clock's (t, dt); the
// coefficients are the reference's expressions of dt
template <class Ex>
void step_rk2imex(Ex &ex, const GridBufs &s)
{
  const double gamma = 1. - 1. / sqrt(2.);
  const Coef one{Coef::CONST, 1.}, idtg{Coef::C_OVER_DTG, 1.}, midtg{Coef::C_OVER_DTG, -1.};
  const Coef idt{Coef::C_OVER_DT, 1.}, midt{Coef::C_OVER_DT, -1.}, dt1{Coef::DT_TIMES_C, 1.};
  const Coef dtc{Coef::DT_TIMES_C, 1. - 2. * gamma}, dth{Coef::DT_OVER_C, 2.};
  // fused combinations: statements of one launch run in the order listed,
  // R(c, q) is the result of statement q of the same launch
  auto A = [](Coef c, const double *x) { return CombTerm{c, x, 0}; };
  auto R = [](Coef c, int q) { return CombTerm{c, nullptr, q}; };
  const CombTerm none{{Coef::CONST, 0.}, nullptr, 0};
  // each seq() precedes a u2p and also clears the four list counters of the
  // u2p and of the explicit operator that follows it (no reader of them runs
  // in between: the floor passes of the previous sweeps were joined)
  // the statements of one launch (at most COMB_MAXST)
  auto make = [&](std::initializer_list<CombStmt> l) {
    OpCombSeq op{s.G, s.P.nv, (int)l.size(), {}, s.clock, gamma};
    int q = 0;
    for (const CombStmt &st : l) op.st[q++] = st;
    if (op.size() >= 4) {
      op.zero = s.fcnt;
      op.nzero = 4;
    }
    return op;
  };
  auto seq = [&](std::initializer_list<CombStmt> l) {
    const OpCombSeq op = make(l);
    if (op.nzero == 0) ex.each(OpZeroU32{s.fcnt, 4});
    each_comb(ex, s, op);
  };
  auto explicit_then = [&](std::initializer_list<CombStmt> l) {
    // u2p ended with set_bc; without a polar correction p and u are unchanged
    explicit_op(ex, s, !polar_correction_active(s), true);
    seq(l);
  };
  // Without radiation implicit_op is empty, so u does not change between
  // the first copies: ut1 = u = ut0 (1 u, the same bits), and drt1 =
  // (u - ut0) / (gamma dt) = idtg ut0 + midtg ut0, which the launches that
  // read drt1 recompute from ut0 with the same operations instead of
  // storing it.  Neither ut1 (first stage) nor drt1 is then stored or read.
  const bool norad = !s.P.rad;
  if (norad) {
    seq({CombStmt{s.ut0, 1, {A(one, s.u), none, none}}});
  } else {
    comb(ex, s, gamma, s.ut0, one, s.u);
    implicit_op(ex, s, gamma);
    // drt1 = (u - ut0) / (gamma dt); ut1 = u
    seq({CombStmt{s.drt1, 2, {A(idtg, s.u), A(midtg, s.ut0), none}}, CombStmt{s.ut1, 1, {A(one, s.u), none, none}}});
  }
  u2p(ex, s, true);
  do_correct(ex, s);
  // explicit; dut1 = (u - ut1) / dt; u = ut0 + dt dut1 + (1 - 2 gamma) dt drt1; uforget = u
  if (norad)
    explicit_then({CombStmt{nullptr, 2, {A(idtg, s.ut0), A(midtg, s.ut0), none}},
                   CombStmt{s.dut1, 2, {A(idt, s.u), A(midt, s.ut0), none}},
                   CombStmt{s.u, 3, {A(one, s.ut0), R(dt1, 1), R(dtc, 0)}},
                   CombStmt{s.uforget, 1, {R(one, 2), none, none}}});
  else
    explicit_then({CombStmt{s.dut1, 2, {A(idt, s.u), A(midt, s.ut1), none}},
                   CombStmt{s.u, 3, {A(one, s.ut0), R(dt1, 0), A(dtc, s.drt1)}},
                   CombStmt{s.uforget, 1, {R(one, 1), none, none}}});
  u2p(ex, s, true);
  do_correct(ex, s);
  implicit_op(ex, s, gamma);
  // drt2 = (u - uforget) / (gamma dt); ut2 = u
  seq({CombStmt{s.drt2, 2, {A(idtg, s.u), A(midtg, s.uforget), none}},
       CombStmt{s.ut1, 1, {A(one, s.u), none, none}}});
  u2p(ex, s, true);
  do_correct(ex, s);
  // explicit; dut2 = (u - ut2) / dt; u = ut0 + dt/2 (dut1 + dut2); u = u + dt/2 (drt1 + drt2)
  if (norad)
    explicit_then({CombStmt{nullptr, 2, {A(idtg, s.ut0), A(midtg, s.ut0), none}},
                   CombStmt{s.dut2, 2, {A(idt, s.u), A(midt, s.ut1), none}},
                   CombStmt{nullptr, 3, {A(one, s.ut0), A(dth, s.dut1), R(dth, 1)}},
                   CombStmt{s.u, 3, {R(one, 2), R(dth, 0), A(dth, s.drt2)}}});
  else
    explicit_then({CombStmt{s.dut2, 2, {A(idt, s.u), A(midt, s.ut1), none}},
                   CombStmt{nullptr, 3, {A(one, s.ut0), A(dth, s.dut1), R(dth, 0)}},
                   CombStmt{s.u, 3, {R(one, 1), A(dth, s.drt1), A(dth, s.drt2)}}});
  u2p(ex, s, true);
  do_correct(ex, s);
}

}  // namespace kd
