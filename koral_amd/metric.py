"""Tabulated metric (host side, computed once), restating metric.c.

KORAL precomputes, for every cell centre and face, the 20 numbers of
``struct geometry`` (ko.h:336-353, calc_metric metric.c:3866-4080):

  g20[i*5+j] = g_{ij} (i,j<4),  g20[i*5+4] = d ln sqrt(-g) / dx^{i+1} (i<3),
  g20[3*5+4] = sqrt(-g)
  G20[i*5+j] = g^{ij},          G20[3*5+4] = g_tt + 1 (calc_gttpert, metric.c)

plus the Christoffel symbols Gamma^i_{jk} at cell centres (64 numbers,
gKr[i*16+j*4+k]) with the MODYFIKUJKRZYSIE correction (metric.c:828-905).

Coordinates supported: MINKCOORDS (Cartesian), KSCOORDS and MKS1COORDS
(r = exp(x1) + R0, theta = x2, phi = x3; metric.c coco_MKS12KS).  Metric
derivatives are taken with the complex-step method, exact to rounding for
these analytic metrics.  All supported metrics are independent of x3, so grid
tables are stored per (x1, x2) only -- a layout choice (KORAL's
METRICAXISYMMETRIC, ko.h:263-270) that leaves every value unchanged.
"""
import numpy as np

MINK, KS, MKS1 = 4, 2, 7  # mnemonics.h coordinate ids


def _ks_g(r, th, a):
    """Kerr-Schild g_{mu nu} (t, r, th, ph) as a (..., 4, 4) array."""
    s = np.sin(th)
    c = np.cos(th)
    Sig = r * r + a * a * c * c
    z = 2. * r / Sig
    shape = np.broadcast(r, th).shape
    g = np.zeros(shape + (4, 4), dtype=np.result_type(r, th, np.float64))
    g[..., 0, 0] = -(1. - z)
    g[..., 0, 1] = g[..., 1, 0] = z
    g[..., 0, 3] = g[..., 3, 0] = -z * a * s * s
    g[..., 1, 1] = 1. + z
    g[..., 1, 3] = g[..., 3, 1] = -a * s * s * (1. + z)
    g[..., 2, 2] = Sig
    g[..., 3, 3] = s * s * (Sig + a * a * s * s * (1. + z))
    return g


def _ks_G(r, th, a):
    s = np.sin(th)
    c = np.cos(th)
    Sig = r * r + a * a * c * c
    Delta = r * r - 2. * r + a * a
    z = 2. * r / Sig
    shape = np.broadcast(r, th).shape
    G = np.zeros(shape + (4, 4))
    G[..., 0, 0] = -(1. + z)
    G[..., 0, 1] = G[..., 1, 0] = z
    G[..., 1, 1] = Delta / Sig
    G[..., 1, 3] = G[..., 3, 1] = a / Sig
    G[..., 2, 2] = 1. / Sig
    G[..., 3, 3] = 1. / (Sig * s * s)
    return G


class Coords:
    """A coordinate system: code coordinates x -> metric tensors."""

    def __init__(self, kind, spin=0.0, r0=0.0):
        self.kind, self.a, self.r0 = kind, float(spin), float(r0)

    def g(self, x1, x2):
        """Covariant metric in code coordinates; accepts complex arguments."""
        if self.kind == MINK:
            sh = np.broadcast(x1, x2).shape
            g = np.zeros(sh + (4, 4), dtype=np.result_type(x1, x2, np.float64))
            g[..., 0, 0] = -1.
            g[..., 1, 1] = g[..., 2, 2] = g[..., 3, 3] = 1.
            return g
        if self.kind == KS:
            return _ks_g(x1, x2, self.a)
        r = np.exp(x1) + self.r0
        drdx = np.exp(x1)
        g = _ks_g(r, x2, self.a)
        J = np.ones(g.shape[:-2] + (4,), dtype=g.dtype)
        J[..., 1] = drdx
        return g * J[..., :, None] * J[..., None, :]

    def G(self, x1, x2):
        if self.kind == MINK:
            return np.real(self.g(x1, x2))
        if self.kind == KS:
            return _ks_G(x1, x2, self.a)
        r = np.exp(x1) + self.r0
        drdx = np.exp(x1)
        G = _ks_G(r, x2, self.a)
        J = np.ones(G.shape[:-2] + (4,))
        J[..., 1] = 1. / drdx
        return G * J[..., :, None] * J[..., None, :]

    def gdet(self, x1, x2):
        if self.kind == MINK:
            return np.ones(np.broadcast(x1, x2).shape, dtype=np.result_type(x1, x2, np.float64))
        if self.kind == KS:
            r, th = x1, x2
            return (r * r + self.a * self.a * np.cos(th) ** 2) * np.sin(th)
        r = np.exp(x1) + self.r0
        th = x2
        return (r * r + self.a * self.a * np.cos(th) ** 2) * np.sin(th) * np.exp(x1)

    def gttpert(self, x1, x2):
        """g_tt + 1 without cancellation (calc_gttpert_arb)."""
        if self.kind == MINK:
            return np.zeros(np.broadcast(x1, x2).shape)
        r = x1 if self.kind == KS else np.exp(x1) + self.r0
        return 2. * r / (r * r + self.a * self.a * np.cos(x2) ** 2)

    def to_rth(self, x1, x2):
        """Boyer-Lindquist/Kerr-Schild radius and polar angle of code coords."""
        if self.kind == KS:
            return x1, x2
        if self.kind == MKS1:
            return np.exp(x1) + self.r0, x2
        raise ValueError("no spherical radius for Minkowski")


_H = 1.e-30


def _dg(coords, x1, x2):
    """d g_{mu nu} / dx^k for k=0..3 as (..., 4(k), 4, 4) by complex step."""
    x1 = np.asarray(x1, dtype=np.float64)
    x2 = np.asarray(x2, dtype=np.float64)
    sh = np.broadcast(x1, x2).shape
    d = np.zeros(sh + (4, 4, 4))
    d[..., 1, :, :] = np.imag(coords.g(x1 + 1j * _H, x2 + 0j)) / _H
    d[..., 2, :, :] = np.imag(coords.g(x1 + 0j, x2 + 1j * _H)) / _H
    return d


def _dlgdet(coords, x1, x2):
    x1 = np.asarray(x1, dtype=np.float64)
    x2 = np.asarray(x2, dtype=np.float64)
    gd = coords.gdet(x1, x2)
    out = np.zeros(np.broadcast(x1, x2).shape + (3,))
    if coords.kind == MINK:
        return out
    out[..., 0] = np.imag(coords.gdet(x1 + 1j * _H, x2 + 0j)) / _H / gd
    out[..., 1] = np.imag(coords.gdet(x1 + 0j, x2 + 1j * _H)) / _H / gd
    return out


def point_tables(coords, x1, x2):
    """(g20, G20) for arbitrary points: arrays of shape (..., 20)."""
    x1 = np.asarray(x1, dtype=np.float64)
    x2 = np.asarray(x2, dtype=np.float64)
    sh = np.broadcast(x1, x2).shape
    g = np.real(coords.g(x1, x2)).reshape(sh + (4, 4))
    G = coords.G(x1, x2).reshape(sh + (4, 4))
    g20 = np.zeros(sh + (4, 5))
    G20 = np.zeros(sh + (4, 5))
    g20[..., :, :4] = g
    G20[..., :, :4] = G
    g20[..., :3, 4] = _dlgdet(coords, x1, x2)
    g20[..., 3, 4] = coords.gdet(x1, x2)
    G20[..., 3, 4] = coords.gttpert(x1, x2)
    return g20.reshape(sh + (20,)), G20.reshape(sh + (20,))


def christoffel(coords, x1, x2):
    """Gamma^i_{jk} (calc_Krzysie_arb_ana) as (..., 4, 4, 4)."""
    dg = _dg(coords, x1, x2)  # [k][l][m] = d_k g_lm
    G = coords.G(np.asarray(x1, float), np.asarray(x2, float))
    # Gamma^i_jk = 1/2 g^il (d_j g_lk + d_k g_lj - d_l g_jk)
    t = dg.transpose(*range(dg.ndim - 3), dg.ndim - 2, dg.ndim - 3, dg.ndim - 1)  # [l][j][k] = d_j g_lk
    term = t + np.swapaxes(t, -1, -2) - dg  # [l][j][k]
    return 0.5 * np.einsum("...il,...ljk->...ijk", G, term)


class Grid:
    """Uniform grid in code coordinates with NG ghosts (set_grid,
    finite.c:1629-1665; calc_xb problem.c:10-38)."""

    def __init__(self, coords, nx, ny, nz, ng, x1lim, x2lim, x3lim):
        self.coords = coords
        self.nx, self.ny, self.nz, self.ng = nx, ny, nz, ng
        self.ngx = ng
        self.ngy = ng if ny > 1 else 0
        self.ngz = ng if nz > 1 else 0
        self.sx = nx + 2 * self.ngx
        self.sy = ny + 2 * self.ngy
        self.sz = nz + 2 * self.ngz
        self.lims = (x1lim, x2lim, x3lim)
        self.xb, self.xc = [], []
        for n, (c0, c1) in zip((nx, ny, nz), self.lims):
            dc = (c1 - c0) / float(n)
            i = np.arange(-ng, n + ng + 1)
            xb = c0 + i.astype(np.float64) * dc
            xc = .5 * (xb[1:] + xb[:-1])
            self.xb.append(xb)
            self.xc.append(xc)

    def face(self, d, i):
        """x^{d+1} of the lower face of cell i (i may be negative)."""
        return self.xb[d][i + self.ng]

    def centre(self, d, i):
        return self.xc[d][i + self.ng]

    def size(self, d, i):
        """get_size_x: finite.c:2248-2252."""
        return self.xb[d][i + 1 + self.ng] - self.xb[d][i + self.ng]

    def cell_index_range(self, d):
        n = (self.nx, self.ny, self.nz)[d]
        g = (self.ngx, self.ngy, self.ngz)[d]
        return np.arange(-g, n + g)

    def widths(self, d):
        """Cell widths for ghost-inclusive indices along dimension d."""
        idx = self.cell_index_range(d)
        return np.array([self.size(d, int(i)) for i in idx])


def grid_tables(grid):
    """Metric tables on the (x1, x2) plane of a grid, including ghosts.

    Returns a dict of float64 arrays indexed [comp, iy, ix] (SoA):
      gc, Gc   (20, SY, SX)     cell centres
      kr       (64, SY, SX)     Christoffels at centres, corrected
      gfx, Gfx (20, SY, SX+1)   x-faces (face ix between cells ix-1, ix)
      gfy, Gfy (20, SY+1, SX)   y-faces
      gfz, Gfz (20, SY, SX)     z-faces (equal to centres: no x3 dependence)
    """
    c = grid.coords
    ix = grid.cell_index_range(0)
    iy = grid.cell_index_range(1)
    x1c = np.array([grid.centre(0, int(i)) for i in ix])
    x2c = np.array([grid.centre(1, int(j)) for j in iy]) if grid.ny > 1 else np.array([grid.centre(1, 0)])
    X1, X2 = np.meshgrid(x1c, x2c)  # [iy, ix]
    gc, Gc = point_tables(c, X1, X2)
    # x faces
    x1f = np.array([grid.face(0, int(i)) for i in range(-grid.ngx, grid.nx + grid.ngx + 1)])
    F1, F2 = np.meshgrid(x1f, x2c)
    gfx, Gfx = point_tables(c, F1, F2)
    if grid.ny > 1:
        x2f = np.array([grid.face(1, int(j)) for j in range(-grid.ngy, grid.ny + grid.ngy + 1)])
    else:
        x2f = np.array([grid.face(1, 0), grid.face(1, 1)])
    Y1, Y2 = np.meshgrid(x1c, x2f)
    gfy, Gfy = point_tables(c, Y1, Y2)
    kr = christoffel(c, X1, X2)
    if c.kind != MINK:
        kr = _modify_christoffel(grid, kr, X1, X2)
    T = lambda a: np.ascontiguousarray(np.moveaxis(a, -1, 0))
    return dict(gc=T(gc), Gc=T(Gc), kr=T(kr.reshape(kr.shape[:-3] + (64,))),
                gfx=T(gfx), Gfx=T(Gfx), gfy=T(gfy), Gfy=T(Gfy), gfz=T(gc), Gfz=T(Gc))


def _modify_christoffel(grid, kr, X1, X2):
    """MODYFIKUJKRZYSIE, metric.c:828-905: rescale Gamma^mu_{mu k} so that
    their sum equals the finite-difference d_k ln sqrt(-g) across the cell."""
    c = grid.coords
    kr = kr.copy()
    org = kr.copy()
    ny_idx = grid.cell_index_range(1) if grid.ny > 1 else np.array([0])
    ix = grid.cell_index_range(0)
    gd0 = c.gdet(X1, X2)
    for kap in (1, 2, 3):
        Ck = sum(org[..., m, kap, m] for m in range(4))
        Sk = 1.e-300 + sum(np.abs(org[..., m, kap, m]) for m in range(4))
        if kap == 1:
            up = np.array([grid.face(0, int(i) + 1) for i in ix])[None, :] + 0 * X1
            lo = np.array([grid.face(0, int(i)) for i in ix])[None, :] + 0 * X1
            gu, gl = c.gdet(up, X2), c.gdet(lo, X2)
            dxk = np.array([grid.size(0, int(i)) for i in ix])[None, :] + 0 * X1
        elif kap == 2:
            up = np.array([grid.face(1, int(j) + 1) for j in ny_idx])[:, None] + 0 * X2
            lo = np.array([grid.face(1, int(j)) for j in ny_idx])[:, None] + 0 * X2
            gu, gl = c.gdet(X1, up), c.gdet(X1, lo)
            dxk = np.array([grid.size(1, int(j)) for j in ny_idx])[:, None] + 0 * X2
        else:
            gu = gl = gd0
            dxk = grid.size(2, 0) + 0 * X1
        D = (gu - gl) / (dxk * gd0)
        for mu in range(4):
            W = np.abs(org[..., mu, kap, mu]) / Sk
            kr[..., mu, kap, mu] = kr[..., mu, kap, mu] + (D - Ck) * W
            kr[..., mu, mu, kap] = kr[..., mu, kap, mu]
    return kr
