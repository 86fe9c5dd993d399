"""Grid description shared by the device path and its host-side set-up.

A grid is NX x NY x NZ interior cells (a collapsed dimension has size 1 and
no ghosts, ko.h:225-245) with NG ghosts (3 for PPM, 4 for MP5, 2 otherwise,
ko.h:209-222).  Device state is SoA: x[iv, iz, iy, ix] over the ghost-inclusive
box SZ x SY x SX.  Metric tables are per (x1, x2) point.
"""
import numpy as np

from . import metric as M

BC_COPY, BC_PERIODIC, BC_REFLECT = 0, 1, 2


def ng_for(int_order):
    return {0: 2, 1: 2, 2: 3, 4: 4}[int_order]


class GridSpec:
    def __init__(self, coords, n, lims, ng, bc):
        self.coords = coords
        self.nx, self.ny, self.nz = n
        self.ng = ng
        self.bc = list(bc)
        self.g = M.Grid(coords, self.nx, self.ny, self.nz, ng, *lims)
        self.sx, self.sy, self.sz = self.g.sx, self.g.sy, self.g.sz
        self.ngx, self.ngy, self.ngz = self.g.ngx, self.g.ngy, self.g.ngz
        self.tables = M.grid_tables(self.g)  # SoA [comp, y, x]
        self.xb = [np.ascontiguousarray(x) for x in self.g.xb]

    @property
    def shape(self):
        return (self.sz, self.sy, self.sx)

    @property
    def ncell(self):
        return self.sx * self.sy * self.sz

    def interior(self):
        """Slices of the interior inside a ghost-inclusive [z, y, x] box."""
        return (slice(self.ngz, self.ngz + self.nz), slice(self.ngy, self.ngy + self.ny),
                slice(self.ngx, self.ngx + self.nx))

    def centres(self):
        """Code coordinates of all cell centres, arrays [z, y, x]."""
        x1 = self.g.xc[0][self.ng - self.ngx: self.ng + self.nx + self.ngx]
        x2 = self.g.xc[1][self.ng - self.ngy: self.ng + self.ny + self.ngy]
        x3 = self.g.xc[2][self.ng - self.ngz: self.ng + self.nz + self.ngz]
        Z, Y, X = np.meshgrid(x3, x2, x1, indexing="ij")
        return X, Y, Z

    def aos_tables(self):
        """Metric tables as [point, comp] (point = iy*SX+ix) for the CPU oracle."""
        t = self.tables
        return {k: np.ascontiguousarray(np.moveaxis(v, 0, -1).reshape(-1, v.shape[0])) for k, v in t.items()}

    def cell_tables(self):
        """(g20, G20) per ghost-inclusive cell as [z, y, x, 20] (no x3 dependence)."""
        gc = np.moveaxis(self.tables["gc"], 0, -1)
        Gc = np.moveaxis(self.tables["Gc"], 0, -1)
        return (np.broadcast_to(gc[None], (self.sz,) + gc.shape),
                np.broadcast_to(Gc[None], (self.sz,) + Gc.shape))


def to_aos(x):
    """SoA [nv, z, y, x] -> AoS [z, y, x, nv] (KORAL's layout)."""
    return np.ascontiguousarray(np.moveaxis(x, 0, -1))


def to_soa(x):
    return np.ascontiguousarray(np.moveaxis(x, -1, 0))
