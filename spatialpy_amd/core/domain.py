"""Particle domain mirroring spatialpy/core/domain.py (points, per-particle
properties, lattice fills, find_h, type mapping)."""
from collections import OrderedDict

import numpy as np

from .errors import DomainError


class Domain:
    """core/domain.py:37-1240 (the parts the engine consumes)."""

    def __init__(self, numpoints, xlim, ylim, zlim, rho0=1.0, c0=10, P0=None, gravity=None):
        self.vertices = np.zeros((numpoints, 3), dtype=float)
        self.vol = np.zeros(numpoints)
        self.mass = np.zeros(numpoints)
        self.type_id = np.array(["type_UnAssigned"] * numpoints, dtype=object)
        self.nu = np.zeros(numpoints)
        self.c = np.zeros(numpoints)
        self.rho = np.zeros(numpoints)
        self.fixed = np.zeros(numpoints, dtype=bool)
        self.rho0 = rho0
        self.c0 = c0
        self.P0 = P0 if P0 is not None else rho0 * c0 ** 2       # domain.py:92-95
        self.gravity = gravity
        self.xlim, self.ylim, self.zlim = tuple(xlim), tuple(ylim), tuple(zlim)
        self.typeNdxMapping = None
        self.listOfTypeIDs = []
        self.dimensions = None

    def add_point(self, point, vol=1, mass=1, type_id="UnAssigned", nu=0, fixed=False, rho=None, c=10):
        """domain.py:203-255."""
        if vol < 0:
            raise DomainError("Volume must be a positive value.")
        if isinstance(type_id, int) and type_id <= 0:
            raise DomainError("Type_id must be a non-zero positive integer or a string.")
        if rho is None:
            rho = mass / vol
        self.type_id = np.append(self.type_id, f"type_{type_id}")
        self.vol = np.append(self.vol, vol)
        self.mass = np.append(self.mass, mass)
        self.nu = np.append(self.nu, nu)
        self.c = np.append(self.c, c)
        self.rho = np.append(self.rho, rho)
        self.fixed = np.append(self.fixed, fixed)
        self.vertices = np.append(self.vertices, [point], axis=0)

    def add_points(self, points, vol, mass, type_id, nu, fixed, rho=None, c=0):
        """Vectorised add_point for many particles with the same properties."""
        points = np.asarray(points, dtype=float).reshape(-1, 3)
        m = points.shape[0]
        if rho is None:
            rho = mass / vol
        self.type_id = np.concatenate([self.type_id, np.array([f"type_{type_id}"] * m, dtype=object)])
        self.vol = np.concatenate([self.vol, np.full(m, vol, float)])
        self.mass = np.concatenate([self.mass, np.full(m, mass, float)])
        self.nu = np.concatenate([self.nu, np.full(m, nu, float)])
        self.c = np.concatenate([self.c, np.full(m, c, float)])
        self.rho = np.concatenate([self.rho, np.full(m, rho, float)])
        self.fixed = np.concatenate([self.fixed, np.full(m, bool(fixed))])
        self.vertices = np.concatenate([self.vertices, points])

    def fill_lattice(self, xmin, xmax, dx, ymin, ymax, dy, zmin=0.0, zmax=0.0, dz=0.0, inside=None, **props):
        """CartesianLattice fill (core/lattice.py:167-231)."""
        from ..workloads import cartesian_lattice
        pts = cartesian_lattice(xmin, xmax, dx, ymin, ymax, dy, zmin, zmax, dz, inside=inside)
        if "vol" not in props or props["vol"] is None:
            vol = 4 / 3 * np.pi * (dx / 2) * (dy / 2) * (dz / 2) if dz > 0 else np.pi * (dx / 2) * (dy / 2)
            props["vol"] = vol
        props.setdefault("mass", 1.0)
        props.setdefault("nu", 1.0)
        props.setdefault("fixed", False)
        props.setdefault("type_id", 1)
        self.add_points(pts, props["vol"], props["mass"], props["type_id"], props["nu"], props["fixed"],
                        props.get("rho"), props.get("c", 0))
        lo, hi = self.vertices.min(axis=0), self.vertices.max(axis=0)
        self.xlim, self.ylim, self.zlim = (lo[0], hi[0]), (lo[1], hi[1]), (lo[2], hi[2])
        return pts.shape[0]

    @classmethod
    def create_2D_domain(cls, xlim, ylim, numx, numy, type_id=1, mass=1.0, nu=1.0, rho=None, c=0,
                         fixed=False, **kwargs):
        """domain.py:508-572."""
        xl = np.linspace(xlim[0], xlim[1], numx)
        yl = np.linspace(ylim[0], ylim[1], numy)
        dx = xlim[1] - xlim[0] if len(xl) <= 1 else xl[1] - xl[0]
        dy = ylim[1] - ylim[0] if len(yl) <= 1 else yl[1] - yl[0]
        vol = abs(xlim[1] - xlim[0]) * abs(ylim[1] - ylim[0]) / (numx * numy)
        obj = cls(0, xlim, ylim, (0, 0), **kwargs)
        obj.fill_lattice(xlim[0], xlim[1], dx, ylim[0], ylim[1], dy, 0.0, 0.0, 0.0, type_id=type_id, vol=vol,
                         mass=mass, nu=nu, rho=rho, c=c, fixed=fixed)
        return obj

    @classmethod
    def create_3D_domain(cls, xlim, ylim, zlim, numx, numy, numz, type_id=1, mass=1.0, nu=1.0, rho=None, c=0,
                         fixed=False, **kwargs):
        """domain.py:574-645."""
        xl, yl, zl = (np.linspace(a[0], a[1], m) for a, m in ((xlim, numx), (ylim, numy), (zlim, numz)))
        vol = abs(xlim[1] - xlim[0]) * abs(ylim[1] - ylim[0]) * abs(zlim[1] - zlim[0]) / (numx * numy * numz)
        obj = cls(0, xlim, ylim, zlim, **kwargs)
        obj.fill_lattice(xlim[0], xlim[1], xl[1] - xl[0], ylim[0], ylim[1], yl[1] - yl[0], zlim[0], zlim[1],
                         zl[1] - zl[0], type_id=type_id, vol=vol, mass=mass, nu=nu, rho=rho, c=c, fixed=fixed)
        return obj

    def coordinates(self):
        return self.vertices

    def get_num_voxels(self):
        return self.vertices.shape[0]

    def find_h(self):
        """domain.py:751-769."""
        from ..workloads import find_h
        return find_h(self.vertices)

    def compile_prep(self):
        """domain.py:479-496 + _get_type_mappings (:137-146).  Types are
        numbered in order of first appearance (the reference iterates a
        Python set, whose order depends on the string hash seed)."""
        if list(self.type_id).count("type_UnAssigned") > 0:
            raise DomainError("Particles must be assigned a type_id.")
        if np.count_nonzero(self.rho) < len(self.rho):
            raise DomainError("Rho must be a positive value.")
        self.typeNdxMapping = OrderedDict({"type_UnAssigned": 0})
        for t in self.type_id:
            if t not in self.typeNdxMapping:
                self.typeNdxMapping[t] = len(self.typeNdxMapping)
        self.listOfTypeIDs = list(self.typeNdxMapping.values())
