"""Geometry ingestion over the C ABI: the reference's Grid3D built from its
Shape3D (triangle mesh) or Shape2D (polylines extruded in depth) input files
(src/FluidSolver3D/Grid3D.cpp:373-981, src/FluidSolver2D/Grid2D.cpp:105-356).

    g = Grid3D("shape3d", "box_pipe_3D_data.txt", dx=0.01, dy=0.01, dz=0.01, baseT=1.0)
    dom = g.Prepare(0.0)          # a cmcx.domains.Domain3D with the node arrays
    solver = make_solver(dom, "f32", path=1)

The walls, outside and fluid nodes are built on the GPU (csrc/geometry.hip).
Grid2D is the 2D solvers' grid of a Shape2D input (Grid2D::Prepare):

    g = Grid2D("box_pipe_data.txt", dx=0.0007, dy=0.0007)
    dom = g.Prepare(0.0)          # a cmcx.domains.Domain2D
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .domains import Domain2D, Domain3D


class Grid3D:
    def __init__(self, kind: str, path: str, dx: float, dy: float, dz: float, baseT: float = 1.0,
                 depth: float = 0.0, depth_var: float = 0.0, align: bool = False):
        """kind: "shape3d" (Grid3D(dx, dy, dz, baseT) + Load3DShape) or "shape2d"
        (Grid3D(dx, dy, dz, depth, depth_var, baseT) + Grid2D::LoadFromFile)."""
        k = _lib.CMCX_SHAPE3D if kind == "shape3d" else _lib.CMCX_SHAPE2D
        sp = (ctypes.c_double * 3)(dx, dy, dz)
        h = _lib._vp(None)
        check(lib().cmcx_geometry_load(k, str(path).encode(), sp, float(baseT), float(depth),
                                       float(depth_var), 1 if align else 0, ctypes.byref(h)),
              "cmcx_geometry_load")
        self.h = h
        self.dx, self.dy, self.dz, self.baseT = dx, dy, dz, baseT
        dims = (ctypes.c_int32 * 3)()
        bmin, bmax = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        nf, cyc = ctypes.c_int(0), ctypes.c_double(0)
        check(lib().cmcx_geometry_info(self.h, dims, bmin, bmax, ctypes.byref(nf), ctypes.byref(cyc)),
              "cmcx_geometry_info")
        self.dimx, self.dimy, self.dimz = (int(x) for x in dims)
        self.bbox_min = tuple(bmin)
        self.bbox_max = tuple(bmax)
        self.num_frames = nf.value
        self.cycle_length = cyc.value

    @property
    def shape(self):
        return (self.dimx, self.dimy, self.dimz)

    def arrays(self, time: float = 0.0):
        """Node arrays (type, bc_vel, bc_temp, vx, vy, vz, T) of the frame at `time`."""
        n = self.dimx * self.dimy * self.dimz
        u8 = [np.zeros(n, np.uint8) for _ in range(3)]
        f8 = [np.zeros(n, np.float64) for _ in range(4)]
        check(lib().cmcx_geometry_build(self.h, float(time), *[a.ctypes.data_as(_lib._u8p) for a in u8],
                                        *[a.ctypes.data_as(_lib._dp) for a in f8]), "cmcx_geometry_build")
        return (*u8, *f8)

    def Prepare(self, time: float = 0.0, **run) -> Domain3D:
        """Grid3D::Prepare(time) as a Domain3D (run settings: re, dt, steps, ...)."""
        t, bv, bt, vx, vy, vz, T = self.arrays(time)
        d = Domain3D(self.dimx, self.dimy, self.dimz, self.dx, self.dy, self.dz, self.baseT, t, bv, bt,
                     vx, vy, vz, T)
        for k, v in run.items():
            setattr(d, k, v)
        return d

    def close(self):
        if self.h and self.h.value:
            lib().cmcx_geometry_destroy(self.h)
            self.h = _lib._vp(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid2D:
    """The reference's Grid2D (src/FluidSolver2D/Grid2D.cpp) of a Shape2D input,
    as the 2D solvers read it: Grid2D(dx, dy, startT, bc_noslip = NoSlip) +
    LoadFromFile, then Prepare(time) per step."""

    def __init__(self, path: str, dx: float, dy: float, startT: float = 1.0, align: bool = False):
        sp = (ctypes.c_double * 3)(dx, dy, dx)
        h = _lib._vp(None)
        check(lib().cmcx_geometry_load(_lib.CMCX_SHAPE2D, str(path).encode(), sp, float(startT), 0.0, 0.0,
                                       1 if align else 0, ctypes.byref(h)), "cmcx_geometry_load")
        self.h = h
        self.dx, self.dy, self.startT = dx, dy, startT
        dims = (ctypes.c_int32 * 3)()
        bmin, bmax = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        nf, cyc = ctypes.c_int(0), ctypes.c_double(0)
        check(lib().cmcx_geometry_info(self.h, dims, bmin, bmax, ctypes.byref(nf), ctypes.byref(cyc)),
              "cmcx_geometry_info")
        self.dimx, self.dimy = int(dims[0]), int(dims[1])
        self.bbox_min = tuple(bmin)[:2]
        self.bbox_max = tuple(bmax)[:2]
        self.num_frames = nf.value
        self.cycle_length = cyc.value

    @property
    def shape(self):
        return (self.dimx, self.dimy)

    def arrays(self, time: float = 0.0):
        """Cell arrays (type, bc, vx, vy, T) of the frame at `time`, y fastest."""
        n = self.dimx * self.dimy
        t, bc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        f8 = [np.zeros(n, np.float64) for _ in range(3)]
        check(lib().cmcx_geometry_build2d(self.h, float(time), t.ctypes.data_as(_lib._u8p),
                                          bc.ctypes.data_as(_lib._u8p),
                                          *[a.ctypes.data_as(_lib._dp) for a in f8]), "cmcx_geometry_build2d")
        return (t, bc, *f8)

    def Prepare(self, time: float = 0.0, **run) -> Domain2D:
        """Grid2D::Prepare(time) as a Domain2D (run settings: dt, steps, ...)."""
        t, bc, vx, vy, T = self.arrays(time)
        d = Domain2D(self.dimx, self.dimy, self.dx, self.dy, self.startT, t, bc, vx, vy, T)
        for k, v in run.items():
            setattr(d, k, v)
        return d

    def close(self):
        if self.h and self.h.value:
            lib().cmcx_geometry_destroy(self.h)
            self.h = _lib._vp(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
