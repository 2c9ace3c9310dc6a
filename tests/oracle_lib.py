"""ctypes access to the oracle (oracle/_build/liboracle.so) and to the
reference-header pieces (oracle/_ref/libref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libref.so")

_c_u8 = ctypes.POINTER(ctypes.c_uint8)
_c_d = ctypes.POINTER(ctypes.c_double)
_c_f = ctypes.POINTER(ctypes.c_float)
_c_i32 = ctypes.POINTER(ctypes.c_int32)

_lib = None
_ref = None


def _ptr(a, t):
    return a.ctypes.data_as(t)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-C", ORACLE_DIR], stdout=subprocess.DEVNULL)
    L = ctypes.CDLL(ORACLE_SO)
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             _c_u8, _c_u8, _c_u8, _c_d, _c_d, _c_d, _c_d,
                             ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_int, ctypes.c_int]
    for name in ("orc_destroy", "orc_create_segments", "orc_update_boundaries"):
        getattr(L, name).argtypes = [ctypes.c_void_p]
        getattr(L, name).restype = None
    L.orc_num_segments.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_num_segments.restype = ctypes.c_int
    L.orc_get_segments.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_i32]
    L.orc_get_segments.restype = None
    L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, _c_d]
    L.orc_step.restype = ctypes.c_int
    L.orc_get_layer_raw.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_d, _c_d, _c_d, _c_d]
    L.orc_get_layer_raw.restype = None
    L.orc_set_fields.argtypes = [ctypes.c_void_p, _c_d, _c_d, _c_d, _c_d]
    L.orc_set_boundary.argtypes = [ctypes.c_void_p, _c_u8, _c_u8, _c_d, _c_d, _c_d, _c_d]
    L.orc_set_boundary.restype = None
    L.orc_set_fields.restype = None
    L.orc_div_error_cur.argtypes = [ctypes.c_void_p]
    L.orc_div_error_cur.restype = ctypes.c_double
    L.orc_get_layer.argtypes = [ctypes.c_void_p, _c_f, _c_d, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_get_layer.restype = None
    L.orc_thomas_f64.argtypes = [_c_d] * 5 + [ctypes.c_int]
    L.orc_thomas_f32.argtypes = [_c_f] * 5 + [ctypes.c_int]
    L.orc_segments_count.argtypes = [_c_u8, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_segments_count.restype = ctypes.c_int
    L.orc_segments.argtypes = [_c_u8, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_i32]
    L.orc_adj_record.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.orc_adj_record.restype = None
    L.orc_adj_backward.argtypes = [ctypes.c_void_p, _c_d, _c_d, _c_d, _c_d]
    L.orc_adj_backward.restype = None
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_max_threads.restype = ctypes.c_int
    # 2D
    L.orc2_create.restype = ctypes.c_void_p
    L.orc2_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, _c_u8, _c_u8, _c_d, _c_d, _c_d,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.orc2_destroy.argtypes = [ctypes.c_void_p]
    L.orc2_update_boundaries.argtypes = [ctypes.c_void_p]
    L.orc2_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, _c_d,
                            ctypes.POINTER(ctypes.c_int)]
    L.orc2_step.restype = ctypes.c_int
    L.orc2_get_layer_raw.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_d, _c_d, _c_d]
    L.orc2_segments.argtypes = [ctypes.c_void_p, ctypes.c_int, _c_i32]
    L.orc2_segments.restype = ctypes.c_int
    L.orc2_set_method.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc2_set_grid.argtypes = [ctypes.c_void_p, _c_u8, _c_u8, _c_d, _c_d, _c_d]
    L.orc2_stable_solve_uv.argtypes = [ctypes.c_void_p, ctypes.c_double] + [_c_d] * 6
    L.orc2_set_poisson.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.orc2_poisson_sweeps.argtypes = [ctypes.c_void_p, _c_i32, ctypes.c_int]
    L.orc2_poisson_sweeps.restype = ctypes.c_int
    L.orc2_stable_project.argtypes = [ctypes.c_void_p, _c_d, _c_d, _c_d, _c_d, _c_d,
                                      ctypes.POINTER(ctypes.c_int)]
    L.orc2_stable_project.restype = ctypes.c_int
    # geometry ingestion
    L.orc_geo_load.restype = ctypes.c_void_p
    L.orc_geo_load.argtypes = [ctypes.c_int, ctypes.c_char_p] + [ctypes.c_double] * 6 + [ctypes.c_int]
    L.orc_geo_dims.argtypes = [ctypes.c_void_p, _c_i32]
    L.orc_geo_build.argtypes = [ctypes.c_void_p, ctypes.c_double, _c_u8, _c_u8, _c_u8, _c_d, _c_d, _c_d, _c_d]
    L.orc_geo_grid2d.argtypes = [ctypes.c_void_p, _c_i32, _c_u8]
    L.orc_geo_build2d.argtypes = [ctypes.c_void_p, ctypes.c_double, _c_u8, _c_u8, _c_d, _c_d, _c_d]
    L.orc_geo_free.argtypes = [ctypes.c_void_p]
    _lib = L
    return L


def load_ref():
    """The reference's own Thomas / FluidParams compiled from its headers, or None."""
    global _ref
    if _ref is not None:
        return _ref
    if not os.path.exists(REF_SO):
        if os.path.isdir("/root/reference/src/Common"):
            subprocess.call(["make", "-C", ORACLE_DIR, "ref"], stdout=subprocess.DEVNULL)
        if not os.path.exists(REF_SO):
            return None
    R = ctypes.CDLL(REF_SO)
    R.ref_thomas_f32.argtypes = [_c_f] * 5 + [ctypes.c_int]
    R.ref_thomas_f64.argtypes = [_c_d] * 5 + [ctypes.c_int]
    R.ref_fluid_params_normalized.argtypes = [ctypes.c_double] * 3 + [_c_f]
    R.ref_fluid_params_dimensional.argtypes = [ctypes.c_double] * 5 + [_c_f]
    _ref = R
    return R


def segments(type_u8: np.ndarray, shape, axis: int, nblockz: int = 1) -> np.ndarray:
    L = load()
    t = np.ascontiguousarray(type_u8, dtype=np.uint8)
    nx, ny, nz = shape
    n = L.orc_segments_count(_ptr(t, _c_u8), nx, ny, nz, axis, nblockz)
    out = np.zeros((max(n, 1), 7), np.int32)
    L.orc_segments(_ptr(t, _c_u8), nx, ny, nz, axis, nblockz, _ptr(out, _c_i32))
    return out[:n]


class Oracle3D:
    """CPU restatement of AdiSolver3D (CPU backend) for one domain."""

    def __init__(self, dom, dtype: str = "f64", merge_mode: int = 0, params=None, nblockz: int = 1):
        L = load()
        self.L = L
        self.dom = dom
        self.shape = dom.shape
        p = params or dom.params()
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in (dom.type, dom.bc_vel, dom.bc_temp)]
        vals = [np.ascontiguousarray(a, dtype=np.float64) for a in (dom.vx, dom.vy, dom.vz, dom.T)]
        self._keep = arrs + vals
        self.h = L.orc_create(0 if dtype == "f32" else 1, dom.nx, dom.ny, dom.nz, dom.dx, dom.dy,
                              dom.dz, *[_ptr(a, _c_u8) for a in arrs], *[_ptr(a, _c_d) for a in vals],
                              p.v_T, p.v_vis, p.t_vis, p.t_phi, merge_mode, nblockz)
        self.p = p

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def create_segments(self):
        self.L.orc_create_segments(self.h)

    def get_segments(self, axis: int) -> np.ndarray:
        n = self.L.orc_num_segments(self.h, axis)
        out = np.zeros((max(n, 1), 7), np.int32)
        self.L.orc_get_segments(self.h, axis, _ptr(out, _c_i32))
        return out[:n]

    def update_boundaries(self):
        self.L.orc_update_boundaries(self.h)

    def step(self, dt, G=4, L=1, compute_err=True):
        err = ctypes.c_double(0.0)
        r = self.L.orc_step(self.h, dt, G, L, 1 if compute_err else 0, ctypes.byref(err))
        return r, err.value

    def layer(self, which: int = 0):
        n = self.dom.n
        out = [np.zeros(n) for _ in range(4)]
        self.L.orc_get_layer_raw(self.h, which, *[_ptr(a, _c_d) for a in out])
        return out

    def fields(self):
        return self.layer(0)

    def set_fields(self, u, v, w, T):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, T)]
        self.L.orc_set_fields(self.h, *[_ptr(a, _c_d) for a in arrs])

    def set_boundary(self, bc_vel, bc_temp, vx, vy, vz, T):
        bv = np.ascontiguousarray(bc_vel, dtype=np.uint8)
        bt = np.ascontiguousarray(bc_temp, dtype=np.uint8)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy, vz, T)]
        self.L.orc_set_boundary(self.h, _ptr(bv, _c_u8), _ptr(bt, _c_u8),
                                *[_ptr(a, _c_d) for a in arrs])

    def div_error(self):
        return self.L.orc_div_error_cur(self.h)

    def get_layer(self, ox=0, oy=0, oz=0):
        nx, ny, nz = self.shape
        ox, oy, oz = ox or nx, oy or ny, oz or nz
        v = np.zeros((ox * oy * oz, 3), np.float32)
        T = np.zeros(ox * oy * oz)
        self.L.orc_get_layer(self.h, _ptr(v, _c_f), _ptr(T, _c_d), ox, oy, oz)
        return v, T

    def adj_record(self, steps, dt, G):
        """Runs `steps` recorded steps (UpdateBoundaries + TimeStep, L = 1)."""
        self.L.orc_adj_record(self.h, steps, dt, G)

    def adj_backward(self, seed):
        """seed: adjoint of the final state, shape (4, N).  Returns (gbar (4, N),
        cbar0 (4, N), pbar (4,) for v_T, v_vis, t_vis, t_phi)."""
        n = self.dom.n
        sd = np.ascontiguousarray(seed, dtype=np.float64).reshape(4 * n)
        gb = np.zeros(4 * n)
        cb = np.zeros(4 * n)
        pb = np.zeros(4)
        self.L.orc_adj_backward(self.h, _ptr(sd, _c_d), _ptr(gb, _c_d), _ptr(cb, _c_d), _ptr(pb, _c_d))
        return gb.reshape(4, n), cb.reshape(4, n), pb

    def run(self, steps, dt=None, G=None, L=None, compute_every=10):
        d = self.dom
        dt = d.dt if dt is None else dt
        G = d.num_global if G is None else G
        L = d.num_local if L is None else L
        for i in range(steps):
            self.update_boundaries()
            r, err = self.step(dt, G, L, (i % compute_every == 0) or (i == steps - 1))
            if r != 0:
                return r, err
        return 0, err


class Oracle2D:
    """The 2D solvers: method "adi" (AdiSolver2D) or "stable" (StableSolver2D)."""

    def __init__(self, dom, dtype: str = "f64", params=None, method: str = "adi"):
        L = load()
        self.L = L
        self.dom = dom
        from cmcx.params import FluidParams
        p = params or FluidParams.normalized(dom.re, 0.72, 1.4)
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in (dom.type, dom.bc)]
        vals = [np.ascontiguousarray(a, dtype=np.float64) for a in (dom.vx, dom.vy, dom.T)]
        self._keep = arrs + vals
        self.h = L.orc2_create(0 if dtype == "f32" else 1, dom.nx, dom.ny, dom.dx, dom.dy,
                               dom.base_t, *[_ptr(a, _c_u8) for a in arrs],
                               *[_ptr(a, _c_d) for a in vals], p.v_T, p.v_vis, p.t_vis, p.t_phi)
        if method == "stable":
            L.orc2_set_method(self.h, 2)

    def set_grid(self, t, bc, vx, vy, T):
        """Grid2D::Prepare(t) between steps: new cell types, BC kinds, node values."""
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in (t, bc)]
        vals = [np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy, T)]
        self.L.orc2_set_grid(self.h, *[_ptr(a, _c_u8) for a in arrs], *[_ptr(a, _c_d) for a in vals])

    def solve_uv(self, dt, cu, cv, tu, tv, ou=None, ov=None):
        """SolveU / SolveV alone: (ou, ov) on NODE_IN cells from cur (cu, cv) and temp (tu, tv)."""
        n = self.dom.nx * self.dom.ny
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (cu, cv, tu, tv)]
        ou = np.zeros(n) if ou is None else np.array(ou, dtype=np.float64)
        ov = np.zeros(n) if ov is None else np.array(ov, dtype=np.float64)
        self.L.orc2_stable_solve_uv(self.h, dt, *[_ptr(a, _c_d) for a in ins], _ptr(ou, _c_d), _ptr(ov, _c_d))
        return ou, ov

    def set_poisson(self, tol=1e-2, cap=1 << 30):
        self.L.orc2_set_poisson(self.h, tol, cap)

    def poisson_sweeps(self):
        n = self.L.orc2_poisson_sweeps(self.h, ctypes.cast(None, _c_i32), 0)
        out = np.zeros(max(n, 1), np.int32)
        self.L.orc2_poisson_sweeps(self.h, _ptr(out, _c_i32), n)
        return out[:n]

    def project(self, wu, wv):
        """One project of the Stable solver on (wu, wv): (q, pu, pv, sweeps, capped)."""
        n = self.dom.nx * self.dom.ny
        wu = np.ascontiguousarray(wu, dtype=np.float64)
        wv = np.ascontiguousarray(wv, dtype=np.float64)
        q, pu, pv = (np.zeros(n) for _ in range(3))
        sw = ctypes.c_int(0)
        r = self.L.orc2_stable_project(self.h, _ptr(wu, _c_d), _ptr(wv, _c_d), _ptr(q, _c_d),
                                       _ptr(pu, _c_d), _ptr(pv, _c_d), ctypes.byref(sw))
        return q, pu, pv, sw.value, r

    def __del__(self):
        try:
            self.L.orc2_destroy(self.h)
        except Exception:
            pass

    def update_boundaries(self):
        self.L.orc2_update_boundaries(self.h)

    def step(self, dt, G=4, L=1):
        err = ctypes.c_double(0.0)
        it = ctypes.c_int(0)
        r = self.L.orc2_step(self.h, dt, G, L, ctypes.byref(err), ctypes.byref(it))
        return r, err.value, it.value

    def layer(self, which=0):
        n = self.dom.nx * self.dom.ny
        out = [np.zeros(n) for _ in range(3)]
        self.L.orc2_get_layer_raw(self.h, which, *[_ptr(a, _c_d) for a in out])
        return out

    def segments(self, axis):
        n = self.L.orc2_segments(self.h, axis, ctypes.cast(None, _c_i32))
        out = np.zeros((max(n, 1), 5), np.int32)
        self.L.orc2_segments(self.h, axis, _ptr(out, _c_i32))
        return out[:n]


class OracleGeometry:
    """CPU restatement of the reference's geometry loaders (Shape3D triangle
    meshes, Shape2D polylines extruded in depth)."""

    def __init__(self, kind: str, path: str, dx, dy, dz, base_t=1.0, depth=0.0, depth_var=0.0,
                 align=False):
        L = load()
        self.L = L
        k = 3 if kind == "shape3d" else 2
        self.h = L.orc_geo_load(k, str(path).encode(), dx, dy, dz, depth, depth_var, base_t,
                                1 if align else 0)
        if not self.h:
            raise RuntimeError(f"oracle could not load {path}")
        d = np.zeros(3, np.int32)
        L.orc_geo_dims(self.h, _ptr(d, _c_i32))
        self.shape = tuple(int(x) for x in d)

    def build(self, time=0.0):
        n = int(np.prod(self.shape))
        t, bv, bt = (np.zeros(n, np.uint8) for _ in range(3))
        vx, vy, vz, T = (np.zeros(n) for _ in range(4))
        self.L.orc_geo_build(self.h, time, *[_ptr(a, _c_u8) for a in (t, bv, bt)],
                             *[_ptr(a, _c_d) for a in (vx, vy, vz, T)])
        return t, bv, bt, vx, vy, vz, T

    def build2d(self, time=0.0):
        """Grid2D::Prepare(time): (type, bc, vx, vy, T) over the 2D grid, y fastest."""
        d = np.zeros(2, np.int32)
        self.L.orc_geo_grid2d(self.h, _ptr(d, _c_i32), ctypes.cast(None, _c_u8))
        n = int(d[0] * d[1])
        t, bc = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
        vx, vy, T = (np.zeros(n) for _ in range(3))
        self.L.orc_geo_build2d(self.h, time, _ptr(t, _c_u8), _ptr(bc, _c_u8), _ptr(vx, _c_d),
                               _ptr(vy, _c_d), _ptr(T, _c_d))
        self.dims2d = (int(d[0]), int(d[1]))
        return t, bc, vx, vy, T

    def grid2d(self):
        d = np.zeros(2, np.int32)
        self.L.orc_geo_grid2d(self.h, _ptr(d, _c_i32), ctypes.cast(None, _c_u8))
        c = np.zeros(int(d[0] * d[1]), np.uint8)
        self.L.orc_geo_grid2d(self.h, _ptr(d, _c_i32), _ptr(c, _c_u8))
        return c.reshape(int(d[0]), int(d[1]))

    def close(self):
        if self.h:
            self.L.orc_geo_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
