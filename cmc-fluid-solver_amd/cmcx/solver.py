"""Host-side mirror of the reference's solver interface over the C ABI.

``AdiSolver3D`` keeps the method names, argument meaning and error behaviour
of src/FluidSolver3D/AdiSolver3D.h:49-114 and Solver3D.h:24-49:

    solver = AdiSolver3D()
    solver.Init(GPU, False, grid, params)       # AdiSolver3D::Init
    solver.CreateSegments()                     # AdiSolver3D::CreateSegments
    for i in range(steps):
        solver.UpdateBoundaries()               # AdiSolver3D::UpdateBoundaries
        solver.TimeStep(dt, G, L, i % 10 == 0)  # AdiSolver3D::TimeStep, raises DivergenceError
    vel, T = solver.GetLayer()                  # Solver3D::GetLayer (lags one step, as the reference)

``grid`` is a :class:`cmcx.domains.Domain3D` (node arrays of the reference's
Grid3D).  The only backend is the GPU; asking for the CPU backend raises.
Beyond the reference interface, :meth:`fields` returns the current state
(u, v, w, p, T) with p = v_T * T.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .params import FluidParams

CPU, GPU = 0, 1
X, Y, Z = 0, 1, 2

_NP = {_lib.CMCX_F32: np.float32, _lib.CMCX_F64: np.float64}


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class AdiSolver3D:
    def __init__(self, dtype: str = "f32"):
        self.h = ctypes.c_void_p(None)
        self.dtype = _lib.CMCX_F32 if dtype in ("f32", "float32", np.float32) else _lib.CMCX_F64
        self.np_dtype = _NP[self.dtype]
        self.grid = None
        self.params = None
        self._keep = []

    # -- reference interface -------------------------------------------------------
    def Init(self, backend, csv, grid, params: FluidParams, useBlocking: bool = False,
             nblockZ: int = 1, stream=None):
        """AdiSolver3D::Init (AdiSolver3D.cpp:166-268)."""
        if backend != GPU:
            raise ValueError("only the GPU backend exists (no CPU fallback)")
        if useBlocking:
            raise NotImplementedError("the Z-blocked multi-GPU X/Y pipeline is not built (one GPU)")
        L = lib()
        self.grid = grid
        self.params = params
        arrs = [_u8(grid.type), _u8(grid.bc_vel), _u8(grid.bc_temp)]
        vals = [_f64(grid.vx), _f64(grid.vy), _f64(grid.vz), _f64(grid.T)]
        self._keep = arrs + vals
        dom = _lib.cmcx_domain()
        dom.dims[:] = [grid.nx, grid.ny, grid.nz]
        dom.h[:] = [grid.dx, grid.dy, grid.dz]
        dom.base_t = grid.base_t
        dom.type, dom.bc_vel, dom.bc_temp = (a.ctypes.data_as(_lib._u8p) for a in arrs)
        for c in range(3):
            dom.v[c] = vals[c].ctypes.data_as(_lib._dp)
        dom.T = vals[3].ctypes.data_as(_lib._dp)
        prm = _lib.cmcx_params(params.v_T, params.v_vis, params.t_vis, params.t_phi)
        h = ctypes.c_void_p(None)
        check(L.cmcx_create(ctypes.byref(dom), ctypes.byref(prm), self.dtype,
                            ctypes.c_void_p(stream), ctypes.byref(h)), "cmcx_create")
        self.h = h
        if nblockZ != 1:
            check(L.cmcx_set_nblockz(self.h, nblockZ), "cmcx_set_nblockz")
        return self

    def SetOptionsGPU(self, transpose: bool = False, decompose: bool = False):
        """AdiSolver3D::SetOptionsGPU.  `transpose` (coalesced Z sweeps in the
        reference) selects the fused line-tiled sweeps, which stage every axis
        through LDS; `decompose` is parsed and unused, as in the reference."""
        self.set_options(path=_lib.CMCX_PATH_FUSED if transpose else _lib.CMCX_PATH_PLAIN)

    def CreateSegments(self):
        check(lib().cmcx_build_segments(self.h), "cmcx_build_segments")

    def UpdateBoundaries(self):
        check(lib().cmcx_update_boundaries(self.h), "cmcx_update_boundaries")

    def TimeStep(self, dt: float, num_global: int, num_local: int, computeError: bool) -> float:
        """AdiSolver3D::TimeStep; raises DivergenceError like the reference's throw."""
        err = ctypes.c_double(0.0)
        check(lib().cmcx_step(self.h, float(dt), int(num_global), int(num_local),
                              1 if computeError else 0, ctypes.byref(err)), "cmcx_step")
        return err.value

    def GetLayer(self, outdimx: int = 0, outdimy: int = 0, outdimz: int = 0):
        """Solver3D::GetLayer: (vel[n,3], T[n]) of the layer the reference outputs."""
        g = self.grid
        ox, oy, oz = outdimx or g.nx, outdimy or g.ny, outdimz or g.nz
        vel = np.zeros((ox * oy * oz, 3), self.np_dtype)
        T = np.zeros(ox * oy * oz, np.float64)
        check(lib().cmcx_get_layer(self.h, vel.ctypes.data, T.ctypes.data_as(_lib._dp), ox, oy, oz),
              "cmcx_get_layer")
        return vel, T

    def sum_layer(self, ch) -> float:
        return 0.0  # the reference returns 0 (AdiSolver3D.cpp:30-58)

    def debug(self, flag: bool):
        pass

    # -- extensions --------------------------------------------------------------------
    def set_options(self, path: int = _lib.CMCX_PATH_PLAIN, merge_order: int = _lib.CMCX_MERGE_CPU):
        check(lib().cmcx_set_options(self.h, path, merge_order), "cmcx_set_options")

    def num_segments(self, axis: int) -> int:
        n = ctypes.c_int64(0)
        check(lib().cmcx_num_segments(self.h, axis, ctypes.byref(n)), "cmcx_num_segments")
        return n.value

    def segments(self, axis: int) -> np.ndarray:
        n = self.num_segments(axis)
        out = np.zeros((max(n, 1), 7), np.int32)
        check(lib().cmcx_get_segments(self.h, axis, out.ctypes.data_as(_lib._i32p), n), "cmcx_get_segments")
        return out[:n]

    def segment_nodes(self, axis: int):
        n = self.num_segments(axis)
        flags = np.zeros((max(n, 1), 6), np.uint8)
        vals = np.zeros((max(n, 1), 8), np.float64)
        check(lib().cmcx_get_segment_nodes(self.h, axis, flags.ctypes.data_as(_lib._u8p),
                                           vals.ctypes.data_as(_lib._dp), n), "cmcx_get_segment_nodes")
        return flags[:n], vals[:n]

    def fields(self):
        """Current state (u, v, w, p, T) as numpy arrays of the field type."""
        n = self.grid.n
        out = [np.zeros(n, self.np_dtype) for _ in range(5)]
        check(lib().cmcx_get_fields(self.h, *[a.ctypes.data for a in out]), "cmcx_get_fields")
        return tuple(out)

    def set_fields(self, u, v, w, T):
        arrs = [np.ascontiguousarray(a, dtype=self.np_dtype) for a in (u, v, w, T)]
        check(lib().cmcx_set_fields(self.h, *[a.ctypes.data for a in arrs]), "cmcx_set_fields")

    def fields_device(self, u, v, w, p, T):
        """Copy the state into device buffers given as integer addresses (e.g. torch data_ptr())."""
        check(lib().cmcx_get_fields_device(self.h, u, v, w, p, T), "cmcx_get_fields_device")

    def state_ptrs(self):
        arr = (ctypes.c_void_p * 4)()
        check(lib().cmcx_state_device_ptrs(self.h, arr), "cmcx_state_device_ptrs")
        return [arr[i] for i in range(4)]

    def div_error(self) -> float:
        e = ctypes.c_double(0.0)
        check(lib().cmcx_div_error(self.h, ctypes.byref(e)), "cmcx_div_error")
        return e.value

    def run(self, steps: int, dt: float, num_global: int = 4, num_local: int = 1) -> float:
        e = ctypes.c_double(0.0)
        check(lib().cmcx_run(self.h, float(dt), steps, num_global, num_local, ctypes.byref(e)), "cmcx_run")
        return e.value

    def set_boundary(self, bc_vel, bc_temp, vx, vy, vz, T):
        a = [_u8(bc_vel), _u8(bc_temp)]
        v = [_f64(vx), _f64(vy), _f64(vz), _f64(T)]
        check(lib().cmcx_set_boundary(self.h, *[x.ctypes.data_as(_lib._u8p) for x in a],
                                      *[x.ctypes.data_as(_lib._dp) for x in v]), "cmcx_set_boundary")

    def profile(self, enable: bool):
        check(lib().cmcx_profile(self.h, 1 if enable else 0), "cmcx_profile")

    def profile_read(self, reset: bool = True):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        check(lib().cmcx_profile_read(self.h, ctypes.byref(ms), ctypes.byref(n), 1 if reset else 0),
              "cmcx_profile_read")
        return ms.value, n.value

    def step_bytes(self, num_global: int, num_local: int = 1) -> float:
        b = ctypes.c_double(0.0)
        check(lib().cmcx_step_bytes(self.h, num_global, num_local, ctypes.byref(b)), "cmcx_step_bytes")
        return b.value

    def autotune(self, reps: int = 3):
        """Pick the fastest strided-tile variant per axis on this domain."""
        ch = (ctypes.c_int * 3)()
        check(lib().cmcx_autotune(self.h, reps, ch), "cmcx_autotune")
        return [ch[i] for i in range(3)]

    def set_tile_variant(self, axis: int, variant: int):
        check(lib().cmcx_set_tile_variant(self.h, axis, variant), "cmcx_set_tile_variant")

    def set_params(self, params: FluidParams):
        prm = _lib.cmcx_params(params.v_T, params.v_vis, params.t_vis, params.t_phi)
        check(lib().cmcx_set_params(self.h, ctypes.byref(prm)), "cmcx_set_params")
        self.params = params

    # -- reverse mode ------------------------------------------------------------------
    def adjoint_forward(self, steps: int, dt: float, num_global: int = 4, checkpoint_every: int = 20):
        """Rollout of `steps` steps keeping checkpoints for :meth:`adjoint_backward`."""
        check(lib().cmcx_adjoint_forward(self.h, float(dt), int(steps), int(num_global),
                                         int(checkpoint_every)), "cmcx_adjoint_forward")

    def adjoint_backward(self, seed):
        """seed: dL/d(final u, v, w, T), array-like (4, N) (None entries = 0).
        Returns (gbar (4, N), sbar (4, N), pbar (4,)): gradients of the node
        values, of the initial state and of (v_T, v_vis, t_vis, t_phi)."""
        n = self.grid.n
        keep = []
        sp = (_lib._vp * 4)()
        for f in range(4):
            if seed is None or seed[f] is None:
                sp[f] = None
            else:
                a = np.ascontiguousarray(seed[f], dtype=self.np_dtype).reshape(n)
                keep.append(a)
                sp[f] = a.ctypes.data
        gb = np.zeros((4, n), self.np_dtype)
        sb = np.zeros((4, n), self.np_dtype)
        gp = (_lib._vp * 4)(*[gb[f].ctypes.data for f in range(4)])
        bp = (_lib._vp * 4)(*[sb[f].ctypes.data for f in range(4)])
        pb = np.zeros(4, np.float64)
        check(lib().cmcx_adjoint_backward(self.h, sp, gp, bp, pb.ctypes.data_as(_lib._dp)),
              "cmcx_adjoint_backward")
        return gb, sb, pb

    def adjoint_backward_device(self, seed_ptrs, gbar_ptrs, sbar_ptrs):
        """adjoint_backward with device buffers: each argument is 4 device
        addresses (ints, None = absent) of N elements of the field type.
        Returns pbar (4,)."""
        def arr(ps):
            return None if ps is None else (_lib._vp * 4)(*[p if p else None for p in ps])
        pb = np.zeros(4, np.float64)
        check(lib().cmcx_adjoint_backward_device(self.h, arr(seed_ptrs), arr(gbar_ptrs),
                                                 arr(sbar_ptrs), pb.ctypes.data_as(_lib._dp)),
              "cmcx_adjoint_backward_device")
        return pb

    def adjoint_release(self):
        check(lib().cmcx_adjoint_release(self.h), "cmcx_adjoint_release")

    def adjoint_generation(self) -> int:
        """Id of the current recording (0 = none); changes with every
        recording and whenever a parameter, node-value or option change clears it."""
        g = ctypes.c_uint64(0)
        check(lib().cmcx_adjoint_generation(self.h, ctypes.byref(g)), "cmcx_adjoint_generation")
        return g.value

    def set_graph(self, enable: bool = True):
        """Replay each step as a HIP graph (captured on first use)."""
        check(lib().cmcx_set_graph(self.h, 1 if enable else 0), "cmcx_set_graph")

    def synchronize(self):
        check(lib().cmcx_synchronize(self.h), "cmcx_synchronize")

    def close(self):
        if self.h and self.h.value:
            lib().cmcx_destroy(self.h)
            self.h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AdiSolver2D:
    """The reference's 2D ADI solver (src/FluidSolver2D/AdiSolver2D.h, Solver2D.h)
    over the C ABI: Init, UpdateBoundaries, TimeStep (returns (err, iters); raises
    cmcx._lib.CmcxError with CMCX_ERR_PICARD where the reference exits), plus
    fields() = (u, v, p, T)."""

    method = 1  # CMCX2D_ADI

    def __init__(self, dtype: str = "f64"):
        self.h = ctypes.c_void_p(None)
        self.dtype = _lib.CMCX_F32 if dtype in ("f32", "float32", np.float32) else _lib.CMCX_F64
        self.np_dtype = _NP[self.dtype]

    def Init(self, grid, params: FluidParams, stream=None):
        arrs = [_u8(grid.type), _u8(grid.bc)]
        vals = [_f64(grid.vx), _f64(grid.vy), _f64(grid.T)]
        self._keep = arrs + vals
        dom = _lib.cmcx2d_domain()
        dom.dims[:] = [grid.nx, grid.ny]
        dom.h[:] = [grid.dx, grid.dy]
        dom.start_t = grid.base_t
        dom.type, dom.bc = (a.ctypes.data_as(_lib._u8p) for a in arrs)
        dom.v[0] = vals[0].ctypes.data_as(_lib._dp)
        dom.v[1] = vals[1].ctypes.data_as(_lib._dp)
        dom.T = vals[2].ctypes.data_as(_lib._dp)
        prm = _lib.cmcx_params(params.v_T, params.v_vis, params.t_vis, params.t_phi)
        h = ctypes.c_void_p(None)
        check(lib().cmcx2d_create_method(ctypes.byref(dom), ctypes.byref(prm), self.method, self.dtype,
                                         ctypes.c_void_p(stream), ctypes.byref(h)), "cmcx2d_create_method")
        self.h = h
        self.grid = grid
        self.params = params
        return self

    def UpdateBoundaries(self):
        check(lib().cmcx2d_update_boundaries(self.h), "cmcx2d_update_boundaries")

    def SetGrid(self, t, bc, vx, vy, T):
        """Grid2D::Prepare(t) between steps: new cell types, BC kinds and node values."""
        arrs = [_u8(t), _u8(bc)]
        vals = [_f64(vx), _f64(vy), _f64(T)]
        check(lib().cmcx2d_set_grid(self.h, *[a.ctypes.data_as(_lib._u8p) for a in arrs],
                                    *[a.ctypes.data_as(_lib._dp) for a in vals]), "cmcx2d_set_grid")

    def TimeStep(self, dt: float, num_global: int, num_local: int = 1):
        err = ctypes.c_double(0.0)
        it = ctypes.c_int(0)
        check(lib().cmcx2d_step(self.h, float(dt), int(num_global), int(num_local), ctypes.byref(err),
                                ctypes.byref(it)), "cmcx2d_step")
        return err.value, it.value

    def fields(self):
        n = self.grid.nx * self.grid.ny
        out = [np.zeros(n, self.np_dtype) for _ in range(4)]
        check(lib().cmcx2d_get_fields(self.h, *[a.ctypes.data for a in out]), "cmcx2d_get_fields")
        return tuple(out)

    def close(self):
        if self.h and self.h.value:
            lib().cmcx2d_destroy(self.h)
            self.h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StableSolver2D(AdiSolver2D):
    """The reference's 2D Stable solver (src/FluidSolver2D/StableSolver2D.h):
    explicit advection-diffusion of u and v, then a projection whose Poisson
    equation is relaxed on the GPU in the reference's cell order.  Same
    interface as AdiSolver2D (num_local is unused, as in the reference), plus
    the relaxation's settings and sweep counts."""

    method = 2  # CMCX2D_STABLE

    def set_poisson(self, tol: float = 1e-2, max_sweeps: int = 1000000):
        check(lib().cmcx2d_set_poisson(self.h, float(tol), int(max_sweeps)), "cmcx2d_set_poisson")

    def poisson_sweeps(self):
        """Relaxation sweeps of each project of the last step."""
        n = ctypes.c_int(0)
        check(lib().cmcx2d_poisson_sweeps(self.h, None, 0, ctypes.byref(n)), "cmcx2d_poisson_sweeps")
        out = np.zeros(max(n.value, 1), np.int32)
        check(lib().cmcx2d_poisson_sweeps(self.h, out.ctypes.data_as(_lib._i32p), n.value, ctypes.byref(n)),
              "cmcx2d_poisson_sweeps")
        return out[:n.value]


def _domain_struct(grid, keep: list):
    arrs = [_u8(grid.type), _u8(grid.bc_vel), _u8(grid.bc_temp)]
    vals = [_f64(grid.vx), _f64(grid.vy), _f64(grid.vz), _f64(grid.T)]
    keep.extend(arrs + vals)
    dom = _lib.cmcx_domain()
    dom.dims[:] = [grid.nx, grid.ny, grid.nz]
    dom.h[:] = [grid.dx, grid.dy, grid.dz]
    dom.base_t = grid.base_t
    dom.type, dom.bc_vel, dom.bc_temp = (a.ctypes.data_as(_lib._u8p) for a in arrs)
    for c in range(3):
        dom.v[c] = vals[c].ctypes.data_as(_lib._dp)
    dom.T = vals[3].ctypes.data_as(_lib._dp)
    return dom


class BatchSolver3D(AdiSolver3D):
    """Several independent grids advanced together (one fused launch per sweep
    covers all of them).  Per-grid calls (fields, segments, GetLayer, ...)
    address the grid chosen with :meth:`select`."""

    def InitBatch(self, grids, params_list, stream=None):
        if len(grids) != len(params_list) or not grids:
            raise ValueError("one FluidParams per grid")
        keep = []
        doms = (_lib.cmcx_domain * len(grids))(*[_domain_struct(g, keep) for g in grids])
        prms = (_lib.cmcx_params * len(grids))(
            *[_lib.cmcx_params(p.v_T, p.v_vis, p.t_vis, p.t_phi) for p in params_list])
        h = ctypes.c_void_p(None)
        check(lib().cmcx_create_batch(doms, prms, len(grids), self.dtype, ctypes.c_void_p(stream),
                                      ctypes.byref(h)), "cmcx_create_batch")
        self.h = h
        self._keep = keep
        self.grids = list(grids)
        self.grid = self.grids[0]
        self.params = params_list[0]
        return self

    def __len__(self):
        n = ctypes.c_int(0)
        check(lib().cmcx_batch_size(self.h, ctypes.byref(n)), "cmcx_batch_size")
        return n.value

    def select(self, i: int):
        check(lib().cmcx_select(self.h, int(i)), "cmcx_select")
        self.grid = self.grids[i]
        return self

    def grid_errors(self) -> np.ndarray:
        out = np.zeros(len(self.grids), np.float64)
        check(lib().cmcx_grid_errors(self.h, out.ctypes.data_as(_lib._dp)), "cmcx_grid_errors")
        return out

    def fields_of(self, i: int):
        return self.select(i).fields()


def make_batch(doms, dtype: str = "f32", path: int = _lib.CMCX_PATH_FUSED,
               params_list=None) -> BatchSolver3D:
    """InitBatch + SetOptions + CreateSegments for a list of domains."""
    s = BatchSolver3D(dtype)
    s.InitBatch(doms, params_list or [d.params() for d in doms])
    s.set_options(path, _lib.CMCX_MERGE_CPU)
    s.CreateSegments()
    return s


def make_solver(dom, dtype: str = "f32", path: int = _lib.CMCX_PATH_PLAIN,
                merge_order: int = _lib.CMCX_MERGE_CPU, params: FluidParams | None = None,
                nblockz: int = 1, variant: int | None = None) -> AdiSolver3D:
    """Init + SetOptions + CreateSegments for a domain in one call; `variant`
    forces the fused path's X/Y tile variant (the Z sweep takes min(variant, 7))."""
    s = AdiSolver3D(dtype)
    s.Init(GPU, False, dom, params or dom.params(), nblockZ=nblockz)
    s.set_options(path, merge_order)
    s.CreateSegments()
    if variant is not None and path == _lib.CMCX_PATH_FUSED:
        for axis in range(3):
            s.set_tile_variant(axis, min(variant, 7) if axis == Z else variant)
    return s
