#!/usr/bin/env python3
"""Benchmark of the ADI time step on one MI355X.

Headline workload: the 256^3 irregular domain with 40 seeded ellipsoids
removed (FP32, inlet u=1, outlet BC_FREE, Re=1000, dt=2e-4, 4 Picard
iterations, 1 local sweep).  A step is UpdateBoundaries + TimeStep over the
whole grid, with the state resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--path fused|plain]

The timed region is bracketed by device synchronisations of the HIP runtime the
solver runs on (torch is not imported: it bundles a second HIP runtime).

Prints one JSON line:
  value / ms_per_step  cell-updates/s of the headline path (fused sweeps; the
                       faster of direct launches and graph replay)
  roofline             the dominant kernel, the sweep: algorithmic bytes
                       (16w+1)*S_d per sweep over the sweep-kernel time from HIP
                       events on the solver's stream; traffic from the committed
                       PMC summary
  cpu_baseline         the oracle (C++/OpenMP restatement of the reference CPU
                       path) on a bounded sample, on the host cores
  plain                the per-segment, per-axis-launch GPU path
  batch                64 independent boxes advanced together (one launch per
                       sweep) against the plain path on the same grids
  adjoint              gradient of a probe's kinetic energy after 200 steps of
                       a 96^3 channel w.r.t. a 32x32 inlet control and the
                       viscosity: backward/forward time and finite-difference check
  channel2d            the 2D ADI path on its 256 x 128 channel, GPU beside the
                       oracle on the host
  stable2d             the 2D Stable (projection) path on the reference's box_pipe
                       example: the Poisson relaxation's time per sweep and per
                       wavefront step, GPU beside the oracle on the host
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cmc-fluid-solver_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "ADI cell-updates/s + %HBM peak, batched vs plain on 1 GPU; adjoint/forward time ratio"


def _hip():
    """The HIP runtime libcmcx.so is linked against (same process-wide instance)."""
    import ctypes

    from cmcx._lib import lib

    lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipDeviceSynchronize.restype = ctypes.c_int
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _sync():
    # device-wide barrier on the runtime the solver uses (one GPU, one process)
    rc = _hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


class DeviceFields:
    """`k` device arrays of `n` elements (hipMalloc in the solver's runtime)."""

    def __init__(self, k: int, n: int, dtype):
        import ctypes

        self.h = _hip()
        self.n, self.dtype = n, np.dtype(dtype)
        self.ptrs = []
        for _ in range(k):
            p = ctypes.c_void_p()
            assert self.h.hipMalloc(ctypes.byref(p), n * self.dtype.itemsize) == 0
            self.ptrs.append(p.value)

    def upload(self, arrays):
        for p, a in zip(self.ptrs, arrays):
            a = np.ascontiguousarray(a, dtype=self.dtype)
            assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return self.ptrs

    def download(self, i):
        out = np.empty(self.n, self.dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptrs[i], out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def copy_calibration(mib: int):
    """A plain device-to-device copy of `mib` MiB (3 times), the reference kernel
    for scaling the PMC byte counters."""
    import ctypes

    h = _hip()
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(a), n) == 0 and h.hipMalloc(ctypes.byref(b), n) == 0
    assert h.hipMemset(a, 1, n) == 0
    for _ in range(3):
        assert h.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    _sync()
    h.hipFree(a)
    h.hipFree(b)


def run_plain_steps(dom, path: int, steps: int, warmup: int, dtype: str = "f32", variants=None):
    """Steps only, no HIP events and no autotuning (for PMC counter runs, where the
    profiler does not support event timing); `variants` fixes the tile variants."""
    from cmcx.solver import make_solver

    s = make_solver(dom, dtype, path=path)
    if variants and path == 1:
        for a, v in enumerate(variants):
            s.set_tile_variant(a, v)
    for i in range(warmup + steps):
        s.UpdateBoundaries()
        s.TimeStep(dom.dt, dom.num_global, dom.num_local, False)
    s.synchronize()
    s.close()


def time_solver(s, cells: int, dt: float, G: int, L: int, steps: int, warmup: int,
                tune: bool = True, graph: bool = False):
    """Times `steps` steps (boundary fill + TimeStep) of a built solver: one pass
    with HIP events around every sweep launch (kernel time), one wall-clock pass
    without events (direct launches or graph replays)."""
    variants = s.autotune(3) if tune else None
    s.set_graph(graph)
    for _ in range(warmup):
        s.UpdateBoundaries()
        s.TimeStep(dt, G, L, False)
    s.synchronize()
    _sync()
    # kernel timing pass (HIP events around each sweep launch on the solver's stream)
    s.profile(True)
    s.profile_read(reset=True)
    for _ in range(steps):
        s.UpdateBoundaries()
        s.TimeStep(dt, G, L, False)
    sweep_ms, launches = s.profile_read(reset=True)
    s.profile(False)
    s.synchronize()
    _sync()
    # wall-clock pass without events
    t0 = time.perf_counter()
    if graph:
        s.run(steps, dt, G, L)  # graph of boundary fill + step, error every 10th step
    else:
        for i in range(steps):
            s.UpdateBoundaries()
            s.TimeStep(dt, G, L, i % 10 == 0 or i == steps - 1)
    s.synchronize()
    _sync()
    wall = time.perf_counter() - t0
    step_bytes = s.step_bytes(G, L)
    return {
        "ms_per_step": wall / steps * 1e3,
        "cell_updates_per_s": cells * steps / wall,
        "sweep_ms_per_launch": sweep_ms / max(launches, 1),
        "sweep_launches": launches,
        "sweep_bytes_per_step": step_bytes,
        "sweep_achieved_gbs": step_bytes * steps / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else 0.0,
        "variants": variants,
    }


def time_path(dom, path: int, steps: int, warmup: int, dtype: str = "f32", tune: bool = True,
              graph: bool = False):
    from cmcx.solver import make_solver

    s = make_solver(dom, dtype, path=path)
    r = time_solver(s, dom.n, dom.dt, dom.num_global, dom.num_local, steps, warmup,
                    tune=tune and path == 1, graph=graph)
    r["div_error"] = s.div_error()
    s.close()
    return r


def time_batch(steps: int, warmup: int, count: int = 64):
    """The multi-grid workload: `count` seeded boxes (edges from {32,48,64,96})
    advanced together; fused (one launch per sweep, graph replay) against the
    plain per-segment kernels (one launch per grid and sweep)."""
    from cmcx import domains
    from cmcx.solver import make_batch

    doms = domains.batch_boxes(count=count, seed=0)
    cells = sum(d.n for d in doms)
    d0 = doms[0]
    out = {"grids": count, "cells": cells, "seed": 0}
    s = make_batch(doms, "f32", path=1)
    out["fused"] = time_solver(s, cells, d0.dt, d0.num_global, 1, steps, warmup, tune=True,
                               graph=True)
    s.close()
    s = make_batch(doms, "f32", path=0)
    out["plain"] = time_solver(s, cells, d0.dt, d0.num_global, 1, max(2, steps // 4), 1,
                               tune=False, graph=False)
    s.close()
    return out


def _probe(shape, size: int = 16):
    """Probe box behind the cylinder (whose centre lies in the upstream half):
    x in [nx/2+1, nx/2+1+size), centred in y and z."""
    nx, ny, nz = shape
    m = np.zeros(shape, bool)
    x0 = nx // 2 + 1
    m[x0:x0 + size, ny // 2 - size // 2:ny // 2 + size // 2, nz // 2 - size // 2:nz // 2 + size // 2] = True
    return m.ravel()


def _probe_ke(fields, mask):
    n = mask.sum()
    u, v, w = (np.asarray(fields[k], np.float64) for k in range(3))
    seed = np.zeros((4, u.size))
    for k, a in enumerate((u, v, w)):
        seed[k][mask] = a[mask] / n
    return 0.5 * float(np.sum((u * u + v * v + w * w)[mask])) / n, seed


def adjoint_workload(ndir: int = 16, steps: int = 200, every: int = 20, n: int = 96):
    """The inlet-control gradient workload: 96^3 channel with a seeded cylinder,
    loss = mean kinetic energy over a 16^3 downstream probe after 200 steps,
    gradient w.r.t. the 32x32 inlet control and nu; FP32 forward and adjoint
    timed (adjoint = checkpointed backward including the recompute; one reversed
    step is timed as the difference of two fully taped backwards), gradients
    checked against central differences (eps 1e-3) of the FP64 forward along
    `ndir` seeded directions, and against the FP64 adjoint."""
    from cmcx import domains
    from cmcx.solver import make_solver

    d, control = domains.cylinder_channel(n=n, steps=steps)
    G = d.num_global
    mask = _probe(d.shape)
    ids = np.arange(d.ny * d.nz)
    nc = control.shape[0]

    def inlet(c):
        vx = d.vx.copy()
        vx[ids] = np.kron(c, np.ones((3, 3)))[:d.ny, :d.nz].ravel()
        return vx

    def block_sum(g):
        g = g[ids].reshape(d.ny, d.nz)
        return g[:3 * nc, :3 * nc].reshape(nc, 3, nc, 3).sum(axis=(1, 3))

    def grads(dtype, timed):
        s = make_solver(d, dtype, path=1)
        if timed:
            s.autotune(3)
        init = s.fields()
        st0 = [init[0], init[1], init[2], init[4]]
        out = {}
        if timed:
            for _ in range(3):  # warm-up
                s.UpdateBoundaries()
                s.TimeStep(d.dt, G, 1, False)
            s.set_fields(*st0)
            s.synchronize()
            _sync()
            t0 = time.perf_counter()
            for _ in range(steps):
                s.UpdateBoundaries()
                s.TimeStep(d.dt, G, 1, False)
            s.synchronize()
            _sync()
            out["forward_s"] = time.perf_counter() - t0
            s.set_fields(*st0)
        _sync()
        t0 = time.perf_counter()
        s.adjoint_forward(steps, d.dt, G, every)
        s.synchronize()
        _sync()
        out["recorded_forward_s"] = time.perf_counter() - t0
        loss, seed = _probe_ke(s.fields(), mask)
        # backward with the seed and the gradients in device memory (a caller
        # whose loss lives on the GPU); the host-array call is timed beside it
        dev = DeviceFields(4, d.n, s.np_dtype)
        sd = dev.upload(seed)
        gdev = DeviceFields(4, d.n, s.np_dtype)
        _sync()
        t0 = time.perf_counter()
        pb = s.adjoint_backward_device(sd, gdev.ptrs, None)
        _sync()
        out["backward_s"] = time.perf_counter() - t0
        gb0 = gdev.download(0)
        dev.free()
        gdev.free()
        t0 = time.perf_counter()
        gbh, _, pbh = s.adjoint_backward(seed)
        _sync()
        out["backward_host_s"] = time.perf_counter() - t0
        assert np.array_equal(gbh[0], gb0) and np.array_equal(pbh, pb)
        out["loss"] = loss
        out["gc"] = block_sum(gb0.astype(np.float64))
        out["gnu"] = float(pb[1])
        s.adjoint_release()
        if timed:
            # the reverse steps alone: every step taped during the forward (no
            # checkpoint recompute), backwards of `steps` and `steps // 2` steps;
            # their difference removes the per-call part (seed upload, gradient
            # download), leaving the cost of one reversed step
            def taped_backward(n):
                s.set_fields(*st0)
                s.adjoint_forward(n, d.dt, G, n)
                s.synchronize()
                _sync()
                t0 = time.perf_counter()
                s.adjoint_backward(seed)
                _sync()
                t = time.perf_counter() - t0
                s.adjoint_release()
                return t

            half_n = steps // 2
            t_full, t_half = taped_backward(steps), taped_backward(half_n)
            out["reverse_step_s"] = (t_full - t_half) / (steps - half_n)
            out["reverse_fixed_s"] = t_full - steps * out["reverse_step_s"]

        def L(c, nu):
            p = d.params()
            p.v_vis = nu
            s.set_params(p)
            s.set_boundary(d.bc_vel, d.bc_temp, inlet(c), d.vy, d.vz, d.T)
            s.set_fields(*st0)
            for _ in range(steps):
                s.UpdateBoundaries()
                s.TimeStep(d.dt, G, 1, False)
            return _probe_ke(s.fields(), mask)[0]

        out["L"] = L
        out["solver"] = s
        return out

    r32 = grads("f32", True)
    r64 = grads("f64", False)
    nu0 = d.params().v_vis
    rng = np.random.default_rng(0)
    eps = 1e-3
    e32, e64 = [float("nan")], [float("nan")]
    if ndir:
        e32, e64 = [], []
    for _ in range(ndir):
        dc = rng.normal(size=control.shape)
        dn = rng.normal() * nu0
        fd = (r64["L"](control + eps * dc, nu0 + eps * dn)
              - r64["L"](control - eps * dc, nu0 - eps * dn)) / (2 * eps)
        a64 = float(np.sum(r64["gc"] * dc) + r64["gnu"] * dn)
        a32 = float(np.sum(r32["gc"] * dc) + r32["gnu"] * dn)
        e64.append(abs(a64 - fd) / abs(fd))
        e32.append(abs(a32 - fd) / abs(fd))
    for r in (r32, r64):
        r["solver"].close()
    return {
        "workload": f"{n}^3 channel, seeded z-cylinder r=12, inlet {nc}x{nc} control (3x3 node "
                    f"blocks), nu={nu0:g}, dt={d.dt:g}, {steps} steps, G={G}, checkpoint every {every}, "
                    "loss = mean KE over a 16^3 probe",
        "forward_ms": r32["forward_s"] * 1e3,
        "recorded_forward_ms": r32["recorded_forward_s"] * 1e3,
        "backward_ms": r32["backward_s"] * 1e3,
        "backward_host_arrays_ms": r32["backward_host_s"] * 1e3,
        "adjoint_forward_ratio": r32["backward_s"] / r32["forward_s"],
        "reverse_step_ms": r32["reverse_step_s"] * 1e3,
        "forward_step_ms": r32["forward_s"] / steps * 1e3,
        "adjoint_step_ratio": r32["reverse_step_s"] / (r32["forward_s"] / steps),
        "backward_call_fixed_ms": r32["reverse_fixed_s"] * 1e3,
        "ratio_note": "adjoint_forward_ratio = checkpointed backward (every 20, every interval "
                      "but the last recomputed; seed and gradients in device memory) / forward "
                      "(200 steps, direct launches); backward_host_arrays_ms = the same backward "
                      "taking and returning host arrays; adjoint_step_ratio = one reversed step "
                      "(all steps taped; difference of backwards over 200 and 100 steps) / one "
                      "forward step",
        "loss_fp32": r32["loss"],
        "loss_fp64": r64["loss"],
        "fd_directions": ndir,
        "fd_eps": eps,
        "max_rel_err_fp64_adjoint_vs_fd": max(e64),
        "max_rel_err_fp32_adjoint_vs_fd": max(e32),
        "median_rel_err_fp32_adjoint_vs_fd": float(np.median(e32)),
    }


def channel2d_workload():
    """The 2D channel (256 x 128, cylinder r = 16, parabolic inlet, Re 100,
    dt 1e-3, 200 steps, G = 4, FP64): the GPU 2D path timed, against the 2D
    oracle (the reference's path runs on the CPU) timed on the host, with the
    final-state error between them."""
    import oracle_lib
    from cmcx import domains
    from cmcx.params import FluidParams
    from cmcx.solver import AdiSolver2D

    d = domains.channel2d()
    p = FluidParams.normalized(d.re, 0.72, 1.4)
    s = AdiSolver2D("f64").Init(d, p)
    _sync()
    t0 = time.perf_counter()
    iters = 0
    for _ in range(d.steps):
        s.UpdateBoundaries()
        iters += s.TimeStep(d.dt, d.num_global, 1)[1]
    _sync()
    gpu_s = time.perf_counter() - t0
    o = oracle_lib.Oracle2D(d, "f64", params=p)
    t0 = time.perf_counter()
    for _ in range(d.steps):
        o.update_boundaries()
        o.step(d.dt, d.num_global, 1)
    cpu_s = time.perf_counter() - t0
    u, v, _, T = s.fields()
    ou, ov, oT = o.layer(0)
    err = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
              for a, b in ((u, ou), (v, ov), (T, oT)))
    s.close()
    cells = d.nx * d.ny
    return {
        "workload": f"2D channel {d.nx}x{d.ny}, cylinder r=16, Re {d.re:g}, dt {d.dt:g}, "
                    f"{d.steps} steps, G={d.num_global}, fp64",
        "gpu_ms_per_step": gpu_s / d.steps * 1e3,
        "gpu_cell_updates_per_s": cells * d.steps / gpu_s,
        "cpu_ms_per_step": cpu_s / d.steps * 1e3,
        "cpu_kind": "port (oracle, 1 thread)",
        "picard_iterations": iters,
        "max_rel_l2_vs_oracle": err,
    }


def stable2d_workload(steps: int = 2):
    """The reference's 2D box_pipe example with its config (Stable solver,
    120 x 135 cells, dx 0.7 mm, 30 steps per 0.035 s frame, num_global 2,
    viscosity 0.05, density 1000, FP64): the first steps of its driver loop
    (Prepare(t), UpdateBoundaries, TimeStep) on the GPU, grid built on the GPU,
    and the same steps of the oracle on the host; states compared bit for bit.
    The relaxation is bound by its chain of nx + ny + 1 barrier-separated
    wavefront steps per sweep, not by HBM or MFMA."""
    import oracle_lib
    from cmcx.geometry import Grid2D
    from cmcx.params import FluidParams
    from cmcx.solver import StableSolver2D

    path = os.path.join(ROOT, "tests", "golden", "geometry2d", "box_pipe_data.txt")
    g = Grid2D(path, 0.0007, 0.0007, startT=1.0)
    dt = g.cycle_length / (g.num_frames * 30)
    d = g.Prepare(0.0)
    p = FluidParams.dimensional(0.05, 1000.0, 287.0, 0.0257, 1000.0)
    grids = [g.arrays((k + 1) * dt) for k in range(steps)]
    s = StableSolver2D("f64").Init(d, p)
    gpu_s, sweeps = 0.0, 0
    for k in range(steps):
        s.SetGrid(*grids[k])
        s.UpdateBoundaries()
        _sync()
        t0 = time.perf_counter()
        s.TimeStep(dt, 2, 1)
        _sync()
        gpu_s += time.perf_counter() - t0
        sweeps += int(s.poisson_sweeps().sum())
    o = oracle_lib.Oracle2D(d, "f64", params=p, method="stable")
    cpu_s = 0.0
    for k in range(steps):
        o.set_grid(*grids[k])
        o.update_boundaries()
        t0 = time.perf_counter()
        o.step(dt, 2, 1)
        cpu_s += time.perf_counter() - t0
    u, v, _, T = s.fields()
    ou, ov, oT = o.layer(0)
    same = bool(np.array_equal(u, ou) and np.array_equal(v, ov) and np.array_equal(T, oT))
    wave = g.dimx + g.dimy + 1
    s.close()
    g.close()
    return {
        "workload": f"box_pipe 2D example {g.dimx}x{g.dimy}, Stable solver, dt {dt:.4g}, "
                    f"first {steps} steps, G=2, fp64",
        "relaxation_sweeps": sweeps,
        "gpu_ms_per_step": gpu_s / steps * 1e3,
        "gpu_us_per_sweep": gpu_s / sweeps * 1e6,
        "wavefront_steps_per_sweep": wave,
        "gpu_ns_per_wavefront_step": gpu_s / sweeps / wave * 1e9,
        "cpu_ms_per_step": cpu_s / steps * 1e3,
        "cpu_us_per_sweep": cpu_s / sweeps * 1e6,
        "cpu_kind": "port (oracle, 1 thread; the reference's relaxation is serial)",
        "bitwise_equal_to_oracle": same,
        "bound": "nx + ny + 1 barrier-separated wavefront steps per sweep, sweeps in groups of up to 8 in one work-group",
    }


def measured_traffic(size: int, path: str):
    """HBM bytes per sweep launch from the committed PMC summary of this
    workload (FETCH_SIZE and WRITE_SIZE runs, scaled by a plain copy), or None."""
    fn = os.path.join(ROOT, "profiles", "pmc_sweeps_ellipsoids256_fp32.md")
    if size != 256 or path != "fused" or not os.path.exists(fn):
        return None
    for line in open(fn):
        if line.startswith("traffic_bytes_per_sweep_launch:"):
            return float(line.split(":")[1])
    return None


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _nproc() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(dom, budget_s: float = 20.0):
    """Oracle (C++/OpenMP port of the reference CPU path) on a bounded sample:
    whole FP32 time steps of the same workload on all the OpenMP threads until
    about `budget_s` is spent, then one step on one thread."""
    import oracle_lib

    L = oracle_lib.load()
    threads = int(L.orc_max_threads())
    o = oracle_lib.Oracle3D(dom, "f32")
    o.create_segments()
    n = 0
    t0 = time.perf_counter()
    while True:
        o.update_boundaries()
        o.step(dom.dt, dom.num_global, dom.num_local, False)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    wall = time.perf_counter() - t0
    # one more step on one thread
    L.orc_set_threads(1)
    t1 = time.perf_counter()
    o.update_boundaries()
    o.step(dom.dt, dom.num_global, dom.num_local, False)
    wall1 = time.perf_counter() - t1
    L.orc_set_threads(threads)
    o.close()
    return {
        "value": dom.n * n / wall,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} full FP32 time step(s) of the same 256^3 workload (G=4, L=1), {wall:.1f} s, "
                  f"OpenMP on {threads} threads; then 1 step on 1 thread, {wall1:.1f} s",
        "value_1thread": dom.n / wall1,
        "nproc": _nproc(),
        "cpu_model": _cpu_model(),
    }


def headline_parity(dom, variants, steps: int = 3):
    """The headline path (FP32, the timed tile variants) against the FP64 oracle
    after `steps` steps from the same initial state: relative L2 errors per
    component, each relative to its own signal (u, v, w; T - baseT; p - v_T
    baseT) and velocity-norm-scaled."""
    import oracle_lib
    from cmcx.solver import make_solver
    from gpu_util import ROUNDOFF_FP32, component_errors

    s = make_solver(dom, "f32", path=1)
    for a, v in enumerate(variants):
        s.set_tile_variant(a, v)
    for _ in range(steps):
        s.UpdateBoundaries()
        s.TimeStep(dom.dt, dom.num_global, dom.num_local, False)
    u, v, w, _, T = s.fields()
    s.close()
    o = oracle_lib.Oracle3D(dom, "f64")
    o.create_segments()
    for _ in range(steps):
        o.update_boundaries()
        o.step(dom.dt, dom.num_global, dom.num_local, False)
    ref = o.fields()
    o.close()
    e = component_errors((u, v, w, T), ref, dom.base_t, dom.params().v_T, ROUNDOFF_FP32)
    return {
        "steps": steps,
        "tile_variants": list(variants),
        "own": {k: x[0] for k, x in e.items()},
        "scaled": {k: x[1] for k, x in e.items()},
        "max_rel_l2_vs_fp64_oracle": max(e[k][1] for k in ("u", "v", "w", "p")),
        "note": "scaled: velocity components over the velocity-field norm, T and p over their own "
                "norms (the north-star metric, bar 1e-5); own: each component over its own signal "
                "(T - baseT, p - v_T baseT); None = zero up to FP32 rounding",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", choices=["fused", "plain"], default=None)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-2d", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the inlet-control gradient workload and print its JSON")
    ap.add_argument("--adjoint-dirs", type=int, default=16,
                    help="finite-difference directions of the gradient check")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--pmc-mode", action="store_true",
                    help="only run the steps (no events, no autotuning, no JSON line)")
    ap.add_argument("--tile-variants", default="0,1,0",
                    help="pmc-mode: fused kernel variants of the X, Y, Z sweeps (the autotuner's "
                         "choice on the 256^3 workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (reference kernel for "
                         "PMC byte counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the whole domain on a single "
              "MI355X (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cmcx import domains
    from cmcx._lib import CMCX_PATH_FUSED, CMCX_PATH_PLAIN, lib

    lib()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.adjoint_only:
        print(json.dumps(adjoint_workload(ndir=args.adjoint_dirs)), flush=True)
        return
    dom = domains.ellipsoids(args.size)
    fused_ok = True
    try:
        from cmcx.solver import make_solver

        probe = make_solver(domains.box(8, 8, 8), "f32", path=CMCX_PATH_FUSED)
        probe.close()
    except Exception:
        fused_ok = False
    head_path = args.path or ("fused" if fused_ok else "plain")
    if args.pmc_mode:
        run_plain_steps(dom, CMCX_PATH_FUSED if head_path == "fused" else CMCX_PATH_PLAIN,
                        args.steps, args.warmup,
                        variants=[int(v) for v in args.tile_variants.split(",")])
        print(f"pmc-mode: {args.steps} steps of the {head_path} path done", flush=True)
        return
    head_id = CMCX_PATH_FUSED if head_path == "fused" else CMCX_PATH_PLAIN
    direct = time_path(dom, head_id, args.steps, args.warmup)
    graphed = time_path(dom, head_id, args.steps, args.warmup, tune=False, graph=True)
    head = graphed if graphed["cell_updates_per_s"] >= direct["cell_updates_per_s"] else direct
    head["variants"] = direct["variants"]
    plain = None
    if head_path == "fused" and not args.no_plain:
        plain = time_path(dom, CMCX_PATH_PLAIN, max(2, args.steps // 4), 1)
    batch = None if args.no_batch else time_batch(max(4, args.steps // 2), 2)
    adjoint = None if args.no_adjoint else adjoint_workload(ndir=args.adjoint_dirs)
    two_d = None if args.no_2d else channel2d_workload()
    stable = None if args.no_2d else stable2d_workload()
    cpu = None if args.no_cpu else cpu_baseline(dom, args.cpu_budget)
    parity = (None if args.no_parity or head_path != "fused"
              else headline_parity(dom, head["variants"]))

    achieved = head["sweep_achieved_gbs"]
    line = {
        "metric": METRIC,
        "value": head["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"ellipsoids_{args.size}^3_fp32",
            "grid": [dom.nx, dom.ny, dom.nz],
            "ellipsoids": 40,
            "seed": 0,
            "solid_fraction": round(dom.solid_fraction(), 4),
            "Re": dom.re,
            "dt": dom.dt,
            "num_global": dom.num_global,
            "num_local": dom.num_local,
            "path": head_path,
            "tile_variants": head["variants"],
            "graph": head is graphed,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.size, head_path),
            "traffic_source": "profiles/pmc_sweeps_ellipsoids256_fp32.md",
            "kernel": "sweep",
            "bytes_per_step": head["sweep_bytes_per_step"],
            "sweep_ms_per_launch": head["sweep_ms_per_launch"],
        },
        "cpu_baseline": cpu,
    }
    line["step_launch"] = {
        "direct_ms_per_step": direct["ms_per_step"],
        "graph_ms_per_step": graphed["ms_per_step"],
    }
    if batch is not None:
        bf, bp = batch["fused"], batch["plain"]
        line["batch"] = {
            "workload": f"{batch['grids']} boxes, edges from {{32,48,64,96}}, seed {batch['seed']}, fp32",
            "cells": batch["cells"],
            "value": bf["cell_updates_per_s"],
            "unit": "cell-updates/s",
            "ms_per_step": bf["ms_per_step"],
            "sweep_achieved_gbs": bf["sweep_achieved_gbs"],
            "hbm_frac": bf["sweep_achieved_gbs"] / HBM_PEAK_GBS,
            "tile_variants": bf["variants"],
            "plain_value": bp["cell_updates_per_s"],
            "speedup_over_plain": bf["cell_updates_per_s"] / bp["cell_updates_per_s"],
        }
    if adjoint is not None:
        line["adjoint"] = adjoint
    if two_d is not None:
        line["channel2d"] = two_d
    if stable is not None:
        line["stable2d"] = stable
    if parity is not None:
        line["parity"] = parity
    if plain is not None:
        line["plain"] = {
            "value": plain["cell_updates_per_s"],
            "ms_per_step": plain["ms_per_step"],
            "sweep_achieved_gbs": plain["sweep_achieved_gbs"],
            "speedup_fused_over_plain": head["cell_updates_per_s"] / plain["cell_updates_per_s"],
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
