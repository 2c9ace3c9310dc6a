"""COARDS result files over the C ABI: the reference's OutputNetCDF3D_header and
OutputNetCDF3D_layer (src/Common/IO.h:136-388) with the same arguments.

    OutputNetCDF3D_header(path, bbox_min, bbox_max, dt * out_time_steps, finaltime,
                          ox, oy, oz, ["u", "v", "w", "T"])
    vel, T = solver.GetLayer(ox, oy, oz)
    OutputNetCDF3D_layer(path, vel, T, out_layer, ox, oy, oz, ["u", "v", "w", "T"])

The file is netCDF classic with 64-bit offsets (readable by any netCDF reader,
e.g. scipy.io.netcdf_file); see include/cmcx.h for the data model.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def _vars(vars_) -> bytes:
    if isinstance(vars_, str):
        vars_ = vars_.split()
    return " ".join(vars_).encode()


def OutputNetCDF3D_header(path: str, bbox_min, bbox_max, timestep: float, time: float, outdimx: int,
                          outdimy: int, outdimz: int, vars_, xy_degree_units: bool = False,
                          depth=None) -> None:
    """Creates the file: dimensions, variables and attributes, axis values, and
    time/timestep records (time = i * timestep).  depth: d at the output
    resolution (outdimx, outdimy) when "d" is among vars_."""
    h = _lib.cmcx_coards_header()
    h.bbox_min[:] = [float(x) for x in bbox_min]
    h.bbox_max[:] = [float(x) for x in bbox_max]
    h.timestep = float(timestep)
    h.time = float(time)
    h.outdims[:] = [int(outdimx), int(outdimy), int(outdimz)]
    h.vars = _vars(vars_)
    h.xy_degree_units = 1 if xy_degree_units else 0
    keep = None
    if depth is not None:
        keep = np.ascontiguousarray(depth, dtype=np.float32).reshape(-1)
        if keep.size != outdimx * outdimy:
            raise ValueError("depth must have outdimx * outdimy values")
        h.depth = keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    check(lib().cmcx_coards_write_header(str(path).encode(), ctypes.byref(h)), "cmcx_coards_write_header")


def OutputNetCDF3D_layer(path: str, vel, T, time: int, dimx: int, dimy: int, dimz: int, vars_) -> None:
    """Writes output layer `time` (a record index) from GetLayer's arrays:
    vel (n, 3) of float32 or float64, T (n,) float64."""
    vel = np.ascontiguousarray(vel)
    if vel.dtype not in (np.float32, np.float64):
        vel = vel.astype(np.float64)
    if vel.size != 3 * dimx * dimy * dimz:
        raise ValueError("vel must hold 3 values per output node")
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
    if T.size != dimx * dimy * dimz:
        raise ValueError("T must hold one value per output node")
    dt = _lib.CMCX_F32 if vel.dtype == np.float32 else _lib.CMCX_F64
    check(lib().cmcx_coards_write_layer(str(path).encode(), vel.ctypes.data, dt,
                                        T.ctypes.data_as(_lib._dp), int(time), int(dimx),
                                        int(dimy), int(dimz), _vars(vars_)),
          "cmcx_coards_write_layer")
