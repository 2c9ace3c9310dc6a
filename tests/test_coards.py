"""The COARDS writer (cmcx_coards_write_header / _layer, the reference's
OutputNetCDF3D_header / OutputNetCDF3D_layer, IO.h:136-388) read back with an
independent netCDF reader (scipy.io.netcdf_file): dimensions, variables,
types, attributes, axis and time values, fill values, layer data, records
appended past the announced ones, and argument errors.  Host code only."""
import numpy as np
import pytest

from cmcx import _lib
from cmcx.io import OutputNetCDF3D_header, OutputNetCDF3D_layer

netcdf = pytest.importorskip("scipy.io").netcdf_file

FILL = 9.9692099683868690e+36


def _read(path):
    return netcdf(str(path), "r", mmap=False, maskandscale=False)


def test_header_layout(tmp_path):
    p = tmp_path / "res.nc"
    OutputNetCDF3D_header(p, (0.0, -1.0, 0.0), (2.0, 1.0, 0.5), 0.25, 1.0, 4, 5, 3, ["u", "v", "w", "T"])
    with _read(p) as f:
        assert f.version_byte == 2
        assert f.dimensions == {"x": 4, "y": 5, "z": 3, "t": None}
        assert list(f.variables) == ["x", "y", "z", "time", "u", "v", "w", "T"]
        x, y, z, t = (f.variables[k] for k in ("x", "y", "z", "time"))
        assert x.data.dtype.kind == "f" and x.data.dtype.itemsize == 4
        np.testing.assert_array_equal(x.data, np.float32(0) + np.float32(0.5) * np.arange(4, dtype=np.float32))
        np.testing.assert_array_equal(y.data, np.float32(-1) + np.float32(0.4) * np.arange(5, dtype=np.float32))
        assert x.actual_range.tolist() == [0.0, 2.0] and x.units == b"metres"
        assert x.long_name == b"x coord" and z.long_name == b"z coord"
        assert t.units == b"s" and t.actual_range.tolist() == [0.0, 1.0]
        np.testing.assert_array_equal(t.data, [0.0, 0.25, 0.5, 0.75])
        for k, lng, units in (("u", b"x-velocity", b"m/s"), ("T", b"temperature", b"tmp")):
            v = f.variables[k]
            assert v.dimensions == ("t", "x", "y", "z")
            assert v.data.dtype.itemsize == 8 and v.data.shape == (4, 4, 5, 3)
            assert v.long_name == lng and v.units == units and v.var_desc == k.encode()
            assert np.atleast_1d(v.missing_value).tolist() == [99999.0]
            assert v.valid_range.tolist() == [-1.0, 1.0]
            assert np.all(v.data == FILL)
        assert f.Conventions == b"COARDS"
        assert f.title == b"cmc-fluid-solver results"


def test_layers_round_trip(tmp_path):
    p = tmp_path / "res.nc"
    nx, ny, nz = 3, 4, 5
    # the header announces int(time / timestep) records, as the reference
    OutputNetCDF3D_header(p, (0, 0, 0), (1, 1, 1), 0.25, 0.75, nx, ny, nz, "u w T")
    rng = np.random.default_rng(0)
    layers = []
    for t in range(5):  # records 3 and 4 are appended past the announced three
        vel = rng.normal(size=(nx * ny * nz, 3)).astype(np.float32)
        T = 1.0 + 0.1 * rng.normal(size=nx * ny * nz)
        OutputNetCDF3D_layer(p, vel, T, t, nx, ny, nz, ["u", "w", "T"])
        layers.append((vel, T))
    with _read(p) as f:
        assert "v" not in f.variables
        u, w, T = f.variables["u"].data, f.variables["w"].data, f.variables["T"].data
        assert u.shape == (5, nx, ny, nz)
        for t, (vel, Tt) in enumerate(layers):
            np.testing.assert_array_equal(u[t].ravel(), vel[:, 0].astype(np.float64))
            np.testing.assert_array_equal(w[t].ravel(), vel[:, 2].astype(np.float64))
            np.testing.assert_array_equal(T[t].ravel(), Tt)
        time = f.variables["time"].data
        np.testing.assert_array_equal(time[:3], [0.0, 0.25, 0.5])
        assert np.all(time[3:] == FILL)


def test_layer_gap_and_depth(tmp_path):
    p = tmp_path / "res.nc"
    depth = -np.arange(6, dtype=np.float32).reshape(2, 3)
    OutputNetCDF3D_header(p, (0, 0, -5), (1, 1, 0), 1.0, 1.0, 2, 3, 2, ["T", "d"], xy_degree_units=True,
                          depth=depth)
    vel = np.zeros((12, 3), np.float64)
    T = np.arange(12, dtype=np.float64)
    OutputNetCDF3D_layer(p, vel, T, 2, 2, 3, 2, ["T", "d"])  # record 1 stays at the fill value
    with _read(p) as f:
        np.testing.assert_array_equal(f.variables["d"].data, depth)
        assert f.variables["d"].dimensions == ("x", "y")
        assert f.variables["x"].units == b"degree_north" and f.variables["y"].units == b"degree_east"
        T3 = f.variables["T"].data
        assert T3.shape[0] == 3
        assert np.all(T3[1] == FILL)
        np.testing.assert_array_equal(T3[2].ravel(), T)


def test_errors(tmp_path):
    p = tmp_path / "res.nc"
    with pytest.raises(_lib.CmcxError):
        OutputNetCDF3D_header(p, (0, 0, 0), (1, 1, 1), 0.1, 1.0, 2, 2, 2, ["u", "q"])
    with pytest.raises(_lib.CmcxError):
        OutputNetCDF3D_header(p, (0, 0, 0), (1, 1, 1), 0.0, 1.0, 2, 2, 2, ["u"])
    OutputNetCDF3D_header(p, (0, 0, 0), (1, 1, 1), 0.5, 1.0, 2, 2, 2, ["u"])
    with pytest.raises(_lib.CmcxError):  # v was not defined in the header
        OutputNetCDF3D_layer(p, np.zeros((8, 3)), np.zeros(8), 0, 2, 2, 2, ["v"])
    with pytest.raises(_lib.CmcxError):  # wrong layer shape
        OutputNetCDF3D_layer(p, np.zeros((27, 3)), np.zeros(27), 0, 3, 3, 3, ["u"])
