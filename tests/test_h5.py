"""The reference's HDF5 model and kernel files (host code, CPU only).

The layout is checked with the HDF5 tools' own reader (h5dump), not only by
reading back through the library that wrote it: the reference's Fortran
writers (src/shared/hdf5_interface.f90:362-417) store a Fortran array of
shape (n1, ..., nk) as a dataset of C dimensions (nk, ..., n1) in memory
order, and every model array goes through transpose_k first
(src/shared/utils.f90:820-850), so /vel of svel(nperiod, nx, ny)
(src/acqui_2d.f90:242) has C dimensions (nperiod, nx, ny) with element
[p][i][j] = svel(p, i, j).  The test arrays hold distinct values per index
(100p + 10i + j), so a swapped or unreversed axis fails."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import satt_h5io as H

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
pytestmark = pytest.mark.skipif(not os.path.exists(H.LIB_PATH), reason="libsatt_h5.so not built")


def dump(path, name):
    """(C dims, values in C order) of a dataset, from h5dump."""
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    out = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", "-m", "%.17g", path],
                         capture_output=True, text=True, check=True).stdout
    dims = tuple(int(v) for v in re.search(r"DATASPACE\s+SIMPLE\s*\{\s*\(([^)]*)\)", out)
                 .group(1).split(","))
    data = out[out.index("DATA {") + 6:]
    data = data[:data.index("}")]
    vals = np.array([float(v) for v in re.split(r"[,\s]+", data.strip()) if v])
    return dims, vals


def idx3(a, b, c):
    return np.fromfunction(lambda p, i, j: 100.0 * p + 10.0 * i + j, (a, b, c))


def test_model_2d_layout(tmp_path):
    np_, nx, ny = 3, 5, 4
    svel = np.asfortranarray(idx3(np_, nx, ny))
    periods, lon, lat = np.array([5.0, 10.0, 20.0]), 100.0 + 0.1 * np.arange(nx), 20.0 + 0.1 * np.arange(ny)
    stlo, stla = np.array([100.1, 100.2]), np.array([20.1, 20.3])
    f = str(tmp_path / "final_model_ph.h5")
    H.write_model_2d(f, svel, periods, lon, lat, stlo, stla)
    dims, vals = dump(f, "/vel")
    assert dims == (np_, nx, ny)
    assert np.array_equal(vals.reshape(dims), svel)          # [p][i][j] = svel(p, i, j)
    for name, ref in (("/period", periods), ("/lon", lon), ("/lat", lat), ("/stlo", stlo),
                      ("/stla", stla)):
        d, v = dump(f, name)
        assert d == (len(ref),) and np.array_equal(v, ref), name
    with H.File(f, "old", "read") as h:
        assert h.shape("/vel") == (ny, nx, np_)              # the Fortran view: transpose_3(svel)
        assert np.array_equal(h.get("/vel"), svel.transpose(2, 1, 0))


def test_target_model_has_no_stations(tmp_path):
    f = str(tmp_path / "target_model_ph.h5")
    H.write_model_2d(f, np.ones((2, 3, 4)), [4.0, 8.0], np.arange(3.0), np.arange(4.0))
    with H.File(f, "old", "read") as h:
        assert not h.exist("/stlo") and not h.exist("/stla")
        assert h.exist("/vel") and h.exist("/period")


def test_model_iter_file(tmp_path):
    f = str(tmp_path / "model_iter_ph.h5")
    sv0, sv1 = idx3(2, 3, 4), idx3(2, 3, 4) + 0.5
    ax = dict(periods=[6.0, 12.0], lon=np.arange(3.0), lat=np.arange(4.0))
    with H.File(f, "new", "write") as h:
        h.write_iter_2d(0, "ph", sv0, stlo=[1.0], stla=[2.0], **ax)
        h.write_iter_2d(1, "ph", sv1, **ax)
        with pytest.raises(H.H5Error) as e:           # a record is written once
            h.write_iter_2d(1, "ph", sv1, **ax)
        assert e.value.status == H.E_ARG
    d0, v0 = dump(f, "/velph_000")
    d1, v1 = dump(f, "/velph_001")
    assert d0 == d1 == (2, 3, 4)
    assert np.array_equal(v0.reshape(d0), sv0) and np.array_equal(v1.reshape(d1), sv1)
    assert dump(f, "/stlo")[1].tolist() == [1.0]


def test_model_3d_file_and_initial_model(tmp_path):
    nx, ny, nz, nker = 4, 3, 5, 3
    vs = np.asfortranarray(idx3(nx, ny, nz))
    f = str(tmp_path / "model_iter.h5")
    with H.File(f, "new", "write") as h:
        h.write_model_3d_init(np.arange(nx) + 100.0, np.arange(ny) + 20.0, np.arange(nz) * 2.0, vs,
                              vs * 1.7, vs * 0.8)
        h.write_model_3d_iter(1, vs + 1)
        g = np.fromfunction(lambda k, i, j, l: 1000.0 * k + 100.0 * i + 10.0 * j + l,
                            (nker, nx, ny, nz))
        h.write_gradient_3d(0, g)
        h.write_kdensity_3d(0, "ph", g * 2)
    dims, vals = dump(f, "/vs_000")
    assert dims == (nx, ny, nz)
    # transpose_3(vs3d) written from Fortran: C order [i][j][k] = vs3d(i, j, k)
    assert np.array_equal(vals.reshape(dims), vs)
    assert dump(f, "/vp_000")[0] == (nx, ny, nz) and dump(f, "/dep")[0] == (nz,)
    dg, vg = dump(f, "/gradient_000")
    assert dg == (nker, nx, ny, nz) and np.array_equal(vg.reshape(dg), g)
    assert np.array_equal(dump(f, "/kdensity_ph_000")[1].reshape(dg), 2 * g)
    # init_model_type 2 reads a /vs of this layout back as vs3d (nx, ny, nz)
    assert np.array_equal(H.read_model_3d(f, (nx, ny, nz), "/vs_001"), vs + 1)
    with pytest.raises(H.H5Error) as e:       # a shape that does not match: the reference stops
        H.read_model_3d(f, (ny, nx, nz), "/vs_001")
    assert e.value.status == H.E_ARG
    with pytest.raises(H.H5Error):
        H.read_model_3d(f, (nx, ny, nz), "/vs")  # no such dataset


def test_generic_add_is_fortran_order(tmp_path):
    f = str(tmp_path / "x.h5")
    x = np.asfortranarray(np.arange(6.0).reshape(2, 3))       # x(2,3)
    with H.File(f, "new", "write") as h:
        h.add("/grp/sub/x", x)                                 # parent groups created
    dims, vals = dump(f, "/grp/sub/x")
    assert dims == (3, 2)                                      # reversed
    assert np.array_equal(vals, x.ravel(order="F"))            # memory order


def test_library_exports_every_header_symbol():
    import ctypes
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "include", "satt_h5.h")).read()
    names = sorted(set(re.findall(r"^int (satt_h5_\w+)\(", hdr, re.M)))
    assert len(names) == 14
    lib = ctypes.CDLL(H.LIB_PATH)
    assert [n for n in names if not hasattr(lib, n)] == []
