"""The C ABI library (hpg_amd/libhpgx.so): loads, exports every symbol that
include/hpgx.h declares, and record layouts agree -- CPU only, no GPU calls
beyond device discovery."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hpgx.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"HPGX_API\s+[\w\s\*]+?\b(hpgx_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 30
    for must in ["hpgx_create", "hpgx_set_convolution_function",
                 "hpgx_grid_visibilities", "hpgx_normalize_by_weights",
                 "hpgx_apply_grid_fft", "hpgx_shift_grid", "hpgx_degrid_adjoint",
                 "hpgx_grid_values", "hpgx_grid_weights", "hpgx_fence"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    from hpg_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding describes every declared symbol, and no others
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared_symbols()
    _lib.lib()  # applies every signature


def test_visdata_layouts_agree():
    import hpg_amd
    from oracle import pyoracle
    for n, size in [(1, 88), (2, 96), (3, 112), (4, 120)]:
        a, b = hpg_amd.visdata_dtype(n), pyoracle.visdata_dtype(n)
        assert a.itemsize == b.itemsize == size
        assert [a.fields[k][1] for k in a.names] == [b.fields[k][1] for k in b.names]

    # the C struct of the header, restated with ctypes
    class V(ctypes.Structure):
        _fields_ = [("vis", ctypes.c_float * 2 * 4), ("w", ctypes.c_float * 4),
                    ("freq", ctypes.c_double), ("phase", ctypes.c_double),
                    ("uvw", ctypes.c_double * 3), ("grid_cube", ctypes.c_uint),
                    ("cf_index", ctypes.c_uint * 2), ("grad", ctypes.c_double * 2)]
    d = hpg_amd.visdata_dtype(4)
    assert ctypes.sizeof(V) == d.itemsize
    assert V.freq.offset == d.fields["frequency"][1]
    assert V.grid_cube.offset == d.fields["grid_cube"][1]
    assert V.grad.offset == d.fields["cf_phase_gradient"][1]


def test_create_argument_errors_without_compute():
    """Validation that precedes any device work (src/hpg.cc:238-263)."""
    from hpg_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    gs = (ctypes.c_uint * 4)(8, 8, 2, 1)
    sc = (ctypes.c_double * 2)(1.0, 1.0)
    m = (ctypes.c_int * 1)(0)
    # one Mueller row for a grid with two mrows
    rc = L.hpgx_create(ctypes.byref(h), 0, 10, gs, sc, 1, 1, m, m, 0, 0, 0, None)
    assert rc == 4  # 1 + ErrorType::InvalidNumberMuellerIndexRows
    assert "Mueller" in _lib.last_create_error()
    if L.hpgx_device_count() == 0:
        gs1 = (ctypes.c_uint * 4)(8, 8, 1, 1)
        rc = L.hpgx_create(ctypes.byref(h), 0, 10, gs1, sc, 1, 1, m, m, 0, 0, 0, None)
        assert rc == 1  # DisabledDevice


def test_python_api_errors_without_compute():
    import hpg_amd
    with pytest.raises(hpg_amd.Error) as e:
        hpg_amd.GridderState.create(hpg_amd.Device.Hip, 0, 10, None, (8, 8, 2, 1),
                                    (1.0, 1.0), [[0]])
    assert e.value.type == hpg_amd.ErrorType.InvalidNumberMuellerIndexRows
    with pytest.raises(hpg_amd.Error) as e:
        hpg_amd.GridderState.create(hpg_amd.Device.Cuda, 0, 10, None, (8, 8, 1, 1),
                                    (1.0, 1.0), [[0]])
    assert e.value.type == hpg_amd.ErrorType.DisabledDevice


def test_workload_shapes():
    from hpg_amd import workloads
    cf = workloads.gaussian_cf()
    assert cf.extents(0) == (220, 220, 1, 1)
    assert cf.radii(0) == (3, 3)
    v = cf.values(0)[0, 0]
    # peak at the centre tap (100, 100), separable and symmetric
    assert np.unravel_index(np.argmax(np.abs(v)), v.shape) == (100, 100)
    assert v[100, 100] == 1.0
    np.testing.assert_allclose(v[100, 130].real, np.exp(-1.5 ** 2 / (2 * 1.5 ** 2)), rtol=1e-6)
    rec = workloads.config1_visibilities(1000)
    assert rec.dtype.itemsize == 88
    assert np.all(np.abs(rec["uvw"][:, :2]) <= 0.45 * 1024)
    assert np.all((rec["weights"] > 0.5) & (rec["weights"] <= 1.5))
