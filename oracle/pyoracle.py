"""ctypes wrapper of the CPU oracle (oracle/libhpg_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, as the checker (or, for the OpenMP form,
as the timed CPU baseline).  The product package never imports it.

Array conventions (all host numpy arrays):
  * visibilities: structured arrays of hpg::VisData<N> records
    (src/hpg.hpp:236-311), see ``visdata_dtype``;
  * CF group values: complex64, C-order shape (cube, mueller, y, x), i.e.
    Layout::Left over the CFArray axes (x, y, mueller, cube);
  * grid / model: complex128, C-order shape (cube, mrow, y, x);
  * weights: float64, C-order shape (cube, mrow).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhpg_oracle.so")
_lib = None


def _align8(n):
    return (n + 7) & ~7


def visdata_dtype(npol):
    """numpy dtype with the memory layout of hpg::VisData<npol>."""
    off_f = _align8(12 * npol)
    return np.dtype({
        "names": ["vis", "weights", "frequency", "phase", "uvw", "grid_cube",
                  "cf_index", "cf_phase_gradient"],
        "formats": [("<c8", (npol,)), ("<f4", (npol,)), "<f8", "<f8",
                    ("<f8", (3,)), "<u4", ("<u4", (2,)), ("<f8", (2,))],
        "offsets": [0, 8 * npol, off_f, off_f + 8, off_f + 16, off_f + 40,
                    off_f + 44, off_f + 56],
        "itemsize": off_f + 72,
    })


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"{_LIB_PATH} is missing; run `make -C oracle` first")
        L = ctypes.CDLL(_LIB_PATH)
        vp, u, d, i = ctypes.c_void_p, ctypes.c_uint, ctypes.c_double, ctypes.c_int
        sz = ctypes.c_size_t
        L.hpgo_create.restype = vp
        L.hpgo_create.argtypes = [vp, vp, u, vp, vp]
        L.hpgo_destroy.argtypes = [vp]
        L.hpgo_set_cf.argtypes = [vp, u, u, vp, vp]
        L.hpgo_set_model.argtypes = [vp, vp]
        L.hpgo_grid_visibilities.restype = i
        L.hpgo_grid_visibilities.argtypes = [vp, vp, sz, u, i, i, i, i]
        L.hpgo_normalize_by_weights.argtypes = [vp, d]
        L.hpgo_fft.argtypes = [vp, i, d, i]
        L.hpgo_shift.argtypes = [vp, i, i]
        L.hpgo_reset_grid.argtypes = [vp]
        L.hpgo_grid_values.argtypes = [vp, vp]
        L.hpgo_grid_weights.argtypes = [vp, vp]
        L.hpgo_grid_window.argtypes = [vp, u, u, u, u, u, u, vp]
        L.hpgo_model_values.argtypes = [vp, vp]
        L.hpgo_cone_cf.argtypes = [u, u, u, vp]
        L.hpgo_num_threads.restype = i
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def cone_cf(nmueller, oversampling, radius):
    """ConeCFArray of the reference's tests (GridderStateUnitTests.cc:198-253):
    one group, one cube; returns (extents, values[cube, mueller, y, x])."""
    w = 2 * (radius + 2) * oversampling + 1
    out = np.zeros((1, nmueller, w, w), dtype=np.complex64)
    lib().hpgo_cone_cf(nmueller, oversampling, radius, _ptr(out))
    return (w, w, nmueller, 1), out


def num_threads():
    return lib().hpgo_num_threads()


class Oracle:
    """FP64 Serial restatement of one hpg GridderState (CPU)."""

    def __init__(self, grid_size, grid_scale, mueller, conj_mueller=None):
        if conj_mueller is None:
            conj_mueller = mueller
        self.grid_size = tuple(int(g) for g in grid_size)
        m = np.ascontiguousarray(np.asarray(mueller, dtype=np.int32))
        mc = np.ascontiguousarray(np.asarray(conj_mueller, dtype=np.int32))
        assert m.shape == mc.shape and m.shape[0] == self.grid_size[2]
        self.npol = m.shape[1]
        gs = np.asarray(self.grid_size, dtype=np.uint32)
        sc = np.asarray(grid_scale, dtype=np.float64)
        self._h = lib().hpgo_create(_ptr(gs), _ptr(sc), self.npol, _ptr(m), _ptr(mc))
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            lib().hpgo_destroy(self._h)
            self._h = None

    def set_cf(self, oversampling, groups):
        """groups: list of (extents(x, y, mueller, cube), values[cube,mueller,y,x])"""
        ext = np.asarray([list(e) for e, _ in groups], dtype=np.uint32).reshape(-1)
        vals = [np.ascontiguousarray(v, dtype=np.complex64) for _, v in groups]
        for (e, _), v in zip(groups, vals):
            assert v.shape == (e[3], e[2], e[1], e[0]), (v.shape, e)
        ptrs = (ctypes.c_void_p * len(vals))(*[v.ctypes.data for v in vals])
        lib().hpgo_set_cf(self._h, oversampling, len(vals), _ptr(ext), ptrs)

    def set_model(self, model):
        if model is None:
            lib().hpgo_set_model(self._h, None)
            return
        m = np.ascontiguousarray(model, dtype=np.complex128)
        assert m.shape == self.grid_shape
        lib().hpgo_set_model(self._h, _ptr(m))

    @property
    def grid_shape(self):
        nx, ny, nr, nc = self.grid_size
        return (nc, nr, ny, nx)

    def grid_visibilities(self, records, update_grid_weights=True,
                          do_degrid=False, do_grid=True, openmp=False):
        """Grids (and/or degrids) a copy of `records`; returns the copy, with
        residual or predicted values where the reference writes them."""
        dt = visdata_dtype(self.npol)
        recs = np.asarray(records)
        assert recs.dtype.names == dt.names, "records are not VisData<%d>" % self.npol
        # (np.concatenate and friends pack structured dtypes: restore the layout)
        recs = np.ascontiguousarray(recs.astype(dt, copy=True))
        rc = lib().hpgo_grid_visibilities(
            self._h, _ptr(recs), recs.shape[0], self.npol, int(update_grid_weights),
            int(do_degrid), int(do_grid), int(openmp))
        if rc != 0:
            raise RuntimeError("oracle grid_visibilities failed")
        return recs

    def normalize_by_weights(self, wfactor=1.0):
        lib().hpgo_normalize_by_weights(self._h, float(wfactor))

    def apply_fft(self, norm=1.0, sign="POSITIVE", which="grid"):
        lib().hpgo_fft(self._h, 1 if which == "model" else 0, float(norm),
                       1 if sign == "NEGATIVE" else 0)

    def shift(self, direction="FORWARD", which="grid"):
        lib().hpgo_shift(self._h, 1 if which == "model" else 0,
                         1 if direction == "BACKWARD" else 0)

    def reset_grid(self):
        lib().hpgo_reset_grid(self._h)

    def grid_values(self):
        out = np.empty(self.grid_shape, dtype=np.complex128)
        lib().hpgo_grid_values(self._h, _ptr(out))
        return out

    def grid_window(self, x0, y0, nx, ny, mrow=0, cube=0):
        out = np.empty((ny, nx), dtype=np.complex128)
        lib().hpgo_grid_window(self._h, x0, y0, nx, ny, mrow, cube, _ptr(out))
        return out

    def grid_weights(self):
        out = np.empty((self.grid_size[3], self.grid_size[2]), dtype=np.float64)
        lib().hpgo_grid_weights(self._h, _ptr(out))
        return out

    def model_values(self):
        out = np.zeros(self.grid_shape, dtype=np.complex128)
        lib().hpgo_model_values(self._h, _ptr(out))
        return out
