"""hpg_amd -- HPG's gridding hot path as an MI355X (gfx950) device.

Python mirror of the reference's hpg::GridderState API (src/hpg.hpp:1103-2162)
over the C ABI of include/hpgx.h (hpg_amd/libhpgx.so).  Differences from the
C++ API are only those of the language: methods update the state in place
and raise ``hpg_amd.Error`` (carrying the hpg::ErrorType) instead of
returning ``rval_t``; ``copy()`` gives the reference's value-copy semantics.

Arrays
  * visibilities: numpy structured arrays of ``visdata_dtype(N)`` records
    (the memory layout of hpg::VisData<N>), or a torch uint8 tensor holding
    such records on the GPU (device-resident input, no host copy);
  * CF values: complex64 arrays of C-order shape (cube, mueller, y, x), i.e.
    Layout::Left over the CFArray axes (x, y, mueller, cube);
  * grid and model values: complex128 of C-order shape (cube, mrow, y, x);
  * weights: float64 of C-order shape (cube, mrow).
"""
import ctypes
import enum

import numpy as np

from . import _lib

__all__ = [
    "Device", "ErrorType", "Error", "FFTSign", "ShiftDirection", "GridMode",
    "Layout", "CFArray", "GridderState", "visdata_dtype", "initialize",
    "finalize", "devices", "host_devices",
]


class Device(enum.IntEnum):
    """hpg::Device (src/hpg.hpp:127-131) plus the HIP device."""
    Serial = 0
    OpenMP = 1
    Cuda = 2
    Hip = 3


class ErrorType(enum.IntEnum):
    """hpg::ErrorType (src/hpg_error.hpp:26-38)."""
    DisabledDevice = 0
    DisabledHostDevice = 1
    OutOfBoundsCFIndex = 2
    InvalidNumberMuellerIndexRows = 3
    InvalidNumberPolarizations = 4
    InvalidCFLayout = 5
    InvalidModelGridSize = 6
    ExcessiveNumberVisibilities = 7
    UpdateWeightsWithoutGridding = 8
    CFSupportExceedsGrid = 9
    Other = 10


class Error(Exception):
    """hpg::Error (src/hpg_error.hpp:42-66)."""

    def __init__(self, message, type=ErrorType.Other):
        super().__init__(message)
        self.type = ErrorType(type)

    def message(self):
        return self.args[0]


class FFTSign(enum.IntEnum):
    """hpg::FFTSign (src/hpg.hpp:965-968); value = HPGX_FFT_*."""
    POSITIVE = 0
    NEGATIVE = 1


grid_fft_sign_dflt = FFTSign.POSITIVE   # src/hpg.hpp:972
model_fft_sign_dflt = FFTSign.NEGATIVE  # src/hpg.hpp:976-979


class ShiftDirection(enum.IntEnum):
    """hpg::ShiftDirection (src/hpg.hpp:989-992)."""
    FORWARD = 0
    BACKWARD = 1


class Layout(enum.IntEnum):
    """hpg::Layout (src/hpg.hpp:617-620); value = HPGX_LAYOUT_*."""
    Right = 0
    Left = 1


class GridMode(enum.IntEnum):
    """Grid value representation on the device (HPGX_GRID_*)."""
    FP64 = 0
    FP32_COMPENSATED = 1


def _align8(n):
    return (n + 7) & ~7


def visdata_dtype(npol):
    """numpy dtype with the memory layout of hpg::VisData<npol>
    (src/hpg.hpp:236-311): 88, 96, 112, 120 bytes for npol = 1..4."""
    if npol not in (1, 2, 3, 4):
        raise ValueError("npol must be 1..4")
    off_f = _align8(12 * npol)
    return np.dtype({
        "names": ["visibilities", "weights", "frequency", "phase", "uvw",
                  "grid_cube", "cf_index", "cf_phase_gradient"],
        "formats": [("<c8", (npol,)), ("<f4", (npol,)), "<f8", "<f8",
                    ("<f8", (3,)), "<u4", ("<u4", (2,)), ("<f8", (2,))],
        "offsets": [0, 8 * npol, off_f, off_f + 8, off_f + 16, off_f + 40,
                    off_f + 44, off_f + 56],
        "itemsize": off_f + 72,
    })


_initialized = False


def initialize(device_id=-1):
    """hpg::initialize (src/hpg.cc:62-75): selects the HIP device."""
    global _initialized
    rc = _lib.lib().hpgx_initialize(int(device_id))
    if rc != 0:
        raise Error(_lib.last_create_error() or "initialization failed", rc - 1)
    _initialized = True
    return True


def finalize():
    """hpg::finalize (src/hpg.cc:77-80)."""
    _lib.lib().hpgx_finalize()


def is_initialized():
    return _initialized


def devices():
    """hpg::devices (src/hpg.cc:82-95): the HIP device when a GPU is visible."""
    return {Device.Hip} if _lib.lib().hpgx_device_count() > 0 else set()


def host_devices():
    """hpg::host_devices (src/hpg.cc:97-109): host-side copy/layout paths."""
    return {Device.Serial, Device.OpenMP}


class CFArray:
    """Convolution function (hpg::CFArray, src/hpg.hpp:623-721) held as arrays.

    groups: list of complex64 arrays of shape (cube, mueller, y, x), the
    extents (x, y, mueller, cube) of each group including the CF padding
    (``padding`` major cells of ``oversampling`` taps on every edge)."""

    padding = 2  # src/hpg.hpp:655

    def __init__(self, oversampling, groups):
        self._os = int(oversampling)
        self._groups = []
        for g in groups:
            if _is_device_tensor(g):
                import torch
                if g.dtype != torch.complex64 or not g.is_contiguous():
                    raise ValueError("device CF groups must be contiguous complex64 tensors")
            else:
                g = np.ascontiguousarray(g, dtype=np.complex64)
            if g.ndim != 4:
                raise ValueError("CF group arrays must have shape (cube, mueller, y, x)")
            self._groups.append(g)
        kinds = {_is_device_tensor(g) for g in self._groups}
        if len(kinds) > 1:
            raise ValueError("CF groups must all be host arrays or all device tensors")

    def oversampling(self):
        return self._os

    def num_groups(self):
        return len(self._groups)

    def extents(self, grp):
        c, m, y, x = self._groups[grp].shape
        return (x, y, m, c)

    def values(self, grp):
        return self._groups[grp]

    def radii(self, grp):
        """CFArray::radii (src/hpg.hpp:687-694)."""
        os_ = self._os
        ext = self.extents(grp)
        return (((ext[0] - 2 * self.padding * os_) // os_) // 2,
                ((ext[1] - 2 * self.padding * os_) // os_) // 2)


def _is_device_tensor(a):
    return hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)


def _order_after_torch(device=None):
    """Makes work queued on torch's current stream (which may still be
    writing a device buffer handed to the C ABI) complete before the state's
    own stream reads it."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def _check(rc, state):
    if rc != 0:
        raise Error(_lib.error_message(state), rc - 1)


def _as_records(vis, npol):
    """Returns (pointer, count, memory kind, keepalive) for visibilities."""
    if hasattr(vis, "data_ptr") and hasattr(vis, "is_cuda"):
        if not vis.is_cuda:
            vis = vis.numpy()
        else:
            itemsize = visdata_dtype(npol).itemsize
            nbytes = vis.numel() * vis.element_size()
            if not vis.is_contiguous() or nbytes % itemsize:
                raise ValueError("device visibilities must be contiguous VisData records")
            # the state's stream reads the records in place: work queued on
            # torch's stream that writes them must be done first
            _order_after_torch(vis.device)
            return ctypes.c_void_p(vis.data_ptr()), nbytes // itemsize, 1, vis
    arr = np.asarray(vis)
    dt = visdata_dtype(npol)
    if arr.dtype != dt:
        if arr.dtype.names != dt.names or arr.dtype["visibilities"].shape != (npol,):
            raise Error("Number of visibility polarizations does not match Mueller matrix",
                        ErrorType.InvalidNumberPolarizations)
        # numpy packs structured dtypes in some operations (np.concatenate):
        # restore the VisData<N> layout
        arr = arr.astype(dt)
    arr = np.ascontiguousarray(arr)
    return ctypes.c_void_p(arr.ctypes.data if arr.size else 0), arr.shape[0], 0, arr


class GridderState:
    """hpg::GridderState (src/hpg.hpp:1103-2162) on the HIP device."""

    def __init__(self, handle):
        self._h = handle
        # device arrays read by queued work, kept alive until the next fence
        self._inflight = []

    # -- construction (GridderState::create, src/hpg.cc:226-264) -----------
    @classmethod
    def create(cls, device, max_added_tasks, max_visibility_batch_size,
               init_cf_shape, grid_size, grid_scale, mueller_indexes,
               conjugate_mueller_indexes=None, grid_mode=GridMode.FP64):
        if conjugate_mueller_indexes is None:
            conjugate_mueller_indexes = mueller_indexes
        m = np.ascontiguousarray(np.asarray(mueller_indexes, dtype=np.int32))
        mc = np.ascontiguousarray(np.asarray(conjugate_mueller_indexes, dtype=np.int32))
        gsz = (ctypes.c_uint * 4)(*[int(g) for g in grid_size])
        if m.ndim != 2 or mc.shape != m.shape or m.shape[0] != gsz[2]:
            raise Error("Number of rows of Mueller indexes does not equal grid mrow extent",
                        ErrorType.InvalidNumberMuellerIndexRows)
        if Device(device) != Device.Hip:
            raise Error("Requested device is not enabled", ErrorType.DisabledDevice)
        sc = (ctypes.c_double * 2)(*[float(s) for s in grid_scale])
        os_, ng, ext = 0, 0, None
        if init_cf_shape is not None and init_cf_shape.num_groups() > 0:
            os_ = init_cf_shape.oversampling()
            ng = init_cf_shape.num_groups()
            ext = (ctypes.c_uint * (4 * ng))(
                *[e for g in range(ng) for e in init_cf_shape.extents(g)])
        h = ctypes.c_void_p()
        rc = _lib.lib().hpgx_create(
            ctypes.byref(h), int(max_added_tasks), int(max_visibility_batch_size),
            gsz, sc, m.shape[1], m.shape[0], m.ctypes.data_as(ctypes.c_void_p),
            mc.ctypes.data_as(ctypes.c_void_p), int(grid_mode), os_, ng, ext)
        if rc != 0:
            raise Error(_lib.last_create_error(), rc - 1)
        return cls(h)

    def __del__(self):
        self.release()

    def release(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().hpgx_destroy(h)
            self._h = None
        self._inflight = []

    def copy(self):
        """Value copy (GridderState copy ctor, src/hpg.cc:266-306)."""
        h = ctypes.c_void_p()
        _check(_lib.lib().hpgx_clone(self._h, ctypes.byref(h)), self._h)
        return GridderState(h)

    # -- accessors (src/hpg.cc:308-345) -------------------------------------
    def device(self):
        return Device.Hip

    def max_added_tasks(self):
        return _lib.lib().hpgx_max_added_tasks(self._h)

    def max_visibility_batch_size(self):
        return _lib.lib().hpgx_max_visibility_batch_size(self._h)

    def grid_size(self):
        g = (ctypes.c_uint * 4)()
        _lib.lib().hpgx_grid_size(self._h, g)
        return tuple(g)

    def grid_scale(self):
        s = (ctypes.c_double * 2)()
        _lib.lib().hpgx_grid_scale(self._h, s)
        return tuple(s)

    def num_polarizations(self):
        return _lib.lib().hpgx_num_polarizations(self._h)

    def grid_mode(self):
        return GridMode(_lib.lib().hpgx_grid_mode(self._h))

    def is_null(self):
        return not self._h

    @property
    def grid_shape(self):
        nx, ny, nr, nc = self.grid_size()
        return (nc, nr, ny, nx)

    # -- convolution function ---------------------------------------------
    def convolution_function_region_size(self, shape=None):
        if shape is None:
            return _lib.lib().hpgx_convolution_function_region_size(self._h, 0, 0, None)
        ng = shape.num_groups()
        ext = (ctypes.c_uint * (4 * max(ng, 1)))(
            *[e for g in range(ng) for e in shape.extents(g)])
        return _lib.lib().hpgx_convolution_function_region_size(
            self._h, shape.oversampling(), ng, ext)

    def allocate_convolution_function_region(self, shape):
        if shape is None:
            _check(_lib.lib().hpgx_allocate_convolution_function_region(
                self._h, 0, 0, None), self._h)
            return self
        ng = shape.num_groups()
        ext = (ctypes.c_uint * (4 * max(ng, 1)))(
            *[e for g in range(ng) for e in shape.extents(g)])
        _check(_lib.lib().hpgx_allocate_convolution_function_region(
            self._h, shape.oversampling(), ng, ext), self._h)
        return self

    def set_convolution_function(self, host_device, cf):
        """StateT::set_convolution_function (src/hpg_runtime.hpp:890-915)."""
        if Device(host_device) not in host_devices():
            raise Error("Requested host device is not enabled", ErrorType.DisabledHostDevice)
        ng = cf.num_groups()
        ext = (ctypes.c_uint * (4 * max(ng, 1)))(
            *[int(e) for g in range(ng) for e in cf.extents(g)])
        keep = []
        ptrs = (ctypes.c_void_p * max(ng, 1))()
        kinds = {_is_device_tensor(cf.values(g)) for g in range(ng)}
        if len(kinds) > 1:
            raise ValueError("CF groups must all be host arrays or all device tensors")
        memory = 1 if kinds == {True} else 0
        for g in range(ng):
            v = cf.values(g)
            if memory:
                import torch
                if (v.dtype != torch.complex64 or not v.is_contiguous()
                        or v.numel() != int(np.prod(cf.extents(g)))):
                    raise ValueError(
                        f"device CF group {g} must be a contiguous complex64 tensor of "
                        f"shape (cube, mueller, y, x) = {tuple(cf.extents(g))[::-1]}")
                ptrs[g] = v.data_ptr()
            else:
                v = np.ascontiguousarray(v, dtype=np.complex64)
                if v.size != int(np.prod(cf.extents(g))):
                    raise ValueError(f"CF group {g} does not match its extents")
                ptrs[g] = v.ctypes.data
            keep.append(v)
        if memory:
            _order_after_torch(keep[0].device)
        _check(_lib.lib().hpgx_set_convolution_function(
            self._h, cf.oversampling(), ng, ext, ptrs, memory), self._h)
        return self

    # -- model ----------------------------------------------------------------
    def set_model(self, host_device, values, layout=Layout.Left):
        """StateT::set_model (src/hpg_runtime.hpp:917-963); values is a
        complex128 array of shape (cube, mrow, y, x) (Layout.Left) or
        (x, y, mrow, cube) (Layout.Right), or a torch complex128 GPU tensor."""
        if Device(host_device) not in host_devices():
            raise Error("Requested host device is not enabled", ErrorType.DisabledHostDevice)
        if _is_device_tensor(values):
            import torch
            keep = values.to(torch.complex128).contiguous()
            shape = tuple(keep.shape)
            ptr, memory = ctypes.c_void_p(keep.data_ptr()), 1
            # the conversion (or the caller's own kernels) must have written
            # the values before the state's stream copies them
            _order_after_torch(keep.device)
        else:
            keep = np.ascontiguousarray(values, dtype=np.complex128)
            shape = keep.shape
            ptr, memory = ctypes.c_void_p(keep.ctypes.data), 0
        if len(shape) != 4:
            raise Error("model must have rank 4", ErrorType.InvalidModelGridSize)
        msz = shape[::-1] if Layout(layout) == Layout.Left else shape
        msz = (ctypes.c_uint * 4)(*[int(s) for s in msz])
        _check(_lib.lib().hpgx_set_model(self._h, msz, ptr, int(layout), memory), self._h)
        del keep
        return self

    def reset_model(self):
        _check(_lib.lib().hpgx_reset_model(self._h), self._h)
        return self

    # -- gridding ---------------------------------------------------------------
    def grid_visibilities_base(self, host_device, visibilities, update_grid_weights,
                               do_degrid, return_visibilities, do_grid):
        """GridderState::grid_visibilities_base (src/hpg.cc:434-478), with the
        implementation's positional meaning of the flags.  Returns the
        records (residual or predicted values) when return_visibilities."""
        if Device(host_device) not in host_devices():
            raise Error("Requested host device is not enabled", ErrorType.DisabledHostDevice)
        npol = self.num_polarizations()
        ptr, n, memory, keep = _as_records(visibilities, npol)
        out = None
        out_ptr = None
        if return_visibilities:
            if memory == 1:
                import torch
                out = torch.empty_like(keep)
                out_ptr = ctypes.c_void_p(out.data_ptr())
            else:
                out = np.empty(n, dtype=visdata_dtype(npol))
                out_ptr = ctypes.c_void_p(out.ctypes.data if n else 0)
        _check(_lib.lib().hpgx_grid_visibilities(
            self._h, ptr, n, npol, int(bool(update_grid_weights)), int(bool(do_degrid)),
            int(bool(return_visibilities)), int(bool(do_grid)), out_ptr, memory), self._h)
        if memory == 1:
            self._inflight.append(keep)
        if return_visibilities:
            self.fence()
            return out
        return None

    def grid_visibilities(self, host_device, visibilities, update_grid_weights=True):
        return self.grid_visibilities_base(
            host_device, visibilities, update_grid_weights, False, False, True)

    def degrid_grid_visibilities(self, host_device, visibilities, update_grid_weights=True):
        return self.grid_visibilities_base(
            host_device, visibilities, update_grid_weights, True, False, True)

    def degrid_grid_get_residual_visibilities(self, host_device, visibilities,
                                              update_grid_weights=True):
        return self.grid_visibilities_base(
            host_device, visibilities, update_grid_weights, True, True, True)

    def degrid_get_predicted_visibilities(self, host_device, visibilities):
        return self.grid_visibilities_base(
            host_device, visibilities, False, True, True, False)

    def degrid_forward(self, visibilities):
        """Forward operator of the differentiable form (include/hpgx.h
        hpgx_degrid_forward): FP64 predictions (n, npol) of the current model.
        Returns a numpy complex128 array, or a torch complex128 GPU tensor
        when the records are a torch GPU tensor."""
        npol = self.num_polarizations()
        ptr, n, memory, keep = _as_records(visibilities, npol)
        if memory == 1:
            import torch
            out = torch.empty((n, npol), dtype=torch.complex128, device=keep.device)
            optr = ctypes.c_void_p(out.data_ptr())
        else:
            out = np.empty((n, npol), dtype=np.complex128)
            optr = ctypes.c_void_p(out.ctypes.data if n else 0)
        _check(_lib.lib().hpgx_degrid_forward(self._h, ptr, optr, n, npol, memory), self._h)
        if memory == 1:
            self.fence()
        return out

    def degrid_adjoint(self, visibilities, residuals):
        """Adjoint of ``degrid_forward`` (include/hpgx.h hpgx_degrid_adjoint):
        adds the gridding of r * phasor / conj(W) into the grid.  residuals:
        complex128 (n, npol), numpy or torch GPU (with GPU records)."""
        npol = self.num_polarizations()
        ptr, n, memory, keep = _as_records(visibilities, npol)
        if hasattr(residuals, "data_ptr") and getattr(residuals, "is_cuda", False):
            if memory != 1:
                raise ValueError("residuals and visibilities must be in the same memory")
            import torch
            rkeep = residuals.to(torch.complex128).contiguous()
            rptr = ctypes.c_void_p(rkeep.data_ptr())
            if rkeep.numel() != n * npol:
                raise ValueError("residuals must hold npol values per record")
            _order_after_torch(rkeep.device)
        else:
            if memory != 0:
                raise ValueError("residuals and visibilities must be in the same memory")
            rkeep = np.ascontiguousarray(residuals, dtype=np.complex128)
            rptr = ctypes.c_void_p(rkeep.ctypes.data if rkeep.size else 0)
        _check(_lib.lib().hpgx_degrid_adjoint(self._h, ptr, rptr, n, npol, memory), self._h)
        self.fence()
        return self

    def fence(self):
        try:
            _check(_lib.lib().hpgx_fence(self._h), self._h)
        finally:
            self._inflight.clear()
        return self

    # -- finishing passes -----------------------------------------------------
    def normalize_by_weights(self, wfactor=1.0):
        _check(_lib.lib().hpgx_normalize_by_weights(self._h, float(wfactor)), self._h)
        return self

    def apply_grid_fft(self, norm=1.0, sign=grid_fft_sign_dflt, in_place=True):
        _check(_lib.lib().hpgx_apply_grid_fft(
            self._h, float(norm), int(sign), int(bool(in_place))), self._h)
        return self

    def apply_model_fft(self, norm=1.0, sign=model_fft_sign_dflt, in_place=True):
        _check(_lib.lib().hpgx_apply_model_fft(
            self._h, float(norm), int(sign), int(bool(in_place))), self._h)
        return self

    def shift_grid(self, direction=ShiftDirection.FORWARD):
        _check(_lib.lib().hpgx_shift_grid(self._h, int(direction)), self._h)
        return self

    def shift_model(self, direction=ShiftDirection.FORWARD):
        _check(_lib.lib().hpgx_shift_model(self._h, int(direction)), self._h)
        return self

    def reset_grid(self):
        _check(_lib.lib().hpgx_reset_grid(self._h), self._h)
        return self

    # -- readback ---------------------------------------------------------------
    def grid_values(self, layout=Layout.Left):
        shape = self.grid_shape if Layout(layout) == Layout.Left else self.grid_shape[::-1]
        out = np.empty(shape, dtype=np.complex128)
        _check(_lib.lib().hpgx_grid_values(
            self._h, out.ctypes.data_as(ctypes.c_void_p), int(layout)), self._h)
        return out

    def grid_weights(self, layout=Layout.Left):
        nx, ny, nr, nc = self.grid_size()
        out = np.empty((nc, nr) if Layout(layout) == Layout.Left else (nr, nc),
                       dtype=np.float64)
        _check(_lib.lib().hpgx_grid_weights(
            self._h, out.ctypes.data_as(ctypes.c_void_p), int(layout)), self._h)
        return out

    def model_values(self, layout=Layout.Left):
        shape = self.grid_shape if Layout(layout) == Layout.Left else self.grid_shape[::-1]
        out = np.empty(shape, dtype=np.complex128)
        _check(_lib.lib().hpgx_model_values(
            self._h, out.ctypes.data_as(ctypes.c_void_p), int(layout)), self._h)
        return out

    def grid_window(self, x0, y0, nx, ny, mrow=0, cube=0):
        """Window [x0, x0+nx) x [y0, y0+ny) of one plane, shape (ny, nx)."""
        out = np.empty((ny, nx), dtype=np.complex128)
        _check(_lib.lib().hpgx_grid_window(
            self._h, x0, y0, nx, ny, mrow, cube, out.ctypes.data_as(ctypes.c_void_p)),
            self._h)
        return out

    def grid_device_ptr(self):
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(_lib.lib().hpgx_grid_device_ptr(self._h, ctypes.byref(p), ctypes.byref(n)),
               self._h)
        return p.value, n.value

    # -- timing on the state's stream -----------------------------------------
    def event_record(self, slot):
        _check(_lib.lib().hpgx_event_record(self._h, int(slot)), self._h)

    def event_elapsed_ms(self, slot0, slot1):
        ms = ctypes.c_float()
        _check(_lib.lib().hpgx_event_elapsed(self._h, int(slot0), int(slot1),
                                             ctypes.byref(ms)), self._h)
        return ms.value
