"""The allocation-free gridding entry point (hpgx_grid_visibilities_noalloc)
against hpgx_grid_visibilities, both called through the C ABI with
device-resident records.

Bars: FP64-grid mode, the same per-tap FP32 products summed in FP64 in
another order: 1e-12 of max|grid|; FP32-compensated mode: 1e-6 of max|grid|
(each is within 1e-6 of the FP64 oracle); weights (FP64 sums): 1e-12.
Each case runs two calls in a row, the second on a smaller batch, so the
second call finds the first call's contents in the scratch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device_records(rec):
    import torch
    return torch.from_numpy(rec.view(np.uint8).copy()).cuda()


def _noalloc(hpg, gs, dev, npol, update=True):
    import torch
    from hpg_amd import _lib
    torch.cuda.synchronize(dev.device)
    n = dev.numel() // hpg.visdata_dtype(npol).itemsize
    return _lib.lib().hpgx_grid_visibilities_noalloc(
        gs._h, ctypes.c_void_p(dev.data_ptr()), n, npol, int(update))


def _compare(a, b, wa, wb, mode):
    bar = 1e-12 if mode == 0 else 1e-6
    scale = np.abs(a).max()
    assert scale > 0
    assert np.abs(b - a).max() <= bar * scale
    np.testing.assert_allclose(wb, wa, rtol=1e-12, atol=0)


# (npol, grid size, Mueller rows, radii, mueller elements, cubes, os):
# haloed tiles (supports 5-13: the quadrant kernel for FP32 grids), wider
# supports (haloed tiles with LDS atomics), supports beyond a tile
# (block-local bins of clipped footprints), and a many-key grid whose bins
# take several passes through the scratch
CASES = [
    (1, (256, 192, 1, 1), [[0]], [2, 6], 1, 1, 20),
    (2, (192, 160, 2, 2), [[0, 1], [2, 3]], [3, 9], 4, 2, 8),
    (1, (160, 160, 1, 2), [[0]], [18, 20], 1, 1, 4),
]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_matches_grid_visibilities(hpg, case, mode):
    from hpg_amd import workloads
    npol, gsz, M, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=40 + case)
    rec = workloads.random_visibilities(9000, npol, gsz, (1.0, 1.0), cf, seed=50 + case,
                                        phase=True)
    print(f"noalloc case {case} mode {mode}: grid {gsz} npol {npol} radii {radii}", flush=True)
    ref = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, gsz, (1.0, 1.0), M,
                                  grid_mode=mode)
    ref.set_convolution_function(hpg.Device.Serial, cf)
    na = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, gsz, (1.0, 1.0), M,
                                 grid_mode=mode)
    na.set_convolution_function(hpg.Device.Serial, cf)
    for n in (rec.shape[0], 2500):
        dev = _device_records(rec[:n])
        ref.grid_visibilities(hpg.Device.Serial, dev)
        ref.fence()
        assert _noalloc(hpg, na, dev, npol) == 0, hpg._lib.error_message(na._h)
        na.fence()
        _compare(ref.grid_values(), na.grid_values(), ref.grid_weights(), na.grid_weights(),
                 mode)
        ref.reset_grid()
        na.reset_grid()


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_several_passes(hpg, mode):
    """8192 (plane, tile) keys: the halo values take most of the scratch, so
    the 600,000 records are binned in three passes, against one pass of
    hpgx_grid_visibilities; then a second, smaller call."""
    from hpg_amd import workloads
    gsz = (2048, 2048, 2, 1)
    M = [[0], [0]]
    cf = workloads.random_cf(10, [3], nmueller=1, ncube=1, seed=61)
    rec = workloads.random_visibilities(600000, 1, gsz, (1.0, 1.0), cf, seed=62)
    print(f"noalloc passes mode {mode}: grid {gsz} records {rec.shape[0]}", flush=True)
    ref = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, gsz, (1.0, 1.0), M,
                                  grid_mode=mode)
    ref.set_convolution_function(hpg.Device.Serial, cf)
    na = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, gsz, (1.0, 1.0), M,
                                 grid_mode=mode)
    na.set_convolution_function(hpg.Device.Serial, cf)
    for n in (rec.shape[0], 200000):
        dev = _device_records(rec[:n])
        ref.grid_visibilities(hpg.Device.Serial, dev)
        ref.fence()
        assert _noalloc(hpg, na, dev, 1) == 0, hpg._lib.error_message(na._h)
        na.fence()
        _compare(ref.grid_values(), na.grid_values(), ref.grid_weights(), na.grid_weights(),
                 mode)
        ref.reset_grid()
        na.reset_grid()


def test_noalloc_rejects_what_does_not_fit(hpg):
    """A grid with more (plane, tile) keys than the entry point bins, and a
    grid whose halo values alone exceed the scratch, give HPGX_ERR_OTHER
    with a message; the grid is left untouched."""
    from hpg_amd import workloads
    # 4096^2: 16384 keys; 2048^2 x 2 with 25-pixel supports: 8192 keys whose
    # halo values (2112 pixels each) need 277 MB
    for gsz, radius, words in (((4096, 4096, 1, 1), 3, "keys are not supported"),
                               ((2048, 2048, 2, 1), 12, "bytes of scratch")):
        cfx = workloads.random_cf(4, [radius], nmueller=1, ncube=1, seed=72)
        M = [[0]] * gsz[2]
        gs = hpg.GridderState.create(hpg.Device.Hip, 0, 1000, cfx, gsz, (1.0, 1.0), M)
        gs.set_convolution_function(hpg.Device.Serial, cfx)
        rec = workloads.random_visibilities(1000, 1, gsz, (1.0, 1.0), cfx, seed=73)
        dev = _device_records(rec)
        rc = _noalloc(hpg, gs, dev, 1)
        assert rc == 11, rc  # HPGX_ERR_OTHER
        assert words in hpg._lib.error_message(gs._h)
        gs.fence()
        assert not gs.grid_weights().any()
