"""Parity of the HIP device (through the C ABI) with the CPU oracle.

Bars (stated per test):
  * FP64-grid mode, no phase gradient: the per-tap FP32 products are the
    reference's bit for bit; only the order of FP64 accumulation differs, so
    grids agree to 1e-12 of max|grid| and single-visibility grids exactly.
  * phase gradients: the device's FP64 sin/cos may round the FP32 screen
    differently in rare taps: 1e-6 of max|grid|.
  * FP32-compensated mode: e_max, e_rms <= 1e-6 against the FP64 oracle.
  * degridded visibilities: FP64 sums rounded to FP32: 1e-6 relative.
  * normalize / shift: exact; FFT (hipFFT vs the oracle's FP64 FFT): 1e-12.
"""
import numpy as np
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu


def oracle_view(rec, oracle_mod):
    return np.ascontiguousarray(rec).view(oracle_mod.visdata_dtype(rec.dtype["visibilities"].shape[0]))


def make_pair(hpg, oracle_mod, grid_size, grid_scale, M, cf, mode=0, Mc=None,
              max_batch=100000):
    Mc = M if Mc is None else Mc
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, max_batch, cf, grid_size,
                                 grid_scale, M, Mc, grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    o = oracle_mod.Oracle(grid_size, grid_scale, M, Mc)
    o.set_cf(cf.oversampling(),
             [(cf.extents(g), cf.values(g)) for g in range(cf.num_groups())])
    return gs, o


def rel_errors(got, ref):
    scale = np.abs(ref).max()
    d = np.abs(got - ref)
    return d.max() / scale, np.linalg.norm(d) / np.linalg.norm(ref)


def cone_cfarray(hpg, oracle_mod, nmueller, os_, radius):
    _, vals = oracle_mod.cone_cf(nmueller, os_, radius)
    return hpg.CFArray(os_, [vals])


@pytest.mark.parametrize("mode", [0, 1])
def test_gridone(hpg, oracle_mod, gridone_golden, mode):
    """GridOne on a 16384^2 grid: bit-exact with the oracle (FP64 mode, and
    FP32-compensated mode since each pixel holds one FP32 product)."""
    g = rc.GRIDONE
    cf = cone_cfarray(hpg, oracle_mod, 1, g["oversampling"], g["radius"])
    gs, o = make_pair(hpg, oracle_mod, g["grid_size"], g["grid_scale"], [[0]], cf, mode)
    rec = rc.one_record(hpg.visdata_dtype(1), g["vis"], g["weight"], g["freq"], g["uvw"])
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    x0, y0 = g["ll"][0] - 8, g["ll"][1] - 8
    got = gs.grid_window(x0, y0, 107, 107)
    exp = o.grid_window(x0, y0, 107, 107)
    assert np.array_equal(got, exp)
    win = gs.grid_window(g["ll"][0], g["ll"][1], 91, 91).T
    assert np.array_equal(win != 0, gridone_golden != 0)
    assert (win[:90, :90] == gridone_golden[:90, :90]).mean() > 0.98
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-14)


@pytest.mark.parametrize("mode", [0, 1])
def test_gridding_sweep(hpg, oracle_mod, mode):
    """Gridding test (GridderStateUnitTests.cc:1273-1341), exact."""
    g = rc.GRIDDING
    cf = cone_cfarray(hpg, oracle_mod, 1, g["oversampling"], g["radius"])
    fine = g["grid_scale"][0] * g["oversampling"]
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, 1, cf, g["grid_size"],
                                 g["grid_scale"], [[0]], grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    for i in range(g["oversampling"] + 1):
        for j in range(g["oversampling"] + 1):
            uvw = (i * fine, j * fine, 0.0)
            gs.reset_grid()
            rec = rc.one_record(hpg.visdata_dtype(1), g["vis"], g["weight"], g["freq"], uvw)
            gs.grid_visibilities(hpg.Device.Serial, rec, True)
            exp = rc.single_vis_expectation(cf.values(0)[0, 0], g["grid_size"],
                                            g["grid_scale"], g["freq"], uvw, g["radius"],
                                            g["oversampling"], g["vis"])
            assert np.array_equal(gs.grid_values()[0, 0], exp), (i, j)


CASES = [
    # npol, grid_size, M, Mc, radii, nmueller, ncube, os
    (1, (40, 36, 1, 1), [[0]], [[0]], [3], 1, 1, 10),
    (2, (48, 40, 2, 2), [[0, -1], [-1, 0]], [[0, -1], [-1, 0]], [2, (4, 3)], 1, 2, 8),
    (2, (64, 64, 2, 1), [[0, 1], [2, 3]], [[1, 0], [3, 2]], [1, 5], 4, 1, 6),
    (4, (70, 66, 4, 2), [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]],
     [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]], [2, 7, 17], 16, 2, 5),
    (3, (50, 50, 2, 1), [[0, -1, 1], [2, 0, -1]], [[1, -1, 0], [-1, 2, 0]], [3, 3], 3, 1, 7),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("gradient", [False, True])
def test_random_vs_oracle_fp64(hpg, oracle_mod, case, gradient):
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=10 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), M, cf, 0, Mc)
    rec = workloads.random_visibilities(700, npol, gsz, (0.9, 1.1), cf, seed=20 + case,
                                        gradient=gradient, phase=True, freq_scale=1.1)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    assert np.abs(exp).max() > 0
    e_max, _ = rel_errors(got, exp)
    assert e_max <= (1e-6 if gradient else 1e-12), e_max
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_random_vs_oracle_fp32_compensated(hpg, oracle_mod, case):
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=10 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), M, cf, 1, Mc)
    rec = workloads.random_visibilities(3000, npol, gsz, (0.9, 1.1), cf, seed=30 + case,
                                        phase=True)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    for c in range(gsz[3]):
        for r in range(gsz[2]):
            if np.abs(exp[c, r]).max() == 0:
                continue
            e_max, e_rms = rel_errors(got[c, r], exp[c, r])
            assert e_max <= 1e-6 and e_rms <= 1e-6, (c, r, e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


@pytest.mark.parametrize("gridder", [2, 3, 4, 5])
@pytest.mark.parametrize("case", [0, 2, 3, 4])
@pytest.mark.parametrize("mode", [0, 1])
def test_grid_without_weights_vs_oracle(hpg, oracle_mod, monkeypatch, gridder, case, mode):
    """grid_vis_no_weights (src/hpg_core.hpp:995-1068, 1221-1251): gridding
    with update_grid_weights false adds to the grid exactly as the oracle
    does and leaves the weights bit for bit as they were (here: nonzero,
    from a first call with weights).  FP64 mode 1e-12 of max|grid| per
    plane, FP32-compensated e_max, e_rms <= 1e-6."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=140 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), M, cf, mode, Mc)
    first = workloads.random_visibilities(800, npol, gsz, (0.9, 1.1), cf, seed=150 + case,
                                          phase=True, margin=-3)
    second = workloads.random_visibilities(1200, npol, gsz, (0.9, 1.1), cf, seed=160 + case,
                                           phase=True, margin=-3)
    print(f"no weights: gridder {gridder} case {case} mode {mode}", flush=True)
    gs.grid_visibilities(hpg.Device.Serial, first, True)
    o.grid_visibilities(oracle_view(first, oracle_mod), True)
    before = gs.grid_weights()
    assert np.abs(before).max() > 0
    gs.grid_visibilities(hpg.Device.Serial, second, False)
    o.grid_visibilities(oracle_view(second, oracle_mod), False)
    assert np.array_equal(gs.grid_weights(), before)
    np.testing.assert_allclose(before, o.grid_weights(), rtol=1e-12)
    got, exp = gs.grid_values(), o.grid_values()
    tol = 1e-12 if mode == 0 else 1e-6
    for c in range(gsz[3]):
        for r in range(gsz[2]):
            if np.abs(exp[c, r]).max() == 0:
                assert np.abs(got[c, r]).max() == 0
                continue
            e_max, e_rms = rel_errors(got[c, r], exp[c, r])
            assert e_max <= tol and e_rms <= max(tol, 1e-12), (c, r, e_max, e_rms)


QUAD_CASES = [
    # npol, grid_size, M, Mc, radii, nmueller, ncube, os: supports of at most
    # 8 pixels, one support for all groups or several
    (1, (40, 36, 1, 1), [[0]], [[0]], [3], 1, 1, 10),
    (2, (96, 80, 2, 2), [[0, 1], [2, 3]], [[1, 0], [3, 2]], [1, (3, 2), 2], 4, 2, 6),
    (4, (70, 66, 4, 1), [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]],
     [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]], [3, 1], 16, 1, 5),
    (3, (50, 50, 2, 1), [[0, -1, 1], [2, 0, -1]], [[1, -1, 0], [-1, 2, 0]], [3, 3], 3, 1, 7),
]


@pytest.mark.parametrize("sub", ["1", "0"])
@pytest.mark.parametrize("case", range(len(QUAD_CASES)))
def test_quadrant_kernels_vs_oracle(hpg, oracle_mod, monkeypatch, sub, case):
    """FP32-compensated grids with supports up to 8 pixels through the
    quadrant-key kernel (k_sub, HPGX_SUB_TILES=1, the default) and through
    the round-based quadrant kernel (k_quad, HPGX_SUB_TILES=0), with phase
    gradients on a third of the records: e_max, e_rms <= 1e-6 per plane."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_SUB_TILES", sub)
    npol, gsz, M, Mc, radii, nm, nc, os_ = QUAD_CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=170 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), M, cf, 1, Mc)
    rec = workloads.random_visibilities(2500, npol, gsz, (0.9, 1.1), cf, seed=180 + case,
                                        phase=True, margin=-3, gradient=True)
    rec["cf_phase_gradient"][1::3] = 0.0
    rec["cf_phase_gradient"][2::3] = 0.0
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    for c in range(gsz[3]):
        for r in range(gsz[2]):
            if np.abs(exp[c, r]).max() == 0:
                assert np.abs(got[c, r]).max() == 0
                continue
            e_max, e_rms = rel_errors(got[c, r], exp[c, r])
            assert e_max <= 1e-6 and e_rms <= 1e-6, (c, r, e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


ONE_SUPPORT_CASES = [
    # grid_size, radii (one support for every group), ncube, os, records
    ((40, 36, 1, 1), [0], 1, 10, 6000),
    ((64, 64, 1, 2), [1, 1], 2, 6, 9000),
    ((100, 70, 1, 1), [2], 1, 8, 4000),
    ((96, 96, 1, 3), [(3, 2), (3, 2)], 2, 5, 12000),
    ((130, 97, 1, 1), [(1, 3)], 1, 20, 3000),
    ((33, 65, 1, 1), [3], 1, 9, 7000),
]


@pytest.mark.parametrize("case", range(len(ONE_SUPPORT_CASES)))
def test_quadrant_kernel_one_support_vs_oracle(hpg, oracle_mod, case):
    """The quadrant-key kernel (k_sub) with one polarization and one support
    of 1 to 7 pixels a side for every CF group (its lanes of the support
    fixed for the call) against the oracle: quadrant lists of several
    64-entry batches, phase gradients on a third of the records, records off
    the grid, ragged grid edges: e_max, e_rms <= 1e-6 per plane and the
    weights to 1e-12."""
    from hpg_amd import workloads
    gsz, radii, nc, os_, nrec = ONE_SUPPORT_CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=1, ncube=nc, seed=230 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), [[0]], cf, 1)
    rec = workloads.random_visibilities(nrec, 1, gsz, (0.9, 1.1), cf, seed=240 + case,
                                        phase=True, margin=-3, gradient=True)
    rec["cf_phase_gradient"][1::3] = 0.0
    rec["cf_phase_gradient"][2::3] = 0.0
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    for c in range(gsz[3]):
        assert np.abs(exp[c, 0]).max() > 0
        e_max, e_rms = rel_errors(got[c, 0], exp[c, 0])
        assert e_max <= 1e-6 and e_rms <= 1e-6, (c, e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


DENSE_CASES = [
    # npol, M, Mc, nmueller
    (1, [[0]], [[0]], 1),
    (2, [[0, 1], [2, 3]], [[1, 0], [3, 2]], 4),
    (4, [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]],
     [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]], 16),
]


@pytest.mark.parametrize("case", range(len(DENSE_CASES)))
def test_quadrant_kernel_dense_folds_vs_oracle(hpg, oracle_mod, case):
    """The quadrant-key kernel on quadrants of over a thousand entries each
    (24000 records on a 64 x 64 grid: about ten rounds of batches, so the
    shared fold runs inside the round loop as well as after it), one, two
    and four polarizations, phase gradients on a third of the records:
    e_max, e_rms <= 1e-6 per plane, weights to 1e-12."""
    from hpg_amd import workloads
    npol, M, Mc, nm = DENSE_CASES[case]
    gsz = (64, 64, len(M), 1)
    cf = workloads.random_cf(8, [3, (2, 3)], nmueller=nm, ncube=1, seed=260 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 1, Mc)
    rec = workloads.random_visibilities(24000, npol, gsz, (1.0, 1.0), cf, seed=270 + case,
                                        phase=True, margin=4, gradient=True)
    rec["cf_phase_gradient"][1::3] = 0.0
    rec["cf_phase_gradient"][2::3] = 0.0
    print(f"dense quadrants npol {npol}", flush=True)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    for r in range(gsz[2]):
        assert np.abs(exp[0, r]).max() > 0
        e_max, e_rms = rel_errors(got[0, r], exp[0, r])
        assert e_max <= 1e-6 and e_rms <= 1e-6, (r, e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


GLOBAL_SUB_RADII = [[3], [1, (3, 2), 2], [1]]


@pytest.mark.parametrize("case", range(len(GLOBAL_SUB_RADII)))
def test_quadrant_kernel_global_bins_vs_oracle(hpg, oracle_mod, case):
    """The quadrant-key kernel on globally counted quadrant bins: a
    one-polarization FP32-compensated grid of 320 x 288 x 1 x 100 (9000
    (plane, tile) keys, more than one workgroup's histogram holds), one
    support or several, phase gradients on a third of the records, records
    off the grid: e_max, e_rms <= 1e-6 per plane, weights to 1e-12."""
    from hpg_amd import workloads
    gsz = (320, 288, 1, 100)
    cf = workloads.random_cf(7, GLOBAL_SUB_RADII[case], nmueller=1, ncube=2, seed=190 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), [[0]], cf, 1)
    rec = workloads.random_visibilities(20000, 1, gsz, (0.9, 1.1), cf, seed=195 + case,
                                        phase=True, margin=-3, gradient=True)
    rec["cf_phase_gradient"][1::3] = 0.0
    rec["cf_phase_gradient"][2::3] = 0.0
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    planes = 0
    for c in range(gsz[3]):
        if np.abs(exp[c, 0]).max() == 0:
            assert np.abs(got[c, 0]).max() == 0
            continue
        planes += 1
        e_max, e_rms = rel_errors(got[c, 0], exp[c, 0])
        assert e_max <= 1e-6 and e_rms <= 1e-6, (c, e_max, e_rms)
    assert planes >= 90
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


@pytest.mark.parametrize("gridder", [2, 3, 4, 5])
@pytest.mark.parametrize("case", [0, 2, 3, 4])
@pytest.mark.parametrize("mode", [0, 1])
def test_each_gridder_vs_oracle(hpg, oracle_mod, monkeypatch, gridder, case, mode):
    """Every gridding pipeline (HPGX_GRIDDER, read when a state is created:
    2 globally binned entries, 3 block-local bins with
    clipped tiles, 4 block-local bins with haloed tiles updated by LDS
    atomics -- which falls back to 3 for case 3, whose 35-pixel support
    exceeds a tile -- and 5 the haloed tiles added per quadrant wave, one
    lane per tap (FP32-compensated grids: the quadrant-key kernel for
    supports up to 8 pixels, the round-based one up to 14), falling back to
    4 where the quadrant regions do not fit LDS) against the oracle, with phase gradients on a third of the records and some records
    off the grid.  FP64 mode:
    1e-6 of max|grid| (gradient screens); FP32-compensated: e_max, e_rms
    <= 1e-6 per plane."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=40 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), M, cf, mode, Mc)
    rec = workloads.random_visibilities(1500, npol, gsz, (0.9, 1.1), cf, seed=50 + case,
                                        phase=True, margin=-3, gradient=True)
    rec["cf_phase_gradient"][1::3] = 0.0
    rec["cf_phase_gradient"][2::3] = 0.0
    print(f"gridder {gridder} case {case} mode {mode}", flush=True)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    for c in range(gsz[3]):
        for r in range(gsz[2]):
            if np.abs(exp[c, r]).max() == 0:
                assert np.abs(got[c, r]).max() == 0
                continue
            e_max, e_rms = rel_errors(got[c, r], exp[c, r])
            assert e_max <= 1e-6 and e_rms <= 1e-6, (c, r, e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("gridder", [4, 5])
def test_dense_tile_vs_oracle(hpg, oracle_mod, monkeypatch, gridder, mode):
    """4000 single-pol records on one (plane, tile) key: 85% with their
    support corner in a 6 x 6 box of one quadrant, the rest over the whole
    tile, a few with phase gradients.  One key then takes many entry rounds
    with very unequal quadrant shares (gridder 5) or many scheduling rounds
    (gridder 4).  Against the oracle: e_max, e_rms <= 1e-6 (FP64 mode: the
    gradient screens' FP32 rounding; FP32-compensated: the pairs)."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    gsz = (96, 96, 1, 1)
    cf = workloads.random_cf(10, [3], nmueller=1, ncube=1, seed=61)
    n = 4000
    rec = workloads.random_visibilities(n, 1, gsz, (1.0, 1.0), cf, seed=62, phase=True,
                                        gradient=True)
    keep = rec["cf_phase_gradient"][::50].copy()
    rec["cf_phase_gradient"] = 0.0
    rec["cf_phase_gradient"][::50] = keep
    rng = np.random.default_rng(63)
    dense = rng.random(n) < 0.85
    # grid position = u + 48; tile (1, 1) covers 32..63, its first quadrant's
    # support corners (position - 3) lie in 32..47
    pos = np.where(dense[:, None], rng.uniform(38.0, 44.0, (n, 2)),
                   rng.uniform(36.0, 60.0, (n, 2)))
    rec["uvw"][:, 0] = pos[:, 0] - 48.0
    rec["uvw"][:, 1] = pos[:, 1] - 48.0
    print(f"dense tile gridder {gridder} mode {mode}", flush=True)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), [[0]], cf, mode, [[0]])
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    assert np.abs(exp).max() > 0
    e_max, e_rms = rel_errors(got[0, 0], exp[0, 0])
    assert e_max <= 1e-6 and e_rms <= 1e-6, (e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


@pytest.mark.parametrize("gridder", [4, 5])
@pytest.mark.parametrize("scheme", ["cas", "fp64"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp32_tile_schemes(hpg, oracle_mod, monkeypatch, gridder, scheme, case):
    """Each tile accumulation of the FP32 grid (HPGX_FP32_TILE: FP32 two-sum
    pairs -- under compare-and-swap in gridder 4, plain read-modify-write in
    gridder 5 --, or FP64 tile sums) against the FP64 oracle: e_max, e_rms <= 1e-6 per plane.  The
    records are dense (3000 on grids of 40-70 pixels), so pixels collect
    hundreds of contributions and lanes collide on pixels."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    monkeypatch.setenv("HPGX_FP32_TILE", scheme)
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=60 + case)
    gs, o = make_pair(hpg, oracle_mod, gsz, (0.9, 1.1), M, cf, 1, Mc)
    rec = workloads.random_visibilities(3000, npol, gsz, (0.9, 1.1), cf, seed=70 + case,
                                        phase=True)
    print(f"gridder {gridder} scheme {scheme} case {case}", flush=True)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    for c in range(gsz[3]):
        for r in range(gsz[2]):
            if np.abs(exp[c, r]).max() == 0:
                continue
            e_max, e_rms = rel_errors(got[c, r], exp[c, r])
            assert e_max <= 1e-6 and e_rms <= 1e-6, (c, r, e_max, e_rms)


@pytest.mark.parametrize("gridder", [4, 5])
@pytest.mark.parametrize("radius", [(15, 2), (2, 15), (16, 1), (6, 6), (12, 1), (15, 0), (15, 1),
                                    (0, 0)])
@pytest.mark.parametrize("mode", [0, 1])
def test_support_at_tile_edge(hpg, oracle_mod, monkeypatch, gridder, radius, mode):
    """Largest supports the haloed tiles take (31 pixels on one axis), the
    first they hand to the clipped tiles (33), 13 x 13 (the quadrant kernel's
    largest square regions), wide and flat supports the quadrant kernel takes
    (25 x 3, 31 x 1) and one whose tile + halo exceeds its flush (31 x 3),
    and a 1x1 support (no halo) on a grid whose edges are not tile
    multiples; records cover the grid up to its edges.  No phase gradients:
    FP64 grids to 1e-12 of max|grid|, FP32-compensated to 1e-6."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    gsz = (97, 75, 1, 1)
    cf = workloads.random_cf(4, [radius], seed=7)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), [[0]], cf, mode)
    rec = workloads.random_visibilities(2500, 1, gsz, (1.0, 1.0), cf, seed=8, phase=True,
                                        margin=-2)
    print(f"gridder {gridder} support radius {radius} mode {mode}", flush=True)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    e_max, e_rms = rel_errors(got, exp)
    assert e_max <= (1e-12 if mode == 0 else 1e-6) and e_rms <= 1e-6, (e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)


def test_serialization_and_batches(hpg, oracle_mod):
    """Serialization (GridderStateUnitTests.cc:1126-1218): the grid does not
    depend on the order of the batches; gridding in batches equals one call."""
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[1]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=3)
    a = workloads.random_visibilities(500, npol, gsz, (1.0, 1.0), cf, seed=4)
    b = workloads.random_visibilities(500, npol, gsz, (1.0, 1.0), cf, seed=5)
    gs1, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    gs2, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    gs1.grid_visibilities(hpg.Device.Serial, a)
    gs1.grid_visibilities(hpg.Device.Serial, b)
    gs2.grid_visibilities(hpg.Device.Serial, b)
    gs2.grid_visibilities(hpg.Device.Serial, a)
    g1, g2 = gs1.grid_values(), gs2.grid_values()
    np.testing.assert_allclose(g1, g2, rtol=0, atol=1e-12 * np.abs(g1).max())
    np.testing.assert_allclose(gs1.grid_weights(), gs2.grid_weights(), rtol=1e-13)


@pytest.mark.parametrize("gridder", [3, 4, 5])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_degrid_and_residual_vs_oracle(hpg, oracle_mod, monkeypatch, gridder, case):
    """Predicted values, residuals and the gridded residuals against the
    oracle: 1e-6 relative (FP32 outputs of FP64 sums), grids to 1e-5.  With
    the haloed-tile gridders (4, 5) and supports below 32 pixels the degridder
    is binned (model tiles in LDS, per-record sums, then the finishing pass,
    summed atomically over grid rows when there are several); otherwise it
    runs one lane per record (3; and case 3, whose support is 35 pixels, one
    wave per record).  Phase gradients on half the records, records off the
    grid, both signs of w."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=7 + case)
    rec = workloads.random_visibilities(800, npol, gsz, (1.0, 1.0), cf, seed=8 + case,
                                        gradient=True, phase=True, margin=-2)
    rec["cf_phase_gradient"][::2] = 0.0
    print(f"degrid case {case} gridder {gridder}", flush=True)
    rng = np.random.default_rng(9)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    gs.set_model(hpg.Device.Serial, model)
    o.set_model(model)
    pred = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rec)
    pexp = o.grid_visibilities(oracle_view(rec, oracle_mod), False, do_degrid=True,
                               do_grid=False)
    p, pe = pred["visibilities"], pexp["vis"]
    assert np.abs(pe).max() > 0
    np.testing.assert_allclose(p, pe, rtol=1e-6, atol=1e-6 * np.abs(pe).max())
    # residual + gridding of the residual
    res = gs.degrid_grid_get_residual_visibilities(hpg.Device.Serial, rec)
    rexp = o.grid_visibilities(oracle_view(rec, oracle_mod), True, do_degrid=True,
                               do_grid=True)
    np.testing.assert_allclose(res["visibilities"], rexp["vis"], rtol=1e-6,
                               atol=1e-6 * np.abs(rexp["vis"]).max())
    e_max, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e_max <= 1e-5, e_max


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("radius", [0, 1, 3, (4, 3), (3, 4), 5])
@pytest.mark.parametrize("gridder", [4, 5])
def test_degrid_single_pol_vs_oracle(hpg, oracle_mod, monkeypatch, gridder, radius, mixed):
    """Single-polarization binned degridding against the oracle: predictions
    and residuals to 1e-6 relative, the gridded residuals to 1e-5.  Supports
    of 1 to 63 taps take the degridder's branch-free tap loop when no record
    of a wave has a phase gradient (mixed=False); 121 taps, and waves holding
    one of the records with a gradient (every 97th when mixed), take the
    chunked loop.  Two CF cubes, both signs of w, records off the grid."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    gsz = (48, 40, 1, 1)
    cf = workloads.random_cf(10, [radius], nmueller=1, ncube=2, seed=31)
    rec = workloads.random_visibilities(900, 1, gsz, (1.0, 1.0), cf, seed=32,
                                        gradient=mixed, phase=True, margin=-2)
    if mixed:
        keep = rec["cf_phase_gradient"][::97].copy()
        rec["cf_phase_gradient"] = 0.0
        rec["cf_phase_gradient"][::97] = keep
    print(f"single-pol degrid gridder {gridder} radius {radius} mixed {mixed}", flush=True)
    rng = np.random.default_rng(33)
    model = rng.standard_normal((1, 1, 40, 48)) + 1j * rng.standard_normal((1, 1, 40, 48))
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), [[0]], cf, 0, [[0]])
    gs.set_model(hpg.Device.Serial, model)
    o.set_model(model)
    pred = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rec)
    pexp = o.grid_visibilities(oracle_view(rec, oracle_mod), False, do_degrid=True,
                               do_grid=False)
    p, pe = pred["visibilities"], pexp["vis"]
    assert np.abs(pe).max() > 0
    np.testing.assert_allclose(p, pe, rtol=1e-6, atol=1e-6 * np.abs(pe).max())
    res = gs.degrid_grid_get_residual_visibilities(hpg.Device.Serial, rec)
    rexp = o.grid_visibilities(oracle_view(rec, oracle_mod), True, do_degrid=True,
                               do_grid=True)
    np.testing.assert_allclose(res["visibilities"], rexp["vis"], rtol=1e-6,
                               atol=1e-6 * np.abs(rexp["vis"]).max())
    e_max, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e_max <= 1e-5, e_max


@pytest.mark.parametrize("npol", [1, 2])
def test_degrid_slice_order_vs_oracle(hpg, oracle_mod, npol):
    """The wave-per-record degridder (supports above 32 x 32 taps) takes a
    batch's records in CF-slice order (group, CF cube, fine offsets): three
    groups with supports of 33 to 45 pixels, two CF cubes and two grid cubes,
    records off the grid, phase gradients on half the records.  Predictions
    and residuals against the oracle to 1e-6 relative, the gridded residuals
    to 1e-5, and the batch in reverse order gives every record's prediction
    bit for bit."""
    from hpg_amd import workloads
    if npol == 1:
        M = Mc = [[0]]
        nm = 1
    else:
        M, Mc = [[0, 1], [2, 3]], [[1, 0], [3, 2]]
        nm = 4
    gsz = (128, 112, npol, 2)
    cf = workloads.random_cf(6, [16, (18, 20), 22], nmueller=nm, ncube=2, seed=51)
    rec = workloads.random_visibilities(3000, npol, gsz, (1.0, 1.0), cf, seed=52,
                                        gradient=True, phase=True, margin=-10)
    rec["cf_phase_gradient"][::2] = 0.0
    print(f"slice-ordered degrid npol {npol}", flush=True)
    rng = np.random.default_rng(53)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    gs.set_model(hpg.Device.Serial, model)
    o.set_model(model)
    pred = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rec)
    pexp = o.grid_visibilities(oracle_view(rec, oracle_mod), False, do_degrid=True,
                               do_grid=False)
    p, pe = pred["visibilities"], pexp["vis"]
    assert np.abs(pe).max() > 0
    np.testing.assert_allclose(p, pe, rtol=1e-6, atol=1e-6 * np.abs(pe).max())
    rev = np.ascontiguousarray(rec[::-1])
    prev = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rev)["visibilities"]
    assert np.array_equal(prev[::-1], p)
    res = gs.degrid_grid_get_residual_visibilities(hpg.Device.Serial, rec)
    rexp = o.grid_visibilities(oracle_view(rec, oracle_mod), True, do_degrid=True,
                               do_grid=True)
    np.testing.assert_allclose(res["visibilities"], rexp["vis"], rtol=1e-6,
                               atol=1e-6 * np.abs(rexp["vis"]).max())
    e_max, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e_max <= 1e-5, e_max


def test_residual_without_model_is_input(hpg, oracle_mod):
    """ResidualVisibilities (GridderStateUnitTests.cc:1743-1815)."""
    cf = cone_cfarray(hpg, oracle_mod, 1, 10, 3)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, 1, cf, (20, 20, 2, 1),
                                 (0.0476591, 0.0476591), [[0, -1], [-1, 0]])
    gs.set_convolution_function(hpg.Device.Serial, cf)
    rec = rc.one_record(hpg.visdata_dtype(2), [1 - 1j, -1 + 1j], [1.0, 0.5], 3.693e9,
                        (0.0, 0.0, 0.1))
    res = gs.degrid_grid_get_residual_visibilities(hpg.Device.Serial, rec)
    assert res.tobytes() == rec.tobytes()


@pytest.mark.parametrize("mode", [0, 1])
def test_finishing_passes_vs_oracle(hpg, oracle_mod, mode):
    """normalize_by_weights -> FFT -> shift on a gridded plane set."""
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[1]
    gsz = (64, 48, 2, 2)
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=12)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, mode, Mc)
    rec = workloads.random_visibilities(2000, npol, gsz, (1.0, 1.0), cf, seed=13)
    gs.grid_visibilities(hpg.Device.Serial, rec)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    gs.normalize_by_weights(2.0)
    o.normalize_by_weights(2.0)
    tol = 1e-12 if mode == 0 else 1e-6
    e1, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e1 <= tol, e1
    for sign in (hpg.FFTSign.POSITIVE, hpg.FFTSign.NEGATIVE):
        gs.apply_grid_fft(norm=3.0, sign=sign)
        o.apply_fft(norm=3.0, sign=sign.name)
        e2, _ = rel_errors(gs.grid_values(), o.grid_values())
        assert e2 <= (1e-12 if mode == 0 else 2e-6), (sign, e2)
    for d in (hpg.ShiftDirection.FORWARD, hpg.ShiftDirection.BACKWARD):
        gs.shift_grid(d)
        o.shift(d.name)
        e3, _ = rel_errors(gs.grid_values(), o.grid_values())
        assert e3 <= (1e-12 if mode == 0 else 2e-6), (d, e3)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("shape", [(64, 48), (63, 48), (64, 47)])
def test_finishing_chain_vs_oracle(hpg, oracle_mod, mode, shape):
    """normalize -> FFT -> shift with no read in between (one pass and the
    FFT for even sizes, the shift a (-1)^(x+y) modulation before the FFT;
    the separate passes otherwise), then gridding onto the finished grid,
    another chain in the other order (FFT -> normalize), and a cleared
    chain; each against the oracle running the same calls."""
    from hpg_amd import workloads
    npol, _, M, Mc, radii, nm, nc, os_ = CASES[1]
    gsz = (shape[0], shape[1], 2, 2)
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=14)
    gs, o = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, mode, Mc)
    tol = 1e-12 if mode == 0 else 2e-6
    rec = workloads.random_visibilities(2000, npol, gsz, (1.0, 1.0), cf, seed=15)
    gs.grid_visibilities(hpg.Device.Serial, rec)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    gs.normalize_by_weights(2.0)
    gs.apply_grid_fft(norm=3.0, sign=hpg.FFTSign.NEGATIVE)
    gs.shift_grid(hpg.ShiftDirection.FORWARD)
    o.normalize_by_weights(2.0)
    o.apply_fft(norm=3.0, sign="NEGATIVE")
    o.shift("FORWARD")
    e1, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e1 <= tol, e1
    # a chain still pending when records are gridded runs first
    rec2 = workloads.random_visibilities(500, npol, gsz, (1.0, 1.0), cf, seed=16)
    gs.apply_grid_fft(norm=1.0, sign=hpg.FFTSign.POSITIVE)
    gs.shift_grid(hpg.ShiftDirection.BACKWARD)
    gs.grid_visibilities(hpg.Device.Serial, rec2)
    o.apply_fft(norm=1.0, sign="POSITIVE")
    o.shift("BACKWARD")
    o.grid_visibilities(oracle_view(rec2, oracle_mod), True)
    e2, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e2 <= tol, e2
    # FFT before normalize, then a shift: no fusion across the normalize
    gs.apply_grid_fft(norm=2.0, sign=hpg.FFTSign.NEGATIVE)
    gs.normalize_by_weights(0.5)
    gs.shift_grid(hpg.ShiftDirection.FORWARD)
    o.apply_fft(norm=2.0, sign="NEGATIVE")
    o.normalize_by_weights(0.5)
    o.shift("FORWARD")
    gs.fence()
    e3, _ = rel_errors(gs.grid_values(), o.grid_values())
    assert e3 <= tol, e3
    # a pending chain on a cleared grid leaves it zero
    gs.normalize_by_weights(1.0)
    gs.apply_grid_fft()
    gs.shift_grid()
    gs.reset_grid()
    assert not np.any(gs.grid_values())


@pytest.mark.parametrize("shape", [(20, 20), (19, 19), (18, 19), (19, 18)])
def test_shift_cycle(hpg, shape):
    """ShiftCycle (GridderStateUnitTests.cc:1860-1924) through the model."""
    nx, ny = shape
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, 1, None, (nx, ny, 1, 2), (1.0, 1.0), [[0]])
    rng = np.random.default_rng(1)
    m = rng.standard_normal((2, 1, ny, nx)) + 1j * rng.standard_normal((2, 1, ny, nx))
    gs.set_model(hpg.Device.Serial, m)
    gs.shift_model(hpg.ShiftDirection.FORWARD)
    assert np.array_equal(gs.model_values(), np.fft.fftshift(m, axes=(-2, -1)))
    gs.shift_model(hpg.ShiftDirection.BACKWARD)
    assert np.array_equal(gs.model_values(), m)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("gridder", [3, 5])
@pytest.mark.parametrize("case", [0, 1])
def test_zero_model_leaves_gridding_unchanged(hpg, oracle_mod, monkeypatch, case, gridder,
                                              mode):
    """ZeroModel (GridderStateUnitTests.cc:1523-1596): degridding against a
    zero model and gridding the residuals gives the grid and weights of
    plain gridding (the predictions are zero, so the residuals are the
    records' values).  Gridder 5 degrids binned, 3 one lane per record;
    FP64 and FP32-compensated grids, to 1e-12 and 1e-6 of max|grid|."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=41 + case)
    rec = workloads.random_visibilities(700, npol, gsz, (1.0, 1.0), cf, seed=42 + case,
                                        gradient=True, phase=True, margin=-2)
    print(f"zero model case {case} gridder {gridder} mode {mode}", flush=True)
    gs_d, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, mode, Mc)
    gs_g, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, mode, Mc)
    gs_d.set_model(hpg.Device.Serial,
                   np.zeros((gsz[3], gsz[2], gsz[1], gsz[0]), dtype=np.complex128))
    gs_d.degrid_grid_visibilities(hpg.Device.Serial, rec)
    gs_g.grid_visibilities(hpg.Device.Serial, rec)
    a, b = gs_d.grid_values(), gs_g.grid_values()
    assert np.abs(b).max() > 0
    tol = 1e-12 if mode == 0 else 1e-6
    np.testing.assert_allclose(a, b, rtol=0, atol=tol * np.abs(b).max())
    np.testing.assert_allclose(gs_d.grid_weights(), gs_g.grid_weights(), rtol=1e-13)


def test_model_fft_linearity(hpg):
    """ModelFFT (GridderStateUnitTests.cc:1600-1700): the model FFT divided by
    its norm is linear in the model, and it is numpy's forward 2-D DFT (the
    model's default sign is NEGATIVE) of every plane, to 1e-12."""
    gsz = (64, 48, 2, 2)
    gs1 = hpg.GridderState.create(hpg.Device.Hip, 0, 15, None, gsz, (0.1, 0.1),
                                  [[0, -1], [-1, 0]])
    gs2 = hpg.GridderState.create(hpg.Device.Hip, 0, 15, None, gsz, (0.1, 0.1),
                                  [[0, -1], [-1, 0]])
    rng = np.random.default_rng(43)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    m = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    gs1.set_model(hpg.Device.Serial, m)
    gs1.apply_model_fft(norm=1.0)
    gs2.set_model(hpg.Device.Serial, 2.5 * m)
    gs2.apply_model_fft(norm=2.5)
    f1, f2 = gs1.model_values(), gs2.model_values()
    ref = np.fft.fft2(m, axes=(-2, -1))
    np.testing.assert_allclose(f1, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(f2, f1, rtol=0, atol=1e-12 * np.abs(f1).max())


def test_one_fft_sign(hpg):
    """OneFFT (GridderStateUnitTests.cc:1704-1740)."""
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, 15, None, (2048, 2048, 1, 1),
                                 (0.12, -0.34), [[0]])
    m = np.zeros((1, 1, 2048, 2048), dtype=np.complex128)
    m[0, 0, 1108, 1024] = 0.825
    gs.set_model(hpg.Device.Serial, m)
    gs.apply_model_fft()
    gs.shift_model(hpg.ShiftDirection.FORWARD)
    assert gs.model_values()[0, 0, 10, 1070].imag < 0


def test_layouts_and_copy(hpg, oracle_mod):
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[1]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=14)
    gs, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    rec = workloads.random_visibilities(300, npol, gsz, (1.0, 1.0), cf, seed=15)
    gs.grid_visibilities(hpg.Device.Serial, rec)
    left = gs.grid_values(hpg.Layout.Left)
    right = gs.grid_values(hpg.Layout.Right)
    assert np.array_equal(right, left.transpose(3, 2, 1, 0))
    wl, wr = gs.grid_weights(hpg.Layout.Left), gs.grid_weights(hpg.Layout.Right)
    assert np.array_equal(wr, wl.T)
    cp = gs.copy()
    gs.grid_visibilities(hpg.Device.Serial, rec)
    assert np.array_equal(cp.grid_values(), left)
    cp.grid_visibilities(hpg.Device.Serial, rec)
    assert np.array_equal(cp.grid_values(), gs.grid_values())


def test_error_paths(hpg, oracle_mod):
    E = hpg.ErrorType
    cf = cone_cfarray(hpg, oracle_mod, 1, 10, 3)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, 2, cf, (20, 20, 1, 1), (1.0, 1.0), [[0]])
    gs.set_convolution_function(hpg.Device.Serial, cf)
    rec3 = np.zeros(3, dtype=hpg.visdata_dtype(1))
    with pytest.raises(hpg.Error) as e:
        gs.grid_visibilities(hpg.Device.Serial, rec3)
    assert e.value.type == E.ExcessiveNumberVisibilities
    with pytest.raises(hpg.Error) as e:
        gs.grid_visibilities(hpg.Device.Serial, np.zeros(1, dtype=hpg.visdata_dtype(2)))
    assert e.value.type == E.InvalidNumberPolarizations
    with pytest.raises(hpg.Error) as e:
        gs.grid_visibilities_base(hpg.Device.Serial, rec3[:1], True, True, False, False)
    assert e.value.type == E.UpdateWeightsWithoutGridding
    with pytest.raises(hpg.Error) as e:
        gs.grid_visibilities(hpg.Device.Cuda, rec3[:1])
    assert e.value.type == E.DisabledHostDevice
    big = cone_cfarray(hpg, oracle_mod, 1, 10, 12)  # extent 291 > 20 * 10
    with pytest.raises(hpg.Error) as e:
        gs.set_convolution_function(hpg.Device.Serial, big)
    assert e.value.type == E.CFSupportExceedsGrid
    with pytest.raises(hpg.Error) as e:
        gs.set_model(hpg.Device.Serial, np.zeros((1, 1, 20, 21), dtype=np.complex128))
    assert e.value.type == E.InvalidModelGridSize
    bad = rc.one_record(hpg.visdata_dtype(1), 1 + 0j, 1.0, rc.C_LIGHT, (0.0, 0.0, 0.0),
                        cf_index=(0, 5))
    gs.grid_visibilities(hpg.Device.Serial, bad)
    with pytest.raises(hpg.Error) as e:
        gs.fence()
    assert e.value.type == E.OutOfBoundsCFIndex
    gs.fence()
    assert not gs.grid_values().any()


def test_empty_inputs(hpg, oracle_mod):
    """GridNone / EmptyCF (GridderStateUnitTests.cc:1465-1521, 1817-1857)."""
    cf = cone_cfarray(hpg, oracle_mod, 1, 20, 45)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, 1, cf, (1024, 1024, 1, 1),
                                 (0.0476591, 0.0476591), [[0]])
    gs.grid_visibilities(hpg.Device.Serial, np.zeros(0, dtype=hpg.visdata_dtype(1)))
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.grid_visibilities(hpg.Device.Serial, np.zeros(0, dtype=hpg.visdata_dtype(1)))
    assert not gs.grid_values().any() and not gs.grid_weights().any()


def test_device_resident_records(hpg, oracle_mod):
    """Records handed over as a torch GPU tensor give the same grid."""
    import torch
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[2]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=16)
    rec = workloads.random_visibilities(500, npol, gsz, (1.0, 1.0), cf, seed=17)
    gs1, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    gs2, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
    gs1.grid_visibilities(hpg.Device.Serial, rec)
    dev = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    gs2.grid_visibilities(hpg.Device.Serial, dev)
    a, b = gs1.grid_values(), gs2.grid_values()
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("case", [0, 2, 3])
def test_device_resident_degridding(hpg, oracle_mod, case):
    """Degridding with records in device memory (read in place; the input
    tensor is never written) gives the values and grids of host records:
    predicted values, residuals with gridding, and gridding of residuals
    without returning them.  Case 0: one grid row (predictions finished in
    the tile kernel); case 2: two grid rows (per-record sums); case 3: a
    35-pixel support (staged copy, one wave per record)."""
    import torch
    from hpg_amd import workloads
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=80 + case)
    rec = workloads.random_visibilities(600, npol, gsz, (1.0, 1.0), cf, seed=81 + case,
                                        phase=True, margin=-2)
    rng = np.random.default_rng(82 + case)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dev = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    before = dev.clone()

    def pair():
        out = []
        for _ in range(2):
            gs, _ = make_pair(hpg, oracle_mod, gsz, (1.0, 1.0), M, cf, 0, Mc)
            gs.set_model(hpg.Device.Serial, model)
            out.append(gs)
        return out

    g_host, g_dev = pair()
    p_host = g_host.degrid_get_predicted_visibilities(hpg.Device.Serial, rec)
    p_dev = g_dev.degrid_get_predicted_visibilities(hpg.Device.Serial, dev)
    assert p_dev.cpu().numpy().view(rec.dtype).tobytes() == p_host.tobytes()
    g_host, g_dev = pair()
    r_host = g_host.degrid_grid_get_residual_visibilities(hpg.Device.Serial, rec)
    r_dev = g_dev.degrid_grid_get_residual_visibilities(hpg.Device.Serial, dev)
    assert r_dev.cpu().numpy().view(rec.dtype).tobytes() == r_host.tobytes()
    a, b = g_host.grid_values(), g_dev.grid_values()
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-12 * np.abs(a).max())
    g_host, g_dev = pair()
    g_host.degrid_grid_visibilities(hpg.Device.Serial, rec)
    g_dev.degrid_grid_visibilities(hpg.Device.Serial, dev)
    a, b = g_host.grid_values(), g_dev.grid_values()
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-12 * np.abs(a).max())
    np.testing.assert_allclose(g_dev.grid_weights(), g_host.grid_weights(), rtol=1e-14)
    g_dev.fence()
    assert torch.equal(dev, before)


@pytest.mark.parametrize("gridder", [2, 4, 5])
@pytest.mark.parametrize("shape", [(4096, 2112, 1, 1), (1024, 1024, 2, 5)])
@pytest.mark.parametrize("mode", [0, 1])
def test_large_grids_global_bins(hpg, oracle_mod, monkeypatch, gridder, shape, mode):
    """Grids with more (plane, tile) keys than one workgroup's LDS histogram
    holds (8448 and 10240 keys) are binned through global counters: into
    one entry per touched tile (2), or into the haloed tiles (4, the
    default).  Against the oracle: FP64 to 1e-12 of max|grid| per plane,
    FP32-compensated e_max, e_rms <= 1e-6; weights 1e-12."""
    from hpg_amd import workloads
    monkeypatch.setenv("HPGX_GRIDDER", str(gridder))
    npol = 1 if shape[2] == 1 else 2
    M = [[0]] if npol == 1 else [[0, -1], [-1, 0]]
    cf = workloads.random_cf(6, [3, (2, 4)], nmueller=1, ncube=shape[3], seed=90)
    rec = workloads.random_visibilities(30000, npol, shape, (1.0, 1.0), cf, seed=91,
                                        phase=True, margin=-3)
    print(f"large grid {shape} gridder {gridder} mode {mode}", flush=True)
    gs, o = make_pair(hpg, oracle_mod, shape, (1.0, 1.0), M, cf, mode, max_batch=30000)
    gs.grid_visibilities(hpg.Device.Serial, rec, True)
    o.grid_visibilities(oracle_view(rec, oracle_mod), True)
    got, exp = gs.grid_values(), o.grid_values()
    tol = 1e-12 if mode == 0 else 1e-6
    for c in range(shape[3]):
        for r in range(shape[2]):
            if np.abs(exp[c, r]).max() == 0:
                continue
            e_max, e_rms = rel_errors(got[c, r], exp[c, r])
            assert e_max <= tol and e_rms <= max(tol, 1e-12), (c, r, e_max, e_rms)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)
