"""Precision form of the single-pol workload: the same 1e6 records gridded
into a complex<float> grid with compensated FP32 accumulation, at 1024^2,
2048^2 and 4096^2 (u, v uniform within +-0.45 of each extent), against the
FP64 oracle.  Bar (per polarization plane): e_max = max|g-g64|/max|g64| and
e_rms = ||g-g64||/||g64|| <= 1e-6 after normalize_by_weights and again after
FFT + shift; degridded values from the FP64 oracle's normalized grid agree to
1e-6 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(got, ref):
    d = np.abs(got - ref)
    return d.max() / np.abs(ref).max(), np.linalg.norm(d) / np.linalg.norm(ref)


@pytest.mark.parametrize("grid", [1024, 2048, 4096])
def test_fp32_compensated_vs_fp64(hpg, oracle_mod, grid):
    from hpg_amd import workloads
    cf = workloads.gaussian_cf(3, 20, 1.5)
    rec = workloads.config1_visibilities(1_000_000, grid, seed=100 + grid)
    print(f"grid {grid}: 1e6 records", flush=True)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, (grid, grid, 1, 1),
                                 (1.0, 1.0), [[0]], grid_mode=hpg.GridMode.FP32_COMPENSATED)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.grid_visibilities(hpg.Device.Serial, rec)
    o = oracle_mod.Oracle((grid, grid, 1, 1), (1.0, 1.0), [[0]])
    o.set_cf(20, [(cf.extents(0), cf.values(0))])
    o.grid_visibilities(rec.view(oracle_mod.visdata_dtype(1)), True)
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-13)
    gs.normalize_by_weights(1.0)
    o.normalize_by_weights(1.0)
    g32, g64 = gs.grid_values()[0, 0], o.grid_values()[0, 0]
    e_max, e_rms = rel(g32, g64)
    print(f"grid {grid} after normalize: e_max {e_max:.3e} e_rms {e_rms:.3e}", flush=True)
    assert e_max <= 1e-6 and e_rms <= 1e-6
    # degrid from the FP64 normalized grid: FP32 path vs FP64 oracle
    sub = rec[:20000]
    gs.set_model(hpg.Device.Serial, g64[None, None])
    o.set_model(g64[None, None])
    p = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, sub)["visibilities"]
    pe = o.grid_visibilities(sub.view(oracle_mod.visdata_dtype(1)), False, do_degrid=True,
                             do_grid=False)["vis"]
    pm, pr = rel(p, pe)
    print(f"grid {grid} degrid: e_max {pm:.3e} e_rms {pr:.3e}", flush=True)
    assert pm <= 1e-6 and pr <= 1e-6
    gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE)
    gs.shift_grid(hpg.ShiftDirection.FORWARD)
    o.apply_fft(1.0, "POSITIVE")
    o.shift("FORWARD")
    e_max, e_rms = rel(gs.grid_values()[0, 0], o.grid_values()[0, 0])
    print(f"grid {grid} after FFT+shift: e_max {e_max:.3e} e_rms {e_rms:.3e}", flush=True)
    assert e_max <= 1e-6 and e_rms <= 1e-6


@pytest.mark.parametrize("spread", [0.5, 12.0])
def test_fp32_compensated_coherent_cluster(hpg, oracle_mod, spread):
    """The precision form's hardest case for its FP32 partial sums: 300,000
    records of value 1 and weight 1 within +-spread pixels of one point of a
    1024^2 grid, so that the entries of a quadrant's batches fall on the same
    pixels and every product with the positive Gaussian CF has the same
    sign (no cancellation): e_max, e_rms <= 1e-6 after normalize against the
    FP64 oracle."""
    from hpg_amd import workloads
    n = 300_000
    cf = workloads.gaussian_cf(3, 20, 1.5)
    rec = workloads.config1_visibilities(n, 1024, seed=7)
    rng = np.random.default_rng(8)
    rec["visibilities"][:, 0] = 1.0 + 0.0j
    rec["weights"][:, 0] = 1.0
    rec["uvw"][:, 0] = 37.3 + rng.uniform(-spread, spread, n)
    rec["uvw"][:, 1] = -21.7 + rng.uniform(-spread, spread, n)
    print(f"coherent cluster +-{spread} px: {n} records", flush=True)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, (1024, 1024, 1, 1), (1.0, 1.0),
                                 [[0]], grid_mode=hpg.GridMode.FP32_COMPENSATED)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.grid_visibilities(hpg.Device.Serial, rec)
    o = oracle_mod.Oracle((1024, 1024, 1, 1), (1.0, 1.0), [[0]])
    o.set_cf(20, [(cf.extents(0), cf.values(0))])
    o.grid_visibilities(rec.view(oracle_mod.visdata_dtype(1)), True)
    gs.normalize_by_weights(1.0)
    o.normalize_by_weights(1.0)
    e_max, e_rms = rel(gs.grid_values()[0, 0], o.grid_values()[0, 0])
    print(f"coherent cluster +-{spread} px: e_max {e_max:.3e} e_rms {e_rms:.3e}", flush=True)
    assert e_max <= 1e-6 and e_rms <= 1e-6
