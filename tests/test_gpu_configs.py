"""Scaled-down forms of the reference's multi-polarization workloads against
the FP64 oracle, at sizes the oracle finishes in seconds.

* Full-pol W-projection (the 4096^2 x 4 pol x 8 channel form): VisData<4>
  records through the full 4x4 Mueller matrix, several channels, u, v
  Gaussian (sigma 0.2 of the extent: dense core, sparse edges, records off
  the grid skipped), CF groups chosen by w with uneven supports, taps a
  Gaussian taper times the chirp exp(i pi w rho^2).  Grid + degrid
  (residuals gridded) on a random model; both the haloed-tile gridder
  (supports up to 27 pixels) and the clipped tiles (supports up to 33).
* A/W-projection (the 8192^2 x 4 pol x 16 channel form): supports 17 to 129
  pixels, a random unit-modulus phase screen per Mueller element, weights in
  (0, 1]; grid + degrid round trip.
* Differentiable degridding (the 2048^2 x 4 pol point-source form): the
  chi^2 gradient with respect to the model from the adjoint, against central
  finite differences on 64 random pixels.

Bars: FP64 grids to 1e-12 of max|grid| per plane (no phase gradients, so
the FP32 products are the reference's; sums differ in order only); FP32-
compensated grids e_max, e_rms <= 1e-6 per plane; predicted and residual
values 1e-6 relative; weights 1e-12; finite differences 1e-4 per pixel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_view(rec, oracle_mod):
    npol = rec.dtype["visibilities"].shape[0]
    return np.ascontiguousarray(rec).view(oracle_mod.visdata_dtype(npol))


def per_plane_errors(got, exp):
    out = []
    for c in range(exp.shape[0]):
        for r in range(exp.shape[1]):
            ref = exp[c, r]
            if np.abs(ref).max() == 0:
                assert np.abs(got[c, r]).max() == 0
                continue
            d = np.abs(got[c, r] - ref)
            out.append((d.max() / np.abs(ref).max(), np.linalg.norm(d) / np.linalg.norm(ref)))
    return out


def run_grid_degrid(hpg, oracle_mod, cf, rec, gsz, mode, model_seed):
    from hpg_amd import workloads
    M = workloads.MUELLER_FULL
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, gsz, (1.0, 1.0), M, M,
                                 grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    o = oracle_mod.Oracle(gsz, (1.0, 1.0), M, M)
    o.set_cf(cf.oversampling(), [(cf.extents(g), cf.values(g)) for g in range(cf.num_groups())])
    rng = np.random.default_rng(model_seed)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    gs.set_model(hpg.Device.Serial, model)
    o.set_model(model)
    res = gs.degrid_grid_get_residual_visibilities(hpg.Device.Serial, rec)
    rexp = o.grid_visibilities(oracle_view(rec, oracle_mod), True, do_degrid=True,
                               do_grid=True)
    np.testing.assert_allclose(res["visibilities"], rexp["vis"], rtol=1e-6,
                               atol=1e-6 * np.abs(rexp["vis"]).max())
    np.testing.assert_allclose(gs.grid_weights(), o.grid_weights(), rtol=1e-12)
    errs = per_plane_errors(gs.grid_values(), o.grid_values())
    assert errs
    return errs


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("radii", [(4, 7, 10, 13), (4, 8, 12, 16)])
def test_fullpol_wprojection(hpg, oracle_mod, mode, radii):
    from hpg_amd import workloads
    gsz = (192, 160, 4, 2)
    w_max = 1.0
    w_bins = [(g + 0.5) * w_max / len(radii) * 0.02 for g in range(len(radii))]
    cf = workloads.wproj_cf(4, list(radii), w_bins, nmueller=16, ncube=2, seed=21)
    rec = workloads.gaussian_uv_visibilities(4000, 4, gsz, cf, 0.2, w_max=w_max, seed=22)
    print(f"full-pol W-projection radii {radii} mode {mode}", flush=True)
    errs = run_grid_degrid(hpg, oracle_mod, cf, rec, gsz, mode, 23)
    tol = 1e-12 if mode == 0 else 1e-6
    for e_max, e_rms in errs:
        assert e_max <= tol and e_rms <= max(tol, 1e-12), errs


@pytest.mark.parametrize("mode", [0, 1])
def test_awprojection_large_supports(hpg, oracle_mod, mode):
    from hpg_amd import workloads
    gsz = (320, 288, 4, 2)
    radii = (8, 16, 32, 64)
    cf = workloads.wproj_cf(2, list(radii), [0.001, 0.002, 0.004, 0.008], nmueller=16,
                            ncube=2, seed=31, phase_screen=True)
    rec = workloads.gaussian_uv_visibilities(400, 4, gsz, cf, 0.25, w_max=1.0,
                                             weight_range=(1e-6, 1.0), seed=32)
    print(f"A/W-projection supports 17-129 mode {mode}", flush=True)
    errs = run_grid_degrid(hpg, oracle_mod, cf, rec, gsz, mode, 33)
    tol = 1e-12 if mode == 0 else 1e-6
    for e_max, e_rms in errs:
        assert e_max <= tol and e_rms <= max(tol, 1e-12), errs


def test_differentiable_point_sources(hpg):
    """chi^2 = sum w |degrid(m) - d|^2 for the model of a point-source image
    (4 polarizations, 16 W-CF groups with supports 9-33); d = the operator on
    a perturbed model plus N(0, 0.01) noise.  Autograd gradient (adjoint
    kernel) against central finite differences, step 1e-3, 64 pixels."""
    import torch
    from hpg_amd import workloads
    from hpg_amd.autograd import degrid
    size, npol = 96, 4
    M = [[0, -1, -1, -1], [-1, 1, -1, -1], [-1, -1, 2, -1], [-1, -1, -1, 3]]
    radii = [4 + (12 * g) // 15 for g in range(16)]
    cf = workloads.wproj_cf(4, radii, [0.002 * g for g in range(16)], nmueller=4, ncube=1,
                            seed=41)
    gsz = (size, size, npol, 1)
    rec = workloads.gaussian_uv_visibilities(3000, npol, gsz, cf, 0.15, w_max=1.0, seed=42)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, rec.shape[0], cf, gsz, (1.0, 1.0), M, M)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    img = workloads.point_source_image(size, npol, num_sources=60, seed=43)
    model = np.fft.fft2(img.astype(np.float64))[None].astype(np.complex128)  # (1, npol, y, x)
    rng = np.random.default_rng(44)
    pert = model * (1.0 + 0.01 * rng.standard_normal(model.shape))
    gs.set_model(hpg.Device.Serial, pert)
    data = gs.degrid_forward(rec)
    data = data + 0.01 * (rng.standard_normal(data.shape) + 1j * rng.standard_normal(data.shape))
    wts = rng.uniform(0.5, 1.5, data.shape)
    dev_rec = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    m_t = torch.from_numpy(model).cuda().requires_grad_(True)
    pred = degrid(m_t, gs, dev_rec)
    loss = (torch.from_numpy(wts).cuda() * (pred - torch.from_numpy(data).cuda()).abs() ** 2).sum()
    loss.backward()
    grad = m_t.grad.cpu().numpy()

    def L(m):
        gs.set_model(hpg.Device.Serial, m)
        p = gs.degrid_forward(rec)
        return float((wts * np.abs(p - data) ** 2).sum())

    h = 1e-3
    gmax = np.abs(grad).max()
    assert gmax > 0
    for k in rng.choice(model.size, 64, replace=False):
        idx = np.unravel_index(k, model.shape)
        fd = []
        for unit in (1.0, 1j):
            mp = model.copy()
            mm = model.copy()
            mp[idx] += h * unit
            mm[idx] -= h * unit
            fd.append((L(mp) - L(mm)) / (2 * h))
        g_fd = fd[0] + 1j * fd[1]
        # the loss is quadratic in the model, so central differences are exact
        # up to rounding: per pixel relative 1e-4
        assert abs(g_fd - grad[idx]) <= 1e-4 * abs(grad[idx]), (idx, g_fd, grad[idx])
