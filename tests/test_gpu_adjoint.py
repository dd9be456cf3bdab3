"""The differentiable form: degrid forward operator and its hand-written
adjoint (include/hpgx.h hpgx_degrid_forward / hpgx_degrid_adjoint), and the
torch.autograd.Function over them (hpg_amd/autograd.py).

Bars: dot-product test <A m, r> = <m, A^H r> to 1e-6 relative (FP32 CF x
value products in the adjoint); central finite differences of an FP64
chi^2 loss on 64 random pixels (step 1e-3) agree with the autograd gradient
to 1e-4 of its largest magnitude; the adjoint equals the plain gridder
applied to r / conj(W) with unit weights, W computed here from the CF."""
import numpy as np
import pytest

import refcases as rc

pytestmark = pytest.mark.gpu

M4 = [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]]


def make(hpg, npol=2, gsz=(48, 40, 2, 1), M=((0, 1), (2, 3)), nm=4, radii=(2, 4), os_=8,
         n=300, seed=0):
    from hpg_amd import workloads
    cf = workloads.random_cf(os_, list(radii), nmueller=nm, ncube=1, seed=seed + 1)
    rec = workloads.random_visibilities(n, npol, gsz, (1.0, 1.0), cf, seed=seed + 2,
                                        gradient=True, phase=True)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, gsz, (1.0, 1.0), [list(r) for r in M])
    gs.set_convolution_function(hpg.Device.Serial, cf)
    rng = np.random.default_rng(seed + 3)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return gs, cf, rec, model


def adjoint(hpg, gs, rec, r):
    gs.reset_grid()
    gs.degrid_adjoint(rec, r)
    return gs.grid_values()


@pytest.mark.parametrize("case", [0, 1])
def test_dot_product(hpg, case):
    if case == 0:
        gs, cf, rec, model = make(hpg)
    else:
        gs, cf, rec, model = make(hpg, npol=4, gsz=(64, 64, 4, 2), M=M4, nm=16,
                                  radii=(3, 6, 9), os_=6, n=500, seed=10)
    gs.set_model(hpg.Device.Serial, model)
    am = gs.degrid_forward(rec)
    rng = np.random.default_rng(7)
    r = rng.standard_normal(am.shape) + 1j * rng.standard_normal(am.shape)
    ahr = adjoint(hpg, gs, rec, r)
    lhs = np.vdot(am, r)
    rhs = np.vdot(model, ahr)
    assert abs(lhs) > 0
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), np.linalg.norm(am) * np.linalg.norm(r)), (lhs, rhs)


def test_forward_matches_degridder(hpg):
    """The forward operator is the degridder before the FP32 rounding of its
    results."""
    gs, cf, rec, model = make(hpg, seed=20)
    gs.set_model(hpg.Device.Serial, model)
    am = gs.degrid_forward(rec)
    pred = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rec)["visibilities"]
    # records outside the grid: the operator predicts 0, the degridder leaves
    # their values unchanged (src/hpg_core.hpp:1136-1169)
    inside = np.any(am != 0, axis=1)
    assert inside.sum() > 0.5 * rec.shape[0]
    assert np.array_equal(pred[~inside], rec["visibilities"][~inside])
    assert np.array_equal(am[inside].astype(np.complex64), pred[inside])


def test_finite_differences_via_autograd(hpg):
    import torch
    from hpg_amd.autograd import degrid
    gs, cf, rec, model = make(hpg, seed=30)
    rng = np.random.default_rng(31)
    n, npol = rec.shape[0], 2
    wts = rng.uniform(0.5, 1.5, (n, npol))
    data = rng.standard_normal((n, npol)) + 1j * rng.standard_normal((n, npol))
    dev_rec = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    w_t = torch.from_numpy(wts).cuda()
    d_t = torch.from_numpy(data).cuda()
    m_t = torch.from_numpy(model).cuda().requires_grad_(True)
    pred = degrid(m_t, gs, dev_rec)
    loss = (w_t * (pred - d_t).abs() ** 2).sum()
    loss.backward()
    grad = m_t.grad.cpu().numpy()

    def L(m):
        gs.set_model(hpg.Device.Serial, m)
        p = gs.degrid_forward(rec)
        return float((wts * np.abs(p - data) ** 2).sum())

    # the loss is quadratic in the model, so central differences are exact
    # up to rounding: every checked pixel meets a per-pixel relative 1e-4.
    # Pixels no record touches have a zero gradient; the 64 pixels are drawn
    # from those the footprints cover.
    h = 1e-3
    gmax = np.abs(grad).max()
    touched = np.flatnonzero(np.abs(grad).ravel() > 1e-3 * gmax)
    assert touched.size >= 64
    for k in rng.choice(touched, 64, replace=False):
        idx = np.unravel_index(k, model.shape)
        fd = []
        for unit in (1.0, 1j):
            mp = model.copy()
            mm = model.copy()
            mp[idx] += h * unit
            mm[idx] -= h * unit
            fd.append((L(mp) - L(mm)) / (2 * h))
        g_fd = fd[0] + 1j * fd[1]
        assert abs(g_fd - grad[idx]) <= 1e-4 * abs(grad[idx]), (idx, g_fd, grad[idx])


def test_adjoint_is_gridding_of_scaled_residuals(hpg):
    """A^H r = gridding of r / conj(W_v) with unit weights (M == Mc), W_v the
    sum over grid rows of the record's CF slice sums, computed here."""
    gs, cf, rec, model = make(hpg, seed=40)
    rec = rec.copy()
    rec["cf_phase_gradient"] = 0.0
    rng = np.random.default_rng(41)
    n, npol = rec.shape[0], 2
    r = rng.standard_normal((n, npol)) + 1j * rng.standard_normal((n, npol))
    got = adjoint(hpg, gs, rec, r)
    M = [[0, 1], [2, 3]]
    os_ = cf.oversampling()
    scaled = rec.copy()
    keep = np.zeros(n, dtype=bool)
    for k in range(n):
        grp = rec["cf_index"][k, 1]
        rx, ry = cf.radii(grp)
        vals = cf.values(grp)
        (gx, mx, ix), (gy, my, iy) = rc.coords(
            (48, 40), (1.0, 1.0), rec["frequency"][k], rec["uvw"][k], 0, os_)
        if not (0 <= gx - rx and gx + rx + 1 <= 48 and 0 <= gy - ry and gy + ry + 1 <= 40):
            continue
        keep[k] = True
        pos_w = rec["uvw"][k, 2] > 0
        for v in range(npol):
            W = 0j
            for gpol in range(2):
                m = M[gpol][v]
                sl = vals[0, m, my * os_ + iy::os_, mx * os_ + ix::os_][:2 * ry + 1, :2 * rx + 1]
                s = complex(sl.astype(np.complex128).sum())
                W += s if pos_w else np.conj(s)
            scaled["visibilities"][k, v] = np.complex64(r[k, v] / np.conj(W))
    scaled["weights"] = 1.0
    scaled = scaled[keep]
    gs2 = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, (48, 40, 2, 1), (1.0, 1.0), M)
    gs2.set_convolution_function(hpg.Device.Serial, cf)
    gs2.grid_visibilities(hpg.Device.Serial, scaled, False)
    exp = gs2.grid_values()
    assert np.abs(exp).max() > 0
    assert np.abs(got - exp).max() <= 1e-6 * np.abs(exp).max()


@pytest.mark.parametrize("case", [2, 3])
def test_chunked_gridding_equals_one_pass(hpg, monkeypatch, case):
    """A batch whose bin entries exceed the entry budget is gridded (and
    degridded) in chunks with the same result: case 2 runs the haloed tiles
    and the binned degridder (two grid rows), case 3 (35-pixel support) the
    clipped tiles."""
    from hpg_amd import workloads
    from test_gpu_parity import CASES
    npol, gsz, M, Mc, radii, nm, nc, os_ = CASES[case]
    cf = workloads.random_cf(os_, radii, nmueller=nm, ncube=nc, seed=50)
    rec = workloads.random_visibilities(20000, npol, gsz, (1.0, 1.0), cf, seed=51,
                                        phase=True)
    rng = np.random.default_rng(52)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    states = []
    for budget in (None, "1"):
        if budget:
            monkeypatch.setenv("HPGX_ENTRY_BUDGET_MB", budget)
        gs = hpg.GridderState.create(hpg.Device.Hip, 0, 20000, cf, gsz, (1.0, 1.0), M, Mc)
        gs.set_convolution_function(hpg.Device.Serial, cf)
        gs.grid_visibilities(hpg.Device.Serial, rec)
        gs.set_model(hpg.Device.Serial, model)
        pred = gs.degrid_forward(rec)
        states.append((gs, pred))
    (gs1, p1), (gs2, p2) = states
    a, b = gs1.grid_values(), gs2.grid_values()
    assert np.abs(a).max() > 0
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    np.testing.assert_allclose(gs2.grid_weights(), gs1.grid_weights(), rtol=1e-12)
    assert np.abs(p1).max() > 0
    assert np.abs(p1 - p2).max() <= 1e-12 * np.abs(p1).max()
