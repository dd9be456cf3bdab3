"""The multi-polarization workloads at their single-GPU shapes.

* Full-pol W-projection: 4096^2 grid x 4 polarization rows x 8 channels,
  VisData<4> records through the full 4x4 Mueller matrix, 64 CF groups by w
  with supports 9..65 (radius 4 + round(28 g / 63)), oversampling 20, taps a
  Gaussian taper times the chirp exp(i pi kappa_g rho^2), u, v ~ N(0, (0.2
  extent)^2), 5e7 records; degridding of a random N(0, 1) model with the
  residuals gridded (update_grid_weights).
* A/W-projection, one GPU's share of the 16-channel form: 8192^2 x 4 x 2
  channels, 512 CF groups (32 w bins x 16 parallactic-angle bins, supports
  17..129, oversampling 10, a random unit-modulus phase screen per tap and
  Mueller element), u, v ~ N(0, (0.25 extent)^2), weights in (0, 1], 2.5e7
  records; grid + degrid round trip.
* Differentiable degridding: a 2048^2 x 4 point-source image (200 sources,
  log-uniform flux in [1e-3, 1], N(0, 1e-4) noise) -> model (model FFT,
  sign NEGATIVE, then shift) -> degridding through 16 W-CF groups
  (supports 9..33, oversampling 16) to 2e7 records; the weighted chi^2
  against the operator on a perturbed image plus N(0, 0.01) noise.

Each grid form is checked two ways.  (1) A subset of the records (the first
ones of the batch, so drawn from the same distribution) is degridded and
gridded at the full grid shape and compared with the FP64 oracle per
polarization plane: FP64 grids e_max <= 1e-12 (the FP32 products are the
reference's; sums differ in order only), FP32-compensated grids e_max,
e_rms <= 1e-6, weights 1e-12, residual values 1e-6.  (2) The whole batch
runs through size-independent properties: the batch in one call equals the
batch in two calls in swapped order (FP64 sums reordered: 1e-12 of each
plane's max; weights 1e-12), and the predictions of sampled records inside
the whole batch equal those of the same records degridded alone (the
degridding of a record does not depend on the others: 1e-12) and the
oracle's (1e-6).  The differentiable form's image gradient (adjoint kernel,
adjoint model FFT and shift) is checked against central finite differences
of the loss on 64 random pixels, step 1e-3, per pixel relative 1e-4.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def device_grid(gs):
    """The state's grid as a torch tensor sharing its device memory, shape
    (cube, mrow, y, x, 2) (re, im), FP64 or FP32 by grid mode."""
    import torch
    ptr, count = gs.grid_device_ptr()
    nx, ny, nr, nc = gs.grid_size()
    fp64 = int(gs.grid_mode()) == 0

    class _Array:
        pass

    a = _Array()
    a.__cuda_array_interface__ = {"shape": (count * 2,), "typestr": "<f8" if fp64 else "<f4",
                                  "data": (ptr, False), "version": 3}
    return torch.as_tensor(a, device="cuda").view(nc, nr, ny, nx, 2)


def plane_errors_device(a, b):
    """per (cube, mrow) plane: (max|a - b| / max|b|, ||a - b|| / ||b||) of two
    (cube, mrow, y, x, 2) tensors, computed on the device in FP64"""
    import torch
    out = []
    for c in range(b.shape[0]):
        for r in range(b.shape[1]):
            ref = b[c, r].to(torch.float64)
            d = a[c, r].to(torch.float64) - ref
            dm = torch.sqrt((d * d).sum(-1))
            rm = torch.sqrt((ref * ref).sum(-1))
            scale = float(rm.max())
            if scale == 0:
                assert float(dm.max()) == 0
                continue
            out.append((float(dm.max()) / scale, float(torch.linalg.norm(dm) / torch.linalg.norm(rm))))
    return out


def plane_errors_host(got, exp):
    out = []
    for c in range(exp.shape[0]):
        for r in range(exp.shape[1]):
            ref = exp[c, r]
            scale = np.abs(ref).max()
            if scale == 0:
                assert np.abs(got[c, r]).max() == 0
                continue
            d = np.abs(got[c, r] - ref)
            out.append((d.max() / scale, np.linalg.norm(d) / np.linalg.norm(ref)))
    return out


def host_records(dev, npol, hpg):
    return dev.cpu().numpy().view(hpg.visdata_dtype(npol)).reshape(-1)


def oracle_view(rec, oracle_mod):
    npol = rec.dtype["visibilities"].shape[0]
    return np.ascontiguousarray(rec).view(oracle_mod.visdata_dtype(npol))


def random_model(shape, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
                         torch.randn(shape, generator=g, device="cuda", dtype=torch.float64))


def subset_vs_oracle(hpg, mode, gsz, cf, M, model_dev, sub_dev, oracle_res):
    """Degrid + grid the subset at the full grid shape, returning residuals;
    compare residuals, weights and the grid per plane with the oracle's."""
    o_vis, o_grid, o_wgt = oracle_res
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, sub_dev.shape[0], cf, gsz, (1.0, 1.0),
                                 M, M, grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.set_model(hpg.Device.Serial, model_dev)
    res = gs.degrid_grid_get_residual_visibilities(hpg.Device.Serial, sub_dev)
    got_vis = host_records(res, 4, hpg)["visibilities"]
    np.testing.assert_allclose(got_vis, o_vis, rtol=1e-6, atol=1e-6 * np.abs(o_vis).max())
    np.testing.assert_allclose(gs.grid_weights(), o_wgt, rtol=1e-12)
    errs = plane_errors_host(gs.grid_values(), o_grid)
    gs.release()
    tol = 1e-12 if mode == 0 else 1e-6
    log(f"  mode {mode}: worst plane e_max {max(e[0] for e in errs):.3e} "
        f"e_rms {max(e[1] for e in errs):.3e} over {len(errs)} planes")
    for e_max, e_rms in errs:
        assert e_max <= tol and e_rms <= tol, errs


def full_batch_properties(hpg, oracle, mode, gsz, cf, M, model_dev, rec_dev, sample, npol):
    """The whole batch in one call against two calls in swapped order, and
    the predictions of `sample` records inside the batch against the same
    records alone and against the oracle (set to the same model)."""
    import time
    import torch
    n = rec_dev.shape[0]
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, gsz, (1.0, 1.0), M, M,
                                 grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.set_model(hpg.Device.Serial, model_dev)
    t0 = time.perf_counter()
    gs.degrid_grid_visibilities(hpg.Device.Serial, rec_dev)
    gs.fence()
    log(f"  {n} records, degrid + grid in one call: {time.perf_counter() - t0:.2f} s")
    g1 = device_grid(gs).clone()
    w1 = gs.grid_weights()
    gs.reset_grid()
    half = n // 2
    gs.degrid_grid_visibilities(hpg.Device.Serial, rec_dev[half:])
    gs.degrid_grid_visibilities(hpg.Device.Serial, rec_dev[:half])
    gs.fence()
    np.testing.assert_allclose(gs.grid_weights(), w1, rtol=1e-12)
    errs = plane_errors_device(device_grid(gs), g1)
    log(f"  halves in swapped order: worst plane e_max {max(e[0] for e in errs):.3e}")
    tol = 1e-12 if mode == 0 else 1e-6
    for e_max, _ in errs:
        assert e_max <= tol, errs
    del g1
    torch.cuda.empty_cache()
    # predictions: inside the whole batch, alone, and the oracle's
    pred_all = gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rec_dev)
    idx = torch.as_tensor(sample, device="cuda")
    in_batch = host_records(pred_all[idx], npol, hpg)["visibilities"]
    del pred_all
    alone = host_records(gs.degrid_get_predicted_visibilities(hpg.Device.Serial, rec_dev[idx]),
                         npol, hpg)["visibilities"]
    np.testing.assert_allclose(in_batch, alone, rtol=1e-12, atol=0)
    recs = host_records(rec_dev[idx], npol, hpg)
    exp = oracle(recs)["vis"]
    np.testing.assert_allclose(alone, exp, rtol=1e-6, atol=1e-6 * np.abs(exp).max())
    gs.release()


# ---- full-pol W-projection -------------------------------------------------

CFG2_GRID = (4096, 4096, 4, 8)
CFG2_RECORDS = 50_000_000
CFG2_SUBSET = 20_000


@pytest.fixture(scope="module")
def cfg2(hpg, oracle_mod):
    import torch
    from hpg_amd import workloads
    radii, chirps = workloads.config2_geometry()
    cf = workloads.taper_chirp_cf_device(20, radii, chirps, 16)
    rec = workloads.gaussian_uv_device_records(
        CFG2_RECORDS, 4, CFG2_GRID, lambda w: torch.clamp((w * 64).floor(), max=63), 0.2,
        ncube=8)
    model = random_model((8, 4, 4096, 4096), 7)
    M = workloads.MUELLER_FULL
    log(f"full-pol W-projection: grid {CFG2_GRID}, 64 groups, {CFG2_RECORDS} records")
    o = oracle_mod.Oracle(CFG2_GRID, (1.0, 1.0), M, M)
    o.set_cf(20, [(cf.extents(g), cf.values(g).cpu().numpy()) for g in range(cf.num_groups())])
    o.set_model(model.cpu().numpy())
    sub = rec[:CFG2_SUBSET]
    ores = o.grid_visibilities(oracle_view(host_records(sub, 4, hpg), oracle_mod), True,
                               do_degrid=True, do_grid=True)
    oracle_res = (ores["vis"], o.grid_values(), o.grid_weights())

    def predict(recs):
        o.reset_grid()
        return o.grid_visibilities(oracle_view(recs, oracle_mod), False, do_degrid=True,
                                   do_grid=False)

    yield dict(cf=cf, rec=rec, model=model, M=M, oracle_res=oracle_res, predict=predict)
    del o


@pytest.mark.parametrize("mode", [0, 1])
def test_cfg2_subset_vs_oracle(hpg, cfg2, mode):
    log(f"full-pol W-projection subset of {CFG2_SUBSET} records, mode {mode}")
    subset_vs_oracle(hpg, mode, CFG2_GRID, cfg2["cf"], cfg2["M"], cfg2["model"],
                     cfg2["rec"][:CFG2_SUBSET], cfg2["oracle_res"])


def test_cfg2_full_batch_properties(hpg, cfg2):
    rng = np.random.default_rng(24)
    sample = np.sort(rng.choice(CFG2_RECORDS, 2000, replace=False))
    log("full-pol W-projection, whole batch, FP64 grid")
    full_batch_properties(hpg, cfg2["predict"], 0, CFG2_GRID, cfg2["cf"], cfg2["M"],
                          cfg2["model"], cfg2["rec"], sample, 4)


# ---- A/W-projection, one GPU's share -----------------------------------------

CFG5_GRID = (8192, 8192, 4, 2)
CFG5_RECORDS = 25_000_000
CFG5_SUBSET = 2_000
# the groups of the oracle subset: w bins 0, 10, 20, 31 x parallactic-angle
# bins 0, 15 (supports 17, 53, 89, 129)
CFG5_SUB_GROUPS = [wb * 16 + pa for wb in (0, 10, 20, 31) for pa in (0, 15)]


@pytest.fixture(scope="module")
def cfg5(hpg, oracle_mod):
    import torch
    from hpg_amd import workloads
    radii, chirps = workloads.config5_geometry()
    cf = workloads.taper_chirp_cf_device(10, radii, chirps, 16, phase_screen=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(51)
    pa = torch.randint(0, 16, (CFG5_RECORDS,), generator=g, device="cuda")

    def group_of_w(w):
        return torch.clamp((w * 32).floor(), max=31).to(torch.int64) * 16 + pa

    rec = workloads.gaussian_uv_device_records(CFG5_RECORDS, 4, CFG5_GRID, group_of_w, 0.25,
                                               weight_range=(1e-6, 1.0), ncube=2, seed=52)
    M = workloads.MUELLER_FULL
    log(f"A/W-projection: grid {CFG5_GRID}, 512 groups, {CFG5_RECORDS} records")
    # the subset: records of the batch whose group is one of CFG5_SUB_GROUPS;
    # the oracle holds just those groups, renumbered
    dt_off = hpg.visdata_dtype(4).fields["cf_index"][1]
    grp = rec[:, dt_off + 4:dt_off + 8].contiguous().view(torch.int32)[:, 0]
    sel = torch.zeros_like(grp, dtype=torch.bool)
    for k in CFG5_SUB_GROUPS:
        sel |= grp == k
    sub = rec[torch.nonzero(sel)[:CFG5_SUBSET, 0]].contiguous()
    remap = {k: i for i, k in enumerate(CFG5_SUB_GROUPS)}
    o = oracle_mod.Oracle(CFG5_GRID, (1.0, 1.0), M, M)
    o.set_cf(10, [(cf.extents(k), cf.values(k).cpu().numpy()) for k in CFG5_SUB_GROUPS])
    model = random_model((2, 4, 8192, 8192), 53)
    o.set_model(model.cpu().numpy())

    def remapped(recs):
        out = oracle_view(recs, oracle_mod).copy()
        out["cf_index"][:, 1] = [remap[int(k)] for k in out["cf_index"][:, 1]]
        return out

    ores = o.grid_visibilities(remapped(host_records(sub, 4, hpg)), True, do_degrid=True,
                               do_grid=True)
    oracle_res = (ores["vis"], o.grid_values(), o.grid_weights())

    def predict(recs):
        o.reset_grid()
        return o.grid_visibilities(remapped(recs), False, do_degrid=True, do_grid=False)

    sub_index = torch.nonzero(sel)[:, 0].cpu().numpy()
    yield dict(cf=cf, rec=rec, model=model, M=M, sub=sub, oracle_res=oracle_res,
               predict=predict, sub_index=sub_index)
    del o


@pytest.mark.parametrize("mode", [1, 0])
def test_cfg5_subset_vs_oracle(hpg, cfg5, mode):
    log(f"A/W-projection subset of {cfg5['sub'].shape[0]} records, mode {mode}")
    subset_vs_oracle(hpg, mode, CFG5_GRID, cfg5["cf"], cfg5["M"], cfg5["model"],
                     cfg5["sub"], cfg5["oracle_res"])


def test_cfg5_full_batch_properties(hpg, cfg5):
    rng = np.random.default_rng(54)
    # sampled records of the oracle's groups
    sample = np.sort(rng.choice(cfg5["sub_index"], 500, replace=False))
    log("A/W-projection, whole batch, FP32-compensated grid")
    full_batch_properties(hpg, cfg5["predict"], 1, CFG5_GRID, cfg5["cf"], cfg5["M"],
                          cfg5["model"], cfg5["rec"], sample, 4)


# ---- differentiable degridding from the image ----------------------------------

def test_cfg4_image_gradient(hpg):
    import torch
    from hpg_amd import workloads
    from hpg_amd.autograd import degrid, image_to_model, _model_pass
    size, npol, nrec = 2048, 4, 20_000_000
    gsz = (size, size, npol, 1)
    M = [[0, -1, -1, -1], [-1, 1, -1, -1], [-1, -1, 2, -1], [-1, -1, -1, 3]]
    radii = [4 + (12 * g) // 15 for g in range(16)]
    chirps = [0.002 * g for g in range(16)]
    cf = workloads.taper_chirp_cf_device(16, radii, chirps, 4)
    rec = workloads.gaussian_uv_device_records(
        nrec, npol, gsz, lambda w: torch.clamp((w * 16).floor(), max=15), 0.15, seed=61)
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, nrec, cf, gsz, (1.0, 1.0), M, M)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    log(f"differentiable degridding: image {size}^2 x {npol}, 16 groups, {nrec} records")
    img_np = workloads.point_source_image(size, npol, num_sources=200, seed=62)
    image = torch.from_numpy(img_np.astype(np.float64)[None]).cuda()  # (1, npol, y, x)
    rng = np.random.default_rng(63)
    # target data: the operator on a perturbed image, plus N(0, 0.01) noise
    pert = image * (1.0 + 0.01 * torch.from_numpy(rng.standard_normal(image.shape)).cuda())
    with torch.no_grad():
        data = degrid(image_to_model(pert, gs), gs, rec)
    data = data + 0.01 * torch.complex(torch.from_numpy(rng.standard_normal(data.shape)),
                                       torch.from_numpy(rng.standard_normal(data.shape))).cuda()
    wts = torch.from_numpy(rng.uniform(0.5, 1.5, data.shape)).cuda()

    x = image.clone().requires_grad_(True)
    pred = degrid(image_to_model(x, gs), gs, rec)
    loss = (wts * (pred - data).abs() ** 2).sum()
    loss.backward()
    grad = x.grad.detach()
    p0 = pred.detach()
    resid = p0 - data
    log(f"  loss {float(loss.detach()):.6e}, max |grad| {float(grad.abs().max()):.3e}")

    # central differences of the loss: the predictions are linear in the
    # image, so L(x +- h e_k) = sum w |r +- h a_k|^2 with a_k the operator
    # applied to the unit image e_k
    h = 1e-3
    worst = 0.0
    for k in rng.choice(image.numel(), 64, replace=False):
        idx = np.unravel_index(int(k), image.shape)
        e = torch.zeros_like(image, dtype=torch.complex128)
        e[idx] = 1.0
        gs.set_model(hpg.Device.Serial, _model_pass(gs, e, hpg.FFTSign.NEGATIVE,
                                                    hpg.ShiftDirection.FORWARD, True))
        a = gs.degrid_forward(rec)
        lp = float((wts * (resid + h * a).abs() ** 2).sum())
        lm = float((wts * (resid - h * a).abs() ** 2).sum())
        g_fd = (lp - lm) / (2 * h)
        g = float(grad[idx])
        rel = abs(g_fd - g) / abs(g)
        worst = max(worst, rel)
        assert rel <= 1e-4, (idx, g_fd, g)
    log(f"  finite differences on 64 pixels: worst relative error {worst:.2e}")
    gs.release()
