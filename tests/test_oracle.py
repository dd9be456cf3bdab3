"""Pins of the CPU oracle (oracle/hpg_oracle.cc) -- CPU only.

What pins each oracle function (a plausible mistake -- a dropped term, a
wrong sign or index, a transposed operand -- fails at least one check):

  compute_vis_coord / Vis / skip rule
      GridOne footprint closed form ll=(9523,8522), ur=(9613,8612) and
      golden values (reference tests/gridone.h5); the 11x11 Gridding sweep
      closed form, which includes the lrint tie 7.5 -> 8; records straddling
      the grid edge are skipped; a one-cell shift of u moves the footprint by
      one pixel.
  grid_vis (FP32 products, FP64 sums)
      gridone.h5 (single visibility, exact up to the libm that made the CF);
      brute-force accumulation of several visibilities; w > 0 uses the
      conjugate CF and the mueller_indexes row (conjugation symmetry).
  weights
      closed form sum_v |sum_XY cf| * w_v.
  phase gradient
      a ramp across the footprint: the screen phase at footprint pixel X is
      g * ((X - r) * os + f)  (the ratio of gradient / no-gradient grids).
  degrid_vis
      ResidualVisibilities: no model -> residual equals input exactly;
      ZeroModel: zero model leaves gridding unchanged; a constant CF
      predicts the mean of the model over the footprint; linearity.
      Exact FP64 degrid values (Kokkos complex division rounding) are
      parity unpinned -- the reference has no golden degrid values.
  normalize_by_weights   closed form.
  FFT                    numpy.fft (both signs, power-of-two and not) and the
                         OneFFT sign test.
  shift                  numpy.fft.fftshift / ifftshift and ShiftCycle.
  OpenMP baseline        equals the Serial oracle to FP64 rounding.
"""
import numpy as np
import pytest

import refcases as rc


def test_visdata_layout(oracle_mod):
    for n, size in [(1, 88), (2, 96), (3, 112), (4, 120)]:
        assert oracle_mod.visdata_dtype(n).itemsize == size


def _gridone(oracle_mod, cf_vals):
    g = rc.GRIDONE
    ext, _ = oracle_mod.cone_cf(1, g["oversampling"], g["radius"])
    o = oracle_mod.Oracle(g["grid_size"], g["grid_scale"], [[0]])
    o.set_cf(g["oversampling"], [(ext, cf_vals)])
    rec = rc.one_record(oracle_mod.visdata_dtype(1), g["vis"], g["weight"], g["freq"],
                        g["uvw"])
    o.grid_visibilities(rec, True)
    return o


def test_gridone_footprint_and_golden(oracle_mod, gridone_golden):
    g = rc.GRIDONE
    (gx, _, _), (gy, _, _) = rc.coords(g["grid_size"], g["grid_scale"], g["freq"],
                                       g["uvw"], g["radius"], g["oversampling"])
    assert (gx, gy) == g["ll"]
    assert (gx + 2 * g["radius"], gy + 2 * g["radius"]) == g["ur"]
    _, cf = oracle_mod.cone_cf(1, g["oversampling"], g["radius"])
    o = _gridone(oracle_mod, cf)
    # footprint: exactly the (2r+1)^2 window is touched
    win = o.grid_window(g["ll"][0] - 5, g["ll"][1] - 5, 101, 101)
    nz = np.argwhere(win != 0)
    assert nz[:, 1].min() + g["ll"][0] - 5 >= g["ll"][0]
    assert nz[:, 1].max() + g["ll"][0] - 5 <= g["ur"][0]
    assert nz[:, 0].min() + g["ll"][1] - 5 >= g["ll"][1]
    assert nz[:, 0].max() + g["ll"][1] - 5 <= g["ur"][1]
    got = o.grid_window(g["ll"][0], g["ll"][1], 91, 91).T  # [x][y]
    gold = gridone_golden
    # same non-zero pattern as the golden file
    assert np.array_equal(got != 0, gold != 0)
    # compared range [ll, ur) as in the reference test (line 1433)
    eq = got[:90, :90] == gold[:90, :90]
    assert eq.mean() > 0.98
    # every mismatch is a float product fl32(c * (1-1j)) whose CF tap differs
    # from this host's libm cone value by <= 2 ulp per component
    (gxx, mx, nx_), (gyy, my, ny_) = rc.coords(
        g["grid_size"], g["grid_scale"], g["freq"], g["uvw"], g["radius"],
        g["oversampling"])
    f32 = np.float32

    def nudge(v, k):
        for _ in range(abs(k)):
            v = np.nextafter(v, f32(np.inf) if k > 0 else f32(-np.inf))
        return v

    for x, y in np.argwhere(~eq):
        xf = (x + mx) * g["oversampling"] + nx_
        yf = (y + my) * g["oversampling"] + ny_
        c = cf[0, 0, yf, xf]
        a, b = f32(c.real), f32(c.imag)
        ok = any(
            complex(f32(nudge(a, da) + nudge(b, db)), f32(nudge(b, db) - nudge(a, da)))
            == gold[x, y] for da in range(-2, 3) for db in range(-2, 3))
        assert ok, (x, y)
    # weights: |sum of the footprint taps| * 1
    w = o.grid_weights()[0, 0]
    s = 0j
    for X in range(91):
        for Y in range(91):
            s += complex(cf[0, 0, (Y + my) * 20 + ny_, (X + mx) * 20 + nx_])
    assert w == pytest.approx(abs(s), rel=1e-13)


def test_gridding_sweep(oracle_mod):
    """Gridding test (GridderStateUnitTests.cc:1273-1341): 11x11 sub-pixel
    positions, exact float products at the closed-form CF taps."""
    g = rc.GRIDDING
    ext, cf = oracle_mod.cone_cf(1, g["oversampling"], g["radius"])
    fine = g["grid_scale"][0] * g["oversampling"]
    ties = 0
    for i in range(g["oversampling"] + 1):
        for j in range(g["oversampling"] + 1):
            uvw = (i * fine, j * fine, 0.0)
            o = oracle_mod.Oracle(g["grid_size"], g["grid_scale"], [[0]])
            o.set_cf(g["oversampling"], [(ext, cf)])
            rec = rc.one_record(oracle_mod.visdata_dtype(1), g["vis"], g["weight"],
                                g["freq"], uvw)
            o.grid_visibilities(rec, True)
            got = o.grid_values()[0, 0]
            exp = rc.single_vis_expectation(cf[0, 0], g["grid_size"], g["grid_scale"],
                                            g["freq"], uvw, g["radius"],
                                            g["oversampling"], g["vis"])
            assert np.array_equal(got, exp), (i, j)
            pos = g["grid_scale"][0] * uvw[0] + g["grid_size"][0] / 2.0
            ties += pos == 7.5
    assert ties == 11  # the i = 5 column hits the round-half-even tie


def test_skip_rule_and_translation(oracle_mod):
    """Records whose support leaves the grid are ignored with their weight;
    moving u by one cell moves the footprint by one pixel."""
    os_, r = 10, 3
    ext, cf = oracle_mod.cone_cf(1, os_, r)
    gs = (16, 16, 1, 1)
    dt = oracle_mod.visdata_dtype(1)

    def grid_at(u):
        o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0]])
        o.set_cf(os_, [(ext, cf)])
        o.grid_visibilities(rc.one_record(dt, 1 - 1j, 1.0, rc.C_LIGHT, (u, 0.3, 0.0)))
        return o.grid_values()[0, 0], o.grid_weights()[0, 0]

    g0, w0 = grid_at(0.2)
    g1, w1 = grid_at(1.2)
    assert np.array_equal(np.roll(g0, 1, axis=1), g1)
    assert w0 == w1 > 0
    # support corner at x = 8 - 3 + 5 = 10: 10 + 7 = 17 > 16 -> skipped
    g2, w2 = grid_at(5.2)
    assert not g2.any() and w2 == 0
    # corner at 8 + 4 - 3 = 9: 9 + 7 = 16 <= 16 -> kept
    g3, w3 = grid_at(4.2)
    assert g3.any() and w3 > 0


def _brute_grid(rec, cf_groups, os_, gs, scale, M, Mc):
    """Independent numpy gridding of VisData records (w sign, Mueller rows,
    phase, weights; no phase gradient) in the reference's Serial order."""
    nx, ny, nr, nc = gs
    grid = np.zeros((nc, nr, ny, nx), dtype=np.complex128)
    wts = np.zeros((nc, nr))
    npol = rec["vis"].shape[1]
    for k in range(rec.shape[0]):
        grp = rec["cf_index"][k, 1]
        cube = rec["cf_index"][k, 0]
        vals = cf_groups[grp]
        e = vals.shape
        rx = ((e[3] - 4 * os_) // os_) // 2
        ry = ((e[2] - 4 * os_) // os_) // 2
        (gx, mx, ix), (gy, my, iy) = rc.coords(
            (nx, ny), scale, rec["frequency"][k], rec["uvw"][k], 0, os_)
        gx -= rx
        gy -= ry
        sx, sy = 2 * rx + 1, 2 * ry + 1
        if not (0 <= gx and gx + sx <= nx and 0 <= gy and gy + sy <= ny):
            continue
        ph = rec["phase"][k]
        phasor = complex(np.cos(ph), np.sin(ph))
        pos_w = rec["uvw"][k, 2] > 0
        gcube = rec["grid_cube"][k]
        gvis = []
        for v in range(npol):
            t = complex(rec["vis"][k, v]) * phasor * float(rec["weights"][k, v])
            gvis.append(np.complex64(t))
        for gpol in range(nr):
            row = (M if pos_w else Mc)[gpol]
            wsum = [0j] * npol
            for X in range(sx):
                for Y in range(sy):
                    acc = 0j
                    for v in range(npol):
                        m = row[v]
                        if m < 0:
                            continue
                        c = np.complex64(vals[cube, m, (Y + my) * os_ + iy,
                                              (X + mx) * os_ + ix])
                        if pos_w:
                            c = np.conj(c)
                        p = np.complex64(c * gvis[v])
                        acc += complex(p)
                        wsum[v] += complex(c)
                    grid[gcube, gpol, gy + Y, gx + X] += acc
            wts[gcube, gpol] += sum(abs(wsum[v]) * float(rec["weights"][k, v])
                                    for v in range(npol))
    return grid, wts


def test_brute_force_multi(oracle_mod):
    """Two groups, two CF cubes, two grid cubes, 2 pols, off-diagonal
    Mueller rows, both signs of w, random phases."""
    os_ = 4
    rng = np.random.default_rng(5)
    groups = []
    for r, nm in [(1, 4), (2, 4)]:
        e = (2 * r + 1 + 4) * os_
        v = rng.uniform(-1, 1, (2, nm, e, e, 2)).astype(np.float32)
        groups.append((v[..., 0] + 1j * v[..., 1]).astype(np.complex64))
    gs = (12, 10, 2, 2)
    scale = (0.7, 0.9)
    M = [[0, 1], [2, 3]]
    Mc = [[1, 0], [3, 2]]
    dt = oracle_mod.visdata_dtype(2)
    n = 40
    rec = np.zeros(n, dtype=dt)
    rec["vis"] = (rng.uniform(-1, 1, (n, 2)) + 1j * rng.uniform(-1, 1, (n, 2))).astype(np.complex64)
    rec["weights"] = rng.uniform(0, 1, (n, 2)).astype(np.float32)
    rec["frequency"] = rc.C_LIGHT * 1.3
    rec["phase"] = rng.uniform(-3, 3, n)
    rec["uvw"][:, 0] = rng.uniform(-7, 7, n)
    rec["uvw"][:, 1] = rng.uniform(-5, 5, n)
    rec["uvw"][:, 2] = rng.uniform(-1, 1, n)
    rec["grid_cube"] = rng.integers(0, 2, n)
    rec["cf_index"][:, 0] = rng.integers(0, 2, n)
    rec["cf_index"][:, 1] = rng.integers(0, 2, n)
    o = oracle_mod.Oracle(gs, scale, M, Mc)
    o.set_cf(os_, [((g.shape[3], g.shape[2], g.shape[1], g.shape[0]), g) for g in groups])
    o.grid_visibilities(rec, True)
    exp_g, exp_w = _brute_grid(rec, groups, os_, gs, scale, M, Mc)
    got_g, got_w = o.grid_values(), o.grid_weights()
    assert np.abs(exp_g).max() > 0
    np.testing.assert_allclose(got_g, exp_g, rtol=0, atol=1e-12 * np.abs(exp_g).max())
    np.testing.assert_allclose(got_w, exp_w, rtol=1e-12)


def test_phase_gradient_ramp(oracle_mod):
    """With a phase gradient (gx, gy) the footprint pixel (X, Y) is the
    no-gradient value times fl32(exp(i (gx*((X-r)*os+fx) + gy*((Y-r)*os+fy))))."""
    os_, r = 8, 2
    ext, cf = oracle_mod.cone_cf(1, os_, r)
    gs = (12, 12, 1, 1)
    dt = oracle_mod.visdata_dtype(1)
    uvw = (0.37, -0.81, 0.0)
    grads = (0.013, -0.021)

    def grid(grad):
        o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0]])
        o.set_cf(os_, [(ext, cf)])
        o.grid_visibilities(rc.one_record(dt, 1 + 0j, 1.0, rc.C_LIGHT, uvw, gradient=grad))
        return o.grid_values()[0, 0]

    g0 = grid((0.0, 0.0))
    g1 = grid(grads)
    (gx, _, _), (gy, _, _) = rc.coords(gs, (1.0, 1.0), rc.C_LIGHT, uvw, r, os_)
    fx = int(np.rint((np.rint(uvw[0] + 6) - (uvw[0] + 6)) * os_))
    fy = int(np.rint((np.rint(uvw[1] + 6) - (uvw[1] + 6)) * os_))
    for X in range(2 * r + 1):
        for Y in range(2 * r + 1):
            ph = grads[0] * ((X - r) * os_ + fx) + grads[1] * ((Y - r) * os_ + fy)
            z = g0[gy + Y, gx + X]
            if z == 0:
                continue
            ratio = g1[gy + Y, gx + X] / z
            assert abs(ratio - np.exp(1j * ph)) < 1e-6


def test_conjugation_symmetry(oracle_mod):
    """w > 0 grids through conj(CF) and the mueller_indexes rows; w <= 0
    through CF and the conjugate rows."""
    os_ = 5
    rng = np.random.default_rng(9)
    e = (3 + 4) * os_
    v = rng.uniform(-1, 1, (1, 2, e, e, 2)).astype(np.float32)
    vals = (v[..., 0] + 1j * v[..., 1]).astype(np.complex64)
    dt = oracle_mod.visdata_dtype(1)
    gs = (10, 10, 1, 1)

    def grid(cfv, M, Mc, w):
        o = oracle_mod.Oracle(gs, (1.0, 1.0), M, Mc)
        o.set_cf(os_, [((e, e, 2, 1), cfv)])
        o.grid_visibilities(rc.one_record(dt, 1 + 0j, 1.0, rc.C_LIGHT, (0.2, 0.1, w)))
        return o.grid_values()[0, 0]

    a = grid(vals, [[0]], [[1]], 0.5)            # conj(CF[m=0])
    b = grid(np.conj(vals), [[1]], [[0]], -0.5)  # conj(CF) via the other row
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)


def test_normalize_closed_form(oracle_mod):
    os_ = 4
    ext, cf = oracle_mod.cone_cf(1, os_, 1)
    o = oracle_mod.Oracle((8, 8, 1, 2), (1.0, 1.0), [[0]])
    o.set_cf(os_, [(ext, cf)])
    dt = oracle_mod.visdata_dtype(1)
    recs = np.concatenate([
        rc.one_record(dt, 1 - 1j, 0.7, rc.C_LIGHT, (0.3, 0.1, 0.0), grid_cube=0),
        rc.one_record(dt, 2 + 1j, 0.4, rc.C_LIGHT, (-1.2, 0.8, 0.0), grid_cube=1)])
    o.grid_visibilities(recs, True)
    g, w = o.grid_values(), o.grid_weights()
    o.normalize_by_weights(2.5)
    n = o.grid_values()
    for c in range(2):
        d = 2.5 * w[c, 0]
        assert np.array_equal(n[c, 0].real, g[c, 0].real / d)
        assert np.array_equal(n[c, 0].imag, g[c, 0].imag / d)


@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (5, 7)])
@pytest.mark.parametrize("sign", ["NEGATIVE", "POSITIVE"])
def test_fft_vs_numpy(oracle_mod, shape, sign):
    nx, ny = shape
    gs = (nx, ny, 2, 2)
    o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0], [0]])
    rng = np.random.default_rng(3)
    m = rng.standard_normal((2, 2, ny, nx)) + 1j * rng.standard_normal((2, 2, ny, nx))
    o.set_model(m)
    o.apply_fft(norm=0.5, sign=sign, which="model")
    got = o.model_values()
    if sign == "NEGATIVE":
        exp = np.fft.fft2(m, axes=(-2, -1))
    else:
        exp = np.fft.ifft2(m, axes=(-2, -1)) * (nx * ny)
    np.testing.assert_allclose(got, exp / 0.5, rtol=0, atol=1e-12 * np.abs(exp).max())


def test_one_fft_sign(oracle_mod):
    """OneFFT (GridderStateUnitTests.cc:1704-1740)."""
    gs = (2048, 2048, 1, 1)
    o = oracle_mod.Oracle(gs, (0.12, -0.34), [[0]])
    m = np.zeros((1, 1, 2048, 2048), dtype=np.complex128)
    m[0, 0, 1108, 1024] = 0.825
    o.set_model(m)
    o.apply_fft(sign="NEGATIVE", which="model")
    o.shift("FORWARD", which="model")
    assert o.model_values()[0, 0, 10, 1070].imag < 0


@pytest.mark.parametrize("shape", [(20, 20), (19, 19), (18, 19), (19, 18)])
def test_shift_vs_numpy_and_cycle(oracle_mod, shape):
    nx, ny = shape
    gs = (nx, ny, 1, 2)
    o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0]])
    rng = np.random.default_rng(4)
    m = rng.standard_normal((2, 1, ny, nx)) + 1j * rng.standard_normal((2, 1, ny, nx))
    o.set_model(m)
    o.shift("FORWARD", which="model")
    f = o.model_values()
    assert np.array_equal(f, np.fft.fftshift(m, axes=(-2, -1)))
    o.shift("BACKWARD", which="model")
    assert np.array_equal(o.model_values(), m)
    o.shift("BACKWARD", which="model")
    assert np.array_equal(o.model_values(), np.fft.ifftshift(m, axes=(-2, -1)))


def test_residual_visibilities_and_zero_model(oracle_mod):
    """ResidualVisibilities (1743-1815) and ZeroModel (1523-1596)."""
    os_, r = 10, 3
    ext, cf = oracle_mod.cone_cf(1, os_, r)
    gs = (20, 20, 2, 1)
    M = [[0, -1], [-1, 0]]
    dt = oracle_mod.visdata_dtype(2)
    rec = rc.one_record(dt, [1 - 1j, -1 + 1j], [1.0, 0.5], 3.693e9, (0.0, 0.0, 0.1))
    o = oracle_mod.Oracle(gs, (0.0476591, 0.0476591), M)
    o.set_cf(os_, [(ext, cf)])
    res = o.grid_visibilities(rec, True, do_degrid=True, do_grid=True)
    assert np.array_equal(res, rec)
    g_nomodel = o.grid_values()
    o2 = oracle_mod.Oracle(gs, (0.0476591, 0.0476591), M)
    o2.set_cf(os_, [(ext, cf)])
    o2.set_model(np.zeros(o2.grid_shape, dtype=np.complex128))
    o2.grid_visibilities(rec, True, do_degrid=True, do_grid=True)
    assert np.abs(g_nomodel).max() > 0
    assert np.array_equal(o2.grid_values(), g_nomodel)


def test_degrid_constant_cf_mean(oracle_mod):
    """With every CF tap 1 and no phase, the prediction is the mean of the
    model over the footprint."""
    os_, r = 4, 2
    e = (2 * r + 1 + 4) * os_
    ones = np.ones((1, 1, e, e), dtype=np.complex64)
    gs = (16, 16, 1, 1)
    o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0]])
    o.set_cf(os_, [((e, e, 1, 1), ones)])
    rng = np.random.default_rng(6)
    model = rng.standard_normal((1, 1, 16, 16)) + 1j * rng.standard_normal((1, 1, 16, 16))
    o.set_model(model)
    dt = oracle_mod.visdata_dtype(1)
    for uvw in [(0.3, -1.6, 0.0), (-2.4, 2.2, 0.5)]:
        rec = rc.one_record(dt, 0j, 1.0, rc.C_LIGHT, uvw)
        pred = o.grid_visibilities(rec, False, do_degrid=True, do_grid=False)
        (gx, _, _), (gy, _, _) = rc.coords(gs, (1.0, 1.0), rc.C_LIGHT, uvw, r, os_)
        mean = model[0, 0, gy:gy + 5, gx:gx + 5].mean()
        assert abs(complex(pred["vis"][0, 0]) - mean) < 1e-6 * abs(mean) + 1e-7


def test_degrid_linearity(oracle_mod):
    os_ = 5
    rng = np.random.default_rng(8)
    e = (5 + 4) * os_
    v = rng.uniform(-1, 1, (1, 1, e, e, 2)).astype(np.float32)
    vals = (v[..., 0] + 1j * v[..., 1]).astype(np.complex64)
    gs = (20, 20, 1, 1)
    dt = oracle_mod.visdata_dtype(1)
    recs = np.concatenate([rc.one_record(dt, 0j, 1.0, rc.C_LIGHT, (u, w, w))
                           for u, w in [(0.3, -1.1), (2.2, 0.7), (-3.1, 2.9)]])
    m1 = rng.standard_normal((1, 1, 20, 20)) + 1j * rng.standard_normal((1, 1, 20, 20))
    m2 = rng.standard_normal((1, 1, 20, 20)) + 1j * rng.standard_normal((1, 1, 20, 20))

    def pred(m):
        o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0]])
        o.set_cf(os_, [((e, e, 1, 1), vals)])
        o.set_model(m)
        return o.grid_visibilities(recs, False, do_degrid=True, do_grid=False)["vis"][:, 0]

    p1, p2, p12 = pred(m1), pred(m2), pred(m1 + 2.0 * m2)
    np.testing.assert_allclose(p12, p1 + 2.0 * p2, rtol=0, atol=1e-5)


def test_openmp_baseline_matches_serial(oracle_mod):
    os_ = 8
    rng = np.random.default_rng(11)
    e = (7 + 4) * os_
    v = rng.uniform(-1, 1, (1, 1, e, e, 2)).astype(np.float32)
    vals = (v[..., 0] + 1j * v[..., 1]).astype(np.complex64)
    gs = (64, 64, 1, 1)
    dt = oracle_mod.visdata_dtype(1)
    n = 3000
    rec = np.zeros(n, dtype=dt)
    rec["vis"][:, 0] = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    rec["weights"][:, 0] = rng.uniform(0.5, 1.5, n).astype(np.float32)
    rec["frequency"] = rc.C_LIGHT
    rec["uvw"][:, :2] = rng.uniform(-28, 28, (n, 2))
    rec["cf_phase_gradient"] = rng.uniform(-0.1, 0.1, (n, 2))
    out = []
    for omp in (False, True):
        o = oracle_mod.Oracle(gs, (1.0, 1.0), [[0]])
        o.set_cf(os_, [((e, e, 1, 1), vals)])
        o.grid_visibilities(rec, True, openmp=omp)
        out.append((o.grid_values(), o.grid_weights()))
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=0,
                               atol=1e-12 * np.abs(out[0][0]).max())
    np.testing.assert_allclose(out[1][1], out[0][1], rtol=1e-12)


def _adjoint_case(oracle_mod, M, Mc, seed):
    """Random complex CFs (2 groups, 2 cubes), 2 visibility polarizations,
    2 grid rows, records with both signs of w, random phases and phase
    gradients."""
    rng = np.random.default_rng(seed)
    os_ = 4
    npol = 2
    gs = (24, 20, 2, 2)
    groups = []
    for r in (2, 3):
        e = (2 * r + 1 + 4) * os_
        v = rng.uniform(-1, 1, (2, 4, e, e, 2)).astype(np.float32)
        groups.append(((e, e, 4, 2), (v[..., 0] + 1j * v[..., 1]).astype(np.complex64)))
    dt = oracle_mod.visdata_dtype(npol)
    n = 24
    rec = np.zeros(n, dtype=dt)
    rec["weights"] = 1.0
    rec["frequency"] = rc.C_LIGHT
    rec["phase"] = rng.uniform(-np.pi, np.pi, n)
    rec["uvw"][:, 0] = rng.uniform(-6, 6, n)
    rec["uvw"][:, 1] = rng.uniform(-5, 5, n)
    rec["uvw"][:, 2] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0) * rng.uniform(0.1, 1, n)
    rec["cf_index"][:, 1] = rng.integers(0, 2, n)
    rec["cf_index"][:, 0] = rng.integers(0, 2, n)
    rec["grid_cube"] = rng.integers(0, 2, n)
    rec["cf_phase_gradient"] = rng.uniform(-0.3, 0.3, (n, 2))
    return gs, os_, groups, rec, rng


@pytest.mark.parametrize("conj_differs", [False, True])
def test_degrid_adjoint_identity(oracle_mod, conj_differs):
    """Pins the oracle's degridder (degrid_vis, src/hpg_core.hpp:579-682)
    against its gridder (grid_vis, 791-886) on the CPU: for every record,

        <degrid_{M,Mc}(m), r> = <m, grid_{Mc,M}(r phasor / conj(W))>

    with W_v the degridding normalization.  degrid_vis reads Mueller row
    (w > 0 ? Mc : M) and conj(CF) relative to grid_vis, which reads
    (w > 0 ? M : Mc), so the adjoint of degridding with (M, Mc) is gridding
    with the roles of M and Mc exchanged.  W_v is not taken from the
    degridder: it is the conjugated sum over grid rows and footprint of the
    gridder's own output for a unit value in polarization v (phase 0, no
    phase gradient), so a wrong conjugation, Mueller row, phase screen sign,
    phasor or normalization in either function breaks the identity.  Both
    sides hold FP32-rounded values (predictions, gvis, CF products):
    tolerance 1e-5 of |degrid(m)| |r|."""
    M = [[0, 1], [2, 3]]
    Mc = [[1, 0], [3, 2]] if conj_differs else M
    gs, os_, groups, rec, rng = _adjoint_case(oracle_mod, M, Mc, 90 + int(conj_differs))
    shape = (gs[3], gs[2], gs[1], gs[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    checked = 0
    for i in range(rec.shape[0]):
        one = rec[i:i + 1].copy()
        # degridding normalization from the gridder with the exchanged rows
        W = np.zeros(2, dtype=np.complex128)
        for v in range(2):
            unit = one.copy()
            unit["vis"][0, :] = 0
            unit["vis"][0, v] = 1.0
            unit["phase"] = 0.0
            unit["cf_phase_gradient"] = 0.0
            og = oracle_mod.Oracle(gs, (1.0, 1.0), Mc, M)
            og.set_cf(os_, groups)
            og.grid_visibilities(unit, False)
            W[v] = np.conj(og.grid_values().sum())
        if not np.all(np.abs(W) > 0):
            continue  # record off the grid
        od = oracle_mod.Oracle(gs, (1.0, 1.0), M, Mc)
        od.set_cf(os_, groups)
        od.set_model(model)
        pred = od.grid_visibilities(one, False, do_degrid=True, do_grid=False)["vis"][0]
        r = (rng.standard_normal(2) + 1j * rng.standard_normal(2))
        lhs = np.vdot(pred.astype(np.complex128), r)
        adj = one.copy()
        adj["vis"][0, :] = (r / np.conj(W)).astype(np.complex64)
        oa = oracle_mod.Oracle(gs, (1.0, 1.0), Mc, M)
        oa.set_cf(os_, groups)
        oa.grid_visibilities(adj, False)
        rhs = np.vdot(model, oa.grid_values())
        scale = np.linalg.norm(pred) * np.linalg.norm(r)
        assert abs(lhs - rhs) <= 1e-5 * scale, (i, lhs, rhs)
        checked += 1
    assert checked >= 10


def test_degrid_adjoint_identity_detects_planted_errors(oracle_mod):
    """The identity above fails for plausible mistakes: predictions that
    miss the phasor, or a model read through the wrong phase-screen sign
    (degridding the conjugated-gradient record)."""
    M = [[0, 1], [2, 3]]
    gs, os_, groups, rec, rng = _adjoint_case(oracle_mod, M, M, 97)
    shape = (gs[3], gs[2], gs[1], gs[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    bad_phasor = bad_screen = 0
    for i in range(rec.shape[0]):
        one = rec[i:i + 1].copy()
        W = np.zeros(2, dtype=np.complex128)
        for v in range(2):
            unit = one.copy()
            unit["vis"][0, :] = 0
            unit["vis"][0, v] = 1.0
            unit["phase"] = 0.0
            unit["cf_phase_gradient"] = 0.0
            og = oracle_mod.Oracle(gs, (1.0, 1.0), M, M)
            og.set_cf(os_, groups)
            og.grid_visibilities(unit, False)
            W[v] = np.conj(og.grid_values().sum())
        if not np.all(np.abs(W) > 0):
            continue

        def degrid(rec1):
            od = oracle_mod.Oracle(gs, (1.0, 1.0), M, M)
            od.set_cf(os_, groups)
            od.set_model(model)
            return od.grid_visibilities(rec1, False, do_degrid=True,
                                        do_grid=False)["vis"][0].astype(np.complex128)

        r = rng.standard_normal(2) + 1j * rng.standard_normal(2)
        adj = one.copy()
        adj["vis"][0, :] = (r / np.conj(W)).astype(np.complex64)
        oa = oracle_mod.Oracle(gs, (1.0, 1.0), M, M)
        oa.set_cf(os_, groups)
        oa.grid_visibilities(adj, False)
        rhs = np.vdot(model, oa.grid_values())
        pred = degrid(one)
        scale = np.linalg.norm(pred) * np.linalg.norm(r)
        phasor = np.exp(1j * one["phase"][0])
        if abs(np.vdot(pred * phasor, r) - rhs) > 1e-3 * scale:
            bad_phasor += 1
        flipped = one.copy()
        flipped["cf_phase_gradient"] *= -1
        if abs(np.vdot(degrid(flipped), r) - rhs) > 1e-3 * scale:
            bad_screen += 1
    assert bad_phasor >= 8 and bad_screen >= 8
