"""Synthetic inputs for the gridding workloads (fixed seeds, numpy only).

``gaussian_cf`` / ``config1_visibilities`` build the 1024x1024 single-pol
workload: 1,000,000 VisData<1> records with N(0,1) values, weights in
(0.5, 1.5], u, v uniform within +-0.45 of the grid extent, w = 0, and one
7x7-support separable Gaussian CF (sigma 1.5 px) at oversampling 20.
``random_cf`` / ``random_visibilities`` build small general cases (several
groups and cubes, Mueller matrices, phase gradients, both signs of w) for the
parity tests.
"""
import numpy as np

from . import CFArray, visdata_dtype

C_LIGHT = 299792458.0
PADDING = CFArray.padding


def cf_extent(radius, oversampling):
    """CF extent (oversampled units) of a support of radius `radius`, with
    the CF padding on both edges: radii() of it returns `radius`."""
    return (2 * radius + 1 + 2 * PADDING) * oversampling


def gaussian_cf(radius=3, oversampling=20, sigma=1.5, nmueller=1, ncube=1):
    """Separable Gaussian taper g(tx) g(ty), g(t) = exp(-t^2 / (2 sigma^2)),
    t the distance in pixels from the centre tap; imaginary part 0."""
    ext = cf_extent(radius, oversampling)
    centre = (radius + PADDING) * oversampling
    t = (np.arange(ext) - centre) / oversampling
    g = np.exp(-t * t / (2.0 * sigma * sigma))
    plane = np.outer(g, g).astype(np.complex64)  # [y][x]
    vals = np.broadcast_to(plane, (ncube, nmueller, ext, ext)).copy()
    return CFArray(oversampling, [vals])


def config1_visibilities(num_vis=1_000_000, grid=1024, extent_frac=0.45, seed=20261202):
    """VisData<1> records of the single-pol workload (grid_scale 1, frequency
    c, so a record's grid position is u + grid/2)."""
    rng = np.random.default_rng(seed)
    rec = np.zeros(num_vis, dtype=visdata_dtype(1))
    vals = rng.standard_normal((num_vis, 2), dtype=np.float32)
    rec["visibilities"][:, 0] = vals[:, 0] + 1j * vals[:, 1]
    rec["weights"][:, 0] = (1.5 - rng.random(num_vis, dtype=np.float32)).astype(np.float32)
    lim = extent_frac * grid
    rec["uvw"][:, 0] = rng.uniform(-lim, lim, num_vis)
    rec["uvw"][:, 1] = rng.uniform(-lim, lim, num_vis)
    rec["frequency"] = C_LIGHT
    return rec


def random_cf(oversampling, radii, nmueller=1, ncube=1, seed=1, hermitian=False):
    """CF groups with uniform random complex taps in [-1, 1]^2."""
    rng = np.random.default_rng(seed)
    groups = []
    for r in radii:
        rx, ry = (r, r) if np.isscalar(r) else r
        ex, ey = cf_extent(rx, oversampling), cf_extent(ry, oversampling)
        v = rng.uniform(-1, 1, (ncube, nmueller, ey, ex, 2)).astype(np.float32)
        groups.append((v[..., 0] + 1j * v[..., 1]).astype(np.complex64))
    return CFArray(oversampling, groups)


def random_visibilities(num_vis, npol, grid_size, grid_scale, cf, seed=2,
                        gradient=False, phase=False, w_signs=True, margin=0,
                        freq_scale=1.0):
    """Records spread over the grid (some near or beyond the edges when
    margin < 0), random groups and cubes, optional phase gradients/phases."""
    rng = np.random.default_rng(seed)
    rec = np.zeros(num_vis, dtype=visdata_dtype(npol))
    v = rng.uniform(-1, 1, (num_vis, npol, 2)).astype(np.float32)
    rec["visibilities"] = v[..., 0] + 1j * v[..., 1]
    rec["weights"] = rng.uniform(0, 1, (num_vis, npol)).astype(np.float32)
    freq = C_LIGHT * freq_scale
    rec["frequency"] = freq
    ng = cf.num_groups()
    grp = rng.integers(0, ng, num_vis)
    cubes = np.array([cf.extents(g)[3] for g in range(ng)])
    rec["cf_index"][:, 1] = grp
    rec["cf_index"][:, 0] = (rng.random(num_vis) * cubes[grp]).astype(np.uint32)
    rec["grid_cube"] = rng.integers(0, grid_size[3], num_vis)
    for d in range(2):
        half = grid_size[d] / 2.0 - margin
        pos = rng.uniform(-half, half, num_vis)
        rec["uvw"][:, d] = pos / (grid_scale[d] * freq / C_LIGHT)
    if w_signs:
        rec["uvw"][:, 2] = rng.uniform(-1, 1, num_vis)
    if gradient:
        rec["cf_phase_gradient"] = rng.uniform(-3.141, 3.141, (num_vis, 2))
    if phase:
        rec["phase"] = rng.uniform(-np.pi, np.pi, num_vis)
    return rec


# ---- scaled-down forms of the multi-polarization workloads ----------------
#
# Full 4x4 Mueller matrices: row r (grid polarization) uses CF Mueller
# elements 4r..4r+3 for the four visibility polarizations, conjugates the same.
MUELLER_FULL = [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [12, 13, 14, 15]]


def wproj_cf(oversampling, radii, w_values, nmueller=16, ncube=1, seed=3,
             phase_screen=False):
    """W-projection-style CF groups: group g has support radius radii[g] and
    taps exp(-rho^2 / (2 (r/2.5)^2)) * exp(i pi w_g rho^2) (rho in pixels
    from the centre tap), one slightly different amplitude per Mueller
    element (scaled by 1 / (1 + m)) and cube; with phase_screen, each
    (cube, Mueller element) is also multiplied by a random unit-modulus
    screen exp(i phi(x, y))."""
    rng = np.random.default_rng(seed)
    groups = []
    for r, w in zip(radii, w_values):
        ext = cf_extent(r, oversampling)
        centre = (r + PADDING) * oversampling
        t = (np.arange(ext) - centre) / oversampling
        rho2 = t[None, :] ** 2 + t[:, None] ** 2  # [y][x]
        sigma = max(r, 1) / 2.5
        base = np.exp(-rho2 / (2.0 * sigma * sigma)) * np.exp(1j * np.pi * w * rho2)
        vals = np.empty((ncube, nmueller, ext, ext), dtype=np.complex64)
        for c in range(ncube):
            for m in range(nmueller):
                v = base / (1.0 + m + 0.5 * c)
                if phase_screen:
                    v = v * np.exp(1j * rng.uniform(-np.pi, np.pi, v.shape))
                vals[c, m] = v.astype(np.complex64)
        groups.append(vals)
    return CFArray(oversampling, groups)


def gaussian_uv_visibilities(num_vis, npol, grid_size, cf, sigma_frac, w_max=1.0,
                             weight_range=(0.5, 1.5), seed=5):
    """VisData<npol> records with u, v ~ N(0, (sigma_frac * extent)^2) pixels
    (grid_scale 1, frequency c), w ~ U(0, w_max) selecting the CF group
    floor(ngroups * w / w_max), channel (grid cube and CF cube) uniform,
    values N(0, 1) and weights uniform in weight_range."""
    rng = np.random.default_rng(seed)
    rec = np.zeros(num_vis, dtype=visdata_dtype(npol))
    v = rng.standard_normal((num_vis, npol, 2)).astype(np.float32)
    rec["visibilities"] = v[..., 0] + 1j * v[..., 1]
    lo, hi = weight_range
    rec["weights"] = rng.uniform(lo, hi, (num_vis, npol)).astype(np.float32)
    rec["frequency"] = C_LIGHT
    for d in range(2):
        rec["uvw"][:, d] = rng.normal(0.0, sigma_frac * grid_size[d], num_vis)
    w = rng.uniform(0.0, w_max, num_vis)
    rec["uvw"][:, 2] = w
    ng = cf.num_groups()
    grp = np.minimum((ng * w / w_max).astype(np.int64), ng - 1)
    rec["cf_index"][:, 1] = grp
    ncube = min(grid_size[3], min(cf.extents(g)[3] for g in range(ng)))
    cube = rng.integers(0, ncube, num_vis)
    rec["grid_cube"] = cube
    rec["cf_index"][:, 0] = cube
    return rec


def point_source_image(size, npol, num_sources=200, noise=1e-4, seed=11):
    """(npol, size, size) float32 sky: point sources at random pixels with
    log-uniform flux in [1e-3, 1], plus N(0, noise^2) noise."""
    rng = np.random.default_rng(seed)
    img = (noise * rng.standard_normal((npol, size, size))).astype(np.float32)
    ys = rng.integers(0, size, num_sources)
    xs = rng.integers(0, size, num_sources)
    flux = np.exp(rng.uniform(np.log(1e-3), 0.0, num_sources)).astype(np.float32)
    for p in range(npol):
        img[p, ys, xs] += flux
    return img
