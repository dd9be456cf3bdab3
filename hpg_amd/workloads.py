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


# ---- device-side generators for the full-size workloads ---------------------
#
# The full-size multi-polarization workloads hold 6-50 GB of CF taps and up to
# 5e7 records; they are generated directly in GPU memory with torch (fixed
# seeds) and handed to the state through the device routes of the C ABI.

def _torch():
    import torch
    return torch


def device_records(num_vis, npol, device="cuda"):
    """(records, fields): a zeroed torch uint8 tensor of num_vis
    VisData<npol> records on the GPU, and a dict of typed views of its
    fields (vis_re/vis_im/weights: float32 (n, npol); frequency, phase:
    float64 (n,); u, v, w: float64 (n,); grid_cube, cf_cube, cf_group: int32
    (n,); grad_x, grad_y: float64 (n,))."""
    torch = _torch()
    dt = visdata_dtype(npol)
    size = dt.itemsize
    rec = torch.zeros((num_vis, size), dtype=torch.uint8, device=device)

    def field(off, nbytes, tdt):
        return rec[:, off:off + nbytes].view(tdt)

    off_v = dt.fields["visibilities"][1]
    off_w = dt.fields["weights"][1]
    vis = field(off_v, 8 * npol, torch.float32)  # (n, 2 npol): re, im interleaved
    f = {
        "vis": vis,
        "weights": field(off_w, 4 * npol, torch.float32),
        "frequency": field(dt.fields["frequency"][1], 8, torch.float64)[:, 0],
        "phase": field(dt.fields["phase"][1], 8, torch.float64)[:, 0],
        "uvw": field(dt.fields["uvw"][1], 24, torch.float64),
        "grid_cube": field(dt.fields["grid_cube"][1], 4, torch.int32)[:, 0],
        "cf_index": field(dt.fields["cf_index"][1], 8, torch.int32),
        "gradient": field(dt.fields["cf_phase_gradient"][1], 16, torch.float64),
    }
    return rec, f


def gaussian_uv_device_records(num_vis, npol, grid_size, group_of_w, sigma_frac,
                               w_max=1.0, weight_range=(0.5, 1.5), ncube=1, cf_ncube=1,
                               seed=5, device="cuda"):
    """VisData<npol> records in GPU memory: values N(0, 1), weights uniform
    in weight_range, u, v ~ N(0, (sigma_frac * extent)^2) pixels (grid_scale
    1, frequency c), w ~ U(0, w_max), CF group group_of_w(w) (a torch
    function of the w tensor), grid cube uniform in [0, ncube), CF cube the
    grid cube when the CF has one cube per channel (cf_ncube > 1), else 0."""
    torch = _torch()
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    rec, f = device_records(num_vis, npol, device)
    f["vis"].copy_(torch.randn((num_vis, 2 * npol), generator=g, device=device))
    lo, hi = weight_range
    f["weights"].copy_(lo + (hi - lo) * torch.rand((num_vis, npol), generator=g, device=device))
    f["frequency"].fill_(C_LIGHT)
    uvw = torch.empty((num_vis, 3), dtype=torch.float64, device=device)
    uvw[:, 0] = torch.randn(num_vis, generator=g, device=device, dtype=torch.float64) \
        * (sigma_frac * grid_size[0])
    uvw[:, 1] = torch.randn(num_vis, generator=g, device=device, dtype=torch.float64) \
        * (sigma_frac * grid_size[1])
    w = torch.rand(num_vis, generator=g, device=device, dtype=torch.float64) * w_max
    uvw[:, 2] = w
    f["uvw"].copy_(uvw)
    grp = group_of_w(w).to(torch.int32)
    cube = torch.randint(0, ncube, (num_vis,), generator=g, device=device, dtype=torch.int32)
    f["grid_cube"].copy_(cube)
    f["cf_index"][:, 0] = cube if cf_ncube > 1 else 0
    f["cf_index"][:, 1] = grp
    return rec


def taper_chirp_cf_device(oversampling, radii, chirps, nmueller, ncube=1,
                          phase_screen=False, seed=3, sigma_div=2.5, device="cuda"):
    """Device (torch complex64) CF groups: group g has support radius
    radii[g] and taps exp(-rho^2 / (2 (r/sigma_div)^2)) * exp(i pi chirps[g]
    rho^2) / (1 + m + c/2) (rho in pixels from the centre tap, m the Mueller
    element, c the cube); with phase_screen each tap is also multiplied by a
    random unit-modulus factor exp(i phi), phi ~ U(-pi, pi) (seed + g)."""
    torch = _torch()
    groups = []
    for gi, (r, kappa) in enumerate(zip(radii, chirps)):
        ext = cf_extent(r, oversampling)
        centre = (r + PADDING) * oversampling
        t = (torch.arange(ext, device=device, dtype=torch.float64) - centre) / oversampling
        rho2 = t[None, :] ** 2 + t[:, None] ** 2  # [y][x]
        sigma = max(r, 1) / sigma_div
        base = torch.exp(-rho2 / (2.0 * sigma * sigma)) \
            * torch.exp(1j * np.pi * kappa * rho2)
        scale = torch.tensor([[1.0 / (1.0 + m + 0.5 * c) for m in range(nmueller)]
                              for c in range(ncube)], dtype=torch.float64, device=device)
        vals = (base[None, None] * scale[:, :, None, None]).to(torch.complex64)
        if phase_screen:
            gen = torch.Generator(device=device)
            gen.manual_seed(int(seed) * 100003 + gi)
            phi = (torch.rand(vals.shape, generator=gen, device=device) * 2.0 - 1.0) * np.pi
            vals = vals * torch.polar(torch.ones_like(phi), phi)
        groups.append(vals.contiguous())
    return CFArray(oversampling, groups)


def config2_geometry(ngroups=64, oversampling=20):
    """Full-pol W-projection workload: CF group g has radius
    4 + round(28 g / 63) (supports 9..65 pixels) and chirp coefficient
    0.02 (g + 1/2) / ngroups; records with w in [0, 1) use group
    floor(ngroups w)."""
    radii = [4 + int(round(28 * g / max(ngroups - 1, 1))) for g in range(ngroups)]
    chirps = [0.02 * (g + 0.5) / ngroups for g in range(ngroups)]
    return radii, chirps


def config5_geometry(nw=32, npa=16):
    """A/W-projection workload: 32 w bins x 16 parallactic-angle bins,
    group g = wbin * npa + pa; radius 8 + round(56 wbin / 31) (supports
    17..129 pixels), chirp 0.004 (wbin + 1/2) / nw; records with w in [0, 1)
    fall in w bin floor(nw w)."""
    radii, chirps = [], []
    for wb in range(nw):
        for _ in range(npa):
            radii.append(8 + int(round(56 * wb / max(nw - 1, 1))))
            chirps.append(0.004 * (wb + 0.5) / nw)
    return radii, chirps
