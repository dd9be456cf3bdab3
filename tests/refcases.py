"""Test cases restated from the reference's own unit tests
(tests/GridderStateUnitTests.cc), independent of the oracle and of the HIP
device: inputs, and closed-form expectations where the reference test has
one."""
import numpy as np

C_LIGHT = 299792458.0
PADDING = 2

# GridOne (GridderStateUnitTests.cc:1343-1463)
GRIDONE = dict(
    grid_size=(16384, 16384, 1, 1), grid_scale=(0.0476591, 0.0476591),
    radius=45, oversampling=20, freq=3.693e9, vis=1 - 1j, weight=1.0,
    uvw=(2344.1, 638.066, -1826.55), ll=(9523, 8522), ur=(9613, 8612))

# Gridding (GridderStateUnitTests.cc:1273-1341)
GRIDDING = dict(grid_size=(14, 14, 1, 1), grid_scale=(0.1, 0.1), radius=3,
                oversampling=10, freq=C_LIGHT, vis=1 - 1j, weight=1.0)


def one_record(dtype, vis, weight, freq, uvw, grid_cube=0, cf_index=(0, 0),
               gradient=(0.0, 0.0), phase=0.0):
    rec = np.zeros(1, dtype=dtype)
    rec[rec.dtype.names[0]][0, :] = vis
    rec[rec.dtype.names[1]][0, :] = weight
    rec["frequency"] = freq
    rec["phase"] = phase
    rec["uvw"][0] = uvw
    rec["grid_cube"] = grid_cube
    rec["cf_index"][0] = cf_index
    rec["cf_phase_gradient"][0] = gradient
    return rec


def coords(grid_size, grid_scale, freq, uvw, radius, oversampling):
    """ConeCFArray::coords of the reference tests (GridderStateUnitTests.cc:
    257-330): (support corner, cf_major, cf_minor) per axis, with lrint
    round-half-even."""
    out = []
    for d in range(2):
        position = grid_scale[d] * uvw[d] * (freq / C_LIGHT) + grid_size[d] / 2.0
        g = int(np.rint(position))
        fine = int(np.rint((g - position) * oversampling))
        if fine >= 0:
            major, minor = PADDING, fine
        else:
            major, minor = PADDING - 1, oversampling + fine
        out.append((g - radius, major, minor))
    return out


def single_vis_expectation(cf_plane, grid_size, grid_scale, freq, uvw, radius,
                           oversampling, vis):
    """Grid expected from gridding one visibility with value `vis`, weight 1,
    phase 0, w <= 0 and no phase gradient, through a one-Mueller CF plane
    cf_plane[y][x]: each footprint pixel holds the float product
    fl32(cf * vis) (verify_gridded_vis, GridderStateUnitTests.cc:332-368)."""
    (gx, mx, nx_), (gy, my, ny_) = coords(grid_size, grid_scale, freq, uvw, radius,
                                          oversampling)
    s = 2 * radius + 1
    out = np.zeros((grid_size[1], grid_size[0]), dtype=np.complex128)
    v = np.complex64(vis)
    vr, vi = np.float32(v.real), np.float32(v.imag)
    for X in range(s):
        for Y in range(s):
            xf = (X + mx) * oversampling + nx_
            yf = (Y + my) * oversampling + ny_
            c = cf_plane[yf, xf]
            cr, ci = np.float32(c.real), np.float32(c.imag)
            re = np.float32(cr * vr - ci * vi)
            im = np.float32(cr * vi + ci * vr)
            out[gy + Y, gx + X] = complex(float(re), float(im))
    return out
