"""Tasks (max_added_tasks + 1 record buffers and copy streams in rotation,
src/hpg_runtime.hpp:1496-1526, 1643-1681): host batches gridded back to
back, with the caller refilling its one host buffer for the next batch as
soon as a call returns, give the grid and weights of the same batches on a
state without added tasks (FP64 grids: same products and sums in the same
order per batch, 1e-12 of max|grid|; weights 1e-12); degridding with
returned residuals too.  The task count is clamped to 16 and reported."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_added_tasks_clamped(hpg):
    from hpg_amd import workloads
    cf = workloads.gaussian_cf(3, 20, 1.5)
    for req, got in ((0, 0), (3, 3), (100, 15)):
        gs = hpg.GridderState.create(hpg.Device.Hip, req, 10, cf, (64, 64, 1, 1), (1.0, 1.0),
                                     [[0]])
        assert gs.max_added_tasks() == got
        gs.release()


@pytest.mark.parametrize("tasks", [1, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_pipelined_host_batches(hpg, tasks, mode):
    from hpg_amd import workloads
    npol, gsz = 2, (256, 192, 2, 2)
    M, Mc = [[0, 1], [2, 3]], [[1, 0], [3, 2]]
    cf = workloads.random_cf(8, [3, 6], nmueller=4, ncube=2, seed=91)
    batches = [workloads.random_visibilities(20000, npol, gsz, (1.0, 1.0), cf, seed=92 + b,
                                             phase=True, margin=-2) for b in range(7)]
    ref = hpg.GridderState.create(hpg.Device.Hip, 0, 20000, cf, gsz, (1.0, 1.0), M, Mc,
                                  grid_mode=mode)
    ref.set_convolution_function(hpg.Device.Serial, cf)
    pip = hpg.GridderState.create(hpg.Device.Hip, tasks, 20000, cf, gsz, (1.0, 1.0), M, Mc,
                                  grid_mode=mode)
    pip.set_convolution_function(hpg.Device.Serial, cf)
    assert pip.max_added_tasks() == tasks
    buf = np.empty_like(batches[0])
    for b in batches:
        ref.grid_visibilities(hpg.Device.Serial, b)
        buf[...] = b  # the caller's one buffer, refilled as soon as a call returns
        pip.grid_visibilities(hpg.Device.Serial, buf)
        buf[...] = 0
    a, b = ref.grid_values(), pip.grid_values()
    tol = 1e-12 if mode == 0 else 1e-6
    assert np.abs(b - a).max() <= tol * np.abs(a).max()
    np.testing.assert_allclose(pip.grid_weights(), ref.grid_weights(), rtol=1e-12)
    # degridding with residuals returned, through the rotation
    rng = np.random.default_rng(99)
    shape = (gsz[3], gsz[2], gsz[1], gsz[0])
    model = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for gs in (ref, pip):
        gs.set_model(hpg.Device.Serial, model)
        gs.reset_grid()
    for bt in batches[:3]:
        r0 = ref.degrid_grid_get_residual_visibilities(hpg.Device.Serial, bt)
        buf[...] = bt
        r1 = pip.degrid_grid_get_residual_visibilities(hpg.Device.Serial, buf)
        buf[...] = 0
        np.testing.assert_allclose(r1["visibilities"], r0["visibilities"], rtol=1e-6,
                                   atol=1e-6 * np.abs(r0["visibilities"]).max())
    a, b = ref.grid_values(), pip.grid_values()
    assert np.abs(b - a).max() <= tol * np.abs(a).max()
