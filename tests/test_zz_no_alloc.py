"""The allocation-free reduction entry point vlbm3d_reduce_noalloc (static
device scratch for the block partials, results written to a caller device
array) against vlbm3d_reduce on the same handle, through the C ABI: two
calls in a row, the second on a smaller lattice (its partials overwrite
fewer rows of the shared scratch), both bitwise equal, with and without the
displacement-rate term."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reduce_pair(s, torch, u_prev):
    ref = (C.c_double * 8)()
    prev = None if u_prev is None else int(u_prev.data_ptr())
    assert s._lib.vlbm3d_reduce(s._handle, prev, ref) == 0
    out = torch.full((8,), float("nan"), dtype=torch.float64, device=s.device)
    torch.cuda.synchronize(s.device)
    assert s._lib.vlbm3d_reduce_noalloc(s._handle, prev, out.data_ptr()) == 0
    torch.cuda.synchronize(s.device)
    return np.array(ref[:]), out.cpu().numpy()


def test_gpu_reduce_noalloc_matches_reduce_on_two_lattices():
    torch = pytest.importorskip("torch")
    import gpu_cases
    from curved_lbm.three_d import curved_boundary_warp3d as cb

    results = []
    for n, steps in ((48, 9), (12, 5)):
        g, kw = gpu_cases.traction_shell(n, ramp_steps=4)
        s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
        s.run(steps)
        u_prev = s.u.snapshot()
        s.run(3)
        s.refresh_current()
        for prev in (None, u_prev):
            ref, got = _reduce_pair(s, torch, prev)
            assert np.isfinite(got).all()
            np.testing.assert_array_equal(got, ref)
            results.append((n, prev is not None, got))
        assert results[-1][2][7] > 0.0 and results[-2][2][7] == 0.0
    # the second lattice is smaller: its reduction must not see the first's rows
    assert results[0][2][0] != results[2][2][0]


def test_gpu_reduce_noalloc_rejects_null_output():
    pytest.importorskip("torch")
    import gpu_cases
    from curved_lbm.three_d import curved_boundary_warp3d as cb

    g, kw = gpu_cases.traction_shell(12, ramp_steps=4)
    s = cb.WarpCurvedBoundaryLBM3D(g, **kw)
    assert s._lib.vlbm3d_reduce_noalloc(s._handle, None, None) != 0
    assert b"null" in s._lib.vlbm_last_error()
