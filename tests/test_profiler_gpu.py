"""GPU tests of the library's per-kernel profiler (the figures bench.py's
roofline uses): launch counts per fused ETDRK4 step, kernel-only times that
add up to no more than the wall time of the same steps, and the kept-band
byte counts of the row and column passes."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fused_step_profile(ff):
    import torch
    from fourierflows_hip import _lib

    n, steps = 256, 20
    p = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005, T=np.float32)
    p.sol.copy_(ff.vorticity_ic(p.grid, seed=4))
    ff.stepforward(p, 2)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    _lib.profile_start()
    e0.record()
    ff.stepforward(p, steps)
    e1.record()
    rep = _lib.profile_stop()
    torch.cuda.synchronize()
    wall_ms = e0.elapsed_time(e1)
    k = {r["name"]: r for r in rep["kernels"]}
    # one run: the stage-0 column pass once, then 4 row + 4 column passes per step
    assert k["vort_col_stage0"]["count"] == 1
    assert k["vort_row"]["count"] == 4 * steps
    for s in (1, 2, 3, 4):
        assert k[f"vort_col_stage{s}"]["count"] == steps
    total = sum(r["total_ms"] for r in rep["kernels"])
    assert all(r["total_ms"] > 0 for r in rep["kernels"])
    assert total <= wall_ms * 1.02, (total, wall_ms)  # kernel execution only: no gaps counted
    # kept-band bytes of the row pass: 5 arrays of KX x ny complex values
    lo, hi = p.grid.kralias  # 1-based inclusive aliased kx range
    KX = (n // 2 + 1) - (hi - lo + 1)
    assert k["vort_row"]["bytes"] / k["vort_row"]["count"] == 5 * KX * n * 8
