"""GPU parity of the fused 2D vorticity ETDRK4 path.

The fused path (8 kernels per step, kept band only, blocked layouts) must
agree with the FP64 oracle and with the library's unfused path (the
reference's dataflow) on the same seeded inputs.
"""

import numpy as np
import pytest

from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph
from oracle import timesteppers as ots
from oracle import utils as ou

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x)).cuda()


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def _problem(ff, monkeypatch, fused, **kw):
    if not fused:
        monkeypatch.setenv("FFHIP_DISABLE_FUSED", "1")
    p = ff.TwoDVorticityProblem(**kw)
    monkeypatch.delenv("FFHIP_DISABLE_FUSED", raising=False)
    return p


def _oracle(n, ny, Ly, zh0, L, dt, nsteps):
    o = og.TwoDGrid(n, 2 * np.pi, ny, Ly, x0=0, y0=0)
    eqn = ots.Equation(L, oph.vorticity_calcN, o)
    p = ots.Problem(eqn, "ETDRK4", dt, o)
    p.sol[...] = zh0
    p.stepforward(nsteps)
    return p.sol


@pytest.mark.parametrize("n,ny,Ly", [(64, 64, 2 * np.pi), (128, 128, 2 * np.pi), (256, 128, 4 * np.pi),
                                     (128, 512, 2 * np.pi), (512, 512, 2 * np.pi), (1024, 1024, 2 * np.pi),
                                     (2048, 1024, np.pi), (4096, 64, np.pi / 16), (64, 4096, 32 * np.pi)])
def test_fused_fp64_matches_oracle(ff, n, ny, Ly):
    o = og.TwoDGrid(n, 2 * np.pi, ny, Ly, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=11, k0=min(6.0, n / 10))
    L = oph.vorticity_linear(o, nu=2e-3)
    steps = 30 if n * ny <= 512 * 512 else 6
    ref = _oracle(n, ny, Ly, zh0, L, 0.004, steps)
    prob = ff.TwoDVorticityProblem(nx=n, ny=ny, Ly=Ly, nu=2e-3, dt=0.004)
    assert prob.eqn.calcN.fused
    prob.sol.copy_(_dev(zh0))
    ff.stepforward(prob, steps)
    err = ou.rel_l2(_host(prob.sol), ref)
    assert err < 1e-12, err


@pytest.mark.parametrize("n,ny,Ly", [(64, 64, 2 * np.pi), (256, 128, 4 * np.pi), (128, 512, 2 * np.pi),
                                     (512, 512, 2 * np.pi), (1024, 1024, 2 * np.pi), (2048, 1024, np.pi),
                                     (1024, 2048, 2 * np.pi), (4096, 64, np.pi / 16), (64, 4096, 32 * np.pi)])
def test_fused_fp32_matches_oracle(ff, n, ny, Ly):
    """The FP32 fused passes (paired transforms up to 1024 in the column pass,
    paired inverses in every row pass) against the FP64 oracle: relative L2
    at FP32 roundoff after the steps."""
    o = og.TwoDGrid(n, 2 * np.pi, ny, Ly, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=12, k0=min(6.0, n / 10))
    L = oph.vorticity_linear(o, nu=2e-3)
    steps = 30 if n * ny <= 512 * 512 else 6
    ref = _oracle(n, ny, Ly, zh0, L, 0.004, steps)
    prob = ff.TwoDVorticityProblem(nx=n, ny=ny, Ly=Ly, nu=2e-3, dt=0.004, T=np.float32)
    assert prob.eqn.calcN.fused
    prob.sol.copy_(_dev(zh0.astype(np.complex64)))
    ff.stepforward(prob, steps)
    err = ou.rel_l2(_host(prob.sol).astype(np.complex128), ref)
    assert err < 1e-5, err


_THREE_TEAM_REF = {}


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n,ny,Ly", [(1024, 1024, 2 * np.pi), (2048, 1024, np.pi), (1024, 2048, 2 * np.pi)])
def test_fused_three_team_passes(ff, monkeypatch, prec, n, ny, Ly):
    """The three-team row and column passes (FFHIP_COL3 / FFHIP_ROW3 = 2:
    both precisions, 1024- and 2048-point axes), off by default, against
    the FP64 oracle."""
    T = np.float64 if prec == 64 else np.float32
    CT = np.complex128 if prec == 64 else np.complex64
    key = (n, ny, Ly)
    if key not in _THREE_TEAM_REF:
        o = og.TwoDGrid(n, 2 * np.pi, ny, Ly, x0=0, y0=0)
        zh0 = oics.vorticity_ic(o, seed=13, k0=6.0)
        L = oph.vorticity_linear(o, nu=2e-3)
        _THREE_TEAM_REF[key] = (zh0, _oracle(n, ny, Ly, zh0, L, 0.004, 12))
    zh0, ref = _THREE_TEAM_REF[key]
    monkeypatch.setenv("FFHIP_COL3", "2")
    monkeypatch.setenv("FFHIP_ROW3", "2")
    prob = ff.TwoDVorticityProblem(nx=n, ny=ny, Ly=Ly, nu=2e-3, dt=0.004, T=T)
    assert prob.eqn.calcN.fused
    prob.sol.copy_(_dev(zh0.astype(CT)))
    ff.stepforward(prob, 12)  # an eager step, one 10-step graph replay, an eager last step
    err = ou.rel_l2(_host(prob.sol).astype(np.complex128), ref)
    monkeypatch.delenv("FFHIP_COL3")
    monkeypatch.delenv("FFHIP_ROW3")
    assert err < (1e-12 if prec == 64 else 1e-5), err


def test_fused_4096_square(ff, monkeypatch):
    """4096² (fused up to 4096 points per axis in both precisions): the FP64
    fused run against the FP64 plan-based run (the reference's dataflow), and
    the FP32 fused run against the FP64 one, over a few steps."""
    import torch

    n, dt = 4096, 1e-3
    nu8 = oph.hyperviscosity_for(n, dt)
    kw = dict(nx=n, nu_n=nu8, n_hyper=4, mu=0.01, dt=dt)
    p32 = ff.TwoDVorticityProblem(T=np.float32, **kw)
    p64 = ff.TwoDVorticityProblem(T=np.float64, **kw)
    pg = _problem(ff, monkeypatch, False, T=np.float64, **kw)
    assert p32.eqn.calcN.fused and p64.eqn.calcN.fused and not pg.eqn.calcN.fused
    z0 = ff.vorticity_ic(p64.grid, seed=3)
    p64.sol.copy_(z0)
    pg.sol.copy_(z0)
    p32.sol.copy_(z0.to(torch.complex64))
    for p in (p64, pg, p32):
        ff.stepforward(p, 4)
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))  # noqa: E731
    assert rel(p64.sol, pg.sol) < 1e-12
    assert rel(p32.sol.to(torch.complex128), p64.sol) < 1e-5


@pytest.mark.parametrize("prec", [64, 32])
def test_fused_matches_unfused(ff, monkeypatch, prec):
    T = np.float64 if prec == 64 else np.float32
    CT = np.complex128 if prec == 64 else np.complex64
    n = 256
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=5).astype(CT)
    kw = dict(nx=n, nu=1e-3, dt=0.005, T=T)
    pf = _problem(ff, monkeypatch, True, **kw)
    pg = _problem(ff, monkeypatch, False, **kw)
    assert pf.eqn.calcN.fused and not pg.eqn.calcN.fused
    for p in (pf, pg):
        p.sol.copy_(_dev(zh0))
        ff.stepforward(p, 40)
    err = ou.rel_l2(_host(pf.sol).astype(np.complex128), _host(pg.sol).astype(np.complex128))
    assert err < (1e-13 if prec == 64 else 2e-5), err


def test_fused_hyperviscous_2048_short(ff):
    """2048², ν₈|k|⁸ + μ, dt = 1e-3 (dt·L reaches ≈ -410): a few steps vs the oracle."""
    n, dt = 2048, 1e-3
    nu8 = oph.hyperviscosity_for(n, dt)
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=1)
    L = oph.vorticity_linear(o, nu_n=nu8, n_hyper=4, mu=0.01)
    ref = _oracle(n, n, 2 * np.pi, zh0, L, dt, 3)
    prob = ff.TwoDVorticityProblem(nx=n, nu_n=nu8, n_hyper=4, mu=0.01, dt=dt)
    prob.sol.copy_(_dev(zh0))
    ff.stepforward(prob, 3)
    assert ou.rel_l2(_host(prob.sol), ref) < 1e-12
    p32 = ff.TwoDVorticityProblem(nx=n, nu_n=nu8, n_hyper=4, mu=0.01, dt=dt, T=np.float32)
    p32.sol.copy_(_dev(zh0.astype(np.complex64)))
    ff.stepforward(p32, 3)
    assert ou.rel_l2(_host(p32.sol).astype(np.complex128), ref) < 1e-5


def test_fused_aliased_modes_evolve_linearly(ff, monkeypatch):
    """A state with content outside the kept band: the fused run matches the
    unfused stepper bit for bit there (exp(dt L) per step) and to roundoff inside."""
    n = 128
    rng = np.random.default_rng(3)
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=2) + 1e-3 * (rng.standard_normal(o.spec_shape)
                                                 + 1j * rng.standard_normal(o.spec_shape))
    kw = dict(nx=n, nu=1e-3, dt=0.005)
    pf = _problem(ff, monkeypatch, True, **kw)
    pg = _problem(ff, monkeypatch, False, **kw)
    for p in (pf, pg):
        p.sol.copy_(_dev(zh0))
        ff.stepforward(p, 25)
    a, b = _host(pf.sol), _host(pg.sol)
    kept = og.kept_mask(o)
    assert np.array_equal(a[~kept], b[~kept])
    assert ou.rel_l2(a[kept], b[kept]) < 1e-13


def test_fused_repeated_calls_equal_one_call(ff):
    """stepforward!(prob, 10) twice == stepforward!(prob, 20) (run boundaries are exact)."""
    n = 128
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=8)
    p1 = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005)
    p2 = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005)
    p1.sol.copy_(_dev(zh0))
    p2.sol.copy_(_dev(zh0))
    ff.stepforward(p1, 10)
    ff.stepforward(p1, 10)
    ff.stepforward(p2, 20)
    assert np.array_equal(_host(p1.sol), _host(p2.sol))
    assert p1.clock.step == p2.clock.step == 20


@pytest.mark.parametrize("T,stepper", [(np.float32, "ETDRK4"), (np.float64, "ETDRK4"),
                                       (np.float32, "FilteredETDRK4")])
def test_fused_graph_replays_equal_eager_steps(ff, T, stepper):
    """Runs of >= 12 steps replay a captured 10-step graph; one-step calls are
    all eager.  Two 35-step runs (three replays each, with the library's eager
    first and last steps and the import/export kernels between them) must
    equal 70 one-step calls bit for bit."""
    n = 256
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=12).astype(np.complex64 if T == np.float32 else np.complex128)
    p1 = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005, T=T, stepper=stepper)
    p2 = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005, T=T, stepper=stepper)
    assert p1.eqn.calcN.fused
    p1.sol.copy_(_dev(zh0))
    p2.sol.copy_(_dev(zh0))
    for _ in range(2):
        ff.stepforward(p1, 35)
        for _ in range(35):
            ff.stepforward(p2, 1)
        assert np.array_equal(_host(p1.sol), _host(p2.sol))
    assert p1.clock.step == p2.clock.step == 70


def test_fused_deterministic(ff):
    n = 256
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=9).astype(np.complex64)
    outs = []
    for _ in range(2):
        p = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005, T=np.float32)
        p.sol.copy_(_dev(zh0))
        ff.stepforward(p, 20)
        outs.append(_host(p.sol))
    assert np.array_equal(outs[0], outs[1])


def test_fused_2048_full_run_fp32_vs_fp64(ff):
    """The 2048² hyperviscous workload over its full 2000 steps: FP32 stays
    within relative L2 1e-4 of FP64 (field, energy, enstrophy), and both
    runs dissipate -- energy and enstrophy, saved every 200 steps through
    the diagnostics, never grow (the dealiased Jacobian conserves both, so
    only ν₈|k|⁸ + μ act on them)."""
    import torch

    n, dt, nsteps = 2048, 1e-3, 2000
    nu8 = oph.hyperviscosity_for(n, dt)
    o = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=1)
    runs = {}
    for T, CT in ((np.float64, np.complex128), (np.float32, np.complex64)):
        p = ff.TwoDVorticityProblem(nx=n, nu_n=nu8, n_hyper=4, mu=0.01, dt=dt, T=T)
        p.sol.copy_(_dev(zh0.astype(CT)))
        E = ff.Diagnostic(ff.diagnostics.energy, p, freq=200, nsteps=nsteps)
        Z = ff.Diagnostic(ff.diagnostics.enstrophy, p, freq=200, nsteps=nsteps)
        ff.stepforward(p, [E, Z], nsteps)
        runs[T] = (p, np.asarray(E["data"]), np.asarray(Z["data"]))
        assert list(E["steps"]) == list(range(0, nsteps + 1, 200))
        assert np.all(np.diff(runs[T][1]) < 0) and np.all(np.diff(runs[T][2]) < 0)
    p64, e64, z64 = runs[np.float64]
    p32, e32, z32 = runs[np.float32]
    err = float(torch.linalg.vector_norm(p32.sol.to(torch.complex128) - p64.sol) / torch.linalg.vector_norm(p64.sol))
    assert err < 1e-4, err
    assert np.max(np.abs(e32 - e64) / e64) < 1e-4
    assert np.max(np.abs(z32 - z64) / z64) < 1e-4
    # the initial energy and enstrophy are the oracle's
    assert abs(e64[0] - ou.energy(zh0, o)) <= 1e-12 * ou.energy(zh0, o)
    assert abs(z64[0] - ou.enstrophy(zh0, o)) <= 1e-12 * ou.enstrophy(zh0, o)
