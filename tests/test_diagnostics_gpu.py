"""GPU tests of diagnostics (src/diagnostics.jl), following the reference's
diagnostics testset (test/test_diagnostics.jl, test/runtests.jl:322-333) on
the 1D diffusion testbed, plus the built-in vorticity energy / enstrophy
incremented through fused chunks and checked against the oracle."""

import numpy as np
import pytest

from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph
from oracle import timesteppers as ots
from oracle import utils as ou

pytestmark = pytest.mark.gpu


def _diffusion(ff, stepper="RK4", fn=np.sin, nx=6, Lx=2 * np.pi, kappa=1e-2):
    k0 = 2 * np.pi / Lx
    dt = 1e-2 / (kappa * k0**2)
    D = ff.Diffusion
    prob = D.Problem(nx=nx, Lx=Lx, kappa=kappa, dt=dt, stepper=stepper)
    x = np.asarray(prob.grid.x, dtype=np.float64)
    D.set_c(prob, fn(k0 * x))
    return prob, k0, kappa


@pytest.mark.parametrize("nsteps,freq,ndata", [(100, 1, None), (100, 2, None), (100, 9, 20)])
def test_diagnosticsteps(ff, nsteps, freq, ndata):
    """test/test_diagnostics.jl:1-9."""
    prob = ff.Diffusion.Problem(nx=6, Lx=2 * np.pi)
    d = ff.Diagnostic(lambda p: 1, prob, nsteps=nsteps, freq=freq, ndata=ndata)
    ff.stepforward(prob, d, nsteps)
    expected = [0] + list(range(freq, nsteps + 1, freq))
    assert list(d.steps[: d.i]) == expected
    assert list(d["steps"]) == expected


@pytest.mark.parametrize("freq", [1, 2])
def test_scalardiagnostics(ff, freq):
    """test/test_diagnostics.jl:11-31: variance of a decaying sine."""
    prob, k0, kappa = _diffusion(ff)
    Lx = prob.grid.Lx
    var = lambda p: ff.parsevalsum2(p.sol, p.grid) / Lx
    d = ff.Diagnostic(var, prob, nsteps=100, freq=freq)
    ff.stepforward(prob, d, 100)
    expected = 0.5 * np.exp(-kappa * k0**2 * d["t"]) ** 2
    got = np.asarray(d["data"], dtype=np.float64)
    assert np.linalg.norm(got - expected) <= 1.5e-8 * np.linalg.norm(expected)


def test_basicdiagnostics(ff):
    """test/test_diagnostics.jl:33-46."""
    import torch

    prob, _, _ = _diffusion(ff, fn=np.cos)
    d = ff.Diagnostic(lambda p: p.sol.clone(), prob, nsteps=1)
    ff.stepforward(prob, d, 1)
    assert torch.equal(d.data[d.i - 1], prob.sol)
    assert d.i == 2 and d.steps[1] == 1


def test_extenddiagnostic(ff):
    """test/test_diagnostics.jl:48-68."""
    prob, _, _ = _diffusion(ff)
    d1 = ff.Diagnostic(lambda p: p.sol, prob, nsteps=1)
    d2 = ff.Diagnostic(lambda p: p.sol, prob, nsteps=4)
    n1, n2 = len(d1.t), len(d2.t)
    ff.diagnostics.extend(d1, 12)
    ff.diagnostics.extend(d2)
    assert len(d1.t) == n1 + 12 and len(d2.t) == 2 * n2
    assert len(d1.data) == len(d1.steps) == len(d1.t)


def test_incrementdiagnostic_and_getindex(ff):
    """test/test_diagnostics.jl:70-121 (increment! over a list; getindex)."""
    prob, _, _ = _diffusion(ff, stepper="ETDRK4")
    d1 = ff.Diagnostic(lambda p: 2.0, prob, nsteps=10)
    d2 = ff.Diagnostic(lambda p: 2.0j, prob, nsteps=10)
    ff.increment([d1, d2])
    assert d1.data[1] == 2.0 and d2.data[1] == 2.0j
    for _ in range(6):
        ff.increment([d1, d2])
    assert d1[1:5] == [2.0] * 4 and d2[1:5] == [2.0j] * 4
    assert d1.i == 8


def test_diagnostic_grows_past_ndata(ff):
    """update! extends the storage by the initial ndata when it is full."""
    prob = ff.Diffusion.Problem(nx=6, Lx=2 * np.pi)
    d = ff.Diagnostic(lambda p: p.clock.step, prob, nsteps=3, freq=1)
    assert len(d.t) == 4
    ff.stepforward(prob, d, 10)
    assert d.i == 11 and len(d.t) == 12
    assert d["data"] == list(range(11))


@pytest.mark.parametrize("prec", [64, 32])
def test_vorticity_diagnostics_chunked(ff, prec):
    """Energy and enstrophy every 3 and 5 steps through fused chunks equal
    (i) the same run stepped one step at a time with per-step increments
    and (ii) the oracle's energy/enstrophy of its own run at those steps."""
    T = np.float64 if prec == 64 else np.float32
    CT = np.complex128 if prec == 64 else np.complex64
    n, dt, nsteps = 64, 0.005, 31
    o = og.TwoDGrid(n, 2 * np.pi, n, 2 * np.pi, x0=0, y0=0)
    zh0 = oics.vorticity_ic(o, seed=5, k0=6.0)
    L = -1e-3 * o.Krsq

    def make():
        prob = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=dt, T=T)
        import torch

        prob.sol.copy_(torch.as_tensor(zh0.astype(CT)).cuda())
        E = ff.Diagnostic(ff.diagnostics.energy, prob, nsteps=nsteps, freq=3)
        Z = ff.Diagnostic(ff.diagnostics.enstrophy, prob, nsteps=nsteps, freq=5)
        return prob, E, Z

    pa, Ea, Za = make()
    ff.stepforward(pa, [Ea, Za], nsteps)
    pb, Eb, Zb = make()
    for _ in range(nsteps):
        ff.stepforward(pb, 1)
        ff.increment([Eb, Zb])
    assert list(Ea["steps"]) == [0] + list(range(3, nsteps + 1, 3))
    assert list(Za["steps"]) == [0] + list(range(5, nsteps + 1, 5))
    assert list(Ea["steps"]) == list(Eb["steps"]) and list(Za["steps"]) == list(Zb["steps"])
    np.testing.assert_allclose(Ea["t"], Eb["t"], rtol=0, atol=1e-14)
    tol_same = 1e-14 if prec == 64 else 1e-6
    np.testing.assert_allclose(Ea["data"], Eb["data"], rtol=tol_same)
    np.testing.assert_allclose(Za["data"], Zb["data"], rtol=tol_same)
    assert pa.clock.step == nsteps

    eqn = ots.Equation(L, oph.vorticity_calcN, o)
    p = ots.Problem(eqn, "ETDRK4", dt, o)
    p.sol[...] = zh0
    refE, refZ = {0: ou.energy(zh0, o)}, {0: ou.enstrophy(zh0, o)}
    for s in range(1, nsteps + 1):
        p.stepforward(1)
        if s % 3 == 0:
            refE[s] = ou.energy(p.sol, o)
        if s % 5 == 0:
            refZ[s] = ou.enstrophy(p.sol, o)
    tol = 1e-12 if prec == 64 else 2e-5
    np.testing.assert_allclose(Ea["data"], [refE[s] for s in Ea["steps"]], rtol=tol)
    np.testing.assert_allclose(Za["data"], [refZ[s] for s in Za["steps"]], rtol=tol)


def test_tracer_diagnostics_chunked(ff):
    """Tracer variance every 4 steps through fused RK4 chunks equals per-step stepping."""
    import torch

    def make():
        prob = ff.TracerProblem(n=32, kappa=1e-2, dt=0.01, T=np.float64)
        ff.mul(prob.sol, prob.grid.rfftplan, ff.tracer_ic(prob.grid, seed=2, ngauss=4, sigma=0.5))
        V = ff.Diagnostic(ff.diagnostics.tracer_variance, prob, nsteps=12, freq=4)
        return prob, V

    pa, Va = make()
    ff.stepforward(pa, Va, 12)
    pb, Vb = make()
    for _ in range(12):
        ff.stepforward(pb, 1)
        ff.increment(Vb)
    assert list(Va["steps"]) == [0, 4, 8, 12]
    np.testing.assert_allclose(Va["data"], Vb["data"], rtol=1e-14)
    assert torch.equal(pa.sol, pb.sol)
