"""GPU parity of the built-in 2D vorticity operator and its ETDRK4 runs.

The 256² FP64 case is the reference-scale parity case: 1000 ETDRK4 steps of
2D vorticity (ν=1e-3, dt=0.005) from the seeded k⁴exp(-(k/6)²) field, which
must match the FP64 oracle to relative L2 ≤ 1e-10.
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


def _grids(ff, n, T=np.float64, Ly=None, ny=None):
    ny = n if ny is None else ny
    Ly = 2 * np.pi if Ly is None else Ly
    g = ff.TwoDGrid(nx=n, Lx=2 * np.pi, ny=ny, Ly=Ly, x0=0.0, y0=0.0, T=T)
    o = og.TwoDGrid(n, 2 * np.pi, ny, Ly, x0=0, y0=0)
    return g, o


def _oracle_run(o, zh0, L, dt, nsteps):
    eqn = ots.Equation(L, oph.vorticity_calcN, o)
    p = ots.Problem(eqn, "ETDRK4", dt, o)
    p.sol[...] = zh0
    p.stepforward(nsteps)
    return p.sol


@pytest.mark.parametrize("n,ny,Ly", [(64, 64, None), (48, 64, 3 * np.pi), (256, 256, None), (32, 30, 2.0)])
@pytest.mark.parametrize("prec", [64, 32])
def test_calcN_matches_oracle(ff, n, ny, Ly, prec):
    T = np.float64 if prec == 64 else np.float32
    g, o = _grids(ff, n, T, Ly, ny)
    zh = oics.vorticity_ic(o, seed=3, k0=min(6.0, n / 8))
    ref = np.zeros_like(zh)
    oph.vorticity_calcN(ref, zh, 0.0, None, None, None, o)
    op = ff.VorticityOperator(g)
    CT = np.complex128 if prec == 64 else np.complex64
    N = ff.zeros(g.device, CT, g.spec_dims)
    op(N, _dev(zh.astype(CT)), 0.0)
    tol = 1e-12 if prec == 64 else 2e-5
    assert ou.rel_l2(_host(N), ref) < tol


def test_calcN_reads_only_kept_band(ff):
    """Aliased modes of the input are treated as zero (the oracle dealiases a copy)."""
    g, o = _grids(ff, 64)
    rng = np.random.default_rng(0)
    zh = rng.standard_normal(o.spec_shape) + 1j * rng.standard_normal(o.spec_shape)
    ref = np.zeros_like(zh)
    oph.vorticity_calcN(ref, zh, 0.0, None, None, None, o)
    op = ff.VorticityOperator(g)
    N = ff.zeros(g.device, np.complex128, g.spec_dims)
    op(N, _dev(zh), 0.0)
    assert ou.rel_l2(_host(N), ref) < 1e-12
    assert np.all(_host(N)[~og.kept_mask(o)] == 0)


def test_single_shell_is_steady(ff):
    """J(ψ, ζ) = 0 for a single |k| shell: N vanishes and ζ(t) = ζ0 e^{Lt}."""
    g, o = _grids(ff, 64)
    X, Y = np.meshgrid(o.x, o.y)
    z0 = np.cos(3 * X + 4 * Y) + 0.7 * np.sin(5 * X) - 0.3 * np.cos(4 * X - 3 * Y)
    prob = ff.TwoDVorticityProblem(nx=64, nu=1e-2, dt=0.01)
    prob.sol.copy_(_dev(o.mul(z0)))
    ff.stepforward(prob, 50)
    out = o.ldiv(_host(prob.sol))
    assert ou.isapprox(out, z0 * np.exp(-1e-2 * 25 * 0.5), rtol=1e-12)


@pytest.mark.parametrize("n", [64, 48])
def test_etdrk4_short_run_matches_oracle(ff, n):
    g, o = _grids(ff, n)
    L = oph.vorticity_linear(o, nu=1e-3)
    zh0 = oics.vorticity_ic(o, seed=4, k0=5.0)
    ref = _oracle_run(o, zh0, L, 0.005, 50)
    prob = ff.TwoDVorticityProblem(nx=n, nu=1e-3, dt=0.005)
    prob.sol.copy_(_dev(zh0))
    ff.stepforward(prob, 50)
    assert prob.clock.step == 50
    assert ou.rel_l2(_host(prob.sol), ref) < 1e-12


def test_python_calcN_path_matches_fused(ff):
    """A Python calcN! (the operator called through Equation as a plain
    callable) drives the library's pointwise passes; same result as the
    all-in-library run."""
    g, o = _grids(ff, 64)
    zh0 = oics.vorticity_ic(o, seed=5, k0=5.0)
    op = ff.VorticityOperator(g)
    L = ff.vorticity_linear(g, nu=1e-3)
    eqn = ff.Equation(L, lambda N, sol, t, *a: op(N, sol, t), g)
    p1 = ff.Problem(eqn, "ETDRK4", 0.005, g)
    p1.sol.copy_(_dev(zh0))
    ff.stepforward(p1, 20)
    p2 = ff.TwoDVorticityProblem(nx=64, nu=1e-3, dt=0.005)
    p2.sol.copy_(_dev(zh0))
    ff.stepforward(p2, 20)
    assert ou.rel_l2(_host(p1.sol), _host(p2.sol)) < 1e-14


def test_config1_256_fp64_1000_steps(ff):
    """256², FP64, ν=1e-3, dt=0.005, 1000 ETDRK4 steps: rel L2 ≤ 1e-10 vs the oracle."""
    g, o = _grids(ff, 256)
    L = oph.vorticity_linear(o, nu=1e-3)
    zh0 = oics.vorticity_ic(o, seed=1)
    ref = _oracle_run(o, zh0, L, 0.005, 1000)
    prob = ff.TwoDVorticityProblem(nx=256, nu=1e-3, dt=0.005)
    prob.sol.copy_(_dev(zh0))
    ff.stepforward(prob, 1000)
    err = ou.rel_l2(_host(prob.sol), ref)
    print(f"config1 256^2 FP64 1000 steps: rel L2 = {err:.3e}")
    assert err <= 1e-10
    e, z = prob.eqn.calcN.energy_enstrophy(prob.sol)
    assert abs(e - ou.energy(ref, o)) <= 1e-12 * ou.energy(ref, o)
    assert abs(z - ou.enstrophy(ref, o)) <= 1e-12 * ou.enstrophy(ref, o)


def test_fp32_run_tracks_fp64(ff):
    """FP32 bulk with FP64 coefficients stays within 1e-4 of the FP64 oracle."""
    g, o = _grids(ff, 128)
    L = oph.vorticity_linear(o, nu=1e-3)
    zh0 = oics.vorticity_ic(o, seed=2)
    ref = _oracle_run(o, zh0, L, 0.005, 200)
    prob = ff.TwoDVorticityProblem(nx=128, nu=1e-3, dt=0.005, T=np.float32)
    prob.sol.copy_(_dev(zh0.astype(np.complex64)))
    ff.stepforward(prob, 200)
    err = ou.rel_l2(_host(prob.sol).astype(np.complex128), ref)
    print(f"128^2 FP32 200 steps: rel L2 vs FP64 oracle = {err:.3e}")
    assert err < 1e-4


def test_product_ic_matches_oracle_ic(ff):
    g, o = _grids(ff, 128)
    zh = ff.vorticity_ic(g, seed=1)
    ref = oics.vorticity_ic(o, seed=1)
    assert ou.rel_l2(_host(zh), ref) < 1e-13
