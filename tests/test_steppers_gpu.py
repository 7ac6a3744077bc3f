"""GPU parity of dealias!, the ETDRK4/RK4 pointwise passes, Parseval sums and
the reference's timestepper tests (test/test_timesteppers.jl)."""

import numpy as np
import pytest

from oracle import grids as og
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


@pytest.mark.parametrize("f", [1 / 3, 1 / 2, 1 / 4])
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_dealias_matches_oracle(ff, nd, f):
    ns = (16, 32, 34)[:nd]
    G = (ff.OneDGrid, ff.TwoDGrid, ff.ThreeDGrid)[nd - 1]
    OG = (og.OneDGrid, og.TwoDGrid, og.ThreeDGrid)[nd - 1]
    kw = dict(nx=16, Lx=2 * np.pi, aliased_fraction=f)
    okw = [16, 2 * np.pi]
    if nd >= 2:
        kw.update(ny=32, Ly=2 * np.pi)
        okw += [32, 2 * np.pi]
    if nd >= 3:
        kw.update(nz=34, Lz=2 * np.pi)
        okw += [34, 2 * np.pi]
    g = G(**kw)
    o = OG(*okw, aliased_fraction=f)
    rng = np.random.default_rng(0)
    for xlen in (g.nkr, 16):  # rfft layout uses kralias, full layout kalias
        shape = tuple(reversed((xlen,) + ns[1:]))
        a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        d = _dev(a)
        ff.dealias(d, g)
        ref = og.dealias(a.copy(), o)
        assert np.array_equal(_host(d), ref)


def test_dealias_noop_for_zero_fraction(ff):
    g = ff.TwoDGrid(nx=8, Lx=1.0, aliased_fraction=0)
    d = _dev(np.ones((8, 5), complex))
    assert ff.dealias(d, g) is None
    assert np.all(_host(d) == 1)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("complexL", [False, True])
def test_etdrk4_pointwise_passes(ff, prec, complexL):
    """ETDRK4substep12!/substep3!/update! (src/timesteppers.jl:501-516) vs NumPy
    with the oracle's coefficients."""
    g = ff.TwoDGrid(nx=16, Lx=2 * np.pi, T=np.float64 if prec == 64 else np.float32)
    rng = np.random.default_rng(2)
    shape = (16, 9)
    L = -rng.uniform(0, 50, shape)
    if complexL:
        L = L + 1j * rng.uniform(-20, 20, shape)
    CT = np.complex128 if prec == 64 else np.complex64
    eqn = ff.Equation(L, None, g, T=CT)
    dt = 0.01
    ts = ff.ETDRK4TimeStepper(eqn, dt)
    dtT = float(np.dtype(np.float64 if prec == 64 else np.float32).type(dt))
    E, Eh = ots.getexpLs(dtT, L)
    Z, A, B, G = ots.getetdcoeffs(dtT, L)
    arrs = [(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(CT) for _ in range(5)]
    sol, N1, N2, N3, N4 = [a.astype(np.complex128) for a in arrs]
    d = [_dev(a) for a in arrs]
    out = _dev(np.zeros(shape, CT))
    tol = 1e-15 if prec == 64 else 1e-6
    ts.substep12(out, d[0], d[1])
    assert ou.rel_l2(_host(out), Eh * sol + Z * N1) < tol
    ts.substep3(out, d[0], d[1], d[3])
    assert ou.rel_l2(_host(out), Eh * sol + Z * (2 * N3 - N1)) < tol
    ts.update(d[0], d[1], d[2], d[3], d[4])
    ref = E * sol + A * N1 + 2 * B * (N2 + N3) + G * N4
    assert ou.rel_l2(_host(d[0]), ref) < tol


def test_rk4_pointwise_passes(ff):
    g = ff.TwoDGrid(nx=16, Lx=2 * np.pi)
    rng = np.random.default_rng(3)
    shape = (16, 9)
    L = -rng.uniform(0, 50, shape)
    ts = ff.RK4TimeStepper(ff.Equation(L, None, g))
    a = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(5)]
    d = [_dev(x) for x in a]
    ts.addlinearterm(d[1], d[0])
    assert ou.rel_l2(_host(d[1]), a[1] + L * a[0]) < 1e-15
    out = _dev(np.zeros(shape, complex))
    ts.substepsol(out, d[0], d[2], 0.3)
    assert ou.rel_l2(_host(out), a[0] + 0.3 * a[2]) < 1e-15


@pytest.mark.parametrize("prec", [64, 32])
def test_parseval_sums(ff, prec):
    """test/test_utils.jl:82-96 on 64x128, Lx=2π, Ly=3π: ∫f² = parsevalsum2(f̂)."""
    T = np.float64 if prec == 64 else np.float32
    g = ff.TwoDGrid(nx=64, Lx=2 * np.pi, ny=128, Ly=3 * np.pi, T=T)
    o = og.TwoDGrid(64, 2 * np.pi, 128, 3 * np.pi)
    X, Y = o.x[None, :], o.y[:, None]
    rng = np.random.default_rng(7)
    for f in (np.exp(-(X**2 + Y**2) / (2 * 0.5**2)), rng.standard_normal(o.phys_shape)):
        f = f.astype(T)
        fh = ff.zeros(g.device, np.result_type(T, np.complex64), g.spec_dims)
        ff.mul(fh, g.rfftplan, _dev(f))
        host = _host(fh).astype(np.complex128)
        # the reduction itself against the exactly rounded sum of the same input
        assert abs(ff.parsevalsum2(fh, g) - ou.parsevalsum2(host, o)) <= 1e-15 * ou.parsevalsum2(host, o)
        assert abs(ff.parsevalsum(fh, g) - ou.parsevalsum(host, o)) <= 1e-14 * abs(ou.parsevalsum(host, o)) + 1e-300
        integral = float(np.sum(f.astype(np.float64) ** 2) * o.dx * o.dy)
        tol = 1e-13 if prec == 64 else 1e-6
        assert abs(ff.parsevalsum2(fh, g) - integral) <= tol * integral


def test_parseval_nonfinite_detected(ff):
    g = ff.TwoDGrid(nx=8, Lx=1.0)
    a = np.ones((8, 5), complex)
    a[3, 2] = np.nan
    with pytest.raises(ff.NonFiniteError):
        ff.parsevalsum2(_dev(a), g)


def _gaussian(x, t, c0=0.01, sigma=0.2, kappa=1e-2):
    return c0 * sigma / np.sqrt(sigma**2 + 2 * kappa * t) * np.exp(-(x**2) / (2 * (sigma**2 + 2 * kappa * t)))


@pytest.mark.parametrize("stepper", ["ETDRK4", "RK4"])
@pytest.mark.parametrize("varying", [False, True])
def test_diffusion_timestepper(ff, stepper, varying):
    """test/test_timesteppers.jl:37-59: Gaussian diffusion, nx=128, κ=1e-2,
    dt=1e-7, 1000 steps, rtol = step * 1e-12."""
    kappa = 1e-2
    dt = 1e-9 / kappa
    kap = kappa * np.ones(128) if varying else kappa
    prob = ff.Diffusion.Problem(nx=128, Lx=2 * np.pi, kappa=kap, dt=dt, stepper=stepper)
    ff.Diffusion.set_c(prob, _gaussian(prob.grid.x, 0))
    ff.stepforward(prob, 1000)
    ff.Diffusion.updatevars(prob)
    c_final = _gaussian(prob.grid.x, 1000 * dt)
    assert ou.isapprox(c_final, _host(prob.vars.c), rtol=prob.clock.step * 1e-12)


def test_step_until_rejects_etdrk4(ff):
    prob = ff.Diffusion.Problem(nx=32, kappa=1e-2, dt=1e-7, stepper="ETDRK4")
    with pytest.raises(RuntimeError):
        ff.step_until(prob, 1e-4)


def test_step_until_rk4(ff):
    """test/test_timesteppers.jl:61-70 (constant κ, RK4)."""
    kappa = 1e-2
    dt = 1e-7
    prob = ff.Diffusion.Problem(nx=128, kappa=kappa, dt=dt, stepper="RK4")
    t_final = 1000 * dt + 1e-6 / np.pi
    ff.Diffusion.set_c(prob, _gaussian(prob.grid.x, 0))
    ff.step_until(prob, t_final)
    ff.Diffusion.updatevars(prob)
    c_final = _gaussian(prob.grid.x, t_final)
    assert ou.isapprox(c_final, _host(prob.vars.c), rtol=prob.clock.step * 1e-12)
