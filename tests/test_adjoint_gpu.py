"""GPU discrete adjoint: plan transposes, the fused ETDRK4 adjoint, autograd.

* mul!^T / ldiv!^T satisfy the dot-product identities and equal the
  oracle's transposes;
* the FP64 fused adjoint gradient equals the FP64 oracle adjoint gradient
  (same misfit, same checkpoints) to roundoff;
* the 512^2 FP32 gradient (500 steps, 23 checkpoints) passes a central-FD
  Taylor test in 8 seeded directions and agrees with the FP64 gradient;
* the autograd Functions give the same gradient as the explicit calls.
"""

import numpy as np
import pytest

from oracle import adjoint as oa
from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch

    return torch.as_tensor(np.ascontiguousarray(x)).cuda()


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _rand_spec(shape, rng):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("kind,ns", [("rfft", (48, 40)), ("rfft", (64,)), ("rfft", (12, 10, 8)), ("fft", (24, 18))])
def test_plan_transposes(ff, kind, ns):
    import torch

    rng = np.random.default_rng(3)
    plan = ff.Plan(ns, np.float64, kind)
    nkr = ns[0] // 2 + 1 if kind == "rfft" else ns[0]
    sshape = tuple(reversed((nkr,) + ns[1:]))
    pshape = tuple(reversed(ns))
    X = _rand_spec(sshape, rng)  # not Hermitian: Im at kx = 0, nx/2 is ignored by the inverse
    phi = rng.standard_normal(pshape) + (1j * rng.standard_normal(pshape) if kind == "fft" else 0)
    Xd, phid = _dev(X), _dev(phi)
    inv = plan.inverse(torch.zeros_like(phid), Xd)
    iT = plan.inverse_adjoint(torch.zeros_like(Xd), phid)
    fw = plan.forward(torch.zeros_like(Xd), phid)
    fT = plan.forward_adjoint(torch.zeros_like(phid), Xd)
    re = lambda a, b: float(np.sum(a.real * b.real + a.imag * b.imag))  # noqa: E731
    lhs, rhs = re(_host(inv), phi), re(X, _host(iT))
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))
    lhs2, rhs2 = re(_host(fw), X), re(phi, _host(fT))
    assert abs(lhs2 - rhs2) < 1e-12 * max(1.0, abs(lhs2))
    if kind == "rfft":
        g = og.TwoDGrid(ns[0], 2 * np.pi, ns[1], 3.0) if len(ns) == 2 else (
            og.OneDGrid(ns[0], 2 * np.pi) if len(ns) == 1 else og.ThreeDGrid(ns[0], 2 * np.pi, ns[1], 3.0, ns[2], 1.0))
        np.testing.assert_allclose(_host(iT), oa.ldiv_T(phi, g), rtol=0, atol=1e-14 * np.abs(phi).max())
        np.testing.assert_allclose(_host(fT), oa.mul_T(X, g), rtol=0, atol=1e-12 * np.abs(X).max() * np.prod(ns))


def _setup(ff, n, nu, dt, nsteps, T, seed0=1, seed_obs=2):
    """(prob, zeta0, zeta_obs) with zeta_obs a forward run from a second seed."""
    import torch

    prob = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, T=T)
    assert prob.eqn.calcN.fused
    g = prob.grid
    z0 = torch.zeros(g.ny, g.nx, dtype=prob.sol.real.dtype, device="cuda")
    ff.ldiv(z0, g.rfftplan, ff.vorticity_ic(g, seed=seed0))
    prob.sol.copy_(ff.vorticity_ic(g, seed=seed_obs))
    ff.stepforward(prob, nsteps)
    zobs = torch.zeros_like(z0)
    ff.ldiv(zobs, g.rfftplan, prob.sol)
    return prob, z0, zobs


@pytest.mark.parametrize("n,nsteps,ncheck", [(64, 30, 4), (128, 17, 17), (64, 12, 1)])
def test_fp64_gradient_matches_oracle(ff, n, nsteps, ncheck):
    nu, dt = 1e-2, 0.01
    prob, z0, zobs = _setup(ff, n, nu, dt, nsteps, np.float64)
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, ncheck)
    g = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    adj = oa.ETDRK4Adjoint(g, oph.vorticity_linear(g, nu=nu), dt)
    Jr, gr = adj.gradient(_host(z0), _host(zobs), nsteps, ncheck)
    assert abs(J - Jr) < 1e-12 * Jr, (J, Jr)
    err = np.linalg.norm(_host(grad) - gr) / np.linalg.norm(gr)
    assert err < 1e-11, err


def test_fp64_gradient_512_matches_oracle(ff):
    """The 512^2 grid of the FP32 adjoint workload, in FP64, over 40 steps."""
    n, nu, dt, nsteps = 512, 5e-4, 0.002, 40
    prob, z0, zobs = _setup(ff, n, nu, dt, nsteps, np.float64)
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, 5)
    g = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    adj = oa.ETDRK4Adjoint(g, oph.vorticity_linear(g, nu=nu), dt)
    Jr, gr = adj.gradient(_host(z0), _host(zobs), nsteps, 5)
    assert abs(J - Jr) < 1e-12 * Jr, (J, Jr)
    err = np.linalg.norm(_host(grad) - gr) / np.linalg.norm(gr)
    assert err < 1e-11, err


@pytest.mark.parametrize("T,nx,tol", [(np.float64, 2048, 1e-11), (np.float32, 2048, 1e-4), (np.float32, 1024, 1e-4)])
def test_gradient_halflength_rows_matches_oracle(ff, T, nx, tol):
    """Rows of 1024 (FP32) and 2048 points run the half-length row kernels
    (one wavefront per FP32 row, two per FP64 row), whose stored stage fields
    (u + iv and zeta, normalised) feed the adjoint sweep: the gradient of an
    nx x 64 grid against the FP64 oracle adjoint."""
    import torch

    ny, nu, dt, nsteps, ncheck = 64, 1e-2, 1e-3, 4, 2
    prob = ff.TwoDVorticityProblem(nx=nx, ny=ny, nu=nu, dt=dt, T=T)
    assert prob.eqn.calcN.fused
    g = prob.grid
    z0 = torch.zeros(g.ny, g.nx, dtype=prob.sol.real.dtype, device="cuda")
    ff.ldiv(z0, g.rfftplan, ff.vorticity_ic(g, seed=1))
    prob.sol.copy_(ff.vorticity_ic(g, seed=2))
    ff.stepforward(prob, nsteps)
    zobs = torch.zeros_like(z0)
    ff.ldiv(zobs, g.rfftplan, prob.sol)
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, ncheck)
    og_ = og.TwoDGrid(nx, 2 * np.pi, ny, 2 * np.pi, x0=0, y0=0)
    adj = oa.ETDRK4Adjoint(og_, oph.vorticity_linear(og_, nu=nu), dt)
    Jr, gr = adj.gradient(_host(z0).astype(np.float64), _host(zobs).astype(np.float64), nsteps, ncheck)
    assert abs(J - Jr) < max(tol, 1e-12) * Jr * 10, (J, Jr)
    err = np.linalg.norm(_host(grad).astype(np.float64) - gr) / np.linalg.norm(gr)
    assert err < tol, err


def test_adjoint_repeatable_and_checkpoint_independent(ff):
    prob, z0, zobs = _setup(ff, 128, 1e-2, 0.01, 24, np.float64)
    J1, g1 = ff.adjoint.misfit_gradient(prob, z0, zobs, 24, 6)
    J2, g2 = ff.adjoint.misfit_gradient(prob, z0, zobs, 24, 24)
    J3, g3 = ff.adjoint.misfit_gradient(prob, z0, zobs, 24, 6)
    assert J1 == J3 and np.array_equal(_host(g1), _host(g3))
    assert J1 == J2 and np.array_equal(_host(g1), _host(g2))  # recompute is bitwise the forward


def test_adjoint_requires_matching_forward(ff):
    import torch

    prob, z0, zobs = _setup(ff, 64, 1e-2, 0.01, 4, np.float64)
    op, ts = prob.eqn.calcN, prob.timestepper
    s = torch.zeros_like(prob.sol)
    from fourierflows_hip import _lib

    with pytest.raises(ff.FFHIPError):
        _lib.call("ffhip_op_etdrk4_forward_ckpt", op.handle, ts.handle, s.data_ptr(), s.data_ptr(), 10, 11, 0)
    _lib.call("ffhip_op_etdrk4_forward_ckpt", op.handle, ts.handle, s.data_ptr(), s.data_ptr(), 10, 2, 0)
    with pytest.raises(ff.FFHIPError):
        _lib.call("ffhip_op_etdrk4_adjoint", op.handle, ts.handle, s.data_ptr(), s.data_ptr(), 10, 3, 0)
    _lib.call("ffhip_op_etdrk4_adjoint", op.handle, ts.handle, s.data_ptr(), s.data_ptr(), 10, 2, 0)


def test_autograd_matches_explicit(ff):
    import torch

    n, nsteps = 128, 20
    prob, z0, zobs = _setup(ff, n, 1e-2, 0.01, nsteps, np.float64)
    g = prob.grid
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, 4)
    x = z0.clone().requires_grad_(True)
    s = ff.adjoint.dealias(g, ff.adjoint.mul(g, x))
    zT = ff.adjoint.ldiv(g, ff.adjoint.etdrk4_solve(prob, s, nsteps, 4))
    Ja = 0.5 * torch.sum((zT - zobs) ** 2)
    Ja.backward()
    assert abs(float(Ja.detach()) - J) < 1e-12 * J
    err = float(torch.linalg.vector_norm(x.grad - grad) / torch.linalg.vector_norm(grad))
    assert err < 1e-13, err


def test_autograd_chained_solves(ff):
    """Two solves on one operator before backward: the first solve's
    checkpoints are replaced by the second's, and its pullback must rebuild
    them (same nsteps and ncheck, so only the run identifier tells them apart)."""
    import torch

    n, nsteps = 64, 10
    prob, z0, zobs = _setup(ff, n, 1e-2, 0.01, 2 * nsteps, np.float64)
    g = prob.grid
    # 2 nsteps steps with segments of 5: the same recompute schedule as the chained pair
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, 2 * nsteps, 4)
    x = z0.clone().requires_grad_(True)
    s = ff.adjoint.dealias(g, ff.adjoint.mul(g, x))
    s1 = ff.adjoint.etdrk4_solve(prob, s, nsteps, 2)
    s2 = ff.adjoint.etdrk4_solve(prob, s1, nsteps, 2)
    zT = ff.adjoint.ldiv(g, s2)
    Ja = 0.5 * torch.sum((zT - zobs) ** 2)
    Ja.backward()
    assert abs(float(Ja.detach()) - J) < 1e-12 * J
    err = float(torch.linalg.vector_norm(x.grad - grad) / torch.linalg.vector_norm(grad))
    assert err < 1e-12, err


def test_autograd_loss_over_two_solves(ff):
    """J = J(x) + J(2x) through two independent solves: dJ = g(x) + 2 g(2x)."""
    import torch

    n, nsteps = 64, 12
    prob, z0, zobs = _setup(ff, n, 1e-2, 0.01, nsteps, np.float64)
    g = prob.grid
    Ja_ref, ga = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, 3)
    Jb_ref, gb = ff.adjoint.misfit_gradient(prob, 2 * z0, zobs, nsteps, 3)
    x = z0.clone().requires_grad_(True)
    terms = []
    for scale in (1.0, 2.0):
        s = ff.adjoint.dealias(g, ff.adjoint.mul(g, scale * x))
        zT = ff.adjoint.ldiv(g, ff.adjoint.etdrk4_solve(prob, s, nsteps, 3))
        terms.append(0.5 * torch.sum((zT - zobs) ** 2))
    (terms[0] + terms[1]).backward()
    want = ga + 2 * gb
    assert abs(float(terms[0].detach()) - Ja_ref) < 1e-12 * Ja_ref
    assert abs(float(terms[1].detach()) - Jb_ref) < 1e-12 * Jb_ref
    err = float(torch.linalg.vector_norm(x.grad - want) / torch.linalg.vector_norm(want))
    assert err < 1e-12, err


def _taylor(ff, prob, z0, zobs, grad, nsteps, ndir, hs, seed):
    import torch

    g = prob.grid
    out = []
    gen = torch.Generator(device="cuda").manual_seed(seed)

    def J(x):
        s = torch.zeros_like(prob.sol)
        ff.mul(s, g.rfftplan, x)
        ff.dealias(s, g)
        ts = prob.timestepper
        from fourierflows_hip import _lib

        _lib.call("ffhip_op_etdrk4_run", prob.eqn.calcN.handle, ts.handle, s.data_ptr(), nsteps,
                  torch.cuda.current_stream().cuda_stream)
        zT = torch.zeros_like(x)
        ff.ldiv(zT, g.rfftplan, s)
        return 0.5 * float(torch.sum((zT.double() - zobs.double()) ** 2))

    for _ in range(ndir):
        d = torch.randn(z0.shape, generator=gen, device="cuda", dtype=torch.float64).to(z0.dtype)
        gd = float(torch.sum(grad.double() * d.double()))
        errs = [abs((J(z0 + h * d) - J(z0 - h * d)) / (2 * h) - gd) for h in hs]
        out.append((gd, errs))
    return out


def test_taylor_fp64_512_500_steps(ff):
    """The adjoint workload's grid and window in FP64: the central-FD error
    falls as h^2 until roundoff."""
    n, nu, dt, nsteps = 512, 5e-4, 0.002, 500
    prob, z0, zobs = _setup(ff, n, nu, dt, nsteps, np.float64)
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, 23)
    hs = [1e-1, 1e-2, 1e-3, 1e-4]
    for gd, errs in _taylor(ff, prob, z0, zobs, grad, nsteps, 8, hs, seed=7):
        print(f"taylor fp64: <g,d>={gd:.6e} errs={['%.3e' % e for e in errs]}")
        assert errs[1] < 0.03 * errs[0] + 1e-9 * abs(gd), (gd, errs)
        assert errs[2] < 0.03 * errs[1] + 1e-9 * abs(gd), (gd, errs)
        assert min(errs) < 1e-6 * abs(gd), (gd, errs)


def test_fp32_adjoint_512_500_steps(ff):
    """FP32 gradient (500 steps, 23 checkpoints) against the FP64 GPU gradient
    (itself pinned to the FP64 oracle above) and a central-FD Taylor test."""
    import torch

    n, nu, dt, nsteps = 512, 5e-4, 0.002, 500
    p32, z0, zobs = _setup(ff, n, nu, dt, nsteps, np.float32)
    J32, g32 = ff.adjoint.misfit_gradient(p32, z0, zobs, nsteps, 23)
    p64, z064, zobs64 = _setup(ff, n, nu, dt, nsteps, np.float64)
    J64, g64 = ff.adjoint.misfit_gradient(p64, z064, zobs64, nsteps, 23)
    print(f"J fp32 {J32!r} fp64 {J64!r}")
    assert abs(J32 - J64) < 1e-4 * J64, (J32, J64)
    err = float(torch.linalg.vector_norm(g32.double() - g64) / torch.linalg.vector_norm(g64))
    print(f"gradient fp32 vs fp64 rel l2 {err:.3e}")
    assert err < 1e-4, err
    # Taylor test of the FP32 gradient against central differences of the FP64
    # misfit (same fields: FP32 roundoff of J itself would swamp the FD at
    # small h): the error falls as h^2 to the FP32 gradient's own error
    hs = [1e-1, 1e-2, 1e-3, 1e-4]
    for gd, errs in _taylor(ff, p64, z064, zobs64, g32, nsteps, 8, hs, seed=8):
        print(f"taylor fp32 grad: <g,d>={gd:.6e} errs={['%.3e' % e for e in errs]}")
        assert errs[1] < 0.03 * errs[0] + 1e-4 * abs(gd), (gd, errs)
        assert min(errs) < 1e-4 * abs(gd), (gd, errs)
