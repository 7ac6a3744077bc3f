"""Full-length workloads against the FP64 oracle's golden fixtures.

The fixtures (tests/golden/*.npz) were written by tests/golden/make_golden.py
from the FP64 CPU oracle: the 2048² hyperviscous run over its 2000 steps,
three 128² members of the mixed-grid ensemble over 500 steps, and the 512²
misfit gradient over 500 steps with 23 checkpoints.  The FP64 GPU results
must match to relative 1e-10 (field blocks, norms, energy, enstrophy, J and
the gradient), the FP32 results (FP64 coefficients and reductions) to 1e-4.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = {np.float64: 1e-10, np.float32: 1e-4}


def _load(name):
    return np.load(os.path.join(HERE, name))


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.complex128 if t.is_complex() else np.float64)


def _low_block(zh, K):
    return np.concatenate([zh[: K + 1, : K + 1], zh[-K:, : K + 1]], axis=0)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_vorticity_2048_full_run_vs_oracle(ff, T):
    """The 2048² workload (ν₈|k|⁸ + μ = 0.01, dt = 1e-3, seed 1) over 2000
    steps: energy and enstrophy every 200 steps, the |kx|, |l| <= 32 block of
    ζ̂ and ‖ζ̂‖ after 1000 and 2000 steps."""
    import torch

    gd = _load("vort2048_etdrk4_2000.npz")
    n, dt, K = int(gd["n"]), float(gd["dt"]), int(gd["K"])
    p = ff.TwoDVorticityProblem(nx=n, nu_n=float(gd["nu8"]), n_hyper=4, mu=float(gd["mu"]), dt=dt, T=T)
    assert p.eqn.calcN.fused
    zh0 = ff.vorticity_ic(p.grid, seed=int(gd["seed"]))
    p.sol.copy_(zh0 if T == np.float64 else zh0.to(torch.complex64))
    tol = TOL[T]
    steps = [int(s) for s in gd["steps"]]
    for i, k in enumerate(steps):
        if k > 0:
            ff.stepforward(p, k - steps[i - 1])
        e, z = p.eqn.calcN.energy_enstrophy(p.sol)
        assert abs(e - gd["energy"][i]) <= tol * gd["energy"][i], (k, e, gd["energy"][i])
        assert abs(z - gd["enstrophy"][i]) <= tol * gd["enstrophy"][i], (k, z, gd["enstrophy"][i])
        if f"zh_low_{k}" in gd.files:
            zh = _host(p.sol)
            assert _rel(_low_block(zh, K), gd[f"zh_low_{k}"]) <= tol, k
            assert abs(np.linalg.norm(zh) - gd[f"zh_norm_{k}"]) <= tol * gd[f"zh_norm_{k}"], k
    assert p.clock.step == 2000


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_ensemble_members_full_run_vs_oracle(ff, T):
    """The mixed-grid FP32 workload's 64 members stepped together for their
    500 steps (FP64: the same ensemble in FP64); members 0, 15 and 30 (128²,
    own ν, seed and k0) against the oracle."""
    gd = _load("ensemble128_etdrk4_500.npz")
    dt, nsteps, K = float(gd["dt"]), int(gd["nsteps"]), int(gd["K"])
    members = ff.ensemble_members()
    ens = ff.VorticityEnsemble([dict(nx=n, nu=nu) for (n, sd, nu, k0) in members], dt, T=T)
    for s, g, (n, sd, nu, k0) in zip(ens.sols, ens.grids, members):
        s.copy_(ff.vorticity_ic(g, seed=sd, k0=k0, spectrum="k4exp2"))
    ens.run(nsteps)
    tol = TOL[T]
    for m in gd["members"]:
        m = int(m)
        zh = _host(ens.sols[m])
        assert _rel(_low_block(zh, K), gd[f"zh_low_{m}"]) <= tol, m
        assert abs(np.linalg.norm(zh) - gd[f"zh_norm_{m}"]) <= tol * gd[f"zh_norm_{m}"], m


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_adjoint_512_full_window_vs_oracle(ff, T):
    """J and dJ/dζ₀ at 512² over the whole 500-step window, 23 checkpoints."""
    import torch

    gd = _load("adjoint512_etdrk4_500.npz")
    n, nu, dt, nsteps, ncheck, K = (int(gd["n"]), float(gd["nu"]), float(gd["dt"]), int(gd["nsteps"]),
                                    int(gd["ncheck"]), int(gd["K"]))
    prob = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, T=T)
    g = prob.grid
    rt = torch.float64 if T == np.float64 else torch.float32
    z0 = torch.zeros(g.ny, g.nx, dtype=rt, device="cuda")
    ff.ldiv(z0, g.rfftplan, ff.vorticity_ic(g, seed=1))
    prob.sol.copy_(ff.vorticity_ic(g, seed=2))
    ff.stepforward(prob, nsteps)
    zobs = torch.zeros_like(z0)
    ff.ldiv(zobs, g.rfftplan, prob.sol)
    J, grad = ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, ncheck)
    tol = TOL[T]
    assert abs(J - gd["J"]) <= tol * gd["J"], (J, float(gd["J"]))
    gr = _host(grad)
    assert abs(np.linalg.norm(gr) - gd["grad_norm"]) <= tol * gd["grad_norm"]
    assert _rel(_low_block(np.fft.rfft2(gr), K), gd["gradh_low"]) <= tol
