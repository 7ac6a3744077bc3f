"""GPU batched ensembles of 2D vorticity problems on mixed grids.

Each member of an ensemble is the same computation as the single-grid fused
path (itself pinned to the FP64 oracle): members agree with single-grid runs
to a few ulps and with the FP64 oracle to 1e-12; graph replays are bitwise
equal to eager launches; an autotuned variant (another radix schedule)
agrees to roundoff.
"""

import numpy as np
import pytest

from oracle import grids as og
from oracle import ics as oics
from oracle import physics as oph
from oracle import timesteppers as ots
from oracle import utils as ou

pytestmark = pytest.mark.gpu

MEMBERS = [dict(nx=64, nu=3e-3), dict(nx=128, nu=1e-3, mu=0.01), dict(nx=256, nu=5e-4),
           dict(nx=128, ny=64, Ly=np.pi, nu=2e-3), dict(nx=64, ny=256, Ly=8 * np.pi, nu=1e-3),
           dict(nx=512, nu=2e-4), dict(nx=128, nu=4e-3)]


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _singles(ff, members, dt, T, seeds):
    out = []
    for m, sd in zip(members, seeds):
        p = ff.TwoDVorticityProblem(nx=m["nx"], ny=m.get("ny"), Lx=m.get("Lx", 2 * np.pi),
                                    Ly=m.get("Ly", m.get("Lx", 2 * np.pi)), nu=m.get("nu", 0.0),
                                    mu=m.get("mu", 0.0), dt=dt, T=T)
        assert p.eqn.calcN.fused
        p.sol.copy_(ff.vorticity_ic(p.grid, seed=sd, k0=5.0))
        out.append(p)
    return out


@pytest.mark.parametrize("T,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
def test_ensemble_matches_single_grid(ff, T, tol):
    """Same kernels bodies and coefficients as the single-grid fused path; the
    two instantiations may contract multiply-adds differently, so agreement
    is to a few ulps rather than bitwise."""
    dt = 0.004
    seeds = list(range(11, 11 + len(MEMBERS)))
    ens = ff.VorticityEnsemble(MEMBERS, dt, T=T)
    singles = _singles(ff, MEMBERS, dt, T, seeds)
    for s, p in zip(ens.sols, singles):
        s.copy_(p.sol)
    for n in (23, 7, 10):
        ens.run(n)
        for p in singles:
            ff.stepforward(p, n)
        for m, (s, p) in enumerate(zip(ens.sols, singles)):
            a, b = _host(s).astype(np.complex128), _host(p.sol).astype(np.complex128)
            err = np.linalg.norm(a - b) / np.linalg.norm(b)
            assert err < tol, (T, n, m, err)


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_graph_replays_equal_eager_steps(ff, T):
    """Runs of >= 10 steps replay a captured HIP graph; single-step runs launch
    eagerly.  Both must give bitwise the same states, replay after replay,
    with eager work and host synchronisations in between."""
    dt = 0.004
    A = ff.VorticityEnsemble(MEMBERS, dt, T=T)
    B = ff.VorticityEnsemble(MEMBERS, dt, T=T)
    for m, (a, b, g) in enumerate(zip(A.sols, B.sols, A.grids)):
        z = ff.vorticity_ic(g, seed=100 + m, k0=5.0)
        a.copy_(z)
        b.copy_(z)
    done = 0
    for n in (23, 10, 7, 20):
        A.run(n)
        for _ in range(n):
            B.run(1)
        done += n
        for m, (a, b) in enumerate(zip(A.sols, B.sols)):
            assert np.array_equal(_host(a), _host(b)), (T, done, m)


def test_ensemble_fp64_matches_oracle(ff):
    dt, steps = 0.004, 12
    members = MEMBERS[:5]
    seeds = [3, 4, 5, 6, 7]
    ens = ff.VorticityEnsemble(members, dt, T=np.float64)
    singles = _singles(ff, members, dt, np.float64, seeds)
    for s, p in zip(ens.sols, singles):
        s.copy_(p.sol)
    zh0 = [_host(p.sol) for p in singles]
    ens.run(steps)
    for m, (mem, z0) in enumerate(zip(members, zh0)):
        Lx = mem.get("Lx", 2 * np.pi)
        o = og.TwoDGrid(mem["nx"], Lx, mem.get("ny", mem["nx"]), mem.get("Ly", Lx), x0=0, y0=0)
        L = oph.vorticity_linear(o, nu=mem.get("nu", 0.0), mu=mem.get("mu", 0.0))
        p = ots.Problem(ots.Equation(L, oph.vorticity_calcN, o), "ETDRK4", dt, o)
        p.sol[...] = z0
        p.stepforward(steps)
        err = ou.rel_l2(_host(ens.sols[m]), p.sol)
        assert err < 1e-12, (m, err)


@pytest.mark.parametrize("variant", ["128,2048,2048", "256,128,128", "256,512,128", "128,128,512", "256,256,2048",
                                     "128,256,256,1", "256,2048,2048,1"])
def test_fixed_variants(ff, variant, monkeypatch):
    """Every (grid size, pass) body with 8 and with 16 points per thread, in
    both workgroup sizes, one launch per pass or per stage pair, equals the
    single-grid runs (FP64).  A variant is "workgroup size, column threshold,
    row threshold[, stage pair]": sizes from a threshold up take 16 points
    per thread."""
    dt, steps = 0.004, 6
    seeds = list(range(51, 51 + len(MEMBERS)))
    monkeypatch.setenv("FFHIP_ENS_VARIANT", variant)
    ens = ff.VorticityEnsemble(MEMBERS, dt, T=np.float64)
    monkeypatch.delenv("FFHIP_ENS_VARIANT")
    bd, tc, tr, pair = ([int(x) for x in variant.split(",")] + [0])[:4]
    mask = lambda t: sum(1 << s for s in range(5) if (64 << s) >= t)  # noqa: E731
    threads, masks, _ = ens.info()
    assert (threads, masks) == (bd, mask(tc) | (mask(tr) << 8) | (pair << 16))
    singles = _singles(ff, MEMBERS, dt, np.float64, seeds)
    for s, p in zip(ens.sols, singles):
        s.copy_(p.sol)
    ens.run(steps)
    for m, (s, p) in enumerate(zip(ens.sols, singles)):
        ff.stepforward(p, steps)
        assert ou.rel_l2(_host(s), _host(p.sol)) < 1e-12, (m, variant)


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_stage_pair_equals_per_pass(ff, T, monkeypatch):
    """One launch per stage pair (per-member waits inside the launch) gives
    the same bits as one launch per pass with the same radix schedule, over
    graph replays (25 steps) and again in a second run."""
    dt = 0.004
    seeds = list(range(61, 61 + len(MEMBERS)))
    outs = []
    for variant in ("256,256,256,0", "256,256,256,1"):
        monkeypatch.setenv("FFHIP_ENS_VARIANT", variant)
        ens = ff.VorticityEnsemble(MEMBERS, dt, T=T)
        monkeypatch.delenv("FFHIP_ENS_VARIANT")
        for s, p in zip(ens.sols, _singles(ff, MEMBERS, dt, T, seeds)):
            s.copy_(p.sol)
        ens.run(25)
        ens.run(13)
        outs.append([_host(s) for s in ens.sols])
    for m, (a, b) in enumerate(zip(*outs)):
        assert np.array_equal(a, b), m


@pytest.mark.parametrize("T,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_autotuned_ensemble(ff, T, tol):
    dt, steps = 0.004, 20
    seeds = list(range(31, 31 + len(MEMBERS)))
    ens = ff.VorticityEnsemble(MEMBERS, dt, T=T, autotune=True)
    singles = _singles(ff, MEMBERS, dt, T, seeds)
    for s, p in zip(ens.sols, singles):
        s.copy_(p.sol)
    ens.run(steps)
    threads, prule, ms = ens.info()
    assert threads in (128, 256) and 0 <= prule < (1 << 17) and ms > 0
    for m, (s, p) in enumerate(zip(ens.sols, singles)):
        ff.stepforward(p, steps)
        a, b = _host(s).astype(np.complex128), _host(p.sol).astype(np.complex128)
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < tol, (m, err, threads, prule)


def test_mixed_grid_workload_members(ff):
    """The 64-member FP32 workload (32 x 128^2, 24 x 256^2, 8 x 512^2, own nu,
    seed and k0, dt = 0.002): a short run equals per-member single-grid runs."""
    members = ff.ensemble_members()
    dt = 0.002
    spec = [dict(nx=n, nu=nu) for (n, sd, nu, k0) in members]
    ens = ff.VorticityEnsemble(spec, dt, T=np.float32)
    for s, g, (n, sd, nu, k0) in zip(ens.sols, ens.grids, members):
        s.copy_(ff.vorticity_ic(g, seed=sd, k0=k0, spectrum="k4exp2"))
    z0 = [s.clone() for s in ens.sols]
    ens.run(12)
    for m in (0, 31, 32, 55, 56, 63):
        n, sd, nu, k0 = members[m]
        p = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, T=np.float32)
        p.sol.copy_(z0[m])
        ff.stepforward(p, 12)
        a, b = _host(ens.sols[m]).astype(np.complex128), _host(p.sol).astype(np.complex128)
        assert np.linalg.norm(a - b) < 1e-5 * np.linalg.norm(b), m


def test_ensemble_rejects_bad_grids(ff):
    with pytest.raises(ff.FFHIPError):
        ff.VorticityEnsemble([dict(nx=96)], 0.01)
    with pytest.raises(ff.FFHIPError):
        ff.VorticityEnsemble([dict(nx=2048)], 0.01)
    with pytest.raises(ff.FFHIPError):
        ff.VorticityEnsemble([dict(nx=64)], 0.01, aliased_fraction=0)
