"""Write the long-run golden fixtures from the FP64 oracle.

The GPU tests check the library's full-length workloads against these
vectors (the oracle needs minutes to hours on the host for them, so they are
computed once here and committed as small npz files):

* ``vort2048_etdrk4_2000.npz`` -- the 2048² hyperviscous vorticity workload
  (ν₈|k|⁸ with damping 1/dt at 2/3 of k_max, μ = 0.01, dt = 1e-3, seed 1):
  energy and enstrophy every 200 steps over 2000 steps, the low-mode block
  |kx|, |l| ≤ 32 of ζ̂ and ‖ζ̂‖ after 1000 and 2000 steps.
* ``ensemble128_etdrk4_500.npz`` -- three 128² members of the mixed-grid
  ensemble (their own ν, seed and k0; dt = 0.002) after 500 steps: low-mode
  blocks, ‖ζ̂‖, energy and enstrophy.
* ``adjoint512_etdrk4_500.npz`` -- the 512² misfit gradient (ν = 5e-4,
  dt = 0.002, 500 steps, 23 checkpoints, ζ₀ from seed 1, ζ_obs a forward run
  from seed 2): J, ‖∇J‖ and the low-mode block of rfft2(∇J).

    FFORACLE_WORKERS=8 python tests/golden/make_golden.py [vort2048|ensemble|adjoint ...]
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import adjoint as oa  # noqa: E402
from oracle import grids as og  # noqa: E402
from oracle import ics as oics  # noqa: E402
from oracle import physics as oph  # noqa: E402
from oracle import timesteppers as ots  # noqa: E402
from oracle import utils as ou  # noqa: E402

KB = 32  # low-mode block half-width


def low_block(zh, K=KB):
    """Rows l = 0..K and -K..-1, columns kx = 0..K of an rfft-layout array [ny][nkr]."""
    return np.ascontiguousarray(np.concatenate([zh[: K + 1, : K + 1], zh[-K:, : K + 1]], axis=0))


def vort2048():
    n, dt, nsteps, every = 2048, 1e-3, 2000, 200
    nu8 = oph.hyperviscosity_for(n, dt)
    g = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    L = oph.vorticity_linear(g, nu_n=nu8, n_hyper=4, mu=0.01)
    p = ots.Problem(ots.Equation(L, oph.vorticity_calcN, g), "ETDRK4", dt, g)
    p.sol[...] = oics.vorticity_ic(g, seed=1)
    steps, E, Z, out = [0], [ou.energy(p.sol, g)], [ou.enstrophy(p.sol, g)], {}
    t0 = time.time()
    for k in range(every, nsteps + 1, every):
        p.stepforward(every)
        steps.append(k)
        E.append(ou.energy(p.sol, g))
        Z.append(ou.enstrophy(p.sol, g))
        if k in (1000, 2000):
            out[f"zh_low_{k}"] = low_block(p.sol)
            out[f"zh_norm_{k}"] = np.float64(np.linalg.norm(p.sol))
        print(f"vort2048: step {k} ({time.time() - t0:.0f} s)", flush=True)
    np.savez(os.path.join(HERE, "vort2048_etdrk4_2000.npz"), n=n, dt=dt, nu8=nu8, mu=0.01, seed=1, K=KB,
             steps=np.array(steps), energy=np.array(E), enstrophy=np.array(Z), **out)


def ensemble():
    dt, nsteps = 0.002, 500
    members = oics.ensemble_members()
    idx = [m for m, mem in enumerate(members) if mem[0] == 128][:: 15][:3]  # members 0, 15, 30
    out = {"members": np.array(idx), "dt": dt, "nsteps": nsteps, "K": KB}
    for m in idx:
        n, sd, nu, k0 = members[m]
        g = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
        p = ots.Problem(ots.Equation(oph.vorticity_linear(g, nu=nu), oph.vorticity_calcN, g), "ETDRK4", dt, g)
        p.sol[...] = oics.vorticity_ic(g, seed=sd, k0=k0, spectrum="k4exp2")
        p.stepforward(nsteps)
        out[f"zh_low_{m}"] = low_block(p.sol)
        out[f"zh_norm_{m}"] = np.float64(np.linalg.norm(p.sol))
        out[f"energy_{m}"] = np.float64(ou.energy(p.sol, g))
        out[f"enstrophy_{m}"] = np.float64(ou.enstrophy(p.sol, g))
        print(f"ensemble: member {m} (n={n}, nu={nu:.3e}, k0={k0:.2f}) done", flush=True)
    np.savez(os.path.join(HERE, "ensemble128_etdrk4_500.npz"), **out)


def adjoint():
    n, nu, dt, nsteps, ncheck = 512, 5e-4, 0.002, 500, 23
    g = og.TwoDGrid(n, 2 * np.pi, x0=0, y0=0)
    L = oph.vorticity_linear(g, nu=nu)
    z0 = g.ldiv(oics.vorticity_ic(g, seed=1))
    p = ots.Problem(ots.Equation(L, oph.vorticity_calcN, g), "ETDRK4", dt, g)
    p.sol[...] = oics.vorticity_ic(g, seed=2)
    p.stepforward(nsteps)
    zobs = g.ldiv(p.sol)
    t0 = time.time()
    J, grad = oa.ETDRK4Adjoint(g, L, dt).gradient(z0, zobs, nsteps, ncheck)
    print(f"adjoint: J = {J:.6e} ({time.time() - t0:.0f} s)", flush=True)
    gh = np.fft.rfft2(grad)
    np.savez(os.path.join(HERE, "adjoint512_etdrk4_500.npz"), n=n, nu=nu, dt=dt, nsteps=nsteps, ncheck=ncheck,
             K=KB, J=np.float64(J), grad_norm=np.float64(np.linalg.norm(grad)), gradh_low=low_block(gh))


if __name__ == "__main__":
    which = sys.argv[1:] or ["ensemble", "adjoint", "vort2048"]
    for w in which:
        {"vort2048": vort2048, "ensemble": ensemble, "adjoint": adjoint}[w]()
