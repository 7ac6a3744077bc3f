"""Benchmark: ETDRK4 steps/s of 2D vorticity at 2048² on one MI355X.

Workload (the headline configuration): 2D incompressible vorticity on
[0, 2π)², nx = ny = 2048, hyperviscosity ν₈|k|⁸ with ν₈ set so the damping at
2/3 of k_max is 1/dt, linear drag μ = 0.01, ETDRK4 with dt = 1e-3, seeded
k⁴exp(-(k/6)²) initial vorticity at rms 1.  FP32 bulk with FP64 ETDRK4
coefficients is the measured path (`value`); the same run in FP64 is reported
beside it.

A step is one ETDRK4 step of the whole 2048² field (4 nonlinear-term
evaluations).  K steps are timed with HIP events after W warm-up steps, the
state already resident in HBM.  The per-kernel roofline comes from a second,
event-bracketed pass over the same kernels; the CPU baseline is the FP64
NumPy/pocketfft restatement of the reference's algorithm timed on a bounded
sample of steps on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fourierflows.jl_amd"))

METRIC = "ETDRK4 steps/s at 2048² FP32 vs FP64 on GPU; FFT HBM GB/s; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "vorticity_2048_fp32_hbm_traffic.json")
ROCPROF_NAMES = {"vort_row": "k_vrowh<2048>",
                 **{f"vort_col_stage{s}": f"k_vcol<float, 2048, 16, {s}>" for s in range(5)}}
N = 2048
DT = 1e-3
MU = 0.01


def nu8_for(n, dt):
    kcut = (n / 2) * (1 - 1 / 3)
    return 1.0 / (dt * kcut**8)


def make_problem(ff, T):
    import numpy as np

    return ff.TwoDVorticityProblem(nx=N, nu_n=nu8_for(N, DT), n_hyper=4, mu=MU, dt=DT,
                                   T=np.float32 if T == 32 else np.float64)


def time_steps(ff, torch, prob, steps, warmup):
    ff.stepforward(prob, warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ff.stepforward(prob, steps)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1000.0


GOLDEN = os.path.join(ROOT, "tests", "golden", "vort2048_etdrk4_2000.npz")


def accuracy_vs_oracle(ff, torch):
    """The workload's full 2000-step run in FP32 and in FP64 against the FP64
    CPU oracle's fixture (written by tests/golden/make_golden.py): the largest
    relative error of energy and enstrophy over the steps saved every 200, the
    relative L2 error of the |kx|, |l| <= 32 block of the spectral vorticity
    after 1000 and 2000 steps, and the FP32 field against the FP64 field."""
    import numpy as np

    if not os.path.exists(GOLDEN):
        return None
    gd = np.load(GOLDEN)
    K = int(gd["K"])
    steps = [int(x) for x in gd["steps"]]
    out, finals = {"steps": steps[-1], "reference": "FP64 CPU oracle (numpy/pocketfft)"}, {}
    for T in (32, 64):
        p = make_problem(ff, T)
        z0 = ff.vorticity_ic(p.grid, seed=int(gd["seed"]))
        p.sol.copy_(z0)
        e_err = z_err = 0.0
        blocks = {}
        for i, k in enumerate(steps):
            if k > 0:
                ff.stepforward(p, k - steps[i - 1])
            e, z = p.eqn.calcN.energy_enstrophy(p.sol)
            e_err = max(e_err, abs(e - gd["energy"][i]) / gd["energy"][i])
            z_err = max(z_err, abs(z - gd["enstrophy"][i]) / gd["enstrophy"][i])
            if f"zh_low_{k}" in gd.files:
                zh = p.sol.to(torch.complex128).cpu().numpy()
                blk = np.concatenate([zh[: K + 1, : K + 1], zh[-K:, : K + 1]], axis=0)
                ref = gd[f"zh_low_{k}"]
                blocks[str(k)] = float(np.linalg.norm(blk - ref) / np.linalg.norm(ref))
        finals[T] = p.sol.to(torch.complex128)
        out[f"fp{T}"] = {"energy": e_err, "enstrophy": z_err, "low_mode_block": blocks}
    out["fp32_field_vs_fp64_field"] = float(torch.linalg.vector_norm(finals[32] - finals[64]) /
                                            torch.linalg.vector_norm(finals[64]))
    return out


def kernel_profile(ff, prob, steps):
    """Per-kernel averages over `steps` event-bracketed steps."""
    from fourierflows_hip import _lib

    _lib.profile_start()
    ff.stepforward(prob, steps)
    rep = _lib.profile_stop()
    out = {}
    for k in rep["kernels"]:
        avg_ms = k["total_ms"] / k["count"]
        bytes_per = k["bytes"] / k["count"]
        out[k["name"]] = {"count": k["count"], "avg_us": avg_ms * 1e3, "total_ms": k["total_ms"],
                          "bytes_per_launch": bytes_per, "GBps": bytes_per / (avg_ms * 1e-3) / 1e9}
    return out


def adjoint_timing(ff, torch, n=512, nu=5e-4, dt=0.002, nsteps=500, ncheck=23):
    """Adjoint-gradient cost against one forward solve (FP32, 512^2, 500 steps,
    23 checkpoints): J = 1/2 ||zeta(T) - zeta_obs||^2 and dJ/dzeta0."""
    import numpy as np

    prob = ff.TwoDVorticityProblem(nx=n, nu=nu, dt=dt, T=np.float32)
    g = prob.grid
    z0 = torch.zeros(g.ny, g.nx, dtype=torch.float32, device="cuda")
    ff.ldiv(z0, g.rfftplan, ff.vorticity_ic(g, seed=1))
    prob.sol.copy_(ff.vorticity_ic(g, seed=2))
    ff.stepforward(prob, nsteps)
    zobs = torch.zeros_like(z0)
    ff.ldiv(zobs, g.rfftplan, prob.sol)
    ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, ncheck)  # warm-up (allocates checkpoints)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    prob.sol.copy_(ff.vorticity_ic(g, seed=1))
    ff.stepforward(prob, nsteps)
    e[1].record()
    e[2].record()
    ff.adjoint.misfit_gradient(prob, z0, zobs, nsteps, ncheck)
    e[3].record()
    torch.cuda.synchronize()
    tf, tg = e[0].elapsed_time(e[1]), e[2].elapsed_time(e[3])
    return {"workload": f"{n}^2 FP32, {nsteps} ETDRK4 steps, {ncheck} checkpoints", "forward_ms": round(tf, 3),
            "gradient_ms": round(tg, 3), "gradient_over_forward": round(tg / tf, 3)}


def ensemble_timing(ff, torch, nsteps=500, dt=0.002):
    """The mixed-grid ensemble: 64 FP32 members (32 x 128^2, 24 x 256^2,
    8 x 512^2), own seed, nu log-uniform in [1e-4, 1e-2] and spectrum peak
    k0 uniform in [4, 10] (E ~ k^4 exp(-2 (k/k0)^2)), ETDRK4 dt = 0.002,
    `nsteps` steps in one run (autotuned variant, HIP-graph replays)."""
    import numpy as np

    members = ff.ensemble_members()
    ens = ff.VorticityEnsemble([dict(nx=n, nu=nu) for (n, sd, nu, k0) in members], dt, T=np.float32,
                               autotune=True)
    for s, g, (n, sd, nu, k0) in zip(ens.sols, ens.grids, members):
        s.copy_(ff.vorticity_ic(g, seed=sd, k0=k0, spectrum="k4exp2"))
    ens.run(20)  # tunes, captures the graph
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ens.run(nsteps)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1000.0
    threads, prule, tuned_ms = ens.info()
    return {"workload": "64 FP32 members: 32 x 128^2, 24 x 256^2, 8 x 512^2, ETDRK4 dt=0.002",
            "steps": nsteps, "seconds": round(t, 4), "ensemble_steps_per_s": round(nsteps / t, 2),
            "realization_steps_per_s": round(len(members) * nsteps / t, 1),
            "grid_points_per_step": int(sum(n * n for (n, _, _, _) in members)),
            "variant": {"threads": threads,
                        "points_per_thread": {f"{64 << c}": {"cols": 16 if (prule >> c) & 1 else 8,
                                                             "rows": 16 if (prule >> (8 + c)) & 1 else 8}
                                              for c in range(5) if (64 << c) in (128, 256, 512)},
                        "tuned_ms_per_step": round(tuned_ms, 4)}}


def tracer_timing(ff, torch, n=256, nsteps=300, dt=0.0025, kappa=1e-3):
    """3D ABC-flow tracer, 256^3 FP32, RK4: forward steps/s and the cost of the
    adjoint gradient of 1/2 ||c(T) - c_obs||^2 (forward + RK4 adjoint)."""
    import numpy as np

    p = ff.TracerProblem(n=n, kappa=kappa, dt=dt, T=np.float32)
    g = p.grid
    c0 = ff.tracer_ic(g, seed=1, ngauss=16, sigma=0.3)
    cobs = ff.tracer_ic(g, seed=2, ngauss=16, sigma=0.3)
    s = torch.zeros_like(p.sol)
    ff.mul(s, g.rfftplan, c0)
    ff.dealias(s, g)
    p.sol.copy_(s)
    ff.stepforward(p, 2)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    p.sol.copy_(s)
    e[0].record()
    ff.stepforward(p, nsteps)
    e[1].record()
    e[2].record()
    J, grad = ff.adjoint.tracer_misfit_gradient(p, c0, cobs, nsteps)
    e[3].record()
    torch.cuda.synchronize()
    tf, tg = e[0].elapsed_time(e[1]) / 1000.0, e[2].elapsed_time(e[3]) / 1000.0
    return {"workload": f"{n}^3 FP32 ABC tracer, RK4 dt={dt}, kappa={kappa}, {nsteps} steps",
            "steps_per_s": round(nsteps / tf, 2), "forward_s": round(tf, 4), "gradient_s": round(tg, 4),
            "gradient_over_forward": round(tg / tf, 3),
            "note": "dt=0.01 is outside RK4's stability region for this flow at 256^3 (growth 2.4 per step)"}


def vort4096_timing(ff, torch, steps=50):
    """The same hyperviscous vorticity problem at 4096² in FP32 and in FP64
    (the size the reference's docs time, in its default FP64:
    docs/src/assets/CPUversusGPU.jl:11, 23.724 s per 100 GPU steps): fused
    steps/s over `steps` steps."""
    import numpy as np

    n = 4096
    out = {"workload": f"2D vorticity {n}^2, nu8|k|^8 + mu=0.01, dt=1e-3", "steps": steps,
           "reference_published_gpu_steps_per_s": round(100 / 23.724, 3)}
    for T, key in ((np.float32, "fp32"), (np.float64, "fp64")):
        p = ff.TwoDVorticityProblem(nx=n, nu_n=nu8_for(n, DT), n_hyper=4, mu=MU, dt=DT, T=T)
        if not p.eqn.calcN.fused:
            return None
        p.sol.copy_(ff.vorticity_ic(p.grid, seed=1).to(p.sol.dtype))
        t = time_steps(ff, torch, p, steps, 5)
        out[key] = {"steps_per_s": round(steps / t, 2), "ms_per_step": round(1e3 * t / steps, 4)}
        del p
    out["steps_per_s"] = out["fp32"]["steps_per_s"]
    return out


def fft_plan_timing(ff, torch, n=N, reps=50):
    """mul!/ldiv! on the grid's rfft plan (src/domains.jl:5, src/diffusion.jl:137-139)
    at n x n in FP32 and FP64: each transform's HBM bytes are its input plus
    its output (real n^2 and complex (n/2+1) n arrays), over the event-timed
    average of `reps` calls; per-kernel averages from the library profiler."""
    import numpy as np

    from fourierflows_hip import _lib

    out = {}
    for T, rt, ct in ((32, torch.float32, torch.complex64), (64, torch.float64, torch.complex128)):
        plan = ff.Plan((n, n), np.float32 if T == 32 else np.float64, "rfft")
        x = torch.randn(n, n, dtype=rt, device="cuda")
        X = torch.zeros(n, n // 2 + 1, dtype=ct, device="cuda")
        y = torch.zeros_like(x)
        byts = x.numel() * x.element_size() + X.numel() * X.element_size()
        rec = {}
        for name, fn in (("mul", lambda: plan.forward(X, x)), ("ldiv", lambda: plan.inverse(y, X))):
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / 1000.0 / reps
            _lib.profile_start()
            for _ in range(5):
                fn()
            prof = _lib.profile_stop()
            kern = {k["name"]: {"count": k["count"] // 5, "avg_us": round(1e3 * k["total_ms"] / k["count"], 3),
                                "GBps": round(k["bytes"] / k["count"] / (k["total_ms"] / k["count"] * 1e-3) / 1e9, 1)}
                    for k in prof["kernels"]}
            gbps = byts / t / 1e9
            rec[name] = {"us": round(t * 1e6, 3), "io_bytes": byts, "GBps": round(gbps, 1),
                         "frac": round(gbps / HBM_PEAK_GBS, 4), "kernels": kern}
        err = float(torch.linalg.vector_norm(y - x) / torch.linalg.vector_norm(x))
        rec["roundtrip_rel_l2"] = err
        out[f"fp{T}"] = rec
        del plan
    return {"workload": f"rfft plan {n}x{n}: mul! (r2c) and ldiv! (c2r), device arrays", **out}


def algorithmic_bytes(kernel, n, cplx_bytes):
    """Algorithmic bytes of one launch of a fused 2D kernel on an n x n grid,
    from the per-step model of the unfused algorithm: S = (n/2+1) n complex
    values, C = (n/2+1) n FP64 coefficients; per calcN the x pass reads 3 S
    and writes 2 S, the y-inverse pass reads 1 S and writes 3 S, the
    y-forward pass reads 2 S and writes 1 S; the ETDRK4 combinations move
    2S+2C, 2S+2C, 3S+2C and 5S+4C (70 S per FP32 step).  A fused column stage
    is y-forward + combination + y-inverse.  The kernels move less (the kept
    band only): that is kept_band_bytes_per_launch."""
    S = (n // 2 + 1) * n * cplx_bytes
    C = (n // 2 + 1) * n * 8
    table = {"vort_row": 5 * S, "vort_col_stage0": 4 * S, "vort_col_stage1": 3 * S + 2 * S + 2 * C + 4 * S,
             "vort_col_stage2": 3 * S + 2 * S + 2 * C + 4 * S, "vort_col_stage3": 3 * S + 3 * S + 2 * C + 4 * S,
             "vort_col_stage4": 3 * S + 5 * S + 4 * C + 4 * S}
    return float(table[kernel])


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by the counters' reading of a 256 MiB copy), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f).get(ROCPROF_NAMES.get(kernel, ""))
    except (OSError, ValueError):
        return None
    return None if rec is None else round(rec["bytes"])


def cpu_baseline(seconds=12.0):
    """FP64 oracle (NumPy + scipy.fft, reference operation order) on host cores."""
    import numpy as np

    workers = min(16, os.cpu_count() or 1)
    os.environ["FFORACLE_WORKERS"] = str(workers)
    import importlib

    import oracle.grids as og
    importlib.reload(og)
    from oracle import ics as oics
    from oracle import physics as oph
    from oracle import timesteppers as ots

    g = og.TwoDGrid(N, 2 * np.pi, x0=0, y0=0)
    L = oph.vorticity_linear(g, nu_n=nu8_for(N, DT), n_hyper=4, mu=MU)
    eqn = ots.Equation(L, oph.vorticity_calcN, g)
    p = ots.Problem(eqn, "ETDRK4", DT, g)
    p.sol[...] = oics.vorticity_ic(g, seed=1)
    p.stepforward(1)  # warm pocketfft's plan cache
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        p.stepforward(1)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "steps/s", "cores": workers, "kind": "port",
            "sample": f"{n} FP64 ETDRK4 steps of the same 2048² problem (NumPy/pocketfft, "
                      f"{workers} FFT workers), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-cost measurement")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the mixed-grid ensemble measurement")
    ap.add_argument("--no-tracer", action="store_true", help="skip the 256^3 tracer measurement")
    ap.add_argument("--no-fft", action="store_true", help="skip the rfft plan (mul!/ldiv!) measurement")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (calibrates rocprofv3 FETCH/WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the 2048² step is a single-GPU workload; ensembles shard without exchange)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import fourierflows_hip as ff

    ff.init(0)
    if a.calib_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print(f"# bench: 2D vorticity {N}^2 ETDRK4, steps={a.steps} warmup={a.warmup}", file=sys.stderr, flush=True)

    # Both timed problems are set up first (their ETDRK4 coefficients are
    # host work), then the accuracy run (2000 FP32 and FP64 steps against the
    # oracle fixture) and each precision's per-kernel profile (event-bracketed
    # steps of that problem) keep the GPU busy right up to the timed steps, so
    # these run with the GPU and the problem's memory in the state of a long
    # run, not in the ramp after an idle spell.
    p32 = make_problem(ff, 32)
    if not p32.eqn.calcN.fused:
        raise RuntimeError("fused path unavailable for the benchmark grid")
    p32.sol.copy_(ff.vorticity_ic(p32.grid, seed=1))
    k64 = a.fp64_steps or max(1, a.steps // 2)
    p64 = make_problem(ff, 64)
    p64.sol.copy_(ff.vorticity_ic(p64.grid, seed=1))

    accuracy = accuracy_vs_oracle(ff, torch)
    print(f"# accuracy: {accuracy}", file=sys.stderr, flush=True)

    # FP32 (measured path)
    prof32 = kernel_profile(ff, p32, a.profile_steps)
    t32 = time_steps(ff, torch, p32, a.steps, a.warmup)
    sps32 = a.steps / t32
    print(f"# FP32: {sps32:.1f} steps/s ({1e3 * t32 / a.steps:.4f} ms/step)", file=sys.stderr, flush=True)

    # FP64 on the same problem
    prof64 = kernel_profile(ff, p64, a.profile_steps)
    t64 = time_steps(ff, torch, p64, k64, max(1, a.warmup // 2))
    sps64 = k64 / t64
    print(f"# FP64: {sps64:.1f} steps/s", file=sys.stderr, flush=True)

    # dominant kernel of the FP32 step: the one with the largest total time
    dom_name, dom = max(prof32.items(), key=lambda kv: kv[1]["total_ms"])
    alg = algorithmic_bytes(dom_name, N, 8)
    achieved = alg / (dom["avg_us"] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom_name),
                "kernel": dom_name, "avg_us": round(dom["avg_us"], 3), "algorithmic_bytes_per_launch": alg,
                "kept_band_bytes_per_launch": dom["bytes_per_launch"],
                "kept_band_GBps": round(dom["GBps"], 1)}
    step_bytes = sum(v["bytes_per_launch"] * v["count"] for v in prof32.values()) / a.profile_steps

    adj = None
    if not a.no_adjoint:
        adj = adjoint_timing(ff, torch)
        print(f"# adjoint: {adj}", file=sys.stderr, flush=True)

    ens = None
    if not a.no_ensemble:
        ens = ensemble_timing(ff, torch)
        print(f"# ensemble: {ens}", file=sys.stderr, flush=True)
    big = vort4096_timing(ff, torch)
    print(f"# 4096^2: {big}", file=sys.stderr, flush=True)
    fft = None
    if not a.no_fft:
        fft = fft_plan_timing(ff, torch)
        print(f"# fft plan: fp32 mul {fft['fp32']['mul']['GBps']} GB/s, ldiv {fft['fp32']['ldiv']['GBps']} GB/s; "
              f"fp64 mul {fft['fp64']['mul']['GBps']} GB/s, ldiv {fft['fp64']['ldiv']['GBps']} GB/s",
              file=sys.stderr, flush=True)
    trc = None
    if not a.no_tracer:
        trc = tracer_timing(ff, torch)
        print(f"# tracer: {trc}", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a.cpu_seconds)

    line = {
        "metric": METRIC, "value": round(sps32, 2), "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * t32 / a.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 ETDRK4 coefficients)",
        "data": "synthetic (seeded k^4 exp(-(k/6)^2) vorticity, rms 1)",
        "config": {"workload": "2D vorticity 2048^2 ETDRK4, nu8|k|^8 + mu=0.01, dt=1e-3, 2/3 dealiasing",
                   "nx": N, "ny": N, "stepper": "ETDRK4", "dt": DT, "parallelism": "single GPU"},
        "fp64_steps_per_s": round(sps64, 2), "fp32_over_fp64": round(sps32 / sps64, 3),
        "accuracy_vs_fp64_oracle": accuracy,
        "adjoint": adj,
        "ensemble": ens,
        "tracer": trc,
        "fft_plan": fft,
        "vorticity_4096": big,
        "step_kept_band_GB": round(step_bytes / 1e9, 4),
        "step_kept_band_GBps": round(step_bytes / (t32 / a.steps) / 1e9, 1),
        "step_algorithmic_GB": round(70 * (N // 2 + 1) * N * 8 / 1e9, 4),
        "step_algorithmic_GBps": round(70 * (N // 2 + 1) * N * 8 / (t32 / a.steps) / 1e9, 1),
        "roofline": roofline,
        "kernels_fp32": {k: {kk: round(vv, 3) if isinstance(vv, float) else vv for kk, vv in v.items()}
                         for k, v in prof32.items()},
        "kernels_fp64": {k: {kk: round(vv, 3) if isinstance(vv, float) else vv for kk, vv in v.items()}
                         for k, v in prof64.items()},
        "cpu_baseline": cpu,
        "reference_published_gpu_steps_per_s_2048": 18.3,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
