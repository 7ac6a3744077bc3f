"""Output tests (test/test_output.jl): path helpers and the archive layout on
the CPU; problems with device fields on the GPU."""

import os
import types

import numpy as np
import pytest


def _out():
    from fourierflows_hip import output

    return output


def test_withoutjld2():
    o = _out()
    assert o.withoutjld2("blahblah.jld2") == "blahblah"
    assert o.withoutjld2("blahblah") == "blahblah"
    assert o.withoutjld2("blahblah.npz") == "blahblah"


def test_uniquepath(tmp_path):
    o = _out()
    path, p1, p2 = (str(tmp_path / n) for n in ("foo.npz", "foo_1.npz", "foo_2.npz"))
    open(path, "w").close()
    assert o.uniquepath(path) == p1
    open(p1, "w").close()
    assert o.uniquepath(path) == p2


def test_archive_layout_and_append(tmp_path):
    """Group paths as keys; appending keeps earlier members; a second write of
    one key is refused (as for a JLD2 dataset)."""
    o = _out()
    path = str(tmp_path / "a.npz")
    o._write(path, [("snapshots/t/0", np.float64(0.5)), ("grid/x", np.arange(4.0))])
    o._write(path, [("snapshots/t/10", np.float64(1.5))])
    with np.load(path) as f:
        assert sorted(f.files) == ["grid/x", "snapshots/t/0", "snapshots/t/10"]
    d = o.load(path)
    assert d["snapshots"]["t"]["10"] == 1.5 and np.array_equal(d["grid"]["x"], np.arange(4.0))
    with pytest.raises(KeyError):
        o._write(path, [("snapshots/t/0", np.float64(2.0))])


def test_output_host_problem(tmp_path):
    """Output / saveoutput / saveproblem / savediagnostic on a problem whose
    fields are host arrays (the layout of src/output.jl:62-158)."""
    o = _out()
    clock = types.SimpleNamespace(dt=0.1, t=0.3, step=3)
    eqn = types.SimpleNamespace(L=-np.arange(5.0), dims=(5,), T=np.complex128)
    params = types.SimpleNamespace(kappa=0.01, func=np.sin)
    prob = types.SimpleNamespace(clock=clock, eqn=eqn, grid=object(), params=params, c=np.arange(3.0))
    out = o.Output(prob, str(tmp_path / "run.jld2"), ("c", lambda p: p.c))
    assert out.path.endswith("run.npz")
    assert np.array_equal(out["c"], np.arange(3.0))
    o.saveproblem(out)
    o.saveoutput(out)
    d = o.load(out.path)
    assert np.array_equal(d["eqn"]["L"], -np.arange(5.0)) and list(d["eqn"]["dims"]) == [5]
    assert d["clock"]["dt"] == 0.1 and d["params"]["kappa"] == 0.01 and "func" not in d["params"]
    assert d["snapshots"]["t"]["3"] == 0.3 and np.array_equal(d["snapshots"]["c"]["3"], np.arange(3.0))
    diag = types.SimpleNamespace(i=3, steps=np.array([0, 2, 4, 0]), t=np.array([0.0, 0.2, 0.4, 0.0]),
                                 data=[1.0, 1.0, 1.0, None])
    o.savediagnostic(diag, "mydiagnostic", out.path)
    d = o.load(out.path)
    assert list(d["diagnostics"]["mydiagnostic"]["steps"]) == [0, 2, 4]
    assert list(d["diagnostics"]["mydiagnostic"]["data"]) == [1.0, 1.0, 1.0]


def test_background_writes_interleaved_with_synchronous(tmp_path):
    """Background snapshots interleaved with saveproblem, savediagnostic and
    synchronous saveoutput on the same path: one valid archive holding every
    member; the writer is flushed before each synchronous write."""
    o = _out()
    clock = types.SimpleNamespace(dt=0.1, t=0.0, step=0)
    eqn = types.SimpleNamespace(L=-np.arange(5.0), dims=(5,), T=np.complex128)
    prob = types.SimpleNamespace(clock=clock, eqn=eqn, grid=object(), params=None, c=np.arange(4096.0))
    out = o.Output(prob, str(tmp_path / "mixed"), ("c", lambda p: p.c * (1 + p.clock.step)))
    o.saveproblem(out)
    for k in range(1, 25):
        clock.step, clock.t = k, 0.1 * k
        o.saveoutput(out, background=(k % 3 != 0))
        if k % 4 == 0:
            diag = types.SimpleNamespace(i=2, steps=np.array([0, k]), t=np.array([0.0, clock.t]), data=[1.0, 2.0])
            o.savediagnostic(diag, f"d{k}", out.path)
    out.close()
    d = o.load(out.path)
    assert sorted(int(s) for s in d["snapshots"]["c"]) == list(range(1, 25))
    for k in range(1, 25):
        assert np.array_equal(d["snapshots"]["c"][str(k)], np.arange(4096.0) * (1 + k))
        assert d["snapshots"]["t"][str(k)] == 0.1 * k
    assert sorted(d["diagnostics"]) == [f"d{k}" for k in sorted(range(4, 25, 4), key=str)]
    assert list(d["eqn"]["dims"]) == [5]


@pytest.mark.gpu
def test_saveproblem_saveoutput_gpu(ff, tmp_path):
    """test/test_output.jl:56-77 on the diffusion testbed with array kappa."""
    import torch

    nx = 32
    prob = ff.Diffusion.Problem(nx=nx, Lx=2 * np.pi, kappa=1e-2 * np.ones(nx), dt=1e-7, stepper="ForwardEuler")
    ctest = np.zeros(nx)
    ctest[2] = np.pi
    prob.vars.c.copy_(torch.as_tensor(ctest))
    out = ff.Output(prob, str(tmp_path / "testoutput.jld2"), ("c", lambda p: p.vars.c))
    ff.saveproblem(out)
    ff.saveoutput(out)
    d = ff.output.load(out.path)
    assert np.allclose(d["snapshots"]["c"]["0"], ctest, rtol=1e-12)
    assert np.isclose(d["grid"]["Lx"], prob.grid.Lx, rtol=1e-12)
    assert np.allclose(d["params"]["kappa"], 1e-2 * np.ones(nx))


@pytest.mark.gpu
def test_saveproblem_twodgrid_and_diagnostic_gpu(ff, tmp_path):
    """test/test_output.jl:79-130: a 2D grid's Ly and L; a diagnostic stepped
    past its storage saved up to its last entry."""
    prob = ff.TwoDVorticityProblem(nx=32, nu=1e-3, dt=0.01)
    E = ff.Diagnostic(ff.diagnostics.energy, prob, nsteps=100, freq=2)
    ff.stepforward(prob, E, 105)
    out = ff.Output(prob, str(tmp_path / "testoutput"), ("sol", lambda p: p.sol))
    ff.saveproblem(out)
    ff.savediagnostic(E, "energy", out.path)
    d = ff.output.load(out.path)
    assert np.isclose(d["grid"]["Ly"], prob.grid.Ly)
    assert np.allclose(d["eqn"]["L"], np.asarray(prob.eqn.L))
    assert list(d["diagnostics"]["energy"]["steps"]) == [0] + list(range(2, 106, 2))
    assert np.allclose(d["diagnostics"]["energy"]["data"], np.asarray(E["data"]))


@pytest.mark.gpu
def test_background_saveoutput_matches_synchronous(ff, tmp_path):
    """Snapshots taken with background=True (pinned-buffer copies on the stream,
    file writes on a writer thread) while the run continues equal the
    synchronous ones."""
    import torch

    def run(background, path):
        prob = ff.TwoDVorticityProblem(nx=64, nu=1e-3, dt=0.01, T=np.float32)
        prob.sol.copy_(ff.vorticity_ic(prob.grid, seed=3))
        out = ff.Output(prob, path, ("sol", lambda p: p.sol))
        for _ in range(5):
            ff.stepforward(prob, 7)
            ff.saveoutput(out, background=background)
        out.close()
        torch.cuda.synchronize()
        return ff.output.load(out.path)

    a = run(False, str(tmp_path / "sync"))
    b = run(True, str(tmp_path / "async"))
    assert set(a["snapshots"]["sol"]) == set(b["snapshots"]["sol"]) == {str(7 * k) for k in range(1, 6)}
    for k in a["snapshots"]["sol"]:
        assert np.array_equal(a["snapshots"]["sol"][k], b["snapshots"]["sol"][k])
        assert a["snapshots"]["t"][k] == b["snapshots"]["t"][k]
