"""CPU-only checks of the C ABI and the host-side logic (no GPU calls).

* libffhip.so loads and exports every entry point declared in include/ffhip.h
* host numerics of the library (ETDRK4 contour coefficients, aliased index
  ranges) against the oracle
* argument validation returns the documented status codes
* the product's host-side grid and SplitMix64 stream match the oracle
"""

import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import fourierflows_hip as ffh
from fourierflows_hip import _lib
from oracle import grids as og
from oracle import ics as oics
from oracle import timesteppers as ots


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "ffhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ffhip_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = _header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the binding covers every declared entry point
    assert sorted(_lib.SIGNATURES) == syms


def test_version_and_error_string():
    lib = _lib.load()
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert lib.ffhip_version(C.byref(a), C.byref(b), C.byref(c)) == 0
    assert (a.value, b.value) == (0, 1)
    assert isinstance(_lib.last_error(), str)


@pytest.mark.parametrize("hL", [0.0, -1e-6, -0.3, -1.0, -7.5, -410.0, 2.0])
def test_etdcoeffs_match_oracle(hL):
    dt = 0.002
    L = np.array([hL / dt, 0.5 * hL / dt])
    got = ffh.etdcoeffs(dt, L)
    ref = ots.getetdcoeffs(dt, L)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g, r, rtol=1e-13, atol=0)


def test_etdcoeffs_complex_L():
    L = np.array([-3.0 + 40j, -0.1 - 2j])
    got = ffh.etdcoeffs(0.01, L)
    ref = ots.getetdcoeffs(0.01, L)
    for g, r in zip(got, ref):
        assert np.iscomplexobj(g)
        np.testing.assert_allclose(g, r, rtol=1e-13)


def test_etdcoeffs_radius_and_count():
    L = np.linspace(-300, 0, 7)
    got = ffh.etdcoeffs(0.01, L, ncirc=64, rcirc=2.0)
    ref = ots.getetdcoeffs(0.01, L, ncirc=64, rcirc=2.0)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(g, r, rtol=1e-13)


@pytest.mark.parametrize("f", [0, 1 / 3, 1 / 2, 1 / 4, 0.2])
@pytest.mark.parametrize("n", [6, 16, 32, 34, 128, 2048])
def test_aliased_wavenumbers_match_oracle(n, f):
    got = ffh.getaliasedwavenumbers(n, n // 2 + 1, f)
    ref = og.getaliasedwavenumbers(n, n // 2 + 1, f)
    assert got == ref


def test_aliased_fraction_one_rejected():
    with pytest.raises(_lib.FFHIPError) as e:
        ffh.getaliasedwavenumbers(16, 9, 1.0)
    assert e.value.code == _lib.ERR_ARG


def test_status_codes_without_gpu():
    lib = _lib.load()
    h = C.c_void_p()
    # odd nx for a real transform -> DomainError code, before any device work
    rc = lib.ffhip_plan_rfft(C.byref(h), 2, _lib.i64arr([5, 4]), 1, 64)
    assert rc == _lib.ERR_DOMAIN
    assert "even" in _lib.last_error()
    rc = lib.ffhip_plan_rfft(C.byref(h), 4, _lib.i64arr([4, 4, 4, 4]), 1, 64)
    assert rc == _lib.ERR_ARG
    rc = lib.ffhip_plan_rfft(C.byref(h), 2, _lib.i64arr([4, 4]), 1, 16)
    assert rc == _lib.ERR_ARG
    rc = lib.ffhip_plan_forward(None, None, None, None)
    assert rc == _lib.ERR_ARG
    rc = lib.ffhip_op_create(C.byref(h), 99, 64, 2, _lib.i64arr([8, 8]), None, 0.3, None, 0)
    assert rc == _lib.ERR_ARG
    box = (C.c_double * 2)(1.0, 1.0)
    rc = lib.ffhip_op_create(C.byref(h), _lib.OP_VORTICITY2D, 64, 2, _lib.i64arr([7, 8]), box, 0.3, None, 0)
    assert rc == _lib.ERR_DOMAIN


def test_python_odd_grid_raises_domainerror():
    with pytest.raises(ffh.DomainError):
        ffh.TwoDGrid(ffh.CPU(), nx=5, Lx=1.0, ny=4, Ly=2.0)
    with pytest.raises(ffh.DomainError):
        ffh.ThreeDGrid(ffh.CPU(), nx=4, Lx=1.0, ny=6, Ly=2.0, nz=5, Lz=3.0)


def test_host_grid_matches_oracle():
    g = ffh.TwoDGrid(ffh.CPU(), nx=32, Lx=2 * np.pi, ny=64, Ly=3 * np.pi)
    o = og.TwoDGrid(32, 2 * np.pi, 64, 3 * np.pi)
    np.testing.assert_array_equal(g.kr_host, o.kr)
    np.testing.assert_array_equal(g.l_host, o.l)
    np.testing.assert_array_equal(g.invKrsq_host, o.invKrsq)
    np.testing.assert_array_equal(g.x, o.x)
    assert (g.kalias, g.kralias, g.lalias) == (o.kalias, o.kralias, o.lalias)
    g3 = ffh.ThreeDGrid(ffh.CPU(), nx=16, Lx=2 * np.pi, ny=32, Ly=2 * np.pi, nz=34, Lz=2 * np.pi)
    o3 = og.ThreeDGrid(16, 2 * np.pi, 32, 2 * np.pi, 34, 2 * np.pi)
    assert (g3.lalias, g3.malias) == (o3.lalias, o3.malias)
    np.testing.assert_array_equal(g3.Krsq_host, o3.Krsq)


def test_splitmix_stream_matches_oracle():
    a = ffh.splitmix64(99, 1000, offset=5)
    b = oics.splitmix64(99, 1000, offset=5)
    assert np.array_equal(a, b)
    g = ffh.TwoDGrid(ffh.CPU(), nx=64, Lx=2 * np.pi, x0=0.0, y0=0.0)
    from fourierflows_hip.ics import vorticity_spectrum_amplitude

    amp = vorticity_spectrum_amplitude(g, 7)
    # the oracle IC starts from the same host array before its FFTs
    o = og.TwoDGrid(64, 2 * np.pi, x0=0, y0=0)
    for spec in ("k4exp1", "k4exp2"):
        ref = oics.vorticity_spectrum_amplitude(o, 7, 5.5, spec)
        assert np.array_equal(vorticity_spectrum_amplitude(g, 7, 5.5, spec), ref)
    assert np.array_equal(amp, oics.vorticity_spectrum_amplitude(o, 7))


def test_ensemble_members_match_oracle():
    assert ffh.ensemble_members() == oics.ensemble_members()
    m = ffh.ensemble_members(counts=((64, 3), (128, 2)), master_seed=5)
    assert m == oics.ensemble_members(counts=((64, 3), (128, 2)), master_seed=5)
    nus = np.array([x[2] for x in ffh.ensemble_members()])
    k0s = np.array([x[3] for x in ffh.ensemble_members()])
    assert nus.min() >= 1e-4 and nus.max() <= 1e-2 and k0s.min() >= 4 and k0s.max() <= 10
    assert [x[0] for x in ffh.ensemble_members()] == [128] * 32 + [256] * 24 + [512] * 8


def test_oracle_not_imported_by_product():
    pkg = os.path.join(ROOT, "fourierflows.jl_amd", "fourierflows_hip")
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            src = open(os.path.join(pkg, f)).read()
            assert not re.search(r"^\s*(import|from)\s+oracle\b", src, flags=re.M), f


def test_diagnostic_chunking_host():
    """stepforward(prob, diags, n) with a built-in operator runs the steps
    between diagnostic updates as single library calls, and the saved steps
    are those of the reference's step-by-step loop (src/timesteppers.jl:28-35,
    src/diagnostics.jl:86-104)."""
    import types

    from fourierflows_hip import diagnostics as fd
    from fourierflows_hip import timesteppers as fts

    calls = []

    class FakeTS:
        def run_fused(self, op, sol, clock, n):
            calls.append(n)
            clock.step += n
            clock.t += n * 0.5

    clock = types.SimpleNamespace(step=0, t=0.0)
    calcN = types.SimpleNamespace(builtin_operator=object())
    prob = types.SimpleNamespace(clock=clock, timestepper=FakeTS(), sol=None,
                                 eqn=types.SimpleNamespace(calcN=calcN))
    d3 = fd.Diagnostic(lambda p: p.clock.step, prob, nsteps=20, freq=3)
    d4 = fd.Diagnostic(lambda p: p.clock.t, prob, nsteps=20, freq=4)
    fts.stepforward(prob, [d3, d4], 20)
    assert sum(calls) == 20 and calls == [3, 1, 2, 2, 1, 3, 3, 1, 2, 2]
    assert list(d3["steps"]) == [0, 3, 6, 9, 12, 15, 18]
    assert list(d3["data"]) == [0, 3, 6, 9, 12, 15, 18]
    assert list(d4["steps"]) == [0, 4, 8, 12, 16, 20]
    assert list(d4["t"]) == [0.0, 2.0, 4.0, 6.0, 8.0, 10.0]
    # a second call continues from the clock's step
    calls.clear()
    fts.stepforward(prob, d3, 5)
    assert calls == [1, 3, 1] and list(d3["steps"])[-2:] == [21, 24]
    # the default storage (ceil((nsteps+1)/freq)) grows as the reference's extend! does
    assert len(d3.t) == 14 and d3.i == 9
