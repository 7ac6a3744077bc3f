"""GPU parity of the libffhip FFT plans (mul!/ldiv! on rfftplan and fftplan).

Known answers from the reference's own tests (test/createffttestfunctions.jl,
test/test_fft.jl, test/test_ifft.jl, rtol 1e-12 as test/runtests.jl:21) and
seeded random fields against the oracle (pocketfft FP64).
"""

import numpy as np
import pytest

from oracle import grids as og
from oracle import utils as ou

pytestmark = pytest.mark.gpu

RTOL_FFT = 1e-12


def _dev(x, dtype=None):
    import torch

    t = torch.as_tensor(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def _analytic_2d(g):
    X, Y = g.x[None, :], g.y[:, None]
    m, n = 5, 2
    k0, l0 = g.k[0, 1], g.l[1, 0]
    f1 = np.cos(m * k0 * X) * np.cos(n * l0 * Y)
    f2 = np.sin(m * k0 * X + n * l0 * Y)
    f1hr = np.zeros(g.spec_shape, complex)
    f2hr = np.zeros(g.spec_shape, complex)
    for j in range(g.nl):
        for i in range(g.nkr):
            kr, l = g.kr[0, i], g.l[j, 0]
            if abs(kr) == m * k0 and abs(l) == n * l0:
                f1hr[j, i] = -g.nx * g.ny / 4
            if kr == m * k0 and l == n * l0:
                f2hr[j, i] = -g.nx * g.ny / 2
    return f1, f2, f1hr, -1j * f2hr


def test_rfft_known_answers_1d(ff):
    """test/createffttestfunctions.jl:1-33 with nx=32 (mul! and ldiv!)."""
    g = ff.OneDGrid(nx=32, Lx=2 * np.pi)
    og1 = og.OneDGrid(32, 2 * np.pi)
    phi = np.pi / 3
    f = np.cos(5 * og1.k[1] * og1.x + phi)
    ref = np.zeros(g.nkr, complex)
    for i in range(g.nkr):
        if abs(og1.k[i]) == 5 * og1.k[1]:
            ref[i] = -np.exp(np.sign(og1.kr[i]) * 1j * phi) * 32 / 2
    fd = _dev(f)
    fh = ff.zeros(g.device, np.complex128, (g.nkr,))
    ff.mul(fh, g.rfftplan, fd)
    assert ou.isapprox(_host(fh), ref, rtol=RTOL_FFT)
    back = ff.zeros(g.device, np.float64, (32,))
    ff.ldiv(back, g.rfftplan, fh)
    assert ou.isapprox(_host(back), f, rtol=RTOL_FFT)


def test_rfft_known_answers_2d(ff):
    """test/createffttestfunctions.jl:36-87 on 32x64, Lx=2π, Ly=3π."""
    g = ff.TwoDGrid(nx=32, Lx=2 * np.pi, ny=64, Ly=3 * np.pi)
    o = og.TwoDGrid(32, 2 * np.pi, 64, 3 * np.pi)
    f1, f2, f1hr, f2hr = _analytic_2d(o)
    for f, ref in ((f1, f1hr), (f2, f2hr)):
        fh = ff.zeros(g.device, np.complex128, g.spec_dims)
        ff.mul(fh, g.rfftplan, _dev(f))
        assert ou.isapprox(_host(fh), ref, rtol=RTOL_FFT)
        back = ff.zeros(g.device, np.float64, g.ns)
        ff.ldiv(back, g.rfftplan, fh)
        assert ou.isapprox(_host(back), f, rtol=RTOL_FFT)
        # complex plan (grid.fftplan)
        c = ff.zeros(g.device, np.complex128, g.ns)
        ff.mul(c, g.fftplan, _dev(f.astype(complex)))
        assert ou.isapprox(_host(c), o.fft(f), rtol=RTOL_FFT)


def test_rfft_known_answers_3d_mixed_radix(ff):
    """test/createffttestfunctions.jl:90-139 on 32x30x16 (radix 3 and 5 along y)."""
    g = ff.ThreeDGrid(nx=32, Lx=2 * np.pi, ny=30, Ly=3 * np.pi, nz=16, Lz=4.0)
    o = og.ThreeDGrid(32, 2 * np.pi, 30, 3 * np.pi, 16, 4.0)
    X, Y, Z = o.x[None, None, :], o.y[None, :, None], o.z[:, None, None]
    k0, l0, m0 = o.k[0, 0, 1], o.l[0, 1, 0], o.m[1, 0, 0]
    f2 = np.sin(5 * k0 * X + 2 * l0 * Y + 3 * m0 * Z)
    ref = np.zeros(o.spec_shape, complex)
    N = 32 * 30 * 16
    for kk in range(o.nm):
        for j in range(o.nl):
            for i in range(o.nkr):
                if o.kr[0, 0, i] == 5 * k0 and o.l[0, j, 0] == 2 * l0 and o.m[kk, 0, 0] == 3 * m0:
                    ref[kk, j, i] = N / 2
    ref = -1j * ref
    fh = ff.zeros(g.device, np.complex128, g.spec_dims)
    ff.mul(fh, g.rfftplan, _dev(f2))
    assert ou.isapprox(_host(fh), ref, rtol=RTOL_FFT)
    back = ff.zeros(g.device, np.float64, g.ns)
    ff.ldiv(back, g.rfftplan, fh)
    assert ou.isapprox(_host(back), f2, rtol=RTOL_FFT)


# (ny, nx) / (nz, ny, nx): power-of-two rows with nx in 512..2048 and columns
# of 64..2048 run the fast plan kernels, the rest the generic ones
SHAPES = [(8,), (6,), (34,), (2048,), (6, 8), (64, 64), (30, 32), (34, 34), (18, 22),
          (256, 128), (2048, 2048), (4096, 8), (16, 30, 32), (8, 10, 6), (64, 64, 64), (8, 8, 256),
          (512, 512), (1024, 1024), (128, 1024), (2048, 64), (32, 64, 512), (6, 512)]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", SHAPES)
def test_rfft_random_vs_oracle(ff, shape, prec):
    """Seeded random fields: forward vs pocketfft, inverse round trip, input preserved."""
    nd = len(shape)
    ns = tuple(reversed(shape))
    T = np.float64 if prec == 64 else np.float32
    Ls = (2 * np.pi, 3.0, 5.0)[:nd]
    G = (ff.OneDGrid, ff.TwoDGrid, ff.ThreeDGrid)[nd - 1]
    kw = dict(nx=ns[0], Lx=Ls[0], T=T)
    if nd >= 2:
        kw.update(ny=ns[1], Ly=Ls[1])
    if nd >= 3:
        kw.update(nz=ns[2], Lz=Ls[2])
    g = G(**kw)
    rng = np.random.default_rng(hash(shape) % 2**32)
    f = rng.standard_normal(shape).astype(T)
    ref = np.fft.rfftn(f.astype(np.float64))
    fd = _dev(f)
    fh = ff.zeros(g.device, np.result_type(T, np.complex64), g.spec_dims)
    ff.mul(fh, g.rfftplan, fd)
    tol = 1e-13 if prec == 64 else 2e-6
    assert ou.rel_l2(_host(fh), ref) < tol
    # ldiv! keeps its input and inverts
    keep = _host(fh).copy()
    back = ff.zeros(g.device, T, g.ns)
    ff.ldiv(back, g.rfftplan, fh)
    assert np.array_equal(_host(fh), keep)
    assert ou.rel_l2(_host(back), f) < tol
    # non-Hermitian input: inverse equals pocketfft's irfftn semantics
    h = (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)).astype(np.complex128)
    hd = _dev(h.astype(np.result_type(T, np.complex64)))
    ff.ldiv(back, g.rfftplan, hd)
    assert ou.rel_l2(_host(back), np.fft.irfftn(h, s=shape, axes=tuple(range(h.ndim - len(shape), h.ndim)))) < tol


@pytest.mark.parametrize("shape", [(6, 8), (30, 32), (64, 64, 64)])
def test_fft_c2c_random_vs_oracle(ff, shape):
    nd = len(shape)
    ns = tuple(reversed(shape))
    G = (ff.OneDGrid, ff.TwoDGrid, ff.ThreeDGrid)[nd - 1]
    kw = dict(nx=ns[0], Lx=1.0)
    if nd >= 2:
        kw.update(ny=ns[1], Ly=1.0)
    if nd >= 3:
        kw.update(nz=ns[2], Lz=1.0)
    g = G(**kw)
    rng = np.random.default_rng(1)
    f = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    out = ff.zeros(g.device, np.complex128, g.ns)
    ff.mul(out, g.fftplan, _dev(f))
    assert ou.rel_l2(_host(out), np.fft.fftn(f)) < 1e-13
    back = ff.zeros(g.device, np.complex128, g.ns)
    ff.ldiv(back, g.fftplan, out)
    assert ou.rel_l2(_host(back), f) < 1e-13


@pytest.mark.parametrize("shape", [(6, 8), (34, 34), (64, 128), (8, 10, 6)])
def test_global_memory_path(ff, shape, monkeypatch):
    """Plans forced onto the one-pass-per-radix global-memory path."""
    monkeypatch.setenv("FFHIP_FORCE_GLOBAL_FFT", "1")
    nd = len(shape)
    ns = tuple(reversed(shape))
    plan = ff.Plan(ns, np.float64, "rfft")
    rng = np.random.default_rng(5)
    f = rng.standard_normal(shape)
    fh = _dev(np.zeros(np.fft.rfftn(f).shape, complex))
    plan.forward(fh, _dev(f))
    assert ou.rel_l2(_host(fh), np.fft.rfftn(f)) < 1e-13
    back = _dev(np.zeros(shape))
    plan.inverse(back, fh)
    assert ou.rel_l2(_host(back), f) < 1e-13
    cplan = ff.Plan(ns, np.float64, "fft")
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    zo = _dev(np.zeros(shape, complex))
    cplan.forward(zo, _dev(z))
    assert ou.rel_l2(_host(zo), np.fft.fftn(z)) < 1e-13


@pytest.mark.parametrize("prec", [64, 32])
def test_fast_kernels_match_generic(ff, prec, monkeypatch):
    """The power-of-two kernels and the generic Stockham kernels give the same
    transforms (to roundoff) on a batched 2D plan and a 3D plan."""
    import torch

    T = np.float64 if prec == 64 else np.float32
    tol = 1e-13 if prec == 64 else 2e-6
    rng = np.random.default_rng(21)
    for ns, batch in (((1024, 256), 3), ((512, 128, 64), 1)):
        shape = (batch,) + tuple(reversed(ns))
        f = _dev(rng.standard_normal(shape).astype(T))
        outs = []
        for fast in ("1", "0"):
            monkeypatch.setenv("FFHIP_FFT_FAST", fast)
            plan = ff.Plan(ns, T, "rfft", batch=batch)
            monkeypatch.delenv("FFHIP_FFT_FAST")
            fh = torch.zeros(shape[:-1] + (ns[0] // 2 + 1,), dtype=torch.complex128 if prec == 64 else
                             torch.complex64, device="cuda")
            plan.forward(fh, f)
            back = torch.zeros_like(f)
            plan.inverse(back, fh)
            outs.append((_host(fh), _host(back)))
        assert ou.rel_l2(outs[0][0], outs[1][0]) < tol
        assert ou.rel_l2(outs[0][1], outs[1][1]) < tol


def test_batched_plan(ff):
    rng = np.random.default_rng(9)
    f = rng.standard_normal((5, 16, 12))  # 5 transforms of (ny=16, nx=12)
    plan = ff.Plan((12, 16), np.float64, "rfft", batch=5)
    fh = _dev(np.zeros((5, 16, 7), complex))
    plan.forward(fh, _dev(f))
    assert ou.rel_l2(_host(fh), np.fft.rfftn(f, axes=(1, 2))) < 1e-13
    back = _dev(np.zeros_like(f))
    plan.inverse(back, fh)
    assert ou.rel_l2(_host(back), f) < 1e-13
