"""Pins for the FP64 CPU oracle.

Every check here compares the oracle with something independent of it: the
reference's own analytic known answers (test/createffttestfunctions.jl,
test/test_timesteppers.jl, test/test_utils.jl, test/test_grid.jl), a
brute-force DFT, mpmath at 50 digits, or a closed-form solution.  None of
them re-types the oracle's formula.
"""

import math

import mpmath
import numpy as np
import pytest

from oracle import grids, ics, physics, timesteppers as ts, utils

RTOL_FFT = 1e-12          # test/runtests.jl:21
RTOL_UTILS = 1e-13        # test/runtests.jl:23
RTOL_TIMESTEPPERS = 1e-12  # test/runtests.jl:24


# --------------------------------------------------------------- FFT known answers
def _analytic_1d(g, phi=np.pi / 3, m=5):
    """test/createffttestfunctions.jl:1-33."""
    k0 = g.k[1]
    f = np.cos(m * k0 * g.x + phi)
    fhr = np.zeros(g.nkr, complex)
    for i in range(g.nkr):
        if abs(g.k[i]) == m * k0:
            fhr[i] = -np.exp(np.sign(g.kr[i]) * 1j * phi) * g.nx / 2
    return f, fhr


def test_fft_1d_known_answer():
    g = grids.OneDGrid(32, 2 * np.pi)
    f, fhr = _analytic_1d(g)
    assert utils.isapprox(g.mul(f), fhr, rtol=RTOL_FFT)
    assert utils.isapprox(g.ldiv(fhr.copy()), f, rtol=RTOL_FFT)


def _analytic_2d(g):
    """test/createffttestfunctions.jl:36-87 (numpy axis order [ny][nx])."""
    X, Y = g.x[None, :], g.y[:, None]
    m, n = 5, 2
    k0, l0 = g.k[0, 1], g.l[1, 0]
    f1 = np.cos(m * k0 * X) * np.cos(n * l0 * Y) + 0 * X * Y
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


def test_fft_2d_known_answer():
    g = grids.TwoDGrid(32, 2 * np.pi, 64, 3 * np.pi)
    f1, f2, f1hr, f2hr = _analytic_2d(g)
    assert utils.isapprox(g.mul(f1), f1hr, rtol=RTOL_FFT)
    assert utils.isapprox(g.mul(f2), f2hr, rtol=RTOL_FFT)
    assert utils.isapprox(g.ldiv(f1hr), f1, rtol=RTOL_FFT)
    assert utils.isapprox(g.ldiv(f2hr), f2, rtol=RTOL_FFT)


def test_fft_3d_known_answer():
    """test/createffttestfunctions.jl:90-139 on the 32x30x16 grid (ny=30: radix 3 and 5)."""
    g = grids.ThreeDGrid(32, 2 * np.pi, 30, 3 * np.pi, 16, 4.0)
    X, Y, Z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    mx, my, mz = 5, 2, 3
    k0, l0, m0 = g.k[0, 0, 1], g.l[0, 1, 0], g.m[1, 0, 0]
    f1 = np.cos(mx * k0 * X) * np.cos(my * l0 * Y) * np.cos(mz * m0 * Z)
    f2 = np.sin(mx * k0 * X + my * l0 * Y + mz * m0 * Z)
    f1hr = np.zeros(g.spec_shape, complex)
    f2hr = np.zeros(g.spec_shape, complex)
    N = g.nx * g.ny * g.nz
    for kk in range(g.nm):
        for j in range(g.nl):
            for i in range(g.nkr):
                kr, l, m = g.kr[0, 0, i], g.l[0, j, 0], g.m[kk, 0, 0]
                if abs(kr) == mx * k0 and abs(l) == my * l0 and abs(m) == mz * m0:
                    f1hr[kk, j, i] = N / 8
                if kr == mx * k0 and l == my * l0 and m == mz * m0:
                    f2hr[kk, j, i] = N / 2
    f2hr = -1j * f2hr
    assert utils.isapprox(g.mul(f1), f1hr, rtol=RTOL_FFT)
    assert utils.isapprox(g.mul(f2), f2hr, rtol=RTOL_FFT)
    assert utils.isapprox(g.ldiv(f2hr), f2, rtol=RTOL_FFT)


def _dft_matrix(n, sign=-1):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


@pytest.mark.parametrize("shape", [(6,), (8, 6), (4, 6, 10)])
def test_rfft_layout_vs_bruteforce(shape):
    """mul!/ldiv! axis conventions against an O(n^2) DFT per axis (x = last NumPy axis)."""
    rng = np.random.default_rng(3)
    nd = len(shape)
    ns = tuple(reversed(shape))
    if nd == 1:
        g = grids.OneDGrid(ns[0], 1.0)
    elif nd == 2:
        g = grids.TwoDGrid(ns[0], 1.0, ns[1], 2.0)
    else:
        g = grids.ThreeDGrid(ns[0], 1.0, ns[1], 2.0, ns[2], 3.0)
    f = rng.standard_normal(shape)
    F = f.astype(complex)
    for ax in range(nd):
        F = np.moveaxis(np.tensordot(_dft_matrix(shape[ax]), np.moveaxis(F, ax, 0), axes=1), 0, ax)
    expect = F[..., : shape[-1] // 2 + 1]
    assert utils.isapprox(g.mul(f), expect, rtol=1e-13)
    assert utils.isapprox(g.ldiv(expect), f, rtol=1e-13)


def test_c2r_ignores_imaginary_dc_nyquist_after_y_pass():
    """c2r semantics on a non-Hermitian kx=0 column: y-c2c first, then x-c2r
    keeping Re at kx=0 and kx=nx/2 (equivalently: Hermitian part of the column)."""
    g = grids.TwoDGrid(8, 2 * np.pi, 6, 2 * np.pi)
    rng = np.random.default_rng(1)
    fh = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    out = g.ldiv(fh)
    col = np.fft.ifft(fh, axis=0)           # y inverse (1/ny)
    col[:, 0] = col[:, 0].real
    col[:, -1] = col[:, -1].real
    ref = np.zeros((6, 8))
    for j in range(6):
        c = col[j]
        full = np.concatenate([c, np.conj(c[-2:0:-1])])
        ref[j] = np.real(np.fft.ifft(full))
    assert utils.isapprox(out, ref, rtol=1e-13)


# --------------------------------------------------------------- dealias / grids
def test_aliased_fraction_ranges():
    """test/test_grid.jl:203-227 with nx,ny,nz = 16,32,34."""
    for f in (0, 1 / 3, 1 / 2, 1 / 4):
        nx, ny, nz = 16, 32, 34
        g2 = grids.TwoDGrid(nx, 2 * np.pi, ny, 2 * np.pi, aliased_fraction=f)
        g3 = grids.ThreeDGrid(nx, 2 * np.pi, ny, 2 * np.pi, nz, 2 * np.pi, aliased_fraction=f)
        lo = lambda n: math.floor((1 - f) / 2 * n) + 1  # noqa: E731
        hi = lambda n: math.ceil((1 + f) / 2 * n)  # noqa: E731
        if f == 0:
            assert g2.kralias is None and g2.lalias is None and g3.malias is None
            continue
        assert g2.kralias == (lo(nx), nx // 2 + 1)
        assert g2.kalias == (lo(nx), hi(nx))
        assert g2.lalias == (lo(ny), hi(ny))
        assert g3.malias == (lo(nz), hi(nz))


@pytest.mark.parametrize("n,kmax,lrange", [(128, 41, (-42, 41)), (256, 84, (-85, 84)),
                                           (512, 169, (-170, 169)), (2048, 681, (-682, 681))])
def test_dealias_kept_band(n, kmax, lrange):
    """Kept |kx| and l from the index formula (src/domains.jl:408-421)."""
    g = grids.TwoDGrid(n, 2 * np.pi)
    m = grids.kept_mask(g)
    kr = np.rint(g.kr[0]).astype(int)
    l = np.rint(g.l[:, 0]).astype(int)
    kept_k = kr[m.any(axis=0)]
    kept_l = l[m.any(axis=1)]
    assert kept_k.max() == kmax and kept_k.min() == 0
    assert (kept_l.min(), kept_l.max()) == lrange


@pytest.mark.parametrize("shape", [(6, 8), (64, 32), (30, 48)])
def test_dealias_predicate(shape):
    """test/test_grid.jl:95-133: kept exactly when kr < kmax and -lmax <= l < lmax."""
    nx, ny = shape
    g = grids.TwoDGrid(nx, 2 * np.pi, ny, 4 * np.pi)
    fh = np.ones(g.spec_shape, complex)
    grids.dealias(fh, g)
    kmax = round(g.kr.max() * 2 / 3)
    lmax = round(np.abs(g.l).max() * 2 / 3)
    pred = (g.kr < kmax) & (g.l < lmax) & (g.l >= -lmax)
    assert np.all((fh != 0) == pred)


def test_odd_size_raises():
    with pytest.raises(grids.DomainError):
        grids.TwoDGrid(5, 1.0, 4, 2.0)
    with pytest.raises(grids.DomainError):
        grids.ThreeDGrid(4, 1.0, 6, 2.0, 5, 3.0)


def test_wavenumber_ordering():
    """test/test_grid.jl:40-59: k = [0:n/2-1, -n/2:-1], kr Nyquist positive."""
    g = grids.TwoDGrid(6, 2 * np.pi, 8, 4 * np.pi)
    assert np.allclose(g.k[0], [0, 1, 2, -3, -2, -1])
    assert np.allclose(g.kr[0], [0, 1, 2, 3])
    assert np.allclose(g.l[:, 0], 0.5 * np.array([0, 1, 2, 3, -4, -3, -2, -1]))
    assert g.invKrsq[0, 0] == 0


# --------------------------------------------------------------- ETDRK4 coefficients
mpmath.mp.dps = 50


def _phi_mp(z):
    """Exact ETDRK4 coefficient functions (Kassam & Trefethen 2005, eq. 2.5), in mpmath."""
    z = mpmath.mpf(z)
    if z == 0:
        return mpmath.mpf(1) / 2, mpmath.mpf(1) / 6, mpmath.mpf(1) / 6, mpmath.mpf(1) / 6
    ez = mpmath.exp(z)
    zeta = (mpmath.exp(z / 2) - 1) / z
    alpha = (-4 - z + ez * (4 - 3 * z + z**2)) / z**3
    beta = (2 + z + ez * (z - 2)) / z**3
    gamma = (-4 - 3 * z - z**2 + ez * (4 - z)) / z**3
    return zeta, alpha, beta, gamma


@pytest.mark.parametrize("hL", [0.0, -1e-8, -5e-6, -1e-3, -0.5, -0.9, -1.0, -5.0, -40.0,
                                -410.0, 0.7])
def test_etd_coefficients_vs_mpmath(hL):
    dt = 0.005
    zeta, alpha, beta, gamma = ts.getetdcoeffs(dt, np.array([hL / dt]))
    ref = [float(dt * c) for c in _phi_mp(hL)]
    got = [zeta[0], alpha[0], beta[0], gamma[0]]
    # With the reference's unit circle (rcirc=1) one contour node passes close to
    # z = 0 when |hL| is near 1, where the integrands cancel: the method itself is
    # then accurate to ~3e-13 relative.  Elsewhere it is at FP64 roundoff.
    tol = 1e-12 if 0.4 <= abs(hL) <= 1.6 else 2e-15
    for g_, r_ in zip(got, ref):
        assert abs(g_ - r_) <= tol * abs(r_), (hL, g_, r_)


def test_etd_coefficient_limits():
    zeta, alpha, beta, gamma = ts.getetdcoeffs(1e-3, np.array([0.0]))
    assert np.isclose(zeta[0], 1e-3 / 2, rtol=1e-14)
    for c in (alpha, beta, gamma):
        assert np.isclose(c[0], 1e-3 / 6, rtol=1e-14)


def test_etd_coefficients_complex_L():
    L = np.array([-3.0 + 40j])
    zeta, *_ = ts.getetdcoeffs(0.01, L)
    z = mpmath.mpc(-0.03, 0.4)
    ref = complex(0.01 * (mpmath.exp(z / 2) - 1) / z)
    assert np.iscomplexobj(zeta) and abs(zeta[0] - ref) < 1e-15


# --------------------------------------------------------------- time steppers
def _gaussian(x, t, c0=0.01, sigma=0.2, kappa=1e-2):
    """test/test_timesteppers.jl:1."""
    return c0 * sigma / np.sqrt(sigma**2 + 2 * kappa * t) * np.exp(-(x**2) / (2 * (sigma**2 + 2 * kappa * t)))


@pytest.mark.parametrize("stepper", ["ETDRK4", "RK4"])
@pytest.mark.parametrize("varying", [False, True])
def test_diffusion_gaussian(stepper, varying):
    """test/test_timesteppers.jl:37-59: nx=128, κ=1e-2, dt=1e-7, 1000 steps, rtol=1000e-12."""
    kappa = 1e-2
    dt = 1e-9 / kappa
    kap = kappa * np.ones(128) if varying else kappa
    prob = physics.diffusion_problem(nx=128, Lx=2 * np.pi, kappa=kap, dt=dt, stepper=stepper)
    physics.diffusion_set_c(prob, _gaussian(prob.grid.x, 0))
    prob.stepforward(1000)
    physics.diffusion_updatevars(prob)
    c_final = _gaussian(prob.grid.x, 1000 * dt)
    assert utils.isapprox(c_final, prob.vars.c, rtol=prob.clock.step * RTOL_TIMESTEPPERS)


def test_step_until_rejects_etdrk4():
    prob = physics.diffusion_problem(kappa=1e-2, dt=1e-7, stepper="ETDRK4")
    with pytest.raises(RuntimeError):
        ts.step_until(prob, 1e-4)


def test_etdrk4_exact_for_linear():
    """With N = 0, ETDRK4 reproduces exp(L t) to roundoff."""
    g = grids.TwoDGrid(32, 2 * np.pi)
    L = physics.vorticity_linear(g, nu=0.05)
    zero = lambda N, *a: N.__setitem__(Ellipsis, 0)  # noqa: E731
    eqn = ts.Equation(L, zero, g)
    p = ts.Problem(eqn, "ETDRK4", 0.01, g)
    rng = np.random.default_rng(0)
    u0 = rng.standard_normal(g.spec_shape) + 1j * rng.standard_normal(g.spec_shape)
    p.sol[...] = u0
    p.stepforward(100)
    assert utils.isapprox(p.sol, np.exp(L * 1.0) * u0, rtol=1e-13)


def test_single_shell_vorticity_is_steady_nonlinearly():
    """A single |k| shell has J(ψ, ζ) = 0 so ζ(t) = ζ0 exp(L t) exactly."""
    g = grids.TwoDGrid(64, 2 * np.pi, x0=0, y0=0)
    X, Y = np.meshgrid(g.x, g.y)
    z0 = np.cos(3 * X + 4 * Y) + 0.7 * np.sin(5 * X) - 0.3 * np.cos(4 * X - 3 * Y)  # |k| = 5
    L = physics.vorticity_linear(g, nu=1e-2)
    eqn = ts.Equation(L, physics.vorticity_calcN, g)
    p = ts.Problem(eqn, "ETDRK4", 0.01, g)
    p.sol[...] = g.mul(z0)
    N = np.zeros_like(p.sol)
    physics.vorticity_calcN(N, p.sol, 0, None, None, None, g)
    assert np.abs(N).max() < 1e-10 * np.abs(p.sol).max()
    p.stepforward(50)
    assert utils.isapprox(g.ldiv(p.sol), z0 * np.exp(-1e-2 * 25 * 0.5), rtol=1e-12)


def test_etdrk4_fourth_order_on_vorticity():
    """Self-convergence slope of ETDRK4 on the nonlinear 2D vorticity problem is ~4."""
    g = grids.TwoDGrid(32, 2 * np.pi, x0=0, y0=0)
    L = physics.vorticity_linear(g, nu=1e-2)
    eqn = ts.Equation(L, physics.vorticity_calcN, g)
    zh0 = ics.vorticity_ic(g, seed=5, k0=3.0)

    def run(dt, n):
        p = ts.Problem(eqn, "ETDRK4", dt, g)
        p.sol[...] = zh0
        p.stepforward(n)
        return p.sol

    T = 0.8
    ref = run(T / 256, 256)
    errs = [utils.rel_l2(run(T / n, n), ref) for n in (8, 16, 32)]
    slopes = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.6 < s < 4.6 for s in slopes), (errs, slopes)


def test_rk4_fourth_order_on_tracer():
    """Self-convergence slope of RK4 on the 3D tracer problem is ~4."""
    g = grids.ThreeDGrid(16, 2 * np.pi, x0=0, y0=0, z0=0)
    params = physics.TracerParams(kappa=1e-2)
    vars = physics.TracerVars(g, params)
    eqn = ts.Equation(physics.tracer_linear(g, params.kappa), physics.tracer_calcN, g)
    c0 = grids.dealias(g.mul(ics.tracer_ic(g, seed=3, ngauss=2, sigma=0.8)), g)

    def run(dt, n):
        p = ts.Problem(eqn, "RK4", dt, g, vars, params)
        p.sol[...] = c0
        p.stepforward(n)
        return p.sol

    T = 0.6
    ref = run(T / 192, 192)
    errs = [utils.rel_l2(run(T / n, n), ref) for n in (6, 12, 24)]
    slopes = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(3.6 < s < 4.6 for s in slopes), (errs, slopes)


# --------------------------------------------------------------- nonlinear terms
def test_vorticity_N_equals_minus_dealiased_jacobian():
    """N = -dealias(J(ψ, ζ)) with J from the reference's jacobianh (src/utils.jl:190-205)."""
    g = grids.TwoDGrid(64, 2 * np.pi, 64, 2 * np.pi)
    zh = ics.vorticity_ic(g, seed=11, k0=8.0)
    N = np.zeros_like(zh)
    physics.vorticity_calcN(N, zh, 0, None, None, None, g)
    psi = g.ldiv(-g.invKrsq * zh)
    zeta = g.ldiv(zh)
    J = grids.dealias(utils.jacobianh(psi, zeta, g), g)
    assert utils.isapprox(N, -J, rtol=1e-12)


def test_tracer_N_closed_form():
    """c = sin x: N = -u·∇c = -(A sin z + C cos y) cos x (∇·u = 0)."""
    g = grids.ThreeDGrid(16, 2 * np.pi, x0=0, y0=0, z0=0)
    params = physics.TracerParams(A=1.0, B=0.5, C=0.25)
    vars = physics.TracerVars(g, params)
    x, y, z = g.x[None, None, :], g.y[None, :, None], g.z[:, None, None]
    c = np.sin(x) + 0 * y + 0 * z
    N = np.zeros(g.spec_shape, complex)
    physics.tracer_calcN(N, g.mul(c), 0, None, vars, params, g)
    expect = -(params.A * np.sin(z) + params.C * np.cos(y)) * np.cos(x)
    assert utils.isapprox(g.ldiv(N), expect, rtol=1e-12)


# --------------------------------------------------------------- reductions
def test_jacobian_known_answers():
    """test/runtests.jl:256-280 on 64x128, Lx=2π, Ly=3π."""
    g = grids.TwoDGrid(64, 2 * np.pi, 128, 3 * np.pi)
    X, Y = g.x[None, :], g.y[:, None]
    k0, l0 = 2 * np.pi / g.Lx, 2 * np.pi / g.Ly
    k1, l1, k2, l2 = 2 * k0, 6 * l0, 3 * k0, -3 * l0
    s1, s2 = np.sin(k1 * X + l1 * Y), np.sin(k2 * X + l2 * Y)
    J = (k1 * l2 - k2 * l1) * np.cos(k1 * X + l1 * Y) * np.cos(k2 * X + l2 * Y)
    atol = g.nx * g.ny * 10 * np.finfo(float).eps
    assert utils.isapprox(utils.jacobian(s1, s2, g), J, atol=atol)
    assert utils.isapprox(utils.jacobian(s1, s1, g), 0 * s1, atol=atol)
    e1 = np.exp(1j * (k1 * X + l1 * Y))
    e2 = np.exp(1j * (k2 * X + l2 * Y))
    Je = (k2 * l1 - k1 * l2) * np.exp(1j * ((k1 + k2) * X + (l1 + l2) * Y))
    assert utils.isapprox(utils.jacobian(e1, e2, g), Je, atol=atol)


def test_parseval_sums():
    """test/test_utils.jl:82-96: ∫f² = parsevalsum2(f̂) = parsevalsum(|f̂|²)."""
    g = grids.TwoDGrid(64, 2 * np.pi, 128, 3 * np.pi)
    X, Y = g.x[None, :], g.y[:, None]
    rng = np.random.default_rng(7)
    for f in (np.exp(-(X**2 + Y**2) / (2 * 0.5**2)), rng.standard_normal(g.phys_shape)):
        integral = np.sum(f**2) * g.dx * g.dy
        fh = g.mul(f)
        assert abs(utils.parsevalsum2(fh, g) - integral) <= RTOL_UTILS * integral
        assert abs(utils.parsevalsum(np.abs(fh) ** 2, g) - integral) <= RTOL_UTILS * integral
        fhc = g.fft(f)
        assert abs(utils.parsevalsum2(fhc, g) - integral) <= RTOL_UTILS * integral
    g1 = grids.OneDGrid(64, 2 * np.pi)
    f = np.exp(-(g1.x**2) / (2 * 0.5**2))
    assert abs(utils.parsevalsum2(g1.mul(f), g1) - np.sum(f**2) * g1.dx) <= RTOL_UTILS * np.sum(f**2) * g1.dx


def test_enstrophy_energy_of_single_mode():
    """ζ = cos(3x): enstrophy ½⟨ζ²⟩ = 1/4, energy ½⟨|u|²⟩ = 1/(4·9)."""
    g = grids.TwoDGrid(32, 2 * np.pi)
    zh = g.mul(np.cos(3 * g.x[None, :]) + 0 * g.y[:, None])
    assert abs(utils.enstrophy(zh, g) - 0.25) < 1e-15
    assert abs(utils.energy(zh, g) - 0.25 / 9) < 1e-15


# --------------------------------------------------------------- ICs
def test_splitmix64_known_values():
    """SplitMix64 from seed 0, scalar restatement in Python ints."""
    def scalar(seed, n):
        out, s = [], seed
        for _ in range(n):
            s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
            out.append(z ^ (z >> 31))
        return out
    v = ics.splitmix64(0, 4)
    assert [int(a) for a in v] == scalar(0, 4)
    assert int(v[0]) == 0xE220A8397B1DCDAF
    assert [int(a) for a in ics.splitmix64(12345, 3, offset=7)] == scalar(12345, 10)[7:]


def test_vorticity_ic_properties():
    g = grids.TwoDGrid(128, 2 * np.pi, x0=0, y0=0)
    zh = ics.vorticity_ic(g, seed=1)
    z = g.ldiv(zh)
    assert abs(np.sqrt(np.mean(z**2)) - 1.0) < 1e-3
    assert np.all(zh[~grids.kept_mask(g)] == 0)
    assert zh[0, 0] == 0


def _shell_fit(g, zh):
    """Least-squares fit log|zeta^| = a + b log K + c K^2 over the kept modes of
    the interior kx columns (the irfft/rfft round trip keeps them exactly)."""
    K = np.sqrt(g.Krsq)
    m = grids.kept_mask(g).copy()
    m[:, 0] = False
    m[:, -1] = False
    m &= K > 0
    m &= np.abs(zh) > 1e-6 * np.abs(zh).max()  # above the round trip's roundoff
    A = np.stack([np.ones(m.sum()), np.log(K[m]), K[m] ** 2], axis=1)
    coef, res, *_ = np.linalg.lstsq(A, np.log(np.abs(zh[m])), rcond=None)
    resid = np.log(np.abs(zh[m])) - A @ coef
    return coef, float(np.max(np.abs(resid)))


@pytest.mark.parametrize("spectrum,k0", [("k4exp1", 6.0), ("k4exp2", 7.3)])
def test_vorticity_ic_shell_spectrum(spectrum, k0):
    """|zeta^|^2 = K E(K) up to one constant: E = K^4 exp(-(K/k0)^2) gives
    |zeta^| ~ K^(5/2) exp(-K^2/(2 k0^2)), E = K^4 exp(-2 (K/k0)^2) gives
    exp(-K^2/k0^2).  sqrt(E) in place of sqrt(K E) would fit b = 2, and the
    other exponential twice (or half) the c; the residual shows the phases
    carry no amplitude."""
    g = grids.TwoDGrid(128, 2 * np.pi, x0=0, y0=0)
    zh = ics.vorticity_ic(g, seed=3, k0=k0, spectrum=spectrum)
    (a, b, c), resid = _shell_fit(g, zh)
    assert resid < 1e-8, resid
    assert abs(b - 2.5) < 1e-8, b
    want_c = -1 / (2 * k0**2) if spectrum == "k4exp1" else -1 / k0**2
    assert abs(c - want_c) < 1e-8 * abs(want_c), (c, want_c)


@pytest.mark.parametrize("n,sigma", [(32, 0.3), (48, 0.45)])
def test_tracer_ic_single_gaussian(n, sigma):
    """One Gaussian: the field is a rank-1 product of 1D profiles whose log has
    second differences -h^2/sigma^2 (pins the width), and its integral is
    (2 pi)^(3/2) sigma^3 (pins the unit peak amplitude)."""
    g = grids.ThreeDGrid(n, 2 * np.pi, x0=0, y0=0, z0=0)
    c = ics.tracer_ic(g, seed=11, ngauss=1, sigma=sigma)
    h = 2 * np.pi / n
    integral = c.sum() * h**3
    assert abs(integral - (2 * np.pi) ** 1.5 * sigma**3) < 1e-10 * integral
    iz, iy, ix = np.unravel_index(np.argmax(c), c.shape)
    prof = c[iz, iy, :]
    # rank 1: every x line is a multiple of the line through the maximum
    assert np.allclose(c[iz + 1, iy - 1, :] / c[iz + 1, iy - 1, ix], prof / prof[ix], rtol=1e-12)
    # within half a box of the centre the log is an exact parabola
    j = (ix + np.arange(-n // 4, n // 4 + 1)) % n
    d2 = np.diff(np.log(prof[j]), 2)
    assert np.allclose(d2, -(h**2) / sigma**2, rtol=1e-9)


def test_tracer_ic_sum_integral():
    """16 Gaussians of width 0.3: integral 16 (2 pi)^(3/2) sigma^3 and a maximum
    of at least one (each bump peaks at 1 at its centre; the grid misses the
    centres by at most half a cell)."""
    g = grids.ThreeDGrid(40, 2 * np.pi, x0=0, y0=0, z0=0)
    sigma = 0.3
    c = ics.tracer_ic(g, seed=5, ngauss=16, sigma=sigma)
    h = 2 * np.pi / 40
    integral = c.sum() * h**3
    assert abs(integral - 16 * (2 * np.pi) ** 1.5 * sigma**3) < 1e-10 * integral
    assert c.max() > np.exp(-3 * (h / 2) ** 2 / (2 * sigma**2)) - 1e-12
    assert c.min() >= 0
