"""Pins of the oracle that do not depend on the reference: closed forms, invariants,
special cases that reduce to textbook formulas, brute force on tiny inputs.  Each is
chosen so that a plausible slip in the restatement (a dropped term, a wrong sign, a
swapped index or direction) fails at least one of them."""
from __future__ import annotations

import numpy as np
import pytest

import pyoracle as po

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
MODELS = ["bgk", "trt", "rr"]


def moments(f):
    """f: (..., 9) -> rho, jx, jy, Pxx, Pyy, Pxy"""
    return (f.sum(-1), (f * CX).sum(-1), (f * CY).sum(-1), (f * CX * CX).sum(-1),
            (f * CY * CY).sum(-1), (f * CX * CY).sum(-1))


def random_state(nx, ny, seed, amp=0.05, ld=None):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.02 * rng.standard_normal((ny, nx))
    ux, uy = amp * rng.standard_normal((ny, nx)), amp * rng.standard_normal((ny, nx))
    f = po.new_f(nx, ny, ld)
    feq = np.stack([W[q] * rho * (1 + 3 * (CX[q] * ux + CY[q] * uy) + 4.5 * (CX[q] * ux + CY[q] * uy) ** 2
                                  - 1.5 * (ux ** 2 + uy ** 2)) for q in range(9)], axis=2)
    f[:ny] = feq * (1 + 0.01 * rng.standard_normal(feq.shape))  # off-equilibrium
    return f


def test_equilibrium_moments(oracle):
    rng = np.random.default_rng(0)
    for _ in range(50):
        rho, ux, uy = 1 + 0.2 * rng.random(), 0.1 * rng.standard_normal(), 0.1 * rng.standard_normal()
        r, jx, jy, pxx, pyy, pxy = moments(oracle.equilibrium(rho, ux, uy))
        assert r == pytest.approx(rho, rel=1e-15)
        assert jx == pytest.approx(rho * ux, abs=1e-15)
        assert jy == pytest.approx(rho * uy, abs=1e-15)
        assert pxx == pytest.approx(rho * (ux * ux + 1 / 3), rel=1e-14)
        assert pyy == pytest.approx(rho * (uy * uy + 1 / 3), rel=1e-14)
        assert pxy == pytest.approx(rho * ux * uy, abs=1e-15)
    np.testing.assert_allclose(oracle.equilibrium(1.3, 0.0, 0.0), 1.3 * W, rtol=1e-15)


def test_stream_is_the_periodic_shift(oracle):
    """Brute force: f_q(y, x) = f_q(y - cy_q, x - cx_q) with periodic wrap (np.roll)."""
    for nx, ny in [(4, 4), (5, 3), (1, 7), (7, 1), (3, 16)]:
        ld = po.padded_ld(ny)
        src = po.new_f(nx, ny, ld)
        src[:ny] = np.arange(ny * nx * 9, dtype=float).reshape(ny, nx, 9, order="F") + 1
        dst = po.new_f(nx, ny, ld)
        oracle.stream(nx, ny, src, dst)
        for q in range(9):
            want = np.roll(src[:ny, :, q], shift=(CY[q], CX[q]), axis=(0, 1))
            np.testing.assert_array_equal(dst[:ny, :, q], want)
        assert dst[:ny].sum() == src[:ny].sum()
        assert not dst[ny:].any()  # padding rows untouched


@pytest.mark.parametrize("model", MODELS)
def test_collisions_conserve_mass_and_momentum(oracle, model):
    nx, ny = 9, 7
    f = random_state(nx, ny, 1)
    before = moments(f[:ny])
    oracle.collide(model, nx, ny, f, 1.7, 0.25)
    after = moments(f[:ny])
    np.testing.assert_allclose(after[0], before[0], rtol=2e-15)
    if model == "trt":  # incompressible: j = sum c f is conserved directly
        np.testing.assert_allclose(after[1], before[1], atol=2e-16)
        np.testing.assert_allclose(after[2], before[2], atol=2e-16)
    else:
        np.testing.assert_allclose(after[1], before[1], atol=2e-16)
        np.testing.assert_allclose(after[2], before[2], atol=2e-16)


@pytest.mark.parametrize("model", MODELS)
def test_rest_state_is_fixed_point(oracle, model):
    nx, ny = 4, 5
    f = po.new_f(nx, ny)
    f[:ny] = 1.07 * W
    g = f.copy(order="F")
    oracle.collide(model, nx, ny, g, 1.3, 0.25)
    np.testing.assert_allclose(g[:ny], f[:ny], rtol=1e-15)


def test_bgk_omega_one_gives_equilibrium(oracle):
    nx, ny = 6, 5
    f = random_state(nx, ny, 2)
    rho, jx, jy = moments(f[:ny])[:3]
    oracle.collide("bgk", nx, ny, f, 1.0)
    for y in range(ny):
        for x in range(nx):
            want = oracle.equilibrium(rho[y, x], jx[y, x] / rho[y, x], jy[y, x] / rho[y, x])
            np.testing.assert_allclose(f[y, x], want, rtol=1e-14)


def test_bgk_relaxation_formula(oracle):
    """f' = f + omega (feq - f) for general omega (textbook BGK)."""
    nx, ny, om = 5, 4, 1.63
    f = random_state(nx, ny, 3)
    f0 = f.copy(order="F")
    oracle.collide("bgk", nx, ny, f, om)
    rho, jx, jy = moments(f0[:ny])[:3]
    for y in range(ny):
        for x in range(nx):
            feq = oracle.equilibrium(rho[y, x], jx[y, x] / rho[y, x], jy[y, x] / rho[y, x])
            np.testing.assert_allclose(f[y, x], f0[y, x] + om * (feq - f0[y, x]), rtol=1e-13)


def incompressible_feq(rho, jx, jy):
    cu = CX * jx + CY * jy
    return W * (rho + 3 * cu + 4.5 * cu * cu - 1.5 * (jx * jx + jy * jy))


def test_trt_equal_rates_is_incompressible_bgk(oracle):
    """lambda_odd == lambda_even reduces TRT to BGK towards the incompressible feq."""
    nx, ny, le = 5, 6, 1.45
    f = random_state(nx, ny, 4)
    f0 = f.copy(order="F")
    oracle.collide_trt(nx, ny, f, le, le)
    rho, jx, jy = moments(f0[:ny])[:3]
    for y in range(ny):
        for x in range(nx):
            feq = incompressible_feq(rho[y, x], jx[y, x], jy[y, x])
            np.testing.assert_allclose(f[y, x], f0[y, x] + le * (feq - f0[y, x]), rtol=1e-13,
                                       atol=1e-16)


def test_trt_odd_part_relaxes_with_lambda_odd(oracle):
    """Antisymmetric parts relax with lambda_odd, symmetric ones with lambda_even."""
    nx, ny, le, lo = 4, 4, 1.3, 0.55
    f = random_state(nx, ny, 5)
    f0 = f.copy(order="F")
    oracle.collide_trt(nx, ny, f, le, lo)
    rho, jx, jy = moments(f0[:ny])[:3]
    opp = [0, 3, 4, 1, 2, 7, 8, 5, 6]
    for y in range(ny):
        for x in range(nx):
            feq = incompressible_feq(rho[y, x], jx[y, x], jy[y, x])
            a, e = f0[y, x], feq
            sym = 0.5 * (a + a[opp]) - le * 0.5 * ((a + a[opp]) - (e + e[opp]))
            asym = 0.5 * (a - a[opp]) - lo * 0.5 * ((a - a[opp]) - (e - e[opp]))
            np.testing.assert_allclose(f[y, x], sym + asym, rtol=1e-13, atol=1e-16)


def test_magic_parameter_identities(oracle):
    for le in (0.5, 1.0, 1.5, 1.94):
        for L in (1 / 12, 3 / 16, 0.25, 1.0):
            lo = oracle.lambda_d(le, L)
            assert oracle.magic_number(le, lo) == pytest.approx(L, rel=1e-14)
        assert oracle.lambda_d(le, 0.25) == pytest.approx(2 - le, rel=1e-15)


def hermite4_feq(rho, ux, uy):
    """4th-order Hermite equilibrium on D2Q9 (textbook form, w_q rho sum_n H_n:u^n/n!cs^2n)."""
    cs2 = 1 / 3
    out = np.zeros(9)
    for q in range(9):
        c = np.array([CX[q], CY[q]])
        u = np.array([ux, uy])
        H2 = np.outer(c, c) - cs2 * np.eye(2)
        t1 = c @ u / cs2
        t2 = np.sum(H2 * np.outer(u, u)) / (2 * cs2 ** 2)
        hxxy = H2[0, 0] * c[1]
        hxyy = H2[1, 1] * c[0]
        t3 = (hxxy * ux * ux * uy + hxyy * ux * uy * uy) * 3 / (6 * cs2 ** 3)
        t4 = H2[0, 0] * H2[1, 1] * ux * ux * uy * uy * 6 / (24 * cs2 ** 4)
        out[q] = W[q] * rho * (1 + t1 + t2 + t3 + t4)
    return out


def test_rr_omega_one_gives_hermite4_equilibrium(oracle):
    nx, ny = 4, 3
    f = random_state(nx, ny, 6)
    rho, jx, jy = moments(f[:ny])[:3]
    oracle.collide("rr", nx, ny, f, 1.0)
    for y in range(ny):
        for x in range(nx):
            want = hermite4_feq(rho[y, x], jx[y, x] / rho[y, x], jy[y, x] / rho[y, x])
            np.testing.assert_allclose(f[y, x], want, rtol=1e-13)


def test_update_macros_definition(oracle):
    nx, ny = 7, 5
    f = random_state(nx, ny, 7, ld=16)
    r, u, v = oracle.update_macros(nx, ny, f)
    rho, jx, jy = moments(f[:ny])[:3]
    np.testing.assert_allclose(r, rho, rtol=1e-15)
    np.testing.assert_allclose(u, jx / rho, rtol=1e-13, atol=1e-17)
    np.testing.assert_allclose(v, jy / rho, rtol=1e-13, atol=1e-17)


@pytest.mark.parametrize("model", MODELS)
def test_shifted_forms_equal_unshifted_in_fp64(oracle, model):
    """Deviation storage g = f - w rho0 is the same collision (FP64, to roundoff)."""
    nx, ny, om = 6, 6, 1.8
    f = random_state(nx, ny, 8)
    g = f.copy(order="F")
    g[:ny] -= W
    lo = oracle.lambda_d(om, 0.25)
    oracle.collide(model, nx, ny, f, om, 0.25)
    oracle.collide_shifted(model, nx, ny, g, om, lo)
    np.testing.assert_allclose(g[:ny], f[:ny] - W, rtol=0, atol=3e-16)
    rs, us, vs = oracle.update_macros(nx, ny, g, shifted=True)
    r, u, v = oracle.update_macros(nx, ny, f)
    np.testing.assert_allclose(rs, r, rtol=2e-15)
    np.testing.assert_allclose(us, u, atol=1e-16)


def test_shifted_equilibrium(oracle):
    for rho, ux, uy in [(1.02, 0.03, -0.04), (0.97, -0.1, 0.05)]:
        np.testing.assert_allclose(oracle.equilibrium(rho, ux, uy, shifted=True) + W,
                                   oracle.equilibrium(rho, ux, uy), atol=2e-16)


def test_taylor_green_closed_form(oracle):
    """The textbook TGV: divergence-free, decays as exp(-nu (kx^2+ky^2) t)."""
    n, u0, nu = 32, 0.05, 0.01
    k = 2 * np.pi / n
    p0, ux0, uy0 = oracle.tgv(n, n, k, k, u0, nu, 0.0)
    p1, ux1, uy1 = oracle.tgv(n, n, k, k, u0, nu, 500.0)
    decay = np.exp(-nu * 2 * k * k * 500.0)
    np.testing.assert_allclose(ux1, ux0 * decay, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(p1, p0 * decay ** 2, rtol=1e-12, atol=1e-18)
    x = np.arange(n) + 0.5
    X, Y = np.meshgrid(x, x)  # (ny, nx)
    np.testing.assert_allclose(ux0, -u0 * np.cos(k * X) * np.sin(k * Y), atol=1e-16)
    np.testing.assert_allclose(uy0, u0 * np.sin(k * X) * np.cos(k * Y), atol=1e-16)


def test_lbm_decay_matches_analytic(oracle):
    """Discrete TGV decay agrees with exp(-2 nu k^2 t) to O(k^2, Ma^2)."""
    n, u0, nu, steps = 32, 0.02, 0.02, 400
    k = 2 * np.pi / n
    p, ux, uy = oracle.tgv(n, n, k, k, u0, nu, 0.0)
    rho = np.asfortranarray(p * 3 + 1)
    par = oracle.set_properties(nu, 1.0, 0.25)
    for model in MODELS:
        fa, fb = po.new_f(n, n), po.new_f(n, n)
        oracle.set_pdf_to_equilibrium(n, n, rho, ux, uy, fa)
        latest, _ = oracle.run(model, n, n, fa, fb, steps, par["omega"], 0.25)
        _, u, v = oracle.update_macros(n, n, latest)
        _, ua, va = oracle.tgv(n, n, k, k, u0, nu, float(steps))
        assert oracle.l2_norm(u, v, ua, va) < 1e-2, model
        # the decay itself: kinetic energy falls as exp(-2 nu k^2 t) per velocity
        # component pair, i.e. by exp(-4 nu k^2 t) (k = |(k, k)|^2 / 2 per axis)
        ke0 = np.sum(ux ** 2 + uy ** 2)
        ke = np.sum(u ** 2 + v ** 2)
        assert ke / ke0 == pytest.approx(np.exp(-4 * nu * k ** 2 * steps), rel=2e-2), model


# SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c): state += 0x9E3779B97F4A7C15,
# then the output mix.  Published test vector (Rosetta Code "Pseudo-random numbers/
# Splitmix64"): the first five outputs from state 1234567, and for state 987654321 the
# counts of floor(5 x) over 100 000 deviates x = output / 2^64.
SPLITMIX64_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                      4593380528125082431, 16408922859458223821]
SPLITMIX64_987654321_COUNTS = [20027, 19892, 20073, 19978, 20030]


def test_noise_generator_published_vectors(oracle):
    """orc_noise(seed, c) is output c + 1 of SplitMix64 from state `seed`, top 53 bits
    scaled to [0, 1): the published outputs pin the mix constants, shifts and counter
    offset bit for bit; the published bucket counts pin 100 000 consecutive deviates."""
    for c, z in enumerate(SPLITMIX64_1234567):
        assert oracle.noise(1234567, c) == (z >> 11) * 2.0 ** -53
    u = np.array([oracle.noise(987654321, c) for c in range(100000)])
    counts = np.bincount(np.floor(5 * u).astype(int), minlength=5)
    assert counts.tolist() == SPLITMIX64_987654321_COUNTS


def test_l2_norm_closed_forms(oracle):
    """The app's relative L2 norm (app/main_taylor_green.f90:174-212,
    norm2(hypot(u-ua, v-va)) / norm2(hypot(ua, va))): a 3-4-5 / 6-8-10 cell gives exactly
    1/2; a uniform relative perturbation u = (1+e) ua gives |e| (a missing square root
    gives e^2, a wrong sign 2+e); u = 0 gives 1; the norm is scale-invariant."""
    one = lambda v: np.array([[v]])
    assert oracle.l2_norm(one(9.0), one(12.0), one(6.0), one(8.0)) == 0.5
    rng = np.random.default_rng(5)
    ua = np.asfortranarray(rng.standard_normal((12, 9)))
    va = np.asfortranarray(rng.standard_normal((12, 9)))
    for e in (1e-3, -0.25, 2.0):
        got = oracle.l2_norm((1 + e) * ua, (1 + e) * va, ua, va)
        assert got == pytest.approx(abs(e), rel=1e-13)
    assert oracle.l2_norm(0 * ua, 0 * va, ua, va) == pytest.approx(1.0, rel=1e-15)
    u = np.asfortranarray(rng.standard_normal((12, 9)))
    v = np.asfortranarray(rng.standard_normal((12, 9)))
    assert oracle.l2_norm(3 * u, 3 * v, 3 * ua, 3 * va) == pytest.approx(
        oracle.l2_norm(u, v, ua, va), rel=1e-14)


def test_shear_layer_closed_forms(oracle):
    """Shear layer (src/benchmarks/shear_layer.f90:25-52): rho = rho0; uy = U0 delta
    cos(2 pi xd) (= sin(2 pi (xd + 1/4))) depends on x only; ux depends on y only, is even
    about yd = 1/2 (ux(1 - yd) = ux(yd)), odd about yd = 1/4 within the lower half, and
    |ux| <= U0 with ux -> -U0 at yd -> 0.  A swapped axis, a dropped 1/4 shift or a sign
    error fails one of these."""
    for n, length in ((32, 32.0), (96, 96.0), (48, 64.0)):
        u0, k, delta = 0.04, 80.0, 0.05
        rho, ux, uy = oracle.shear(n, n, length, u0, k, delta)
        assert np.all(rho == 1.0)
        xd = (np.arange(n) + 0.5) / length
        yd = (np.arange(n) + 0.5) / length
        np.testing.assert_allclose(uy, np.broadcast_to(u0 * delta * np.cos(2 * np.pi * xd),
                                                       (n, n)), rtol=0, atol=1e-15)
        assert np.all(ux == ux[:, :1]) and np.all(uy == uy[:1, :])
        col = ux[:, 0]
        lower = yd <= 0.5
        np.testing.assert_allclose(col[lower], u0 * np.tanh(k * (yd[lower] - 0.25)),
                                   rtol=0, atol=1e-15)
        if length == n:
            # grid mirror y -> n - 1 - y maps yd to 1 - yd
            np.testing.assert_allclose(col, col[::-1], rtol=0, atol=1e-15)
            q = n // 4  # rows 1/4 +- s about yd = 1/4 when n is a multiple of 4
            np.testing.assert_allclose(col[q:2 * q], -col[q - 1::-1], rtol=0, atol=1e-15)
        assert np.all(np.abs(col) <= u0) and col[0] < -0.99 * u0


def test_rr_nonequilibrium_stress_relaxes_by_one_minus_omega(oracle):
    """Regularized collision: the second-order non-equilibrium moments of the output are
    (1 - omega) times those of the input, and the result is the Hermite-4 equilibrium
    plus a purely regularized (Hermite-projected) non-equilibrium part."""
    nx, ny, om = 5, 4, 1.37
    f = random_state(nx, ny, 9)
    f0 = f.copy(order="F")
    oracle.collide("rr", nx, ny, f, om)
    rho, jx, jy = moments(f0[:ny])[:3]
    for y in range(ny):
        for x in range(nx):
            feq = hermite4_feq(rho[y, x], jx[y, x] / rho[y, x], jy[y, x] / rho[y, x])
            pre = moments(f0[y, x] - feq)[3:]
            post = moments(f[y, x] - feq)[3:]
            np.testing.assert_allclose(post, (1 - om) * np.array(pre), rtol=1e-10, atol=1e-17)


@pytest.mark.parametrize("op", ["fvm_bardow", "fdm_bardow", "fdm_sofonea"])
def test_fd_streaming_is_lax_wendroff_along_axes(oracle, op):
    """Along an axis direction all three schemes reduce to 1-D Lax-Wendroff:
    e^{ikx} -> G e^{ikx}, G = 1 - i c sin k - c^2 (1 - cos k), c = dt (closed form)."""
    nx, ny = 32, 8
    kx = 2 * np.pi * 3 / nx
    x = np.arange(nx)
    for dt in (1.0, 0.6):
        f = po.new_f(nx, ny)
        f[:ny, :, 1] = np.cos(kx * x)[None, :]  # q = 1: c = (1, 0)
        out = po.new_f(nx, ny)
        oracle.stream_op(op, nx, ny, f, out, dt)
        G = 1 - 1j * dt * np.sin(kx) - dt * dt * (1 - np.cos(kx))
        want = np.real(G * np.exp(1j * kx * x))
        np.testing.assert_allclose(out[:ny, :, 1], np.broadcast_to(want, (ny, nx)), atol=1e-14)


@pytest.mark.parametrize("op", ["fvm_bardow", "fdm_bardow", "fdm_sofonea"])
def test_fd_streaming_conserves_and_keeps_constants(oracle, op):
    nx, ny = 9, 7
    rng = np.random.default_rng(4)
    f = po.new_f(nx, ny)
    f[:ny] = rng.random((ny, nx, 9))
    out = po.new_f(nx, ny)
    oracle.stream_op(op, nx, ny, f, out, 0.8)
    np.testing.assert_allclose(out[:ny].sum(axis=(0, 1)), f[:ny].sum(axis=(0, 1)), rtol=1e-13)
    np.testing.assert_array_equal(out[:ny, :, 0], f[:ny, :, 0])  # rest population copied
    c = po.new_f(nx, ny)
    c[:ny] = 0.37
    oracle.stream_op(op, nx, ny, c, out, 0.8)
    np.testing.assert_allclose(out[:ny], 0.37, rtol=1e-15)
    oracle.stream_op(op, nx, ny, f, out, 0.0)  # dt = 0: identity
    np.testing.assert_array_equal(out[:ny], f[:ny])


# ---------------------------------------------------------------- DUGKS


def _random_f(nx, ny, seed):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.02 * rng.standard_normal((ny, nx))
    ux, uy = 0.05 * rng.standard_normal((ny, nx)), 0.05 * rng.standard_normal((ny, nx))
    f = po.new_f(nx, ny)
    feq = np.stack([np.stack([po_feq(r, u, v) for r, u, v in zip(rho[:, x], ux[:, x], uy[:, x])])
                    for x in range(nx)], axis=1)
    f[:ny] = feq * (1 + 0.01 * rng.standard_normal(feq.shape))
    return f


def po_feq(rho, ux, uy):
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    c = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]])
    cu = c[:, 0] * ux + c[:, 1] * uy
    return w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux * ux + uy * uy))


def test_dugks_collision_is_bgk(oracle):
    """kernel_bgk (src/periodic_dugks.F90:80-165) is BGK in the split-weight algebra: it
    agrees with collide_bgk to roundoff, omega = 1 gives the equilibrium, and both
    outputs keep each cell's mass and momentum."""
    nx, ny = 7, 5
    f = _random_f(nx, ny, 11)
    fold, fnew = f.copy(order="F"), po.new_f(nx, ny)
    oracle.dugks_collide(nx, ny, fold, fnew, 1.7, 0.9)
    for om, got in ((1.7, fnew), (0.9, fold)):
        want = f.copy(order="F")
        oracle.collide("bgk", nx, ny, want, om)
        np.testing.assert_allclose(got[:ny], want[:ny], rtol=0, atol=1e-15)
        cx = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
        cy = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
        np.testing.assert_allclose(got[:ny].sum(axis=2), f[:ny].sum(axis=2), rtol=1e-15)
        np.testing.assert_allclose(got[:ny] @ cx, f[:ny] @ cx, atol=1e-16)
        np.testing.assert_allclose(got[:ny] @ cy, f[:ny] @ cy, atol=1e-16)
    one, dummy = f.copy(order="F"), po.new_f(nx, ny)
    oracle.dugks_collide(nx, ny, one, dummy, 1.0, 1.0)
    r, u, v = oracle.update_macros(nx, ny, f)
    for x in range(nx):
        for y in range(ny):
            np.testing.assert_allclose(one[y, x], po_feq(r[y, x], u[y, x], v[y, x]),
                                       rtol=0, atol=2e-16)


def test_dugks_stream_is_fvm_flux_form(oracle):
    """kernel_stream with fp = ft is stream_fvm_bardow's update (same face values, same
    balance), so it matches that reference-pinned operator bit for bit; q = 0 is untouched,
    mass per direction is conserved, constants and dt = 0 leave fp unchanged."""
    nx, ny = 9, 7
    rng = np.random.default_rng(5)
    ft = po.new_f(nx, ny)
    ft[:ny] = rng.random((ny, nx, 9))
    fp = ft.copy(order="F")
    oracle.dugks_stream(nx, ny, ft, fp, 0.8)
    fv = po.new_f(nx, ny)
    oracle.stream_op("fvm_bardow", nx, ny, ft, fv, 0.8)
    np.testing.assert_array_equal(fp[:ny], fv[:ny])
    other = po.new_f(nx, ny)
    other[:ny] = rng.random((ny, nx, 9))
    before = other.copy(order="F")
    oracle.dugks_stream(nx, ny, ft, other, 0.8)
    np.testing.assert_array_equal(other[:ny, :, 0], before[:ny, :, 0])
    np.testing.assert_allclose(other[:ny].sum(axis=(0, 1)), before[:ny].sum(axis=(0, 1)),
                               rtol=1e-13)
    const = po.new_f(nx, ny)
    const[:ny] = 0.37
    other = before.copy(order="F")
    oracle.dugks_stream(nx, ny, const, other, 0.8)
    np.testing.assert_allclose(other[:ny], before[:ny], rtol=1e-14)
    other = before.copy(order="F")
    oracle.dugks_stream(nx, ny, ft, other, 0.0)
    np.testing.assert_array_equal(other[:ny], before[:ny])


@pytest.mark.parametrize("q,c", [(2, (0, 1)), (5, (1, 1)), (6, (-1, 1))])
def test_dugks_stream_is_second_order_advection(oracle, q, c):
    """With fp = ft, one stream step advects a Fourier mode by dt c_q with a local error
    of third order in k (halving k divides it by ~8): a wrong sign or neighbour in a face
    value or cross term drops the order."""
    n, dt = 64, 0.7
    x, y = np.arange(n)[None, :], np.arange(n)[:, None]
    errs = []
    for m in (1, 2, 4):
        kx, ky = 2 * np.pi * m / n, 2 * np.pi * 2 * m / n
        ft = po.new_f(n, n)
        ft[:n, :, q] = np.cos(kx * x + ky * y)
        fp = ft.copy(order="F")
        oracle.dugks_stream(n, n, ft, fp, dt)
        exact = np.cos(kx * (x - c[0] * dt) + ky * (y - c[1] * dt))
        errs.append(np.abs(fp[:n, :, q] - exact).max())
    assert 7.0 < errs[1] / errs[0] < 8.5 and 7.0 < errs[2] / errs[1] < 8.5, errs
