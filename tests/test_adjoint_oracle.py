"""Pins of the CPU oracle of the differentiable scheme (van Albada + HLLE).

Neither piece exists in the reference, so nothing here compares with it:
the limiter is checked against van Albada's phi(r) = (r^2 + r)/(r^2 + 1) form
and its limits, HLLE against Toro's HLL formula (Riemann Solvers, 3rd ed.,
eq. 10.21) with Davis speeds and against its consistency, upwinding and
mirror-symmetry properties, the RK2 scheme against free-stream preservation
and exact periodic conservation, and the tangent-linear model against central
differences of the primal.
"""
import numpy as np
import pytest

import pyoracle as O
from fornaxamr_amd import problems as P

GAMMA = 5.0 / 3.0
XL = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
E2 = 1e-12          # van Albada eps^2 for the limiter checks
E2S = (1.0 / 16) ** 3   # dx^3 on the 16^3 test grids


def _phys_flux(q, gamma):
    rho, u, v, w, p, ye = q
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, u * (E + p), rho * ye * u])


def _cons(q, gamma):
    rho, u, v, w, p, ye = q
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.array([rho, rho * u, rho * v, rho * w, E, rho * ye])


def _rand_q(rng, mach=0.5):
    rho = rng.uniform(0.5, 2.0)
    p = rng.uniform(0.3, 3.0)
    c = np.sqrt(GAMMA * p / rho)
    return np.array([rho, *(mach * c * rng.uniform(-1, 1, 3)), p, rng.uniform(0, 1)])


def test_van_albada_limits():
    rng = np.random.default_rng(1)
    for _ in range(200):
        a, b = rng.uniform(-1, 1, 2) * 10.0 ** rng.uniform(-3, 1, 2)
        assert O.adj_va(a, b, E2) == O.adj_va(b, a, E2)
        assert O.adj_va(-a, -b, E2) == -O.adj_va(a, b, E2)
        r = a / b
        phi = (r * r + r) / (r * r + 1.0)            # eps -> 0 form
        # to first order the regularisation moves the slope by
        # eps^2 (a + b)(a - b)^2 / (a^2 + b^2)^2
        shift = 1e-12 * abs(a + b) * (a - b) ** 2 / (a * a + b * b) ** 2
        assert abs(O.adj_va(a, b, E2) - phi * b) <= 1.01 * shift + 1e-15 * abs(phi * b)
        assert O.adj_va(a, a, E2) == pytest.approx(a, rel=2e-16)
    assert abs(O.adj_va(0.0, 1e-2, E2)) < 1e-8          # flat side: slope ~ eps^2/s
    assert O.adj_va(0.0, 0.0, E2) == 0.0


def test_hlle_consistency_and_upwinding():
    rng = np.random.default_rng(2)
    for _ in range(100):
        q = _rand_q(rng)
        assert np.allclose(O.adj_hlle(q, q, GAMMA), _phys_flux(q, GAMMA), rtol=1e-14, atol=1e-15)
        # supersonic to the right: F = F(qL); to the left: F = F(qR)
        qL, qR = _rand_q(rng), _rand_q(rng)
        for s, side in ((1.0, 0), (-1.0, 1)):
            a, b = qL.copy(), qR.copy()
            for t in (a, b):
                t[1] = s * (np.sqrt(GAMMA * t[4] / t[0]) * rng.uniform(1.1, 3.0))
            F = O.adj_hlle(a, b, GAMMA)
            assert np.allclose(F, _phys_flux((a, b)[side], GAMMA), rtol=1e-14, atol=1e-15)


def test_hlle_matches_toro_hll_with_davis_speeds():
    rng = np.random.default_rng(3)
    for _ in range(200):
        qL, qR = _rand_q(rng, 0.9), _rand_q(rng, 0.9)
        cL = np.sqrt(GAMMA * qL[4] / qL[0])
        cR = np.sqrt(GAMMA * qR[4] / qR[0])
        SL = min(qL[1] - cL, qR[1] - cR)
        SR = max(qL[1] + cL, qR[1] + cR)
        FL, FR = _phys_flux(qL, GAMMA), _phys_flux(qR, GAMMA)
        UL, UR = _cons(qL, GAMMA), _cons(qR, GAMMA)
        if SL >= 0:
            Fx = FL
        elif SR <= 0:
            Fx = FR
        else:
            Fx = (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)
        assert np.allclose(O.adj_hlle(qL, qR, GAMMA), Fx, rtol=1e-12, atol=1e-13)


def test_hlle_mirror_symmetry():
    rng = np.random.default_rng(4)
    flip = np.array([1, -1, 1, 1, 1, 1.0])
    for _ in range(100):
        qL, qR = _rand_q(rng), _rand_q(rng)
        F = O.adj_hlle(qL, qR, GAMMA)
        Fm = O.adj_hlle(qR * flip, qL * flip, GAMMA)
        # mass, energy, tangential momenta and Ye fluxes change sign; normal momentum not
        assert np.allclose(Fm, -F * flip, rtol=1e-13, atol=1e-14)


def _smooth_state(n, seed, amp=0.05):
    pr = P.random_smooth(n=n, seed=seed, bc=(3,) * 6, gamma=GAMMA)
    W = pr.W.copy()
    W[1:4] *= 0.2
    return W


def test_free_stream_and_conservation():
    W = _smooth_state((16, 16, 16), 5)
    U = np.array([W[0], W[0] * W[1], W[0] * W[2], W[0] * W[3],
                  W[0] * W[4] + 0.5 * W[0] * (W[1] ** 2 + W[2] ** 2 + W[3] ** 2), W[0] * W[5]])
    WT = O.adj_forward(W, XL, GAMMA, E2S, 2e-3, 5)
    UT = np.array([WT[0], WT[0] * WT[1], WT[0] * WT[2], WT[0] * WT[3],
                   WT[0] * WT[4] + 0.5 * WT[0] * (WT[1] ** 2 + WT[2] ** 2 + WT[3] ** 2), WT[0] * WT[5]])
    t0, t1 = U.sum(axis=(1, 2, 3)), UT.sum(axis=(1, 2, 3))
    scale = np.array([t0[0]] * 4 + [t0[4], t0[5]])
    assert np.all(np.abs(t1 - t0) / scale < 1e-13)
    Wu = np.broadcast_to(W[:, :1, :1, :1], W.shape).copy()
    assert np.allclose(O.adj_forward(Wu, XL, GAMMA, E2S, 2e-3, 3), Wu, rtol=2e-15, atol=0)


def test_tlm_primal_part_and_linearity():
    W = _smooth_state((16, 16, 16), 6)
    rng = np.random.default_rng(6)
    d = rng.standard_normal(W.shape)
    WT, dWT = O.adj_tlm(W, d, XL, GAMMA, E2S, 2e-3, 3)
    assert np.array_equal(WT, O.adj_forward(W, XL, GAMMA, E2S, 2e-3, 3))
    _, dWT2 = O.adj_tlm(W, 2.0 * d, XL, GAMMA, E2S, 2e-3, 3)
    assert np.array_equal(dWT2, 2.0 * dWT)


def test_tlm_matches_central_differences():
    W = _smooth_state((16, 16, 16), 7)
    rng = np.random.default_rng(7)
    for _ in range(3):
        d = rng.standard_normal(W.shape) * np.array([1, 0.2, 0.2, 0.2, 1, 0.1])[:, None, None, None]
        _, dWT = O.adj_tlm(W, d, XL, GAMMA, E2S, 2e-3, 4)
        # 1e-7: small enough not to straddle the kinks of the Davis min/max
        # (the tangent error then is round-off, ~3e-9 here)
        eps = 1e-7
        Wp = O.adj_forward(W + eps * d, XL, GAMMA, E2S, 2e-3, 4)
        Wm = O.adj_forward(W - eps * d, XL, GAMMA, E2S, 2e-3, 4)
        fd = (Wp - Wm) / (2 * eps)
        assert np.max(np.abs(fd - dWT)) / np.max(np.abs(dWT)) < 1e-7


def test_reverse_sweep_dot_product():
    """The reverse-sweep formulas of the GPU kernels, restated array-wise,
    against the tangent-linear model: <Wbar_0, d> == <Wbar_T, TLM d>."""
    import fxa_reverse as R
    W = P.random_smooth(n=(16, 16, 16), seed=3, bc=(3,) * 6, gamma=1.4).W[:, ::2, ::2, ::2].copy()
    W[1:4] *= 0.3
    e, dt, ns, dx = (1 / 8) ** 3, 2e-3, 3, [1 / 8] * 3
    rng = np.random.default_rng(0)
    gWT = rng.standard_normal(W.shape)
    gW0, WT = R.reverse(W, gWT, 1.4, e, dt, ns, dx)
    assert np.allclose(WT, O.adj_forward(W, XL, 1.4, e, dt, ns), rtol=1e-13, atol=0)
    for _ in range(3):
        d = rng.standard_normal(W.shape)
        _, dWT = O.adj_tlm(W, d, XL, 1.4, e, dt, ns)
        a, b = np.sum(gW0 * d), np.sum(gWT * dWT)
        assert abs(a - b) <= 1e-13 * abs(b)
