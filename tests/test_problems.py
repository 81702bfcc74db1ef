"""The synthetic workloads against the configurations they stand for (CPU).

C1 Sedov 64^3, C2 periodic turbulent box, C3 magnetized Sedov 128^3, C4 the
ragged ensemble, C5 the acoustic box: what each configuration fixes is checked
from the generated arrays by independent means (totals, counts of lattice
points, Fourier spectra, the discrete divergence of the face field), not by
re-running the generator's arithmetic.
"""
import itertools

import numpy as np
import pytest

from fornaxamr_amd import problems as P
from helpers import div_b


def _dx_vol(pr):
    dx = [(pr.xlim[2 * d + 1] - pr.xlim[2 * d]) / pr.n[d] for d in range(3)]
    return dx, dx[0] * dx[1] * dx[2]


def _ball_count(radius_cells):
    """Cells of a grid centred on the origin (centres at half-odd multiples of
    dx) whose centres lie within radius_cells * dx: counted by brute force."""
    h = [i + 0.5 for i in range(-int(radius_cells) - 1, int(radius_cells) + 1)]
    return sum(1 for a, b, c in itertools.product(h, h, h)
               if a * a + b * b + c * c < radius_cells ** 2)


def _injected(pr, p0):
    dx, vol = _dx_vol(pr)
    E = pr.W[0] * pr.W[4]                     # internal energy density (v = 0)
    E0 = p0 / (pr.gamma - 1.0)
    return float(((E - E0) * vol).sum()), int(np.count_nonzero(E > E0 * (1 + 1e-9)))


def test_c1_sedov_64():
    pr = P.sedov(n=(64, 64, 64))
    assert pr.n == (64, 64, 64) and pr.nblocks == 64
    assert pr.xlim == (-1, 1, -1, 1, -1, 1) and pr.bc == (P.OUTFLOW,) * 6
    assert pr.gamma == 5.0 / 3.0
    dx, _ = _dx_vol(pr)
    assert dx == [1.0 / 32] * 3
    assert np.all(pr.W[0] == 1.0) and np.all(pr.W[1:4] == 0.0)
    total, hot = _injected(pr, 1.0e-5)
    # r < 2 dx: the 8 cells at r = 0.87 dx and the 24 at r = 1.66 dx
    assert hot == _ball_count(2.0) == 32
    assert total == pytest.approx(1.0, rel=1e-12)


def test_c2_turbulent_box():
    n = 64
    pr = P.turbulent_box(n=(n, n, n), seed=1)
    assert pr.bc == (P.PERIODIC,) * 6 and pr.xlim == (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    W = pr.W
    p = (pr.gamma - 1.0) * W[0] * W[4]
    assert np.allclose(p, 1.0, rtol=1e-14, atol=0)
    v2 = (W[1:4] ** 2).sum(axis=0)
    assert np.sqrt(v2.mean()) == pytest.approx(0.3, rel=1e-12)
    k = np.fft.fftfreq(n) * n
    kmax = np.maximum(np.maximum(np.abs(k)[:, None, None], np.abs(k)[None, :, None]),
                      np.abs(k)[None, None, :])
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    for d in range(3):
        f = np.abs(np.fft.fftn(W[1 + d]))
        assert f[k2 > 16].max() < 1e-10 * f.max()        # filtered to |k| <= 4
    # rho - 1: 0.1 x a sum of 8 sine modes with integer wavenumbers 1..4
    f = np.fft.fftn(W[0] - 1.0) / n ** 3
    big = np.abs(f) > 1e-12
    assert 2 <= big.sum() <= 16
    assert kmax[big].max() <= 4 and kmax[big].min() >= 1
    # each of the 8 modes puts 0.1 / 2 at +k and at -k; with 8 distinct +-k
    # pairs (as for this seed) every coefficient has exactly that magnitude
    mags = np.abs(f[big])
    assert big.sum() == 16
    assert np.allclose(mags, 0.05, rtol=0, atol=1e-12)


def test_c3_magnetized_sedov_128():
    pr = P.magnetized_sedov(n=(128, 128, 128))
    assert pr.mhd and pr.nblocks == 512 and pr.bc == (P.OUTFLOW,) * 6
    bx, by, bz = pr.bf
    assert bx.shape == (129, 128, 128) and by.shape == (128, 129, 128) and bz.shape == (128, 128, 129)
    assert np.all(bx == 0.1) and np.all(by == 0.0) and np.all(bz == 0.0)
    dx, _ = _dx_vol(pr)
    assert np.all(div_b(pr.bf, dx) == 0.0)
    assert np.all(pr.W[6] == 0.1) and np.all(pr.W[7:9] == 0.0)
    total, hot = _injected(pr, 1.0e-3)
    assert hot == _ball_count(3.0)
    assert total == pytest.approx(1.0, rel=1e-12)


def test_c4_ragged_ensemble():
    ens = P.ragged_ensemble()
    assert len(ens) == 48
    sizes = [p.n[0] for p in ens]
    assert all(p.n == (s, s, s) for p, s in zip(ens, sizes))
    assert set(sizes) <= {32, 48, 64, 96, 128} and len(set(sizes)) >= 4
    assert all(1.3 <= p.gamma <= 1.7 for p in ens)
    assert len({p.gamma for p in ens}) == 48
    assert all(p.bc == (P.PERIODIC,) * 6 for p in ens)
    # each instance its own seed: no two initial states alike
    firsts = {float(p.W[0, 0, 0, 0]) for p in ens}
    assert len(firsts) == 48
    assert sum(p.nblocks for p in ens) == sum((s // 16) ** 3 for s in sizes)


def test_c5_acoustic_box():
    pr = P.acoustic_box(n=(64, 64, 64))
    W = pr.W
    g = pr.gamma
    p = (g - 1.0) * W[0] * W[4]
    assert np.allclose(p, 1.0 / g, rtol=1e-14, atol=0)
    cs0 = np.sqrt(g * (1.0 / g) / 1.0)
    assert cs0 == 1.0
    drho = W[0] - 1.0
    # largest sampled value: centres nearest x = 1/4 and y = 0 sit dx/2 away
    assert np.abs(drho).max() == pytest.approx(1e-3 * np.cos(np.pi / 64) ** 2, rel=1e-12)
    # right-moving linear sound wave: v_x = c_s drho / rho_0, v_y = v_z = 0
    assert np.array_equal(W[1], cs0 * drho)
    assert np.all(W[2:4] == 0.0)
    # rho depends on x and y only, as sin(2 pi x) cos(2 pi y)
    assert np.all(W[0] == W[0][:, :, :1])
