"""Synthetic initial conditions for the RK2-stage workloads.

Each generator returns a :class:`Problem`: global active-cell primitives
``W[var][X][Y][Z]`` (Z fastest; vars rho, v1, v2, v3, eint, ye [, b1, b2, b3]),
for MHD the face fields ``Bx[X+1][Y][Z]``, ``By[X][Y+1][Z]``, ``Bz[X][Y][Z+1]``,
and the mesh description.  All arrays are float64.  Generators are seeded and
deterministic, so the GPU path and the CPU checkers see identical inputs.

Energy deposition follows the reference Sedov set-up: a fixed total energy
spread as a uniform energy density over the cells whose centres lie inside the
injection radius (problems/prj_problem_sedov.c:283-301).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

OUTFLOW, REFLECT, PERIODIC = 0, 1, 3
YE_PASSIVE = 0.1


@dataclass
class Problem:
    name: str
    n: tuple            # active cells per axis (multiples of 16)
    xlim: tuple         # (x1min, x1max, x2min, x2max, x3min, x3max)
    bc: tuple           # six reference boundary codes
    gamma: float
    W: np.ndarray       # (nh, nx, ny, nz)
    bf: list = field(default=None)   # MHD face fields or None
    mhd: bool = False
    eta: float = 0.0
    cfl: float = 0.4

    @property
    def nblocks(self):
        return int(np.prod([v // 16 for v in self.n]))

    @property
    def ncells(self):
        return int(np.prod(self.n))


def _dx(n, xlim):
    # same arithmetic as the mesh: block extent first, then cell width
    return [((xlim[2 * d + 1] - xlim[2 * d]) / (n[d] // 16)) / 16.0 for d in range(3)]


def _centres(n, xlim):
    dx = _dx(n, xlim)
    return [xlim[2 * d] + (np.arange(n[d]) + 0.5) * dx[d] for d in range(3)], dx


def _deposit(W, n, xlim, energy, radius_cells, mhd_b=None):
    """Add `energy` as uniform energy density to cells with r < radius_cells*dx."""
    (x, y, z), dx = _centres(n, xlim)
    r = np.sqrt(x[:, None, None] ** 2 + y[None, :, None] ** 2 + z[None, None, :] ** 2)
    sel = r < radius_cells * dx[0]
    count = int(sel.sum())
    vol = dx[0] * dx[1] * dx[2]
    edens = (1.0 / count) * energy / vol
    rho = W[0]
    # prim -> cons -> add -> prim, velocity zero
    E = rho * W[4]
    E = np.where(sel, E + edens, E)
    W[4] = E / rho
    return count


def sedov(n=(64, 64, 64), xlim=(-1, 1, -1, 1, -1, 1), rho=1.0, p=1.0e-5, energy=1.0,
          radius_cells=2.0, gamma=5.0 / 3.0, bc=(OUTFLOW,) * 6, cfl=0.4):
    """3-D Sedov blast on a uniform mesh (hydro)."""
    W = np.zeros((6,) + tuple(n))
    W[0] = rho
    W[4] = p / ((gamma - 1.0) * rho)
    W[5] = YE_PASSIVE
    _deposit(W, n, xlim, energy, radius_cells)
    return Problem("sedov", tuple(n), tuple(xlim), tuple(bc), gamma, W, cfl=cfl)


def _lowpass_gaussian(rng, n, kmax):
    """White Gaussian noise filtered to |k| <= kmax (integer wavenumbers)."""
    noise = rng.standard_normal(n)
    f = np.fft.rfftn(noise)
    kx = np.fft.fftfreq(n[0]) * n[0]
    ky = np.fft.fftfreq(n[1]) * n[1]
    kz = np.fft.rfftfreq(n[2]) * n[2]
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    f[k2 > kmax * kmax] = 0.0
    return np.fft.irfftn(f, s=n, axes=(0, 1, 2))


def turbulent_box(n=(256, 256, 256), seed=1, gamma=5.0 / 3.0, rho0=1.0, p=1.0,
                  amp=0.1, nmodes=8, vrms=0.3, kmax=4, cfl=0.4):
    """Periodic box: rho = rho0 + amp * (sum of 8 seeded sine modes with integer
    wavenumbers 1..4 and seeded phases), uniform pressure, and a Gaussian random
    velocity filtered to |k| <= kmax and scaled to rms vrms."""
    rng = np.random.default_rng(seed)
    xlim = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    (x, y, z), _ = _centres(n, xlim)
    rho = np.full(tuple(n), rho0)
    for _ in range(nmodes):
        k = rng.integers(1, kmax + 1, size=3) * rng.choice([-1, 1], size=3)
        k[rng.integers(0, 3)] = rng.integers(1, kmax + 1)
        phase = rng.uniform(0.0, 2.0 * np.pi)
        rho += amp * np.sin(2.0 * np.pi * (k[0] * x[:, None, None] + k[1] * y[None, :, None] +
                                           k[2] * z[None, None, :]) + phase)
    v = np.stack([_lowpass_gaussian(rng, tuple(n), kmax) for _ in range(3)])
    v *= vrms / np.sqrt(np.mean(np.sum(v * v, axis=0)))
    W = np.zeros((6,) + tuple(n))
    W[0] = rho
    W[1:4] = v
    W[4] = p / ((gamma - 1.0) * rho)
    W[5] = YE_PASSIVE
    return Problem("turbulent_box", tuple(n), xlim, (PERIODIC,) * 6, gamma, W, cfl=cfl)


def magnetized_sedov(n=(128, 128, 128), xlim=(-1, 1, -1, 1, -1, 1), rho=1.0, p=1.0e-3,
                     b0=0.1, energy=1.0, radius_cells=3.0, gamma=5.0 / 3.0, cfl=0.4):
    """Sedov blast in a uniform field B = (b0, 0, 0), set on faces as the curl of
    A = (0, 0, b0*y) (exact: Bx = b0 on every x-face, By = Bz = 0)."""
    W = np.zeros((9,) + tuple(n))
    W[0] = rho
    W[4] = p / ((gamma - 1.0) * rho)
    W[5] = YE_PASSIVE
    nx, ny, nz = n
    bf = [np.full((nx + 1, ny, nz), b0), np.zeros((nx, ny + 1, nz)), np.zeros((nx, ny, nz + 1))]
    W[6] = b0
    # deposit into total energy; magnetic energy is unchanged, so eint gets it
    _deposit(W, n, xlim, energy, radius_cells)
    return Problem("magnetized_sedov", tuple(n), tuple(xlim), (OUTFLOW,) * 6, gamma, W, bf=bf,
                   mhd=True, cfl=cfl)


def random_smooth(n=(32, 32, 32), seed=0, bc=(PERIODIC,) * 6, gamma=1.4, mhd=False,
                  vamp=0.5, bamp=0.3, cfl=0.4, eta=0.0):
    """Smooth random state with O(1) variations, for parity tests.  For MHD the
    face fields come from a random smooth vector potential (so div B = 0 to
    round-off) plus a uniform field; `eta` is the uniform CT resistivity
    (mesh->mhd_eta)."""
    rng = np.random.default_rng(seed)
    xlim = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    nh = 9 if mhd else 6
    W = np.zeros((nh,) + tuple(n))
    W[0] = 1.0 + 0.3 * _lowpass_gaussian(rng, tuple(n), 3) / 0.3
    W[0] = 1.0 + 0.25 * np.tanh(W[0] - 1.0)
    for d in range(3):
        g = _lowpass_gaussian(rng, tuple(n), 3)
        W[1 + d] = vamp * g / np.std(g)
    g = _lowpass_gaussian(rng, tuple(n), 3)
    p = 1.0 + 0.3 * np.tanh(g / np.std(g))
    W[4] = p / ((gamma - 1.0) * W[0])
    W[5] = 0.1 + 0.05 * np.tanh(_lowpass_gaussian(rng, tuple(n), 2) * 10)
    bf = None
    if mhd:
        nx, ny, nz = n
        dx = _dx(n, xlim)
        # vector potential on edges: A_d lives on d-edges (cell-centred along d)
        A = []
        shapes = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
        for s in shapes:
            a = rng.standard_normal(s)
            f = np.fft.rfftn(a)
            k = [np.fft.fftfreq(s[0]) * s[0], np.fft.fftfreq(s[1]) * s[1], np.fft.rfftfreq(s[2]) * s[2]]
            k2 = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
            f[k2 > 9] = 0.0
            a = np.fft.irfftn(f, s=s, axes=(0, 1, 2))
            A.append(bamp * a / np.std(a) * dx[0] * 4)
        Ax, Ay, Az = A
        # B = curl A on faces (exactly divergence-free in the discrete sense)
        bx = (Az[:, 1:, :] - Az[:, :-1, :]) / dx[1] - (Ay[:, :, 1:] - Ay[:, :, :-1]) / dx[2]
        by = (Ax[:, :, 1:] - Ax[:, :, :-1]) / dx[2] - (Az[1:, :, :] - Az[:-1, :, :]) / dx[0]
        bz = (Ay[1:, :, :] - Ay[:-1, :, :]) / dx[0] - (Ax[:, 1:, :] - Ax[:, :-1, :]) / dx[1]
        bx = bx + 0.2
        bf = [np.ascontiguousarray(bx), np.ascontiguousarray(by), np.ascontiguousarray(bz)]
        W[6] = 0.5 * (bf[0][:-1] + bf[0][1:])
        W[7] = 0.5 * (bf[1][:, :-1] + bf[1][:, 1:])
        W[8] = 0.5 * (bf[2][:, :, :-1] + bf[2][:, :, 1:])
    return Problem("random_smooth", tuple(n), xlim, tuple(bc), gamma, W, bf=bf, mhd=mhd, cfl=cfl,
                   eta=float(eta))


# thresholds of the anomalous CT resistivity (src/prj_timeint.c:224-226)
ETA_RHO_FLOOR = 1.0e13
ETA_YE_THR = 0.25


def dense_resistive(n=(32, 32, 32), seed=0, eta=0.0, ye_amp=1.0e-13, cfl=0.4):
    """MHD state in the regime of the anomalous CT resistivity: the outflow
    random_smooth state with density and pressure scaled by 1e13 (rho within
    a few per cent of the 1e13 floor, above it in about half of the box) and
    the field by 2^22 (a power of two, so the discrete div B is that of the
    unscaled state; plasma beta drops by 1.76) -- velocities and sound speeds
    are those of the unscaled state -- and
    Ye = 0.25 (1 + ye_amp tanh(g)) for a smooth random g, above the 0.25
    threshold in about half of the box.  The per-cell mask
    1e12 log(rho/1e13) log(Ye/0.25) is then positive on the patch where both
    hold and at most about 1e12 * 0.05 * ye_amp (about 0.005 for the default),
    which keeps the resistive step well inside explicit stability."""
    pr = random_smooth(n=n, seed=seed, bc=(OUTFLOW,) * 6, gamma=1.4, mhd=True, eta=eta, cfl=cfl)
    rng = np.random.default_rng(seed + 7919)
    W = pr.W
    W[0] *= 1.0e13             # eint = p / ((gamma - 1) rho) is unchanged
    W[6:9] *= 2.0 ** 22
    bf = [b * 2.0 ** 22 for b in pr.bf]
    g = _lowpass_gaussian(rng, tuple(n), 3)
    W[5] = ETA_YE_THR * (1.0 + ye_amp * np.tanh(2.0 * g / np.std(g)))
    return Problem("dense_resistive", tuple(n), pr.xlim, pr.bc, pr.gamma, W, bf=bf, mhd=True,
                   cfl=cfl, eta=float(eta))


def eta_mask_cells(W):
    """Cells where the anomalous resistivity mask is positive (both log
    arguments above 1)."""
    return (W[0] / ETA_RHO_FLOOR > 1.0) & (W[5] / ETA_YE_THR > 1.0)


def ragged_ensemble(count=48, sizes=(32, 48, 64, 96, 128), seed=2026, gamma_range=(1.3, 1.7),
                    cfl=0.4):
    """Ragged ensemble of independent periodic turbulent boxes: edge sizes drawn
    uniformly from `sizes` (cubed), gamma uniform in gamma_range, each instance
    with its own seed (1..count)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        g = float(rng.uniform(*gamma_range))
        p = turbulent_box(n=(n, n, n), seed=i + 1, gamma=g, cfl=cfl)
        p.name = f"ensemble[{i}]"
        out.append(p)
    return out


def acoustic_box(n=(64, 64, 64), gamma=5.0 / 3.0, amp=1.0e-3, cfl=0.4):
    """Smooth acoustic state in a periodic unit box: rho = 1 + amp sin(2 pi x)
    cos(2 pi y), p = 1/gamma (sound speed 1), v_x = c_s (rho - 1)/rho_0 (the
    right-moving linear sound wave's velocity), v_y = v_z = 0, Ye passive."""
    xlim = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    (x, y, z), _ = _centres(n, xlim)
    rho = 1.0 + amp * np.sin(2.0 * np.pi * x)[:, None, None] * np.cos(2.0 * np.pi * y)[None, :, None] \
        * np.ones(n[2])[None, None, :]
    p = np.full(tuple(n), 1.0 / gamma)
    W = np.zeros((6,) + tuple(n))
    W[0] = rho
    W[1] = 1.0 * (rho - 1.0)
    W[4] = p / ((gamma - 1.0) * rho)
    W[5] = YE_PASSIVE
    return Problem("acoustic_box", tuple(n), xlim, (PERIODIC,) * 6, gamma, W, cfl=cfl)


def prim_from_rho_p(rho, p, v, ye, gamma):
    """W = (rho, v1, v2, v3, eint, Ye) from density, pressure, velocity (3, ...)."""
    return np.stack([rho, v[0], v[1], v[2], p / ((gamma - 1.0) * rho), ye])


def cfl_dt(W, gamma, dx, cfl):
    """CFL step of a state, sound speed only (src/prj_timeint.c:2163-2182):
    cfl / max over cells of sum_d (|v_d| + c_s) / dx_d."""
    rho = W[0]
    p = (gamma - 1.0) * rho * W[4]
    cs = np.sqrt(gamma * p / rho)
    s = sum((np.abs(W[1 + d]) + cs) / dx[d] for d in range(3))
    return cfl / float(np.max(s))


def subbox_mask(n, xlim, lo=0.25, hi=0.5):
    """Cells whose centres lie in [lo, hi]^3."""
    (x, y, z), _ = _centres(n, xlim)
    mx = (x >= lo) & (x <= hi)
    my = (y >= lo) & (y <= hi)
    mz = (z >= lo) & (z <= hi)
    return mx[:, None, None] & my[None, :, None] & mz[None, None, :]
