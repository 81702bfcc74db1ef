"""Shared helpers for the parity tests."""
from __future__ import annotations

import hashlib

import numpy as np


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def max_rel(a, b, floor=1e-300):
    """Max over variables of max|a-b| / max|b| (per conserved/primitive variable)."""
    out = 0.0
    for v in range(a.shape[0]):
        scale = max(np.max(np.abs(b[v])), floor)
        out = max(out, float(np.max(np.abs(a[v] - b[v]))) / scale)
    return out


def ulp_diff(a, b):
    """Max distance in units in the last place between two float64 arrays."""
    ai = np.ascontiguousarray(a, dtype=np.float64).view(np.int64)
    bi = np.ascontiguousarray(b, dtype=np.float64).view(np.int64)
    ai = np.where(ai < 0, np.int64(-(2**63)) - ai, ai)
    bi = np.where(bi < 0, np.int64(-(2**63)) - bi, bi)
    return int(np.max(np.abs(ai - bi))) if ai.size else 0


def cons_totals(W, vol, mhd=False):
    """Sum over cells of the conserved variables times the cell volume."""
    rho, v1, v2, v3, e, ye = W[:6]
    E = rho * e + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3)
    if mhd:
        E = E + 0.5 * (W[6] ** 2 + W[7] ** 2 + W[8] ** 2)
    return np.array([rho.sum(), (rho * v1).sum(), (rho * v2).sum(), (rho * v3).sum(), E.sum(),
                     (rho * ye).sum()]) * vol


def div_b(bf, dx):
    bx, by, bz = bf
    return ((bx[1:] - bx[:-1]) / dx[0] + (by[:, 1:] - by[:, :-1]) / dx[1] +
            (bz[:, :, 1:] - bz[:, :, :-1]) / dx[2])


def glibc_log_misrounded(args):
    """Arguments (float64 > 1) on which the host libm's log -- the reference's
    and the oracle's -- does not return the correctly rounded double.  The GPU
    path takes log correctly rounded, so a parity test through log can only be
    bit-exact where this list is empty.  Screened in long double (64-bit
    significand), with an exact decimal evaluation for the few arguments whose
    log lies close to a rounding boundary."""
    import math
    from decimal import Decimal, localcontext

    a = np.unique(np.asarray(args, dtype=np.float64).ravel())
    if a.size == 0:
        return []
    g = np.log(a)                                   # numpy's log: screened, not trusted
    g = np.array([math.log(x) for x in a])          # the host libm's
    ll = np.log(a.astype(np.longdouble))
    half = np.spacing(g).astype(np.longdouble) / 2
    dist = np.abs(ll - g.astype(np.longdouble))
    margin = np.abs(ll) * np.longdouble(2.0) ** -58
    bad = list(a[dist > half + margin])
    for x, gx in zip(a[np.abs(dist - half) <= margin], g[np.abs(dist - half) <= margin]):
        with localcontext() as c:
            c.prec = 60
            cr = float(format(Decimal(float(x)).ln(), ".45e"))
        if cr != gx:
            bad.append(float(x))
    return sorted(set(bad))


def eta_mask_log_args(W):
    """The log arguments of the anomalous-resistivity mask that matter: rho/1e13
    and Ye/0.25 of the cells where both exceed 1 (elsewhere a factor is 0)."""
    a = W[0] / 1.0e13
    b = W[5] / 0.25
    sel = (a > 1.0) & (b > 1.0)
    return np.concatenate([a[sel], b[sel]])
