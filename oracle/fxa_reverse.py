"""numpy restatement of the reverse sweep of the differentiable scheme.

TEST INFRASTRUCTURE ONLY (imported by tests/).  Mirrors, array-wise, the
per-face HLLE VJP and per-cell slope gather that the GPU reverse kernels
(fornaxamr_amd/csrc/fx_adjoint.hip) run, so that the derivation itself is
pinned on the CPU: the adjoint of a few steps must satisfy the dot-product
identity <Wbar_0, d> = <Wbar_T, TLM(d)> against the C tangent-linear model
(fxa_oracle.c).  The forward here (L_of, cons, prim) follows fxa_body.inc.
"""
import numpy as np

def rot(d):
    return [0, 1 + d, 1 + (d + 1) % 3, 1 + (d + 2) % 3, 4, 5]

def q_of(W, g):
    q = W.copy(); q[4] = (g - 1.0) * W[0] * W[4]; return q

def va(sl, sr, e):
    return (sl * (sr * sr + e) + sr * (sl * sl + e)) / (sl * sl + sr * sr + 2 * e)

def va_d(sl, sr, e):
    sl2, sr2 = sl * sl, sr * sr
    den = sl2 + sr2 + 2 * e
    v = (sl * (sr2 + e) + sr * (sl2 + e)) / den
    return ((sr2 + e) + 2 * sl * sr - v * 2 * sl) / den, (2 * sl * sr + (sl2 + e) - v * 2 * sr) / den

def sh(a, d, s):
    """value at c+s along axis d (periodic)"""
    return np.roll(a, -s, axis=d + 1 if a.ndim == 4 else d)

def face_states(W, g, d, e):
    q = q_of(W, g)
    qs = [sh(q, d, s - 2) for s in range(4)]
    m = rot(d)
    qL = np.empty_like(q); qR = np.empty_like(q)
    for r in range(6):
        v = m[r]
        dl = va(qs[1][v] - qs[0][v], qs[2][v] - qs[1][v], e)
        dr = va(qs[2][v] - qs[1][v], qs[3][v] - qs[2][v], e)
        qL[r] = qs[1][v] + 0.5 * dl; qR[r] = qs[2][v] - 0.5 * dr
    return qL, qR

def side(a, gm1):
    rho, u, v, w, p, y = a
    m = rho * u; v2 = u*u + v*v + w*w; E = p / gm1 + 0.5 * rho * v2
    U = np.array([rho, m, rho * v, rho * w, E, rho * y])
    F = np.array([m, m * u + p, m * v, m * w, (E + p) * u, rho * y * u])
    return U, F

def side_vjp(a, gm1, Ub, Fb):
    rho, u, v, w, p, y = a
    m = rho * u; v2 = u*u + v*v + w*w; E = p / gm1 + 0.5 * rho * v2
    mb = Ub[1] + Fb[0] + Fb[1] * u + Fb[2] * v + Fb[3] * w
    Eb = Ub[4] + Fb[4] * u
    U5b = Ub[5] + Fb[5] * u
    ab = np.zeros_like(a)
    ab[0] = Ub[0] + Ub[2] * v + Ub[3] * w + U5b * y + mb * u + Eb * 0.5 * v2
    ab[1] = mb * rho + Fb[1] * m + Fb[4] * (E + p) + Fb[5] * rho * y + Eb * rho * u
    ab[2] = Ub[2] * rho + Fb[2] * m + Eb * rho * v
    ab[3] = Ub[3] * rho + Fb[3] * m + Eb * rho * w
    ab[4] = Fb[1] + Fb[4] * u + Eb / gm1
    ab[5] = U5b * rho
    return ab

def wmin(a, b): return np.where(a < b, 1.0, np.where(b < a, 0.0, 0.5))
def wmax(a, b): return np.where(a > b, 1.0, np.where(b > a, 0.0, 0.5))

def hlle_vjp(qL, qR, g, Fb):
    gm1 = g - 1
    cL = np.sqrt(g * qL[4] / qL[0]); cR = np.sqrt(g * qR[4] / qR[0])
    a1, a2 = qL[1] - cL, qR[1] - cR; b1, b2 = qL[1] + cL, qR[1] + cR
    SL = np.minimum(a1, a2); SR = np.maximum(b1, b2)
    bm = np.minimum(SL, 0); bp = np.maximum(SR, 0)
    UL, FL = side(qL, gm1); UR, FR = side(qR, gm1)
    bpm = bp * bm; den = bp - bm
    F = ((bp * FL - bm * FR) + bpm * (UR - UL)) / den
    nb = Fb / den
    denb = -(nb * F).sum(0); bpb = (nb * FL).sum(0); bmb = -(nb * FR).sum(0); bpmb = (nb * (UR - UL)).sum(0)
    FLb = nb * bp; FRb = -nb * bm; URb = nb * bpm; ULb = -nb * bpm
    bpb = bpb + bpmb * bm + denb; bmb = bmb + bpmb * bp - denb
    SRb = bpb * wmax(SR, 0); SLb = bmb * wmin(SL, 0)
    wa = wmin(a1, a2); wb = wmax(b1, b2)
    a1b, a2b = SLb * wa, SLb * (1 - wa); b1b, b2b = SRb * wb, SRb * (1 - wb)
    qLb = side_vjp(qL, gm1, ULb, FLb); qRb = side_vjp(qR, gm1, URb, FRb)
    qLb[1] += a1b + b1b; qRb[1] += a2b + b2b
    cLb = b1b - a1b; cRb = b2b - a2b
    qLb[4] += cLb * g / (2 * cL * qL[0]); qLb[0] -= cLb * cL / (2 * qL[0])
    qRb[4] += cRb * g / (2 * cR * qR[0]); qRb[0] -= cRb * cR / (2 * qR[0])
    return qLb, qRb

def jprim_t(w, Wb):
    rho, e, y = w[0], w[4], w[5]; ir = 1 / rho; v2 = w[1]**2 + w[2]**2 + w[3]**2
    Ub = np.empty_like(Wb)
    Ub[0] = Wb[0] + ir * (-(w[1]*Wb[1] + w[2]*Wb[2] + w[3]*Wb[3]) + (0.5*v2 - e)*Wb[4] - y*Wb[5])
    for k in range(1, 4): Ub[k] = ir * (Wb[k] - w[k] * Wb[4])
    Ub[4] = ir * Wb[4]; Ub[5] = ir * Wb[5]
    return Ub

def jcons_t(w, Ub):
    rho, e, y = w[0], w[4], w[5]; v2 = w[1]**2 + w[2]**2 + w[3]**2
    Wb = np.empty_like(Ub)
    Wb[0] = Ub[0] + w[1]*Ub[1] + w[2]*Ub[2] + w[3]*Ub[3] + (e + 0.5*v2)*Ub[4] + y*Ub[5]
    for k in range(1, 4): Wb[k] = rho * (Ub[k] + w[k] * Ub[4])
    Wb[4] = rho * Ub[4]; Wb[5] = rho * Ub[5]
    return Wb

def LT(W, Ub, coef, g, e, dx):
    """Jprim(W)^T dL(W)^T (coef Ub), via face G and per-cell gather."""
    vol = dx[0] * dx[1] * dx[2]; area = [dx[1]*dx[2], dx[0]*dx[2], dx[0]*dx[1]]
    q = q_of(W, g)
    qb = np.zeros_like(W)
    for d in range(3):
        m = rot(d)
        qL, qR = face_states(W, g, d, e)
        k = coef * area[d] / vol
        Fb = np.array([k * (Ub[m[r]] - sh(Ub, d, -1)[m[r]]) for r in range(6)])
        qLb, qRb = hlle_vjp(qL, qR, g, Fb)
        GL = np.empty_like(W); GR = np.empty_like(W)
        for r in range(6): GL[m[r]] = qLb[r]; GR[m[r]] = qRb[r]
        qs = [sh(q, d, s - 2) for s in range(5)]
        for v in range(6):
            gL_c1 = sh(GL[v], d, 1); gR_c0 = GR[v]
            acc = gL_c1 + gR_c0
            gDm = 0.5 * GL[v] - 0.5 * sh(GR[v], d, -1)
            gD0 = 0.5 * gL_c1 - 0.5 * gR_c0
            gDp = 0.5 * sh(GL[v], d, 2) - 0.5 * sh(GR[v], d, 1)
            dl, dr = va_d(qs[1][v] - qs[0][v], qs[2][v] - qs[1][v], e); acc = acc + dr * gDm
            dl, dr = va_d(qs[2][v] - qs[1][v], qs[3][v] - qs[2][v], e); acc = acc + (dl - dr) * gD0
            dl, dr = va_d(qs[3][v] - qs[2][v], qs[4][v] - qs[3][v], e); acc = acc - dl * gDp
            qb[v] += acc
    wb = qb.copy(); gm1 = g - 1
    wb[0] = qb[0] + gm1 * W[4] * qb[4]; wb[4] = gm1 * W[0] * qb[4]
    return jprim_t(W, wb)


def L_of(W, g, e, dx):
    """flux divergence L(W) of the scheme (array form)"""
    vol = dx[0] * dx[1] * dx[2]
    area = [dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1]]
    out = np.zeros_like(W)
    for d in range(3):
        m = rot(d)
        qL, qR = face_states(W, g, d, e)
        gm1 = g - 1
        cL = np.sqrt(g * qL[4] / qL[0])
        cR = np.sqrt(g * qR[4] / qR[0])
        SL = np.minimum(qL[1] - cL, qR[1] - cR)
        SR = np.maximum(qL[1] + cL, qR[1] + cR)
        bm = np.minimum(SL, 0)
        bp = np.maximum(SR, 0)
        UL, FL = side(qL, gm1)
        UR, FR = side(qR, gm1)
        Fr = ((bp * FL - bm * FR) + bp * bm * (UR - UL)) / (bp - bm)
        F = np.empty_like(W)
        for r in range(6):
            F[m[r]] = Fr[r]
        out += -area[d] * (sh(F, d, 1) - F) / vol
    return out


def cons(w):
    return np.array([w[0], w[0] * w[1], w[0] * w[2], w[0] * w[3],
                     w[0] * w[4] + 0.5 * w[0] * (w[1] ** 2 + w[2] ** 2 + w[3] ** 2), w[0] * w[5]])


def prim(u):
    return np.array([u[0], u[1] / u[0], u[2] / u[0], u[3] / u[0],
                     (u[4] - 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0]) / u[0], u[5] / u[0]])


def reverse(W0, gWT, g, e, dt, nsteps, dx):
    """Wbar_0 from Wbar_T for nsteps RK2 steps (all stage inputs kept)."""
    states, W1s = [], []
    S = W0.copy()
    for _ in range(nsteps):
        W1 = prim(cons(S) + dt * L_of(S, g, e, dx))
        states.append(S)
        W1s.append(W1)
        S = prim(0.5 * cons(S) + 0.5 * (cons(W1) + dt * L_of(W1, g, e, dx)))
    Ub2 = jprim_t(S, gWT)
    for k in reversed(range(nsteps)):
        Ub1 = 0.5 * Ub2 + LT(W1s[k], Ub2, 0.5 * dt, g, e, dx)
        Ub2 = 0.5 * Ub2 + Ub1 + LT(states[k], Ub1, dt, g, e, dx)
    return jcons_t(W0, Ub2), S
