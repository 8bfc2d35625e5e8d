"""Soundness of the FP32 decisions of the discrete and swept checks (csrc/band32.hpp, csrc/swept.hip).

The device decides a test in FP32 only when every sign clears its error band (segment-triangle:
`link_tri` + `make_bands` + `mt32`) or when its certificate is sure (capsules: `cap_cert`; quadratic
and cubic sub-pieces: the chord capsule of `quad_cert` / `cubic_cert`); every other test goes to the
FP64 re-decision, which reproduces the oracle.  So each FP32 decision must equal the oracle's.

Here the same float32 arithmetic runs in numpy, operation by operation in the device's order (one
rounding per operation; a fused multiply-add on the device only removes roundings), with the hardware
reciprocal and square root taken within one ulp at random.  The cases are built to sit on the edge of
each decision: segments crossing a triangle near an edge or a vertex, ending near its plane or lying
nearly in it (moved to the world by random link transforms, as the FK gives them); segments passing a
capsule or a quadratic / cubic Bezier sub-piece at distance r(1 +- 1e-9 .. 1e-2).  The oracle decides
each case (orc_seg_tri on the FP64 transform; orc_tube_decide on the exact sub-piece).

Every FP32-decided case must agree with it.  The same cases are replayed with a bound weakened, each
a plausible slip, and those replays must produce wrong decisions, which shows that the cases reach the
bounds the device relies on: the input-error terms of the segment-triangle band left out, the transform
error delta left out, the capsule's evaluation-error term ef left out, a sub-piece's chord deviation
cut by a quarter.  The bounds are worst-case bounds and conservative on real rounding: on these cases a
segment-triangle band halved as a whole, its evaluation term alone, or the capsule's gradient term
alone still decide nothing wrongly, so those are not among the slips the test can see.
"""
import numpy as np
import pytest

from oracle import oracle as O

F = np.float32
EPS = F(2.0 ** -24)
RNG_SEED = 20261204


# ------------------------------------------------------------------------------ float32 helpers
def f32(x):
    return np.asarray(x, dtype=F)


def l1(v):
    return (np.abs(v[:, 0]) + np.abs(v[:, 1])) + np.abs(v[:, 2])


def linf(v):
    return np.maximum(np.maximum(np.abs(v[:, 0]), np.abs(v[:, 1])), np.abs(v[:, 2]))


def dot(a, b):
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def smul(s, v):
    return s[:, None] * v


def round_up(x):
    f = x.astype(F)
    return np.where(f.astype(np.float64) < x, np.nextafter(f, F(np.inf)), f)


def hw(x, rng):
    """A hardware reciprocal or square root result: within one ulp of the correctly rounded one."""
    k = rng.integers(-1, 2, x.shape)
    return np.where(k < 0, np.nextafter(x, F(-np.inf)), np.where(k > 0, np.nextafter(x, F(np.inf)), x)).astype(F)


def rcp(x, rng):
    with np.errstate(divide='ignore'):
        return hw((F(1) / x).astype(F), rng)


def hsqrt(x, rng):
    return hw(np.sqrt(x).astype(F), rng)


def rot_rows(T):
    """FP32 rows [R^T | t] of the FP64 transforms T [n, 3, 4] (as the pose kernel writes them)."""
    r = np.empty((len(T), 3, 4), F)
    r[:, :, :3] = np.transpose(T[:, :, :3], (0, 2, 1)).astype(F)
    r[:, :, 3] = T[:, :, 3].astype(F)
    return r


def rot(r, p):
    out = np.empty(p.shape, F)
    for j in range(3):
        out[:, j] = (r[:, j, 0] * p[:, 0] + r[:, j, 1] * p[:, 1]) + r[:, j, 2] * p[:, 2]
    return out


def random_rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def tiny(rng, n, lo=1e-9, hi=1e-3):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n)) * rng.choice([-1.0, 1.0], n)


# ------------------------------------------------------------------------------ segment vs triangle
def link_tri(rows, wA, wB, wC, mis):
    """csrc/discrete.hip link_tri: the world triangle in the link frame, with its error bound delta."""
    tinf = np.maximum(np.maximum(np.abs(rows[:, 0, 3]), np.abs(rows[:, 1, 3])), np.abs(rows[:, 2, 3]))
    rel = (wA - rows[:, :, 3]).astype(F)
    A = rot(rows, rel)
    e1 = rot(rows, (wB - wA).astype(F))
    e2 = rot(rows, (wC - wA).astype(F))
    en = np.maximum(l1(e1), l1(e2))
    delta = F(1.05) * EPS * ((tinf + F(5) * l1(rel)) + F(8) * en) + F(1e-30)
    return A, e1, e2, (delta * F(mis.get('delta', 1.0))).astype(F)


def tp_mag(Ma, ea, Mb, eb, Mc, ec):
    return ((Ma + F(3) * ea) * (Mb + F(3) * eb)) * (Mc + F(3) * ec)


def tp_band(Ma, ea, Mb, eb, Mc, ec, mis):
    ev = F(6) * EPS * tp_mag(Ma, ea, Mb, eb, Mc, ec) * F(mis.get('eval', 1.0))
    inp = ((ea * Mb) * Mc + (Ma * eb) * Mc) + (Ma * Mb) * ec
    inp = inp + F(3) * (((ea * eb) * Mc + (ea * Mb) * ec) + (Ma * eb) * ec) + F(9) * ((ea * eb) * ec)
    b = ev + inp * F(mis.get('input', 1.0))
    return F(1.25) * F(mis.get('margin', 1.0)) * b + F(1e-35)


def make_bands(D, ad, S, as_, E1, ae1, E2, ae2, mis):
    det = tp_band(E1, ae1, D, ad, E2, ae2, mis)
    u = tp_band(S, as_, D, ad, E2, ae2, mis)
    v = tp_band(D, ad, S, as_, E1, ae1, mis)
    t = tp_band(E2, ae2, S, as_, E1, ae1, mis)
    mdet, mu = tp_mag(E1, ae1, D, ad, E2, ae2), tp_mag(S, as_, D, ad, E2, ae2)
    mv, mt = tp_mag(D, ad, S, as_, E1, ae1), tp_mag(E2, ae2, S, as_, E1, ae1)
    k = F(1.25) * F(mis.get('margin', 1.0))
    w = k * (((det + u) + v) + F(2) * EPS * ((mdet + mu) + mv))
    y = k * ((det + t) + EPS * (mdet + mt))
    return det, u, v, t, w, y


def mt32(P, d, A, e1, e2, bd):
    """csrc/band32.hpp mt32: 0 miss, 1 hit, 2 unsure."""
    bdet, bu, bv, bt, bw, by = bd
    sx, sy, sz = P[:, 0] - A[:, 0], P[:, 1] - A[:, 1], P[:, 2] - A[:, 2]
    pvx = d[:, 1] * e2[:, 2] - d[:, 2] * e2[:, 1]
    pvy = d[:, 2] * e2[:, 0] - d[:, 0] * e2[:, 2]
    pvz = d[:, 0] * e2[:, 1] - d[:, 1] * e2[:, 0]
    det = (e1[:, 0] * pvx + e1[:, 1] * pvy) + e1[:, 2] * pvz
    u = (sx * pvx + sy * pvy) + sz * pvz
    qx = sy * e1[:, 2] - sz * e1[:, 1]
    qy = sz * e1[:, 0] - sx * e1[:, 2]
    qz = sx * e1[:, 1] - sy * e1[:, 0]
    v = (d[:, 0] * qx + d[:, 1] * qy) + d[:, 2] * qz
    t = (e2[:, 0] * qx + e2[:, 1] * qy) + e2[:, 2] * qz
    w = (det - u) - v
    y = det - t
    sg = np.where(det > 0, F(1), F(-1))
    u, v, t, w, y = u * sg, v * sg, t * sg, w * sg, y * sg
    miss = (u < -bu) | (v < -bv) | (w < -bw) | (t < -bt) | (y < -by)
    hit = (u > bu) & (v > bv) & (w > bw) & (t > bt) & (y > by)
    out = np.where(miss, 0, np.where(hit, 1, 2))
    return np.where(np.abs(det) <= bdet, 2, out)


def _crossing_cases(rng, n):
    """Link-frame triangles (float32 vertices) and float64 segments P, Q built to graze them."""
    ta = f32(rng.uniform(-0.12, 0.12, (n, 3)))
    L1 = np.exp(rng.uniform(np.log(0.004), np.log(0.12), n))
    L2 = np.exp(rng.uniform(np.log(0.004), np.log(0.12), n))
    u1 = unit(rng, n)
    u2 = unit(rng, n)
    u2 -= 0.9 * np.sum(u1 * u2, 1, keepdims=True) * u1   # not too close to collinear
    u2 /= np.linalg.norm(u2, axis=1, keepdims=True)
    tb = f32(ta + L1[:, None] * u1)
    tc = f32(ta + L2[:, None] * u2)
    A, B, C = ta.astype(np.float64), tb.astype(np.float64), tc.astype(np.float64)
    # target point: inside (0), near an edge (1), near a vertex (2)
    kind = rng.integers(0, 3, n)
    bar = rng.dirichlet([1, 1, 1], n)
    j = rng.integers(0, 3, n)
    e_small = tiny(rng, n)
    for i in np.where(kind == 1)[0]:
        bar[i, j[i]] = e_small[i]
        rest = 1.0 - bar[i, j[i]]
        k1, k2 = (j[i] + 1) % 3, (j[i] + 2) % 3
        s = rng.uniform()
        bar[i, k1], bar[i, k2] = rest * s, rest * (1 - s)
    for i in np.where(kind == 2)[0]:
        k1, k2 = (j[i] + 1) % 3, (j[i] + 2) % 3
        bar[i, k1], bar[i, k2] = e_small[i], tiny(rng, 1)[0]
        bar[i, j[i]] = 1.0 - bar[i, k1] - bar[i, k2]
    X = bar[:, 0:1] * A + bar[:, 1:2] * B + bar[:, 2:3] * C
    nrm = np.cross(B - A, C - A)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    # direction: generic (0) or nearly in the plane (1)
    dkind = rng.integers(0, 2, n)
    d = unit(rng, n)
    inplane = d - np.sum(d * nrm, 1, keepdims=True) * nrm
    inplane /= np.linalg.norm(inplane, axis=1, keepdims=True)
    d = np.where(dkind[:, None] == 1, inplane + tiny(rng, n, 1e-9, 1e-2)[:, None] * nrm, d)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # where the target lies along the segment: anywhere (0), near the start (1) or the end (2)
    skind = rng.integers(0, 3, n)
    s = np.where(skind == 0, rng.uniform(0.05, 0.95, n),
                 np.where(skind == 1, np.abs(tiny(rng, n)) * rng.choice([-1, 1], n),
                          1.0 + tiny(rng, n)))
    Ls = np.exp(rng.uniform(np.log(0.01), np.log(0.3), n))
    P = X - (s * Ls)[:, None] * d
    Q = P + Ls[:, None] * d
    return ta, tb, tc, P, Q


def _transforms(rng, n):
    T = np.empty((n, 3, 4))
    T[:, :, :3] = random_rotations(rng, n)
    T[:, :, 3] = rng.uniform(-1.2, 1.2, (n, 3))
    return T


def _xf(T, p):
    return np.einsum('nij,nj->ni', T[:, :, :3], p) + T[:, :, 3]


def _inv_xf(T, p):
    return np.einsum('nji,nj->ni', T[:, :, :3], p - T[:, :, 3])


def build_o2r_cases():
    """Obstacle->robot: the scene edge (world) vs a link triangle (link frame)."""
    rng = np.random.default_rng(RNG_SEED)
    n = 120000
    ta, tb, tc, P, Q = _crossing_cases(rng, n)
    T = _transforms(rng, n)
    rows = rot_rows(T)
    wA = f32(_xf(T, P))
    wB = f32(_xf(T, Q))
    wC = f32(_xf(T, P + rng.uniform(-0.1, 0.1, (n, 3))))
    # the oracle: edge A->B of the scene triangle moved into the link frame in FP64
    truth = O.seg_tri_batch(_inv_xf(T, wA.astype(np.float64)), _inv_xf(T, wB.astype(np.float64)), ta, tb, tc)
    return dict(rows=rows, wA=wA, wB=wB, wC=wC, ta=ta, tb=tb, tc=tc, truth=truth)


def o2r_decide(c, mis):
    A, e1, e2, delta = link_tri(c['rows'], c['wA'], c['wB'], c['wC'], mis)
    ta, tb, tc = c['ta'], c['tb'], c['tc']
    # the link is this one triangle: its box centre, vertex radius and edge bounds (tightest bands)
    lo = np.minimum(np.minimum(ta, tb), tc).astype(np.float64)
    hi = np.maximum(np.maximum(ta, tb), tc).astype(np.float64)
    cen = (0.5 * (lo + hi)).astype(F)
    rv = round_up(np.max(np.stack([l1((x.astype(np.float64) - cen)) for x in (ta, tb, tc)]), 0) * (1 + 1e-6))
    e1max = round_up(l1(tb.astype(np.float64) - ta) * (1 + 1e-6))
    e2max = round_up(l1(tc.astype(np.float64) - ta) * (1 + 1e-6))
    P, d = A, e1  # owned edge 0: A -> B
    D = l1(d) * (F(1) + F(8) * EPS)
    ad = F(6.1) * EPS * D
    S = (l1((P - cen).astype(F)) * (F(1) + F(2) * EPS) + F(3) * delta) + rv
    bd = make_bands(D, ad, S, delta + EPS * S, e1max, EPS * e1max, e2max, EPS * e2max, mis)
    return mt32(P, d, ta, (tb - ta).astype(F), (tc - ta).astype(F), bd)


def build_r2o_cases():
    """Robot->obstacle: a link edge (link frame) vs a scene triangle (world)."""
    rng = np.random.default_rng(RNG_SEED + 1)
    n = 120000
    ta, tb, tc, P, Q = _crossing_cases(rng, n)
    T = _transforms(rng, n)
    rows = rot_rows(T)
    # the triangle is the obstacle (moved to the world, rounded), the segment a robot edge (link frame)
    wA, wB, wC = (f32(_xf(T, x.astype(np.float64))) for x in (ta, tb, tc))
    p, q = f32(P), f32(Q)
    truth = O.seg_tri_batch(_xf(T, p.astype(np.float64)), _xf(T, q.astype(np.float64)), wA, wB, wC)
    return dict(rows=rows, wA=wA, wB=wB, wC=wC, p=p, q=q, truth=truth)


def r2o_decide(c, mis):
    A, e1, e2, delta = link_tri(c['rows'], c['wA'], c['wB'], c['wC'], mis)
    p, q = c['p'], c['q']
    cen = (0.5 * (np.minimum(p, q).astype(np.float64) + np.maximum(p, q))).astype(F)
    rv = round_up(np.maximum(l1(p.astype(np.float64) - cen), l1(q.astype(np.float64) - cen)) * (1 + 1e-6))
    dmax = round_up(l1(q.astype(np.float64) - p) * (1 + 1e-6))
    E1 = l1(e1) * (F(1) + F(8) * EPS)
    E2 = l1(e2) * (F(1) + F(8) * EPS)
    S = (l1((A - cen).astype(F)) * (F(1) + F(2) * EPS) + F(3) * delta) + rv
    bd = make_bands(dmax, EPS * dmax, S, delta + EPS * S, E1, F(6.1) * EPS * E1, E2, F(6.1) * EPS * E2, mis)
    return mt32(p, (q - p).astype(F), A, e1, e2, bd)


@pytest.fixture(scope='module')
def o2r_cases():
    return build_o2r_cases()


@pytest.fixture(scope='module')
def r2o_cases():
    return build_r2o_cases()


def _wrong(dec, truth):
    decided = dec != 2
    return int(np.sum(decided & (dec != truth))), int(decided.sum())


@pytest.mark.parametrize('direction', ['o2r', 'r2o'])
def test_segment_triangle_bands_are_sound(direction, o2r_cases, r2o_cases):
    c, decide = (o2r_cases, o2r_decide) if direction == 'o2r' else (r2o_cases, r2o_decide)
    dec = decide(c, {})
    wrong, decided = _wrong(dec, c['truth'])
    unsure = int(np.sum(dec == 2))
    print(f'{direction}: {decided} decided in FP32 ({int(np.sum(dec == 1))} hits), {unsure} to FP64, {wrong} wrong')
    assert wrong == 0
    assert unsure > 0 and decided >= len(dec) // 10   # the cases reach the band; many are still decided
    assert 0.1 < np.mean(c['truth']) < 0.9


@pytest.mark.parametrize('mistake', [{'input': 0.0}, {'delta': 0.0}])
@pytest.mark.parametrize('direction', ['o2r', 'r2o'])
def test_segment_triangle_planted_mistakes_are_caught(direction, mistake, o2r_cases, r2o_cases):
    c, decide = (o2r_cases, o2r_decide) if direction == 'o2r' else (r2o_cases, r2o_decide)
    wrong, _ = _wrong(decide(c, mistake), c['truth'])
    print(f'{direction} {mistake}: {wrong} wrong decisions')
    assert wrong > 0


# ------------------------------------------------------------------------------ swept certificates
def cap_cert(C0, C1, A, B, r_hit, r_miss, ein, rng, mis):
    """csrc/swept.hip cap_cert: 0 miss, 1 hit, 2 unsure (vectorised)."""
    E, D, w0 = (C1 - C0).astype(F), (B - A).astype(F), (C0 - A).astype(F)
    a, b, c, d, e = dot(E, E), dot(E, D), dot(D, D), dot(E, w0), dot(D, w0)
    zero, one = F(0), F(1)
    den = a * c - b * b
    ra = np.where(a > 0, rcp(np.where(a > 0, a, one), rng), zero)
    clamp = lambda x: np.minimum(np.maximum(x, zero), one)
    ok = (den > F(1e-12) * a * c) & (c > 0)
    t = np.where(ok, clamp((b * e - c * d) * rcp(np.where(ok, den, one), rng)), zero)
    s = np.where(c > 0, (b * t + e) * rcp(np.where(c > 0, c, one), rng), zero)
    t_lo = np.where(a > 0, clamp(-d * ra), zero)
    t_hi = np.where(a > 0, clamp((b - d) * ra), zero)
    t = np.where(c > 0, np.where(s < 0, t_lo, np.where(s > 1, t_hi, t)), t_lo)
    s = np.where(c > 0, np.where(s < 0, zero, np.where(s > 1, one, s)), zero)
    v = ((w0 + smul(t, E)) - smul(s, D)).astype(F)
    f = dot(v, v)
    gt, gs = F(2) * dot(E, v), F(-2) * dot(D, v)
    corr = np.where(gt > 0, -gt * t, gt * (one - t)) + np.where(gs > 0, -gs * s, gs * (one - s))
    K = (linf(w0) + linf(E)) + linf(D)
    ev = F(1.01) * (F(3) * ein + F(6) * EPS * K) * F(mis.get('ev', 1.0))
    Vn = (np.abs(v[:, 0]) + np.abs(v[:, 1]) + np.abs(v[:, 2])) + F(3) * ev
    ef = ((F(2) * Vn) * ev + (F(3) * ev) * ev + ((F(4) * EPS) * Vn) * Vn) * F(mis.get('ef', 1.0))
    En = l1(E) + F(6) * ein
    eE = F(2) * ein + EPS * En
    Dn = l1(D)
    eD = EPS * Dn
    eg = (F(2) * ((En * ev + Vn * eE) + ((F(4) * EPS) * En) * Vn) +
          F(2) * ((Dn * ev + Vn * eD) + ((F(4) * EPS) * Dn) * Vn) + (F(2) * EPS) * (np.abs(gt) + np.abs(gs)))
    eg = eg * F(mis.get('eg', 1.0))
    ub = f + F(1.25) * ef
    lb = ((f + corr) - F(1.25) * (ef + eg)) - (F(4) * EPS) * (f + np.abs(corr))
    hit = (r_hit > 0) & (ub < (r_hit * r_hit) * (F(1) - F(4) * EPS))
    miss = lb > (r_miss * r_miss) * (F(1) + F(4) * EPS)
    return np.where(hit, 1, np.where(miss, 0, 2))


def bez(q, t):
    """Point of the Bezier with controls q [n, deg + 1, 3] at parameters t [n] (float64)."""
    deg = q.shape[1] - 1
    from math import comb
    return sum(comb(deg, i) * ((1 - t) ** (deg - i) * t ** i)[:, None] * q[:, i] for i in range(deg + 1))


def bez_d(q, t):
    deg = q.shape[1] - 1
    dq = deg * (q[:, 1:] - q[:, :-1])
    return bez(dq, t) if deg > 1 else dq[:, 0]


def _tube_cases(rng, n, deg):
    """Exact controls q64 [n, deg + 1, 3], a dyadic sub-piece (depth, k), a radius and a segment passing
    the sub-curve at distance r (1 +- 1e-9 .. 1e-2), across (0) or pointing away (1)."""
    base = rng.uniform(-1.0, 1.0, (n, 1, 3))
    q64 = base + rng.uniform(-0.25, 0.25, (n, deg + 1, 3)) * np.exp(rng.uniform(np.log(0.05), 0, (n, 1, 1)))
    depth = rng.integers(0, 9, n) if deg > 1 else np.zeros(n, int)
    k = np.floor(rng.uniform(0, 1, n) * 2.0 ** depth)
    t0, t1 = k / 2.0 ** depth, (k + 1) / 2.0 ** depth
    ts = t0 + rng.uniform(0.02, 0.98, n) * (t1 - t0)
    X = bez(q64, ts)
    Tg = bez_d(q64, ts)
    Tg /= np.maximum(np.linalg.norm(Tg, axis=1, keepdims=True), 1e-300)
    nv = unit(rng, n)
    nv -= np.sum(nv * Tg, 1, keepdims=True) * Tg
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    w = np.cross(Tg, nv)
    r = rng.uniform(0.02, 0.09, n)
    rho = r * (1 + tiny(rng, n, 1e-9, 1e-2))
    Y = X + rho[:, None] * nv
    kind = rng.integers(0, 2, n)
    La, Lb = rng.uniform(0.01, 0.3, n), rng.uniform(0.01, 0.3, n)
    A = np.where(kind[:, None] == 0, Y - La[:, None] * w, Y)
    B = np.where(kind[:, None] == 0, Y + Lb[:, None] * w, Y + Lb[:, None] * nv)
    return q64, depth, t0, t1, f32(r), f32(A), f32(B)


def _controls32(rng, q64, ein_rel):
    """FP32 controls within ein = ein_rel * eps * M of the exact ones (perturbed, then rounded)."""
    M = np.max(np.abs(q64), axis=(1, 2))
    pert = rng.uniform(-1, 1, q64.shape) * ((ein_rel - 1) * float(EPS) * M)[:, None, None]
    q32 = f32(q64 + pert)
    return q32, (F(ein_rel) * EPS * np.max(np.abs(q32), axis=(1, 2)).astype(F)).astype(F)


def blossom2(q, u, v):
    iu, iv = F(1) - u, F(1) - v
    w0, w1, w2 = iu * iv, iu * v + u * iv, u * v
    return ((w0[:, None] * q[:, 0] + w1[:, None] * q[:, 1]) + w2[:, None] * q[:, 2]).astype(F)


def blossom3(q, u, v, w):
    iu, iv, iw = F(1) - u, F(1) - v, F(1) - w
    w0 = (iu * iv) * iw
    w1 = ((u * iv) * iw + (iu * v) * iw) + (iu * iv) * w
    w2 = ((u * v) * iw + (u * iv) * w) + (iu * v) * w
    w3 = (u * v) * w
    return (((w0[:, None] * q[:, 0] + w1[:, None] * q[:, 1]) + w2[:, None] * q[:, 2]) + w3[:, None] * q[:, 3]).astype(F)


def sub_controls64(q64, t0, t1):
    """Exact (FP64) Bezier controls of the sub-piece [t0, t1] by the blossom."""
    deg = q64.shape[1] - 1
    if deg == 1:
        return np.stack([bez(q64, t0), bez(q64, t1)], 1)
    if deg == 2:
        bl = lambda u, v: ((1 - u) * (1 - v))[:, None] * q64[:, 0] + ((1 - u) * v + u * (1 - v))[:, None] * q64[:, 1] + \
            (u * v)[:, None] * q64[:, 2]
        return np.stack([bl(t0, t0), bl(t0, t1), bl(t1, t1)], 1)

    def bl3(u, v, w):
        iu, iv, iw = 1 - u, 1 - v, 1 - w
        return ((iu * iv * iw)[:, None] * q64[:, 0] + (u * iv * iw + iu * v * iw + iu * iv * w)[:, None] * q64[:, 1] +
                (u * v * iw + u * iv * w + iu * v * w)[:, None] * q64[:, 2] + (u * v * w)[:, None] * q64[:, 3])
    return np.stack([bl3(t0, t0, t0), bl3(t0, t0, t1), bl3(t0, t1, t1), bl3(t1, t1, t1)], 1)


def tube_decide(c, deg, mis):
    rng = np.random.default_rng(c['hw_seed'])
    q32, ein, r, A, B = c['q32'], c['ein'], c['r'], c['A'], c['B']
    if deg == 1:
        return cap_cert(q32[:, 0], q32[:, 1], A, B, r, r, ein, rng, mis)
    t0, t1 = f32(c['t0']), f32(c['t1'])
    M = np.max(np.abs(q32), axis=(1, 2)).astype(F)
    if deg == 2:
        e2 = ein + F(8) * EPS * M
        pa, pm, pb = blossom2(q32, t0, t0), blossom2(q32, t0, t1), blossom2(q32, t1, t1)
        dev = (pm - F(0.5) * (pa + pb)).astype(F)
        h = (F(0.5) * hsqrt(dot(dev, dev), rng)) * (F(1) + F(8) * EPS) * F(mis.get('dev', 1.0)) + \
            F(2) * F(1.7321) * e2 * F(mis.get('e2', 1.0))
        return cap_cert(pa, pb, A, B, r - h, r + h, e2, rng, mis)
    e2 = ein + F(12) * EPS * M
    pa, pb = blossom3(q32, t0, t0, t0), blossom3(q32, t0, t0, t1)
    pc, pd = blossom3(q32, t0, t1, t1), blossom3(q32, t1, t1, t1)
    third = F(1) / F(3)
    d1 = (pb - smul(np.full(len(pa), third, F), (F(2) * pa + pd))).astype(F)
    d2 = (pc - smul(np.full(len(pa), third, F), (pa + F(2) * pd))).astype(F)
    h = (F(0.75) * F(mis.get('dev', 1.0)) * hsqrt(np.maximum(dot(d1, d1), dot(d2, d2)), rng)) * (F(1) + F(8) * EPS) + \
        F(4) * F(1.7321) * e2 * F(mis.get('e2', 1.0))
    return cap_cert(pa, pd, A, B, r - h, r + h, e2, rng, mis)


def build_tube_cases(deg):
    rng = np.random.default_rng(RNG_SEED + 10 + deg)
    n = 60000
    q64, depth, t0, t1, r, A, B = _tube_cases(rng, n, deg)
    q32, ein = _controls32(rng, q64, 3.0)
    truth = O.tube_decide_batch(sub_controls64(q64, t0, t1), r.astype(np.float64), A, B)
    return deg, dict(q32=q32, ein=ein, r=r, A=A, B=B, t0=t0, t1=t1, truth=truth, hw_seed=int(rng.integers(1 << 30)))


@pytest.fixture(scope='module', params=[1, 2, 3], ids=['capsule', 'quadratic', 'cubic'])
def tube_cases(request):
    return build_tube_cases(request.param)


def test_swept_certificates_are_sound(tube_cases):
    deg, c = tube_cases
    dec = tube_decide(c, deg, {})
    wrong, decided = _wrong(dec, c['truth'])
    unsure = int(np.sum(dec == 2))
    print(f'degree {deg}: {decided} decided in FP32 ({int(np.sum(dec == 1))} hits), {unsure} to FP64, {wrong} wrong')
    assert wrong == 0
    assert unsure > 0 and decided >= len(dec) // 10
    assert 0.1 < np.mean(c['truth']) < 0.9


TUBE_MISTAKES = {1: [{'ef': 0.0}],
                 2: [{'dev': 0.75}],
                 3: [{'dev': 0.75}]}


def test_swept_planted_mistakes_are_caught(tube_cases):
    deg, c = tube_cases
    for mistake in TUBE_MISTAKES[deg]:
        wrong, _ = _wrong(tube_decide(c, deg, mistake), c['truth'])
        print(f'degree {deg} {mistake}: {wrong} wrong decisions')
        assert wrong > 0, mistake
