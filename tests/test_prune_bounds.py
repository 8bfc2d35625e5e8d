"""Soundness of the robot->obstacle shadow prune (csrc/discrete.hip r2o_begin / r2o_visit).

The walk drops a link-BVH node box when the box, grown by delta (the per-component error bound of the
triangle's vertices in the link frame), lies outside one edge line of the triangle's shadow on the
plane of the two axes other than the normal's largest one.  Here the same float32 arithmetic runs in
numpy (one rounding per operation, as on the GPU; a fused multiply-add there only removes roundings)
on random triangles, slivers and boxes near the triangle, and every pruned box is checked against
the property the prune relies on: no point of the box lies in the shadow of any triangle whose
vertices are within delta of the float ones (corners and random points, vertices moved to the
corners of their delta squares and at random).  The same check with the box corner chosen wrongly or
without the growth by delta fails, so the property test can see those mistakes.
"""
import numpy as np

F = np.float32
EPS = F(5.9604645e-08)


def _pick(p, k):
    return np.where(k == 0, p[:, 0], np.where(k == 1, p[:, 1], p[:, 2])).astype(F)


def _setup(A, e1, e2):
    B, C = (A + e1).astype(F), (A + e2).astype(F)
    n = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], 1).astype(F)
    a = np.abs(n)
    ax = np.where((a[:, 0] >= a[:, 1]) & (a[:, 0] >= a[:, 2]), 0, np.where(a[:, 1] >= a[:, 2], 1, 2))
    iu, iv = np.where(ax == 2, 0, ax + 1), np.where(ax == 0, 2, ax - 1)
    Au, Av = _pick(A, iu), _pick(A, iv)
    p1 = ((_pick(B, iu) - Au) * (_pick(C, iv) - Av)).astype(F)
    p2 = ((_pick(B, iv) - Av) * (_pick(C, iu) - Au)).astype(F)
    area = (p1 - p2).astype(F)
    err = (F(5) * EPS * (np.abs(p1) + np.abs(p2))).astype(F)
    sg = np.where(area > err, F(1), np.where(area < -err, F(-1), F(0))).astype(F)
    return iu, iv, sg


def _misses(A, e1, e2, delta, iu, iv, sg, lo, hi, wrong_corner=False, no_growth=False):
    lu, lv, hu, hv = _pick(lo, iu), _pick(lo, iv), _pick(hi, iu), _pick(hi, iv)
    Au, Av = _pick(A, iu), _pick(A, iv)
    Bu, Bv = (Au + _pick(e1, iu)).astype(F), (Av + _pick(e1, iv)).astype(F)
    Cu, Cv = (Au + _pick(e2, iu)).astype(F), (Av + _pick(e2, iv)).astype(F)

    def outside(Pu, Pv, Qu, Qv):
        du, dv = ((Qu - Pu) * sg).astype(F), ((Qv - Pv) * sg).astype(F)
        if wrong_corner:
            xu = np.where(dv > 0, hu, lu)
        else:
            xu = np.where(dv > 0, lu, hu)
        xv = np.where(du > 0, hv, lv)
        t1, t2 = (du * (xv - Pv)).astype(F), (dv * (xu - Pu)).astype(F)
        m = (F(5) * EPS * (np.abs(t1) + np.abs(t2))).astype(F)
        if not no_growth:
            m = (m + F(1.01) * (np.abs(du) + np.abs(dv)) * delta).astype(F)
        return ((t1 - t2) + m).astype(F) < 0

    out = outside(Au, Av, Bu, Bv) | outside(Bu, Bv, Cu, Cv) | outside(Cu, Cv, Au, Av)
    return out & (sg != 0)


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    U = lambda *s: rng.uniform(-1, 1, s).astype(F)
    sc = (10.0 ** (U(n) * 3)).astype(F)[:, None]
    A = (U(n, 3) * F(0.3)).astype(F)
    e1 = (U(n, 3) * sc * F(0.1)).astype(F)
    e2 = (U(n, 3) * sc * F(0.1)).astype(F)
    sl = np.arange(n) % 3 == 0  # slivers: e2 almost parallel to e1
    e2[sl] = (e1[sl] * F(0.5) + U(int(sl.sum()), 3) * F(1e-7)).astype(F)
    delta = (np.abs(U(n)) * F(1e-6) + F(1e-30)).astype(F)
    c = (U(n, 3) * F(0.4)).astype(F)
    h = (np.abs(U(n, 3)) * np.where(np.arange(n) % 2 == 0, F(0.2), F(0.01))[:, None]).astype(F)
    near = np.arange(n) % 4 == 1  # boxes around a point of the triangle
    s = np.abs(U(n))
    t = (np.abs(U(n)) * (1 - s)).astype(F)
    p = (A + s[:, None] * e1 + t[:, None] * e2).astype(F)
    c[near] = p[near]
    h[near] = (np.abs(U(int(near.sum()), 3)) * F(1e-5)).astype(F)
    edge = np.arange(n) % 4 == 3  # small boxes within a few delta of a point of edge A-B
    s2 = np.abs(U(n))
    q = (A + s2[:, None] * e1 + U(n, 3) * F(3) * delta[:, None]).astype(F)
    c[edge] = q[edge]
    h[edge] = (np.abs(U(int(edge.sum()), 3)) * delta[edge][:, None]).astype(F)
    lo, hi = (c - h).astype(F), (c + h).astype(F)
    return rng, A, e1, e2, delta, lo, hi


def _violations(rng, A, e1, e2, delta, iu, iv, lo, hi, pruned):
    """Pruned boxes that hold a point of the shadow of a triangle within delta of the float one."""
    idx = np.nonzero(pruned)[0]
    A, e1, e2, delta, iu, iv, lo, hi = (x[idx] for x in (A, e1, e2, delta, iu, iv, lo, hi))
    B, C = (A + e1).astype(F), (A + e2).astype(F)
    V = [np.stack([_pick(P, iu), _pick(P, iv)], 1).astype(np.float64) for P in (A, B, C)]
    box_lo = np.stack([_pick(lo, iu), _pick(lo, iv)], 1).astype(np.float64)
    box_hi = np.stack([_pick(hi, iu), _pick(hi, iv)], 1).astype(np.float64)
    d = delta.astype(np.float64)[:, None]
    bad = np.zeros(len(idx), bool)
    for k in range(24):
        if k < 4:
            moves = [np.stack([np.full(len(idx), 1.0 if k & 1 else -1.0), np.full(len(idx), 1.0 if k & 2 else -1.0)], 1)
                     * d for _ in range(3)]
        else:
            moves = [rng.uniform(-1, 1, (len(idx), 2)) * d for _ in range(3)]
        Vm = [v + m for v, m in zip(V, moves)]
        if k < 4:
            q = np.where(np.array([k & 1, k & 2], bool)[None, :], box_hi, box_lo)
        else:
            q = box_lo + (box_hi - box_lo) * rng.uniform(0, 1, (len(idx), 2))

        def E(P, Q):
            return (Q[:, 0] - P[:, 0]) * (q[:, 1] - P[:, 1]) - (Q[:, 1] - P[:, 1]) * (q[:, 0] - P[:, 0])

        a, b, c = E(Vm[0], Vm[1]), E(Vm[1], Vm[2]), E(Vm[2], Vm[0])
        bad |= ((a >= 0) & (b >= 0) & (c >= 0)) | ((a <= 0) & (b <= 0) & (c <= 0))
    return int(bad.sum())


def test_shadow_prune_is_conservative():
    with np.errstate(over='ignore', invalid='ignore'):
        rng, A, e1, e2, delta, lo, hi = _cases(200_000, 7)
        iu, iv, sg = _setup(A, e1, e2)
        pruned = _misses(A, e1, e2, delta, iu, iv, sg, lo, hi)
        assert pruned.sum() > 50_000  # the prune does fire on most far boxes
        assert _violations(rng, A, e1, e2, delta, iu, iv, lo, hi, pruned) == 0


def test_shadow_prune_check_sees_planted_mistakes():
    with np.errstate(over='ignore', invalid='ignore'):
        rng, A, e1, e2, delta, lo, hi = _cases(200_000, 8)
        iu, iv, sg = _setup(A, e1, e2)
        for kw in ({'wrong_corner': True}, {'no_growth': True}):
            pruned = _misses(A, e1, e2, delta, iu, iv, sg, lo, hi, **kw)
            assert _violations(rng, A, e1, e2, delta, iu, iv, lo, hi, pruned) > 0, kw
