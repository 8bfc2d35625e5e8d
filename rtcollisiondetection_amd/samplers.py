"""Pose and trajectory generators (numpy PCG64, ``numpy.random.default_rng(seed)``).

The reference samples Halton poses and KNN trajectories for its benchmarks
(RTCD/Planner/sampler/sampler.h:69-82, RTCD/Utils/Test/poseUtils.h:115-141): ``halton_reference``,
``benchmark_poses_discrete`` and ``knn_trajectories`` restate them; the headline workloads use uniform
joint samples, bounded random walks and clamped quadratic B-splines.
"""
import numpy as np


def uniform_poses(n, lower, upper, seed=0):
    """n joint configurations uniform within [lower, upper], float32 [n, dof]."""
    rng = np.random.default_rng(seed)
    lower = np.asarray(lower, dtype=np.float64)
    upper = np.asarray(upper, dtype=np.float64)
    q = lower + (upper - lower) * rng.random((n, len(lower)))
    return np.ascontiguousarray(q.astype(np.float32))


def halton(n, dof, skip=0):
    """Halton points in [0,1)^dof with the first dof primes as bases (radical inverse)."""
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47][:dof]
    out = np.zeros((n, dof))
    for d, b in enumerate(primes):
        for i in range(n):
            k = i + 1 + skip
            f, r = 1.0, 0.0
            while k > 0:
                f /= b
                r += f * (k % b)
                k //= b
            out[i, d] = r
    return out


_HALTON_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97)


def halton_reference(n, lo, hi, skip=0):
    """RTCD's Halton poses, float32 [n, dof]: sample i, joint d is h(i, b) * (hi[d] - lo[d]) + lo[d] with
    b the (d + skip)-th prime and h the radical inverse computed in float32 exactly as
    localHaltonSingleNumber does (RTCD/Planner/sampler/sampler.h:11-24, createSamplesHalton 69-82):
    integer quotient stored in a float, digit n0 - n1 * b in float, f = 1/b divided by b per digit.
    Sample 0 is lo."""
    lo = np.asarray(lo, dtype=np.float32)
    hi = np.asarray(hi, dtype=np.float32)
    dof = len(lo)
    if skip + dof >= len(_HALTON_BASES):
        raise ValueError('skip + dof must stay below 25')
    out = np.zeros((n, dof), dtype=np.float32)
    for d in range(dof):
        b = _HALTON_BASES[d + skip]
        fb = np.float32(b)
        n0 = np.arange(n, dtype=np.int64)
        hn = np.zeros(n, dtype=np.float32)
        f = np.full(n, np.float32(1) / fb, dtype=np.float32)
        while np.any(n0 > 0):
            act = n0 > 0
            n1 = (n0 // b).astype(np.float32)                       # float n1 = n0 / b (integer division)
            r = (n0.astype(np.float32) - n1 * fb).astype(np.int64)  # int r = n0 - n1 * b (float arithmetic)
            hn = np.where(act, hn + f * r.astype(np.float32), hn).astype(np.float32)
            f = np.where(act, f / fb, f).astype(np.float32)
            n0 = np.where(act, n1.astype(np.int64), n0)
        out[:, d] = hn * (hi[d] - lo[d]) + lo[d]
    return out


def benchmark_poses_discrete(n, lower, upper):
    """The pose set of RTCD's discrete benchmark (RTCD/Benchmark/Discrete/benchmark.cpp:56-57): its
    randomMotions call passes (lowerBound, upperBound) into parameters named (upper, lower)
    (RTCD/Utils/Test/poseUtils.h:16-38), so the Halton samples run from the upper limits down:
    q = h * (lower - upper) + upper."""
    return halton_reference(n, upper, lower)


def knn_trajectories(points, k, n_traj_pts, weight=None):
    """RTCD's KNN trajectories (RTCD/Utils/Test/poseUtils.h:115-141 genCSpaceKNNTraj): for every
    configuration i and each of its k nearest configurations j (weighted squared L2 over the joints,
    accumulated in float32 joint by joint as RTCD/Planner/KNN/knnCUDA.cu:7-76 does, ties to the lower
    index, i itself first at distance 0), the straight line start + span * s, span = (end - start) /
    (n_traj_pts - 1), s = 0..n_traj_pts-1, in float32.  Returns float32 [len(points) * k, n_traj_pts, dof]."""
    P = np.ascontiguousarray(points, dtype=np.float32)
    n, dof = P.shape
    w = np.ones(dof, np.float32) if weight is None else np.asarray(weight, dtype=np.float32)
    w2 = w * w
    idx = np.zeros((n, k), dtype=np.int64)
    for i0 in range(0, n, 512):
        q = P[i0:i0 + 512]
        ssd = np.zeros((len(q), n), dtype=np.float32)
        for d in range(dof):
            t = (P[None, :, d] - q[:, None, d]).astype(np.float32)
            ssd = (ssd + t * t * w2[d]).astype(np.float32)
        idx[i0:i0 + len(q)] = np.argsort(ssd, axis=1, kind='stable')[:, :k]
    start = P[:, None, :]
    end = P[idx]
    span = ((end - start) / np.float32(n_traj_pts - 1)).astype(np.float32)
    s = np.arange(n_traj_pts, dtype=np.float32)
    traj = (start[:, :, None, :] + span[:, :, None, :] * s[None, None, :, None]).astype(np.float32)
    return np.ascontiguousarray(traj.reshape(n * k, n_traj_pts, dof))


def random_walk_trajectories(n, lower, upper, lengths=32, max_step=0.1, seed=4):
    """Random walks in joint space: per step each joint moves U[-max_step, max_step], clipped to the
    limits.  ``lengths`` is an int (fixed) or a (lo, hi) pair for U{lo..hi} waypoint counts.
    Returns (waypoints float32 [sum L, dof], offsets uint32 [n+1])."""
    rng = np.random.default_rng(seed)
    lower = np.asarray(lower, dtype=np.float64)
    upper = np.asarray(upper, dtype=np.float64)
    dof = len(lower)
    if isinstance(lengths, (tuple, list)):
        L = rng.integers(lengths[0], lengths[1] + 1, size=n)
    else:
        L = np.full(n, int(lengths))
    offs = np.zeros(n + 1, dtype=np.uint32)
    offs[1:] = np.cumsum(L)
    Lmax = int(L.max()) if n else 0
    q = lower + (upper - lower) * rng.random((n, dof))
    steps = rng.uniform(-max_step, max_step, size=(max(Lmax - 1, 0), n, dof))
    walk = np.empty((Lmax, n, dof))
    if Lmax:
        walk[0] = q
    for k in range(1, Lmax):
        q = np.clip(q + steps[k - 1], lower, upper)
        walk[k] = q
    # gather the first L_i waypoints of walk i, trajectories concatenated
    idx_t = np.repeat(np.arange(n), L)
    idx_k = np.arange(int(offs[-1])) - np.repeat(offs[:-1].astype(np.int64), L)
    out = walk[idx_k, idx_t]
    return np.ascontiguousarray(out.astype(np.float32)), offs


def clamped_quadratic_basis(n_ctrl, n_samples):
    """Basis matrix [n_samples, n_ctrl] of the clamped uniform quadratic B-spline (knots
    0,0,0,1,..,m,m,m with m = n_ctrl - 2) at n_samples parameters uniform on [0, m] (de Boor)."""
    eye = np.eye(n_ctrl)
    return clamped_quadratic_bspline(eye, n_samples)


def clamped_quadratic_bspline(ctrl, n_samples):
    """Evaluate a clamped uniform quadratic B-spline (knots 0,0,0,1,..,m,m,m with m = nC-2) at
    n_samples parameters uniform on [0, m].  ctrl [nC, dim] -> [n_samples, dim]."""
    ctrl = np.asarray(ctrl, dtype=np.float64)
    nC = len(ctrl)
    m = nC - 2
    knots = np.concatenate([[0, 0], np.arange(0, m + 1), [m, m]]).astype(np.float64)
    us = np.linspace(0.0, m, n_samples)
    out = []
    for u in us:
        k = min(int(np.floor(u)), m - 1) + 2  # knot span: knots[k] <= u < knots[k+1]
        d = [ctrl[j] for j in range(k - 2, k + 1)]
        for r in range(1, 3):  # de Boor
            for j in range(2, r - 1, -1):
                i = j + k - 2
                a = (u - knots[i]) / (knots[i + 3 - r] - knots[i])
                d[j] = (1 - a) * d[j - 1] + a * d[j]
        out.append(d[2])
    return np.asarray(out)


def bspline_trajectories(n, lower, upper, n_ctrl=8, n_wp=16, seed=3):
    """n clamped quadratic B-splines with n_ctrl control points uniform within the limits, sampled at
    n_wp waypoints.  Returns (waypoints float32 [n, n_wp, dof], control points float32 [n, n_ctrl, dof])."""
    rng = np.random.default_rng(seed)
    lower = np.asarray(lower, dtype=np.float64)
    upper = np.asarray(upper, dtype=np.float64)
    ctrl = lower + (upper - lower) * rng.random((n, n_ctrl, len(lower)))
    B = clamped_quadratic_basis(n_ctrl, n_wp)
    wp = np.einsum('wc,ncd->nwd', B, ctrl)
    return np.ascontiguousarray(wp.astype(np.float32)), np.ascontiguousarray(ctrl.astype(np.float32))


def quadratic_fit_weights(n_ctrl, n_wp):
    """Least-squares fit W = pinv(B) (float64 [n_ctrl, n_wp]) of a uniform quadratic B-spline with
    n_ctrl control points to n_wp samples at u = linspace(0, n_ctrl - 2, n_wp).  The reference ships
    these matrices as RTCD/data/weights/quad/normal/{nC}_{nT}.txt (loaded at
    RTCD/Robot/batchCurveRobot.h:560-605)."""
    nseg = n_ctrl - 2
    u = np.linspace(0.0, nseg, n_wp)
    B = np.zeros((n_wp, n_ctrl))
    for i, uu in enumerate(u):
        s = min(int(np.floor(uu)), nseg - 1)
        t = uu - s
        B[i, s] = (1 - t) ** 2 / 2
        B[i, s + 1] = (-2 * t * t + 2 * t + 1) / 2
        B[i, s + 2] = t * t / 2
    return np.ascontiguousarray(np.linalg.pinv(B))
