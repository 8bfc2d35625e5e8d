"""CPU-side checks of the product library: it loads, exports every C-ABI symbol that include/rtcd.h
declares, host-only entry points work, and error paths report status codes without a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, 'tests', 'golden')


def _declared():
    src = open(os.path.join(ROOT, 'include', 'rtcd.h')).read()
    return sorted(set(re.findall(r'\b(rtcd_[a-z_0-9]+)\s*\(', src)))


def test_library_exports_every_declared_symbol():
    from rtcollisiondetection_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


@pytest.mark.parametrize('nc,nt', [(3, 6), (4, 8), (8, 16), (8, 32), (16, 32)])
def test_quadratic_weights_match_reference_files(nc, nt):
    from rtcollisiondetection_amd.detector import quadratic_weights
    W = quadratic_weights(nc, nt)
    ref = np.loadtxt(os.path.join(GOLD, f'weights_quad_normal_{nc}_{nt}.txt')).reshape(nc, nt)
    np.testing.assert_allclose(W, ref, atol=2e-6)


@pytest.mark.parametrize('kind,end,nc,nt', [('quad', 'repeated', 4, 8), ('quad', 'repeated', 8, 16),
                                            ('cubic', 'normal', 4, 16), ('cubic', 'normal', 16, 64),
                                            ('cubic', 'repeated', 4, 16), ('cubic', 'repeated', 8, 32),
                                            ('quad', 'normal', 8, 16)])
def test_curve_weights_match_reference_files(kind, end, nc, nt):
    from rtcollisiondetection_amd.detector import curve_weights
    W = curve_weights(2 if kind == 'quad' else 3, end, nc, nt)
    ref = np.loadtxt(os.path.join(GOLD, f'weights_{kind}_{end}_{nc}_{nt}.txt')).reshape(nc, nt)
    np.testing.assert_allclose(W, ref, atol=2e-6)


@pytest.mark.parametrize('degree', [2, 3])
@pytest.mark.parametrize('end', ['normal', 'repeated'])
def test_curve_weights_invariants(degree, end):
    """W is a left inverse of the basis (W B = I: a fit reproduces any curve of the family exactly), its
    rows sum to 1 (the basis is a partition of unity), and clamped curves start and end on their end
    control points; the basis here is evaluated independently of the library (closed forms per span)."""
    from rtcollisiondetection_amd.detector import curve_weights
    nc, nt = 7, 29
    W = curve_weights(degree, end, nc, nt)
    m = nc - degree
    us = np.linspace(0.0, m, nt)
    B = np.zeros((nt, nc))
    for r, u in enumerate(us):
        if end == 'normal':
            s = min(int(np.floor(u)), m - 1)
            t = u - s
            if degree == 2:
                w = [(1 - t) ** 2 / 2, (-2 * t * t + 2 * t + 1) / 2, t * t / 2]
            else:
                w = [(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t * t + 4) / 6, (-3 * t ** 3 + 3 * t * t + 3 * t + 1) / 6, t ** 3 / 6]
            B[r, s:s + degree + 1] = w
        else:
            # clamped: de Casteljau-free check through the Cox-de Boor definition on explicit knots
            knots = np.concatenate([np.zeros(degree), np.arange(m + 1), np.full(degree, m)]).astype(float)
            def N(i, p, x):
                if p == 0:
                    last = knots[i + 1] == m and x == m and knots[i] < m
                    return 1.0 if (knots[i] <= x < knots[i + 1]) or last else 0.0
                a = 0.0 if knots[i + p] == knots[i] else (x - knots[i]) / (knots[i + p] - knots[i]) * N(i, p - 1, x)
                b = 0.0 if knots[i + p + 1] == knots[i + 1] else (knots[i + p + 1] - x) / (knots[i + p + 1] - knots[i + 1]) * N(i + 1, p - 1, x)
                return a + b
            B[r] = [N(i, degree, u) for i in range(nc)]
    np.testing.assert_allclose(B.sum(axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(W @ B, np.eye(nc), atol=1e-9)
    np.testing.assert_allclose(W.sum(axis=1), 1.0, atol=1e-9)
    if end == 'repeated':
        assert B[0, 0] == 1.0 and B[-1, -1] == 1.0


def test_curve_weights_invalid():
    from rtcollisiondetection_amd import _lib
    from rtcollisiondetection_amd.detector import curve_weights
    for args in ((4, 'normal', 8, 16), (2, 'normal', 8, 4), (3, 'repeated', 3, 16)):
        with pytest.raises(_lib.RtcdError) as e:
            curve_weights(*args)
        assert e.value.status == 1


def test_quadratic_weights_invalid():
    from rtcollisiondetection_amd import _lib
    from rtcollisiondetection_amd.detector import quadratic_weights
    with pytest.raises(_lib.RtcdError) as e:
        quadratic_weights(8, 4)
    assert e.value.status == 1


def test_create_reports_errors_without_crashing():
    from rtcollisiondetection_amd import _lib
    from rtcollisiondetection_amd.detector import Context
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        with pytest.raises(_lib.RtcdError) as e:
            Context(device=999)
        assert e.value.status == 2
    else:
        with pytest.raises(_lib.RtcdError) as e:
            Context()
        assert e.value.status == 2
    L = _lib.lib()
    assert L.rtcd_check_discrete(None, 0, 3, None, 0, None) == 1
    assert L.rtcd_sync(None, None) == 1


def test_cpp_facade_header_compiles():
    """include/rtcd_detector.hpp and its check program compile against include/rtcd.h (no GPU needed)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run(['g++', '-std=c++17', '-fsyntax-only', '-Wall', '-Werror', '-I' + os.path.join(root, 'include'),
                        os.path.join(root, 'tests', 'cpp', 'facade_check.cpp')], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_reference_bridge_compiles_against_rtcd_standins():
    """include/rtcd_bridge.hpp, the code a maintainer adds to RTCD, compiles against stand-ins that declare
    only the members of RTCD's types it cites (tests/cpp/rtcd_standins.hpp): robotConfig<DOF>,
    robot<DOF>::getOrderedLinks, meshModel::mesh / sphereOrigins / sphereRadii, obstacle, and the robot
    wrapper's getPoseBufferPtr / getBatchSize / getDOF.  Its check program (which runs it on the GPU)
    instantiates every template of the bridge."""
    root = ROOT
    r = subprocess.run(['g++', '-std=c++17', '-fsyntax-only', '-Wall', '-Werror', '-D__HIP_PLATFORM_AMD__',
                        '-I/opt/rocm/include', '-I' + os.path.join(root, 'include'),
                        '-I' + os.path.join(root, 'tests', 'cpp'), os.path.join(root, 'tests', 'cpp', 'bridge_check.cpp')],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # the stand-ins declare nothing beyond what the bridge reads: a member the bridge does not use would
    # make the compile above prove less
    src = open(os.path.join(root, 'tests', 'cpp', 'rtcd_standins.hpp')).read()
    for name in ('baseTransform', 'cosMask', 'sinMask', 'oneMask', 'getDof', 'getOrderedLinks', 'sphereOrigins',
                 'sphereRadii', 'vertices', 'indices', 'getPoseBufferPtr', 'getBatchSize', 'getDOF'):
        assert name in open(os.path.join(root, 'include', 'rtcd_bridge.hpp')).read(), name
        assert name in src, name
