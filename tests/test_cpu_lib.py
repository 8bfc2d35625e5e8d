"""CPU-side checks of the product library: it loads, exports every C-ABI symbol that include/rtcd.h
declares, host-only entry points work, and error paths report status codes without a GPU."""
import ctypes as C
import os
import re

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
