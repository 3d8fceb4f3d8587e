"""CPU checks of the C ABI: the library builds, loads and exports every
entry point include/prism_amd.h declares, and its host-side geometry equals
the oracle's.  No kernel is launched here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "prism_amd.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_table():
    from prism_amd import _lib
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_library_exports_every_symbol():
    import prism_amd
    lib = prism_amd.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.pa_version().startswith(b"prism_amd")


@pytest.mark.parametrize("shape", [(64, 64, 64), (384, 384, 256), (512, 512, 512), (33, 47, 129), (17, 1, 1),
                                   (2048, 2048, 1), (10000000, 1, 1), (1, 1, 1), (5, 70, 3), (255, 3, 17)])
def test_layout_equals_oracle(oracle, shape):
    import prism_amd
    a = prism_amd.layout(*shape).as_dict()
    b = oracle.layout(*shape).as_dict()
    assert a == b


def test_layout_random_shapes(oracle):
    import prism_amd
    rng = np.random.default_rng(0)
    for _ in range(300):
        shape = [int(v) for v in rng.integers(1, 400, 3)]
        assert prism_amd.layout(*shape).as_dict() == oracle.layout(*shape).as_dict(), shape


def test_rejects_bad_arguments():
    import ctypes as C
    import prism_amd
    from prism_amd import _lib
    L = _lib.Layout()
    assert prism_amd.lib().pa_layout_init(C.byref(L), 0, 4, 4) == _lib.PA_ERR_ARG
    assert prism_amd.lib().pa_layout_init(C.byref(L), 2048, 2048, 1024) == _lib.PA_ERR_ARG  # 2^32 elements
    assert prism_amd.lib().pa_bitshuffle(None, C.byref(L), 0, None, None) == _lib.PA_ERR_ARG
