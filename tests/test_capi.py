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


def test_noalloc_decode_limit_is_checked_on_the_host():
    """The allocation-free decode rejects a field whose scratch exceeds its
    fixed array before any HIP call (so this runs without a GPU)."""
    import ctypes as C
    import prism_amd
    from prism_amd import _lib
    lib = prism_amd.lib()
    assert lib.pa_noalloc_scratch_bytes() == 1 << 20
    small, big = prism_amd.layout(384, 384, 256), prism_amd.layout(2048, 2048, 128)
    assert lib.pa_decode_workspace(C.byref(small)) <= lib.pa_noalloc_scratch_bytes()
    assert lib.pa_decode_workspace(C.byref(big)) > lib.pa_noalloc_scratch_bytes()
    fake = C.c_void_p(256)  # never dereferenced: the call returns first
    assert lib.pa_lossless_decode_noalloc(fake, 1 << 20, C.byref(big), fake, fake, None) == _lib.PA_ERR_WORKSPACE
    assert lib.pa_lossless_decode_noalloc(None, 0, C.byref(small), fake, fake, None) == _lib.PA_ERR_ARG


# ---- the C++ API library (libprism_cpp.so) and a caller that uses only the
# reference's API names ----

CPP_LIB = os.path.join(ROOT, "prism_amd", "_build", "libprism_cpp.so")
DRIVER = os.path.join(ROOT, "prism_amd", "_build", "reference_driver")


@pytest.mark.parametrize("shape", [(384, 384, 256), (512, 512, 512), (20, 24, 28), (64, 64, 64), (17, 1, 1),
                                   (1, 1, 1), (2048, 2048, 1), (10000000, 1, 1), (33, 47, 129)])
def test_cpp_code_buffer_holds_code_len(shape):
    """Compressor::init's code buffer is exactly L.code_len int32: the
    predictor clears up to code_len and the inverse shuffle writes the
    phantom level-0 codes past the n elements."""
    import ctypes as C
    import prism_amd
    lib = C.CDLL(CPP_LIB)
    lib.prism_code_buffer_bytes.restype = C.c_size_t
    lib.prism_code_buffer_bytes.argtypes = [C.c_uint32] * 3
    L = prism_amd.layout(*shape)
    assert lib.prism_code_buffer_bytes(*shape) == 4 * L.code_len


def _exported(path):
    import subprocess
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", path], capture_output=True, text=True, check=True)
    return out.stdout


def test_cpp_library_exports_the_reference_api():
    """The names a driver of the reference's C++ API links against
    (compressor.hpp, io.hpp, dataloader.hpp, parameters.hpp, prism.hpp)."""
    syms = _exported(CPP_LIB)
    for name in ["prism::extrema_scan<float>(float*, unsigned long, float*)",
                 "prism::extrema_scan<double>(double*, unsigned long, double*)",
                 "prism::StatBuffer<float>::findMinMax()",
                 "prism::Compressor<float, int>::compress_pipeline(", "prism::Compressor<double, int>::init(",
                 "prism::Compressor<double, int>::decompress_progressive_pipeline(",
                 "prism::Buffer::H2D(long long)", "prism::Bitplane::Bitplane(",
                 "void findStrategy_h<int>(unsigned long*, int*, int*, double, double, double&, void*)",
                 "parse_argv(prism_context*, int, char**)", "initialize(prism_context*)"]:
        assert name in syms, name


def test_reference_api_driver_is_built_and_linked():
    """The reference-API driver (prism_amd/cpp/examples) compiled against the
    API headers only and linked with libprism_cpp.so; its GPU run is in
    tests/test_cli.py."""
    import subprocess
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    dyn = subprocess.run(["readelf", "-d", DRIVER], capture_output=True, text=True, check=True).stdout
    assert "libprism_cpp.so" in dyn and "libprism_amd.so" in dyn
    undef = subprocess.run(["nm", "-D", "-C", "--undefined-only", DRIVER], capture_output=True, text=True,
                           check=True).stdout
    for name in ("prism::extrema_scan<float>", "findStrategy_h<int>", "prism::Compressor<float, int>::init"):
        assert name in undef, name
