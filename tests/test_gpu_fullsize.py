"""Full-size workloads on the GPU, checked through properties that do not
need the oracle (which would take minutes at these sizes): the 384 x 384 x
256 float64 density field at rel 1e-6 and the 512^3 float32 Gaussian
random field at rel 1e-4.

- bit-planes survive the codec: decode(stream) equals the shuffled planes;
- codes survive the shuffle: inverse_shuffle(planes) equals the codes;
- every progressive level is within its bound.  Float64: exactly.
  Float32: the method's own float32 arithmetic (the prediction and
  pred + code * 2eb in the quantizer, then one delta sum per round) rounds
  past the bound by a fraction of a unit in the last place; measured on the
  Gaussian random field: 1.00026 x the bound for a plain decompression and
  1.00066 x after the fourth progressive round.  The test prints each
  level's L-inf / bound and allows (k + 1) units of float32 epsilon times
  the field's magnitude after round k (1.0 in the plain case);
- a plain (SM) stream decompresses within its bound (same allowance).
"""
import ctypes as C
import sys

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from prism_amd import _lib, fields, pipeline
    return _lib, fields, pipeline


def P(t):
    return C.c_void_p(t.data_ptr())


def S():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def check_round_trips(_lib, pipeline, c, parts):
    L = parts["layout"]
    ws = torch.empty(_lib.lib().pa_decode_workspace(C.byref(L)), dtype=torch.uint8, device="cuda")
    bp = torch.empty(L.host_size, dtype=torch.uint8, device="cuda")
    _lib.check(_lib.lib().pa_lossless_decode(P(c.stream), c.nbytes, C.byref(L), P(bp), P(ws), ws.numel(), S()))
    assert torch.equal(bp, parts["bp"]), "decode(encode(planes)) differs"
    assert int(ws[:4].view(torch.int32).item()) == 0
    qc = torch.zeros(L.code_len, dtype=torch.int32, device="cuda")  # the anchor slots are not bit-plane data
    mode = _lib.PA_MODE_NB if c.progressive else _lib.PA_MODE_SM
    _lib.check(_lib.lib().pa_inverse_bitshuffle(P(bp), C.byref(L), mode, P(qc), S()))
    assert torch.equal(qc, parts["qc"]), "inverse_shuffle(shuffle(codes)) differs"


@pytest.mark.parametrize("case", ["density_f64", "grf_f32"])
def test_full_size_properties(mods, case):
    _lib, fields, pipeline = mods
    if case == "density_f64":
        f = fields.density(384, 384, 256, seed=2, dtype=torch.float64, device="cuda")
        rel, targets = 1e-6, [1e-2, 1e-3, 1e-4, 1e-5, 1e-6]
    else:
        f = fields.grf(512, 512, 512, seed=1, dtype=torch.float32, device="cuda")
        rel, targets = 1e-4, [1e-1, 1e-2, 1e-3, 1e-4]
    torch.cuda.synchronize()
    announce("full size", case)
    mag = f.abs().max().double().item()
    eps = torch.finfo(f.dtype).eps if f.dtype == torch.float32 else 0.0
    c, parts = pipeline.compress(f, rel, rel=True, progressive=True, return_parts=True)
    check_round_trips(_lib, pipeline, c, parts)
    r = pipeline.Retriever(c)
    for k, t in enumerate(targets):
        out = r.retrieve(t)
        q = pipeline.quality(f, out)
        err, bound = q["linf"], t * c.range
        over = "" if err <= bound or eps == 0 else f"  (over by {(err - bound) / (eps * mag):.2f} eps x max|f|)"
        print(f"LEVEL {case} target {t:g}: linf / bound = {err / bound:.8f}{over}", flush=True)
        assert q["sse"] == q["sse"], "NaN in the reconstruction"
        assert err <= bound + (k + 1) * eps * mag, (case, t, err, bound)
    del r, out, parts
    announce("full size plain stream", case)
    c2, parts2 = pipeline.compress(f, rel, rel=True, progressive=False, return_parts=True)
    check_round_trips(_lib, pipeline, c2, parts2)
    out = pipeline.decompress(c2)
    q = pipeline.quality(f, out)
    err, bound = q["linf"], rel * c2.range
    print(f"LEVEL {case} plain: linf / bound = {err / bound:.8f}", flush=True)
    assert q["sse"] == q["sse"], "NaN in the reconstruction"
    assert err <= bound + eps * mag, (case, err, bound)
