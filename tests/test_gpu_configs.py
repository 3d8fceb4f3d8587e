"""The benchmark workloads themselves, run on the GPU path at their full
sizes:

* the 64^3 float32 field of 8 sinusoids + 4 blobs + noise (seed 0) at
  relative bound 1e-3, plain (sign-magnitude) stream: HIP stream equal to
  the oracle's, decompression equal to the oracle's, bit for bit;
* the 64-field mixed batch (seeds 100+i: 3-D axes to 256, 2-D axes to 2048,
  1-D n to 1e7, half float32): batched streams equal to the per-field
  ones, and every stream retrieved progressively within its bound;
* the serving mix: 16 resident fields (seeds 200+i), 10,000 requests of
  the Poisson stream (seed 5) offered as a burst: every result equal to
  the same request run alone through the plain per-field path (tolerance
  0), compared through 64-bit weighted checksums of the result bytes.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mods():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from prism_amd import fields, pipeline, serve
    return fields, pipeline, serve


def test_smooth_mix_64_plain_equals_oracle(mods, oracle):
    fields, pipeline, _ = mods
    announce("64^3 f32 smooth mix, rel 1e-3, plain stream")
    host = np.ascontiguousarray(fields.smooth_mix(64, 64, 64, seed=0, dtype=torch.float32, device="cpu").numpy())
    c = pipeline.compress(torch.from_numpy(host).cuda(), 1e-3, rel=True, progressive=False)
    ref, _ = oracle.compress(host, 1e-3, mode=1)
    assert np.array_equal(c.stream.cpu().numpy(), ref)
    out = pipeline.decompress(c).cpu().numpy()
    dec = oracle.decompress(ref, host.shape, np.float32, 1e-3)
    assert np.array_equal(out.view(np.uint32), dec.view(np.uint32))
    bound = 1e-3 * (float(host.max()) - float(host.min()))
    assert np.abs(out.astype(np.float64) - host).max() <= bound


class _Digest:
    """64-bit weighted checksum of a tensor's bytes: the bytes as int32
    words, times odd pseudo-random 64-bit weights, summed modulo 2^64 (int64
    arithmetic wraps).  Any single changed word changes the sum (an odd
    weight times a non-zero word below 2^32 is not 0 mod 2^64)."""

    def __init__(self, max_words):
        g = torch.Generator(device="cpu").manual_seed(1234)
        w = torch.randint(-(1 << 62), 1 << 62, (max_words,), generator=g, dtype=torch.int64) * 2 + 1
        self.w = w.cuda()

    def __call__(self, t):
        b = t.reshape(-1).view(torch.uint8)
        pad = (-b.numel()) % 4
        if pad:
            b = torch.cat([b, b.new_zeros(pad)])
        words = b.view(torch.int32).to(torch.int64)
        return torch.stack([(words * self.w[:words.numel()]).sum(), torch.tensor(b.numel(), device=b.device)])


def test_mixed_batch_64_fields(mods):
    fields, pipeline, _ = mods
    announce("mixed batch: 64 fields, seeds 100+i, full size")
    data, ebs = [], []
    for i in range(64):
        t, e = fields.batch_field(i, 100 + i, device="cuda")
        data.append(t)
        ebs.append(e)
    torch.cuda.synchronize()
    want = [pipeline.compress(t, e, rel=True, progressive=True).stream.cpu() for t, e in zip(data, ebs)]
    b = pipeline.BatchCompressor(data, ebs, rel=True, progressive=True, graph=True, autotune=True)
    got = b.run()
    for i, (g, w) in enumerate(zip(got, want)):
        assert torch.equal(g.stream.cpu(), w), f"field {i} differs from its plain stream"
    announce("mixed batch: progressive retrieval of every batched stream")
    for i, (t, c) in enumerate(zip(data, got)):
        r = pipeline.Retriever(c)
        mag = t.abs().max().double().item()
        eps = torch.finfo(t.dtype).eps if t.dtype == torch.float32 else 0.0
        for k, target in enumerate(sorted({max(1e-2, c.eb_input), max(1e-4, c.eb_input), c.eb_input},
                                          reverse=True)):
            out = r.retrieve(target)
            q = pipeline.quality(t, out)
            bound = target * c.range
            # float32: the method's own sums in T may exceed the bound by a
            # few units of T's epsilon times the magnitude per round
            assert q["linf"] <= bound + 4 * (k + 1) * eps * mag, (i, target, q["linf"], bound)
            assert np.isfinite(q["sse"])


def test_serving_mix_alone_equals_in_mix(mods):
    fields, pipeline, serve = mods
    announce("serving: 16 resident fields (seeds 200+i), 10,000 requests as a burst")
    comp = []
    for i in range(16):
        t, e = fields.batch_field(i, 200 + i, device="cuda")
        comp.append(pipeline.compress(t, e, rel=True, progressive=True))
        del t
    pool = [fields.batch_field(j, 300 + j, device="cuda") for j in range(16)]
    torch.cuda.synchronize()
    words = max(max(c.stream.numel() for c in comp), max(t.numel() * t.element_size() for t, _ in pool),
                max(np.prod(c.shape) * (4 if c.dtype == torch.float32 else 8) for c in comp)) // 4 + 1
    digest = _Digest(int(words))
    got = {}

    def on_done(r):
        if r.state == "done":
            got[r.rid] = digest(r.result if r.kind == "retrieve" else r.result.stream)
        else:
            got[r.rid] = r.state

    srv = serve.Server(budget_bytes=100 << 30, max_batch=64, max_inflight=2, on_done=on_done)
    for k, c in enumerate(comp):
        srv.add_field(k, c)
    reqs = serve.poisson_requests(10000, rate=1e6, seed=5, field_ids=list(range(16)), pool=pool)
    srv.run(reqs, pace=False)
    torch.cuda.synchronize()
    assert srv.finished == len(reqs) and srv.held_bytes == 0 and not srv.completed
    assert all(isinstance(v, torch.Tensor) for v in got.values()), {v for v in got.values() if isinstance(v, str)}
    assert srv.batches < len(reqs) // 4  # requests shared batches
    announce("serving: replaying every field's rounds alone")
    for k, c in enumerate(comp):
        alone = pipeline.Retriever(c)
        for rid, target in srv.fields[k].rounds:
            want = digest(alone.retrieve(target, check_stream=False))
            assert torch.equal(want, got[rid]), (k, rid)
        alone.check()
        # begin / end / maxlevel (word 9 names the code hand-off each path took)
        assert torch.equal(alone.state[0:9], srv.fields[k].retriever.state[0:9]), k
    announce("serving: every compression alone")
    for r in reqs:
        if r.kind == "compress":
            want = digest(pipeline.compress(r.data, r.eb, rel=True, progressive=True).stream)
            assert torch.equal(want, got[r.rid]), r.rid
