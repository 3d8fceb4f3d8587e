"""Serving loop: results in the mix equal results alone (tolerance 0),
memory admission, cancellation and failed requests."""
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


LIMITS = dict(max3d=72, max2d=300, max1d=1.5e5)


def resident_set(fields, pipeline, n, seed0):
    out = []
    for i in range(n):
        t, e = fields.batch_field(i, seed0 + i, device="cuda", **LIMITS)
        out.append(pipeline.compress(t, e, rel=True, progressive=True))
    return out


def test_results_alone_equal_in_mix(mods):
    fields, pipeline, serve = mods
    announce("serve: alone vs in the mix")
    comp = resident_set(fields, pipeline, 6, 700)
    pool = [fields.batch_field(i, 800 + i, device="cuda", **LIMITS) for i in range(4)]
    torch.cuda.synchronize()
    srv = serve.Server(budget_bytes=8 << 30, max_batch=16, max_inflight=2)
    for k, c in enumerate(comp):
        srv.add_field(k, c)
    reqs = serve.poisson_requests(160, rate=4000.0, seed=5, field_ids=list(range(6)), pool=pool)
    srv.run(reqs, pace=True)
    torch.cuda.synchronize()
    assert len(srv.completed) == len(reqs)
    assert all(r.state == "done" for r in reqs), {r.state for r in reqs}
    assert srv.batches < len(reqs)  # requests did share batches
    byrid = {r.rid: r for r in reqs}
    # retrievals: each field's rounds replayed alone, in the order they ran
    for k, c in enumerate(comp):
        alone = pipeline.Retriever(c)
        for rid, target in srv.fields[k].rounds:
            want = alone.retrieve(target)
            got = byrid[rid].result
            assert torch.equal(want.view(torch.uint8), got.view(torch.uint8)), (k, rid)
        # begin / end / maxlevel (word 9 names the code hand-off each path took)
        assert torch.equal(alone.state[0:9], srv.fields[k].retriever.state[0:9])
    # compressions: alone through the plain per-field path
    for r in reqs:
        if r.kind == "compress":
            want = pipeline.compress(r.data, r.eb, rel=True, progressive=True)
            assert torch.equal(want.stream.cpu(), r.result.stream.cpu()), r.rid
    lat = sorted(r.latency for r in reqs)
    assert lat[0] >= 0


def test_admission_cancel_and_failure(mods):
    fields, pipeline, serve = mods
    announce("serve: admission, cancellation, failure")
    comp = resident_set(fields, pipeline, 2, 900)
    big, big_eb = fields.batch_field(0, 950, device="cuda", **LIMITS)
    need = serve._compress_need(big)
    srv = serve.Server(budget_bytes=1 << 40)
    for k, c in enumerate(comp):
        srv.add_field(k, c)
    # a budget that holds the resident fields and one compression at a time
    srv.budget = srv.resident_bytes + need + need // 2
    reqs = [serve.Request(rid=i, kind="compress", arrival=0.0, data=big, eb=big_eb) for i in range(3)]
    const = torch.full((20, 20, 20), 2.0, device="cuda", dtype=torch.float32)
    reqs.append(serve.Request(rid=3, kind="compress", arrival=0.0, data=const, eb=1e-3))
    reqs.append(serve.Request(rid=4, kind="retrieve", arrival=0.0, field=1, target=1e-3))
    reqs.append(serve.Request(rid=5, kind="retrieve", arrival=0.0, field=0, target=1e-4))
    reqs.append(serve.Request(rid=6, kind="retrieve", arrival=0.0, field=7, target=1e-4))  # not resident
    huge = torch.zeros(need, dtype=torch.float64, device="cuda")  # more than the budget leaves
    reqs.append(serve.Request(rid=7, kind="compress", arrival=0.0, data=huge, eb=1e-3))
    for r in reqs:
        srv.submit(r)
    assert srv.cancel(5) and not srv.cancel(6)
    srv.drain()
    st = {r.rid: r.state for r in reqs}
    assert st == {0: "done", 1: "done", 2: "done", 3: "failed", 4: "done", 5: "cancelled", 6: "rejected",
                  7: "rejected"}, st
    assert srv.batches >= 3  # the compressions could not share one batch
    assert srv.transient_bytes == 0
    want = pipeline.compress(big, big_eb, rel=True, progressive=True).stream.cpu()
    for i in range(3):
        assert torch.equal(reqs[i].result.stream.cpu(), want)


def test_held_results_count_against_the_budget(mods):
    """Results kept in `completed` count against the budget until
    pop_completed() hands them over; a queue blocked only by held results
    raises MemoryError instead of spinning; with a completion callback the
    server holds nothing.  Results equal the plain path either way."""
    fields, pipeline, serve = mods
    announce("serve: held results and the budget")
    comp = resident_set(fields, pipeline, 2, 910)
    out_bytes = max(int(np.prod(c.shape)) * (4 if c.dtype == torch.float32 else 8) for c in comp)
    targets = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-6, 1e-5, 1e-3]

    def requests():
        return [serve.Request(rid=i, kind="retrieve", arrival=0.0, field=i % 2, target=t)
                for i, t in enumerate(targets)]

    srv = serve.Server(budget_bytes=1 << 40)
    for k, c in enumerate(comp):
        srv.add_field(k, c)
    srv.budget = srv.resident_bytes + 3 * out_bytes  # room for three reconstructions
    reqs = requests()
    for r in reqs:
        srv.submit(r)
    taken, blocked = [], 0
    while srv.queue or srv.inflight:
        try:
            srv.drain()
        except MemoryError:
            blocked += 1
            assert srv.held_bytes > 0 and not srv.inflight
            taken += srv.pop_completed()
            assert srv.held_bytes == 0
    taken += srv.pop_completed()
    assert blocked >= 1 and len(taken) == len(reqs) and all(r.state == "done" for r in reqs)
    alone = [pipeline.Retriever(c) for c in comp]
    for r in reqs:
        want = alone[r.field].retrieve(r.target)
        assert torch.equal(want.view(torch.uint8), r.result.view(torch.uint8)), r.rid
    # the same requests with a callback: nothing is held, nothing blocks
    seen = []
    srv2 = serve.Server(budget_bytes=1 << 40, on_done=lambda r: seen.append((r.rid, r.state, r.result is not None)))
    for k, c in enumerate(comp):
        srv2.add_field(k, c)
    srv2.budget = srv2.resident_bytes + 3 * out_bytes
    for r in requests():
        srv2.submit(r)
    srv2.drain()
    assert sorted(s[0] for s in seen) == list(range(len(targets))) and all(s[1] == "done" and s[2] for s in seen)
    assert srv2.held_bytes == 0 and not srv2.completed


def _corrupted(pipeline, c):
    """The stream with four size-table entries pointing far past its end."""
    bad = c.stream.clone()
    bad[8:16] = 0xFF
    return pipeline.Compressed(stream=bad, shape=c.shape, dtype=c.dtype, eb_input=c.eb_input, rel=c.rel,
                               progressive=c.progressive, range=c.range)


def test_malformed_streams_are_reported(mods):
    """A stream whose size table points outside it: decompress() and
    Retriever.retrieve() raise PrismError, a served retrieval of it fails
    alone (its neighbours in the batch complete)."""
    fields, pipeline, serve = mods
    from prism_amd._lib import PrismError
    announce("malformed streams: plain, progressive, served")
    t, e = fields.batch_field(0, 960, device="cuda", **LIMITS)
    plain = pipeline.compress(t, e, rel=True, progressive=False)
    pipeline.decompress(plain)  # well formed: no error
    with pytest.raises(PrismError, match="malformed"):
        pipeline.decompress(_corrupted(pipeline, plain))
    prog = pipeline.compress(t, e, rel=True, progressive=True)
    with pytest.raises(PrismError, match="malformed"):
        pipeline.Retriever(_corrupted(pipeline, prog)).retrieve(1e-2)
    r = pipeline.Retriever(_corrupted(pipeline, prog))
    r.retrieve(1e-2, check_stream=False)
    assert r.error() != 0
    good = resident_set(fields, pipeline, 1, 970)[0]
    srv = serve.Server(budget_bytes=8 << 30)
    srv.add_field(0, good)
    srv.add_field(1, _corrupted(pipeline, prog))
    reqs = [serve.Request(rid=i, kind="retrieve", arrival=0.0, field=i % 2, target=1e-3) for i in range(4)]
    for q in reqs:
        srv.submit(q)
    srv.drain()
    st = {q.rid: q.state for q in reqs}
    assert st == {0: "done", 1: "failed", 2: "done", 3: "failed"}, st
    assert "malformed" in reqs[1].error and reqs[1].result is None
