"""Batched compression: every field's stream from one batched launch
sequence (eager or replayed from a HIP graph) must equal, byte for byte,
the stream the plain per-field path writes for that field; the plain path
is itself pinned to the oracle by the parity tests.  Tolerance: none."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def announce(*a):
    print("CASE", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pl():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from prism_amd import pipeline
    return pipeline


def mixed_fields(n, seed0, **limits):
    from prism_amd import fields
    out = [fields.batch_field(i, seed0 + i, device="cuda", **limits) for i in range(n)]
    torch.cuda.synchronize()
    return [t for t, _ in out], [e for _, e in out]


def plain_streams(pl, data, ebs, rel=True, progressive=True):
    return [pl.compress(t, e, rel=rel, progressive=progressive).stream.cpu() for t, e in zip(data, ebs)]


def assert_same(got, want, label):
    assert len(got) == len(want)
    for i, (g, w) in enumerate(zip(got, want)):
        gs = g.stream.cpu()
        assert gs.numel() == w.numel(), f"{label}: field {i} length {gs.numel()} != {w.numel()}"
        assert torch.equal(gs, w), f"{label}: field {i} differs at byte {int((gs != w).nonzero()[0])}"


@pytest.mark.parametrize("graph,autotune", [(False, False), (True, False), (True, True)])
def test_mixed_batch_equals_plain(pl, graph, autotune):
    announce("batch mixed 12 fields, graph", graph, "autotune", autotune)
    data, ebs = mixed_fields(12, 300, max3d=80, max2d=300, max1d=2e5)
    want = plain_streams(pl, data, ebs)
    b = pl.BatchCompressor(data, ebs, rel=True, progressive=True, graph=graph, autotune=autotune)
    assert b.launches <= 7
    assert_same(b.run(), want, f"graph={graph} autotune={autotune}")
    p = b.params
    assert p["extrema_batches"] in (2, 4, 8, 16) and p["xcd_order"] in (0, 1) and p["encode_waves"] in (2, 4, 8)
    if autotune:
        assert all(t > 0 for t in p["tuned_ms"]["extrema"] + p["tuned_ms"]["encode"])
        assert_same(b.run(), want, "replay after tuning")


def test_edge_shapes_and_modes(pl):
    """Shapes the tuner skips (an axis under 17), single planes and lines,
    tiny fields, both bit-plane modes and absolute bounds."""
    rng = np.random.default_rng(7)
    shapes = [(1, 1, 100), (1, 1, 7), (1, 20, 33), (16, 16, 16), (17, 17, 17), (18, 40, 65), (5, 70, 3)]
    data = []
    for k, s in enumerate(shapes):
        a = np.cumsum(rng.normal(0, 1, s), axis=-1) * 0.01 + np.sin(np.arange(int(np.prod(s))).reshape(s) / 9.0)
        data.append(torch.from_numpy(a.astype(np.float32 if k % 2 else np.float64)).cuda())
    for rel, progressive in [(True, False), (False, True), (False, False)]:
        announce("batch edge shapes rel", rel, "progressive", progressive)
        ebs = [1e-3 if rel else 2e-3 * (1 + k) for k in range(len(data))]
        want = plain_streams(pl, data, ebs, rel=rel, progressive=progressive)
        b = pl.BatchCompressor(data, ebs, rel=rel, progressive=progressive, graph=True)
        assert_same(b.run(), want, f"rel={rel} progressive={progressive}")


def test_graph_replays_and_rebind(pl):
    """Several replays of one graph with eager work between them, then new
    inputs and bounds bound to the same plan."""
    announce("batch graph replays and rebind")
    data, ebs = mixed_fields(8, 400, max3d=64, max2d=256, max1d=1.5e5)
    want = plain_streams(pl, data, ebs)
    b = pl.BatchCompressor(data, ebs, graph=True)
    for r in range(3):
        assert_same(b.run(), want, f"replay {r}")
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        del junk
    data2 = [t * 1.5 + 0.25 for t in data]
    ebs2 = [e * 3 for e in ebs]
    want2 = plain_streams(pl, data2, ebs2)
    assert_same(b.run(data2, ebs2), want2, "rebind")
    assert_same(b.run(data, ebs), want, "rebind back")


def test_batched_streams_decode(pl):
    """Batched streams retrieve progressively like plain ones: every level
    within its bound, plus the rounding of the method's own arithmetic in
    the field's precision (reconstructions are sums of per-round values
    computed in T, as in the reference): 4 units of T's epsilon times the
    largest magnitude per round so far."""
    announce("batch streams progressive retrieval")
    data, ebs = mixed_fields(4, 500, max3d=64, max2d=256, max1d=1.5e5)
    b = pl.BatchCompressor(data, ebs, graph=True)
    for t, c in zip(data, b.run()):
        r = pl.Retriever(c)
        mag = t.abs().max().double().item()
        eps = torch.finfo(t.dtype).eps
        for k, target in enumerate((1e-2, 1e-4, c.eb_input)):
            target = max(target, c.eb_input)
            out = r.retrieve(target)
            err = (out.double() - t.double()).abs().max().item()
            assert err <= target * c.range + 4 * (k + 1) * eps * mag, (k, err, target * c.range)


def test_constant_field_is_reported(pl):
    announce("batch constant field in REL mode")
    good = torch.linspace(0, 1, 5000, device="cuda", dtype=torch.float64)
    const = torch.full((20, 20, 20), 3.0, device="cuda", dtype=torch.float32)
    b = pl.BatchCompressor([good, const], [1e-3, 1e-3], graph=False)
    b.enqueue()
    r = b.result.cpu().view(-1, 4)
    assert int(r[0, 1]) == 0 and int(r[1, 1]) != 0
    with pytest.raises(Exception):
        b.results()


def test_batched_retrieval_equals_plain(pl):
    """Rounds of many fields in one batched launch sequence against the
    plain per-field Retriever: same reconstructions (bit for bit), same
    begin / end / maxlevel state and bytes read; rounds include looser
    targets after tighter ones, targets under the stream's bound (every
    plane) and fields that sit a round out.  A rough field at a tight bound
    has codes in high planes, so its first round hands over int32 codes."""
    announce("batched retrieval rounds")
    data, ebs = mixed_fields(8, 600, max3d=72, max2d=280, max1d=1.6e5)
    gen = torch.Generator(device="cuda").manual_seed(9)
    data.append(torch.randn((24, 20, 30), generator=gen, dtype=torch.float64, device="cuda"))
    ebs.append(1e-7)
    comp = [pl.compress(t, e, rel=True, progressive=True) for t, e in zip(data, ebs)]
    plain = [pl.Retriever(c) for c in comp]
    batched = [pl.Retriever(c) for c in comp]
    rng = np.random.default_rng(11)
    targets = [1e-2, 1e-3, 1e-4, 1e-5, 1e-6]
    sources = set()  # (3-D field, code source of the round)
    for rnd in range(5):
        items, want = [], []
        for i in range(len(comp)):
            if rnd > 0 and rng.random() < 0.25:
                continue  # this field sits the round out
            t = float(targets[rng.integers(0, 5)]) if rnd < 4 else comp[i].eb_input * 0.5
            want.append((i, plain[i].retrieve(t)))
            items.append((batched[i], t))
        got = pl.retrieve_batch(items, pinned=(rnd % 2 == 1))
        torch.cuda.synchronize()
        for (i, w), g in zip(want, got.outs):
            assert torch.equal(w.view(torch.uint8), g.view(torch.uint8)), (rnd, i)
            # begin / end / maxlevel (word 9 names the code hand-off each path took)
            assert torch.equal(plain[i].state[0:9], batched[i].state[0:9]), (rnd, i)
            assert int(plain[i].bytes_read.item()) == int(batched[i].bytes_read.item()), (rnd, i)
            sources.add((comp[i].shape[2] > 1, int(batched[i].state[9].item())))
    # the batched rounds took every code source: int32 and int16 hand-offs
    # and, on 3-D fields only, the bit-planes
    assert (True, 2) in sources and (False, 2) not in sources, sources
    assert {s for _, s in sources} == {0, 1, 2}, sources
