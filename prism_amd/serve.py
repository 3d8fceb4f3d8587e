"""Serving: a request loop over compressed fields kept resident in HBM.

Requests are progressive retrievals of a resident field (refined from the
field's current level, never restarted) and compressions of new fields.
The loop batches continuously: whatever has arrived when a launch slot is
free is admitted, in arrival order, as far as the memory budget allows, and
goes to the device as one batched retrieval round (pa_retrieve_batch) plus
one batched compression (pa_batch_*) behind the batches already queued on
the stream, so requests that arrive while others run overlap with them.
A field takes at most one round per batch; its next round goes to a later
batch, which the stream runs after it.

Every result equals what the plain per-field path gives for the same
request sequence of that field: retrieval rounds are the batched form of
Retriever.retrieve, compressions the batched form of compress().
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field as dc_field

import torch

from . import pipeline


@dataclass
class Request:
    rid: int
    kind: str                    # "retrieve" or "compress"
    arrival: float               # seconds on the server's clock
    field: int = -1              # retrieve: resident field id
    target: float = 0.0          # retrieve: target bound (relative in REL mode)
    data: torch.Tensor = None    # compress: device field
    eb: float = 0.0              # compress: relative bound
    # set by the server
    state: str = "queued"        # queued, running, done, failed, cancelled, rejected
    admitted: float = None
    done: float = None
    result: object = None        # retrieve: reconstruction; compress: Compressed
    error: str = None
    round: int = -1              # retrieve: the field's round this request ran as
    nbytes: int = 0              # algorithmic bytes (known once done)

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


@dataclass
class Resident:
    compressed: pipeline.Compressed
    retriever: pipeline.Retriever
    nbytes: int
    rounds: list = dc_field(default_factory=list)  # (request id, target) in run order


def _tensor_bytes(*ts):
    return sum(t.numel() * t.element_size() for t in ts if t is not None)


def _compress_need(t: torch.Tensor) -> int:
    """Device bytes a compression of t holds while it runs: output bound and
    batch workspace (codes, bit-planes, encoder status words and tables)."""
    nx, ny, nz = pipeline.shape_xyz(t)
    L = pipeline.layout(nx, ny, nz)
    lib = pipeline.lib()
    bound = lib.pa_compress_bound(pipeline.C.byref(L), pipeline.DTYPES[t.dtype])
    return int(bound + 4 * L.code_len + L.host_size + (L.nchunks + 1) * 8 + L.n4 * t.element_size() + 8192)


def _result_bytes(req: Request) -> int:
    """Device bytes a completed request's result holds: a reconstruction,
    or the stream of a compression (a view of its batch's output buffer,
    which stays allocated as long as any of its streams is held)."""
    res = req.result
    if res is None:
        return 0
    if isinstance(res, torch.Tensor):
        return _tensor_bytes(res)
    return _tensor_bytes(res.stream)


class Server:
    """Continuous-batching request loop on the current HIP stream.

    budget_bytes bounds the device memory the server holds: resident fields
    (stream, decoded bit-planes, codes, current reconstruction), the
    transient memory of admitted requests (new reconstructions, compression
    workspaces and outputs) until they complete, and the results of
    completed requests until they are handed over.  Results are handed over
    in one of two ways:
      - on_done: called with each finished request (done, failed, rejected
        or cancelled) as soon as the loop sees it; the server keeps nothing
        after the call returns (it drops req.result), so the callback must
        consume or copy what it needs;
      - otherwise finished requests collect in `completed`, their results
        counted against the budget (`held_bytes`) until pop_completed()
        hands them to the caller.
    A request that cannot be admitted because held results fill the budget
    while nothing is in flight raises MemoryError instead of waiting for
    ever.  max_inflight batches may be queued on the stream at once;
    max_batch requests form one batch, chosen among the first 4 x max_batch
    queued (a field's later rounds wait their turn)."""

    def __init__(self, budget_bytes: int, max_batch: int = 64, max_inflight: int = 2, clock=time.perf_counter,
                 on_done=None):
        self.budget = int(budget_bytes)
        self.max_batch = max_batch
        self.max_inflight = max_inflight
        self.clock = clock
        self.on_done = on_done
        self.fields: dict[int, Resident] = {}
        self.queue: deque[Request] = deque()
        self.inflight: deque = deque()  # (event, requests, transient bytes, keep-alive)
        self.resident_bytes = 0
        self.transient_bytes = 0
        self.held_bytes = 0
        self.completed: list[Request] = []
        self.finished = 0  # requests finished so far, whichever way they were handed over
        self.batches = 0
        # host seconds spent admitting and launching batches and completing
        # them, and device seconds of the completed batches (start to end
        # event of each batch on the stream)
        self.host_launch_s = 0.0
        self.host_poll_s = 0.0
        self.device_s = 0.0

    def _finish(self, req: Request) -> None:
        self.finished += 1
        if self.on_done is not None:
            self.on_done(req)
            req.result = None
            return
        nb = _result_bytes(req)
        req._held = nb
        self.held_bytes += nb
        self.completed.append(req)

    def pop_completed(self) -> list:
        """Hands the finished requests (and their device results) to the
        caller; the server stops counting their results."""
        out, self.completed = self.completed, []
        for r in out:
            self.held_bytes -= getattr(r, "_held", 0)
            r._held = 0
        return out

    # ------------------------------------------------------------ fields --
    def add_field(self, fid: int, c: pipeline.Compressed) -> None:
        """Makes a compressed (progressive) stream resident."""
        if fid in self.fields:
            raise ValueError(f"field {fid} is already resident")
        r = pipeline.Retriever(c)
        nx, ny, nz = c.shape
        n = nx * ny * nz * torch.empty(0, dtype=c.dtype).element_size()
        nb = _tensor_bytes(c.stream, r.bp, r.qc, r.dws, r.state) + n
        if self.used + nb > self.budget:
            raise MemoryError(f"field {fid} ({nb} B) does not fit the budget")
        self.fields[fid] = Resident(c, r, nb)
        self.resident_bytes += nb

    @property
    def used(self) -> int:
        return self.resident_bytes + self.transient_bytes + self.held_bytes

    # ---------------------------------------------------------- requests --
    def submit(self, req: Request) -> None:
        if req.kind == "retrieve":
            if req.field not in self.fields:
                req.state, req.error = "rejected", f"no resident field {req.field}"
                self._finish(req)
                return
        elif req.kind != "compress":
            raise ValueError(f"unknown request kind {req.kind!r}")
        need = self._need(req)
        if need > self.budget - self.resident_bytes:
            req.state, req.error = "rejected", f"needs {need} B, more than the budget leaves"
            self._finish(req)
            return
        self.queue.append(req)

    def cancel(self, rid: int) -> bool:
        """Cancels a request that has not been admitted yet."""
        for r in self.queue:
            if r.rid == rid:
                self.queue.remove(r)
                r.state = "cancelled"
                self._finish(r)
                return True
        return False

    def _need(self, req: Request) -> int:
        if req.kind == "retrieve":
            c = self.fields[req.field].compressed
            nx, ny, nz = c.shape
            return nx * ny * nz * torch.empty(0, dtype=c.dtype).element_size()
        return _compress_need(req.data)

    # -------------------------------------------------------------- loop --
    def step(self) -> int:
        """Completes finished batches, then admits and launches while a
        launch slot is free.  Returns the number of requests launched."""
        self.poll()
        launched = 0
        t0 = time.perf_counter()
        while self.queue and len(self.inflight) < self.max_inflight:
            batch, need = self._admit()
            if not batch:
                if not self.inflight and self.held_bytes > 0:
                    raise MemoryError(f"request {self.queue[0].rid} waits for memory held by {len(self.completed)} "
                                      f"results not yet taken with pop_completed() ({self.held_bytes} B)")
                break
            self._launch(batch, need)
            launched += len(batch)
        if launched:
            self.host_launch_s += time.perf_counter() - t0
        return launched

    def _admit(self):
        batch, need, seen = [], 0, set()
        room = self.budget - self.used
        keep = deque()
        lookahead = 4 * self.max_batch  # queued requests examined per batch
        while self.queue and len(batch) < self.max_batch and len(batch) + len(keep) < lookahead:
            r = self.queue.popleft()
            if r.kind == "retrieve" and r.field in seen:
                keep.append(r)  # one round per field per batch; the next batch takes it
                continue
            nb = self._need(r)
            if nb > room - need:
                keep.append(r)
                break  # memory admission in arrival order
            need += nb
            batch.append(r)
            if r.kind == "retrieve":
                seen.add(r.field)
        while keep:
            self.queue.appendleft(keep.pop())
        return batch, need

    def _launch(self, batch, need):
        now = self.clock()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        rets = [r for r in batch if r.kind == "retrieve"]
        comps = [r for r in batch if r.kind == "compress"]
        keep = []
        if rets:
            items = []
            for r in rets:
                res = self.fields[r.field]
                r.round = res.retriever.rounds
                res.rounds.append((r.rid, r.target))
                items.append((res.retriever, r.target))
            rb = pipeline.retrieve_batch(items, pinned=True)
            for r, out in zip(rets, rb.outs):
                r.result = out
            keep.append(rb)
        else:
            rb = None
        bc = None
        if comps:
            bc = pipeline.BatchCompressor([r.data for r in comps], [r.eb for r in comps], rel=True,
                                          progressive=True, graph=False)
            bc.enqueue()
            keep.append(bc)
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        for r in batch:
            r.state, r.admitted = "running", now
        self.transient_bytes += need
        self.inflight.append((ev, batch, need, keep, comps, bc, rets, rb, ev0))
        self.batches += 1

    def poll(self) -> int:
        """Marks the requests of completed batches done (in stream order)."""
        n = 0
        while self.inflight and self.inflight[0][0].query():
            t0 = time.perf_counter()
            ev, batch, need, keep, comps, bc, rets, rb, ev0 = self.inflight.popleft()
            self.device_s += ev0.elapsed_time(ev) * 1e-3
            now = self.clock()
            if bc is not None:
                res = bc.results(strict=False)
                for r, c in zip(comps, res):
                    if c is None:
                        r.state, r.error = "failed", "non-positive error bound"
                    else:
                        if self.on_done is None:
                            # a held stream gets its own allocation, so that it
                            # does not keep the whole batch output alive
                            c.stream = c.stream.clone()
                        r.result = c
            if rb is not None:
                for r, err in zip(rets, rb.errors()):
                    if err:
                        r.state, r.error, r.result = "failed", f"malformed compressed stream (decoder status {err})", None
            self.transient_bytes -= need
            for r in batch:
                if r.state == "running":
                    r.state = "done"
                r.done = now
                self._finish(r)
            n += len(batch)
            self.host_poll_s += time.perf_counter() - t0
        return n

    def drain(self) -> None:
        """Runs the loop until every submitted request has completed."""
        while self.queue or self.inflight:
            if not self.step() and self.inflight:
                self.inflight[0][0].synchronize()

    def run(self, requests, pace: bool = True, idle_sleep: float = 5e-5):
        """Feeds `requests` (sorted by arrival, seconds from now) into the
        loop, in real time when `pace`, and runs it until all completed."""
        t0 = self.clock()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            r.arrival += t0
        while pending or self.queue or self.inflight:
            now = self.clock()
            while pending and (not pace or pending[0].arrival <= now):
                self.submit(pending.popleft())
            if not self.step():
                if pending and pace:
                    wait = pending[0].arrival - self.clock()
                    if wait > 0:
                        time.sleep(min(wait, idle_sleep))
                elif self.inflight:
                    time.sleep(idle_sleep)
        return t0


def field_bytes(server: Server, req: Request) -> int:
    """Field-data bytes of a completed request: a compression reads n s
    and writes |Z|; a retrieval round writes n s and, after the field's
    first round, reads the previous n s.  (The payload a round reads is
    counted per field by the retriever's bytes_read.)"""
    if req.state != "done":
        return 0
    if req.kind == "compress":
        return _tensor_bytes(req.data) + req.result.nbytes
    c = server.fields[req.field].compressed
    nx, ny, nz = c.shape
    ns = nx * ny * nz * torch.empty(0, dtype=c.dtype).element_size()
    return ns * (2 if req.round > 0 else 1)


def poisson_requests(n: int, rate: float, seed: int, field_ids, pool, compress_share: float = 0.2,
                     targets=(1e-2, 1e-3, 1e-4, 1e-5, 1e-6)):
    """n requests with exponential inter-arrival gaps (mean 1/rate) from
    seed: a compression of the next field of `pool` ((tensor, rel eb)
    pairs, cycled) with probability compress_share, else a retrieval of a
    uniformly drawn resident field at a uniformly drawn target."""
    import numpy as np

    g = np.random.Generator(np.random.PCG64(seed))
    gaps = g.exponential(1.0 / rate, n)
    kinds = g.random(n) < compress_share
    fsel = g.integers(0, len(field_ids), n)
    tsel = g.integers(0, len(targets), n)
    t, k, out = 0.0, 0, []
    for i in range(n):
        t += float(gaps[i])
        if kinds[i]:
            data, eb = pool[k % len(pool)]
            k += 1
            out.append(Request(rid=i, kind="compress", arrival=t, data=data, eb=eb))
        else:
            out.append(Request(rid=i, kind="retrieve", arrival=t, field=int(field_ids[fsel[i]]),
                               target=float(targets[tsel[i]])))
    return out
