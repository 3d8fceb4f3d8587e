#!/usr/bin/env python3
"""Benchmark of the compress / progressive-retrieve path on one MI355X.

Workload (default): a 384 x 384 x 256 float64 Miranda-like density field
(synthetic), relative error bound 1e-6, compressed as a progressive stream,
then retrieved progressively at 1e-2, 1e-3, 1e-4, 1e-5, 1e-6.

A step is one compress of the field plus one full progressive retrieval
(5 incremental rounds) from a fresh retrieval state, with the field already
resident in HBM.  value = 2 * field_bytes / (compress_time + retrieve_time):
the mean of the compress and decompress throughputs in GB/s (1e9 B/s) of
field data, so a path that compresses and decompresses at X GB/s scores X.

Printed as one JSON line on rank 0.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

METRIC = "cmp/decmp GB/s + CR @rel-eb 1e-6; batched-vs-plain GB/s; p99 retrieval lat"
HBM_PEAK_GBPS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-baseline", choices=["on", "off"], default="on")
    ap.add_argument("--cpu-sample-planes", type=int, default=256,
                    help="z-planes of the field timed on the host cores for the CPU baseline")
    ap.add_argument("--shape", type=int, nargs=3, default=[384, 384, 256], metavar=("NX", "NY", "NZ"))
    ap.add_argument("--rel-eb", type=float, default=1e-6)
    ap.add_argument("--targets", type=float, nargs="+", default=[1e-2, 1e-3, 1e-4, 1e-5, 1e-6])
    ap.add_argument("--grf-steps", type=int, default=5,
                    help="timed steps of the 512^3 float32 Gaussian random field (rel 1e-4); 0 skips it")
    ap.add_argument("--only", choices=["all", "grf"], default="all",
                    help="grf: run only the Gaussian random field workload and print its own JSON line")
    ap.add_argument("--batch-fields", type=int, default=64,
                    help="fields of the mixed batch (batched vs plain per-field launches); 0 skips it")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the serving run (Poisson arrivals); 0 skips it")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="mean arrival rate, requests/s")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="run one device-to-device copy of this many MiB first (a known-size kernel for "
                         "scaling HBM byte counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (prism_amd runs each field on one GPU and shards nothing)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from prism_amd import fields

    torch.cuda.init()
    dev = torch.device("cuda", 0)
    if args.calibrate_copy_mib > 0:
        src = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only == "grf":  # the Gaussian random field workload alone (for profiling), its own JSON line
        print(json.dumps({"grf512": grf_section(args, dev)}))
        return
    nx, ny, nz = args.shape
    field = fields.density(nx, ny, nz, seed=2, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    main = progressive_workload(field, args.rel_eb, args.targets, args.steps, args.warmup)
    s = field.element_size()
    stats = main["stages"]
    dominant = max(stats, key=lambda k: stats[k]["avg_ms"] * stats[k]["launches_per_step"])
    d = stats[dominant]
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": d["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": (d["gbps"] / HBM_PEAK_GBPS) if d["gbps"] else None, "traffic": None,
                "avg_launch_ms": d["avg_ms"], "algorithmic_bytes_per_launch": d["avg_bytes"]}
    roofline.update(measured_traffic(dominant, "density", nx, ny, nz, s))
    del field

    grf = grf_section(args, dev) if args.grf_steps > 0 else None

    batched = batched_section(args) if args.batch_fields > 0 else None
    serving = serving_section(args) if args.serve_requests > 0 else None

    cpu = None
    if args.cpu_baseline == "on" and int(os.environ.get("RANK", "0")) == 0:
        cpu = cpu_baseline(fields.density(nx, ny, nz, seed=2, dtype=torch.float64, device=dev), args)

    line = {
        "metric": METRIC, "value": main["value"], "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": main["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Miranda-like density: 1 + 0.5 tanh(8 plane waves) + N(0,1e-4), seed 2)",
        "config": {"workload": "progressive density field 384x384x256 f64, rel eb 1e-6, retrieve 1e-2..1e-6",
                   "shape_xyz": [nx, ny, nz], "rel_eb": args.rel_eb, "targets": args.targets},
        "compress_GBps": main["compress_GBps"], "decompress_GBps": main["decompress_GBps"],
        "compression_ratio": main["compression_ratio"], "stream_bytes": main["stream_bytes"],
        "levels": main["levels"], "stages": stats, "roofline": roofline,
        "roofline_end_to_end": main["roofline_end_to_end"], "cpu_baseline": cpu, "grf512": grf,
        "batched": batched, "serving": serving,
    }
    print(json.dumps(line))


def grf_section(args, dev):
    """The 512^3 float32 Gaussian random field (power ~ |k|^-5/3, random
    phases seed 1, unit variance) at relative bound 1e-4, retrieved at 1e-1,
    1e-2, 1e-3 and 1e-4."""
    import torch

    from prism_amd import fields

    g = fields.grf(512, 512, 512, seed=1, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    out = progressive_workload(g, 1e-4, [1e-1, 1e-2, 1e-3, 1e-4], args.grf_steps, 1)
    out["config"] = {"workload": "Gaussian random field 512^3 f32 (power ~ |k|^-5/3, seed 1, unit variance), "
                                 "rel eb 1e-4, retrieve 1e-1..1e-4", "shape_xyz": [512, 512, 512]}
    return out


def progressive_workload(field, rel_eb, targets, steps, warmup):
    """One field, compressed as a progressive stream at rel_eb and retrieved
    progressively at each target, with the field resident in HBM.  A step
    is one compress plus a fresh retrieval through every target; compress
    and retrieval are timed separately with host clocks around
    device-synchronised phases.  Reports throughput, per-stage device time
    with algorithmic bytes, per-level quality against the original, and the
    end-to-end roofline fractions on primary bytes (compress: n s + |Z|;
    retrieval round k: |Z_k| + n s written + n s of the previous level read
    from round 2 on)."""
    import torch

    from prism_amd import pipeline
    from prism_amd._lib import layout

    nz, ny, nx = field.shape
    L = layout(nx, ny, nz)
    s = field.element_size()
    field_bytes = field.numel() * s

    def one_step(timer=None):
        c = pipeline.compress(field, rel_eb, rel=True, progressive=True, timer=timer)
        r = pipeline.Retriever(c)
        for t in targets:
            r.retrieve(t, timer=timer, check_stream=False)
        r.check()
        return c, r

    for _ in range(warmup):
        one_step()
    torch.cuda.synchronize()

    # timed region: compress and retrieve timed separately with host clocks
    # around device-synchronised phases, summed per step
    t_cmp = t_dec = 0.0
    torch.cuda.synchronize()
    t_all0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        c = pipeline.compress(field, rel_eb, rel=True, progressive=True)
        torch.cuda.synchronize()
        b = time.perf_counter()
        r = pipeline.Retriever(c)
        for t in targets:
            r.retrieve(t, check_stream=False)
        r.check()  # the decoder's error word, read once after the last round
        torch.cuda.synchronize()
        e = time.perf_counter()
        t_cmp += b - a
        t_dec += e - b
    t_all = time.perf_counter() - t_all0
    del c, r

    # per-stage device time (HIP events on the launch stream), separate pass
    timer = pipeline.StageTimer()
    passes = max(1, min(steps, 5))
    for _ in range(passes):
        one_step(timer)
    stage_ms = timer.results()

    # quality and bytes per level, untimed: statistics against the original
    # with the device kernel (L-inf, PSNR, NRMSE in one pass)
    c = pipeline.compress(field, rel_eb, rel=True, progressive=True)
    stream_bytes = c.nbytes
    r = pipeline.Retriever(c)
    levels = []
    prev = 0
    for t in targets:
        torch.cuda.synchronize()
        a = time.perf_counter()
        out = r.retrieve(t)
        torch.cuda.synchronize()
        lat = (time.perf_counter() - a) * 1e3
        q = pipeline.quality(field, out)
        nread = int(r.bytes_read.item())
        levels.append({"target": t, "bound_abs": t * c.range, "linf": q["linf"],
                       # how the round's codes reached the reconstruction
                       "code_source": ("int32", "int16", "planes")[int(r.narrow.item())],
                       "linf_over_bound": q["linf"] / (t * c.range), "within_bound": q["linf"] <= t * c.range,
                       "sse_finite": math.isfinite(q["sse"]), "psnr_db": q["psnr_db"], "nrmse": q["nrmse"],
                       "payload_bytes_read": nread - prev, "ms": lat,
                       "begin": r.begin.tolist(), "end": r.end.tolist()})
        prev = nread
        del out
    del r, c

    # per-launch algorithmic bytes of each stage
    nrounds = len(targets)
    codes = 4 * (L.n - L.n4)  # int32 codes written by the predictor and read by the shuffle

    def planes_bytes(lv_info):
        return sum((lv_info["end"][3 - lv] - lv_info["begin"][3 - lv]) * L.dev_stride[lv] for lv in range(4))

    def launch_bytes(name, k):
        if name in ("extrema", "autotune"):
            return L.n * s
        if name == "predict":
            return L.n * s + codes + L.n4 * s
        if name == "shuffle":
            return codes + L.host_size
        if name == "encode":
            return L.host_size + stream_bytes
        rnd = levels[k % nrounds]
        if name == "decode":
            return rnd["payload_bytes_read"] + planes_bytes(rnd)
        # the round's code hand-off (int32 or int16 codes), or none: the
        # reconstruction then reads the window's planes itself and the
        # inverse shuffle only its window
        src = rnd["code_source"]
        code_b = {"int32": 4, "int16": 2, "planes": 0}[src]
        if name == "unshuffle":
            return 0 if src == "planes" else planes_bytes(rnd) + code_b * L.code_len
        if name == "reconstruct":
            codes_in = planes_bytes(rnd) if src == "planes" else code_b * (L.n - L.n4)
            return codes_in + L.n * s + (L.n * s if k % nrounds else L.n4 * s)
        return 0

    stats = {}
    for name, times in stage_ms.items():
        byts = [launch_bytes(name, k) for k in range(len(times))]
        avg = sum(times) / len(times)
        stats[name] = {"avg_ms": avg, "launches_per_step": len(times) // passes,
                       "avg_bytes": sum(byts) / len(byts),
                       "gbps": (sum(byts) / (sum(times) * 1e-3) / 1e9) if sum(byts) else None}

    # end-to-end roofline fractions on primary bytes, over the wall-clock
    # time of the timed loop and over the summed device stage times
    cmp_bytes = field_bytes + stream_bytes
    ret_bytes = sum(lv["payload_bytes_read"] + field_bytes + (field_bytes if k else 0) for k, lv in enumerate(levels))
    cmp_stages = [k for k in ("extrema", "autotune", "predict", "shuffle", "encode", "merge") if k in stats]
    ret_stages = [k for k in ("strategy", "decode", "unshuffle", "reconstruct") if k in stats]
    cmp_dev = sum(stats[k]["avg_ms"] * stats[k]["launches_per_step"] for k in cmp_stages) * 1e-3
    ret_dev = sum(stats[k]["avg_ms"] * stats[k]["launches_per_step"] for k in ret_stages) * 1e-3

    def frac(byts, sec):
        return {"bytes": byts, "seconds": sec, "GBps": byts / sec / 1e9, "frac": byts / sec / 1e9 / HBM_PEAK_GBPS}

    e2e = {"unit": "GB/s", "peak": HBM_PEAK_GBPS,
           "compress": {"wall": frac(cmp_bytes, t_cmp / steps), "device": frac(cmp_bytes, cmp_dev)},
           "retrieve": {"wall": frac(ret_bytes, t_dec / steps), "device": frac(ret_bytes, ret_dev),
                        "rounds": nrounds}}
    return {"value": 2 * field_bytes * steps / (t_cmp + t_dec) / 1e9, "ms_per_step": 1e3 * t_all / steps,
            "compress_GBps": field_bytes * steps / t_cmp / 1e9, "decompress_GBps": field_bytes * steps / t_dec / 1e9,
            "compress_ms": 1e3 * t_cmp / steps, "retrieve_ms": 1e3 * t_dec / steps, "steps": steps,
            "compression_ratio": field_bytes / stream_bytes, "stream_bytes": stream_bytes, "field_bytes": field_bytes,
            "levels": levels, "stages": stats, "roofline_end_to_end": e2e}


def batched_section(args):
    """Mixed batch of fields (seeds 100+i; float32 for even i; 3-D, 2-D and
    1-D shapes; density or mode-spectrum values; rel eb log-uniform in
    [1e-6, 1e-2]): one batched launch sequence replayed from a HIP graph
    against a serial loop of plain per-field compressions.  Both produce
    the same streams (checked here, byte for byte); throughput counts the
    input bytes of all fields per wall-clock second, host synchronisations
    included (the plain path needs one per field for the value range and
    one for the stream length)."""
    import torch

    from prism_amd import fields, pipeline

    data, ebs, specs = [], [], []
    for i in range(args.batch_fields):
        t, e = fields.batch_field(i, 100 + i, device="cuda")
        data.append(t)
        ebs.append(e)
        specs.append(list(t.shape))
    torch.cuda.synchronize()
    total_bytes = sum(t.numel() * t.element_size() for t in data)

    def plain():
        return [pipeline.compress(t, e, rel=True, progressive=True) for t, e in zip(data, ebs)]

    b = pipeline.BatchCompressor(data, ebs, rel=True, progressive=True, graph=True, autotune=True)
    be = pipeline.BatchCompressor(data, ebs, rel=True, progressive=True, graph=False)
    ref = [c.stream.cpu() for c in plain()]
    got = b.run()
    identical = all(torch.equal(g.stream.cpu(), r) for g, r in zip(got, ref))
    got_e = be.run()
    identical_eager = all(torch.equal(g.stream.cpu(), r) for g, r in zip(got_e, ref))
    cbytes = sum(r.numel() for r in ref)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(args.batch_steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / args.batch_steps

    t_plain = timed(plain)
    t_graph = timed(b.run)
    t_eager = timed(be.run)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.batch_steps):
        b.enqueue()
    ev1.record()
    torch.cuda.synchronize()
    t_dev = ev0.elapsed_time(ev1) / args.batch_steps * 1e-3
    return {"fields": args.batch_fields, "input_bytes": total_bytes, "compressed_bytes": cbytes,
            "compression_ratio": total_bytes / cbytes, "streams_identical_to_plain": identical,
            "streams_identical_eager": identical_eager,
            "plain_ms": t_plain * 1e3, "batched_graph_ms": t_graph * 1e3, "batched_eager_ms": t_eager * 1e3,
            "batched_device_ms": t_dev * 1e3,
            "plain_GBps": total_bytes / t_plain / 1e9, "batched_GBps": total_bytes / t_graph / 1e9,
            "speedup_vs_plain": t_plain / t_graph, "launches_per_batch": b.launches,
            "launch_params": b.params,
            "shapes_zyx": specs, "dtypes": ["f32" if t.dtype == torch.float32 else "f64" for t in data]}


def serving_section(args):
    """16 resident progressive fields of the mixed-batch kind (seeds
    200+i), a Poisson request stream (exponential gaps, seed 5): 80 %
    retrieval rounds of a random resident field at a target drawn from
    {1e-2 .. 1e-6}, refined from the field's current level, 20 %
    compressions of new fields of the same mix (seeds 300+j).  Memory-based
    admission under a 100 GB budget.  Latency = arrival -> result in HBM
    and observed by the host loop."""
    import numpy as np
    import torch

    from prism_amd import fields, pipeline, serve

    comp = []
    for i in range(16):
        t, e = fields.batch_field(i, 200 + i, device="cuda")
        comp.append(pipeline.compress(t, e, rel=True, progressive=True))
        del t
    pool = [fields.batch_field(j, 300 + j, device="cuda") for j in range(16)]
    torch.cuda.synchronize()
    # warm-up on a throwaway server (allocator, first launches)
    serving_run(comp, serve.poisson_requests(64, rate=5000.0, seed=99, field_ids=list(range(16)), pool=pool))
    torch.cuda.synchronize()
    reqs = serve.poisson_requests(args.serve_requests, rate=args.serve_rate, seed=5, field_ids=list(range(16)),
                                  pool=pool)
    out = serving_run(comp, reqs)
    out["rate_offered"] = args.serve_rate
    # the same request mix offered as a burst (1e6 req/s): what the loop
    # sustains when requests queue up and batches fill
    reqs = serve.poisson_requests(args.serve_requests, rate=1e6, seed=5, field_ids=list(range(16)), pool=pool)
    out["saturated"] = serving_run(comp, reqs)
    out["saturated"]["rate_offered"] = 1e6
    return out


def serving_run(comp, reqs):
    """Runs `reqs` on a fresh server over the resident fields `comp`; each
    finished request is accounted in the server's completion callback (its
    latency and field bytes), after which the server drops its result."""
    import numpy as np
    import torch

    from prism_amd import serve

    acct = []  # (kind, state, latency, field bytes)

    def on_done(r):
        acct.append((r.kind, r.state, r.latency, serve.field_bytes(srv, r), r.done))

    srv = serve.Server(budget_bytes=100e9, max_batch=64, max_inflight=2, on_done=on_done)
    for k, c in enumerate(comp):
        srv.add_field(k, c)
    t0 = srv.run(reqs, pace=True)
    torch.cuda.synchronize()
    done = [a for a in acct if a[1] == "done"]
    lat = np.array([a[2] for a in done]) * 1e3
    wall = max(a[4] for a in done) - t0
    payload = sum(int(res.retriever.bytes_read.item()) for res in srv.fields.values())
    nbytes = sum(a[3] for a in done) + payload

    def pct(a, q):
        return float(np.percentile(a, q)) if len(a) else None

    by = {}
    for kind in ("retrieve", "compress"):
        a = np.array([d[2] for d in done if d[0] == kind]) * 1e3
        by[kind] = {"count": int(len(a)), "p50_ms": pct(a, 50), "p99_ms": pct(a, 99)}
    return {"requests": len(reqs), "completed": len(done),
            "wall_s": wall, "throughput_req_s": len(done) / wall, "GBps": nbytes / wall / 1e9,
            "algorithmic_bytes": nbytes, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "mean_ms": float(lat.mean()),
            "max_ms": float(lat.max()), "by_kind": by, "batches": srv.batches,
            # host time of the loop per completed request (admission, job
            # tables and launches; completion and hand-over), and the share
            # of the wall time the device spent on batches
            "host_us_per_request": 1e6 * (srv.host_launch_s + srv.host_poll_s) / max(len(done), 1),
            "host_launch_us_per_request": 1e6 * srv.host_launch_s / max(len(done), 1),
            "device_busy_fraction": srv.device_s / wall,
            "mean_batch": len(done) / max(srv.batches, 1), "resident_fields": len(srv.fields),
            "resident_bytes": srv.resident_bytes, "held_bytes_at_end": srv.held_bytes, "budget_bytes": srv.budget}


def measured_traffic(stage, workload, nx, ny, nz, s):
    """HBM bytes per launch of a stage's kernel from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE passes, each
    scaled by the same counter on a copy kernel of known size); empty when no
    summary covers this shape and dtype."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"traffic_{workload}_{nx}x{ny}x{nz}_f{8 * s}.json")
    try:
        with open(path) as f:
            prof = json.load(f)
        st = prof["stages"][stage]
    except (OSError, KeyError, ValueError):
        return {}
    return {"traffic": st["traffic_bytes"], "traffic_fetch": st["fetch_bytes"], "traffic_write": st["write_bytes"],
            "traffic_source": os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))}


def cpu_baseline(field, args):
    """The oracle (the reference algorithm restated in C) on the host cores,
    on a slab of the same field: same step (compress + 5-round retrieval)."""
    import numpy as np
    import torch

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, here)
    from oracle import oracle as O

    planes = min(args.cpu_sample_planes, field.shape[0])
    sample = np.ascontiguousarray(field[:planes].cpu().numpy())
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    nthreads = max(1, min(nthreads, os.cpu_count() or 1))

    def step(x, threads):
        t0 = time.perf_counter()
        stream, _ = O.compress(x, args.rel_eb, mode=0, nthreads=threads)
        t1 = time.perf_counter()
        p = O.Progressive(stream, x.shape, x.dtype, args.rel_eb, corrected=True, nthreads=threads)
        for t in args.targets:
            p.retrieve(t)
        return t1 - t0, time.perf_counter() - t1

    step(sample, nthreads)  # warm-up
    runs = sorted((step(sample, nthreads) for _ in range(5)), key=lambda r: r[0] + r[1])
    tc, tr = runs[2]  # median of 5
    b = sample.nbytes
    # one host core on a slab of 32 z-planes (one run)
    slab = np.ascontiguousarray(sample[:min(32, planes)])
    c1, r1 = step(slab, 1)
    return {"value": 2 * b / (tc + tr) / 1e9, "unit": "GB/s", "cores": nthreads, "kind": "port",
            "sample": f"first {planes} z-planes ({sample.shape[2]}x{sample.shape[1]}x{planes} f64) of the same field, "
                      f"one compress + {len(args.targets)}-round retrieval; median of 5 after 1 warm-up",
            "compress_s": tc, "retrieve_s": tr,
            "one_core": {"value": 2 * slab.nbytes / (c1 + r1) / 1e9, "unit": "GB/s",
                         "sample": f"first {slab.shape[0]} z-planes, one run"}}


if __name__ == "__main__":
    main()
