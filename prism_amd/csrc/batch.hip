// batch.hip -- planning and running a batch of fields (host side).
//
// A plan lays out every field's scratch in one caller-provided workspace,
// builds the per-stage prefix tables that let one launch cover all fields,
// uploads the descriptors once and (optionally) captures the fixed launch
// sequence in a HIP graph.  Fields are ordered float32 first so that each
// dtype-specialised kernel covers one contiguous range of fields; results
// are reported in the caller's order.
#include <algorithm>
#include <new>
#include <vector>

#include "batch.hpp"
#include "common.hpp"

struct pa_batch_plan {
    int nf = 0, nf32 = 0;
    std::vector<int> order;  // order[k] = caller index of descriptor k
    std::vector<pa::BatchField> host;
    std::vector<int64_t> pre;  // prefix tables (host copy, source of the upload)
    pa::BatchDev dev{};
    const int64_t* ext_tab[4] = {};  // extrema prefix per ext_batches candidate
    int64_t ext_n[4] = {};
    float tuned_ms[9] = {};          // autotuner timings: extrema x4, spline order x2, encoder waves x3
    uint8_t* table = nullptr;  // descriptors and prefixes inside the workspace
    int launches = 0;
    hipGraphExec_t exec = nullptr;
};

namespace {

using pa::BatchField;

constexpr size_t kAlign = 256;
size_t up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// per-field scratch, in the order the plan lays it out
struct FieldSizes {
    size_t minmax, ebinfo, flags, scores, tune, qc, anchors, bp, encws, cum, outsize;
    size_t total() const {
        return minmax + ebinfo + flags + scores + tune + qc + anchors + bp + encws + cum + outsize;
    }
};

FieldSizes field_sizes(const BatchField& f) {
    const pa_layout& L = f.L;
    FieldSizes z;
    z.minmax = up(4 * sizeof(double));
    z.ebinfo = up(10 * sizeof(double) + 8 * sizeof(float));
    z.flags = up(8);
    z.scores = up(15 * sizeof(double));
    z.tune = up((size_t)f.tnblocks * 15 * sizeof(double));
    z.qc = up((size_t)L.code_len * 4);
    z.anchors = up((size_t)L.n4 * f.dtype);
    z.bp = up((size_t)L.host_size);
    z.encws = up(((size_t)L.nchunks + 1) * 8);
    z.cum = up(128 * sizeof(uint64_t));
    z.outsize = up(sizeof(uint64_t));
    return z;
}

constexpr int kPrefixes = 10;  // ext (one per ext_batches candidate) x4, tune, gap, tile, shuf, chunk, merge

size_t table_bytes(int nf) { return up((size_t)nf * sizeof(BatchField)) + up((size_t)kPrefixes * (nf + 1) * 8); }

// descriptor of one caller field: layout and host-side set-up (no pointers)
int describe(const pa_batch_field& in, BatchField& f) {
    if (!in.data || !in.out || (in.dtype != PA_F32 && in.dtype != PA_F64) ||
        (in.mode != PA_MODE_NB && in.mode != PA_MODE_SM) || !(in.eb > 0))
        return PA_ERR_ARG;
    if (((uintptr_t)in.data & 15) != 0) return PA_ERR_ARG;  // 16-byte loads in the extrema pass
    f = BatchField{};
    f.data = in.data;
    f.dtype = in.dtype;
    f.mode = in.mode;
    f.rel = in.rel ? 1 : 0;
    f.eb = in.eb;
    f.out = in.out;
    const int st = pa_layout_init(&f.L, in.nx, in.ny, in.nz);
    if (st != PA_OK) return st;
    f.cg = pa::make_chunk_geom(&f.L);
    f.lv = pa::make_levels(&f.L);
    pa::batch_predict_setup(f);
    return PA_OK;
}

// field order: float32 first, each dtype in caller order
int plan_order(const pa_batch_field* fields, int nf, std::vector<int>& order) {
    order.clear();
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < nf; ++i)
            if ((fields[i].dtype == PA_F32) == (pass == 0)) order.push_back(i);
    int nf32 = 0;
    for (int i = 0; i < nf; ++i) nf32 += fields[i].dtype == PA_F32;
    return nf32;
}

int enqueue(const pa_batch_plan* p, hipStream_t st) {
    int e;
    if ((e = pa::batch_front(p->dev, st)) != PA_OK) return e;
    if ((e = pa::batch_spline(p->dev, st)) != PA_OK) return e;
    if ((e = pa::batch_shuffle(p->dev, st)) != PA_OK) return e;
    return pa::batch_encode(p->dev, st);
}

int count_launches(const pa::BatchDev& b) {
    return 2                                        // front (tuner, extrema, gaps, status clear), finish
           + (b.n_tile32 > 0) + (b.n_tile > b.n_tile32)  // spline, per dtype
           + (b.n_shuf > 0)                         // shuffle
           + (b.n_chunk > 0) + 1;                   // encode, trailer
}

// descriptors (with device pointers) of every field, in plan order; fails
// when a field differs in shape, dtype, mode or bound kind from `ref`
int build_descriptors(const pa_batch_field* fields, int nf, const std::vector<int>& order, uint8_t* ws,
                      uint64_t* d_result, std::vector<BatchField>& out, const std::vector<BatchField>* ref) {
    out.resize(nf);
    size_t off = 0;
    for (int k = 0; k < nf; ++k) {
        const int i = order[k];
        BatchField& f = out[k];
        const int st = describe(fields[i], f);
        if (st != PA_OK) return st;
        if (ref) {
            const BatchField& r = (*ref)[k];
            if (r.dtype != f.dtype || r.mode != f.mode || r.rel != f.rel || r.L.nx != f.L.nx || r.L.ny != f.L.ny ||
                r.L.nz != f.L.nz)
                return PA_ERR_ARG;
        }
        const FieldSizes z = field_sizes(f);
        uint8_t* b = ws + off;
        f.minmax = reinterpret_cast<double*>(b);
        b += z.minmax;
        f.ebinfo = reinterpret_cast<double*>(b);
        b += z.ebinfo;
        f.flags = b;
        b += z.flags;
        f.scores = reinterpret_cast<double*>(b);
        b += z.scores;
        f.tune_part = reinterpret_cast<double*>(b);
        b += z.tune;
        f.qc = reinterpret_cast<int32_t*>(b);
        b += z.qc;
        f.anchors = b;
        b += z.anchors;
        f.bp = b;
        b += z.bp;
        f.encws = reinterpret_cast<uint64_t*>(b);
        b += z.encws;
        f.cum = reinterpret_cast<uint64_t*>(b);
        b += z.cum;
        f.outsize = reinterpret_cast<uint64_t*>(b);
        off += z.total();
        f.result = d_result + 4 * (size_t)i;
    }
    return PA_OK;
}

// descriptor and prefix tables into the workspace, ordered on `st`
int upload(pa_batch_plan* p, hipStream_t st) {
    const int nf = p->nf;
    std::vector<int64_t>& pre = p->pre;
    pre.assign((size_t)kPrefixes * (nf + 1), 0);
    int64_t* P[kPrefixes];
    for (int s = 0; s < kPrefixes; ++s) P[s] = pre.data() + (size_t)s * (nf + 1);
    for (int k = 0; k < nf; ++k) {
        const BatchField& f = p->host[k];
        const int64_t cnt[kPrefixes] = {f.rel ? pa::batch_ext_blocks(&f.L, f.dtype, pa::kExtChoices[0]) : 0,
                                        f.rel ? pa::batch_ext_blocks(&f.L, f.dtype, pa::kExtChoices[1]) : 0,
                                        f.rel ? pa::batch_ext_blocks(&f.L, f.dtype, pa::kExtChoices[2]) : 0,
                                        f.rel ? pa::batch_ext_blocks(&f.L, f.dtype, pa::kExtChoices[3]) : 0,
                                        f.tnblocks,
                                        pa::batch_gap_blocks(f),
                                        pa::batch_tile_units(&f.L),
                                        pa::batch_shuf_blocks(&f.L),
                                        (f.L.nchunks + 7) / 8 * 8,  // encoder launch slots: whole workgroups
                                        pa::batch_merge_blocks(&f.L, f.dtype)};
        for (int s = 0; s < kPrefixes; ++s) P[s][k + 1] = P[s][k] + cnt[s];
    }
    pa::BatchDev& d = p->dev;
    const size_t desc = up((size_t)nf * sizeof(BatchField));
    d.fields = reinterpret_cast<BatchField*>(p->table);
    d.nf = nf;
    d.nf32 = p->nf32;
    const int64_t* dp = reinterpret_cast<const int64_t*>(p->table + desc);
    for (int i = 0; i < 4; ++i) {
        p->ext_tab[i] = dp + i * (nf + 1);
        p->ext_n[i] = P[i][nf];
    }
    d.ext_batches = pa::kExtChoices[2];  // defaults until the autotuner runs
    d.encode_waves = 4;
    d.xcd_order = 1;
    d.p_ext = p->ext_tab[2];
    d.n_ext = p->ext_n[2];
    d.p_tune = dp + 4 * (nf + 1);
    d.p_gap = dp + 5 * (nf + 1);
    d.p_tile = dp + 6 * (nf + 1);
    d.p_shuf = dp + 7 * (nf + 1);
    d.p_chunk = dp + 8 * (nf + 1);
    d.p_merge = dp + 9 * (nf + 1);
    d.n_tune = P[4][nf];
    d.n_gap = P[5][nf];
    d.n_tile = P[6][nf];
    d.n_shuf = P[7][nf];
    d.n_chunk = P[8][nf];
    d.n_merge = P[9][nf];
    d.n_tune32 = P[4][p->nf32];
    d.n_tile32 = P[6][p->nf32];
    const int64_t lim = 0x7fffffff;
    // the front launch's grid holds every tuner, extrema (most for the
    // smallest load batch), gap and status-clear workgroup
    const int64_t front = 15 * d.n_tune + p->ext_n[0] + d.n_gap + (d.n_chunk + nf + 255) / 256;
    if (front > lim || d.n_tile > lim || d.n_shuf > lim || d.n_chunk + d.nf > lim || d.n_merge > lim)
        return PA_ERR_ARG;
    if (hipMemcpyAsync(p->table, p->host.data(), (size_t)nf * sizeof(BatchField), hipMemcpyHostToDevice, st) !=
        hipSuccess)
        return PA_ERR_HIP;
    if (hipMemcpyAsync(p->table + desc, pre.data(), pre.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess)
        return PA_ERR_HIP;
    const int e = pa::batch_keys_init(d, st);  // the finish kernel resets them after every run
    if (e != PA_OK) return e;
    p->launches = count_launches(d);
    return PA_OK;
}

int capture(pa_batch_plan* p) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return PA_ERR_HIP;
    int st = PA_OK;
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
        st = PA_ERR_HIP;
    } else {
        st = enqueue(p, s);
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (st == PA_OK && ec != hipSuccess) st = PA_ERR_HIP;
    }
    if (st == PA_OK && hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        p->exec = nullptr;
        st = PA_ERR_HIP;
    }
    if (g) hipGraphDestroy(g);
    hipStreamDestroy(s);
    return st;
}

}  // namespace

extern "C" size_t pa_batch_workspace(const pa_batch_field* fields, int nf) {
    if (!fields || nf <= 0) return 0;
    size_t s = table_bytes(nf);
    for (int i = 0; i < nf; ++i) {
        BatchField f;
        if (describe(fields[i], f) != PA_OK) return 0;
        s += field_sizes(f).total();
    }
    return s;
}

extern "C" int pa_batch_plan_create(const pa_batch_field* fields, int nf, void* d_ws, size_t ws_bytes,
                                    uint64_t* d_result, int use_graph, void* stream, pa_batch_plan** plan) {
    if (!fields || nf <= 0 || !d_result || !plan) return PA_ERR_ARG;
    *plan = nullptr;
    const size_t need = pa_batch_workspace(fields, nf);
    if (need == 0) return PA_ERR_ARG;
    if (!d_ws || ws_bytes < need) return PA_ERR_WORKSPACE;
    pa_batch_plan* p = new (std::nothrow) pa_batch_plan;
    if (!p) return PA_ERR_ARG;
    p->nf = nf;
    p->nf32 = plan_order(fields, nf, p->order);
    uint8_t* ws = static_cast<uint8_t*>(d_ws);
    int st = build_descriptors(fields, nf, p->order, ws + table_bytes(nf), d_result, p->host, nullptr);
    if (st == PA_OK) {
        p->table = ws;
        st = upload(p, pa::as_stream(stream));
    }
    if (st == PA_OK && use_graph) st = capture(p);
    if (st != PA_OK) {
        delete p;
        return st;
    }
    *plan = p;
    return PA_OK;
}

extern "C" int pa_batch_plan_rebind(pa_batch_plan* p, const pa_batch_field* fields, int nf, void* stream) {
    if (!p || !fields || nf != p->nf) return PA_ERR_ARG;
    hipStream_t hs = pa::as_stream(stream);
    // earlier uploads and runs read the tables being replaced
    if (hipStreamSynchronize(hs) != hipSuccess) return PA_ERR_HIP;
    std::vector<int> order;
    plan_order(fields, nf, order);
    if (order != p->order) return PA_ERR_ARG;
    // the descriptors keep their scratch: same shapes give the same offsets
    std::vector<BatchField> fs;
    uint8_t* scratch = reinterpret_cast<uint8_t*>(p->host[0].minmax);
    uint64_t* d_result = p->host[0].result - 4 * (size_t)p->order[0];
    const int st = build_descriptors(fields, nf, order, scratch, d_result, fs, &p->host);
    if (st != PA_OK) return st;
    p->host.swap(fs);
    if (hipMemcpyAsync(p->table, p->host.data(), (size_t)nf * sizeof(BatchField), hipMemcpyHostToDevice, hs) !=
        hipSuccess)
        return PA_ERR_HIP;
    return PA_OK;
}

extern "C" int pa_batch_run(pa_batch_plan* p, void* stream) {
    if (!p) return PA_ERR_ARG;
    hipStream_t st = pa::as_stream(stream);
    if (p->exec) return hipGraphLaunch(p->exec, st) == hipSuccess ? PA_OK : PA_ERR_HIP;
    return enqueue(p, st);
}

extern "C" int pa_batch_plan_launches(const pa_batch_plan* p) { return p ? p->launches : 0; }

// Launch-parameter autotuner: times each candidate of every tunable stage
// on the plan's own fields (the stages are idempotent once a full run has
// left their inputs in place), keeps the fastest, and captures the graph
// again with the choices.
extern "C" int pa_batch_plan_autotune(pa_batch_plan* p, int reps, void* stream) {
    if (!p || reps < 1) return PA_ERR_ARG;
    hipStream_t st = pa::as_stream(stream);
    int e = enqueue(p, st);
    if (e != PA_OK) return e;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return PA_ERR_HIP;
    int status = PA_OK;
    auto time = [&](int (*launch)(const pa::BatchDev&, hipStream_t)) -> float {
        if (status != PA_OK) return 0.f;
        int rc = launch(p->dev, st);  // warm-up
        if (rc == PA_OK && hipEventRecord(e0, st) != hipSuccess) rc = PA_ERR_HIP;
        for (int r = 0; r < reps && rc == PA_OK; ++r) rc = launch(p->dev, st);
        float ms = 0.f;
        if (rc == PA_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = PA_ERR_HIP;
        status = rc;
        return ms / reps;
    };
    pa::BatchDev& d = p->dev;
    int best = 2;
    for (int i = 0; i < 4; ++i) {
        d.ext_batches = pa::kExtChoices[i];
        d.p_ext = p->ext_tab[i];
        d.n_ext = p->ext_n[i];
        p->tuned_ms[i] = time(pa::batch_front);
        if (p->tuned_ms[i] < p->tuned_ms[best] || i == 0) best = i;
    }
    d.ext_batches = pa::kExtChoices[best];
    d.p_ext = p->ext_tab[best];
    d.n_ext = p->ext_n[best];
    for (int x = 1; x >= 0; --x) {
        d.xcd_order = x;
        p->tuned_ms[4 + (1 - x)] = time(pa::batch_spline);
    }
    d.xcd_order = p->tuned_ms[4] <= p->tuned_ms[5] ? 1 : 0;
    best = 0;
    for (int i = 0; i < 3; ++i) {
        d.encode_waves = pa::kEncodeChoices[i];
        p->tuned_ms[6 + i] = time(pa::batch_encode_chunks);
        if (p->tuned_ms[6 + i] < p->tuned_ms[6 + best]) best = i;
    }
    d.encode_waves = pa::kEncodeChoices[best];
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (status != PA_OK) return status;
    if (p->exec) {
        hipGraphExecDestroy(p->exec);
        p->exec = nullptr;
        return capture(p);
    }
    return PA_OK;
}

extern "C" int pa_batch_plan_params(const pa_batch_plan* p, int32_t* params3, float* timings9) {
    if (!p || !params3) return PA_ERR_ARG;
    params3[0] = p->dev.ext_batches;
    params3[1] = p->dev.xcd_order;
    params3[2] = p->dev.encode_waves;
    if (timings9)
        for (int i = 0; i < 9; ++i) timings9[i] = p->tuned_ms[i];
    return PA_OK;
}

extern "C" int pa_batch_plan_destroy(pa_batch_plan* p) {
    if (!p) return PA_OK;
    int st = PA_OK;
    if (p->exec && hipGraphExecDestroy(p->exec) != hipSuccess) st = PA_ERR_HIP;
    delete p;
    return st;
}

// ---------------------------------------------------------------- retrieval --
namespace {

using pa::RetrJob;

struct RetrTables {
    size_t jobs, pre, total;  // bytes: job table, two prefix tables
};
RetrTables retr_tables(int nj) {
    RetrTables t;
    t.jobs = up((size_t)nj * sizeof(RetrJob));
    t.pre = up((size_t)2 * (nj + 1) * 8);
    t.total = t.jobs + t.pre;
    return t;
}
size_t retr_keys_bytes(int nj) { return up((size_t)nj * 33 * 33 * sizeof(unsigned long long)); }

int retr_describe(const pa_retrieve_job& in, RetrJob& j) {
    if (!in.stream || !in.state || !in.bp || !in.qc || !in.dws || !in.out || (in.dtype != PA_F32 && in.dtype != PA_F64) ||
        !(in.eb > 0) || !(in.rel_eb >= 0) || in.round < 0 || (in.round > 0 && !in.old))
        return PA_ERR_ARG;
    j = RetrJob{};
    int st = pa_layout_init(&j.L, in.nx, in.ny, in.nz);
    if (st != PA_OK) return st;
    j.g = pa::make_geom(&j.L);
    j.cg = pa::make_chunk_geom(&j.L);
    j.lv = pa::make_levels(&j.L);
    j.dtype = in.dtype;
    // trailer: flags8 | range | cum[4][32] | anchors (compressor.cu:139-151)
    const int64_t t0 = in.stream_len - (j.L.trailer_fixed + j.L.n4 * in.dtype);
    if (t0 < 8 + 2 * j.L.nchunks || t0 % 8) return PA_ERR_STREAM;
    j.stream = in.stream;
    j.stream_len = in.stream_len;
    j.flags = in.stream + t0;
    j.cum = reinterpret_cast<const uint64_t*>(in.stream + t0 + 16);
    j.anchors = in.stream + t0 + 1040;
    // findStrategy_h 181, 193 (as pa_find_strategy)
    if (in.target * 0.954 >= in.eb) {
        const double q = in.target * 0.954 / in.eb;
        if (!(q < 1.8e19)) return PA_ERR_ARG;
        j.search = 1;
        j.target_cost = (uint64_t)__builtin_floor(q) - 1;
    }
    j.offsets_ready = in.round > 0;
    j.first = in.round == 0;
    pa::retr_level_ebs(j, in.eb, in.rel_eb);
    j.state = in.state;
    j.bp = in.bp;
    j.qc = in.qc;
    j.dws = static_cast<uint8_t*>(in.dws);
    j.bytes_read = in.bytes_read;
    j.old = in.old;
    j.out = in.out;
    return PA_OK;
}

}  // namespace

extern "C" size_t pa_retrieve_table_bytes(int nj) { return nj > 0 ? retr_tables(nj).total : 0; }

extern "C" size_t pa_retrieve_workspace(int nj) {
    if (nj <= 0) return 0;
    return retr_tables(nj).total + retr_keys_bytes(nj) + up((size_t)(nj + 1) * 8);
}

extern "C" int pa_retrieve_batch(const pa_retrieve_job* jobs, int nj, void* h_table, size_t table_bytes, void* d_ws,
                                 size_t ws_bytes, void* stream) {
    if (!jobs || nj <= 0 || nj > 65535) return PA_ERR_ARG;
    if (!d_ws || ws_bytes < pa_retrieve_workspace(nj)) return PA_ERR_WORKSPACE;
    const RetrTables tb = retr_tables(nj);
    if (h_table && table_bytes < tb.total) return PA_ERR_WORKSPACE;
    std::vector<uint8_t> local;
    uint8_t* ht = static_cast<uint8_t*>(h_table);
    if (!ht) {
        local.resize(tb.total);
        ht = local.data();
    }
    uint8_t* ws = static_cast<uint8_t*>(d_ws);
    RetrJob* hj = reinterpret_cast<RetrJob*>(ht);
    int64_t* p_unshuf = reinterpret_cast<int64_t*>(ht + tb.jobs);
    int64_t* p_unit = p_unshuf + (nj + 1);
    unsigned long long* keys = reinterpret_cast<unsigned long long*>(ws + tb.total);
    int64_t* p_chunk = reinterpret_cast<int64_t*>(ws + tb.total + retr_keys_bytes(nj));
    // float32 jobs first
    int k = 0, nj32 = 0;
    int64_t max_chunks = 0;
    p_unshuf[0] = p_unit[0] = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < nj; ++i) {
            if ((jobs[i].dtype == PA_F32) != (pass == 0)) continue;
            RetrJob& j = hj[k];
            const int st = retr_describe(jobs[i], j);
            if (st != PA_OK) return st;
            j.keys = keys + (size_t)k * 33 * 33;
            p_unshuf[k + 1] = p_unshuf[k] + pa::retr_unshuf_blocks(&j.L);
            p_unit[k + 1] = p_unit[k] + pa::batch_tile_units(&j.L);
            max_chunks += j.L.nchunks;
            nj32 += pass == 0;
            ++k;
        }
    pa::RetrDev r;
    r.jobs = reinterpret_cast<const RetrJob*>(ws);
    r.nj = nj;
    r.nj32 = nj32;
    r.p_unshuf = reinterpret_cast<const int64_t*>(ws + tb.jobs);
    r.p_unit = r.p_unshuf + (nj + 1);
    r.p_chunk = p_chunk;
    r.n_unshuf = p_unshuf[nj];
    r.n_unit = p_unit[nj];
    r.n_unit32 = p_unit[nj32];
    r.max_chunks = max_chunks;
    if (r.n_unshuf > 0x7fffffff || r.n_unit > 0x7fffffff) return PA_ERR_ARG;
    hipStream_t hs = pa::as_stream(stream);
    if (h_table) {
        if (hipMemcpyAsync(ws, ht, tb.total, hipMemcpyHostToDevice, hs) != hipSuccess) return PA_ERR_HIP;
    } else {
        if (hipStreamSynchronize(hs) != hipSuccess) return PA_ERR_HIP;
        if (hipMemcpy(ws, ht, tb.total, hipMemcpyHostToDevice) != hipSuccess) return PA_ERR_HIP;
    }
    int e;
    if ((e = pa::retr_strategy(r, hs)) != PA_OK) return e;
    if ((e = pa::retr_decode(r, hs)) != PA_OK) return e;
    if ((e = pa::retr_unshuffle(r, hs)) != PA_OK) return e;
    return pa::retr_reconstruct(r, hs);
}
