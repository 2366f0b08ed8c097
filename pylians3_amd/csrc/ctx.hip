// Context, error reporting, scratch memory, stage timers and resident rocFFT plans.
#include "pyl_internal.h"

#include <cstring>
#include <mutex>

namespace pyl {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

SharedTables &shared_tables() {
    static SharedTables *t = new SharedTables();   // never destroyed: used until exit
    return *t;
}

size_t held_bytes(const Ctx *c) {
    size_t s = 0;
    for (int i = 0; i < BUF_COUNT_; ++i) s += c->bufs[i].cap;
    for (auto &kv : c->fft_plans) s += kv.second.work_bytes;
    return s;
}

// Grow-only scratch: reallocates only when the request exceeds the capacity.
int ensure(Ctx *c, BufId id, size_t bytes) {
    DevBuf &b = c->bufs[id];
    if (bytes <= b.cap) return PYL_OK;
    PYL_CHECK(!c->capturing, PYL_EINVAL,
              "scratch buffer %d would grow while a graph is captured: run the same work once "
              "on the context before capturing it", (int)id);
    size_t need = bytes + (bytes >> 4);  // 6% headroom so slowly growing requests do not thrash
    size_t others = held_bytes(c) - b.cap;
    PYL_CHECK(others + need <= c->budget || others + bytes <= c->budget, PYL_EBUDGET,
              "request of %zu bytes exceeds the context HBM budget (%zu held, %zu budget)",
              bytes, others, c->budget);
    if (others + need > c->budget) need = bytes;
    if (b.ptr) {
        PYL_HIP(hipStreamSynchronize(c->stream));
        PYL_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
    }
    hipError_t e = hipMalloc(&b.ptr, need);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        b.gen = 0;
        set_error("hipMalloc(%zu) failed: %s", need, hipGetErrorString(e));
        return PYL_ENOMEM;
    }
    b.cap = need;
    b.gen = ++c->alloc_gen;
    return PYL_OK;
}

// A stage's time spans its first begin and its last end within one API call.
void reset_stages(Ctx *c) {
    for (int s = 0; s < PYL_STAGE_COUNT; ++s) c->ev_used[s] = false;
}

// Adds a recorded pair to the stage totals (waits for its end event).
// A pair whose call failed before its end event was recorded is dropped.
static int harvest(Ctx *c, int stage, int slot) {
    if (!c->ev_pending[stage][slot]) return PYL_OK;
    c->ev_pending[stage][slot] = false;
    PYL_HIP(hipEventSynchronize(c->ev[stage][slot][1]));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[stage][slot][0], c->ev[stage][slot][1]) != hipSuccess) {
        (void)hipGetLastError();
        return PYL_OK;
    }
    c->ev_total_ms[stage] += ms;
    c->ev_count[stage] += 1;
    return PYL_OK;
}

int stage_begin(Ctx *c, int stage) {
    if (!c->timing || !((c->timing_mask >> stage) & 1u) || c->ev_used[stage]) return PYL_OK;
    const int slot = (c->ev_slot[stage] + 1) % Ctx::kEvRing;
    PYL_TRY(harvest(c, stage, slot));             // recorded kEvRing calls ago
    PYL_HIP(hipEventRecord(c->ev[stage][slot][0], c->stream));
    c->ev_slot[stage] = slot;
    c->ev_used[stage] = true;
    c->ev_pending[stage][slot] = true;
    return PYL_OK;
}

int stage_end(Ctx *c, int stage) {
    if (!c->timing || !c->ev_used[stage]) return PYL_OK;
    PYL_HIP(hipEventRecord(c->ev[stage][c->ev_slot[stage]][1], c->stream));
    return PYL_OK;
}

int finish(Ctx *c, int flags) {
    if (c->capturing) return PYL_OK;
    if ((flags & PYL_PTR_DEVICE) && (flags & PYL_NO_SYNC)) return PYL_OK;
    PYL_HIP(hipStreamSynchronize(c->stream));
    return check_total(c);
}

int pinned(Ctx *c, PinnedBuf &b, size_t bytes, void **out) {
    PYL_CHECK(!c->capturing || bytes <= b.cap, PYL_EINVAL,
              "pinned staging would grow while a graph is captured");
    if (b.pending && !c->capturing) {
        PYL_HIP(hipEventSynchronize(b.done));
        b.pending = false;
    }
    if (bytes > b.cap) {
        if (b.ptr) PYL_HIP(hipHostFree(b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
        const size_t need = bytes + (bytes >> 3);
        hipError_t e = hipHostMalloc(&b.ptr, need, hipHostMallocDefault);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            set_error("hipHostMalloc(%zu) failed: %s", need, hipGetErrorString(e));
            return PYL_ENOMEM;
        }
        b.cap = need;
    }
    *out = b.ptr;
    return PYL_OK;
}

int pinned_done(Ctx *c, PinnedBuf &b) {
    if (c->capturing) return PYL_OK;      // a replay is waited for before its host steps
    PYL_HIP(hipEventRecord(b.done, c->stream));
    b.pending = true;
    return PYL_OK;
}

// total_expect: -1 nothing pending; >= 0 the expected record total; -2 the total
// of the device particle ranges, read back into h_total[2] (pyl_ma_batch_dev).
int check_total(Ctx *c) {
    if (c->total_expect == -1 || c->capturing) return PYL_OK;
    const long long mode = c->total_expect;
    c->total_expect = -1;
    PYL_HIP(hipEventSynchronize(c->ev_total));
    const unsigned long long expect = mode == -2 ? c->h_total[2] : (unsigned long long)mode;
    PYL_CHECK(c->h_total[0] == expect, PYL_EHIP, "tile sort lost particles (%llu of %llu)",
              c->h_total[0], expect);
    return PYL_OK;
}

static void destroy_plan(FFTPlan &p) {
    if (p.info) rocfft_execution_info_destroy(p.info);
    if (p.plan) rocfft_plan_destroy(p.plan);
    if (p.work) (void)hipFree(p.work);
    p = FFTPlan();
}

static std::once_flag g_rocfft_once;

// Resident rocFFT plan: forward, single precision, out-of-place r2c of `batch`
// contiguous dims^3 cubes.  Created once per (dims, batch) and kept.
int get_fft_plan(Ctx *c, int dims, int batch, FFTPlan **out, int kind) {
    auto key = std::make_tuple(dims, batch, kind);
    auto it = c->fft_plans.find(key);
    if (it != c->fft_plans.end()) {
        *out = &it->second;
        return PYL_OK;
    }
    std::call_once(g_rocfft_once, [] { rocfft_setup(); });
    FFTPlan p;
    p.dims = dims;
    p.batch = batch;
    size_t lengths[3] = {(size_t)dims, (size_t)dims, (size_t)dims};  // fastest first
    const size_t rank = (kind == FFT_R2C_2D) ? 2 : (kind == FFT_R2C_1D ? 1 : 3);
    const rocfft_transform_type type = (kind == FFT_C2R_3D) ? rocfft_transform_type_real_inverse
                                                            : rocfft_transform_type_real_forward;
    rocfft_status st = rocfft_plan_create(&p.plan, rocfft_placement_notinplace, type,
                                          rocfft_precision_single, rank, lengths, (size_t)batch,
                                          nullptr);
    PYL_CHECK(st == rocfft_status_success, PYL_EFFT, "rocfft_plan_create(%d^%d, kind %d) failed (%d)",
              dims, (int)rank, kind, (int)st);
    st = rocfft_plan_get_work_buffer_size(p.plan, &p.work_bytes);
    if (st != rocfft_status_success) {
        destroy_plan(p);
        set_error("rocfft_plan_get_work_buffer_size failed (%d)", (int)st);
        return PYL_EFFT;
    }
    st = rocfft_execution_info_create(&p.info);
    if (st != rocfft_status_success) {
        destroy_plan(p);
        set_error("rocfft_execution_info_create failed (%d)", (int)st);
        return PYL_EFFT;
    }
    if (p.work_bytes) {
        if (held_bytes(c) + p.work_bytes > c->budget) {
            size_t wb = p.work_bytes;
            destroy_plan(p);
            set_error("rocFFT work buffer of %zu bytes exceeds the HBM budget", wb);
            return PYL_EBUDGET;
        }
        hipError_t e = hipMalloc(&p.work, p.work_bytes);
        if (e != hipSuccess) {
            destroy_plan(p);
            set_error("hipMalloc(rocFFT work) failed: %s", hipGetErrorString(e));
            return PYL_ENOMEM;
        }
        rocfft_execution_info_set_work_buffer(p.info, p.work, p.work_bytes);
    }
    auto ins = c->fft_plans.emplace(key, p);
    *out = &ins.first->second;
    return PYL_OK;
}

int run_fft(Ctx *c, FFTPlan *p, void *in, void *out) {
    rocfft_status st = rocfft_execution_info_set_stream(p->info, c->stream);
    PYL_CHECK(st == rocfft_status_success, PYL_EFFT, "rocfft set_stream failed (%d)", (int)st);
    void *ib[1] = {in};
    void *ob[1] = {out};
    st = rocfft_execute(p->plan, ib, ob, p->info);
    PYL_CHECK(st == rocfft_status_success, PYL_EFFT, "rocfft_execute failed (%d)", (int)st);
    return PYL_OK;
}

}  // namespace pyl

using namespace pyl;

extern "C" {

const char *pyl_last_error(void) { return g_err.c_str(); }

int pyl_version(void) { return 100; }

int pyl_ctx_create(pyl_ctx **out, int device, size_t hbm_budget) {
    PYL_CHECK(out != nullptr, PYL_EINVAL, "pyl_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    PYL_HIP(hipGetDeviceCount(&ndev));
    PYL_CHECK(device >= 0 && device < ndev, PYL_EINVAL, "device %d not present (%d devices)",
              device, ndev);
    PYL_HIP(hipSetDevice(device));
    std::unique_ptr<pyl_ctx> h(new pyl_ctx());
    Ctx *c = &h->c;
    c->device = device;
    PYL_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
    PYL_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    if (hbm_budget == 0) {
        size_t fr = 0, tot = 0;
        PYL_HIP(hipMemGetInfo(&fr, &tot));
        hbm_budget = (size_t)(0.9 * (double)fr);
    }
    c->budget = hbm_budget;
    for (int s = 0; s < PYL_STAGE_COUNT; ++s)
        for (int k = 0; k < Ctx::kEvRing; ++k) {
            PYL_HIP(hipEventCreate(&c->ev[s][k][0]));
            PYL_HIP(hipEventCreate(&c->ev[s][k][1]));
        }
    c->ev_ready = true;
    PYL_HIP(hipEventCreateWithFlags(&c->pin_up.done, hipEventDisableTiming));
    PYL_HIP(hipEventCreateWithFlags(&c->pin_down.done, hipEventDisableTiming));
    PYL_HIP(hipEventCreateWithFlags(&c->pin_scales.done, hipEventDisableTiming));
    PYL_HIP(hipEventCreateWithFlags(&c->ev_total, hipEventDisableTiming));
    PYL_HIP(hipEventCreateWithFlags(&c->ev_switch, hipEventDisableTiming));
    PYL_HIP(hipHostMalloc((void **)&c->h_total, 4 * sizeof(unsigned long long), hipHostMallocDefault));
    c->h_total[1] = c->h_total[2] = c->h_total[3] = 0;
    PYL_TRY(ma_noalloc_setup(c));
    *out = h.release();
    return PYL_OK;
}

int pyl_ctx_trim(pyl_ctx *ctx) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    PYL_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < BUF_COUNT_; ++i) {
        if (c->bufs[i].ptr) PYL_HIP(hipFree(c->bufs[i].ptr));
        c->bufs[i] = DevBuf();
    }
    for (auto &kv : c->fft_plans) destroy_plan(kv.second);
    c->fft_plans.clear();
    for (PinnedBuf *b : {&c->pin_up, &c->pin_down, &c->pin_scales}) {
        if (b->ptr) (void)hipHostFree(b->ptr);
        b->ptr = nullptr;
        b->cap = 0;
        b->pending = false;
    }
    return PYL_OK;
}

int pyl_ctx_destroy(pyl_ctx *ctx) {
    if (!ctx) return PYL_OK;
    int s = pyl_ctx_trim(ctx);
    Ctx *c = &ctx->c;
    if (c->ev_ready)
        for (int i = 0; i < PYL_STAGE_COUNT; ++i)
            for (int k = 0; k < Ctx::kEvRing; ++k) {
                if (c->ev[i][k][0]) (void)hipEventDestroy(c->ev[i][k][0]);
                if (c->ev[i][k][1]) (void)hipEventDestroy(c->ev[i][k][1]);
            }
    for (hipEvent_t e : {c->pin_up.done, c->pin_down.done, c->pin_scales.done, c->ev_total, c->ev_switch})
        if (e) (void)hipEventDestroy(e);
    if (c->h_total) (void)hipHostFree(c->h_total);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete ctx;
    return s;
}

int pyl_ctx_set_stream(pyl_ctx *ctx, void *hip_stream) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    hipStream_t next = hip_stream ? (hipStream_t)hip_stream : c->own_stream;
    if (next != c->stream) {
        // The scratch buffers are shared by every call on the context: the new
        // stream must not start on them while the old stream's kernels still run.
        PYL_HIP(hipSetDevice(c->device));
        PYL_HIP(hipEventRecord(c->ev_switch, c->stream));
        PYL_HIP(hipStreamWaitEvent(next, c->ev_switch, 0));
    }
    c->stream = next;
    return PYL_OK;
}

int pyl_ctx_synchronize(pyl_ctx *ctx) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    PYL_HIP(hipSetDevice(ctx->c.device));
    PYL_HIP(hipStreamSynchronize(ctx->c.stream));
    return check_total(&ctx->c);
}

int pyl_ctx_memory(pyl_ctx *ctx, size_t *bytes_out) {
    PYL_CHECK(ctx && bytes_out, PYL_EINVAL, "NULL argument");
    *bytes_out = held_bytes(&ctx->c);
    return PYL_OK;
}

int pyl_ctx_enable_timing(pyl_ctx *ctx, int enable) {
    return pyl_ctx_enable_stage_timing(ctx, enable ? ~0u : 0u);
}

int pyl_ctx_enable_stage_timing(pyl_ctx *ctx, unsigned stage_mask) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    PYL_HIP(hipStreamSynchronize(c->stream));   // no recorded pair is still in flight
    c->timing = stage_mask != 0;
    c->timing_mask = stage_mask;
    for (int s = 0; s < PYL_STAGE_COUNT; ++s) {
        c->ev_used[s] = false;
        c->ev_total_ms[s] = 0.0;
        c->ev_count[s] = 0;
        for (int k = 0; k < Ctx::kEvRing; ++k) c->ev_pending[s][k] = false;
    }
    return PYL_OK;
}

int pyl_ctx_stage_times(pyl_ctx *ctx, double *ms_out) {
    PYL_CHECK(ctx && ms_out, PYL_EINVAL, "NULL argument");
    Ctx *c = &ctx->c;
    PYL_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < PYL_STAGE_COUNT; ++s) {
        ms_out[s] = 0.0;
        if (!c->timing || !c->ev_used[s]) continue;
        const int k = c->ev_slot[s];
        float ms = 0.f;
        PYL_HIP(hipEventElapsedTime(&ms, c->ev[s][k][0], c->ev[s][k][1]));
        ms_out[s] = ms;
    }
    return PYL_OK;
}

int pyl_ctx_stage_totals(pyl_ctx *ctx, double *ms_out, long long *count_out) {
    PYL_CHECK(ctx && ms_out && count_out, PYL_EINVAL, "NULL argument");
    Ctx *c = &ctx->c;
    PYL_HIP(hipStreamSynchronize(c->stream));
    for (int s = 0; s < PYL_STAGE_COUNT; ++s) {
        for (int k = 0; k < Ctx::kEvRing; ++k) PYL_TRY(harvest(c, s, k));
        ms_out[s] = c->ev_total_ms[s];
        count_out[s] = c->ev_count[s];
    }
    return PYL_OK;
}

// ---- HIP graph capture -------------------------------------------------------

struct pyl_graph {
    pyl_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::function<int()>> deferred;   // host steps after each replay, in order
    std::vector<void *> host;                     // pinned buffers read or filled by the graph
};

static void drop_capture_state(Ctx *c) {
    for (void *p : c->cap_host) (void)hipHostFree(p);
    c->cap_host.clear();
    c->deferred.clear();
    c->capturing = false;
}

int pyl_graph_begin(pyl_ctx *ctx) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    PYL_CHECK(!c->capturing, PYL_EINVAL, "a graph is already being captured on this context");
    PYL_HIP(hipSetDevice(c->device));
    PYL_HIP(hipStreamSynchronize(c->stream));
    PYL_TRY(check_total(c));
    for (PinnedBuf *b : {&c->pin_up, &c->pin_down, &c->pin_scales}) b->pending = false;
    c->box_miss_n = -1;
    c->cap_spill_n = -1;
    c->timing = false;
    drop_capture_state(c);
    PYL_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    return PYL_OK;
}

int pyl_graph_end(pyl_ctx *ctx, pyl_graph **out) {
    PYL_CHECK(ctx != nullptr && out != nullptr, PYL_EINVAL, "NULL argument");
    *out = nullptr;
    Ctx *c = &ctx->c;
    PYL_CHECK(c->capturing, PYL_EINVAL, "no graph is being captured on this context");
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (e != hipSuccess || graph == nullptr) {
        drop_capture_state(c);
        set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
        return PYL_EHIP;
    }
    std::unique_ptr<pyl_graph> g(new pyl_graph());
    g->ctx = ctx;
    g->graph = graph;
    g->deferred = std::move(c->deferred);
    g->host = std::move(c->cap_host);
    c->deferred.clear();
    c->cap_host.clear();
    c->capturing = false;
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        g->exec = nullptr;
        pyl_graph_destroy(g.release());
        set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
        return PYL_EHIP;
    }
    *out = g.release();
    return PYL_OK;
}

int pyl_graph_abort(pyl_ctx *ctx) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    if (!c->capturing) return PYL_OK;
    hipGraph_t graph = nullptr;
    (void)hipStreamEndCapture(c->stream, &graph);
    (void)hipGetLastError();
    if (graph) (void)hipGraphDestroy(graph);
    drop_capture_state(c);
    return PYL_OK;
}

int pyl_graph_launch(pyl_graph *g) {
    TraceRange trace_("pyl_graph_launch");
    PYL_CHECK(g != nullptr && g->exec != nullptr, PYL_EINVAL, "NULL graph");
    Ctx *c = &g->ctx->c;
    PYL_CHECK(!c->capturing, PYL_EINVAL, "the context is capturing another graph");
    PYL_HIP(hipSetDevice(c->device));
    PYL_HIP(hipGraphLaunch(g->exec, c->stream));
    PYL_HIP(hipStreamSynchronize(c->stream));
    for (auto &f : g->deferred) PYL_TRY(f());
    return PYL_OK;
}

int pyl_graph_destroy(pyl_graph *g) {
    if (!g) return PYL_OK;
    if (g->ctx) {
        (void)hipSetDevice(g->ctx->c.device);
        (void)hipStreamSynchronize(g->ctx->c.stream);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (void *p : g->host) (void)hipHostFree(p);
    delete g;
    return PYL_OK;
}

int pyl_ctx_last_sort_window(pyl_ctx *ctx, int *window) {
    PYL_CHECK(ctx != nullptr && window != nullptr, PYL_EINVAL, "NULL argument");
    *window = ctx->c.last_window;
    return PYL_OK;
}

int pyl_ctx_set_tuning(pyl_ctx *ctx, int knob, int value) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    switch (knob) {
        case PYL_TUNE_BIN_ROWS:
            PYL_CHECK(value >= 0 && value <= 256, PYL_EINVAL, "bin rows must be 0..256");
            c->tune_bin_rows = value;
            return PYL_OK;
        case PYL_TUNE_DEPOSIT_ORDER:
            PYL_CHECK(value >= -1 && value <= 1, PYL_EINVAL, "deposit order must be -1, 0 or 1");
            c->tune_order = value;
            return PYL_OK;
        case PYL_TUNE_SORT_WINDOW:
            PYL_CHECK(value >= -1 && value <= 4, PYL_EINVAL, "sort window must be -1 .. 4");
            c->tune_window = value;
            return PYL_OK;
    }
    set_error("unknown tuning knob %d", knob);
    return PYL_EINVAL;
}

}  // extern "C"
