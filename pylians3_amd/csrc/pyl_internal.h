// Internal declarations shared by the libpylians_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pylians_hip.h"

namespace pyl {

// ---- error handling: thread-local message, int status codes --------------
void set_error(const char *fmt, ...);

#define PYL_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            ::pyl::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,            \
                             hipGetErrorString(_e));                               \
            return (_e == hipErrorOutOfMemory) ? PYL_ENOMEM : PYL_EHIP;            \
        }                                                                          \
    } while (0)

#define PYL_CHECK(cond, code, ...)                                                 \
    do {                                                                           \
        if (!(cond)) {                                                             \
            ::pyl::set_error(__VA_ARGS__);                                         \
            return (code);                                                         \
        }                                                                          \
    } while (0)

#define PYL_TRY(expr)                                                              \
    do {                                                                           \
        int _s = (expr);                                                           \
        if (_s != PYL_OK) return _s;                                               \
    } while (0)

// ---- roctx range around a C-ABI call (seen by rocprofv3 --marker-trace) ----
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

// ---- grow-only device buffer owned by a context ---------------------------
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    unsigned long long gen = 0;   // allocation number (Ctx::alloc_gen) of ptr; 0: none
};

// Pinned host staging (hipHostMalloc): small per-call uploads and result
// downloads go through it, so the copies are plain DMA that do not block the
// host.  `done` is recorded after the copies that read or fill it; the next
// user waits for it before touching the memory.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

// ---- resident plans --------------------------------------------------------
struct FFTPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    void *work = nullptr;
    size_t work_bytes = 0;
    int dims = 0;
    int batch = 0;
};

// Exact mode counts of the reference loop for one (dims, axis), DC bins kept.
struct ModePlan {
    int dims = 0, axis = 0;
    std::vector<double> N1D, N2D, N3D, k1D, k3D;
};

// Rows of the half-complex cube grouped by binning key and cut into work items
// (device-resident).
struct BinPlan {
    int dims = 0, axis = 0, R = 0;
    int nitems = 0, nkeys = 0;
    int *d_rows = nullptr;       // kxx*dims + kyy, grouped by key
    int4 *d_items = nullptr;     // (first row, nrows, key, 0)
    int *d_key_items = nullptr;  // [nkeys+1] item range of each key
};

enum BufId {
    BUF_POS = 0,     // staged positions (host-pointer path)
    BUF_W,           // staged weights
    BUF_GRID,        // staged grid
    BUF_TILE_CNT,    // per-tile counters / cursors / offsets / weight bounds
    BUF_RECORDS,     // tile-sorted particle records
    BUF_SCAN_TMP,
    BUF_FFT_IN,      // real FFT input (copy of delta)
    BUF_FFT_OUT,     // half-complex FFT output
    BUF_BINS,        // per-bin accumulators
    BUF_MAS_TAB,     // MAS correction tables
    BUF_VEL,
    BUF_PARTIAL,     // per-tile halo partial sums of the tiled deposit
    BUF_TILE_BOX,    // per-tile int64 boxes of a split deposit
    BUF_BUCKETED,    // records of the two-level sort's first pass
    BUF_PK_AMAX,     // per-cube largest |Re|, |Im| of the transforms being binned
    BUF_CAPREC,      // per-tile record slots of the capacity sort
    BUF_SPILL,       // its spill list, when the bucketed records hold the input
    BUF_COUNT_
};

struct Ctx {
    int device = 0;
    int ncu = 0;                                             // compute units of the device
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    size_t budget = 0;
    DevBuf bufs[BUF_COUNT_];
    std::map<std::tuple<int, int, int>, FFTPlan> fft_plans; // (dims, batch, kind)
    unsigned long long alloc_gen = 0;                        // device allocations so far
    // The chunk table and tile origins last uploaded into BUF_TILE_CNT, and what they
    // were built from: a call with the same particle ranges and grid reuses them.
    struct TileTables {
        unsigned long long gen = 0;          // BUF_TILE_CNT allocation they live in
        const void *at = nullptr;            // the chunk table's address in it (the layout
                                             // before it depends on the sort form)
        long long n = -1, nchunk = -1, target = -1;
        int nreal = 0, dims = 0, axes = 0, mas = 0;
        std::vector<long long> offs;         // realisation offsets (empty: one realisation)
    } tile_tables;
    PinnedBuf pin_up;                                        // MA chunk table + tile origins
    PinnedBuf pin_down;                                      // Pk bin sums
    PinnedBuf pin_scales;                                    // Pk batch scale/shift upload
    // Tile-sort consistency check of the latest tiled MA, read back without a
    // synchronisation of its own: checked at the next synchronising call.
    // pinned: [0] record total, [1] box misses, [2] particle total of device offsets,
    // [3] particles the capacity sort spilled
    unsigned long long *h_total = nullptr;
    long long box_miss_n = -1;       // particles of the call whose box misses h_total[1] will hold
    std::pair<int, int> box_miss_key{0, 0};           // its (dims, MAS)
    std::map<std::pair<int, int>, bool> shuffled_hint; // per (dims, MAS): the latest call's particles
                                                        // were out of order (two-level sort next)
    long long cap_spill_n = -1;      // particles of the capacity call whose spill h_total[3] will hold
    std::pair<int, int> cap_key{0, 0};                // its (dims, MAS)
    std::map<std::pair<int, int>, int> cap_bad;       // per (dims, MAS): calls left without the
                                                      // capacity form after one spilled too much
    int last_window = -1;            // sort window of the latest tiled MA (PYL_WINDOW_*)
    long long total_expect = -1;
    hipEvent_t ev_total = nullptr;
    // Recorded on the outgoing stream when pyl_ctx_set_stream changes streams; the
    // incoming stream waits for it, so work of the next call (which reuses the same
    // scratch buffers) starts only after the previous call's work is done.
    hipEvent_t ev_switch = nullptr;
    // Fixed device scratch of pyl_ma_noalloc: while such a call runs, `arena` is
    // this device's address of the library's fixed array and MA scratch is cut
    // from it (arena_used bytes so far) instead of the grow-only buffers.
    char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    char *noalloc_base = nullptr;    // the array's address on this context's device (set-up)
    // HIP graph capture (pyl_graph_begin .. pyl_graph_end): calls only enqueue work;
    // host steps that need the device's results are deferred and run, in order,
    // after each replay.  Scratch may not grow while capturing.
    bool capturing = false;
    std::vector<std::function<int()>> deferred;
    std::vector<void *> cap_host;    // pinned buffers the captured work reads or fills
    // Batched MA with particle ranges in device memory (pyl_ma_batch_dev): the sort
    // chunks are cut on the device, chunks_per_real per realisation.
    const long long *dev_offs = nullptr;
    int chunks_per_real = 0;
    // Launch parameters a tuner may set (0: the built-in choice).
    int tune_bin_rows = 0;           // binning rows per work item
    int tune_order = -1;             // largest-first deposit order: -1 auto, 0 off, 1 on
    int tune_window = -1;            // MA sort: -1 adaptive, 0 one-pass, 1 wide, 2 two-level, 3 capacity,
                                     // 4 two-level with a capacity second pass
    // Stage timing: a ring of event pairs per stage, so that timing many calls
    // needs no host synchronisation between them.  A pair is read back when its
    // ring slot comes round again (kEvRing calls later) or on request.
    static constexpr int kEvRing = 64;
    bool timing = false;
    unsigned timing_mask = 0;                          // stages recorded when timing
    hipEvent_t ev[PYL_STAGE_COUNT][kEvRing][2] = {};
    int ev_slot[PYL_STAGE_COUNT] = {};                 // slot of the latest call
    bool ev_used[PYL_STAGE_COUNT] = {};                // stage ran in the latest call
    bool ev_pending[PYL_STAGE_COUNT][kEvRing] = {};    // recorded, not yet added to the totals
    double ev_total_ms[PYL_STAGE_COUNT] = {};
    long long ev_count[PYL_STAGE_COUNT] = {};
    bool ev_ready = false;
};

// Tables that depend only on the grid, built once per process and shared by every
// context (the device tables once per device), so that the slots of a server or the
// threads of a pipeline do not each rebuild them: mode counts per (dims, axis),
// binning plans per (device, dims, axis, rows per item), MAS corrections per (device,
// dims, p) and FFT twiddles exp(-2 pi i k/N) per (device, N).  Built under `mu` and
// never changed or freed afterwards (they live until the process ends).
struct SharedTables {
    std::mutex mu;
    std::map<std::pair<int, int>, ModePlan> mode;
    std::map<std::tuple<int, int, int, int>, BinPlan> bin;
    std::map<std::tuple<int, int, int>, void *> mas;
    std::map<std::pair<int, int>, void *> twiddles;
};
SharedTables &shared_tables();

size_t held_bytes(const Ctx *c);
// Looks up this device's address of the fixed MA scratch array (once per device,
// with the device current); called by pyl_ctx_create.
int ma_noalloc_setup(Ctx *c);
int ensure(Ctx *c, BufId id, size_t bytes);
void reset_stages(Ctx *c);
int stage_begin(Ctx *c, int stage);
int stage_end(Ctx *c, int stage);
int finish(Ctx *c, int flags);   // synchronise unless PYL_NO_SYNC on device pointers
// Waits for the previous user of `b`, grows it to `bytes` and returns it.
int pinned(Ctx *c, PinnedBuf &b, size_t bytes, void **out);
int pinned_done(Ctx *c, PinnedBuf &b);   // records b.done after the copies just enqueued
int check_total(Ctx *c);                 // reports a tile sort that lost particles

// FFT plan (resident, one per (dims, batch))
// Transform kinds of the resident rocFFT plans (all single precision, out of place).
enum FFTKind { FFT_R2C_3D = 0, FFT_C2R_3D = 1, FFT_R2C_2D = 2, FFT_R2C_1D = 3 };
int get_fft_plan(Ctx *c, int dims, int batch, FFTPlan **out, int kind = FFT_R2C_3D);
// Forward r2c of `batch` contiguous cubes (own passes for power-of-two sizes).
// Rows of N/2+1 modes zs apart (zs = 0: dense); fft_zstride(dims): the padded
// stride of the library's own half-complex buffers.
int fft_r2c_3d(Ctx *c, float *in, float2 *out, int dims, int batch, long long zs = 0,
               unsigned *amax = nullptr);
// Largest |Re| or |Im| of ncube half-complex cubes (rows zs apart, cube stride
// cstride) as float bits, maxed into 64 words per cube (zeroed by the caller).
int complex_absmax(Ctx *c, const float2 *dk, int dims, long long zs, int ncube, long long cstride,
                   unsigned *amax);
long long fft_zstride(int dims);
bool fft_own_size(int dims);
int run_fft(Ctx *c, FFTPlan *p, void *in, void *out);

// Host-side MAS tables: out[kk] = MAS_correction(pi/dims * k(kk), p) for the
// signed frequency k(kk) of FFT index kk, with the reference's libm calls
// (Pk_library.pyx:83-84, prefact :310).
inline void mas_table(int dims, int p, double *out) {
    const int m = dims / 2;
    const double prefact = M_PI / dims;
    for (int kk = 0; kk < dims; ++kk) {
        int k = kk > m ? kk - dims : kk;
        double x = prefact * k;
        out[kk] = (x == 0.0) ? 1.0 : std::pow(x / std::sin(x), (double)p);
    }
}

// Redshift-space shift of one coordinate, redshift_space_library.pyx:40-46:
// p + v*factor (float), "while p < 0: p += BoxSize", "if p > BoxSize: p = p % BoxSize"
// (C fmod under cdivision).  The while loop is bounded: after 64 box lengths the
// remainder is folded with fmodf.
__device__ __forceinline__ float rsd_shift(float p, float v, float factor, float box) {
    p = p + v * factor;
    int guard = 0;
    while (p < 0.0f && guard < 64) {
        p += box;
        ++guard;
    }
    if (p < 0.0f) p = fmodf(p, box) + box;
    if (p > box) p = fmodf(p, box);
    return p;
}

// Mode-count plan (host computed, cached)
const ModePlan *get_mode_plan(Ctx *c, int dims, int axis);
void compute_mode_plan(int dims, int axis, ModePlan *mp);

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

}  // namespace pyl

struct pyl_ctx {
    pyl::Ctx c;
};
