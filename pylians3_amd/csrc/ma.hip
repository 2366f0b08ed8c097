// Mass assignment (NGP/CIC/TSC/PCS, optional weights, 2D/3D) on gfx950.
//
// Reference semantics: library/MAS_library/MAS_library.pyx:57-545 (MA and the
// per-scheme kernels) and library/MAS_library/MAS_c.c:8-266 (C kernels).  The
// per-axis footprint below reproduces the reference's float/double expression
// types exactly (file compiled with -ffp-contract=off).
//
// Two algorithms:
//  * atomic: one thread per particle, one global float atomic per footprint
//    node (the plain GPU form; summation order is not reproducible).
//  * tiled (default): the grid is cut into tiles (16^3 in 3D, 64^2 in 2D).
//    A particle belongs to the tile of its first footprint node (its home
//    tile).  A counting sort groups the particles by home tile; one workgroup
//    per tile then adds the full footprints of its particles into an LDS box
//    that is the tile plus an upper halo of F-1 nodes per axis, so no node
//    test is needed.  Box cells that no other tile touches (the tile minus its
//    lower band of F-1 nodes) are written to the grid directly; the band and
//    the halo go to a partial-sum buffer, and a finalize pass adds, for every
//    band cell, the tile's own value and the halo values of its lower
//    neighbours in a fixed order.
//    Inside a box, contributions are truncated to integer multiples of a
//    power-of-two quantum 2^q (2^-30 of the realisation's largest |W|) and
//    summed with 64-bit integer LDS atomics: exact, so the box does not
//    depend on the order of the atomics.  With the fixed-order finalize the
//    grid is bitwise reproducible, and identical whether a realisation is
//    deposited alone or in a batch.  The truncation is below 2^-30 of the
//    weight scale per contribution, well under float32 resolution of a cell.
#include "pyl_internal.h"

namespace pyl {
namespace {

constexpr int kBlock = 256;
constexpr int kDepBlock = 512;     // deposit workgroup (one tile)

__host__ __device__ __forceinline__ int wrapi(int i, int d) {
    int r = i % d;
    return r < 0 ? r + d : r;
}

// raw mod d: two selects for the values a footprint of a position in [-L, 2L)
// produces, a division only for anything further out.
__host__ __device__ __forceinline__ int wrap_node(int raw, int d) {
    int r = raw + (raw < 0 ? d : 0);
    r -= (r >= d ? d : 0);
    if ((unsigned)r >= (unsigned)d) r = wrapi(raw, d);
    return r;
}

// (base + j) mod d for base in [0,d) and 0 <= j < 4; the second test only
// fires on grids smaller than the footprint (d < 4).
__host__ __device__ __forceinline__ int node_add(int base, int j, int d) {
    int r = base + j;
    if (r >= d) r -= d;
    if (r >= d) r %= d;
    return r;
}

// First footprint node along one axis, before the periodic wrap, from
// dist = pos * inv_cell_size (a float product, as in the reference).
//
// The reference takes floor() of (double)dist - 1.5 (TSC) / - 2.0 (PCS) and
// truncates (double)dist + 0.5 (NGP).  With k = floorf(dist) and the exact
// fraction f = dist - k (exact for any float), these equal k - 2 + (f >= 0.5),
// k - 2 and k + (f >= 0.5) (for dist >= -0.5) in integer arithmetic, for every
// |dist| < 2^31.  The footprint nodes are raw, raw+1, ..., raw+F-1.
template <int MAS>
__host__ __device__ __forceinline__ int first_node(float dist) {
    if constexpr (MAS == PYL_NGP) {
        // MAS_library.pyx:288-289: <int>(pos*inv + 0.5), the 0.5 is a double
        if (!(dist >= -0.5f)) return (int)((double)dist + 0.5);   // truncation toward 0
        const float k = floorf(dist);
        return (int)k + ((dist - k) >= 0.5f ? 1 : 0);
    } else if constexpr (MAS == PYL_CIC) {
        return (int)dist;                                   // MAS_library.pyx:151-153
    } else if constexpr (MAS == PYL_TSC) {
        const float k = floorf(dist);                       // floor(dist - 1.5) + 1
        return (int)k - ((dist - k) >= 0.5f ? 0 : 1);
    } else {
        return (int)floorf(dist) - 1;                       // floor(dist - 2.0) + 1
    }
}

// TSC weights of the nodes raw, raw+1, raw+2 as (1.5-d0)^2, 0.75-d1^2, (1.5-d2)^2:
// the outer two still take the factor 0.5 (MAS_library.pyx:392-399).
// (float)(raw + j) == (float)raw + j exactly for |raw| < 2^24.
__host__ __device__ __forceinline__ void tsc_parts(float dist, int raw, float &q0, float &c1, float &q2) {
    const float fr = (float)raw;
    const float d0 = fabsf(fr - dist);
    const float d1 = fabsf((fr + 1.0f) - dist);
    const float d2 = fabsf((fr + 2.0f) - dist);
    const float a0 = 1.5f - d0, a2 = 1.5f - d2;
    q0 = a0 * a0;
    c1 = 0.75f - d1 * d1;
    q2 = a2 * a2;
}

// Footprint weights of the nodes raw .. raw+F-1.
template <int MAS>
__host__ __device__ __forceinline__ void foot_weights(float dist, int raw, float *C) {
    if constexpr (MAS == PYL_NGP) {
        C[0] = 1.0f;
    } else if constexpr (MAS == PYL_CIC) {
        // MAS_library.pyx:151-156: u = dist - <int>dist, d = 1.0 - u
        const float u = dist - (float)raw;
        C[1] = u;
        C[0] = 1.0f - u;
    } else if constexpr (MAS == PYL_TSC) {
        // MAS_library.pyx:392-399.  With raw = floor(dist - 0.5) the distances
        // to the three nodes are d0, d2 in [0.5, 1.5] and d1 in [0, 0.5], also
        // after float rounding, so node 1 always takes the inner branch and
        // nodes 0, 2 the outer one (at a branch boundary both formulas give the
        // same float).  diff*diff is a float product, then 0.75 - it;
        // 0.5*(1.5-d)^2 rounds once like the reference's double expression
        // (1.5 - d is exact, its square fits in a double).
        float q0, q2;
        tsc_parts(dist, raw, q0, C[1], q2);
        C[0] = 0.5f * q0;
        C[2] = 0.5f * q2;
    } else {
        // PCS, MAS_library.pyx:485-492 (double polynomials, rounded to float).
        // With raw = floor(dist) - 1, d0 and d3 lie in [1, 2] and d1, d2 in
        // [0, 1] (after float rounding too); at d = 1 and d = 2 both branches
        // of the reference give the same double.
        float d[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = fabsf(((float)raw + (float)j) - dist);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double dd = (double)d[j];
            if (j == 1 || j == 2)
                C[j] = (float)((4.0 - 6.0 * dd * dd + 3.0 * dd * dd * dd) / 6.0);
            else
                C[j] = (float)((2.0 - dd) * (2.0 - dd) * (2.0 - dd) / 6.0);
        }
    }
}

// Per-axis footprint: writes the F weights, returns the first node wrapped into
// [0, dims); node j of the footprint is (first + j) mod dims.
template <int MAS>
__device__ __forceinline__ int axis_foot(float x, float inv, int dims, float *C) {
    const float dist = x * inv;
    const int raw = first_node<MAS>(dist);
    foot_weights<MAS>(dist, raw, C);
    return wrap_node(raw, dims);
}

// ---- tile geometry -----------------------------------------------------------
// Tiles split each axis into nt pieces; tile i covers the nodes with
// floor(node*nt/dims) == i, i.e. [origin(i), origin(i+1)) with
// origin(i) = ceil(i*dims/nt): all extents are dims/nt rounded up or down, so
// none is shorter than the footprint's reach when nt >= 2.  2D grids have one
// node along z.  A batch holds nreal grids; tile keys are r*ntiles + tile.
constexpr int kT3 = 16;        // nominal 3D tile extent
constexpr int kT2 = 64;        // nominal 2D tile extent

struct TileGeom {
    int dims;          // nodes per axis
    int nt;            // tiles along x and y (and z in 3D)
    int ntz;           // tiles along z: nt (3D) or 1 (2D)
    int shift;         // log2(tile extent) when dims % T == 0, else -1
    long long ntiles;  // tiles per grid
    long long ncell;   // nodes per grid
    int pslot;         // partial values stored per tile
    int wmode;         // sort-chunk key window (kWinDirect / kWinBox / kWinLinear)
    const int *orig;   // device table of tile_origin(i), i = 0..nt (same on every axis)
};

__host__ __device__ __forceinline__ int tile_origin(int i, int nt, int dims) {
    return (int)(((long long)i * dims + nt - 1) / nt);
}


// Number of values of a box (extents e + F - 1 per axis, z extent 1 in 2D)
// outside its interior [F-1, e) per axis, and the compact index of such a cell.
// The layout runs over three slabs: x outside the interior; x inside and y
// outside; x and y inside and z outside (3D only).
template <int F, int AXES>
__host__ __device__ __forceinline__ int halo_count(int ex, int ey, int ez) {
    const int bx = ex + F - 1, by = ey + F - 1;
    const int ix = ex - F + 1 > 0 ? ex - F + 1 : 0;
    const int iy = ey - F + 1 > 0 ? ey - F + 1 : 0;
    if (AXES == 2) return bx * by - ix * iy;
    const int bz = ez + F - 1;
    const int iz = ez - F + 1 > 0 ? ez - F + 1 : 0;
    return bx * by * bz - ix * iy * iz;
}

template <int F, int AXES>
__host__ __device__ __forceinline__ int halo_index(int x, int y, int z, int ex, int ey, int ez) {
    const int by = ey + F - 1;
    const int bz = (AXES == 3) ? ez + F - 1 : 1;
    const int ix = ex - F + 1 > 0 ? ex - F + 1 : 0;
    const int iy = ey - F + 1 > 0 ? ey - F + 1 : 0;
    const bool xin = x >= F - 1 && x < ex;
    const bool yin = y >= F - 1 && y < ey;
    if (!xin) {
        const int xi = x < F - 1 ? x : x - ix;
        return (xi * by + y) * bz + z;
    }
    const int sa = (ex + F - 1 - ix) * by * bz;
    if (!yin) {
        const int yi = y < F - 1 ? y : y - iy;
        return sa + ((x - (F - 1)) * (by - iy) + yi) * bz + z;
    }
    const int sb = ix * (by - iy) * bz;
    const int iz = ez - F + 1 > 0 ? ez - F + 1 : 0;
    const int zi = z < F - 1 ? z : z - iz;
    return sa + sb + ((x - (F - 1)) * iy + (y - (F - 1))) * (bz - iz) + zi;
}

// ---- counting sort by home tile -----------------------------------------------
// Each sort workgroup owns a chunk of consecutive particles of one realisation.
// It counts its particles per home tile in an LDS window of kWin consecutive
// keys (cyclic within the realisation's key range, so the periodic wrap in x
// stays close) centred on its first particle's key (coherent input -- a lattice, a
// snapshot in ID order, a sorted catalogue -- keeps almost every key in the
// window); keys outside the window take a global atomic on a separate
// overflow counter.  The window counts are flushed with one global atomic per
// non-zero key, which also hands the chunk its base inside the key's
// segment; the scatter pass reloads those bases and assigns slots with LDS
// atomics only.
constexpr int kSortBlock = 1024;
constexpr int kWin = 16384;                    // 64 KB of LDS
constexpr int kWinWide = 32768;                // 128 KB of LDS (a gfx950 workgroup may hold 160 KB)
constexpr bool kUseWide = true;

struct Chunk {
    long long c0, c1;   // particle range
    int r;              // realisation
    int wbase;          // window placement (count pass): w0 (linear) or wx * nt + wy (box, block)
    int smin, smax;     // first and last non-empty window slot (count pass; smin > smax: none)
    int wz, pad1;       // block window: first tile along z
};

// Key windows of a sort chunk (TileGeom::wmode, chosen on the host):
//  kWinDirect  ntiles <= kWin: slot = key - rbase, every key counted in LDS;
//              also the wide window (kWinWide slots, 3D grids of up to 32768
//              tiles), taken with chunks of 4-16 particles per tile so that
//              particles in any order are counted in LDS;
//  kWinBox     the 8 x 8 columns of tiles (all z) starting 4 tiles below the
//              chunk's middle particle along x and y, periodic: slot =
//              (dx * 8 + dy) * ntz + tz, 64 * ntz <= kWin slots.  A lattice or
//              ID-ordered chunk spans a few tiles in x and y whatever the grid
//              size, so its displaced particles stay inside;
//  kWinLinear  (key - rbase - w0) mod ntiles < kWin, centred on the middle
//              particle's key (grids too thin or too deep for the box);
//  kWinBlock   the kBkt x kBkt x kBktZ tiles of one bucket (second pass of the two-level
//              sort: every particle of the chunk has its home tile there).
constexpr int kWinDirect = 0, kWinBox = 1, kWinLinear = 2, kWinBlock = 3;
constexpr int kBoxW = 8;
constexpr int kBkt = 8;                         // tiles per bucket along x and y
constexpr int kBktZ = 4;                        // and along z
constexpr int kBktSlots = kBkt * kBkt * kBktZ;  // block window slots

// Home tile along one axis; POW2: every tile has extent 2^shift (or nt == 1).
template <int MAS, bool POW2>
__host__ __device__ __forceinline__ int home_tile(float dist, const TileGeom &g) {
    const int node = wrap_node(first_node<MAS>(dist), g.dims);
    if constexpr (POW2) return node >> g.shift;
    else return (int)(((long long)node * g.nt) / g.dims);
}

// Optional redshift-space shift applied to one coordinate as the position is
// read (redshift_space_library.pyx:29-46 fused into the MA load; pos itself is
// not modified).  When off, vel points at pos with vstride = axes (a valid
// address) and neither the velocity load nor the shift is made.
struct RSD {
    const float *vel;
    float factor;      // (1 + z) / H, rounded to float as in the reference
    float box;
    int axis;
    int vstride;
    int on;
};

template <int AXES>
__device__ __forceinline__ float3 load_dist(const float *__restrict__ pos, const RSD &rsd,
                                            long long i, float inv) {
    const float *p = pos + (long long)AXES * i;
    float x = p[0], y = p[1], z = (AXES == 3) ? p[2] : 0.0f;
    if (rsd.on) {                                // uniform across the launch
        const float v = rsd.vel[(long long)rsd.vstride * i + rsd.axis];
        if (rsd.axis == 0) x = rsd_shift(x, v, rsd.factor, rsd.box);
        else if (rsd.axis == 1) y = rsd_shift(y, v, rsd.factor, rsd.box);
        else z = rsd_shift(z, v, rsd.factor, rsd.box);
    }
    float3 d;
    d.x = x * inv;
    d.y = y * inv;
    d.z = z * inv;
    return d;
}

template <int MAS, int AXES, bool POW2>
__device__ __forceinline__ int3 dist_tiles(const float3 &d, const TileGeom &g) {
    int3 t;
    t.x = home_tile<MAS, POW2>(d.x, g);
    t.y = home_tile<MAS, POW2>(d.y, g);
    t.z = (AXES == 3) ? home_tile<MAS, POW2>(d.z, g) : 0;
    return t;
}

__host__ __device__ __forceinline__ unsigned tiles_key(const int3 &t, const TileGeom &g, long long rbase) {
    return (unsigned)(rbase + ((long long)t.x * g.nt + t.y) * g.ntz + t.z);
}

template <int MAS, int AXES, bool POW2>
__device__ __forceinline__ unsigned dist_key(const float3 &d, const TileGeom &g, long long rbase) {
    return tiles_key(dist_tiles<MAS, AXES, POW2>(d, g), g, rbase);
}

// Chunk window: mode, placement (w0, or first tile along x, y, z) and number of slots.
struct Win {
    int mode, a, b, c, slots;
};

__host__ __device__ __forceinline__ Win make_win(const TileGeom &g, const int3 &t0, long long k0, long long rbase) {
    Win w{g.wmode, 0, 0, 0, (int)g.ntiles};
    if (g.wmode == kWinBlock) {
        w.a = t0.x / kBkt * kBkt;
        w.b = t0.y / kBkt * kBkt;
        w.c = t0.z / kBktZ * kBktZ;
        w.slots = kBktSlots;
    } else if (g.wmode == kWinBox) {
        w.a = t0.x - kBoxW / 2;
        if (w.a < 0) w.a += g.nt;
        w.b = t0.y - kBoxW / 2;
        if (w.b < 0) w.b += g.nt;
        w.slots = kBoxW * kBoxW * g.ntz;
    } else if (g.wmode == kWinLinear) {
        // taken cyclically so that particles wrapped across the periodic x
        // boundary stay inside
        long long s = (k0 - rbase - kWin / 2) % g.ntiles;
        w.a = (int)(s < 0 ? s + g.ntiles : s);
        w.slots = kWin;
    }
    return w;
}

// Window slot of a particle's tiles (>= w.slots: outside, counted globally).
__host__ __device__ __forceinline__ unsigned win_slot(const int3 &t, unsigned key, const Win &w,
                                             const TileGeom &g, long long rbase) {
    if (w.mode == kWinBlock) {
        const unsigned dx = (unsigned)(t.x - w.a), dy = (unsigned)(t.y - w.b), dz = (unsigned)(t.z - w.c);
        if ((dx | dy) >= (unsigned)kBkt || dz >= (unsigned)kBktZ) return 0xffffffffu;
        return (dx * kBkt + dy) * kBktZ + dz;
    }
    if (w.mode == kWinBox) {
        int dx = t.x - w.a, dy = t.y - w.b;
        if (dx < 0) dx += g.nt;
        if (dy < 0) dy += g.nt;
        if (dx >= kBoxW || dy >= kBoxW) return 0xffffffffu;
        return (unsigned)((dx * kBoxW + dy) * g.ntz + t.z);
    }
    long long s = (long long)key - rbase - (w.mode == kWinLinear ? w.a : 0);
    if (s < 0) s += g.ntiles;
    return (unsigned)s;
}

__host__ __device__ __forceinline__ long long win_key(int s, const Win &w, const TileGeom &g,
                                             long long rbase) {
    if (w.mode == kWinBlock) {
        const int dz = s % kBktZ, dy = (s / kBktZ) % kBkt, dx = s / (kBkt * kBktZ);
        return rbase + ((long long)(w.a + dx) * g.nt + (w.b + dy)) * g.ntz + (w.c + dz);
    }
    if (w.mode == kWinBox) {
        const int tz = s % g.ntz, c = s / g.ntz;
        int tx = w.a + c / kBoxW, ty = w.b + c % kBoxW;
        if (tx >= g.nt) tx -= g.nt;
        if (ty >= g.nt) ty -= g.nt;
        return rbase + ((long long)tx * g.nt + ty) * g.ntz + tz;
    }
    long long k = (long long)(w.mode == kWinLinear ? w.a : 0) + s;
    if (k >= g.ntiles) k -= g.ntiles;
    return rbase + k;
}

// Particles per thread and loop trip: the loads of kUnroll particles are issued
// (unconditionally, at clamped indices) before any of them is used, so their
// latencies overlap.
constexpr int kUnroll = 4;       // count pass (measured best of 2, 4, 8 at 512^3)
constexpr int kUnrollS = 8;      // scatter pass (8: 0.80 -> 0.71 ms at 512^3)

template <int MAS, int AXES, bool POW2, int WS>
__global__ void __launch_bounds__(kSortBlock)
tile_count_kernel(const float *__restrict__ pos, RSD rsd, float inv, TileGeom g,
                  Chunk *__restrict__ chunks, unsigned *__restrict__ counts,
                  unsigned *__restrict__ ocounts, unsigned *__restrict__ dump,
                  unsigned *__restrict__ box_miss) {
    __shared__ unsigned hist[WS];
    const Chunk ch = chunks[blockIdx.x];
    const long long rbase = (long long)ch.r * g.ntiles;
    // every thread derives the window from the chunk's middle particle
    int3 t0 = make_int3(0, 0, 0);
    if (ch.c1 > ch.c0)
        t0 = dist_tiles<MAS, AXES, POW2>(load_dist<AXES>(pos, rsd, ch.c0 + (ch.c1 - ch.c0) / 2, inv), g);
    const Win w = make_win(g, t0, tiles_key(t0, g, rbase), rbase);
    if (threadIdx.x == 0) {
        chunks[blockIdx.x].wbase = (w.mode == kWinBox || w.mode == kWinBlock) ? w.a * g.nt + w.b : w.a;
        chunks[blockIdx.x].wz = w.c;
        if (ch.c1 <= ch.c0) {                  // empty chunk: no slot for the scatter
            chunks[blockIdx.x].smin = WS;
            chunks[blockIdx.x].smax = -1;
        }
    }
    if (ch.c1 <= ch.c0) return;                // (uniform over the workgroup)
    // particles the 8 x 8 tile-column box around t0 would miss (box_miss: the
    // host's choice between the box and the wide window for the next call)
    TileGeom gb = g;
    gb.wmode = kWinBox;
    const Win wb = make_win(gb, t0, 0, rbase);
    unsigned miss = 0;
    for (int s = threadIdx.x; s < w.slots; s += blockDim.x) hist[s] = 0u;
    __syncthreads();
    const long long step = (long long)blockDim.x * kUnroll;
    const long long last = ch.c1 - 1;
    for (long long base = ch.c0 + threadIdx.x; base < ch.c1; base += step) {
        float3 d[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            d[u] = load_dist<AXES>(pos, rsd, min(base + (long long)u * blockDim.x, last), inv);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (base + (long long)u * blockDim.x < ch.c1) {
                const int3 t = dist_tiles<MAS, AXES, POW2>(d[u], g);
                const unsigned key = tiles_key(t, g, rbase);
                const unsigned s = win_slot(t, key, w, g, rbase);
                if (s < (unsigned)w.slots) atomicAdd(&hist[s], 1u);
                else atomicAdd(&ocounts[key], 1u);
                // in the box window itself the miss is the overflow just taken
                if (box_miss) miss += w.mode == kWinBox ? s >= (unsigned)w.slots
                                                        : win_slot(t, key, wb, gb, rbase) >= (unsigned)wb.slots;
            }
        }
    }
    __syncthreads();
    unsigned *dmp = dump + (long long)blockIdx.x * WS;
    int smin = WS, smax = -1;
    for (int s = threadIdx.x; s < w.slots; s += blockDim.x) {
        const unsigned v = hist[s];
        if (v) {
            dmp[s] = atomicAdd(&counts[win_key(s, w, g, rbase)], v);
            smin = min(smin, s);
            smax = max(smax, s);
        }
    }
    // the scatter pass reloads only the slots in [smin, smax]
    for (int off = 32; off > 0; off >>= 1) {
        smin = min(smin, __shfl_xor(smin, off));
        smax = max(smax, __shfl_xor(smax, off));
        if (box_miss) miss += (unsigned)__shfl_xor((int)miss, off);
    }
    __syncthreads();                           // hist is reused for the wave results
    int *red = (int *)hist;
    const int nw = (int)(blockDim.x >> 6);
    if ((threadIdx.x & 63) == 0) {
        red[2 * (threadIdx.x >> 6)] = smin;
        red[2 * (threadIdx.x >> 6) + 1] = smax;
        red[2 * nw + (threadIdx.x >> 6)] = (int)miss;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = WS, b = -1;
        unsigned m = 0;
        for (int k = 0; k < nw; ++k) {
            a = min(a, red[2 * k]);
            b = max(b, red[2 * k + 1]);
            m += (unsigned)red[2 * nw + k];
        }
        chunks[blockIdx.x].smin = a;
        chunks[blockIdx.x].smax = b;
        // one atomic per workgroup on the single tally word
        if (box_miss && m) atomicAdd(box_miss, m);
    }
}

// One record (dist_x, dist_y, dist_z, w) per particle at offsets[key] + slot:
// window keys take slot = chunk base + LDS cursor, overflow keys slot =
// counts[key] + global cursor.  Weighted input also records max|W| per
// realisation (float bits of the largest finite |W|).
template <int MAS, int AXES, bool POW2, bool WT, int WS>
__global__ void __launch_bounds__(kSortBlock)
tile_scatter_kernel(const float *__restrict__ pos, RSD rsd, const float *__restrict__ W, float inv,
                    TileGeom g, const Chunk *__restrict__ chunks,
                    const unsigned *__restrict__ dump, const unsigned long long *__restrict__ offsets,
                    const unsigned *__restrict__ counts, unsigned *__restrict__ ocursors,
                    unsigned *__restrict__ wmax_bits, float *__restrict__ records,
                    unsigned long long rcap) {
    __shared__ unsigned cur[WS];
    const Chunk ch = chunks[blockIdx.x];
    if (ch.c1 <= ch.c0) return;                // empty chunk (uniform over the workgroup)
    const long long rbase = (long long)ch.r * g.ntiles;
    // the count pass's window of this chunk
    Win win{g.wmode, 0, 0, 0, (int)g.ntiles};
    if (g.wmode == kWinBox || g.wmode == kWinBlock) {
        win.a = ch.wbase / g.nt;
        win.b = ch.wbase - win.a * g.nt;
        win.c = ch.wz;
        win.slots = g.wmode == kWinBox ? kBoxW * kBoxW * g.ntz : kBktSlots;
    } else if (g.wmode == kWinLinear) {
        win.a = ch.wbase;
        win.slots = kWin;
    }
    const unsigned *dmp = dump + (long long)blockIdx.x * WS;
    // only slots in [smin, smax] were counted; no other slot is looked up
    for (int s = ch.smin + threadIdx.x; s <= ch.smax; s += blockDim.x) cur[s] = dmp[s];
    __syncthreads();
    unsigned m = 0;
    const long long step = (long long)blockDim.x * kUnrollS;
    const long long last = ch.c1 - 1;
    for (long long base = ch.c0 + threadIdx.x; base < ch.c1; base += step) {
        float3 d[kUnrollS];
        float w[kUnrollS];
        unsigned key[kUnrollS];
        unsigned long long slot[kUnrollS];
#pragma unroll
        for (int u = 0; u < kUnrollS; ++u) {
            const long long i = min(base + (long long)u * blockDim.x, last);
            d[u] = load_dist<AXES>(pos, rsd, i, inv);
            w[u] = WT ? W[i] : 1.0f;
        }
#pragma unroll
        for (int u = 0; u < kUnrollS; ++u) {
            const int3 t = dist_tiles<MAS, AXES, POW2>(d[u], g);
            key[u] = tiles_key(t, g, rbase);
            slot[u] = 0;
            if (base + (long long)u * blockDim.x < ch.c1) {
                const unsigned s = win_slot(t, key[u], win, g, rbase);
                if (s < (unsigned)win.slots) slot[u] = atomicAdd(&cur[s], 1u);
                else slot[u] = (unsigned long long)counts[key[u]] + atomicAdd(&ocursors[key[u]], 1u);
            }
        }
#pragma unroll
        for (int u = 0; u < kUnrollS; ++u) {
            if (base + (long long)u * blockDim.x < ch.c1) {
                const unsigned long long at = offsets[key[u]] + slot[u];
                if (at >= rcap) continue;               // only after a miscount (reported)
                if (WT) {
                    reinterpret_cast<float4 *>(records)[at] = make_float4(d[u].x, d[u].y, d[u].z, w[u]);
                } else {
                    float *o = records + 3 * at;
                    o[0] = d[u].x;
                    o[1] = d[u].y;
                    o[2] = d[u].z;
                }
                if (WT) {
                    const unsigned b = __float_as_uint(fabsf(w[u]));
                    if (b < 0x7f800000u) m = max(m, b);
                }
            }
        }
    }
    if (WT) {
        // block max through the (no longer needed) cursor table
        for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
        __syncthreads();
        if ((threadIdx.x & 63) == 0) cur[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned mm = 0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) mm = max(mm, cur[k]);
            atomicMax(&wmax_bits[ch.r], mm);
        }
    }
}

// ---- capacity sort: one read of the particles ----------------------------------
// For coherent input and one 3D grid.  Tile t owns `cap` record slots at t*cap in
// crec (a few times the mean occupancy).  Each sub-chunk of kCapSub particles per
// thread ranks its particles per tile in LDS (the chunk's window: every tile, or
// the 8 x 8 tile columns around its middle particle), claims the run of slots it
// needs in each tile it touches with one global atomic on the tile's cursor, and
// writes its records there.  Particles outside the window, and those past a full
// tile's capacity, go to the spill list (dist_x, dist_y, dist_z | W), which the
// regular count / scan / scatter then order by tile into a second record segment.
// No count pass: the positions are read once.  cursor[t] ends as the number of
// particles that claimed a slot of t (min(cursor, cap) were placed).
constexpr int kCapSub = 8;          // particles per thread per sub-chunk

template <int MAS, bool POW2, bool WT>
__global__ void __launch_bounds__(kSortBlock)
cap_scatter_kernel(const float *__restrict__ pos, RSD rsd, const float *__restrict__ W, float inv,
                   TileGeom g, const Chunk *__restrict__ chunks, unsigned *__restrict__ cursor,
                   unsigned cap, float *__restrict__ crec, float *__restrict__ sxyz,
                   float *__restrict__ sw, unsigned long long *__restrict__ nspill,
                   unsigned long long *__restrict__ placed, unsigned *__restrict__ wmax_bits,
                   unsigned *__restrict__ box_miss) {
    extern __shared__ unsigned cap_lds[];                   // hist[slots] | claimed base[slots]
    __shared__ unsigned red[3 * (kSortBlock / 64)];
    const Chunk ch = chunks[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63;
    int3 t0 = make_int3(0, 0, 0);
    if (ch.c1 > ch.c0)
        t0 = dist_tiles<MAS, 3, POW2>(load_dist<3>(pos, rsd, ch.c0 + (ch.c1 - ch.c0) / 2, inv), g);
    const Win w = make_win(g, t0, tiles_key(t0, g, 0), 0);
    unsigned *hist = cap_lds, *gbase = cap_lds + w.slots;
    // misses of the 8 x 8 tile-column box (the host's order finding), as in the count pass
    TileGeom gb = g;
    gb.wmode = kWinBox;
    const Win wb = make_win(gb, t0, 0, 0);
    unsigned m = 0, miss = 0, nplaced = 0;
    constexpr int SUB = kSortBlock * kCapSub;
    for (long long s0 = ch.c0; s0 < ch.c1; s0 += SUB) {
        const int cnt = (int)min((long long)SUB, ch.c1 - s0);
        for (int s = t; s < w.slots; s += kSortBlock) hist[s] = 0u;
        __syncthreads();
        float3 d[kCapSub];
        float wt[kCapSub];
        unsigned key[kCapSub], rk[kCapSub];
        int sl[kCapSub];                                    // window slot; -1 none, -2 outside
#pragma unroll
        for (int k = 0; k < kCapSub; ++k) {
            const long long i = s0 + min(t + k * kSortBlock, cnt - 1);
            d[k] = load_dist<3>(pos, rsd, i, inv);
            wt[k] = WT ? W[i] : 1.0f;
        }
#pragma unroll
        for (int k = 0; k < kCapSub; ++k) {
            sl[k] = -1;
            key[k] = 0u;
            rk[k] = 0u;
            if (t + k * kSortBlock < cnt) {
                const int3 tt = dist_tiles<MAS, 3, POW2>(d[k], g);
                key[k] = tiles_key(tt, g, 0);
                const unsigned s = win_slot(tt, key[k], w, g, 0);
                if (s < (unsigned)w.slots) {
                    sl[k] = (int)s;
                    rk[k] = atomicAdd(&hist[s], 1u);
                } else {
                    sl[k] = -2;
                }
                if (box_miss) miss += w.mode == kWinBox ? s >= (unsigned)w.slots
                                                        : win_slot(tt, key[k], wb, gb, 0) >= (unsigned)wb.slots;
                if (WT) {
                    const unsigned b = __float_as_uint(fabsf(wt[k]));
                    if (b < 0x7f800000u) m = max(m, b);
                }
            }
        }
        __syncthreads();
        // one claim per tile the sub-chunk touches
        for (int s = t; s < w.slots; s += kSortBlock) {
            const unsigned h = hist[s];
            if (h) gbase[s] = atomicAdd(&cursor[win_key(s, w, g, 0)], h);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kCapSub; ++k) {
            bool spill = sl[k] == -2;
            if (sl[k] >= 0) {
                const unsigned p = gbase[sl[k]] + rk[k];
                if (p < cap) {
                    const unsigned long long at = (unsigned long long)key[k] * cap + p;
                    if (WT) {
                        reinterpret_cast<float4 *>(crec)[at] = make_float4(d[k].x, d[k].y, d[k].z, wt[k]);
                    } else {
                        float *o = crec + 3 * at;
                        o[0] = d[k].x;
                        o[1] = d[k].y;
                        o[2] = d[k].z;
                    }
                    ++nplaced;
                } else {
                    spill = true;
                }
            }
            // spill-list slots: one atomic per wave
            const unsigned long long bal = __ballot(spill);
            if (bal) {
                const int first = __ffsll((long long)bal) - 1;
                unsigned long long b0 = 0;
                if (lane == first) b0 = atomicAdd(nspill, (unsigned long long)__popcll(bal));
                b0 = (unsigned long long)__shfl((long long)b0, first);
                if (spill) {
                    const unsigned long long at = b0 + (unsigned long long)__popcll(bal & ((1ull << lane) - 1ull));
                    sxyz[3 * at] = d[k].x;
                    sxyz[3 * at + 1] = d[k].y;
                    sxyz[3 * at + 2] = d[k].z;
                    if (WT) sw[at] = wt[k];
                }
            }
        }
        // (the next sub-chunk clears hist only after every thread has passed the
        // claims, and rewrites gbase only after every thread has written its records)
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = max(m, (unsigned)__shfl_xor((int)m, off));
        miss += (unsigned)__shfl_xor((int)miss, off);
        nplaced += (unsigned)__shfl_xor((int)nplaced, off);
    }
    const int nw = kSortBlock / 64;
    if (lane == 0) {
        red[t >> 6] = m;
        red[nw + (t >> 6)] = miss;
        red[2 * nw + (t >> 6)] = nplaced;
    }
    __syncthreads();
    if (t == 0) {
        unsigned mm = 0, ms = 0;
        unsigned long long pl = 0;
        for (int k = 0; k < nw; ++k) {
            mm = max(mm, red[k]);
            ms += red[nw + k];
            pl += red[2 * nw + k];
        }
        if (WT) atomicMax(&wmax_bits[0], mm);
        if (box_miss && ms) atomicAdd(box_miss, ms);
        if (pl) atomicAdd(placed, pl);
    }
}

// ---- small spill lists ----------------------------------------------------------
// A capacity sort that spilled at most kSpillAtomic particles orders them by tile
// without the windowed count and scatter, whose per-chunk LDS histograms of every
// tile (or tile box) cost more than the particles of a short list: each wave counts
// its particles per tile with one global atomic per distinct tile among its lanes,
// then writes every record at its tile's offset plus a rank from a second such
// atomic.  The order of the records within a tile is whatever the atomics give; the
// deposit's integer sums do not depend on it.  Longer spill lists take the windowed
// passes (spill_count_kernel writes their chunks; these kernels then return at
// once).  512^3 TSC (0.1% spilled): 0.068 -> 0.049 ms; 1024^3 PCS: 0.20 -> 0.08 ms.
constexpr unsigned long long kSpillAtomic = 1ull << 22;
constexpr int kSpillBlocks = 2048;        // workgroups of the atomic spill kernels (grid-stride)

// Adds, for each distinct key among the active lanes of the wave, the number of
// lanes holding it to ctr[key] (one atomic, by the lowest such lane), and returns
// to every active lane the old value plus its rank among those lanes.
__device__ __forceinline__ unsigned wave_group_add(unsigned *__restrict__ ctr, unsigned key, bool active) {
    const int lane = (int)(threadIdx.x & 63u);
    unsigned long long todo = __ballot(active);
    unsigned mine = 0;
    while (todo) {                                   // wave-uniform
        const int leader = __ffsll((long long)todo) - 1;
        const unsigned lk = (unsigned)__shfl((int)key, leader);
        const bool in = active && key == lk;
        const unsigned long long grp = __ballot(in);
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&ctr[lk], (unsigned)__popcll(grp));
        base = (unsigned)__shfl((int)base, leader);
        if (in) mine = base + (unsigned)__popcll(grp & ((1ull << lane) - 1ull));
        todo &= ~grp;
    }
    return mine;
}

// Spill records per tile into ocounts (the scan adds counts + ocounts; counts stay
// zero here).  sxyz holds (dist_x, dist_y, dist_z) per spilled particle.  The first
// threads also write the windowed passes' cpr chunks of the list (empty when this
// kernel takes it) and the g.nt + 1 tile origins.
template <int MAS, bool POW2>
__global__ void __launch_bounds__(kBlock)
spill_count_kernel(const float *__restrict__ sxyz, const unsigned long long *__restrict__ nspill,
                   TileGeom g, unsigned *__restrict__ ocounts, Chunk *__restrict__ ch, int cpr,
                   int *__restrict__ orig) {
    const unsigned long long ns = *nspill;
    const int gi = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gi < cpr) {
        const long long len = ns > kSpillAtomic ? (long long)ns : 0;
        ch[gi] = Chunk{len * gi / cpr, len * (gi + 1) / cpr, 0, 0, 0, 0, 0, 0};
    }
    if (gi <= g.nt) orig[gi] = tile_origin(gi, g.nt, g.dims);
    if (ns > kSpillAtomic) return;
    const RSD off{sxyz, 0.0f, 0.0f, 0, 3, 0};
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long b = (unsigned long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); b < ns;
         b += step) {
        const unsigned long long i = b + (threadIdx.x & 63u);
        const bool active = i < ns;
        unsigned key = 0;
        if (active) key = dist_key<MAS, 3, POW2>(load_dist<3>(sxyz, off, (long long)i, 1.0f), g, 0);
        (void)wave_group_add(ocounts, key, active);
    }
}

// One record per spilled particle at offsets[key] + its rank (ocursors zeroed).  Also
// the capacity sort's record total: total = sorted spill (offsets[ntiles]) + placed.
template <int MAS, bool POW2, bool WT>
__global__ void __launch_bounds__(kBlock)
spill_scatter_kernel(const float *__restrict__ sxyz, const float *__restrict__ sw,
                     const unsigned long long *__restrict__ nspill, TileGeom g,
                     const unsigned long long *__restrict__ offsets, unsigned *__restrict__ ocursors,
                     float *__restrict__ records, unsigned long long rcap,
                     const unsigned long long *__restrict__ placed, unsigned long long *__restrict__ total) {
    if (blockIdx.x == 0 && threadIdx.x == 0) total[0] = offsets[g.ntiles] + placed[0];
    const unsigned long long ns = *nspill;
    if (ns > kSpillAtomic) return;
    const RSD off{sxyz, 0.0f, 0.0f, 0, 3, 0};
    const unsigned long long step = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long b = (unsigned long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); b < ns;
         b += step) {
        const unsigned long long i = b + (threadIdx.x & 63u);
        const bool active = i < ns;
        float3 d = make_float3(0.0f, 0.0f, 0.0f);
        float w = 1.0f;
        unsigned key = 0;
        if (active) {
            d = load_dist<3>(sxyz, off, (long long)i, 1.0f);
            key = dist_key<MAS, 3, POW2>(d, g, 0);
            if (WT) w = sw[i];
        }
        const unsigned r = wave_group_add(ocursors, key, active);
        if (!active) continue;
        const unsigned long long at = offsets[key] + r;
        if (at >= rcap) continue;                        // only after a miscount (reported)
        if (WT) {
            reinterpret_cast<float4 *>(records)[at] = make_float4(d.x, d.y, d.z, w);
        } else {
            float *o = records + 3 * at;
            o[0] = d.x;
            o[1] = d.y;
            o[2] = d.z;
        }
    }
}

// Exclusive scan of per-tile counts into 64-bit offsets (offsets[ntiles] = total).
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ unsigned long long block_exclusive_scan(unsigned long long v,
                                                                   unsigned long long *sh,
                                                                   unsigned long long &total) {
    // Hillis-Steele over the block in LDS (1024 entries).
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        unsigned long long a = (t >= off) ? sh[t - off] : 0ull;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    total = sh[kScanBlock - 1];
    unsigned long long incl = sh[t];
    __syncthreads();
    return incl - v;
}

// The scanned value of element i is in[i] + in2[i].
__global__ void __launch_bounds__(kScanBlock)
scan_partial_kernel(const unsigned *__restrict__ in, const unsigned *__restrict__ in2,
                    long long n, unsigned long long *__restrict__ block_sums) {
    __shared__ unsigned long long sh[kScanBlock];
    long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    unsigned long long s = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += (unsigned long long)in[base + k] + in2[base + k];
    unsigned long long total;
    block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanBlock)
scan_blocksums_kernel(unsigned long long *__restrict__ block_sums, int nblocks) {
    // Single block: nblocks <= kScanTile.
    __shared__ unsigned long long sh[kScanBlock];
    int t = threadIdx.x;
    unsigned long long v[kScanItems];
    unsigned long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        int idx = t * kScanItems + k;
        v[k] = idx < nblocks ? block_sums[idx] : 0ull;
        s += v[k];
    }
    unsigned long long total;
    unsigned long long ex = block_exclusive_scan(s, sh, total);
    for (int k = 0; k < kScanItems; ++k) {
        int idx = t * kScanItems + k;
        if (idx < nblocks) block_sums[idx] = ex;
        ex += v[k];
    }
}

__global__ void __launch_bounds__(kScanBlock)
scan_final_kernel(const unsigned *__restrict__ in, const unsigned *__restrict__ in2, long long n,
                  const unsigned long long *__restrict__ block_prefix,
                  unsigned long long *__restrict__ out) {
    __shared__ unsigned long long sh[kScanBlock];
    long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    unsigned long long v[kScanItems];
    unsigned long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? (unsigned long long)in[base + k] + in2[base + k] : 0ull;
        s += v[k];
    }
    unsigned long long total;
    unsigned long long ex = block_exclusive_scan(s, sh, total) + block_prefix[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanBlock - 1) out[n] = ex;
}

// ---- two-level sort (3D, particles in any order) -------------------------------
// A chunk of particles in random order spreads over every tile of a large grid, so
// the one-pass scatter writes each record alone into its own cache line.  The
// two-level sort first moves every particle, as its (dist_x, dist_y, dist_z) record
// and weight, into its bucket of kBkt x kBkt x kBktZ tiles, then sorts each bucket by tile
// with the block window (512 LDS slots per chunk).  Both passes count, then scatter
// through LDS in slot order (staged_scatter_kernel).  The second pass reads the
// bucketed records with inv = 1 (dist * 1 = dist exactly) and no redshift shift,
// so every record and every tile sum is the same as in one pass.
constexpr int kMaxBuckets = kScanTile;           // the bucket plan is one scan workgroup

// Buckets along x and y (nbt) and z (nbtz) of a grid of nt x nt x ntz tiles.
__host__ __device__ __forceinline__ int bucket_counts_xy(int nt) { return (nt + kBkt - 1) / kBkt; }
__host__ __device__ __forceinline__ int bucket_counts_z(int ntz) { return (ntz + kBktZ - 1) / kBktZ; }

__host__ __device__ __forceinline__ int tile_bucket(const int3 &t, int nbt, int nbtz) {
    return ((t.x / kBkt) * nbt + t.y / kBkt) * nbtz + t.z / kBktZ;
}

// Pass 1, count: particles per bucket in LDS, written as the chunk's column of the
// bucket-major count matrix bmat[b * nchunk + chunk] (plain stores: every chunk
// has particles in every bucket when the input is out of order, so returning
// atomics on the bucket totals would queue nchunk deep on each address).  With few
// buckets the lanes of a wave often hit the same LDS counter, and LDS atomics to
// one address in one instruction are serialised: each bucket has `copies`
// counters, lane l adding to copy l mod copies, summed at the flush.  Tallies the
// particles the 8 x 8 tile-column box around the chunk's middle particle misses,
// as the one-pass count does (one global atomic per workgroup).
constexpr int kBucketHist = 16384;               // counters of the count pass (64 KB)

template <int MAS, bool POW2>
__global__ void __launch_bounds__(kSortBlock)
bucket_count_kernel(const float *__restrict__ pos, RSD rsd, float inv, TileGeom g, int nbt,
                    const Chunk *__restrict__ chunks, unsigned *__restrict__ bmat,
                    unsigned *__restrict__ box_miss) {
    __shared__ unsigned hist[kBucketHist];
    __shared__ unsigned wmiss[kSortBlock / 64];
    const int nbtz = bucket_counts_z(g.ntz);
    const int nb = nbt * nbt * nbtz;
    int copies = 1;                                  // a power of two, nb * copies <= kBucketHist
    while (copies < 16 && nb * copies * 2 <= kBucketHist) copies *= 2;
    const int cmask = copies - 1;
    const Chunk ch = chunks[blockIdx.x];
    int3 t0 = make_int3(0, 0, 0);
    if (ch.c1 > ch.c0)
        t0 = dist_tiles<MAS, 3, POW2>(load_dist<3>(pos, rsd, ch.c0 + (ch.c1 - ch.c0) / 2, inv), g);
    TileGeom gb = g;
    gb.wmode = kWinBox;
    const Win wb = make_win(gb, t0, 0, 0);
    unsigned miss = 0;
    for (int s = threadIdx.x; s < nb * copies; s += blockDim.x) hist[s] = 0u;
    __syncthreads();
    const int mycopy = (int)threadIdx.x & cmask;
    const long long step = (long long)blockDim.x * kUnroll;
    const long long last = ch.c1 - 1;
    for (long long base = ch.c0 + threadIdx.x; base < ch.c1; base += step) {
        float3 d[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            d[u] = load_dist<3>(pos, rsd, min(base + (long long)u * blockDim.x, last), inv);
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (base + (long long)u * blockDim.x < ch.c1) {
                const int3 t = dist_tiles<MAS, 3, POW2>(d[u], g);
                atomicAdd(&hist[tile_bucket(t, nbt, nbtz) * copies + mycopy], 1u);
                if (box_miss) miss += win_slot(t, 0u, wb, gb, 0) >= (unsigned)wb.slots;
            }
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nb; s += blockDim.x) {
        unsigned v = 0;
        for (int k = 0; k < copies; ++k) v += hist[s * copies + k];
        bmat[(long long)s * gridDim.x + blockIdx.x] = v;
    }
    if (!box_miss) return;
    for (int off = 32; off > 0; off >>= 1) miss += (unsigned)__shfl_xor((int)miss, off);
    if ((threadIdx.x & 63) == 0) wmiss[threadIdx.x >> 6] = miss;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m += wmiss[k];
        if (m) atomicAdd(box_miss, m);
    }
}

// Column scan of the count matrix, one workgroup per bucket: the chunk's base
// inside the bucket (bdump[chunk * nb + b], the layout the scatter reads) and the
// bucket's total (bcounts[b]).
__global__ void __launch_bounds__(kScanBlock)
bucket_scan_kernel(const unsigned *__restrict__ bmat, int nchunk, int nb,
                   unsigned *__restrict__ bdump, unsigned *__restrict__ bcounts) {
    __shared__ unsigned long long sh[kScanBlock];
    const int b = blockIdx.x;
    const unsigned *col = bmat + (long long)b * nchunk;
    unsigned long long carry = 0;
    for (int c0 = 0; c0 < nchunk; c0 += kScanTile) {
        const int i0 = c0 + threadIdx.x * kScanItems;
        unsigned v[kScanItems];
        unsigned long long s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = i0 + k < nchunk ? col[i0 + k] : 0u;
            s += v[k];
        }
        unsigned long long tot;
        unsigned long long ex = block_exclusive_scan(s, sh, tot) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < nchunk) bdump[(long long)(i0 + k) * nb + b] = (unsigned)ex;
            ex += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) bcounts[b] = (unsigned)carry;
}

// Bucket offsets (exclusive scan of the counts; boffsets[nb] = total) and the chunk
// table of pass 2: bucket b is cut into ceil(len_b / target) chunks of at most
// target records, so no chunk straddles two buckets; the table has room for
// n / target + nb chunks and the entries past the last one are empty.
__global__ void __launch_bounds__(kScanBlock)
bucket_plan_kernel(const unsigned *__restrict__ bcounts, int nb, long long target,
                   unsigned long long *__restrict__ boffsets, Chunk *__restrict__ ch, int nchunk) {
    __shared__ unsigned long long sh[kScanBlock];
    const int b0 = threadIdx.x * kScanItems;
    unsigned long long len[kScanItems], s = 0, nc = 0;
    for (int k = 0; k < kScanItems; ++k) {
        len[k] = b0 + k < nb ? bcounts[b0 + k] : 0ull;
        s += len[k];
        nc += (len[k] + target - 1) / target;
    }
    unsigned long long tot, ctot;
    unsigned long long off = block_exclusive_scan(s, sh, tot);
    unsigned long long ci = block_exclusive_scan(nc, sh, ctot);
    for (int k = 0; k < kScanItems; ++k) {
        if (b0 + k >= nb) break;
        boffsets[b0 + k] = off;
        const unsigned long long e = off + len[k];
        for (unsigned long long c0 = off; c0 < e; c0 += target, ++ci)
            if (ci < (unsigned long long)nchunk)
                ch[ci] = Chunk{(long long)c0, (long long)(c0 + target < e ? c0 + target : e), 0, 0, 0, 0, 0, 0};
        off = e;
    }
    if (threadIdx.x == 0) boffsets[nb] = tot;
    for (long long i = (long long)ctot + threadIdx.x; i < nchunk; i += blockDim.x)
        ch[i] = Chunk{0, 0, 0, 0, 0, 0, 0, 0};
}

// LDS-staged scatter of the two-level sort.  Pass 1 (MODE 0) moves particles into
// their buckets, pass 2 (MODE 1) the bucketed records into their tiles (block
// window slots).  A chunk is taken STAGE particles at a time: they are counted
// per slot in LDS, placed in LDS in slot order, and written out by consecutive
// lanes to consecutive records of each slot's run, so a run of about STAGE /
// slots records leaves in whole cache lines instead of one record per lane to
// scattered lines.  The chunk's run bases come from its count pass (dump) as in
// the one-pass scatter; the runs of a chunk's sub-chunks follow one another.
constexpr int kStage1 = 4096;      // particles per sub-chunk, pass 1 (buckets)
constexpr int kStage2 = 8192;      // and pass 2 (tiles)

// Exclusive scan of cnt[0..ns) into ofs[0..ns) over the workgroup (kSortBlock
// threads, ns <= NS); returns the total.  wsum: kSortBlock / 64 + 1 words of LDS.
template <int NS>
__device__ __forceinline__ unsigned stage_scan(const unsigned *cnt, unsigned *ofs, int ns, unsigned *wsum) {
    constexpr int E = (NS + kSortBlock - 1) / kSortBlock;
    constexpr int NW = kSortBlock / 64;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    unsigned v[E], s = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = t * E + e;
        v[e] = i < ns ? cnt[i] : 0u;
        s += v[e];
    }
    unsigned x = s;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = (unsigned)__shfl_up((int)x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (t == 0) {
        unsigned a = 0;
        for (int k = 0; k < NW; ++k) {
            const unsigned b = wsum[k];
            wsum[k] = a;
            a += b;
        }
        wsum[NW] = a;
    }
    __syncthreads();
    unsigned ex = x - s + wsum[wave];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = t * E + e;
        if (i < ns) ofs[i] = ex;
        ex += v[e];
    }
    const unsigned total = wsum[NW];
    __syncthreads();
    return total;
}

template <int MAS, bool POW2, bool WT, int MODE, int NS, int STAGE>
__global__ void __launch_bounds__(kSortBlock)
staged_scatter_kernel(const float *__restrict__ pos, RSD rsd, const float *__restrict__ W, float inv,
                      TileGeom g, int nbt, const Chunk *__restrict__ chunks,
                      const unsigned *__restrict__ dump, const unsigned long long *__restrict__ base,
                      const unsigned *__restrict__ counts, unsigned *__restrict__ ocursors,
                      float *__restrict__ out, float *__restrict__ out_w,
                      unsigned *__restrict__ wmax_bits, unsigned long long rcap) {
    constexpr int kStagePer = STAGE / kSortBlock;
    __shared__ float sx[STAGE], sy[STAGE], sz[STAGE];
    __shared__ float sw[WT ? STAGE : 1];
    __shared__ unsigned short sslot[STAGE];
    __shared__ unsigned hist[NS], lofs[NS];
    __shared__ unsigned long long gcur[NS];
    __shared__ unsigned wsum[kSortBlock / 64 + 1];
    const Chunk ch = chunks[blockIdx.x];
    const long long rbase = (long long)ch.r * g.ntiles;
    const int t = threadIdx.x;
    const int nbtz = bucket_counts_z(g.ntz);
    int ns = nbt * nbt * nbtz;
    Win win{kWinBlock, 0, 0, 0, kBktSlots};
    if (MODE == 1) {
        win.a = ch.wbase / g.nt;
        win.b = ch.wbase - win.a * g.nt;
        win.c = ch.wz;
        ns = kBktSlots;
    }
    // run bases: pass 1, bucket offset + the chunk's base in the bucket; pass 2, tile
    // offset + the chunk's base in the tile (only the slots the count found filled
    // have a base; in pass 2 they lie in [smin, smax], and a slot of the block past
    // the grid's last tile is never filled)
    const unsigned *dmp = dump + (long long)blockIdx.x * (MODE == 0 ? ns : NS);
    if (MODE == 0) {
        for (int s = t; s < ns; s += blockDim.x) gcur[s] = base[s] + dmp[s];
    } else {
        for (int s = ch.smin + t; s <= ch.smax; s += blockDim.x) {
            const int dz = s % kBktZ, dy = (s / kBktZ) % kBkt, dx = s / (kBkt * kBktZ);
            if (win.a + dx < g.nt && win.b + dy < g.nt && win.c + dz < g.ntz)
                gcur[s] = base[win_key(s, win, g, rbase)] + dmp[s];
        }
    }
    unsigned m = 0;
    for (long long s0 = ch.c0; s0 < ch.c1; s0 += STAGE) {
        const int cnt = (int)min((long long)STAGE, ch.c1 - s0);
        for (int s = t; s < ns; s += blockDim.x) hist[s] = 0u;
        __syncthreads();
        float3 d[kStagePer];
        float w[kStagePer];
        int sl[kStagePer];
        unsigned rk[kStagePer];
#pragma unroll
        for (int k = 0; k < kStagePer; ++k) {
            const long long i = s0 + min(t + k * kSortBlock, cnt - 1);
            d[k] = load_dist<3>(pos, rsd, i, inv);
            w[k] = WT ? W[i] : 1.0f;
        }
#pragma unroll
        for (int k = 0; k < kStagePer; ++k) {
            sl[k] = -1;
            if (t + k * kSortBlock < cnt) {
                const int3 tt = dist_tiles<MAS, 3, POW2>(d[k], g);
                if (MODE == 0) {
                    sl[k] = tile_bucket(tt, nbt, nbtz);
                } else {
                    const unsigned key = tiles_key(tt, g, rbase);
                    const unsigned s = win_slot(tt, key, win, g, rbase);
                    if (s < (unsigned)kBktSlots) {
                        sl[k] = (int)s;
                    } else {
                        // outside the bucket's block (not produced by pass 1): the
                        // overflow slots of the count pass
                        const unsigned long long at =
                            base[key] + counts[key] + atomicAdd(&ocursors[key], 1u);
                        if (at < rcap) {
                            if (WT) {
                                reinterpret_cast<float4 *>(out)[at] = make_float4(d[k].x, d[k].y, d[k].z, w[k]);
                            } else {
                                out[3 * at] = d[k].x;
                                out[3 * at + 1] = d[k].y;
                                out[3 * at + 2] = d[k].z;
                            }
                        }
                    }
                }
                if (sl[k] >= 0) rk[k] = atomicAdd(&hist[sl[k]], 1u);
                if (MODE == 1 && WT) {
                    const unsigned b = __float_as_uint(fabsf(w[k]));
                    if (b < 0x7f800000u) m = max(m, b);
                }
            }
        }
        __syncthreads();
        const unsigned nst = stage_scan<NS>(hist, lofs, ns, wsum);
#pragma unroll
        for (int k = 0; k < kStagePer; ++k) {
            if (sl[k] >= 0) {
                const unsigned p = lofs[sl[k]] + rk[k];
                sx[p] = d[k].x;
                sy[p] = d[k].y;
                sz[p] = d[k].z;
                if (WT) sw[p] = w[k];
                sslot[p] = (unsigned short)sl[k];
            }
        }
        __syncthreads();
        for (int j = t; j < (int)nst; j += blockDim.x) {
            const int s = sslot[j];
            const unsigned long long at = gcur[s] + (unsigned long long)(j - (int)lofs[s]);
            if (at >= rcap) continue;                 // only after a miscount (reported)
            if (MODE == 1 && WT) {
                reinterpret_cast<float4 *>(out)[at] = make_float4(sx[j], sy[j], sz[j], sw[j]);
            } else {
                out[3 * at] = sx[j];
                out[3 * at + 1] = sy[j];
                out[3 * at + 2] = sz[j];
                if (MODE == 0 && WT) out_w[at] = sw[j];
            }
        }
        __syncthreads();
        for (int s = t; s < ns; s += blockDim.x) gcur[s] += hist[s];
        __syncthreads();
    }
    if (MODE == 1 && WT) {
        for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
        if ((t & 63) == 0) wsum[t >> 6] = m;
        __syncthreads();
        if (t == 0) {
            unsigned mm = 0;
            for (int k = 0; k < (int)(blockDim.x >> 6); ++k) mm = max(mm, wsum[k]);
            atomicMax(&wmax_bits[ch.r], mm);
        }
    }
}

// Capacity form of the two-level sort's second pass.  Every chunk lies in one
// bucket (the block window), but its records are in random tile order, so they are
// staged like staged_scatter_kernel's: a sub-chunk of STAGE records is ranked per
// tile in LDS, placed in LDS in tile order, claims the run of tile slots it needs in
// each tile with one global atomic on the tile's cursor, and is written by
// consecutive lanes to consecutive slots.  Records past a full tile's capacity go to
// the spill list (one atomic per wave); there are no misses (pass 1 put every record
// in its chunk's bucket).  Records are read as (dist_x, dist_y, dist_z) | w.
template <int MAS, bool POW2, bool WT, int STAGE>
__global__ void __launch_bounds__(kSortBlock)
staged_cap_scatter_kernel(const float *__restrict__ bxyz, const float *__restrict__ bw, TileGeom g,
                          const Chunk *__restrict__ chunks, unsigned *__restrict__ cursor, unsigned cap,
                          float *__restrict__ crec, float *__restrict__ sxyz, float *__restrict__ sw,
                          unsigned long long *__restrict__ nspill, unsigned long long *__restrict__ placed,
                          unsigned *__restrict__ wmax_bits) {
    constexpr int NS = kBktSlots;
    constexpr int kPer = STAGE / kSortBlock;
    __shared__ float sx[STAGE], sy[STAGE], sz[STAGE];
    __shared__ float ssw[WT ? STAGE : 1];
    __shared__ unsigned short sslot[STAGE];
    __shared__ unsigned hist[NS], lofs[NS], grem[NS];
    __shared__ unsigned long long gbase[NS];
    __shared__ unsigned wsum[kSortBlock / 64 + 1];
    const Chunk ch = chunks[blockIdx.x];
    const int t = threadIdx.x, lane = t & 63;
    const RSD off{bxyz, 0.0f, 0.0f, 0, 3, 0};
    int3 t0 = make_int3(0, 0, 0);
    if (ch.c1 > ch.c0)
        t0 = dist_tiles<MAS, 3, POW2>(load_dist<3>(bxyz, off, ch.c0 + (ch.c1 - ch.c0) / 2, 1.0f), g);
    const Win win = make_win(g, t0, 0, 0);                      // g.wmode == kWinBlock
    unsigned m = 0, np = 0;
    for (long long s0 = ch.c0; s0 < ch.c1; s0 += STAGE) {
        const int cnt = (int)min((long long)STAGE, ch.c1 - s0);
        for (int s = t; s < NS; s += kSortBlock) hist[s] = 0u;
        __syncthreads();
        float3 d[kPer];
        float w[kPer];
        int sl[kPer];
        unsigned rk[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const long long i = s0 + min(t + k * kSortBlock, cnt - 1);
            d[k] = load_dist<3>(bxyz, off, i, 1.0f);
            w[k] = WT ? bw[i] : 1.0f;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            sl[k] = -1;
            rk[k] = 0u;
            if (t + k * kSortBlock < cnt) {
                const int3 tt = dist_tiles<MAS, 3, POW2>(d[k], g);
                const unsigned s = win_slot(tt, tiles_key(tt, g, 0), win, g, 0);
                if (s < (unsigned)NS) {
                    sl[k] = (int)s;
                    rk[k] = atomicAdd(&hist[s], 1u);
                }
                if (WT) {
                    const unsigned b = __float_as_uint(fabsf(w[k]));
                    if (b < 0x7f800000u) m = max(m, b);
                }
            }
        }
        __syncthreads();
        const unsigned nst = stage_scan<NS>(hist, lofs, NS, wsum);
        for (int s = t; s < NS; s += kSortBlock) {
            const unsigned h = hist[s];
            if (h) {
                const unsigned long long key = (unsigned long long)win_key(s, win, g, 0);
                const unsigned claim = atomicAdd(&cursor[key], h);
                gbase[s] = key * cap + claim;
                grem[s] = claim < cap ? cap - claim : 0u;
            }
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (sl[k] >= 0) {
                const unsigned q = lofs[sl[k]] + rk[k];
                sx[q] = d[k].x;
                sy[q] = d[k].y;
                sz[q] = d[k].z;
                if (WT) ssw[q] = w[k];
                sslot[q] = (unsigned short)sl[k];
            }
        }
        __syncthreads();
        for (int j = t; j < (int)nst; j += kSortBlock) {
            const int s = sslot[j];
            const unsigned r = (unsigned)j - lofs[s];
            const bool spill = r >= grem[s];
            if (!spill) {
                const unsigned long long at = gbase[s] + r;
                if (WT) {
                    reinterpret_cast<float4 *>(crec)[at] = make_float4(sx[j], sy[j], sz[j], ssw[j]);
                } else {
                    crec[3 * at] = sx[j];
                    crec[3 * at + 1] = sy[j];
                    crec[3 * at + 2] = sz[j];
                }
                ++np;
            }
            const unsigned long long bal = __ballot(spill);
            if (bal) {
                const int first = __ffsll((long long)bal) - 1;
                unsigned long long b0 = 0;
                if (lane == first) b0 = atomicAdd(nspill, (unsigned long long)__popcll(bal));
                b0 = (unsigned long long)__shfl((long long)b0, first);
                if (spill) {
                    const unsigned long long at = b0 + (unsigned long long)__popcll(bal & ((1ull << lane) - 1ull));
                    sxyz[3 * at] = sx[j];
                    sxyz[3 * at + 1] = sy[j];
                    sxyz[3 * at + 2] = sz[j];
                    if (WT) sw[at] = ssw[j];
                }
            }
        }
        __syncthreads();
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = max(m, (unsigned)__shfl_xor((int)m, o));
        np += (unsigned)__shfl_xor((int)np, o);
    }
    if (lane == 0) {
        wsum[t >> 6] = m;
        lofs[t >> 6] = np;
    }
    __syncthreads();
    if (t == 0) {
        unsigned mm = 0;
        unsigned long long pl = 0;
        for (int k = 0; k < kSortBlock / 64; ++k) {
            mm = max(mm, wsum[k]);
            pl += lofs[k];
        }
        if (WT) atomicMax(&wmax_bits[0], mm);
        if (pl) atomicAdd(placed, pl);
    }
}

// ---- deposit -----------------------------------------------------------------
// Records of the capacity sort (cap_scatter_kernel): tile `key` has min(count[key],
// cap) records at rec + key*cap; rec == nullptr: none (the sorted records alone).
// nrec >= 0: the realisation's record total (otherwise taken from the offsets).
struct CapRecs {
    const float *rec;
    const unsigned *count;
    unsigned cap;
    long long nrec;
};

__device__ __forceinline__ unsigned long long cap_count(const CapRecs &cr, long long key) {
    return cr.rec ? (unsigned long long)min(cr.count[key], cr.cap) : 0ull;
}

// Output transform of a final node value v: overwrite ? v/scale - shift : grid + v.
struct OutMode {
    int overwrite;
    float scale, shift;
    float prescale;     // > 0: the grid is first divided by it (2D renormalisation), host side only
};

__device__ __forceinline__ void store_node(float *__restrict__ grid, long long node, float v,
                                           const OutMode &om) {
    grid[node] = om.overwrite ? v / om.scale - om.shift : grid[node] + v;
}

struct TileBox {
    int tx, ty, tz;        // tile coordinates
    int ox, oy, oz;        // origin
    int ex, ey, ez;        // extent
};

template <int AXES>
__host__ __device__ __forceinline__ TileBox tile_box(int tile, const TileGeom &g) {
    TileBox b;
    b.tz = tile % g.ntz;
    b.ty = (tile / g.ntz) % g.nt;
    b.tx = tile / (g.ntz * g.nt);
    b.ox = g.orig[b.tx];
    b.oy = g.orig[b.ty];
    b.ex = g.orig[b.tx + 1] - b.ox;
    b.ey = g.orig[b.ty + 1] - b.oy;
    if (AXES == 3) {
        b.oz = g.orig[b.tz];
        b.ez = g.orig[b.tz + 1] - b.oz;
    } else {
        b.oz = 0;
        b.ez = 1;
    }
    return b;
}

// Sorted particle records: (dist_x, dist_y, dist_z, W) as float4 when the
// particles carry weights, (dist_x, dist_y, dist_z) packed in 12 bytes
// otherwise (w reads as 1).
template <bool RW>
__device__ __forceinline__ float4 load_record(const float *__restrict__ records,
                                              unsigned long long i) {
    if (RW) return reinterpret_cast<const float4 *>(records)[i];
    const float *p = records + 3 * i;
    return make_float4(p[0], p[1], p[2], 1.0f);
}

// Record j of a wave-uniform set starting at `set` (32-bit lane offsets).
template <bool RW>
__device__ __forceinline__ float4 load_set_record(const float *__restrict__ set, unsigned j) {
    if (RW) return reinterpret_cast<const float4 *>(set)[j];
    const float *p = set + 3u * j;
    return make_float4(p[0], p[1], p[2], 1.0f);
}

// Box offset of a record's first footprint node (records of this tile satisfy
// 0 <= l < e on every axis; the clamp only guards the LDS box).
template <int MAS, int AXES>
__host__ __device__ __forceinline__ int box_base(const float4 &rec, const TileGeom &g, const TileBox &tb) {
    constexpr int F = MAS;
    constexpr int T = (AXES == 3) ? kT3 : kT2;
    constexpr int BY = T + F - 1;
    constexpr int BZ = (AXES == 3) ? T + F - 1 : 1;
    const int lx = min(max(wrap_node(first_node<MAS>(rec.x), g.dims) - tb.ox, 0), tb.ex - 1);
    const int ly = min(max(wrap_node(first_node<MAS>(rec.y), g.dims) - tb.oy, 0), tb.ey - 1);
    int lz = 0;
    if (AXES == 3) lz = min(max(wrap_node(first_node<MAS>(rec.z), g.dims) - tb.oz, 0), tb.ez - 1);
    return (lx * BY + ly) * BZ + lz;
}

// Uniform 3D tiles (dims % 16 == 0, 16 nodes per tile): tile origins and dims
// are multiples of 16, so the box coordinate of a record of this tile, (first
// node wrapped into [0, dims)) - origin, equals the unwrapped first node mod 16
// on every axis.  The mask also keeps any record inside the box.
template <int MAS>
__host__ __device__ __forceinline__ int box_base_uni(const float4 &rec) {
    constexpr int B = kT3 + MAS - 1;
    const int lx = first_node<MAS>(rec.x) & (kT3 - 1);
    const int ly = first_node<MAS>(rec.y) & (kT3 - 1);
    const int lz = first_node<MAS>(rec.z) & (kT3 - 1);
    return (lx * B + ly) * B + lz;
}

// Bank class of a record for dealing it to an atomic round: its box offset mod
// 16 up to a constant of the tile, from its first nodes up to a constant
// ((int)(d + 0.5f) for NGP and TSC, (int)d for CIC and PCS; a record whose
// approximation is off, or that wraps around the periodic edge, only lands in
// another class).  The class steers bank balance alone: every record is
// deposited at its exact box offset whatever class it was dealt with.
template <int MAS, int AXES>
__host__ __device__ __forceinline__ unsigned bank_class(const float4 &rec) {
    constexpr unsigned T = (AXES == 3) ? kT3 : kT2;
    constexpr unsigned BY = T + MAS - 1, BZ = (AXES == 3) ? T + MAS - 1 : 1;
    constexpr float h = (MAS == PYL_NGP || MAS == PYL_TSC) ? 0.5f : 0.0f;
    const unsigned fx = (unsigned)(int)(rec.x + h), fy = (unsigned)(int)(rec.y + h);
    if (AXES == 2) return (fx * (BY & 15u) + fy) & 15u;
    const unsigned fz = (unsigned)(int)(rec.z + h);
    return (fx * ((BY * BZ) & 15u) + fy * (BZ & 15u) + fz) & 15u;
}

// Adds one record's footprint to the LDS box at compile-time offsets from `a`.
// 64-bit fixed point of s (|s| < 2^62, in quanta): hf = s * 2^-32 rounded to an
// integer and lf = s - hf * 2^32 are exact floats (|lf| <= 2^31: s and hf * 2^32
// are multiples of s's ulp, and the difference has at most 23 significant bits), lf
// is truncated toward zero, so the result is s to within one quantum.  A tie (s an
// odd multiple of 2^31 whose hf rounds down to even) gives lf = 2^31, which int32
// cannot hold: it is taken as 2^31 - 128, 2^-24 of |s| (s's own float resolution).
// Six instructions (mul, rndne, fma, min, two conversions) and a 64-bit add of the
// sign of lf.
__host__ __device__ __forceinline__ unsigned long long fix64(float s) {
    const float hf = rintf(s * 0x1p-32f);
    const float lf = fminf(fmaf(hf, -0x1p32f, s), 0x1.fffffep30f);
    const int lo = (int)lf;
    return ((unsigned long long)(unsigned)((int)hf + (lo >> 31)) << 32) | (unsigned)lo;
}

// ((Cx*Cy)*Cz)*(W*2^-q) == 2^-q * (((Cx*Cy)*Cz)*W) exactly (power-of-two
// scaling commutes with float rounding), so the scaled product is the
// reference's float contribution in units of the quantum; it is truncated
// toward zero to an integer (|error| < 1 quantum = 2^-30 of the weight bound,
// one conversion instruction).  With per-particle weights (FX) the quantum is
// 2^-61 of the bound times the realisation's record count, rounded up to a power of two, and
// each contribution is converted to 64 bits (fix64), so a cell's error stays
// below 1e-5 of the mean cell value even for weights spanning 5 decades.  Without weights (Cx*Cy)*(Cz*2^-q) is the same
// number: it can differ only where (Cx*Cy)*Cz is subnormal, far below one
// quantum, and both truncate to 0.
template <int MAS, int AXES, bool WT, bool FX>
__device__ __forceinline__ void deposit_record(unsigned long long *a, const float4 &rec,
                                               float to_q, float weight_factor) {
    constexpr int F = MAS;
    constexpr int T = (AXES == 3) ? kT3 : kT2;
    constexpr int BY = T + F - 1;
    constexpr int BZ = (AXES == 3) ? T + F - 1 : 1;
    float Cx[4], Cy[4], Cz[4];
    foot_weights<MAS>(rec.x, first_node<MAS>(rec.x), Cx);
    foot_weights<MAS>(rec.y, first_node<MAS>(rec.y), Cy);
    if (AXES == 3) foot_weights<MAS>(rec.z, first_node<MAS>(rec.z), Cz);
    if (WT) {
        const float w = (rec.w * weight_factor) * to_q;
#pragma unroll
        for (int l = 0; l < F; ++l) {
#pragma unroll
            for (int m = 0; m < F; ++m) {
                const float cxy = Cx[l] * Cy[m];
                if (AXES == 3) {
#pragma unroll
                    for (int k = 0; k < F; ++k) {
                        const float c = (cxy * Cz[k]) * w;
                        atomicAdd(a + (l * BY + m) * BZ + k,
                                  FX ? fix64(c) : (unsigned long long)(long long)(int)c);
                    }
                } else {
                    const float c = cxy * w;   // the reference multiplies by C[2][n] = 1.0
                    atomicAdd(a + l * BY + m, FX ? fix64(c) : (unsigned long long)(long long)(int)c);
                }
            }
        }
    } else {
        float Cs[4];                         // last axis' weights in quanta (>= 0)
        if constexpr (MAS == PYL_TSC && AXES == 3) {
            // 0.5*q*to_q == q*(0.5*to_q) exactly (power-of-two scaling; a
            // subnormal q is far below one quantum either way)
            float q0, q2;
            tsc_parts(rec.z, first_node<MAS>(rec.z), q0, Cs[1], q2);
            const float h_q = 0.5f * to_q;
            Cs[0] = q0 * h_q;
            Cs[1] *= to_q;
            Cs[2] = q2 * h_q;
        } else {
#pragma unroll
            for (int k = 0; k < F; ++k) Cs[k] = (AXES == 3 ? Cz[k] : Cy[k]) * to_q;
        }
#pragma unroll
        for (int l = 0; l < F; ++l) {
            if (AXES == 3) {
#pragma unroll
                for (int m = 0; m < F; ++m) {
                    const float cxy = Cx[l] * Cy[m];
#pragma unroll
                    for (int k = 0; k < F; ++k)
                        atomicAdd(a + (l * BY + m) * BZ + k,
                                  (unsigned long long)(unsigned)(cxy * Cs[k]));
                }
            } else {
#pragma unroll
                for (int m = 0; m < F; ++m)
                    atomicAdd(a + l * BY + m, (unsigned long long)(unsigned)(Cx[l] * Cs[m]));
            }
        }
    }
}

// Records each wave stages per round set in the bank-balanced deposit: KR per
// lane (fewer for the PCS 3D box, so that two workgroups still fit one CU).
template <int MAS, int AXES>
constexpr int dep_kr() { return (MAS == 4 && AXES == 3) ? 2 : 4; }

// ---- bank-balanced dealing of a wave's records to atomic rounds -------------
// A wave instruction ds_add_u64 is served as four 16-lane passes (lanes 0-15,
// 16-31, 32-47, 48-63).  An 8-byte word at box offset e sits in the bank pair
// e mod 16, so a pass costs the largest number of its lanes whose offsets agree
// mod 16 (more when they hit the same word), and the instruction costs about
// max(6 clk, sum of its four passes) on MI355X: one 2-way pass in each of two
// groups is free, a 2-way conflict in all four costs 8.3 clk, random offsets
// 12.6 clk.
//
// A wave deals KR*64 records to KR rounds, i.e. to G = 4*KR "layers" of 16
// lanes.  Records are classed by e mod 16 (c_j records in class j).  Copy k < G
// of class j goes to layer k, lane j: every layer holds at most one record per
// class.  Layer k lacks class j when c_j <= k < G; those slots are holes, and
// the copies k >= G (excess) fill them, holes listed class by class (class j's
// holes lie in distinct layers, so one excess class rarely meets itself twice
// in a layer).  Layer L runs in round L % KR as lane group L / KR: the high,
// hole-bearing layers of a round are in different groups.  Holes left empty
// (a set of fewer than KR*64 records) are marked kNoRecord.
constexpr unsigned short kNoRecord = 0xffff;

template <int KR>
__host__ __device__ __forceinline__ int deal_slot(unsigned layer, unsigned lane) {
    return (int)((layer % KR) * 64 + (layer / KR) * 16 + lane);
}

// Slot (layer L, lane j) when class j has cj records: a hole if cj <= L < G, with
// hole index H + L - cj (H = holes of the classes before j).  The first E holes
// of the wave (E = number of excess copies) receive an excess copy through the
// hole table ht (entry layer*16 + lane), the others are marked empty in sidx.
template <int KR>
__host__ __device__ __forceinline__ void deal_hole(unsigned L, unsigned j, unsigned cj, unsigned H,
                                                   unsigned E, unsigned char *ht, unsigned short *sidx) {
    constexpr unsigned G = 4 * KR;
    if (L < cj || L >= G) return;
    const unsigned i = H + (L - cj);
    if (i < E) ht[i] = (unsigned char)((L << 4) | j);
    else sidx[deal_slot<KR>(L, j)] = kNoRecord;
}

// Slot of copy `rank` of class `cls`; X = index of the class's first excess copy.
template <int KR>
__host__ __device__ __forceinline__ int deal_record_slot(unsigned cls, unsigned rank, unsigned X,
                                                         const unsigned char *ht) {
    constexpr unsigned G = 4 * KR;
    if (rank < G) return deal_slot<KR>(rank, cls);
    const unsigned e = ht[X + rank - G];
    return deal_slot<KR>(e >> 4, e & 15u);
}

// Quantum 2^q of a realisation's box sums: footprint weights are <= 1, so
// |contribution| <= weight_factor * max finite |W| of the realisation < 2^e.
// With q = e - 30 every rounded contribution is an int32 and every box sum an
// int64 for any count below 2^33: exact integer atomics, independent of their
// order.
template <bool WT>
__device__ __forceinline__ void box_quantum(const unsigned *__restrict__ wmax_bits, int real,
                                            float weight_factor, unsigned long long nrec,
                                            float &to_q, double &from_q) {
    const float wmax = (WT && wmax_bits) ? __uint_as_float(wmax_bits[real]) : 1.0f;
    const double bound = (double)wmax * (double)weight_factor;
    int e = -1000;
    if (bound > 0.0) frexp(bound, &e);
    // contributions below 2^30 quanta (int32); with per-particle weights (records
    // carry W, wmax_bits set) below 2^(61 - lg) quanta, lg = ceil(log2(nrec)), so
    // that any nrec of them sum below 2^61 in magnitude (int64)
    int q = e - 30;
    if (WT && wmax_bits) q = e - 61 + (nrec > 1 ? 64 - __clzll((long long)(nrec - 1)) : 0);
    if (q < -127) q = -127;                 // keep 2^-q a normal float
    if (q > 127) q = 127;
    to_q = ldexpf(1.0f, -q);
    from_q = ldexp(1.0, q);
}

// Writes a tile's box of integer sums (LDS or global): the nodes no other tile
// touches go to the grid (grid already offset to the realisation), the band and
// halo to the tile's partial slots (halo_index order).
template <int MAS, int AXES, bool UNI>
__device__ __forceinline__ void flush_box(const unsigned long long *acc, int key, const TileBox &tb,
                                          const TileGeom &g, float *__restrict__ grid,
                                          float *__restrict__ partials, const OutMode &om,
                                          double from_q, int ex, int ey, int ez) {
    constexpr int F = MAS;
    constexpr int T = (AXES == 3) ? kT3 : kT2;
    constexpr int BY = T + F - 1;
    constexpr int BZ = (AXES == 3) ? T + F - 1 : 1;
    constexpr int BX = T + F - 1;
    constexpr int NB = BX * BY * BZ;
    float *part = partials + (long long)key * g.pslot;
    if constexpr (UNI && AXES == 3) {
        // Uniform tile: the interior nodes go to the grid, the halo to the
        // partial slots in halo_index order, region by region (x outside the
        // interior: whole planes; x inside, y outside: rows of B; x, y inside, z
        // outside: H nodes), so that every loop has constant extents and the
        // partial stores are contiguous.
        constexpr int B = BX, H = 2 * (F - 1), I = T - (F - 1);
        const unsigned long long node0 =
            ((unsigned long long)tb.ox * g.dims + tb.oy) * g.dims + tb.oz;
        const unsigned dd = (unsigned)g.dims * (unsigned)g.dims;
        for (int i = threadIdx.x; i < I * I * I; i += blockDim.x) {
            const int x = F - 1 + i / (I * I), y = F - 1 + (i / I) % I, z = F - 1 + i % I;
            const float v = (float)((double)(long long)acc[(x * B + y) * B + z] * from_q);
            store_node(grid, (long long)(node0 + (unsigned)x * dd + (unsigned)(y * g.dims + z)), v, om);
        }
        for (int i = threadIdx.x; i < H * B * B; i += blockDim.x) {
            const int xi = i / (B * B), yz = i - xi * (B * B);
            const int x = xi < F - 1 ? xi : xi + I;
            part[i] = (float)((double)(long long)acc[x * B * B + yz] * from_q);
        }
        constexpr int SA = H * B * B;
        for (int i = threadIdx.x; i < I * H * B; i += blockDim.x) {
            const int x = F - 1 + i / (H * B), r = i % (H * B), yi = r / B, z = r - yi * B;
            const int y = yi < F - 1 ? yi : yi + I;
            part[SA + i] = (float)((double)(long long)acc[(x * B + y) * B + z] * from_q);
        }
        constexpr int SB = SA + I * H * B;
        for (int i = threadIdx.x; i < I * I * H; i += blockDim.x) {
            const int x = F - 1 + i / (I * H), r = i % (I * H), y = F - 1 + r / H, zi = r % H;
            const int z = zi < F - 1 ? zi : zi + I;
            part[SB + i] = (float)((double)(long long)acc[(x * B + y) * B + z] * from_q);
        }
        return;
    }
    for (int c = threadIdx.x; c < NB; c += blockDim.x) {
        const int bz = c % BZ, by = (c / BZ) % BY, bx = c / (BZ * BY);
        if (!UNI && (bx >= ex + F - 1 || by >= ey + F - 1 || (AXES == 3 && bz >= ez + F - 1)))
            continue;
        const float v = (float)((double)(long long)acc[c] * from_q);
        const bool inner = bx >= F - 1 && bx < ex && by >= F - 1 && by < ey &&
                           (AXES == 2 || (bz >= F - 1 && bz < ez));
        if (inner) {
            const long long node = AXES == 3
                ? ((long long)(tb.ox + bx) * g.dims + (tb.oy + by)) * g.dims + (tb.oz + bz)
                : (long long)(tb.ox + bx) * g.dims + (tb.oy + by);
            store_node(grid, node, v, om);
        } else {
            part[halo_index<F, AXES>(bx, by, bz, ex, ey, ez)] = v;
        }
    }
}

// One workgroup per tile.  LDS box of BX*BY*BZ 64-bit integers (the tile plus
// the upper halo) holding multiples of 2^q; every record adds its whole
// footprint at fixed offsets from its first node.  WT = false: unit weights
// and weight_factor 1, so every contribution is >= 0 and the power-of-two
// scale can be folded into the last axis' weights.
//
// Bank balancing (F >= 2): a footprint's F^AXES atomics all share the lanes'
// box offsets up to a constant, so their LDS bank conflicts are fixed by the
// offsets mod 16.  Each wave takes KR*64 consecutive records, ranks them by
// class (LDS counters), deals them to rounds as described at deal_slot, stages
// the dealt record indices (2 B each) in LDS and reloads the records (from
// cache).  The sums are integers, so the order of the atomics does not change
// any bit of the result.  With `order` (order_place_kernel), workgroup i takes
// key order[i] instead of key i.
template <int MAS, int AXES, bool WT, bool RW, bool UNI>
__global__ void __launch_bounds__(kDepBlock)
tile_deposit_kernel(const float *__restrict__ records, unsigned long long rcap,
                    const unsigned long long *__restrict__ offsets, TileGeom g,
                    float weight_factor, const unsigned *__restrict__ wmax_bits,
                    float *__restrict__ grid, float *__restrict__ partials, OutMode om,
                    const unsigned *__restrict__ order, int nsplit,
                    unsigned long long *__restrict__ gbox, CapRecs cr) {
    constexpr int F = MAS;
    constexpr int T = (AXES == 3) ? kT3 : kT2;
    constexpr int BY = T + F - 1;
    constexpr int BZ = (AXES == 3) ? T + F - 1 : 1;
    constexpr int BX = T + F - 1;
    constexpr int NB = BX * BY * BZ;
    constexpr bool BAL = F >= 2;
    constexpr int KR = BAL ? dep_kr<MAS, AXES>() : 1;
    constexpr int NW = kDepBlock / 64;                 // waves
    constexpr int WB = KR * 64;                        // records per wave round set
    __shared__ unsigned long long acc[NB];
    __shared__ unsigned short sidx[BAL ? NW * WB : 1];    // dealt record indices
    __shared__ unsigned char htab[BAL ? NW * WB : 1];      // hole table
    __shared__ unsigned whist[BAL ? NW * 16 : 1];          // class counters
    // key r * ntiles + tile; with nsplit > 1, workgroup i takes part i % nsplit of
    // key i / nsplit's records and adds its box into gbox (tile_box_flush_kernel
    // writes the grid)
    const int key = nsplit > 1 ? (int)(blockIdx.x / (unsigned)nsplit)
                               : (order ? (int)order[blockIdx.x] : (int)blockIdx.x);
    const int real = key / (int)g.ntiles;
    const TileBox tb = tile_box<AXES>(key - real * (int)g.ntiles, g);
    grid += real * g.ncell;
    // record range, clamped to the buffer (a miscount is reported by the host)
    unsigned long long r1 = min(offsets[key + 1], rcap), r0 = min(offsets[key], r1);
    if (nsplit > 1) {
        const unsigned long long part = (r1 - r0 + nsplit - 1) / nsplit;
        r0 = min(r0 + part * (blockIdx.x % (unsigned)nsplit), r1);
        r1 = min(r0 + part, r1);
        if (r0 == r1) return;                     // no records in this part
    }
    // Record segments: the tile's capacity slots (capacity sort), then its run of
    // the sorted records.  recs, s0, s1: the segment being deposited.
    unsigned long long c0 = 0, c1 = 0;
    if (cr.rec) {
        c0 = (unsigned long long)key * cr.cap;
        c1 = c0 + min(cr.count[key], cr.cap);
    }
    const bool two_seg = c1 > c0 && r1 > r0;
    const float *recs = c1 > c0 ? cr.rec : records;
    unsigned long long s0 = c1 > c0 ? c0 : r0, s1 = c1 > c0 ? c1 : r1;
    // UNI: every tile has extent T (dims % T == 0, two tiles or more per axis)
    const int ex = UNI ? T : tb.ex, ey = UNI ? T : tb.ey, ez = UNI ? (AXES == 3 ? T : 1) : tb.ez;

    if (s0 == s1) {
        // Empty tile: its interior is zero; the finalize reads no halo from it.
        if (om.overwrite) {
            const int n = ex * ey * ez;
            for (int c = threadIdx.x; c < n; c += blockDim.x) {
                const int lz = c % ez, ly = (c / ez) % ey, lx = c / (ez * ey);
                if (lx < F - 1 || ly < F - 1 || (AXES == 3 && lz < F - 1)) continue;
                const long long node = AXES == 3
                    ? ((long long)(tb.ox + lx) * g.dims + (tb.oy + ly)) * g.dims + (tb.oz + lz)
                    : (long long)(tb.ox + lx) * g.dims + (tb.oy + ly);
                store_node(grid, node, 0.0f, om);
            }
        }
        return;
    }
    // the first round set of every wave is loaded before the box is cleared
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    constexpr int RS = RW ? 4 : 3;                     // floats per record
    // records b .. b+WB-1 of a set, clamped to the tile's last record, through a
    // wave-uniform pointer and 32-bit lane offsets
    auto load_set = [&](unsigned long long b, float4 *dst) {
        const unsigned long long sb = min(b, s1 - 1);
        const float *set = recs + (unsigned long long)RS * sb;
        const unsigned lim = (unsigned)min(s1 - 1 - sb, (unsigned long long)WB);
#pragma unroll
        for (int k = 0; k < KR; ++k) dst[k] = load_set_record<RW>(set, min((unsigned)(ln + 64 * k), lim));
    };
    float4 nxt[KR];
    if constexpr (BAL) load_set(s0 + (unsigned long long)wv * WB, nxt);
    for (int c = threadIdx.x; c < NB; c += blockDim.x) acc[c] = 0ull;

    float to_q;
    double from_q;
    // the realisation's record count, so that a batch and single calls agree bitwise
    const unsigned long long nrec = cr.nrec >= 0 ? (unsigned long long)cr.nrec
        : offsets[(long long)(real + 1) * g.ntiles] - offsets[(long long)real * g.ntiles];
    box_quantum<WT>(wmax_bits, real, weight_factor, nrec, to_q, from_q);
    __syncthreads();

    if constexpr (!BAL) {
        for (int seg = 0;; ++seg) {
            // the next record is loaded while the current one is deposited
            unsigned long long r = s0 + threadIdx.x;
            float4 nxt = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < s1) nxt = load_record<RW>(recs, r);
            for (; r < s1; r += blockDim.x) {
                const float4 rec = nxt;
                if (r + blockDim.x < s1) nxt = load_record<RW>(recs, r + blockDim.x);
                deposit_record<MAS, AXES, WT, RW>(acc + box_base<MAS, AXES>(rec, g, tb), rec, to_q,
                                                  weight_factor);
            }
            if (seg > 0 || !two_seg) break;
            recs = records;
            s0 = r0;
            s1 = r1;
        }
    } else {
        unsigned *wh = whist + wv * 16;
        unsigned short *si = sidx + wv * WB;
        unsigned char *ht = htab + wv * WB;
        constexpr unsigned G = 4 * KR;                 // layers of 16 lanes
        // wave wv takes the round sets starting at s0 + (wv + NW*j)*WB; the next
        // set's records are loaded (clamped into the segment) while this one runs
        for (int seg = 0;; ++seg) {
            unsigned long long b0 = s0 + (unsigned long long)wv * WB;
            for (; b0 < s1; b0 += (unsigned long long)NW * WB) {
                float4 rec[KR];
#pragma unroll
                for (int k = 0; k < KR; ++k) rec[k] = nxt[k];
                const int nb = (int)min((unsigned long long)WB, s1 - b0);   // records in this set
                // rank within class (LDS ops of one wave run in issue order)
                if (ln < 16) wh[ln] = 0u;
                __builtin_amdgcn_wave_barrier();
                unsigned cls[KR], rank[KR];
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    cls[k] = bank_class<MAS, AXES>(rec[k]);
                    rank[k] = (ln + 64 * k < nb) ? atomicAdd(&wh[cls[k]], 1u) : 0u;
                }
                __builtin_amdgcn_wave_barrier();
                // lanes 0..15 hold class j = ln: c_j records, G - c_j holes or c_j - G
                // excess copies.  Exclusive prefixes of both (excess in the high half)
                // by an inclusive DPP scan over row 0 (row_shr 1, 2, 4, 8).
                const unsigned cj = (ln < 16) ? wh[ln] : 0u;
                const unsigned v = (ln < 16) ? ((cj > G ? cj - G : 0u) << 16) | (cj < G ? G - cj : 0u) : 0u;
                unsigned incl = v;
                incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x111, 0xf, 0xf, false);
                incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x112, 0xf, 0xf, false);
                incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x114, 0xf, 0xf, false);
                incl += (unsigned)__builtin_amdgcn_update_dpp(0, (int)incl, 0x118, 0xf, 0xf, false);
                const unsigned excl = incl - v;
                const unsigned E = (unsigned)__builtin_amdgcn_readlane((int)incl, 15) >> 16;
                __builtin_amdgcn_wave_barrier();
                // class j's count, first hole and first excess index (9 bits each)
                if (ln < 16) wh[ln] = cj | (excl & 0xffffu) << 9 | (excl >> 16) << 18;
                __builtin_amdgcn_wave_barrier();
                {
                    // the G*16 (layer, lane) slots, KR per lane: lane ln takes class ln % 16
                    const unsigned cw = wh[ln & 15];
#pragma unroll
                    for (int k = 0; k < KR; ++k)
                        deal_hole<KR>((unsigned)(ln >> 4) + 4u * k, (unsigned)(ln & 15), cw & 511u,
                                      (cw >> 9) & 511u, E, ht, si);
                }
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int k = 0; k < KR; ++k) {
                    if (ln + 64 * k < nb) {
                        const unsigned X = rank[k] < G ? 0u : wh[cls[k]] >> 18;
                        si[deal_record_slot<KR>(cls[k], rank[k], X, ht)] = (unsigned short)(ln + 64 * k);
                    }
                }
                __builtin_amdgcn_wave_barrier();
                // the dealt records of all rounds are reloaded (from cache) before the
                // first atomic, and the next set is loaded only after them: loads
                // complete in issue order, so the atomics wait for the reloads alone
                // and the next set's memory latency overlaps this set's atomics.  The
                // next-set loads are issued unconditionally (past the tile's end they
                // all read its last record) so that the number of loads in flight is
                // the same on every path and the waits before the atomics stay exact.
                unsigned short di[KR];
                float4 dealt[KR];
#pragma unroll
                for (int r = 0; r < KR; ++r) di[r] = si[r * 64 + ln];
                const float *set = recs + (unsigned long long)RS * b0;
#pragma unroll
                for (int r = 0; r < KR; ++r)
                    dealt[r] = load_set_record<RW>(set, di[r] != kNoRecord ? di[r] : 0u);
                __builtin_amdgcn_sched_barrier(0);         // keep the reloads ahead of the next set
                load_set(b0 + (unsigned long long)NW * WB, nxt);
                __builtin_amdgcn_sched_barrier(0);
                // every reloaded record is consumed here, on all lanes: the reloads
                // are not sunk into the per-round branches below (where they would be
                // issued after the next set's loads), and one wait covers them
#pragma unroll
                for (int r = 0; r < KR; ++r) {
                    if constexpr (RW)
                        asm volatile("" ::"v"(dealt[r].x), "v"(dealt[r].y), "v"(dealt[r].z), "v"(dealt[r].w));
                    else
                        asm volatile("" ::"v"(dealt[r].x), "v"(dealt[r].y), "v"(dealt[r].z));
                }
#pragma unroll
                for (int r = 0; r < KR; ++r) {
                    if (di[r] != kNoRecord) {
                        int bb;
                        if constexpr (UNI && AXES == 3) bb = box_base_uni<MAS>(dealt[r]);
                        else bb = box_base<MAS, AXES>(dealt[r], g, tb);
                        deposit_record<MAS, AXES, WT, RW>(acc + bb, dealt[r], to_q, weight_factor);
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (seg > 0 || !two_seg) break;
            // the tile's sorted records: the first set is loaded before the loop, as above
            recs = records;
            s0 = r0;
            s1 = r1;
            load_set(s0 + (unsigned long long)wv * WB, nxt);
        }
    }
    __syncthreads();
    if (nsplit > 1) {
        // integer sums: the parts of a tile add up exactly in any order
        unsigned long long *gb = gbox + (long long)key * NB;
        for (int c = threadIdx.x; c < NB; c += blockDim.x) {
            const unsigned long long v = acc[c];
            if (v) atomicAdd(&gb[c], v);
        }
        return;
    }
    flush_box<MAS, AXES, UNI>(acc, key, tb, g, grid, partials, om, from_q, ex, ey, ez);
}

// Split deposit, second pass: one workgroup per key writes the summed box of its
// parts (all keys, also empty ones: their box is zero).
template <int MAS, int AXES, bool WT, bool UNI>
__global__ void __launch_bounds__(kDepBlock)
tile_box_flush_kernel(const unsigned long long *__restrict__ gbox, TileGeom g, float weight_factor,
                      const unsigned *__restrict__ wmax_bits,
                      const unsigned long long *__restrict__ offsets, float *__restrict__ grid,
                      float *__restrict__ partials, OutMode om) {
    constexpr int T = (AXES == 3) ? kT3 : kT2;
    constexpr int NB = (T + MAS - 1) * (T + MAS - 1) * ((AXES == 3) ? T + MAS - 1 : 1);
    const int key = (int)blockIdx.x;
    const int real = key / (int)g.ntiles;
    const TileBox tb = tile_box<AXES>(key - real * (int)g.ntiles, g);
    const int ex = UNI ? T : tb.ex, ey = UNI ? T : tb.ey, ez = UNI ? (AXES == 3 ? T : 1) : tb.ez;
    float to_q;
    double from_q;
    const unsigned long long nrec = offsets[(long long)(real + 1) * g.ntiles] - offsets[(long long)real * g.ntiles];
    box_quantum<WT>(wmax_bits, real, weight_factor, nrec, to_q, from_q);
    flush_box<MAS, AXES, UNI>(gbox + (long long)key * NB, key, tb, g, grid + real * g.ncell, partials, om,
                              from_q, ex, ey, ez);
}

// Deposit order: keys by decreasing record count (16 buckets per octave), so
// that the largest tiles start first and the smallest fill the end; in key order
// the big tiles of a dense region can start last and leave most CUs idle while
// they finish.  The order within a bucket is arbitrary, and no result depends on
// it (every tile's sums are its own integers).
constexpr int kOrderBlock = 1024;
constexpr int kOrderBuckets = 512;
constexpr long long kOrderMinKeys = 4096;
constexpr long long kOrderMinRecords = 512;      // mean records per key
// Split deposit: below kSplitMaxKeys tiles, a tile holding more than
// kSplitRecords records per workgroup on average is shared by up to kSplitMax
// workgroups.
constexpr long long kSplitMaxKeys = 2048;
constexpr long long kSplitRecords = 16384;
constexpr long long kSplitMax = 256;
// Capacity sort: tile slots kCapFactor times the mean occupancy; adaptive use from
// kCapMinMean particles per tile on average; a spill above 1/kCapSpillDen of the
// particles turns it off for the (dims, MAS); the spill list is sorted in
// kSpillChunks chunks.
constexpr long long kCapFactor = 3;
constexpr long long kCapMinMean = 256;
constexpr unsigned long long kCapSpillDen = 64;
constexpr int kSpillChunks = 512;
constexpr int kCapRetry = 16;                   // calls without the capacity form after a heavy spill

// 16 buckets per octave of the record count (0 alone in bucket 0); distinct
// buckets also spread the LDS histogram's atomics over more addresses
__host__ __device__ __forceinline__ int count_bucket(unsigned long long n) {
    if (n == 0) return 0;
    const int lg = 63 - __builtin_clzll(n);           // floor(log2 n), n > 0
    const int q = lg >= 4 ? (int)((n >> (lg - 4)) & 15ull) : (int)((n << (4 - lg)) & 15ull);
    return min(kOrderBuckets - 1, 1 + 16 * lg + q);
}

// Records of key k: its run of the sorted records plus its capacity slots.
__device__ __forceinline__ unsigned long long key_records(const unsigned long long *__restrict__ offsets,
                                                          long long k, const CapRecs &cr) {
    return offsets[k + 1] - offsets[k] + cap_count(cr, k);
}

// The order is built in two launches of order_wgs(nt) workgroups, each thread
// taking up to kOrderPer keys (k = its global index + j * threads): bucket sizes
// (order_hist_kernel), then the keys placed at their bucket's start plus a run
// claimed per workgroup (order_place_kernel).  Counters sit in LDS per workgroup
// and only their totals go to the global ones (ghist, gcur: zeroed by the call).
constexpr int kOrderPer = 8;
constexpr int kOrderMinWGs = 32;
inline int order_wgs(long long nt) {
    return (int)std::max<long long>(kOrderMinWGs, (nt + (long long)kOrderBlock * kOrderPer - 1) /
                                                      ((long long)kOrderBlock * kOrderPer));
}

__global__ void __launch_bounds__(kOrderBlock)
order_hist_kernel(const unsigned long long *__restrict__ offsets, long long nt, CapRecs cr,
                  unsigned *__restrict__ ghist) {
    __shared__ unsigned hist[kOrderBuckets];
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) hist[b] = 0u;
    __syncthreads();
    const long long k0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long step = (long long)gridDim.x * blockDim.x;
    unsigned long long c[kOrderPer];
#pragma unroll
    for (int j = 0; j < kOrderPer; ++j) {        // all loads before the first atomic
        const long long k = k0 + j * step;
        c[j] = k < nt ? key_records(offsets, k, cr) : 0ull;
    }
#pragma unroll
    for (int j = 0; j < kOrderPer; ++j)
        if (k0 + j * step < nt) atomicAdd(&hist[count_bucket(c[j])], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
        if (hist[b]) atomicAdd(&ghist[b], hist[b]);
}

__global__ void __launch_bounds__(kOrderBlock)
order_place_kernel(const unsigned long long *__restrict__ offsets, long long nt, CapRecs cr,
                   const unsigned *__restrict__ ghist, unsigned *__restrict__ gcur,
                   unsigned *__restrict__ order) {
    __shared__ unsigned start[kOrderBuckets];   // first slot of each bucket, largest first
    __shared__ unsigned lrun[kOrderBuckets];    // this workgroup's keys per bucket, then its cursor
    if (threadIdx.x < 64) {                      // exclusive scan of ghist from the top bucket down
        constexpr int PB = kOrderBuckets / 64;   // buckets per lane
        const int top = kOrderBuckets - 1 - (int)threadIdx.x * PB;
        unsigned h[PB], sum = 0;
#pragma unroll
        for (int j = 0; j < PB; ++j) { h[j] = ghist[top - j]; sum += h[j]; }
        unsigned incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned t = __shfl_up(incl, o);
            if ((int)threadIdx.x >= o) incl += t;
        }
        unsigned run = incl - sum;
#pragma unroll
        for (int j = 0; j < PB; ++j) { start[top - j] = run; run += h[j]; }
    }
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) lrun[b] = 0u;
    __syncthreads();
    const long long k0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long step = (long long)gridDim.x * blockDim.x;
    int bk[kOrderPer];
#pragma unroll
    for (int j = 0; j < kOrderPer; ++j) {
        const long long k = k0 + j * step;
        bk[j] = k < nt ? count_bucket(key_records(offsets, k, cr)) : 0;
    }
#pragma unroll
    for (int j = 0; j < kOrderPer; ++j)
        if (k0 + j * step < nt) atomicAdd(&lrun[bk[j]], 1u);
    __syncthreads();
    // one claim per bucket this workgroup holds keys of
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
        if (lrun[b]) lrun[b] = start[b] + atomicAdd(&gcur[b], lrun[b]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOrderPer; ++j) {
        const long long k = k0 + j * step;
        if (k < nt) order[atomicAdd(&lrun[bk[j]], 1u)] = (unsigned)k;
    }
}

// Images along one axis of a tile t: the tiles whose boxes cover its nodes and
// the offset from a tile-local coordinate l to their box coordinate.  Image 0
// is t itself (offset 0); with nt >= 2, image 1 is the lower neighbour, whose
// halo starts at its extent e(t-1); with one tile along the axis the tile's own
// halo wraps onto it at offsets dims, 2*dims, ...  An image applies to l when
// l + offset lies inside its box, i.e. < extent + F - 1.
template <int F>
__host__ __device__ __forceinline__ int axis_images(int t, int nt, int dims, const int *orig,
                                                    int *it, int *io, int *ie) {
    it[0] = t;
    io[0] = 0;
    ie[0] = orig[t + 1] - orig[t];
    if (F == 1) return 1;
    if (nt >= 2) {
        const int tp = (t == 0) ? nt - 1 : t - 1;
        it[1] = tp;
        io[1] = orig[tp + 1] - orig[tp];
        ie[1] = io[1];
        return 2;
    }
    int n = 1;
    for (int off = dims; off < dims + F - 1 && n < 4; off += dims) {
        it[n] = 0;
        io[n] = off;
        ie[n] = dims;
        ++n;
    }
    return n;
}

// Workgroups are dealt to the 8 XCDs round-robin by launch index, and each XCD
// has its own L2.  The finalize of a tile reads halo partials that its
// neighbours along z (adjacent keys) and y (nt keys away) also read, a few
// bytes apart; this remap gives each XCD a contiguous range of keys, in key
// order, so that those shared cache lines come from HBM once.
__device__ __forceinline__ int xcd_contiguous_block() {
    const unsigned total = gridDim.x, per = total / 8u, L = blockIdx.x;
    return (int)((L < per * 8u) ? (L % 8u) * per + L / 8u : L);
}

// Band cells (l < F-1 along some axis) of every tile: own band value plus the
// halo values of the lower neighbours, summed in double in a fixed order (the
// image order), then written once.  Empty tiles contribute nothing.
template <int MAS, int AXES>
__global__ void __launch_bounds__(kBlock)
tile_finalize_kernel(const unsigned long long *__restrict__ offsets, TileGeom g,
                     const float *__restrict__ partials, float *__restrict__ grid, OutMode om,
                     CapRecs cr) {
    constexpr int F = MAS;
    __shared__ long long s_pbase[64];   // partial base of image combination, -1 if empty
    const int key = xcd_contiguous_block();
    const int real = key / (int)g.ntiles;
    const long long kbase = (long long)real * g.ntiles;
    const TileBox tb = tile_box<AXES>(key - (int)kbase, g);
    grid += real * g.ncell;
    int itx[4], iox[4], iex[4], ity[4], ioy[4], iey[4], itz[4], ioz[4], iez[4];
    const int nx = axis_images<F>(tb.tx, g.nt, g.dims, g.orig, itx, iox, iex);
    const int ny = axis_images<F>(tb.ty, g.nt, g.dims, g.orig, ity, ioy, iey);
    int nz = 1;
    itz[0] = 0; ioz[0] = 0; iez[0] = 1;
    if (AXES == 3) nz = axis_images<F>(tb.tz, g.ntz, g.dims, g.orig, itz, ioz, iez);
    const int ncomb = nx * ny * nz;
    if ((int)threadIdx.x < ncomb) {
        const int c = threadIdx.x % nz, b = (threadIdx.x / nz) % ny, a = threadIdx.x / (nz * ny);
        const long long vkey = kbase + ((long long)itx[a] * g.nt + ity[b]) * g.ntz + itz[c];
        s_pbase[threadIdx.x] =
            (offsets[vkey] == offsets[vkey + 1] && cap_count(cr, vkey) == 0) ? -1 : vkey * g.pslot;
    }
    __syncthreads();
    const int ex = tb.ex, ey = tb.ey, ez = tb.ez;
    const int nbx = min(F - 1, ex), nby = min(F - 1, ey), nbz = (AXES == 3) ? min(F - 1, ez) : 0;
    const int sa = nbx * ey * ez;
    const int sb = (ex - nbx) * nby * ez;
    const int sc = (AXES == 3) ? (ex - nbx) * (ey - nby) * nbz : 0;
    const int nband = sa + sb + sc;
    for (int t = threadIdx.x; t < nband; t += blockDim.x) {
        int lx, ly, lz;
        if (t < sa) {
            lz = t % ez; ly = (t / ez) % ey; lx = t / (ez * ey);
        } else if (t < sa + sb) {
            const int u = t - sa;
            lz = u % ez; ly = (u / ez) % nby; lx = nbx + u / (ez * nby);
        } else {
            const int u = t - sa - sb;
            lz = u % nbz; ly = nby + (u / nbz) % (ey - nby); lx = nbx + u / (nbz * (ey - nby));
        }
        double v = 0.0;
        for (int a = 0; a < nx; ++a) {
            const int bx = lx + iox[a];
            if (bx >= iex[a] + F - 1) continue;
            for (int b = 0; b < ny; ++b) {
                const int by = ly + ioy[b];
                if (by >= iey[b] + F - 1) continue;
                for (int c = 0; c < nz; ++c) {
                    const int bz = lz + ioz[c];
                    if (AXES == 3 && bz >= iez[c] + F - 1) continue;
                    const long long pb = s_pbase[(a * ny + b) * nz + c];
                    if (pb < 0) continue;
                    v += (double)partials[pb + halo_index<F, AXES>(bx, by, bz, iex[a], iey[b], iez[c])];
                }
            }
        }
        const long long node = AXES == 3
            ? ((long long)(tb.ox + lx) * g.dims + (tb.oy + ly)) * g.dims + (tb.oz + lz)
            : (long long)(tb.ox + lx) * g.dims + (tb.oy + ly);
        store_node(grid, node, (float)v, om);
    }
}

// Band cell t of an equal tile of T = kT3 nodes (t < NA: lx < L; then lx >= L,
// ly < L; then lx, ly >= L, lz < L; L = F-1) and the partial slot of its image
// (a, b, c), i.e. halo_index of (lx + a*T, ly + b*T, lz + c*T) in equal tiles,
// as base + a*dA + b*dB + c*dC (B = T + L box nodes, I = T - L interior and
// H = 2L halo nodes per axis, SA = H*B*B, SB = I*H*B):
//   lx < L:           ((lx*B + ly)*B + lz) + a*L*B*B + b*T*B + c*T
//   lx >= L, ly < L:  SA + ((lx-L)*H + ly)*B + lz + b*L*B + c*T    (a = 0)
//   lx, ly >= L:      SA + SB + ((lx-L)*I + ly-L)*H + lz + c*L     (a = b = 0)
// Image (a, b, c) contributes when each of a, b, c is 0 or its coordinate is
// below L.
template <int MAS>
__host__ __device__ __forceinline__ void band_cell(int t, int &lx, int &ly, int &lz, int &base, int &dA,
                                                   int &dB, int &dC) {
    constexpr int F = MAS, T = kT3, L = F - 1;
    constexpr int B = T + L, I = T - L, H = 2 * L, SA = H * B * B, SB = I * H * B;
    constexpr int NA = L * T * T, NB = (T - L) * L * T;
    constexpr int LL = L > 0 ? L : 1;                    // NGP has no band cells
    if (t < NA) {
        lz = t % T; ly = (t / T) % T; lx = t / (T * T);
        base = (lx * B + ly) * B + lz;
        dA = L * B * B; dB = T * B; dC = T;
    } else if (t < NA + NB) {
        const int u = t - NA;
        lz = u % T; ly = (u / T) % LL; lx = L + u / (T * LL);
        base = SA + ((lx - L) * H + ly) * B + lz;
        dA = 0; dB = L * B; dC = T;
    } else {
        const int u = t - NA - NB;
        lz = u % LL; ly = L + (u / LL) % (T - L); lx = L + u / (LL * (T - L));
        base = SA + SB + ((lx - L) * I + (ly - L)) * H + lz;
        dA = 0; dB = 0; dC = L;
    }
}

// The same finalize for 3D grids cut into equal tiles of kT3 nodes with at
// least two tiles per axis (dims % kT3 == 0, dims > kT3).  Every lower
// neighbour then has extent kT3, so each band cell (lx, ly, lz) takes image 1
// along an axis exactly when its coordinate there is below F-1, and the
// partial slots take the closed forms of band_cell.  The up to 8 loads of a
// cell are issued before any is summed; the sum runs in the image order of the
// general kernel (a, b, c nested), so both kernels give the same bits.
constexpr int kFinBlock = 256;      // uniform finalize workgroup (one tile)

template <int MAS>
__global__ void __launch_bounds__(kFinBlock)
tile_finalize_uniform_kernel(const unsigned long long *__restrict__ offsets, TileGeom g,
                             const float *__restrict__ partials, float *__restrict__ grid,
                             OutMode om, CapRecs cr) {
    constexpr int F = MAS, T = kT3, L = F - 1;
    constexpr int NA = L * T * T;                       // lx < L
    constexpr int NB = (T - L) * L * T;                 // lx >= L, ly < L
    constexpr int NC = (T - L) * (T - L) * L;           // lx, ly >= L, lz < L
    constexpr int NBAND = NA + NB + NC;
    __shared__ long long s_pbase[8];                    // image (a, b, c) at (a*2 + b)*2 + c
    const int key = xcd_contiguous_block();
    const int real = key / (int)g.ntiles;
    const long long kbase = (long long)real * g.ntiles;
    const int tile = key - (int)kbase;
    const int tz = tile % g.nt, ty = (tile / g.nt) % g.nt, tx = tile / (g.nt * g.nt);
    if (threadIdx.x < 8) {
        const int a = threadIdx.x >> 2, b = (threadIdx.x >> 1) & 1, c = threadIdx.x & 1;
        const int ix = a ? (tx == 0 ? g.nt - 1 : tx - 1) : tx;
        const int iy = b ? (ty == 0 ? g.nt - 1 : ty - 1) : ty;
        const int iz = c ? (tz == 0 ? g.nt - 1 : tz - 1) : tz;
        const long long vkey = kbase + ((long long)ix * g.nt + iy) * g.nt + iz;
        s_pbase[threadIdx.x] =
            (offsets[vkey] == offsets[vkey + 1] && cap_count(cr, vkey) == 0) ? -1 : vkey * g.pslot;
    }
    __syncthreads();
    grid += real * g.ncell;
    const float *pim[8];
    bool has[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const long long v = s_pbase[i];
        has[i] = v >= 0;
        pim[i] = partials + (v >= 0 ? v : 0);
    }
    const unsigned long long node0 =
        (((unsigned long long)tx * T * g.dims + (unsigned long long)ty * T) * g.dims) + (unsigned long long)tz * T;
    const unsigned dd = (unsigned)g.dims * (unsigned)g.dims;
    for (int t = threadIdx.x; t < NBAND; t += kFinBlock) {
        int lx, ly, lz, base, dA, dB, dC;
        band_cell<MAS>(t, lx, ly, lz, base, dA, dB, dC);
        float x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int a = i >> 2, b = (i >> 1) & 1, c = i & 1;
            const bool on = (a == 0 || lx < L) && (b == 0 || ly < L) && (c == 0 || lz < L) && has[i];
            x[i] = 0.0f;
            if (on) x[i] = pim[i][base + a * dA + b * dB + c * dC];
        }
        // all loads are issued before the first wait: the values are consumed
        // here, not inside the per-image branches
        asm volatile("" ::"v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]),
                     "v"(x[7]));
        // the sum runs in the image order of the general kernel (a, b, c nested)
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int a = i >> 2, b = (i >> 1) & 1, c = i & 1;
            const bool on = (a == 0 || lx < L) && (b == 0 || ly < L) && (c == 0 || lz < L) && has[i];
            if (on) v += (double)x[i];
        }
        store_node(grid, (long long)(node0 + (unsigned)lx * dd + (unsigned)(ly * g.dims + lz)), (float)v, om);
    }
}

// Plain form: one thread per particle, one global float atomic per node.
template <int MAS, int AXES>
__global__ void __launch_bounds__(kBlock)
atomic_deposit_kernel(const float *__restrict__ pos, RSD rsd, const float *__restrict__ W,
                      long long n, float inv, int dims, float weight_factor,
                      float *__restrict__ grid) {
    constexpr int F = MAS;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float Cx[4], Cy[4], Cz[4];
        const float3 d = load_dist<AXES>(pos, rsd, i, inv);
        const int rx = first_node<MAS>(d.x), ry = first_node<MAS>(d.y);
        foot_weights<MAS>(d.x, rx, Cx);
        foot_weights<MAS>(d.y, ry, Cy);
        const int bx = wrap_node(rx, dims), by = wrap_node(ry, dims);
        int bz = 0;
        if (AXES == 3) {
            const int rz = first_node<MAS>(d.z);
            foot_weights<MAS>(d.z, rz, Cz);
            bz = wrap_node(rz, dims);
        }
        float w = (W ? W[i] : 1.0f) * weight_factor;
        const bool weighted = w != 1.0f;
#pragma unroll
        for (int l = 0; l < F; ++l) {
            const int nx = node_add(bx, l, dims);
#pragma unroll
            for (int m = 0; m < F; ++m) {
                const int ny = node_add(by, m, dims);
                float cxy = Cx[l] * Cy[m];
                if (AXES == 3) {
#pragma unroll
                    for (int k = 0; k < F; ++k) {
                        const int nz = node_add(bz, k, dims);
                        float c = cxy * Cz[k];
                        if (weighted) c = c * w;
                        unsafeAtomicAdd(&grid[((long long)nx * dims + ny) * dims + nz], c);
                    }
                } else {
                    float c = cxy;
                    if (weighted) c = c * w;
                    unsafeAtomicAdd(&grid[(long long)nx * dims + ny], c);
                }
            }
        }
    }
}

// CIC interpolation (gather), MAS_library.pyx:558-599: one thread per position,
// the 8 products summed left to right in the reference's order.
__global__ void __launch_bounds__(kBlock)
cic_interp_kernel(const float *__restrict__ density, int dims, float inv,
                  const float *__restrict__ pos, long long n, float *__restrict__ den) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float Cx[4], Cy[4], Cz[4];
        const int bx = axis_foot<PYL_CIC>(pos[3 * i + 0], inv, dims, Cx);
        const int by = axis_foot<PYL_CIC>(pos[3 * i + 1], inv, dims, Cy);
        const int bz = axis_foot<PYL_CIC>(pos[3 * i + 2], inv, dims, Cz);
        const long long x0 = bx, x1 = node_add(bx, 1, dims);
        const long long y0 = by, y1 = node_add(by, 1, dims);
        const long long z0 = bz, z1 = node_add(bz, 1, dims);
        const long long D = dims;
        auto G = [&](long long a, long long b, long long c) { return density[(a * D + b) * D + c]; };
        float s = G(x0, y0, z0) * Cx[0] * Cy[0] * Cz[0];
        s = s + G(x0, y0, z1) * Cx[0] * Cy[0] * Cz[1];
        s = s + G(x0, y1, z0) * Cx[0] * Cy[1] * Cz[0];
        s = s + G(x0, y1, z1) * Cx[0] * Cy[1] * Cz[1];
        s = s + G(x1, y0, z0) * Cx[1] * Cy[0] * Cz[0];
        s = s + G(x1, y0, z1) * Cx[1] * Cy[0] * Cz[1];
        s = s + G(x1, y1, z0) * Cx[1] * Cy[1] * Cz[0];
        s = s + G(x1, y1, z1) * Cx[1] * Cy[1] * Cz[1];
        den[i] = s;
    }
}

__global__ void scale_grid_kernel(float *__restrict__ grid, long long n, float divisor) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        grid[i] = grid[i] / divisor;
}

// grid = grid / scale - shift (the output transform of the atomic form).
__global__ void delta_grid_kernel(float *__restrict__ grid, long long n, float scale, float shift) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        grid[i] = grid[i] / scale - shift;
}

int grid_blocks(long long n, int block) {
    long long b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

// Clears nwords 4-byte words at p (a kernel rather than hipMemsetAsync, so that
// a caller may capture the call in a HIP graph); 16-byte stores when p allows.
template <bool V16>
__global__ void zero_words_kernel(unsigned *__restrict__ p, long long nwords) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (V16) {
        const long long n16 = nwords >> 2;
        for (long long i = t; i < n16; i += stride) ((uint4 *)p)[i] = make_uint4(0u, 0u, 0u, 0u);
        if (t < (nwords & 3)) p[n16 * 4 + t] = 0u;
    } else {
        for (long long i = t; i < nwords; i += stride) p[i] = 0u;
    }
}

// The 2D renormalisation's division of the existing grid (OutMode::prescale).
int prescale_grid(Ctx *c, float *grid, long long ncell, const OutMode &om) {
    if (om.prescale <= 0.0f) return PYL_OK;
    scale_grid_kernel<<<grid_blocks(ncell, kBlock), kBlock, 0, c->stream>>>(grid, ncell, om.prescale);
    PYL_HIP(hipGetLastError());
    return PYL_OK;
}

int zero_bytes(Ctx *c, void *p, size_t bytes) {
    const long long nwords = (long long)(bytes / 4);      // every region here is whole words
    if (nwords == 0) return PYL_OK;
    if (((uintptr_t)p & 15) == 0)
        zero_words_kernel<true><<<grid_blocks((nwords + 3) / 4, kBlock), kBlock, 0, c->stream>>>(
            (unsigned *)p, nwords);
    else
        zero_words_kernel<false><<<grid_blocks(nwords, kBlock), kBlock, 0, c->stream>>>(
            (unsigned *)p, nwords);
    PYL_HIP(hipGetLastError());
    return PYL_OK;
}

// Chunk table and tile origins of one realisation, written on the device (the
// fixed-scratch path has no pinned staging to upload them from).
__global__ void chunk_table_kernel(Chunk *__restrict__ ch, long long n, long long target, int nchunk,
                                   int *__restrict__ orig, int nt, int dims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nchunk) {
        const long long s = (long long)i * target;
        ch[i] = Chunk{s, s + target < n ? s + target : n, 0, 0, 0, 0, 0, 0};
    }
    if (i <= nt) orig[i] = tile_origin(i, nt, dims);
}

// Chunk table of a batch whose particle ranges are in device memory: realisation r
// (particles offs[r] .. offs[r+1]) is cut into cpr chunks of near-equal length
// (empty ones when it is short), so the chunk count does not depend on the counts.
__global__ void chunk_table_dev_kernel(Chunk *__restrict__ ch, const long long *__restrict__ offs,
                                       int nreal, int cpr, int *__restrict__ orig, int nt, int dims) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nreal * cpr) {
        const int r = i / cpr, j = i - r * cpr;
        const long long b = offs[r], len = offs[r + 1] - b;
        ch[i] = Chunk{b + len * j / cpr, b + len * (j + 1) / cpr, r, 0, 0, 0, 0, 0};
    }
    if (i <= nt) orig[i] = tile_origin(i, nt, dims);
}

// Scratch piece `id` of a call: the context's grow-only buffer, or, while
// pyl_ma_noalloc runs (c->arena set), the next 256-byte-aligned piece of its
// fixed device array.
int take(Ctx *c, BufId id, size_t bytes, void **out) {
    if (c->arena) {
        const size_t at = (c->arena_used + 255) & ~(size_t)255;
        PYL_CHECK(at + bytes <= c->arena_cap, PYL_EBUDGET,
                  "pyl_ma_noalloc: the call needs more than the %zu bytes of fixed scratch "
                  "(%zu bytes at least)", c->arena_cap, at + bytes);
        *out = c->arena + at;
        c->arena_used = at + bytes;
        return PYL_OK;
    }
    PYL_TRY(ensure(c, id, bytes));
    *out = c->bufs[id].ptr;
    return PYL_OK;
}

TileGeom make_geom(int dims, int axes, int mas) {
    TileGeom g{};
    const int T = axes == 3 ? kT3 : kT2;
    g.dims = dims;
    g.nt = dims <= T ? 1 : (dims + T - 1) / T;
    g.ntz = axes == 3 ? g.nt : 1;
    g.shift = -1;
    if (g.nt == 1) g.shift = 30;                       // every node in tile 0
    else if (dims % T == 0) g.shift = (axes == 3) ? 4 : 6;
    g.ntiles = (long long)g.nt * g.nt * g.ntz;
    g.ncell = (long long)dims * dims * (axes == 3 ? dims : 1);
    const int emax = (dims + g.nt - 1) / g.nt;
    const int ez = axes == 3 ? emax : 1;
    int ps = 0;
    switch (mas) {
        case PYL_NGP: ps = axes == 3 ? halo_count<1, 3>(emax, emax, ez) : halo_count<1, 2>(emax, emax, ez); break;
        case PYL_CIC: ps = axes == 3 ? halo_count<2, 3>(emax, emax, ez) : halo_count<2, 2>(emax, emax, ez); break;
        case PYL_TSC: ps = axes == 3 ? halo_count<3, 3>(emax, emax, ez) : halo_count<3, 2>(emax, emax, ez); break;
        default: ps = axes == 3 ? halo_count<4, 3>(emax, emax, ez) : halo_count<4, 2>(emax, emax, ez); break;
    }
    g.pslot = ps > 0 ? ps : 1;
    // sort window: every key when they fit, else the box of 8 x 8 tile columns
    // (all z) when it fits and the grid has at least 8 tiles per axis
    if (g.ntiles <= kWin) g.wmode = kWinDirect;
    else if (g.nt >= kBoxW && (long long)kBoxW * kBoxW * g.ntz <= kWin) g.wmode = kWinBox;
    else g.wmode = kWinLinear;
    return g;
}

// The latest order finding per (device, dims, MAS) over every context of the
// process: a context that has not sorted such a grid yet starts from it (requests
// from one source -- a server's slots, a pipeline's threads -- tend to share their
// particle order).
std::mutex g_hint_mu;
std::map<std::tuple<int, int, int>, bool> g_shuffled_hint;

template <int MAS>
int run_tiled(Ctx *c, const float *pos, const RSD &rsd, long long n, const long long *h_offs,
              int nreal, int axes, const float *W, float *grid, int dims, float inv,
              float weight_factor, const OutMode &om) {
    TileGeom g = make_geom(dims, axes, MAS);
    const bool pow2 = g.shift >= 0;
    const long long nt = g.ntiles * nreal;      // keys over the whole batch
    const int nscan_blocks = (int)((nt + kScanTile - 1) / kScanTile);
    PYL_CHECK(nscan_blocks <= kScanTile, PYL_EINVAL, "too many tiles (%lld)", nt);
    PYL_CHECK(nt < (1ll << 31), PYL_EINVAL, "too many tiles (%lld)", nt);
    // Sort forms for particles out of order.  Coherent input (a lattice, a snapshot
    // in ID order, a sorted catalogue) keeps each sort chunk inside a few tiles, so
    // the one-pass sort writes its records in runs.  Particles in random order
    // spread each chunk over the whole grid:
    //  * the wide window (3D realisations of kWin < ntiles <= kWinWide tiles) counts
    //    every key in LDS with chunks of 2-16 particles per tile, so no particle
    //    takes a global atomic, but each record is still written alone;
    //  * the two-level sort (one 3D grid of at least 16 tiles per axis) moves the
    //    particles into buckets of kBkt^2 x kBktZ tiles first and then sorts each bucket, both
    //    passes writing in runs.
    // The count pass tallies the particles its chunks' 8 x 8 tile-column boxes miss
    // (read back without waiting: until the copy has landed the previous choice
    // stands); when more than a quarter missed, the next call with the same (dims,
    // MAS) on this context takes the two-level sort.  Every form gives the same
    // bits (integer tile sums).
    // particle ranges in device memory (pyl_ma_batch_dev): chunks cut on the device
    const bool devt = c->dev_offs != nullptr;
    // (the readback needs the host between calls: not with device ranges or under
    // graph capture; the tally is a 32-bit counter: not for 2^32 particles or more)
    const bool tally = axes == 3 && g.nt >= 16 && !devt && !c->capturing && n < (1ll << 32);
    const bool wide_ok = kUseWide && axes == 3 && g.ntiles > kWin && g.ntiles <= kWinWide;
    const int nbt = bucket_counts_xy(g.nt);
    const int nbk = nbt * nbt * bucket_counts_z(g.ntz);
    const bool two_ok = axes == 3 && g.nt >= 16 && nreal == 1 && !devt && !c->capturing &&
                        !c->arena && nbk <= kMaxBuckets && n > 0 && n < (1ll << 32);
    if ((c->box_miss_n > 0 || c->cap_spill_n > 0) && !c->capturing) {
        const hipError_t q = hipEventQuery(c->ev_total);
        if (q == hipSuccess) {
            if (c->box_miss_n > 0) {
                const bool shuffled = c->h_total[1] * 4 > (unsigned long long)c->box_miss_n;
                c->shuffled_hint[c->box_miss_key] = shuffled;
                std::lock_guard<std::mutex> lk(g_hint_mu);
                g_shuffled_hint[std::make_tuple(c->device, c->box_miss_key.first, c->box_miss_key.second)] =
                    shuffled;
            }
            // a capacity sort that spilled more than 1/64 of its particles: the next
            // kCapRetry calls with this (dims, MAS) that would take it sort without it
            if (c->cap_spill_n > 0 && c->h_total[3] * kCapSpillDen > (unsigned long long)c->cap_spill_n)
                c->cap_bad[c->cap_key] = kCapRetry;
            c->box_miss_n = -1;
            c->cap_spill_n = -1;
        } else if (q == hipErrorNotReady) {
            (void)hipGetLastError();          // pending, not an error: keep it out of later checks
        } else {
            PYL_HIP(q);
        }
    }
    const std::pair<int, int> hkey(dims, MAS);
    // the context's own finding for this (dims, MAS), else the process's latest
    bool shuffled = false;
    if (tally) {
        const auto it = c->shuffled_hint.find(hkey);
        if (it != c->shuffled_hint.end()) {
            shuffled = it->second;
        } else {
            std::lock_guard<std::mutex> lk(g_hint_mu);
            const auto gt = g_shuffled_hint.find(std::make_tuple(c->device, dims, MAS));
            shuffled = gt != g_shuffled_hint.end() && gt->second;
        }
    }
    // Few tiles with many records each (a small grid and many particles): each
    // tile's records are split over nsplit workgroups, which add their boxes into
    // per-tile int64 boxes, and a second pass writes the grid.
    int nsplit = 1;
    if (nt < kSplitMaxKeys && n > kSplitRecords * nt)
        nsplit = (int)std::min<long long>(kSplitMax, (n + kSplitRecords * nt - 1) / (kSplitRecords * nt));
    // The capacity sort (cap_scatter_kernel): coherent particles of one 3D grid whose
    // sort chunks have every tile or the 8 x 8 tile-column box in their window;
    // adaptively when the tiles hold kCapMinMean particles or more on average and no
    // earlier capacity call of this context on such a grid spilled too much.  Its
    // slots take about kCapFactor times the records' bytes.
    const bool cap_ok = axes == 3 && nreal == 1 && !devt && !c->arena && !c->capturing && nsplit == 1 &&
                        (g.wmode == kWinDirect || g.wmode == kWinBox) && n > 0 && n < (1ll << 32);
    const size_t rec_bytes = W ? sizeof(float4) : 3 * sizeof(float);
    // (the capacity form also serves the two-level sort's second pass: its chunks lie
    // in one bucket, so only full tiles spill)
    const bool dense = n >= kCapMinMean * g.ntiles;
    const auto cap_good = [&]() {                  // counts the calls it turns down
        const auto it = c->cap_bad.find(hkey);
        if (it == c->cap_bad.end() || it->second <= 0) return true;
        --it->second;
        return false;
    };
    bool wide = false, two = false, capm = false;
    if (c->tune_window == 1) wide = wide_ok;
    else if (c->tune_window == 2) two = two_ok;
    else if (c->tune_window == 3) capm = cap_ok;
    else if (c->tune_window == 4) capm = two = two_ok;
    else if (c->tune_window < 0 && tally && shuffled) {
        two = two_ok;
        wide = !two && wide_ok;          // batches of host ranges, fixed scratch
        capm = two && dense && cap_good();
    } else if (c->tune_window < 0 && tally && cap_ok && dense) {
        capm = cap_good();
    }
    // tile capacity: kCapFactor times the mean, in whole 64-record runs
    const unsigned long long cap64 =
        capm ? ((unsigned long long)(kCapFactor * n + g.ntiles - 1) / g.ntiles + 63) / 64 * 64 : 0;
    void *cv = nullptr, *sv = nullptr;
    // the spill list: n (dist_x, dist_y, dist_z) | n weights, the weights 256-byte
    // aligned (in the bucketed-record buffer, or a buffer of its own when that holds
    // the two-level sort's first pass)
    const size_t pxyz_bytes = ((size_t)n * 3 * sizeof(float) + 255) & ~(size_t)255;
    const size_t spill_bytes = pxyz_bytes + (W ? (size_t)n * sizeof(float) : 0);
    if (capm) {
        const size_t bytes = (size_t)g.ntiles * cap64 * rec_bytes;
        const size_t extra = two ? spill_bytes - std::min(spill_bytes, (size_t)c->bufs[BUF_SPILL].cap) : 0;
        // within the context's budget, else the form without it
        capm = cap64 < (1ull << 31) &&
               held_bytes(c) - c->bufs[BUF_CAPREC].cap + bytes + (bytes >> 4) + extra <= c->budget &&
               take(c, BUF_CAPREC, bytes, &cv) == PYL_OK &&
               (!two || take(c, BUF_SPILL, spill_bytes, &sv) == PYL_OK);
    }
    const unsigned cap = (unsigned)cap64;
    if (wide) g.wmode = kWinDirect;
    // the spill sort of the capacity form: every tile in LDS when they fit
    TileGeom gs = g;
    int sws = kWin;
    if (g.ntiles <= kWin) gs.wmode = kWinDirect;
    else if (kUseWide && g.ntiles <= kWinWide) {
        gs.wmode = kWinDirect;
        sws = kWinWide;
    }
    // kWinDirect/Box/Linear == PYL_WINDOW_DIRECT/BOX/LINEAR
    c->last_window = (capm && two) ? PYL_WINDOW_TWOLEVEL_CAPACITY
                   : capm ? PYL_WINDOW_CAPACITY : two ? PYL_WINDOW_TWOLEVEL : wide ? PYL_WINDOW_WIDE : g.wmode;
    const int ws = capm ? sws : two ? kBktSlots : wide ? kWinWide : kWin;
    // Chunks of consecutive particles, never straddling two realisations.
    const long long target = wide ? std::max(2 * g.ntiles, std::min(16 * g.ntiles, (n + 2047) / 2048))
                                  : std::max(8192LL, std::min(65536LL, (n + 4095) / 4096));
    long long nchunk_ll = 0;
    if (devt) nchunk_ll = (long long)nreal * c->chunks_per_real;
    else
        for (int r = 0; r < nreal; ++r) {
            const long long b = h_offs ? h_offs[r] : 0, e = h_offs ? h_offs[r + 1] : n;
            if (e > b) nchunk_ll += (e - b + target - 1) / target;
        }
    PYL_CHECK(nchunk_ll < (1 << 30), PYL_EINVAL, "too many particle chunks");
    const int nchunk = (int)nchunk_ll;
    // two-level: pass-2 chunks of up to target2 records, none straddling two buckets
    const long long target2 = 32768;
    const long long nchunk2_ll = two ? (n + target2 - 1) / target2 + nbk : 0;
    PYL_CHECK(nchunk2_ll < (1 << 30), PYL_EINVAL, "too many particle chunks");
    const int nchunk2 = (int)nchunk2_ll;
    // chunks of the tile sort (pass 2 of the two-level sort, the capacity form's spill
    // sort, or the only pass)
    const int nsort = capm ? kSpillChunks : two ? nchunk2 : nchunk;
    // tile buffer: counts[nt] | ocounts[nt] | ocursors[nt] | wmax[nreal] | box_miss |
    //              (two-level) bucket counts[nbk] | pad | offsets[nt+1] | block sums |
    //              chunk table | tile origins | tile order | per-chunk window bases |
    //              (two-level) bucket offsets[nbk+1] | pass-2 chunk table | bucket bases
    //              (capacity) tile cursors[nt] | pad | spill range {0, count}, placed, total (u64)
    //              (ordered) order bucket sizes[512] | bucket cursors[512]
    // Largest tiles first, when there are enough tiles, and records per tile, for the
    // dispatch tail to outweigh the ordering passes.
    const bool ordered = c->tune_order >= 0
        ? (nsplit == 1 && c->tune_order == 1)
        : (nsplit == 1 && nt >= kOrderMinKeys && n >= kOrderMinRecords * nt);
    size_t nwords = (size_t)nt * 3 + (size_t)nreal + 1 + (two ? (size_t)nbk : 0);
    const size_t cur_at = nwords;
    if (capm) nwords += (size_t)nt;
    const size_t capw_at = (nwords + 1) & ~(size_t)1;     // 8-byte aligned
    if (capm) nwords = capw_at + 8;
    const size_t ohist_at = nwords;                        // (ordered) bucket sizes | cursors
    if (ordered) nwords += 2 * (size_t)kOrderBuckets;
    const size_t cnt_bytes = (nwords * sizeof(unsigned) + 255) & ~(size_t)255;
    const size_t off_bytes = (((size_t)nt + 1) * sizeof(unsigned long long) + 255) & ~(size_t)255;
    const size_t bs_bytes = ((size_t)nscan_blocks * sizeof(unsigned long long) + 255) & ~(size_t)255;
    const size_t ch_bytes = ((size_t)std::max(nchunk, 1) * sizeof(Chunk) + 255) & ~(size_t)255;
    const size_t org_bytes = ((size_t)(g.nt + 1) * sizeof(int) + 255) & ~(size_t)255;
    const size_t ord_bytes = ((size_t)nt * sizeof(unsigned) + 255) & ~(size_t)255;
    const size_t dump_bytes = ((size_t)std::max(nsort, 1) * ws * sizeof(unsigned) + 255) & ~(size_t)255;
    const size_t boff_bytes = two ? (((size_t)nbk + 1) * sizeof(unsigned long long) + 255) & ~(size_t)255 : 0;
    const size_t ch2_bytes = two ? ((size_t)nchunk2 * sizeof(Chunk) + 255) & ~(size_t)255 : 0;
    const size_t chs_bytes = capm ? ((size_t)kSpillChunks * sizeof(Chunk) + 255) & ~(size_t)255 : 0;
    const size_t bdump_bytes = two ? ((size_t)std::max(nchunk, 1) * nbk * sizeof(unsigned) + 255) & ~(size_t)255 : 0;
    const int T = axes == 3 ? kT3 : kT2;
    const size_t nbox = (size_t)(T + MAS - 1) * (T + MAS - 1) * (axes == 3 ? T + MAS - 1 : 1);
    // all scratch of the call is taken before the first launch
    void *tbv = nullptr, *pv = nullptr, *rv = nullptr, *bv = nullptr, *kv = nullptr;
    PYL_TRY(take(c, BUF_TILE_CNT, cnt_bytes + off_bytes + bs_bytes + ch_bytes + org_bytes +
                                      ord_bytes + dump_bytes + boff_bytes + ch2_bytes + 2 * bdump_bytes +
                                      chs_bytes, &tbv));
    PYL_TRY(take(c, BUF_PARTIAL, (size_t)nt * g.pslot * sizeof(float), &pv));
    PYL_TRY(take(c, BUF_RECORDS, (size_t)(n ? n : 1) * rec_bytes, &rv));
    // bucketed records (or the one-pass capacity form's spill list): n (dist_x, dist_y,
    // dist_z) | n weights, the weights 256-byte aligned
    const size_t bxyz_bytes = pxyz_bytes;
    if (two || capm) PYL_TRY(take(c, BUF_BUCKETED, bxyz_bytes + (W ? (size_t)n * sizeof(float) : 0), &kv));
    if (nsplit > 1)
        PYL_TRY(take(c, BUF_TILE_BOX, (size_t)nt * nbox * sizeof(unsigned long long), &bv));
    char *tb = (char *)tbv;
    unsigned *counts = (unsigned *)tb;
    unsigned *ocounts = counts + nt;
    unsigned *ocursors = ocounts + nt;
    unsigned *wmax = ocursors + nt;
    unsigned *box_miss = wmax + nreal;
    unsigned *bcounts = box_miss + 1;
    unsigned long long *offsets = (unsigned long long *)(tb + cnt_bytes);
    unsigned long long *bsums = (unsigned long long *)(tb + cnt_bytes + off_bytes);
    Chunk *dchunks = (Chunk *)(tb + cnt_bytes + off_bytes + bs_bytes);
    int *dorig = (int *)(tb + cnt_bytes + off_bytes + bs_bytes + ch_bytes);
    unsigned *order = (unsigned *)(tb + cnt_bytes + off_bytes + bs_bytes + ch_bytes + org_bytes);
    unsigned *dump = (unsigned *)(tb + cnt_bytes + off_bytes + bs_bytes + ch_bytes + org_bytes + ord_bytes);
    char *tb2 = (char *)dump + dump_bytes;
    unsigned long long *boffsets = (unsigned long long *)tb2;
    Chunk *dchunks2 = (Chunk *)(tb2 + boff_bytes);
    unsigned *bdump = (unsigned *)(tb2 + boff_bytes + ch2_bytes);
    unsigned *bmat = (unsigned *)(tb2 + boff_bytes + ch2_bytes + bdump_bytes);
    Chunk *dchunksS = (Chunk *)(tb2 + boff_bytes + ch2_bytes + 2 * bdump_bytes);   // spill chunks
    float *bxyz = (float *)kv;
    float *bw = ((two || capm) && W) ? (float *)((char *)kv + bxyz_bytes) : nullptr;
    unsigned *cursor = counts + cur_at;
    unsigned *ohist = counts + ohist_at;
    unsigned long long *capw = (unsigned long long *)(counts + capw_at);
    float *crec = (float *)cv;
    float *pxyz = two ? (float *)sv : bxyz;
    float *pw = W ? (two ? (float *)((char *)sv + pxyz_bytes) : bw) : nullptr;
    const int nsrc = two ? nchunk2 : nchunk;          // chunks the capacity scatter reads
    g.orig = dorig;
    float *partials = (float *)pv;
    float *records = (float *)rv;
    unsigned long long *gbox = (unsigned long long *)bv;

    PYL_TRY(prescale_grid(c, grid, g.ncell * nreal, om));
    // the whole 256-byte-rounded region: one aligned fill instead of a body and a tail
    PYL_TRY(zero_bytes(c, counts, cnt_bytes));
    // Chunk table and tile origins: the kernels only read their particle ranges and
    // origins (the count pass rewrites each chunk's window fields), so a call with
    // the same ranges and grid in the same allocation reuses the uploaded tables.
    // The fixed-scratch path writes them on the device every call.
    Ctx::TileTables &tt = c->tile_tables;
    if (devt) {
        const int nfill = std::max(nchunk, g.nt + 1);
        chunk_table_dev_kernel<<<ceil_div(nfill, kBlock), kBlock, 0, c->stream>>>(
            dchunks, c->dev_offs, nreal, c->chunks_per_real, dorig, g.nt, dims);
        PYL_HIP(hipGetLastError());
    } else if (c->arena) {
        PYL_CHECK(nreal == 1, PYL_EINVAL, "fixed-scratch MA takes one grid");
        const int nfill = std::max(nchunk, g.nt + 1);
        chunk_table_kernel<<<ceil_div(nfill, kBlock), kBlock, 0, c->stream>>>(dchunks, n, target, nchunk,
                                                                           dorig, g.nt, dims);
        PYL_HIP(hipGetLastError());
    }
    bool same = !c->arena && !devt && tt.gen == c->bufs[BUF_TILE_CNT].gen && tt.at == (const void *)dchunks &&
                tt.n == n && tt.nchunk == nchunk &&
                tt.target == target &&
                tt.nreal == nreal && tt.dims == dims && tt.axes == axes && tt.mas == MAS &&
                tt.offs.size() == (h_offs ? (size_t)nreal + 1 : 0);
    for (size_t i = 0; same && i < tt.offs.size(); ++i) same = tt.offs[i] == h_offs[i];
    if (!same && !c->arena && !devt) {
        tt.gen = 0;                               // valid again once both are queued
        static thread_local std::vector<Chunk> chunks;   // reused: no fresh pages per call
        chunks.clear();
        for (int r = 0; r < nreal; ++r) {
            const long long b = h_offs ? h_offs[r] : 0, e = h_offs ? h_offs[r + 1] : n;
            for (long long s = b; s < e; s += target) chunks.push_back({s, std::min(e, s + target), r, 0});
        }
        static thread_local std::vector<int> orig;
        orig.assign(g.nt + 1, 0);
        for (int i = 0; i <= g.nt; ++i) orig[i] = tile_origin(i, g.nt, dims);
        // through pinned staging (asynchronous DMA)
        void *stage = nullptr;
        const size_t ch_up = (size_t)nchunk * sizeof(Chunk), org_up = orig.size() * sizeof(int);
        PYL_TRY(pinned(c, c->pin_up, ch_up + org_up, &stage));
        if (ch_up) std::memcpy(stage, chunks.data(), ch_up);
        std::memcpy((char *)stage + ch_up, orig.data(), org_up);
        PYL_HIP(hipMemcpyAsync(dorig, (char *)stage + ch_up, org_up, hipMemcpyHostToDevice, c->stream));
        if (nchunk > 0)
            PYL_HIP(hipMemcpyAsync(dchunks, stage, ch_up, hipMemcpyHostToDevice, c->stream));
        PYL_TRY(pinned_done(c, c->pin_up));
        tt.n = n;
        tt.nchunk = nchunk;
        tt.target = target;
        tt.nreal = nreal;
        tt.dims = dims;
        tt.axes = axes;
        tt.mas = MAS;
        tt.offs.assign(h_offs ? h_offs : nullptr, h_offs ? h_offs + nreal + 1 : nullptr);
        tt.at = dchunks;
        tt.gen = c->bufs[BUF_TILE_CNT].gen;
    }
    // inputs of the tile sort: the particles, or the bucketed records of pass 1
    const float *spos = pos, *sW = W;
    RSD srsd = rsd;
    float sinv = inv;
    Chunk *schunks = dchunks;
    if (two) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_BUCKET));
        auto kb = pow2 ? bucket_count_kernel<MAS, true> : bucket_count_kernel<MAS, false>;
        kb<<<nchunk, kSortBlock, 0, c->stream>>>(pos, rsd, inv, g, nbt, dchunks, bmat,
                                                 tally ? box_miss : nullptr);
        bucket_scan_kernel<<<nbk, kScanBlock, 0, c->stream>>>(bmat, nchunk, nbk, bdump, bcounts);
        bucket_plan_kernel<<<1, kScanBlock, 0, c->stream>>>(bcounts, nbk, target2, boffsets, dchunks2,
                                                            nchunk2);
        // (bucket tables of kBktSlots entries when they suffice: less LDS, more workgroups per CU)
        auto kbs = nbk <= kBktSlots
            ? (W ? (pow2 ? staged_scatter_kernel<MAS, true, true, 0, kBktSlots, kStage1>
                         : staged_scatter_kernel<MAS, false, true, 0, kBktSlots, kStage1>)
                 : (pow2 ? staged_scatter_kernel<MAS, true, false, 0, kBktSlots, kStage1>
                         : staged_scatter_kernel<MAS, false, false, 0, kBktSlots, kStage1>))
            : (W ? (pow2 ? staged_scatter_kernel<MAS, true, true, 0, kMaxBuckets, kStage1>
                         : staged_scatter_kernel<MAS, false, true, 0, kMaxBuckets, kStage1>)
                 : (pow2 ? staged_scatter_kernel<MAS, true, false, 0, kMaxBuckets, kStage1>
                         : staged_scatter_kernel<MAS, false, false, 0, kMaxBuckets, kStage1>));
        kbs<<<nchunk, kSortBlock, 0, c->stream>>>(pos, rsd, W, inv, g, nbt, dchunks, bdump, boffsets,
                                                  counts, ocursors, bxyz, bw, nullptr,
                                                  (unsigned long long)n);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_BUCKET));
        spos = bxyz;
        sW = bw;
        srsd = RSD{bxyz, 0.0f, 0.0f, 0, 3, 0};
        sinv = 1.0f;
        schunks = dchunks2;
        g.wmode = kWinBlock;
    }
    if (capm) {
        // one read of the particles (or, after the two-level sort's first pass, of the
        // bucketed records) into the tile slots; the rest into the spill list
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCATTER));
        auto kcs = W ? (pow2 ? cap_scatter_kernel<MAS, true, true> : cap_scatter_kernel<MAS, false, true>)
                     : (pow2 ? cap_scatter_kernel<MAS, true, false> : cap_scatter_kernel<MAS, false, false>);
        const int slots = g.wmode == kWinBlock ? kBktSlots
                        : g.wmode == kWinBox   ? kBoxW * kBoxW * g.ntz : (int)g.ntiles;
        const size_t lds = 2 * (size_t)slots * sizeof(unsigned);
        if (lds > 65536)
            PYL_HIP(hipFuncSetAttribute((const void *)kcs, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (two) {
            // records of one bucket per chunk, in random tile order: staged writes
            auto kss = W ? (pow2 ? staged_cap_scatter_kernel<MAS, true, true, kStage2>
                                 : staged_cap_scatter_kernel<MAS, false, true, kStage2>)
                         : (pow2 ? staged_cap_scatter_kernel<MAS, true, false, kStage2>
                                 : staged_cap_scatter_kernel<MAS, false, false, kStage2>);
            if (nsrc > 0)
                kss<<<nsrc, kSortBlock, 0, c->stream>>>(bxyz, bw, g, schunks, cursor, cap, crec, pxyz, pw,
                                                        capw + 1, capw + 2, wmax);
        } else if (nsrc > 0) {
            kcs<<<nsrc, kSortBlock, lds, c->stream>>>(spos, srsd, sW, sinv, g, schunks, cursor, cap, crec, pxyz,
                                                      pw, capw + 1, capw + 2, wmax, tally ? box_miss : nullptr);
        }
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SCATTER));
        // the spill list (capw[0..1] = {0, count}) ordered by tile: up to kSpillAtomic
        // particles by the atomic kernels (the windowed passes then see empty chunks),
        // more by the windowed count and scatter in kSpillChunks chunks cut on the device
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SPILL));
        static_assert((long long)kSpillBlocks * kBlock >= kSpillChunks, "spill chunk table");
        PYL_CHECK((long long)g.nt + 1 <= (long long)kSpillBlocks * kBlock, PYL_EINVAL, "grid too large");
        (pow2 ? spill_count_kernel<MAS, true> : spill_count_kernel<MAS, false>)
            <<<kSpillBlocks, kBlock, 0, c->stream>>>(pxyz, capw + 1, gs, ocounts, dchunksS, kSpillChunks,
                                                     dorig);
        const RSD off{pxyz, 0.0f, 0.0f, 0, 3, 0};
        auto kc = sws == kWinWide ? (pow2 ? tile_count_kernel<MAS, 3, true, kWinWide>
                                          : tile_count_kernel<MAS, 3, false, kWinWide>)
                                  : (pow2 ? tile_count_kernel<MAS, 3, true, kWin>
                                          : tile_count_kernel<MAS, 3, false, kWin>);
        kc<<<kSpillChunks, kSortBlock, 0, c->stream>>>(pxyz, off, 1.0f, gs, dchunksS, counts, ocounts, dump,
                                                       nullptr);
        scan_partial_kernel<<<nscan_blocks, kScanBlock, 0, c->stream>>>(counts, ocounts, nt, bsums);
        scan_blocksums_kernel<<<1, kScanBlock, 0, c->stream>>>(bsums, nscan_blocks);
        scan_final_kernel<<<nscan_blocks, kScanBlock, 0, c->stream>>>(counts, ocounts, nt, bsums, offsets);
        auto ks = sws == kWinWide
            ? (W ? (pow2 ? tile_scatter_kernel<MAS, 3, true, true, kWinWide>
                         : tile_scatter_kernel<MAS, 3, false, true, kWinWide>)
                 : (pow2 ? tile_scatter_kernel<MAS, 3, true, false, kWinWide>
                         : tile_scatter_kernel<MAS, 3, false, false, kWinWide>))
            : (W ? (pow2 ? tile_scatter_kernel<MAS, 3, true, true, kWin>
                         : tile_scatter_kernel<MAS, 3, false, true, kWin>)
                 : (pow2 ? tile_scatter_kernel<MAS, 3, true, false, kWin>
                         : tile_scatter_kernel<MAS, 3, false, false, kWin>));
        ks<<<kSpillChunks, kSortBlock, 0, c->stream>>>(pxyz, off, pw, 1.0f, gs, dchunksS, dump, offsets, counts,
                                                       ocursors, wmax, records, (unsigned long long)n);
        auto kss2 = W ? (pow2 ? spill_scatter_kernel<MAS, true, true> : spill_scatter_kernel<MAS, false, true>)
                      : (pow2 ? spill_scatter_kernel<MAS, true, false> : spill_scatter_kernel<MAS, false, false>);
        kss2<<<kSpillBlocks, kBlock, 0, c->stream>>>(pxyz, pw, capw + 1, gs, offsets, ocursors, records,
                                                     (unsigned long long)n, capw + 2, capw + 3);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SPILL));
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCAN));
        if (ordered) {
            const CapRecs ocr{crec, cursor, cap, (long long)n};
            order_hist_kernel<<<order_wgs(nt), kOrderBlock, 0, c->stream>>>(offsets, nt, ocr, ohist);
            order_place_kernel<<<order_wgs(nt), kOrderBlock, 0, c->stream>>>(offsets, nt, ocr, ohist,
                                                                              ohist + kOrderBuckets, order);
        }
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SCAN));
    }
    if (nsort > 0 && !capm) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_COUNT));
        auto kc = two ? (pow2 ? tile_count_kernel<MAS, 3, true, kBktSlots>
                              : tile_count_kernel<MAS, 3, false, kBktSlots>)
                : wide ? (pow2 ? tile_count_kernel<MAS, 3, true, kWinWide>
                               : tile_count_kernel<MAS, 3, false, kWinWide>)
                : (axes == 3) ? (pow2 ? tile_count_kernel<MAS, 3, true, kWin>
                                      : tile_count_kernel<MAS, 3, false, kWin>)
                              : (pow2 ? tile_count_kernel<MAS, 2, true, kWin>
                                      : tile_count_kernel<MAS, 2, false, kWin>);
        kc<<<nsort, kSortBlock, 0, c->stream>>>(spos, srsd, sinv, g, schunks, counts, ocounts, dump,
                                                (tally && !two) ? box_miss : nullptr);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_COUNT));
    }
    if (!capm) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCAN));
        scan_partial_kernel<<<nscan_blocks, kScanBlock, 0, c->stream>>>(counts, ocounts, nt, bsums);
        scan_blocksums_kernel<<<1, kScanBlock, 0, c->stream>>>(bsums, nscan_blocks);
        scan_final_kernel<<<nscan_blocks, kScanBlock, 0, c->stream>>>(counts, ocounts, nt, bsums,
                                                                       offsets);
        if (ordered) {
            const CapRecs ocr{nullptr, nullptr, 0, -1};
            order_hist_kernel<<<order_wgs(nt), kOrderBlock, 0, c->stream>>>(offsets, nt, ocr, ohist);
            order_place_kernel<<<order_wgs(nt), kOrderBlock, 0, c->stream>>>(offsets, nt, ocr, ohist,
                                                                              ohist + kOrderBuckets, order);
        }
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SCAN));
    }
    // offsets[nt] (capacity form: the sorted spill plus the placed records) must
    // equal n; it is read back asynchronously and checked at the next synchronising
    // call on this context.  Until then the scatter and the deposit keep every
    // record index below n (capacity slots: below the tile's capacity), so a bad
    // count cannot write or read outside the record buffers.
    const unsigned long long *total_word = capm ? capw + 3 : offsets + nt;
    PYL_TRY(check_total(c));                    // the previous call's check, if still pending
    if (c->capturing) {
        // each captured call reads its totals into pinned words of its own, compared
        // after every replay
        unsigned long long *ht = nullptr;
        PYL_HIP(hipHostMalloc((void **)&ht, 2 * sizeof(unsigned long long), hipHostMallocDefault));
        c->cap_host.push_back(ht);
        ht[1] = (unsigned long long)n;
        PYL_HIP(hipMemcpyAsync(ht, total_word, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               c->stream));
        if (devt)
            PYL_HIP(hipMemcpyAsync(ht + 1, c->dev_offs + nreal, sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, c->stream));
        c->deferred.push_back([ht]() {
            PYL_CHECK(ht[0] == ht[1], PYL_EHIP, "tile sort lost particles (%llu of %llu)", ht[0], ht[1]);
            return PYL_OK;
        });
    } else {
        PYL_HIP(hipMemcpyAsync(c->h_total, total_word, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, c->stream));
        if (devt)
            PYL_HIP(hipMemcpyAsync(c->h_total + 2, c->dev_offs + nreal, sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, c->stream));
        c->box_miss_n = -1;
        if (tally && nchunk > 0) {
            c->h_total[1] = 0;                 // the copy fills the low word
            PYL_HIP(hipMemcpyAsync(c->h_total + 1, box_miss, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   c->stream));
            c->box_miss_n = n;
            c->box_miss_key = hkey;
        }
        c->cap_spill_n = -1;
        if (capm) {
            PYL_HIP(hipMemcpyAsync(c->h_total + 3, capw + 1, sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, c->stream));
            c->cap_spill_n = n;
            c->cap_key = hkey;
        }
        PYL_HIP(hipEventRecord(c->ev_total, c->stream));
        c->total_expect = devt ? -2 : n;      // -2: against the device offsets' total
    }
    const unsigned long long rcap = (unsigned long long)n;
    if (capm) {
        // sorted above
    } else if (nsort > 0 && two) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCATTER));
        auto ks = W ? (pow2 ? staged_scatter_kernel<MAS, true, true, 1, kBktSlots, kStage2>
                            : staged_scatter_kernel<MAS, false, true, 1, kBktSlots, kStage2>)
                    : (pow2 ? staged_scatter_kernel<MAS, true, false, 1, kBktSlots, kStage2>
                            : staged_scatter_kernel<MAS, false, false, 1, kBktSlots, kStage2>);
        ks<<<nsort, kSortBlock, 0, c->stream>>>(spos, srsd, sW, sinv, g, nbt, schunks, dump, offsets,
                                                counts, ocursors, records, nullptr, wmax, rcap);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SCATTER));
    } else if (nsort > 0) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCATTER));
        auto ks = wide
            ? (W ? (pow2 ? tile_scatter_kernel<MAS, 3, true, true, kWinWide>
                         : tile_scatter_kernel<MAS, 3, false, true, kWinWide>)
                 : (pow2 ? tile_scatter_kernel<MAS, 3, true, false, kWinWide>
                         : tile_scatter_kernel<MAS, 3, false, false, kWinWide>))
            : W ? ((axes == 3) ? (pow2 ? tile_scatter_kernel<MAS, 3, true, true, kWin>
                                       : tile_scatter_kernel<MAS, 3, false, true, kWin>)
                               : (pow2 ? tile_scatter_kernel<MAS, 2, true, true, kWin>
                                       : tile_scatter_kernel<MAS, 2, false, true, kWin>))
                : ((axes == 3) ? (pow2 ? tile_scatter_kernel<MAS, 3, true, false, kWin>
                                       : tile_scatter_kernel<MAS, 3, false, false, kWin>)
                               : (pow2 ? tile_scatter_kernel<MAS, 2, true, false, kWin>
                                       : tile_scatter_kernel<MAS, 2, false, false, kWin>));
        ks<<<nsort, kSortBlock, 0, c->stream>>>(spos, srsd, sW, sinv, g, schunks, dump, offsets, counts,
                                                ocursors, wmax, records, rcap);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SCATTER));
    }
    const bool unit = W == nullptr && weight_factor == 1.0f;
    PYL_TRY(stage_begin(c, PYL_STAGE_MA_DEPOSIT));
    const bool uni = g.nt >= 2 && dims % (axes == 3 ? kT3 : kT2) == 0;
    auto kd = (axes == 3)
        ? (uni ? (unit ? tile_deposit_kernel<MAS, 3, false, false, true>
                       : (W ? tile_deposit_kernel<MAS, 3, true, true, true>
                            : tile_deposit_kernel<MAS, 3, true, false, true>))
               : (unit ? tile_deposit_kernel<MAS, 3, false, false, false>
                       : (W ? tile_deposit_kernel<MAS, 3, true, true, false>
                            : tile_deposit_kernel<MAS, 3, true, false, false>)))
        : (uni ? (unit ? tile_deposit_kernel<MAS, 2, false, false, true>
                       : (W ? tile_deposit_kernel<MAS, 2, true, true, true>
                            : tile_deposit_kernel<MAS, 2, true, false, true>))
               : (unit ? tile_deposit_kernel<MAS, 2, false, false, false>
                       : (W ? tile_deposit_kernel<MAS, 2, true, true, false>
                            : tile_deposit_kernel<MAS, 2, true, false, false>)));
    if (nsplit > 1) PYL_TRY(zero_bytes(c, gbox, (size_t)nt * nbox * sizeof(unsigned long long)));
    const CapRecs cr = capm ? CapRecs{crec, cursor, cap, (long long)n} : CapRecs{nullptr, nullptr, 0, -1};
    kd<<<(unsigned)(nt * nsplit), kDepBlock, 0, c->stream>>>(records, rcap, offsets, g, weight_factor,
                                                            W ? wmax : nullptr, grid, partials, om,
                                                            ordered ? order : nullptr, nsplit, gbox, cr);
    PYL_HIP(hipGetLastError());
    if (nsplit > 1) {
        auto kf = (axes == 3)
            ? (uni ? (unit ? tile_box_flush_kernel<MAS, 3, false, true> : tile_box_flush_kernel<MAS, 3, true, true>)
                   : (unit ? tile_box_flush_kernel<MAS, 3, false, false> : tile_box_flush_kernel<MAS, 3, true, false>))
            : (uni ? (unit ? tile_box_flush_kernel<MAS, 2, false, true> : tile_box_flush_kernel<MAS, 2, true, true>)
                   : (unit ? tile_box_flush_kernel<MAS, 2, false, false> : tile_box_flush_kernel<MAS, 2, true, false>));
        kf<<<(unsigned)nt, kDepBlock, 0, c->stream>>>(gbox, g, weight_factor, W ? wmax : nullptr, offsets,
                                                      grid, partials, om);
        PYL_HIP(hipGetLastError());
    }
    PYL_TRY(stage_end(c, PYL_STAGE_MA_DEPOSIT));
    if (MAS > 1) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_FINAL));
        if (axes == 3 && g.nt >= 2 && dims % kT3 == 0)
            tile_finalize_uniform_kernel<MAS><<<(unsigned)nt, kFinBlock, 0, c->stream>>>(
                offsets, g, partials, grid, om, cr);
        else if (axes == 3)
            tile_finalize_kernel<MAS, 3><<<(unsigned)nt, kBlock, 0, c->stream>>>(offsets, g, partials,
                                                                                grid, om, cr);
        else
            tile_finalize_kernel<MAS, 2><<<(unsigned)nt, kBlock, 0, c->stream>>>(offsets, g, partials,
                                                                                grid, om, cr);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_FINAL));
    }
    return PYL_OK;
}

template <int MAS>
int run_atomic(Ctx *c, const float *pos, const RSD &rsd, long long n, const long long *h_offs,
               int nreal, int axes, const float *W, float *grid, int dims, float inv,
               float weight_factor, const OutMode &om) {
    const long long ncell = (long long)dims * dims * (axes == 3 ? dims : 1);
    PYL_TRY(prescale_grid(c, grid, ncell * nreal, om));
    if (om.overwrite) PYL_TRY(zero_bytes(c, grid, (size_t)ncell * nreal * sizeof(float)));
    PYL_TRY(stage_begin(c, PYL_STAGE_MA_ATOMIC));
    for (int r = 0; r < nreal; ++r) {
        const long long b = h_offs ? h_offs[r] : 0, e = h_offs ? h_offs[r + 1] : n;
        if (e <= b) continue;
        const float *p = pos + (long long)axes * b;
        const float *w = W ? W + b : nullptr;
        RSD rr = rsd;
        rr.vel = rsd.vel + (long long)rsd.vstride * b;
        float *gr = grid + (long long)r * ncell;
        const int nb = grid_blocks(e - b, kBlock);
        if (axes == 3)
            atomic_deposit_kernel<MAS, 3><<<nb, kBlock, 0, c->stream>>>(p, rr, w, e - b, inv, dims,
                                                                       weight_factor, gr);
        else
            atomic_deposit_kernel<MAS, 2><<<nb, kBlock, 0, c->stream>>>(p, rr, w, e - b, inv, dims,
                                                                       weight_factor, gr);
        PYL_HIP(hipGetLastError());
    }
    if (om.overwrite && (om.scale != 1.0f || om.shift != 0.0f)) {
        delta_grid_kernel<<<grid_blocks(ncell * nreal, kBlock), kBlock, 0, c->stream>>>(
            grid, ncell * nreal, om.scale, om.shift);
        PYL_HIP(hipGetLastError());
    }
    PYL_TRY(stage_end(c, PYL_STAGE_MA_ATOMIC));
    return PYL_OK;
}

}  // namespace

// Device-pointer MA used by the C ABI entry points, single and batched.
// h_offs: host particle ranges of nreal grids (nreal+1 entries, NULL when nreal == 1);
// grid holds nreal contiguous grids.  With PYL_MA_OVERWRITE the grid receives
// deposit/scale - shift (scale = 1, shift = 0: the plain deposit).
int ma_device(Ctx *c, const float *pos, long long n, const long long *h_offs, int nreal,
              int axes, const float *W, float *grid, int dims, float BoxSize, int mas,
              int flags, float scale, float shift, const float *vel, float rsd_factor,
              int rsd_axis) {
    RSD rsd{pos, 0.0f, BoxSize, 0, axes, 0};
    if (vel) rsd = RSD{vel, rsd_factor, BoxSize, rsd_axis, 3, 1};
    // MAS_library.pyx:135 inv_cell_size = dims/BoxSize in float (MAS_c.c:19 computes it
    // in double and stores it to float, which rounds to the same value).
    const float inv = (float)dims / BoxSize;
    float weight_factor = 1.0f;
    const long long ncell = (long long)dims * dims * (axes == 3 ? dims : 1) * nreal;
    OutMode om{(flags & PYL_MA_OVERWRITE) ? 1 : 0, scale, shift, 0.0f};
    if (axes == 2 && (flags & PYL_MA_CYTHON_2D) && mas != PYL_NGP) {
        // MA() runs the 3D kernel on a (d,d,1) view, so each deposit lands F times
        // in the same cell; with renormalize_2D the whole grid is then divided by F
        // (MAS_library.pyx:84-110).  Equivalent: grid/F + deposit, or grid + F*deposit.
        // The division is launched by the algorithm once its scratch is in hand.
        if (flags & PYL_MA_RENORM_2D) {
            if (!om.overwrite) om.prescale = (float)mas;
        } else {
            weight_factor = (float)mas;
        }
    }
    const bool atomic = (flags & PYL_MA_ATOMIC) != 0;
    switch (mas) {
#define PYL_MA_CASE(M)                                                                      \
    case M:                                                                                 \
        return atomic ? run_atomic<M>(c, pos, rsd, n, h_offs, nreal, axes, W, grid, dims, inv, \
                                      weight_factor, om)                                    \
                      : run_tiled<M>(c, pos, rsd, n, h_offs, nreal, axes, W, grid, dims, inv,  \
                                     weight_factor, om);
        PYL_MA_CASE(PYL_NGP)
        PYL_MA_CASE(PYL_CIC)
        PYL_MA_CASE(PYL_TSC)
        PYL_MA_CASE(PYL_PCS)
#undef PYL_MA_CASE
    }
    set_error("unknown MAS id %d", mas);
    return PYL_EINVAL;
}

// ---- fixed scratch of pyl_ma_noalloc ------------------------------------------
// One array per device, resident from the library's load until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: only the
// address from hipGetSymbolAddress (kept per device below) is passed to kernels.
static __device__ __align__(256) unsigned char scratch[PYL_MA_NOALLOC_SCRATCH];

static std::mutex g_noalloc_mu;
static char *g_noalloc_ptr[64] = {};

int ma_noalloc_setup(Ctx *c) {
    PYL_CHECK(c->device >= 0 && c->device < 64, PYL_EINVAL, "device %d out of range", c->device);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[c->device]) {
        void *ptr = nullptr;
        PYL_HIP(hipSetDevice(c->device));
        PYL_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PYL_CHECK(ptr != nullptr, PYL_EHIP, "no device address for the fixed MA scratch");
        g_noalloc_ptr[c->device] = (char *)ptr;
    }
    c->noalloc_base = g_noalloc_ptr[c->device];
    return PYL_OK;
}

}  // namespace pyl

using namespace pyl;

static int ma_entry(pyl_ctx *ctx, const float *pos, int64_t n, const long long *offs, int nreal,
                    int axes, const float *W, float *grid, int dims, float BoxSize, int mas,
                    int flags, float scale = 1.0f, float shift = 0.0f,
                    const float *vel = nullptr, float rsd_factor = 0.0f, int rsd_axis = 0) {
    TraceRange trace_("pyl_ma");
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    PYL_CHECK(axes == 2 || axes == 3, PYL_EINVAL, "pos must have 2 or 3 columns (got %d)", axes);
    PYL_CHECK(dims >= 1, PYL_EINVAL, "dims must be positive (got %d)", dims);
    PYL_CHECK(mas >= PYL_NGP && mas <= PYL_PCS, PYL_EINVAL, "option not valid (MAS id %d)", mas);
    PYL_CHECK(n >= 0, PYL_EINVAL, "negative particle count");
    PYL_CHECK(grid != nullptr && (pos != nullptr || n == 0), PYL_EINVAL, "NULL array");
    PYL_CHECK(BoxSize > 0.0f, PYL_EINVAL, "BoxSize must be positive");
    PYL_CHECK(scale != 0.0f, PYL_EINVAL, "scale must be non-zero");
    const long long ncell = (long long)dims * dims * (axes == 3 ? dims : 1) * nreal;
    PYL_CHECK(ncell < (1ll << 34), PYL_EINVAL, "grid too large");
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    if (flags & PYL_PTR_DEVICE) {
        PYL_TRY(ma_device(c, pos, n, offs, nreal, axes, W, grid, dims, BoxSize, mas, flags, scale,
                          shift, vel, rsd_factor, rsd_axis));
        return finish(c, flags);
    }
    // Host pointers: stage through device scratch.
    PYL_TRY(ensure(c, BUF_POS, (size_t)(n ? n : 1) * axes * sizeof(float)));
    PYL_TRY(ensure(c, BUF_GRID, (size_t)ncell * sizeof(float)));
    if (W) PYL_TRY(ensure(c, BUF_W, (size_t)(n ? n : 1) * sizeof(float)));
    float *dpos = (float *)c->bufs[BUF_POS].ptr;
    float *dgrid = (float *)c->bufs[BUF_GRID].ptr;
    float *dW = W ? (float *)c->bufs[BUF_W].ptr : nullptr;
    const float *dvel = nullptr;
    if (vel) {
        PYL_TRY(ensure(c, BUF_VEL, (size_t)(n ? n : 1) * 3 * sizeof(float)));
        dvel = (const float *)c->bufs[BUF_VEL].ptr;
        if (n) PYL_HIP(hipMemcpyAsync((void *)dvel, vel, (size_t)n * 3 * sizeof(float),
                                      hipMemcpyHostToDevice, c->stream));
    }
    if (n) PYL_HIP(hipMemcpyAsync(dpos, pos, (size_t)n * axes * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (W && n) PYL_HIP(hipMemcpyAsync(dW, W, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (!(flags & PYL_MA_OVERWRITE))
        PYL_HIP(hipMemcpyAsync(dgrid, grid, (size_t)ncell * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PYL_TRY(ma_device(c, dpos, n, offs, nreal, axes, dW, dgrid, dims, BoxSize, mas,
                      flags | PYL_PTR_DEVICE, scale, shift, dvel, rsd_factor, rsd_axis));
    PYL_HIP(hipMemcpyAsync(grid, dgrid, (size_t)ncell * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PYL_HIP(hipStreamSynchronize(c->stream));
    return PYL_OK;
}

extern "C" int pyl_ma(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
                      float *grid, int dims, float BoxSize, int mas, int flags) {
    return ma_entry(ctx, pos, n, nullptr, 1, axes, W, grid, dims, BoxSize, mas, flags);
}

extern "C" int pyl_ma_noalloc(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
                              float *grid, int dims, float BoxSize, int mas, int flags) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    PYL_CHECK(flags & PYL_PTR_DEVICE, PYL_EINVAL, "pyl_ma_noalloc takes device pointers only");
    PYL_CHECK(!(flags & PYL_MA_ATOMIC), PYL_EINVAL, "pyl_ma_noalloc runs the tiled algorithm only");
    PYL_CHECK(c->noalloc_base != nullptr, PYL_EINVAL, "context has no fixed scratch");
    c->arena = c->noalloc_base;
    c->arena_cap = PYL_MA_NOALLOC_SCRATCH;
    c->arena_used = 0;
    const int s = ma_entry(ctx, pos, n, nullptr, 1, axes, W, grid, dims, BoxSize, mas, flags);
    c->arena = nullptr;
    c->arena_cap = c->arena_used = 0;
    return s;
}

extern "C" int pyl_ma_delta(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
                            float *delta, int dims, float BoxSize, int mas, float scale,
                            float shift, int flags) {
    return ma_entry(ctx, pos, n, nullptr, 1, axes, W, delta, dims, BoxSize, mas,
                    flags | PYL_MA_OVERWRITE, scale, shift);
}

extern "C" int pyl_ma_rsd(pyl_ctx *ctx, const float *pos, const float *vel, int64_t n,
                          const float *W, float *grid, int dims, float BoxSize, int mas,
                          float Hubble, float redshift, int axis, float scale, float shift,
                          int flags) {
    PYL_CHECK(vel != nullptr || n == 0, PYL_EINVAL, "NULL velocities");
    PYL_CHECK(axis >= 0 && axis <= 2, PYL_EINVAL, "axis must be 0, 1 or 2");
    PYL_CHECK(Hubble != 0.0f, PYL_EINVAL, "Hubble must be non-zero");
    PYL_CHECK((flags & PYL_MA_OVERWRITE) || (scale == 1.0f && shift == 0.0f), PYL_EINVAL,
              "scale/shift need PYL_MA_OVERWRITE");
    // redshift_space_library.pyx:36: factor = (1.0 + redshift)/Hubble in double,
    // stored to a float
    const float factor = (float)((1.0 + (double)redshift) / (double)Hubble);
    return ma_entry(ctx, pos, n, nullptr, 1, 3, W, grid, dims, BoxSize, mas, flags, scale, shift,
                    vel ? vel : pos, factor, axis);
}

extern "C" int pyl_ma_batch(pyl_ctx *ctx, const float *pos, const int64_t *offsets, int nreal,
                            int axes, const float *W, float *grids, int dims, float BoxSize,
                            int mas, int flags) {
    PYL_CHECK(offsets != nullptr && nreal >= 1, PYL_EINVAL, "need nreal >= 1 and offsets");
    PYL_CHECK(offsets[0] == 0, PYL_EINVAL, "offsets[0] must be 0");
    for (int r = 0; r < nreal; ++r)
        PYL_CHECK(offsets[r + 1] >= offsets[r], PYL_EINVAL, "offsets must be non-decreasing");
    static_assert(sizeof(int64_t) == sizeof(long long), "int64_t");
    return ma_entry(ctx, pos, offsets[nreal], (const long long *)offsets, nreal, axes, W, grids,
                    dims, BoxSize, mas, flags);
}

extern "C" int pyl_ma_batch_dev(pyl_ctx *ctx, const float *pos, const int64_t *d_offsets, int nreal,
                                int64_t n_cap, int chunks_per_real, int axes, const float *W,
                                float *grids, int dims, float BoxSize, int mas, int flags) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    PYL_CHECK(d_offsets != nullptr && nreal >= 1, PYL_EINVAL, "need nreal >= 1 and device offsets");
    PYL_CHECK(flags & PYL_PTR_DEVICE, PYL_EINVAL, "pyl_ma_batch_dev takes device pointers only");
    PYL_CHECK(!(flags & PYL_MA_ATOMIC), PYL_EINVAL, "pyl_ma_batch_dev runs the tiled algorithm only");
    PYL_CHECK(chunks_per_real >= 1 && chunks_per_real <= 4096, PYL_EINVAL,
              "chunks_per_real must be in 1..4096 (got %d)", chunks_per_real);
    PYL_CHECK((long long)nreal * chunks_per_real < (1 << 30), PYL_EINVAL, "too many particle chunks");
    Ctx *c = &ctx->c;
    c->dev_offs = (const long long *)d_offsets;
    c->chunks_per_real = chunks_per_real;
    static_assert(sizeof(int64_t) == sizeof(long long), "int64_t");
    const int s = ma_entry(ctx, pos, n_cap, nullptr, nreal, axes, W, grids, dims, BoxSize, mas, flags);
    c->dev_offs = nullptr;
    c->chunks_per_real = 0;
    return s;
}

extern "C" int pyl_cic_interp(pyl_ctx *ctx, const float *density, int dims, float BoxSize,
                              const float *pos, int64_t n, float *den, int flags) {
    TraceRange trace_("pyl_cic_interp");
    PYL_CHECK(ctx && density && den && (pos || n == 0), PYL_EINVAL, "NULL argument");
    PYL_CHECK(dims >= 1 && n >= 0 && BoxSize > 0.0f, PYL_EINVAL, "bad arguments");
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    if (n == 0) return PYL_OK;
    const float inv = (float)dims / BoxSize;
    const long long ncell = (long long)dims * dims * dims;
    const float *dd = density, *dp = pos;
    float *dden = den;
    if (!(flags & PYL_PTR_DEVICE)) {
        PYL_TRY(ensure(c, BUF_GRID, (size_t)ncell * sizeof(float)));
        PYL_TRY(ensure(c, BUF_POS, (size_t)n * 3 * sizeof(float)));
        PYL_TRY(ensure(c, BUF_W, (size_t)n * sizeof(float)));
        dd = (const float *)c->bufs[BUF_GRID].ptr;
        dp = (const float *)c->bufs[BUF_POS].ptr;
        dden = (float *)c->bufs[BUF_W].ptr;
        PYL_HIP(hipMemcpyAsync((void *)dd, density, (size_t)ncell * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PYL_HIP(hipMemcpyAsync((void *)dp, pos, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    cic_interp_kernel<<<grid_blocks(n, kBlock), kBlock, 0, c->stream>>>(dd, dims, inv, dp, n, dden);
    PYL_HIP(hipGetLastError());
    if (!(flags & PYL_PTR_DEVICE))
        PYL_HIP(hipMemcpyAsync(den, dden, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return finish(c, flags);
}

// ---- MAS_c.h:3-10 compatible entry points --------------------------------
namespace {
pyl_ctx *default_ctx() {
    static pyl_ctx *ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (pyl_ctx_create(&ctx, 0, 0) != PYL_OK) ctx = nullptr;
    });
    return ctx;
}

void mas_c_entry(int mas, float *pos, float *number, float *W, long particles, int dims,
                 int axes, float BoxSize) {
    pyl_ctx *ctx = default_ctx();
    int s = ctx ? pyl_ma(ctx, pos, particles, axes, W, number, dims, BoxSize, mas, 0)
                : PYL_EHIP;
    if (s != PYL_OK) fprintf(stderr, "pylians_hip: %s\n", pyl_last_error());
}
}  // namespace

extern "C" {
void NGP(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads) {
    (void)threads;
    mas_c_entry(PYL_NGP, pos, number, W, particles, dims, axes, BoxSize);
}
void CIC(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads) {
    (void)threads;
    mas_c_entry(PYL_CIC, pos, number, W, particles, dims, axes, BoxSize);
}
void TSC(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads) {
    (void)threads;
    mas_c_entry(PYL_TSC, pos, number, W, particles, dims, axes, BoxSize);
}
void PCS(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads) {
    (void)threads;
    mas_c_entry(PYL_PCS, pos, number, W, particles, dims, axes, BoxSize);
}
}
