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
//  * tiled (default): the grid is cut into tiles; a counting sort places a
//    copy of each particle in every tile its footprint overlaps; one workgroup
//    per tile sums the footprints of its records into an LDS accumulator and
//    writes each node of its tile exactly once.  Contributions are rounded to
//    an integer multiple of a per-tile power-of-two quantum 2^q chosen so
//    that every partial sum is an integer below 2^53: the float64 LDS adds are
//    then exact, so the result does not depend on the order of the atomics and
//    is bitwise reproducible.  The quantum is ~2^-52 of the tile's weight bound
//    (count x max|W|), far below float32 resolution.
#include "pyl_internal.h"

namespace pyl {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int wrapi(int i, int d) {
    int r = i % d;
    return r < 0 ? r + d : r;
}

// (base + j) mod d for base in [0,d) and 0 <= j < 4; the second test only
// fires on grids smaller than the footprint (d < 4).
__device__ __forceinline__ int node_add(int base, int j, int d) {
    int r = base + j;
    if (r >= d) r -= d;
    if (r >= d) r %= d;
    return r;
}

// Per-axis footprint.  Writes the first node (wrapped into [0,dims)) and the F
// weights; node j of the footprint is (base + j) mod dims.
//
// The reference takes floor() of (double)dist - 1.5 (TSC) / - 2.0 (PCS) and
// truncates (double)dist + 0.5 (NGP).  With k = floorf(dist) and the exact
// fraction f = dist - k (exact for any float), these equal k - 2 + (f >= 0.5),
// k - 2 and k + (f >= 0.5) (for dist >= -0.5) in integer arithmetic, for every
// |dist| < 2^31.
template <int MAS>
__device__ __forceinline__ int axis_foot(float x, float inv, int dims, float *C) {
    if constexpr (MAS == PYL_NGP) {
        // MAS_library.pyx:288-289: <int>(pos*inv + 0.5), the 0.5 is a double
        C[0] = 1.0f;
        const float v = x * inv;
        if (!(v >= -0.5f)) return wrapi((int)((double)v + 0.5), dims);  // truncation toward 0
        const float k = floorf(v);
        return wrapi((int)k + ((v - k) >= 0.5f ? 1 : 0), dims);
    } else if constexpr (MAS == PYL_CIC) {
        // MAS_library.pyx:151-156: u = dist - <int>dist, d = 1.0 - u
        float dist = x * inv;
        int i = (int)dist;
        float u = dist - (float)i;
        C[1] = u;
        C[0] = 1.0f - u;
        return wrapi(i, dims);
    } else if constexpr (MAS == PYL_TSC) {
        // MAS_library.pyx:392-399
        float dist = x * inv;
        const float kf = floorf(dist);
        int minimum = (int)kf - ((dist - kf) >= 0.5f ? 1 : 2);   // floor(dist - 1.5)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float diff = fabsf((float)(minimum + j + 1) - dist);
            float c;
            if (diff < 0.5f) {
                float d2 = diff * diff;          // float product, then 0.75 - it
                c = 0.75f - d2;
            } else if (diff < 1.5f) {
                float a = 1.5f - diff;           // exact
                c = 0.5f * (a * a);              // == (float)(0.5*(1.5-d)*(1.5-d))
            } else {
                c = 0.0f;
            }
            C[j] = c;
        }
        return wrapi(minimum + 1, dims);
    } else {
        // PCS, MAS_library.pyx:485-492 (double polynomial, rounded to float)
        float dist = x * inv;
        int minimum = (int)floorf(dist) - 2;                       // floor(dist - 2.0)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float diff = fabsf((float)(minimum + j + 1) - dist);
            double dd = (double)diff;
            float c;
            if (diff < 1.0f)
                c = (float)((4.0 - 6.0 * dd * dd + 3.0 * dd * dd * dd) / 6.0);
            else if (diff < 2.0f)
                c = (float)((2.0 - dd) * (2.0 - dd) * (2.0 - dd) / 6.0);
            else
                c = 0.0f;
            C[j] = c;
        }
        return wrapi(minimum + 1, dims);
    }
}

// Geometry of the tiled deposit.  2D grids are 3D grids with one z plane.  A
// batch holds nreal grids of the same size; tile keys are r*ntiles + tile.
struct TileGeom {
    int dims;          // nodes per axis (x, y and, for 3D, z)
    int dz;            // nodes along z: dims (3D) or 1 (2D)
    int T[3];          // tile extent per axis
    int nt[3];         // tiles per axis
    long long ntiles;  // tiles per grid
    long long ncell;   // nodes per grid
    int nreal;         // grids in the batch
};

// Realisation owning particle i: offs[r] <= i < offs[r+1] (offs has nreal+1 entries).
__device__ __forceinline__ int realisation_of(const long long *__restrict__ offs, int nreal,
                                              long long i) {
    int lo = 0, hi = nreal;  // invariant: offs[lo] <= i < offs[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (offs[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Distinct tiles touched along one axis by the F consecutive nodes starting at
// base (mod d).  Returns their count (1..3) in order.
template <int F>
__device__ __forceinline__ int axis_tiles(int base, int d, int T, int *t) {
    int n = 0;
    int last = -1;
#pragma unroll
    for (int j = 0; j < F; ++j) {
        int node = node_add(base, j, d);
        int ti = node / T;
        if (ti != last) {
            bool seen = false;
            for (int k = 0; k < n; ++k) seen |= (t[k] == ti);
            if (!seen) t[n++] = ti;
            last = ti;
        }
    }
    return n;
}

// ---- counting sort by tile -------------------------------------------------
// Each workgroup owns a contiguous chunk of particles and counts its (particle,
// tile) copies in an LDS hash table (open addressing, linear probing, no
// deletion).  At the end one global atomic per distinct key reserves the chunk's
// range inside the tile's segment; the (key, base) table is kept in global
// memory so the scatter pass can hand out slots with LDS atomics only.  Keys
// that do not fit in the table (more than kHashCap distinct tiles in a chunk)
// take per-group global atomics on a separate overflow counter.  Lanes of a
// wave that share a key are grouped first, so coherent input (a lattice, a
// sorted catalogue) costs one table operation per wave and distinct tile.
constexpr int kSortBlock = 512;
constexpr int kHashBits = 13;
constexpr int kHashCap = 1 << kHashBits;      // 8192 entries, 64 KB of LDS (key + value)
constexpr int kProbe = 32;
constexpr unsigned kEmpty = 0xffffffffu;

__device__ __forceinline__ unsigned hslot(unsigned key) {
    return (key * 0x9E3779B1u) >> (32 - kHashBits);
}

__device__ __forceinline__ int tab_insert(unsigned *tkey, unsigned *tval, unsigned key,
                                          unsigned n) {
    const unsigned h = hslot(key);
    for (int p = 0; p < kProbe; ++p) {
        const unsigned s = (h + p) & (kHashCap - 1);
        const unsigned prev = atomicCAS(&tkey[s], kEmpty, key);
        if (prev == kEmpty || prev == key) {
            atomicAdd(&tval[s], n);
            return (int)s;
        }
    }
    return -1;
}

__device__ __forceinline__ int tab_find(const unsigned *tkey, unsigned key) {
    const unsigned h = hslot(key);
    for (int p = 0; p < kProbe; ++p) {
        const unsigned s = (h + p) & (kHashCap - 1);
        const unsigned k = tkey[s];
        if (k == key) return (int)s;
        if (k == kEmpty) return -1;
    }
    return -1;
}

template <int MAS>
__device__ __forceinline__ void particle_tiles(const float *pos, long long i, int axes,
                                               float inv, const TileGeom &g, int tx[3],
                                               int ty[3], int tz[3], int &nx, int &ny,
                                               int &nz) {
    constexpr int F = MAS;
    float C[4];
    int bx = axis_foot<MAS>(pos[(long long)axes * i + 0], inv, g.dims, C);
    int by = axis_foot<MAS>(pos[(long long)axes * i + 1], inv, g.dims, C);
    nx = axis_tiles<F>(bx, g.dims, g.T[0], tx);
    ny = axis_tiles<F>(by, g.dims, g.T[1], ty);
    if (axes == 3) {
        int bz = axis_foot<MAS>(pos[(long long)axes * i + 2], inv, g.dims, C);
        nz = axis_tiles<F>(bz, g.dims, g.T[2], tz);
    } else {
        nz = 1;
        tz[0] = 0;
    }
}

__device__ __forceinline__ void chunk_range(long long n, long long chunk, long long &c0,
                                            long long &c1) {
    c0 = (long long)blockIdx.x * chunk;
    c1 = min(n, c0 + chunk);
}

// Pass 1: per-tile counts.  counts[key] receives the table-counted copies (and
// hands each chunk its base), ocounts[key] the overflowed ones.
template <int MAS>
__global__ void __launch_bounds__(kSortBlock)
tile_count_kernel(const float *__restrict__ pos, long long n, int axes, float inv, TileGeom g,
                  const long long *__restrict__ offs, long long chunk,
                  unsigned *__restrict__ counts, unsigned *__restrict__ ocounts,
                  uint2 *__restrict__ dump) {
    __shared__ unsigned tkey[kHashCap];
    __shared__ unsigned tval[kHashCap];
    for (int s = threadIdx.x; s < kHashCap; s += blockDim.x) {
        tkey[s] = kEmpty;
        tval[s] = 0;
    }
    __syncthreads();
    const int lane = __lane_id();
    long long c0, c1;
    chunk_range(n, chunk, c0, c1);
    for (long long base = c0; base < c1; base += blockDim.x) {
        const long long i = base + threadIdx.x;
        const bool valid = i < c1;
        int tx[3], ty[3], tz[3], nx = 0, ny = 0, nz = 0;
        long long rbase = 0;
        if (valid) {
            particle_tiles<MAS>(pos, i, axes, inv, g, tx, ty, tz, nx, ny, nz);
            if (g.nreal > 1) rbase = (long long)realisation_of(offs, g.nreal, i) * g.ntiles;
        }
        const int nc = valid ? nx * ny * nz : 0;
        for (int j = 0;; ++j) {
            const bool act = j < nc;
            if (!__any(act)) break;
            unsigned key = kEmpty;
            if (act) {
                int a = j % nx, b = (j / nx) % ny, c = j / (nx * ny);
                key = (unsigned)(rbase + ((long long)tx[a] * g.nt[1] + ty[b]) * g.nt[2] + tz[c]);
            }
            unsigned long long pend = __ballot(act);
            while (pend) {
                const int leader = __ffsll((unsigned long long)pend) - 1;
                const unsigned lk = __shfl(key, leader);
                const unsigned long long grp = __ballot(act && key == lk) & pend;
                if (lane == leader) {
                    const unsigned cnt = (unsigned)__popcll(grp);
                    if (tab_insert(tkey, tval, lk, cnt) < 0) atomicAdd(&ocounts[lk], cnt);
                }
                pend &= ~grp;
            }
        }
    }
    __syncthreads();
    uint2 *dmp = dump + (long long)blockIdx.x * kHashCap;
    for (int s = threadIdx.x; s < kHashCap; s += blockDim.x) {
        const unsigned k = tkey[s];
        uint2 d = make_uint2(kEmpty, 0u);
        if (k != kEmpty) {
            d.x = k;
            d.y = atomicAdd(&counts[k], tval[s]);
        }
        dmp[s] = d;
    }
}

// Pass 3: write one record (x, y, z, w) per (particle, overlapped tile).  Record
// index = tile offset + chunk base + LDS cursor (table keys), or tile offset +
// table count + global overflow cursor.
template <int MAS>
__global__ void __launch_bounds__(kSortBlock)
tile_scatter_kernel(const float *__restrict__ pos, const float *__restrict__ W, long long n,
                    int axes, float inv, TileGeom g, const long long *__restrict__ offs,
                    long long chunk, const uint2 *__restrict__ dump,
                    const unsigned long long *__restrict__ offsets,
                    const unsigned *__restrict__ counts, unsigned *__restrict__ ocursors,
                    float4 *__restrict__ records) {
    __shared__ unsigned tkey[kHashCap];
    __shared__ unsigned tcur[kHashCap];
    const uint2 *dmp = dump + (long long)blockIdx.x * kHashCap;
    for (int s = threadIdx.x; s < kHashCap; s += blockDim.x) {
        const uint2 d = dmp[s];
        tkey[s] = d.x;
        tcur[s] = d.y;
    }
    __syncthreads();
    const int lane = __lane_id();
    long long c0, c1;
    chunk_range(n, chunk, c0, c1);
    for (long long base = c0; base < c1; base += blockDim.x) {
        const long long i = base + threadIdx.x;
        const bool valid = i < c1;
        int tx[3], ty[3], tz[3], nx = 0, ny = 0, nz = 0;
        long long rbase = 0;
        float4 rec = make_float4(0.f, 0.f, 0.f, 1.f);
        if (valid) {
            particle_tiles<MAS>(pos, i, axes, inv, g, tx, ty, tz, nx, ny, nz);
            if (g.nreal > 1) rbase = (long long)realisation_of(offs, g.nreal, i) * g.ntiles;
            rec.x = pos[(long long)axes * i + 0];
            rec.y = pos[(long long)axes * i + 1];
            rec.z = (axes == 3) ? pos[(long long)axes * i + 2] : 0.f;
            rec.w = W ? W[i] : 1.0f;
        }
        const int nc = valid ? nx * ny * nz : 0;
        for (int j = 0;; ++j) {
            const bool act = j < nc;
            if (!__any(act)) break;
            unsigned key = kEmpty;
            if (act) {
                int a = j % nx, b = (j / nx) % ny, c = j / (nx * ny);
                key = (unsigned)(rbase + ((long long)tx[a] * g.nt[1] + ty[b]) * g.nt[2] + tz[c]);
            }
            unsigned long long pend = __ballot(act);
            unsigned long long slot = 0;
            while (pend) {
                const int leader = __ffsll((unsigned long long)pend) - 1;
                const unsigned lk = __shfl(key, leader);
                const unsigned long long grp = __ballot(act && key == lk) & pend;
                unsigned long long gb = 0;
                if (lane == leader) {
                    const unsigned cnt = (unsigned)__popcll(grp);
                    const int s = tab_find(tkey, lk);
                    if (s >= 0) gb = offsets[lk] + atomicAdd(&tcur[s], cnt);
                    else gb = offsets[lk] + counts[lk] + atomicAdd(&ocursors[lk], cnt);
                }
                gb = __shfl(gb, leader);
                if ((grp >> lane) & 1ull) slot = gb + __popcll(grp & ((1ull << lane) - 1ull));
                pend &= ~grp;
            }
            if (act) records[slot] = rec;
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

// Pass 4: one workgroup per tile.  LDS accumulator of T0*T1*T2 doubles holding
// integer multiples of 2^q (exact sums), then each node written once.
template <int MAS, int AXES>
__global__ void __launch_bounds__(kBlock)
tile_deposit_kernel(const float4 *__restrict__ records,
                    const unsigned long long *__restrict__ offsets, float inv, TileGeom g,
                    int weighted, float weight_factor, float *__restrict__ grid, int overwrite) {
    constexpr int F = MAS;
    extern __shared__ double acc[];
    __shared__ unsigned wred[kBlock / 64];
    const long long key = blockIdx.x;                 // r * ntiles + tile
    const long long real = key / g.ntiles;
    const long long tile = key - real * g.ntiles;
    grid += real * g.ncell;
    const int tz = (int)(tile % g.nt[2]);
    const int ty = (int)((tile / g.nt[2]) % g.nt[1]);
    const int tx = (int)(tile / ((long long)g.nt[2] * g.nt[1]));
    const int ox = tx * g.T[0], oy = ty * g.T[1], oz = tz * g.T[2];
    const int ex = min(g.T[0], g.dims - ox);
    const int ey = min(g.T[1], g.dims - oy);
    const int ez = (AXES == 3) ? min(g.T[2], g.dims - oz) : 1;
    const int ncell = ex * ey * ez;
    const unsigned long long r0 = offsets[key], r1 = offsets[key + 1];

    if (r0 == r1) {
        if (overwrite) {
            for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
                int lz = c % ez, ly = (c / ez) % ey, lx = c / (ez * ey);
                long long node = AXES == 3
                    ? ((long long)(ox + lx) * g.dims + (oy + ly)) * g.dims + (oz + lz)
                    : (long long)(ox + lx) * g.dims + (oy + ly);
                grid[node] = 0.0f;
            }
        }
        return;
    }
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) acc[c] = 0.0;

    // Per-tile bound on |contribution|: B-spline weights are <= 1, so every
    // |c| <= weight_factor * max finite |W| of the tile's records.
    float wmax = 1.0f;
    if (weighted) {
        unsigned m = 0;
        for (unsigned long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
            unsigned b = __float_as_uint(fabsf(records[r].w));
            if (b < 0x7f800000u) m = max(m, b);
        }
        for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
        if ((threadIdx.x & 63) == 0) wred[threadIdx.x >> 6] = m;
        __syncthreads();
        m = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = max(m, wred[i]);
        wmax = __uint_as_float(m);
    }
    // Quantum 2^q: sum |c| <= count*wmax*factor < 2^e and 2^e/2^q = 2^52, so all
    // partial sums of rint(c/2^q) are integers below 2^53: exact in float64.
    const double bound = (double)(r1 - r0) * (double)wmax * (double)weight_factor;
    int e = -1000;
    if (bound > 0.0) frexp(bound, &e);
    int q = e - 52;
    if (q < -127) q = -127;                 // keep 2^-q a normal float
    const float to_q = ldexpf(1.0f, -q);
    const double from_q = ldexp(1.0, q);
    __syncthreads();

    for (unsigned long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const float4 rec = records[r];
        float Cx[4], Cy[4], Cz[4];
        const int bx = axis_foot<MAS>(rec.x, inv, g.dims, Cx);
        const int by = axis_foot<MAS>(rec.y, inv, g.dims, Cy);
        int bz = 0;
        if (AXES == 3) bz = axis_foot<MAS>(rec.z, inv, g.dims, Cz);
        // ((Cx*Cy)*Cz)*(W*2^-q) == 2^-q * (((Cx*Cy)*Cz)*W) exactly (power-of-two
        // scaling commutes with float rounding), so the scaled product is the
        // reference's float contribution in units of the quantum.
        const float w = (rec.w * weight_factor) * to_q;
#pragma unroll
        for (int l = 0; l < F; ++l) {
            const int lx = node_add(bx, l, g.dims) - ox;
            if (lx < 0 || lx >= ex) continue;
#pragma unroll
            for (int m = 0; m < F; ++m) {
                const int ly = node_add(by, m, g.dims) - oy;
                if (ly < 0 || ly >= ey) continue;
                const float cxy = Cx[l] * Cy[m];
                if (AXES == 3) {
                    const int row = (lx * ey + ly) * ez;
#pragma unroll
                    for (int k = 0; k < F; ++k) {
                        const int lz = node_add(bz, k, g.dims) - oz;
                        if (lz < 0 || lz >= ez) continue;
                        const float c = (cxy * Cz[k]) * w;
                        atomicAdd(&acc[row + lz], (double)rintf(c));
                    }
                } else {
                    const float c = cxy * w;   // the reference multiplies by C[2][n] = 1.0
                    atomicAdd(&acc[lx * ey + ly], (double)rintf(c));
                }
            }
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncell; c += blockDim.x) {
        int lz = c % ez, ly = (c / ez) % ey, lx = c / (ez * ey);
        long long node = AXES == 3
            ? ((long long)(ox + lx) * g.dims + (oy + ly)) * g.dims + (oz + lz)
            : (long long)(ox + lx) * g.dims + (oy + ly);
        float v = (float)(acc[c] * from_q);
        grid[node] = overwrite ? v : grid[node] + v;
    }
}

// Plain form: one thread per particle, one global float atomic per node.
template <int MAS, int AXES>
__global__ void __launch_bounds__(kBlock)
atomic_deposit_kernel(const float *__restrict__ pos, const float *__restrict__ W, long long n,
                      float inv, int dims, float weight_factor, float *__restrict__ grid) {
    constexpr int F = MAS;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float Cx[4], Cy[4], Cz[4];
        int bx = axis_foot<MAS>(pos[(long long)AXES * i + 0], inv, dims, Cx);
        int by = axis_foot<MAS>(pos[(long long)AXES * i + 1], inv, dims, Cy);
        int bz = 0;
        if (AXES == 3) bz = axis_foot<MAS>(pos[(long long)AXES * i + 2], inv, dims, Cz);
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

int grid_blocks(long long n, int block) {
    long long b = (n + block - 1) / block;
    if (b < 1) b = 1;
    if (b > 65536) b = 65536;
    return (int)b;
}

TileGeom make_geom(int dims, int axes, int nreal) {
    TileGeom g{};
    g.dims = dims;
    g.dz = axes == 3 ? dims : 1;
    if (axes == 3) {
        g.T[0] = g.T[1] = g.T[2] = dims < 16 ? dims : 16;
    } else {
        g.T[0] = g.T[1] = dims < 64 ? dims : 64;
        g.T[2] = 1;
    }
    g.nt[0] = ceil_div(dims, g.T[0]);
    g.nt[1] = ceil_div(dims, g.T[1]);
    g.nt[2] = axes == 3 ? ceil_div(dims, g.T[2]) : 1;
    g.ntiles = (long long)g.nt[0] * g.nt[1] * g.nt[2];
    g.ncell = (long long)dims * dims * (axes == 3 ? dims : 1);
    g.nreal = nreal;
    return g;
}

// Particle ranges of a batch on the device (offs[0..nreal]); NULL for one grid.
int upload_offsets(Ctx *c, const long long *h_offs, int nreal, const long long **d_offs) {
    *d_offs = nullptr;
    if (nreal <= 1) return PYL_OK;
    PYL_TRY(ensure(c, BUF_SCAN_TMP, sizeof(long long) * (size_t)(nreal + 1)));
    long long *d = (long long *)c->bufs[BUF_SCAN_TMP].ptr;
    PYL_HIP(hipMemcpyAsync(d, h_offs, sizeof(long long) * (size_t)(nreal + 1),
                           hipMemcpyHostToDevice, c->stream));
    PYL_HIP(hipStreamSynchronize(c->stream));   // h_offs may be a temporary
    *d_offs = d;
    return PYL_OK;
}

template <int MAS>
int run_tiled(Ctx *c, const float *pos, long long n, const long long *h_offs, int nreal, int axes,
              const float *W, float *grid, int dims, float inv, float weight_factor,
              int overwrite) {
    TileGeom g = make_geom(dims, axes, nreal);
    const long long nt = g.ntiles * nreal;      // keys over the whole batch
    const int nscan_blocks = (int)((nt + kScanTile - 1) / kScanTile);
    PYL_CHECK(nscan_blocks <= kScanTile, PYL_EINVAL, "too many tiles (%lld)", nt);
    PYL_CHECK(nt < (1ll << 31), PYL_EINVAL, "too many tiles (%lld)", nt);
    const long long *d_offs = nullptr;
    PYL_TRY(upload_offsets(c, h_offs, nreal, &d_offs));
    // one chunk of particles per sort workgroup
    const int nchunk = (int)std::max(1LL, std::min(512LL, (n + 65535) / 65536));
    const long long chunk = (n + nchunk - 1) / nchunk;
    // tile buffer: counts[nt] | ocounts[nt] | ocursors[nt] | pad | offsets[nt+1] |
    //              block sums | per-chunk tables
    const size_t nwords = (size_t)nt * 3;
    const size_t cnt_bytes = (nwords * sizeof(unsigned) + 255) & ~(size_t)255;
    const size_t off_bytes = (((size_t)nt + 1) * sizeof(unsigned long long) + 255) & ~(size_t)255;
    const size_t bs_bytes = ((size_t)nscan_blocks * sizeof(unsigned long long) + 255) & ~(size_t)255;
    const size_t dump_bytes = (size_t)nchunk * kHashCap * sizeof(uint2);
    PYL_TRY(ensure(c, BUF_TILE_CNT, cnt_bytes + off_bytes + bs_bytes + dump_bytes));
    char *tb = (char *)c->bufs[BUF_TILE_CNT].ptr;
    unsigned *counts = (unsigned *)tb;
    unsigned *ocounts = counts + nt;
    unsigned *ocursors = ocounts + nt;
    unsigned long long *offsets = (unsigned long long *)(tb + cnt_bytes);
    unsigned long long *bsums = (unsigned long long *)(tb + cnt_bytes + off_bytes);
    uint2 *dump = (uint2 *)(tb + cnt_bytes + off_bytes + bs_bytes);

    PYL_HIP(hipMemsetAsync(counts, 0, nwords * sizeof(unsigned), c->stream));
    if (n > 0) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_COUNT));
        tile_count_kernel<MAS><<<nchunk, kSortBlock, 0, c->stream>>>(pos, n, axes, inv, g, d_offs,
                                                                     chunk, counts, ocounts, dump);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_COUNT));
    }
    PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCAN));
    scan_partial_kernel<<<nscan_blocks, kScanBlock, 0, c->stream>>>(counts, ocounts, nt, bsums);
    scan_blocksums_kernel<<<1, kScanBlock, 0, c->stream>>>(bsums, nscan_blocks);
    scan_final_kernel<<<nscan_blocks, kScanBlock, 0, c->stream>>>(counts, ocounts, nt, bsums,
                                                                   offsets);
    PYL_HIP(hipGetLastError());
    PYL_TRY(stage_end(c, PYL_STAGE_MA_SCAN));
    unsigned long long total = 0;
    PYL_HIP(hipMemcpyAsync(&total, offsets + nt, sizeof(total), hipMemcpyDeviceToHost, c->stream));
    PYL_HIP(hipStreamSynchronize(c->stream));
    PYL_TRY(ensure(c, BUF_RECORDS, (size_t)(total ? total : 1) * sizeof(float4)));
    float4 *records = (float4 *)c->bufs[BUF_RECORDS].ptr;
    if (n > 0) {
        PYL_TRY(stage_begin(c, PYL_STAGE_MA_SCATTER));
        tile_scatter_kernel<MAS><<<nchunk, kSortBlock, 0, c->stream>>>(
            pos, W, n, axes, inv, g, d_offs, chunk, dump, offsets, counts, ocursors, records);
        PYL_HIP(hipGetLastError());
        PYL_TRY(stage_end(c, PYL_STAGE_MA_SCATTER));
    }
    const size_t lds = (size_t)g.T[0] * g.T[1] * g.T[2] * sizeof(double);
    const int weighted = W != nullptr;
    PYL_TRY(stage_begin(c, PYL_STAGE_MA_DEPOSIT));
    if (axes == 3)
        tile_deposit_kernel<MAS, 3><<<(unsigned)nt, kBlock, lds, c->stream>>>(
            records, offsets, inv, g, weighted, weight_factor, grid, overwrite);
    else
        tile_deposit_kernel<MAS, 2><<<(unsigned)nt, kBlock, lds, c->stream>>>(
            records, offsets, inv, g, weighted, weight_factor, grid, overwrite);
    PYL_HIP(hipGetLastError());
    PYL_TRY(stage_end(c, PYL_STAGE_MA_DEPOSIT));
    return PYL_OK;
}

template <int MAS>
int run_atomic(Ctx *c, const float *pos, long long n, const long long *h_offs, int nreal,
               int axes, const float *W, float *grid, int dims, float inv, float weight_factor,
               int overwrite) {
    const long long ncell = (long long)dims * dims * (axes == 3 ? dims : 1);
    if (overwrite)
        PYL_HIP(hipMemsetAsync(grid, 0, (size_t)ncell * nreal * sizeof(float), c->stream));
    if (n == 0) return PYL_OK;
    PYL_TRY(stage_begin(c, PYL_STAGE_MA_ATOMIC));
    for (int r = 0; r < nreal; ++r) {
        const long long b = h_offs ? h_offs[r] : 0, e = h_offs ? h_offs[r + 1] : n;
        if (e <= b) continue;
        const float *p = pos + (long long)axes * b;
        const float *w = W ? W + b : nullptr;
        float *gr = grid + (long long)r * ncell;
        const int nb = grid_blocks(e - b, kBlock);
        if (axes == 3)
            atomic_deposit_kernel<MAS, 3><<<nb, kBlock, 0, c->stream>>>(p, w, e - b, inv, dims,
                                                                       weight_factor, gr);
        else
            atomic_deposit_kernel<MAS, 2><<<nb, kBlock, 0, c->stream>>>(p, w, e - b, inv, dims,
                                                                       weight_factor, gr);
        PYL_HIP(hipGetLastError());
    }
    PYL_TRY(stage_end(c, PYL_STAGE_MA_ATOMIC));
    return PYL_OK;
}

}  // namespace

// Device-pointer MA used by the C ABI and by the batched driver.
// h_offs: host particle ranges of nreal grids (nreal+1 entries, NULL when nreal == 1);
// grid holds nreal contiguous grids.
int ma_device(Ctx *c, const float *pos, long long n, const long long *h_offs, int nreal,
              int axes, const float *W, float *grid, int dims, float BoxSize, int mas,
              int flags) {
    // MAS_library.pyx:135 inv_cell_size = dims/BoxSize in float (MAS_c.c:19 computes it
    // in double and stores it to float, which rounds to the same value).
    const float inv = (float)dims / BoxSize;
    float weight_factor = 1.0f;
    const long long ncell = (long long)dims * dims * (axes == 3 ? dims : 1) * nreal;
    const int overwrite = (flags & PYL_MA_OVERWRITE) ? 1 : 0;
    if (axes == 2 && (flags & PYL_MA_CYTHON_2D) && mas != PYL_NGP) {
        // MA() runs the 3D kernel on a (d,d,1) view, so each deposit lands F times
        // in the same cell; with renormalize_2D the whole grid is then divided by F
        // (MAS_library.pyx:84-110).  Equivalent: grid/F + deposit, or grid + F*deposit.
        if (flags & PYL_MA_RENORM_2D) {
            if (!overwrite) {
                scale_grid_kernel<<<grid_blocks(ncell, kBlock), kBlock, 0, c->stream>>>(
                    grid, ncell, (float)mas);
                PYL_HIP(hipGetLastError());
            }
        } else {
            weight_factor = (float)mas;
        }
    }
    const bool atomic = (flags & PYL_MA_ATOMIC) != 0;
    switch (mas) {
#define PYL_MA_CASE(M)                                                                      \
    case M:                                                                                 \
        return atomic ? run_atomic<M>(c, pos, n, h_offs, nreal, axes, W, grid, dims, inv,    \
                                      weight_factor, overwrite)                             \
                      : run_tiled<M>(c, pos, n, h_offs, nreal, axes, W, grid, dims, inv,     \
                                     weight_factor, overwrite);
        PYL_MA_CASE(PYL_NGP)
        PYL_MA_CASE(PYL_CIC)
        PYL_MA_CASE(PYL_TSC)
        PYL_MA_CASE(PYL_PCS)
#undef PYL_MA_CASE
    }
    set_error("unknown MAS id %d", mas);
    return PYL_EINVAL;
}

}  // namespace pyl

using namespace pyl;

static int ma_entry(pyl_ctx *ctx, const float *pos, int64_t n, const long long *offs, int nreal,
                    int axes, const float *W, float *grid, int dims, float BoxSize, int mas,
                    int flags) {
    PYL_CHECK(ctx != nullptr, PYL_EINVAL, "NULL context");
    Ctx *c = &ctx->c;
    PYL_CHECK(axes == 2 || axes == 3, PYL_EINVAL, "pos must have 2 or 3 columns (got %d)", axes);
    PYL_CHECK(dims >= 1, PYL_EINVAL, "dims must be positive (got %d)", dims);
    PYL_CHECK(mas >= PYL_NGP && mas <= PYL_PCS, PYL_EINVAL, "option not valid (MAS id %d)", mas);
    PYL_CHECK(n >= 0, PYL_EINVAL, "negative particle count");
    PYL_CHECK(grid != nullptr && (pos != nullptr || n == 0), PYL_EINVAL, "NULL array");
    PYL_CHECK(BoxSize > 0.0f, PYL_EINVAL, "BoxSize must be positive");
    const long long ncell = (long long)dims * dims * (axes == 3 ? dims : 1) * nreal;
    PYL_CHECK(ncell < (1ll << 34), PYL_EINVAL, "grid too large");
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    if (flags & PYL_PTR_DEVICE) {
        PYL_TRY(ma_device(c, pos, n, offs, nreal, axes, W, grid, dims, BoxSize, mas, flags));
        return finish(c, flags);
    }
    // Host pointers: stage through device scratch.
    PYL_TRY(ensure(c, BUF_POS, (size_t)(n ? n : 1) * axes * sizeof(float)));
    PYL_TRY(ensure(c, BUF_GRID, (size_t)ncell * sizeof(float)));
    if (W) PYL_TRY(ensure(c, BUF_W, (size_t)(n ? n : 1) * sizeof(float)));
    float *dpos = (float *)c->bufs[BUF_POS].ptr;
    float *dgrid = (float *)c->bufs[BUF_GRID].ptr;
    float *dW = W ? (float *)c->bufs[BUF_W].ptr : nullptr;
    if (n) PYL_HIP(hipMemcpyAsync(dpos, pos, (size_t)n * axes * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (W && n) PYL_HIP(hipMemcpyAsync(dW, W, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (!(flags & PYL_MA_OVERWRITE))
        PYL_HIP(hipMemcpyAsync(dgrid, grid, (size_t)ncell * sizeof(float), hipMemcpyHostToDevice, c->stream));
    PYL_TRY(ma_device(c, dpos, n, offs, nreal, axes, dW, dgrid, dims, BoxSize, mas,
                      flags | PYL_PTR_DEVICE));
    PYL_HIP(hipMemcpyAsync(grid, dgrid, (size_t)ncell * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    PYL_HIP(hipStreamSynchronize(c->stream));
    return PYL_OK;
}

extern "C" int pyl_ma(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
                      float *grid, int dims, float BoxSize, int mas, int flags) {
    return ma_entry(ctx, pos, n, nullptr, 1, axes, W, grid, dims, BoxSize, mas, flags);
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

extern "C" int pyl_cic_interp(pyl_ctx *ctx, const float *density, int dims, float BoxSize,
                              const float *pos, int64_t n, float *den, int flags) {
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
