// Forward 3D real-to-complex FFT of power-of-two cubes (16..1024 per side).
//
// Same transform as FFT3Dr_f (library/Pk_library/Pk_library.pyx:117-130, FFTW
// forward, unnormalised, e^{-i}): out[x][y][kz], kz in [0, N/2].  Three passes
// over HBM:
//   z: r2c of the contiguous rows (N/2-point complex FFTs of the even/odd
//      pairs, then the split into the N/2+1 output modes);
//   y: complex FFTs along y, in place;
//   x: complex FFTs along x, in place.
// The y and x passes use the same column kernel: a workgroup loads C
// consecutive kz columns (C*8 bytes contiguous per point) of one (x or y) line
// set into LDS, runs C Stockham FFTs of length N with radix-8 stages (radix 4
// or 2 for the remainder) in registers, exchanging through LDS between stages,
// and writes the block back.  Twiddles exp(-2 pi i k/N) are computed on the
// host in double and rounded to float.  Other sizes use a rocFFT 3D plan.
#include "pyl_internal.h"

namespace pyl {
namespace {

constexpr int kFftBlock = 256;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (-i)
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

template <int R>
__device__ __forceinline__ void dft(float2 *v) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const float2 s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    } else {
        // radix 8 = two radix-4 halves (even, odd) and the twiddles w8^k
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        dft<4>(e);
        dft<4>(o);
        const float h = 0.70710678118654752f;
        // o[k] *= exp(-2 pi i k/8)
        o[1] = make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        o[2] = mul_mi(o[2]);
        o[3] = make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
}

// One Stockham stage of radix R over C interleaved columns held in LDS
// (element (n, c) at buf[n*CP + c], CP >= C a padded stride).  Span Ns:
// sub-transforms of length Ns are complete; this stage makes them Ns*R long.
// BLK threads; every thread reads its inputs before any thread writes.
template <int N, int C, int CP, int R, int BLK>
__device__ __forceinline__ void stage(float2 *buf, const float2 *__restrict__ tw, int Ns) {
    constexpr int NB = (N / R) * C;                          // butterflies
    constexpr int PER = (NB + BLK - 1) / BLK;                // per thread
    float2 v[PER][R];
    int dst[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = threadIdx.x + p * BLK;
        dst[p] = -1;
        if (b < NB) {
            const int c = b % C, j = b / C;
            const int k = j % Ns;
            const int tstep = (N / (Ns * R)) * k;
#pragma unroll
            for (int m = 0; m < R; ++m) {
                float2 x = buf[(j + m * (N / R)) * CP + c];
                if (m > 0) x = cmul(x, tw[(m * tstep) & (N - 1)]);
                v[p][m] = x;
            }
            dft<R>(v[p]);
            dst[p] = ((j / Ns) * Ns * R + k) * CP + c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        if (dst[p] >= 0) {
#pragma unroll
            for (int m = 0; m < R; ++m) buf[dst[p] + m * Ns * CP] = v[p][m];
        }
    }
    __syncthreads();
}

// First and last stage of a column FFT straight from and to global memory: the
// first stage (span 1, radix R, no twiddles) reads its R inputs of butterfly b from
// base[n*nstride + c] (c clamped to the cw columns present; the copies of the last
// column in the others are never stored) and writes the LDS block; the last stage
// reads the LDS block and stores its outputs of the columns c < cw, returning the
// largest |Re|, |Im| it stored.  Same butterflies and twiddles as stage<>, so the
// same bits as staging the whole block through LDS.
template <int N, int C, int CP, int R, int BLK>
__device__ __forceinline__ void stage_first(const float2 *base, long long nstride, int cw, float2 *buf) {
    constexpr int NB = (N / R) * C;
    constexpr int PER = (NB + BLK - 1) / BLK;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = min((int)threadIdx.x + p * BLK, NB - 1);
        const int c = b % C, j = b / C;
#pragma unroll
        for (int m = 0; m < R; ++m) v[p][m] = base[(long long)(j + m * (N / R)) * nstride + min(c, cw - 1)];
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = threadIdx.x + p * BLK;
        if (b < NB) {
            const int c = b % C, j = b / C;
            dft<R>(v[p]);
#pragma unroll
            for (int m = 0; m < R; ++m) buf[(j * R + m) * CP + c] = v[p][m];
        }
    }
    __syncthreads();
}

template <int N, int C, int CP, int R, int BLK>
__device__ __forceinline__ float stage_last(const float2 *buf, const float2 *__restrict__ tw, int Ns,
                                            float2 *base, long long nstride, int cw) {
    constexpr int NB = (N / R) * C;
    constexpr int PER = (NB + BLK - 1) / BLK;
    float mx = 0.0f;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int b = threadIdx.x + p * BLK;
        if (b < NB) {
            const int c = b % C, j = b / C;
            const int k = j % Ns;
            const int tstep = (N / (Ns * R)) * k;
            float2 v[R];
#pragma unroll
            for (int m = 0; m < R; ++m) {
                float2 x = buf[(j + m * (N / R)) * CP + c];
                if (m > 0) x = cmul(x, tw[(m * tstep) & (N - 1)]);
                v[m] = x;
            }
            dft<R>(v);
            if (c < cw) {
                const int d = (j / Ns) * Ns * R + k;
#pragma unroll
                for (int m = 0; m < R; ++m) {
                    base[(long long)(d + m * Ns) * nstride + c] = v[m];
                    mx = fmaxf(mx, fmaxf(fabsf(v[m].x), fabsf(v[m].y)));
                }
            }
        }
    }
    return mx;
}

// Column FFTs with the first stage read from and the last stage written to global
// memory (stage_first, stage_last); radix-8 stages in between, the last stage of
// radix 8, 4 or 2.
template <int N, int C, int BLK>
__device__ __forceinline__ float fft_columns_global(float2 *base, long long nstride, int cw, float2 *buf,
                                                    const float2 *__restrict__ tw) {
    constexpr int RL = (N == 16 || N == 128 || N == 1024) ? 2 : (N == 32 || N == 256) ? 4 : 8;
    stage_first<N, C, C, 8, BLK>(base, nstride, cw, buf);
    int Ns = 8;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        if (Ns * 8 * RL <= N) {
            stage<N, C, C, 8, BLK>(buf, tw, Ns);
            Ns *= 8;
        }
    }
    return stage_last<N, C, C, RL, BLK>(buf, tw, Ns, base, nstride, cw);
}

template <int N, int C, int CP, int BLK>
__device__ __forceinline__ void fft_columns(float2 *buf, const float2 *__restrict__ tw) {
    int Ns = 1;
    int rem = N;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (rem >= 8) {
            stage<N, C, CP, 8, BLK>(buf, tw, Ns);
            Ns *= 8;
            rem /= 8;
        }
    }
    if (rem == 4) stage<N, C, CP, 4, BLK>(buf, tw, Ns);
    else if (rem == 2) stage<N, C, CP, 2, BLK>(buf, tw, Ns);
}

// In-place FFT of length N along one axis of a batch of complex arrays:
// element n of column (o, z) of batch b at data[b*bstride + o*ostride + n*nstride + z],
// z in [0, zlen).  Grid: (ceil(zlen/C), nother, nbatch).
// amax (optional, the last pass): the largest |Re| or |Im| of the transform of cube
// b, as float bits, atomically maxed into one of 64 words amax[b*64 + (group & 63)]
// (an integer max, so the words do not depend on the order of the atomics).
__device__ __forceinline__ void block_absmax(float m, unsigned *__restrict__ amax_words, unsigned slot) {
    __shared__ unsigned wmax[16];
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = __float_as_uint(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned mm = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) mm = max(mm, wmax[k]);
        if (mm) atomicMax(&amax_words[slot], mm);
    }
}

template <int N, int C, int BLK>
__global__ void __launch_bounds__(BLK)
fft_col_kernel(float2 *__restrict__ data, long long nstride, long long ostride, long long bstride,
               int zlen, int nother, const float2 *__restrict__ tw, unsigned *__restrict__ amax) {
    static_assert((N * C) % BLK == 0, "block shape");
    __shared__ float2 buf[N * C];
    // Workgroups are dealt to the 8 XCDs round-robin by launch index, and each
    // XCD has its own L2.  Rows are 8*zlen bytes long, so a group's 128-byte
    // column segment straddles two cache lines shared with the neighbouring
    // column groups: logical index w puts all column groups of one line set on
    // the same XCD, next to each other in time, and the shared lines are
    // fetched from HBM once.
    const unsigned ngz = (unsigned)((zlen + C - 1) / C);
    const unsigned total = gridDim.x, per = total / 8u;
    const unsigned L = blockIdx.x;
    const unsigned w = (L < per * 8u) ? (L % 8u) * per + L / 8u : L;
    const unsigned zg = w % ngz, t = w / ngz;
    const unsigned line = t % (unsigned)nother, b = t / (unsigned)nother;
    const int z0 = (int)zg * C;
    const int cw = min(C, zlen - z0);
    float2 *base = data + (long long)b * bstride + (long long)line * ostride + z0;
    const float m = fft_columns_global<N, C, BLK>(base, nstride, cw, buf, tw);
    if (amax) block_absmax(m, amax, b * 64u + (L & 63u));
}

// The same maximum over nb cubes of half-complex rows (zlen modes, zs apart, cube
// stride cstride), for transforms not made by the passes above.
__global__ void __launch_bounds__(256)
absmax_kernel(const float2 *__restrict__ dk, long long nrows, int zlen, long long zs,
              long long cstride, unsigned *__restrict__ amax) {
    const long long cube = blockIdx.y;
    const float2 *d = dk + cube * cstride;
    const long long n = nrows * zlen;
    float m = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long r = i / zlen, k = i - r * zlen;
        const float2 x = d[r * zs + k];
        m = fmaxf(m, fmaxf(fabsf(x.x), fabsf(x.y)));
    }
    block_absmax(m, amax, (unsigned)cube * 64u + (blockIdx.x & 63u));
}

// Real-to-complex FFT of rows of N floats (contiguous), R rows per workgroup:
// z[j] = x[2j] + i x[2j+1] is transformed with N/2-point Stockham FFTs, then
// X[k] = E[k] + e^{-2 pi i k/N} O[k] with E[k] = (Z[k] + conj(Z[N/2-k]))/2 and
// O[k] = -i (Z[k] - conj(Z[N/2-k]))/2, k = 0..N/2 (indices mod N/2).  Output
// rows of N/2+1 complex.
template <int N, int R>
__global__ void __launch_bounds__(kFftBlock)
fft_r2c_rows_kernel(const float *__restrict__ in, float2 *__restrict__ out, long long nrows,
                    const float2 *__restrict__ tw_half, const float2 *__restrict__ tw_full,
                    long long zs) {
    constexpr int H = N / 2;
    constexpr int RP = R + 1;                      // padded LDS stride: rows land in distinct banks
    constexpr int PER = H * R / kFftBlock;
    static_assert(PER * kFftBlock == H * R, "block shape");
    __shared__ float2 buf[H * RP];
    const long long row0 = (long long)blockIdx.x * R;
    const int rw = (int)min((long long)R, nrows - row0);
    // R consecutive rows are N*R contiguous floats: read them as float2 pairs
    const float2 *src = (const float2 *)(in + row0 * N);
    float2 v[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int t = threadIdx.x + p * kFftBlock;        // pair index within the block
        const int r = min(t / H, rw - 1);
        v[p] = src[(long long)r * H + t % H];
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int t = threadIdx.x + p * kFftBlock;
        const int r = t / H, j = t % H;
        buf[j * RP + r] = v[p];                           // column r = row r
    }
    __syncthreads();
    fft_columns<H, R, RP, kFftBlock>(buf, tw_half);
    float2 *dst = out + row0 * zs;                        // output rows zs >= H+1 apart
    for (int t = threadIdx.x; t < (H + 1) * R; t += kFftBlock) {
        const int r = t / (H + 1), k = t % (H + 1);
        if (r >= rw) continue;
        const float2 a = buf[(k % H) * RP + r];
        const float2 b = buf[((H - k) % H) * RP + r];     // Z[N/2 - k]
        const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
        // O = -i/2 (a - conj(b)) = ((a.y + b.y)/2, -(a.x - b.x)/2)
        const float2 o = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
        const float2 w = tw_full[k];                      // k <= N/2 < N
        dst[(long long)r * zs + k] = cadd(e, cmul(w, o));
    }
}

template <int N>
int launch_cols(Ctx *c, float2 *data, long long nstride, long long ostride, long long bstride,
                int zlen, int nother, int nbatch, const float2 *tw, unsigned *amax) {
    // 16 complex (128 B) contiguous per point; 64 KB of LDS at most
    constexpr int C = (N >= 1024) ? 8 : 16;
    constexpr int BLK = (N >= 512) ? 512 : 256;
    const long long nwg = (long long)ceil_div(zlen, C) * nother * nbatch;
    PYL_CHECK(nwg < (1ll << 31), PYL_EINVAL, "FFT column launch too large (%lld workgroups)", nwg);
    fft_col_kernel<N, C, BLK><<<(unsigned)nwg, BLK, 0, c->stream>>>(data, nstride, ostride, bstride,
                                                                     zlen, nother, tw, amax);
    PYL_HIP(hipGetLastError());
    return PYL_OK;
}

int cols(Ctx *c, int n, float2 *data, long long nstride, long long ostride, long long bstride,
         int zlen, int nother, int nbatch, const float2 *tw, unsigned *amax = nullptr) {
    switch (n) {
        case 16: return launch_cols<16>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
        case 32: return launch_cols<32>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
        case 64: return launch_cols<64>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
        case 128: return launch_cols<128>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
        case 256: return launch_cols<256>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
        case 512: return launch_cols<512>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
        case 1024: return launch_cols<1024>(c, data, nstride, ostride, bstride, zlen, nother, nbatch, tw, amax);
    }
    set_error("no column FFT of length %d", n);
    return PYL_EINVAL;
}

template <int N>
int launch_rows(Ctx *c, const float *in, float2 *out, long long nrows, const float2 *twh,
                const float2 *twf, long long zs) {
    constexpr int R = (N >= 512) ? 8 : 32;
    const long long nb = (nrows + R - 1) / R;
    fft_r2c_rows_kernel<N, R><<<(unsigned)nb, kFftBlock, 0, c->stream>>>(in, out, nrows, twh, twf, zs);
    PYL_HIP(hipGetLastError());
    return PYL_OK;
}

int rows(Ctx *c, int n, const float *in, float2 *out, long long nrows, const float2 *twh,
         const float2 *twf, long long zs) {
    switch (n) {
        case 16: return launch_rows<16>(c, in, out, nrows, twh, twf, zs);
        case 32: return launch_rows<32>(c, in, out, nrows, twh, twf, zs);
        case 64: return launch_rows<64>(c, in, out, nrows, twh, twf, zs);
        case 128: return launch_rows<128>(c, in, out, nrows, twh, twf, zs);
        case 256: return launch_rows<256>(c, in, out, nrows, twh, twf, zs);
        case 512: return launch_rows<512>(c, in, out, nrows, twh, twf, zs);
        case 1024: return launch_rows<1024>(c, in, out, nrows, twh, twf, zs);
    }
    set_error("no row FFT of length %d", n);
    return PYL_EINVAL;
}

}  // namespace

bool fft_own_size(int dims) {
    return dims >= 16 && dims <= 1024 && (dims & (dims - 1)) == 0;
}

// Twiddle table exp(-2 pi i k/N), k < N, resident per (device, N) for the process.
int get_twiddles(Ctx *c, int n, const float2 **out) {
    SharedTables &st = shared_tables();
    std::lock_guard<std::mutex> lk(st.mu);
    const auto key = std::make_pair(c->device, n);
    auto it = st.twiddles.find(key);
    if (it == st.twiddles.end()) {
        PYL_CHECK(!c->capturing, PYL_EINVAL, "FFT twiddles created while a graph is captured");
        std::vector<float2> h(n);
        for (int k = 0; k < n; ++k) {
            const double a = -2.0 * M_PI * (double)k / (double)n;
            h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        void *d = nullptr;
        PYL_HIP(hipMalloc(&d, sizeof(float2) * n));
        const hipError_t e = hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            set_error("FFT twiddle upload failed: %s", hipGetErrorString(e));
            return PYL_EHIP;
        }
        it = st.twiddles.emplace(key, d).first;
    }
    *out = (const float2 *)it->second;
    return PYL_OK;
}

// Row stride (complex elements) of the half-complex output the library keeps for
// itself: N/2+1 rounded up to 16 (128 bytes) for the own passes, so that every
// 16-column group of the column passes is one aligned 128-byte segment per row;
// the dense N/2+1 otherwise.
long long fft_zstride(int dims) {
    const long long z = dims / 2 + 1;
    return fft_own_size(dims) ? (z + 15) / 16 * 16 : z;
}

int complex_absmax(Ctx *c, const float2 *dk, int dims, long long zs, int ncube, long long cstride,
                   unsigned *amax) {
    if (zs <= 0) zs = dims / 2 + 1;
    const long long nrows = (long long)dims * dims;
    const int nbx = (int)std::max<long long>(64, std::min<long long>(1024, nrows * (dims / 2 + 1) / 4096));
    absmax_kernel<<<dim3(nbx, ncube), 256, 0, c->stream>>>(dk, nrows, dims / 2 + 1, zs, cstride, amax);
    PYL_HIP(hipGetLastError());
    return PYL_OK;
}

// out = r2c FFT of `batch` contiguous (N,N,N) float cubes, rows of N/2+1 modes zs
// apart (zs = 0: dense, N/2+1), N*N*zs per cube.  `in` may be overwritten.  amax
// (optional, 64 zeroed words per cube): the largest |Re| or |Im| of each cube's
// transform, float bits (see fft_col_kernel).
int fft_r2c_3d(Ctx *c, float *in, float2 *out, int dims, int batch, long long zs, unsigned *amax) {
    if (zs <= 0) zs = dims / 2 + 1;
    if (!fft_own_size(dims)) {
        PYL_CHECK(zs == dims / 2 + 1, PYL_EINVAL, "padded rows need the own FFT passes");
        PYL_CHECK(!c->capturing, PYL_EINVAL,
                  "graph capture takes the own FFT passes only (power-of-two sides 16..1024)");
        FFTPlan *plan = nullptr;
        PYL_TRY(get_fft_plan(c, dims, batch, &plan, FFT_R2C_3D));
        PYL_TRY(run_fft(c, plan, in, out));
        if (amax)
            PYL_TRY(complex_absmax(c, out, dims, zs, batch, (long long)dims * dims * zs, amax));
        return PYL_OK;
    }
    const long long N = dims, Z = dims / 2 + 1;
    const float2 *tw = nullptr, *twh = nullptr;
    PYL_TRY(get_twiddles(c, dims, &tw));
    PYL_TRY(get_twiddles(c, dims / 2, &twh));
    // z: real rows -> half-complex rows (zs apart)
    PYL_TRY(rows(c, dims, in, out, N * N * batch, twh, tw, zs));
    // y: n = y (stride zs), other = x (stride N*zs); Z columns of each row
    PYL_TRY(cols(c, dims, out, zs, N * zs, N * N * zs, (int)Z, dims, batch, tw));
    // x: n = x (stride N*zs), other = y (stride zs)
    PYL_TRY(cols(c, dims, out, N * zs, zs, N * N * zs, (int)Z, dims, batch, tw, amax));
    return PYL_OK;
}

}  // namespace pyl
