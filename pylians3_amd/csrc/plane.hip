// 2D spectra of maps (Pk_plane, XPk_plane) and the MAS deconvolution of a field
// (correct_MAS) on gfx950.
//
// Reference semantics: library/Pk_library/Pk_library.pyx:435-511 (Pk_plane),
// :1093-1220 (XPk_plane), :1882-1939 (correct_MAS), :102-114
// (check_number_modes_2D).
//
// Pk_plane binning is a gather: one workgroup per |k| bin b walks the rows kx
// that cross the annulus b <= |k| < b+1 and, per row, the exact range of ky
// in it (integer square roots), so each mode is visited once, no atomics are
// used and every bin is reduced in a fixed order.
#include "pyl_internal.h"

namespace pyl {
namespace {

constexpr int kPlaneBlock = 256;

// floor(sqrt(v)) for 0 <= v < 2^52, exact.
__device__ __forceinline__ long long isqrt_floor(long long v) {
    long long r = (long long)sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// Smallest r >= 0 with r*r >= v (v >= 0).
__device__ __forceinline__ long long isqrt_ceil(long long v) {
    long long r = isqrt_floor(v);
    return (r * r == v) ? r : r + 1;
}

// One workgroup per bin.  out[b*(2+S) + {0: sum k, 1: modes, 2..: spectra}].
// Mode rule and per-mode arithmetic of Pk_library.pyx:466-497 / :1155-1203:
// float MAS factor (double product of the axis corrections), complex64 scaling,
// double squares.
template <int NF>
__global__ void __launch_bounds__(kPlaneBlock)
pk_plane_kernel(const float2 *__restrict__ dk0, const float2 *__restrict__ dk1, int g,
                const double *__restrict__ c0, const double *__restrict__ c1,
                double *__restrict__ out) {
    constexpr int S = (NF == 1) ? 1 : 3;
    constexpr int NV = 2 + S;
    __shared__ double red[NV][kPlaneBlock];
    const int b = blockIdx.x;
    const int m = g / 2;
    const bool even = (g % 2) == 0;
    const int zlen = m + 1;
    const long long lo2 = (long long)b * b, hi2 = (long long)(b + 1) * (b + 1);
    const int kxmin = even ? -(m - 1) : -m;            // signed rows of an FFT of size g
    const int xa = max(kxmin, -(b + 1)), xb = min(m, b + 1);
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    for (int kx = xa + (int)threadIdx.x; kx <= xb; kx += blockDim.x) {
        const long long kx2 = (long long)kx * kx;
        if (kx2 >= hi2) continue;
        const int kxx = kx < 0 ? kx + g : kx;
        long long ky0 = isqrt_ceil(lo2 > kx2 ? lo2 - kx2 : 0);
        long long ky1 = isqrt_floor(hi2 - kx2 - 1);
        if (ky1 > m) ky1 = m;
        const float2 *row0 = dk0 + (long long)kxx * zlen;
        const float2 *row1 = (NF == 2) ? dk1 + (long long)kxx * zlen : nullptr;
        const double cx0 = c0[kxx];
        const double cx1 = (NF == 2) ? c1[kxx] : 0.0;
        for (long long ky = ky0; ky <= ky1; ++ky) {
            if ((ky == 0 || (ky == m && even)) && kx < 0) continue;
            const double k = sqrt((double)(kx2 + ky * ky));
            const float mf0 = (float)(cx0 * c0[ky]);
            const float2 a = row0[ky];
            const double r0 = (double)(a.x * mf0), i0 = (double)(a.y * mf0);
            acc[0] += k;
            acc[1] += 1.0;
            acc[2] += r0 * r0 + i0 * i0;
            if constexpr (NF == 2) {
                const float mf1 = (float)(cx1 * c1[ky]);
                const float2 bb = row1[ky];
                const double r1 = (double)(bb.x * mf1), i1 = (double)(bb.y * mf1);
                acc[3] += r1 * r1 + i1 * i1;
                acc[4] += r0 * r1 + i0 * i1;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = kPlaneBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int v = 0; v < NV; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < NV) out[(long long)b * NV + threadIdx.x] = red[threadIdx.x][0];
}

// correct_MAS mode scaling (Pk_library.pyx:1910-1929): independent modes times
// the float MAS factor in complex64; the redundant half of the kz = 0 and
// kz = middle planes is skipped.
__global__ void __launch_bounds__(256)
mas_scale_kernel(float2 *__restrict__ dk, int dims, const double *__restrict__ corr) {
    const int m = dims / 2;
    const bool even = (dims % 2) == 0;
    const long long zlen = m + 1;
    const long long total = (long long)dims * dims * zlen;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const int kzz = (int)(e % zlen);
        const int kyy = (int)((e / zlen) % dims);
        const int kxx = (int)(e / (zlen * dims));
        const int kx = kxx > m ? kxx - dims : kxx;
        const int ky = kyy > m ? kyy - dims : kyy;
        const int kz = kzz;
        if (kz == 0 || (kz == m && even)) {
            if (kx < 0) continue;
            if ((kx == 0 || (kx == m && even)) && ky < 0) continue;
        }
        const float mf = (float)(corr[kxx] * corr[kyy] * corr[kzz]);
        float2 v = dk[e];
        v.x = v.x * mf;
        v.y = v.y * mf;
        dk[e] = v;
    }
}

// Hermitian part of the kz = z0 plane: H(k) = (X(k) + conj(X(-k)))/2, written to
// both k and -k (one thread per pair).  This is what a real inverse transform
// reads from a plane whose two halves disagree.
__global__ void __launch_bounds__(256)
hermitian_plane_kernel(float2 *__restrict__ dk, int dims, int kzz) {
    const long long zlen = dims / 2 + 1;
    const long long n = (long long)dims * dims;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const int kxx = (int)(t / dims), kyy = (int)(t % dims);
        const int pxx = (dims - kxx) % dims, pyy = (dims - kyy) % dims;
        const long long p = (long long)pxx * dims + pyy;
        if (p < t) continue;                       // the pair is handled by its lower index
        float2 *a = dk + t * zlen + kzz;
        float2 *b = dk + p * zlen + kzz;
        const float2 x = *a, y = *b;
        float2 h;
        h.x = (x.x + y.x) * 0.5f;
        h.y = (x.y - y.y) * 0.5f;
        *a = h;
        if (p != t) *b = make_float2(h.x, -h.y);
        else a->y = 0.0f;
    }
}

int plane_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (int)std::max<long long>(1, std::min<long long>(b, 65536));
}

}  // namespace
}  // namespace pyl

using namespace pyl;

extern "C" {

int pyl_pk_plane_nbins(int grid) {
    if (grid < 2) return 0;
    const int m = grid / 2;
    return (int)std::sqrt((double)(2 * m * m));   // kmax; the DC bin is dropped
}

int pyl_pk_plane(pyl_ctx *ctx, const float *const *deltas, int nfields, const int *mas, int grid,
                 double BoxSize, int flags, double *k, double *Nmodes, double *Pk) {
    TraceRange trace_("pyl_pk_plane");
    PYL_CHECK(ctx && deltas && mas && k && Nmodes && Pk, PYL_EINVAL, "NULL argument");
    PYL_CHECK(nfields == 1 || nfields == 2, PYL_EINVAL, "nfields must be 1 or 2");
    PYL_CHECK(grid >= 2 && grid <= (1 << 16) && BoxSize > 0.0, PYL_EINVAL, "bad arguments");
    for (int f = 0; f < nfields; ++f) {
        PYL_CHECK(deltas[f] != nullptr, PYL_EINVAL, "field %d is NULL", f);
        PYL_CHECK(mas[f] >= 0 && mas[f] <= 4, PYL_EINVAL, "bad MAS id %d", mas[f]);
    }
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    const int m = grid / 2;
    const int kmax = (int)std::sqrt((double)(2 * m * m));
    const long long n = (long long)grid * grid;
    const long long nk = (long long)grid * (m + 1);
    const int S = (nfields == 1) ? 1 : 3;
    const int NV = 2 + S;
    const bool dev = (flags & PYL_PTR_DEVICE) != 0;
    PYL_TRY(ensure(c, BUF_FFT_IN, (size_t)n * nfields * sizeof(float)));
    PYL_TRY(ensure(c, BUF_FFT_OUT, (size_t)nk * nfields * sizeof(float2)));
    PYL_TRY(ensure(c, BUF_BINS, sizeof(double) * ((size_t)(kmax + 1) * NV + 2 * (size_t)grid)));
    float *in = (float *)c->bufs[BUF_FFT_IN].ptr;
    float2 *dk = (float2 *)c->bufs[BUF_FFT_OUT].ptr;
    double *bins = (double *)c->bufs[BUF_BINS].ptr;
    double *dtab = bins + (size_t)(kmax + 1) * NV;
    std::vector<double> tabs((size_t)2 * grid, 1.0);
    for (int f = 0; f < nfields; ++f) mas_table(grid, mas[f], &tabs[(size_t)f * grid]);
    PYL_HIP(hipMemcpyAsync(dtab, tabs.data(), tabs.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_PREP));
    for (int f = 0; f < nfields; ++f)
        PYL_HIP(hipMemcpyAsync(in + (size_t)f * n, deltas[f], (size_t)n * sizeof(float),
                               dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    PYL_TRY(stage_end(c, PYL_STAGE_PK_PREP));
    FFTPlan *plan = nullptr;
    PYL_TRY(get_fft_plan(c, grid, nfields, &plan, FFT_R2C_2D));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_FFT));
    PYL_TRY(run_fft(c, plan, in, dk));
    PYL_TRY(stage_end(c, PYL_STAGE_PK_FFT));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_BIN));
    if (nfields == 1)
        pk_plane_kernel<1><<<kmax + 1, kPlaneBlock, 0, c->stream>>>(dk, nullptr, grid, dtab, nullptr,
                                                                    bins);
    else
        pk_plane_kernel<2><<<kmax + 1, kPlaneBlock, 0, c->stream>>>(dk, dk + nk, grid, dtab,
                                                                    dtab + grid, bins);
    PYL_HIP(hipGetLastError());
    PYL_TRY(stage_end(c, PYL_STAGE_PK_BIN));
    std::vector<double> h((size_t)(kmax + 1) * NV);
    PYL_HIP(hipMemcpyAsync(h.data(), bins, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    PYL_HIP(hipStreamSynchronize(c->stream));
    if (nfields == 1) {
        // check_number_modes_2D (Pk_library.pyx:102-114)
        double sumN = 0.0;
        for (int b = 0; b <= kmax; ++b) sumN += h[(size_t)b * NV + 1];
        const long long own = (grid % 2 == 1) ? 1 : 4;
        const long long expected = ((long long)grid * grid - own) / 2 + own;
        PYL_CHECK((long long)sumN == expected, PYL_EMODES,
                  "WARNING: Not all modes counted (counted %lld, expected %lld)", (long long)sumN,
                  expected);
    }
    // Pk_library.pyx:499-505 / :1190-1203: drop the DC bin, mean k, (Pk/N)*(L/grid^2)^2
    const double kF = 2.0 * M_PI / BoxSize;
    const double q = BoxSize / ((double)grid * (double)grid);
    const double fact = q * q;
    for (int b = 1; b <= kmax; ++b) {
        const double *v = &h[(size_t)b * NV];
        const double N = v[1];
        k[b - 1] = (v[0] / N) * kF;
        Nmodes[b - 1] = N;
        for (int s = 0; s < S; ++s) Pk[(size_t)(b - 1) * S + s] = (v[2 + s] / N) * fact;
    }
    return PYL_OK;
}

int pyl_correct_mas(pyl_ctx *ctx, const float *delta, float *out, int dims, double BoxSize, int mas,
                    int flags) {
    TraceRange trace_("pyl_correct_mas");
    PYL_CHECK(ctx && delta && out, PYL_EINVAL, "NULL argument");
    PYL_CHECK(dims >= 2 && dims <= 4096 && BoxSize > 0.0, PYL_EINVAL, "bad arguments");
    PYL_CHECK(mas >= 0 && mas <= 4, PYL_EINVAL, "bad MAS id %d", mas);
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    const int m = dims / 2;
    const long long n = (long long)dims * dims * dims;
    const long long nk = (long long)dims * dims * (m + 1);
    const bool dev = (flags & PYL_PTR_DEVICE) != 0;
    PYL_TRY(ensure(c, BUF_FFT_IN, (size_t)n * sizeof(float)));
    PYL_TRY(ensure(c, BUF_FFT_OUT, (size_t)nk * sizeof(float2)));
    PYL_TRY(ensure(c, BUF_MAS_TAB, (size_t)dims * sizeof(double)));
    float *in = (float *)c->bufs[BUF_FFT_IN].ptr;
    float2 *dk = (float2 *)c->bufs[BUF_FFT_OUT].ptr;
    double *dtab = (double *)c->bufs[BUF_MAS_TAB].ptr;
    std::vector<double> tab(dims);
    mas_table(dims, mas, tab.data());
    PYL_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    PYL_HIP(hipMemcpyAsync(in, delta, (size_t)n * sizeof(float),
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    FFTPlan *inv = nullptr;
    PYL_TRY(get_fft_plan(c, dims, 1, &inv, FFT_C2R_3D));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_FFT));
    PYL_TRY(fft_r2c_3d(c, in, dk, dims, 1));
    PYL_TRY(stage_end(c, PYL_STAGE_PK_FFT));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_BIN));
    mas_scale_kernel<<<plane_blocks(nk), 256, 0, c->stream>>>(dk, dims, dtab);
    hermitian_plane_kernel<<<plane_blocks((long long)dims * dims), 256, 0, c->stream>>>(dk, dims, 0);
    if (dims % 2 == 0)
        hermitian_plane_kernel<<<plane_blocks((long long)dims * dims), 256, 0, c->stream>>>(dk, dims,
                                                                                           m);
    PYL_HIP(hipGetLastError());
    PYL_TRY(stage_end(c, PYL_STAGE_PK_BIN));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_FFT));
    PYL_TRY(run_fft(c, inv, dk, in));
    PYL_TRY(stage_end(c, PYL_STAGE_PK_FFT));
    PYL_HIP(hipMemcpyAsync(out, in, (size_t)n * sizeof(float),
                           dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    return finish(c, flags);
}

}  // extern "C"
