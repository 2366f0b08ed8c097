// Power spectra on gfx950: rocFFT r2c + one fused binning pass.
//
// Reference semantics: library/Pk_library/Pk_library.pyx:263-420 (Pk) and
// :529-793 (XPk).  Per mode of the half-complex FFT output the binning pass
// does, in one read of delta_k: the half-space skip rule (:324-327), integer
// bin assignment (k_index = isqrt(k^2), k_per = isqrt(k_perp^2), exact for the
// reference's (int)sqrt of integer squares), MAS deconvolution with a float
// factor equal to the double product of the per-axis corrections (:351-352),
// |delta_k|^2 in double (:355-357), mu-Legendre weights for l = 0, 2, 4 and the
// phase (:358), and the 1D / 2D / 3D accumulations (:361-378).
//
// Work decomposition.  Rows (kxx, kyy) of the half-complex cube are grouped by
// a key that fixes the 2D bins a row can reach: |kx| (axis 0), |ky| (axis 1)
// or k_per = isqrt(kx^2+ky^2) (axis 2).  A work item is up to R rows with one
// key; one workgroup per item accumulates into LDS: the 3D bins, one "line" of
// 2D bins (all k_per at fixed k_par for axes 0/1, all k_par at fixed k_per for
// axis 2) and the 1D bins.  Each lane handles runs of 4 consecutive kz and
// merges equal bins in registers before its LDS atomics.  Items write their
// LDS partials with plain stores; a reduce kernel sums them in item order, so
// the result is deterministic and no global atomics are used.
//
// Mode counts and k sums do not depend on the data: they are computed exactly
// on the host once per (dims, axis) from Hermitian-pair counting and cached.
#include "pyl_internal.h"
#include "fixsum.h"

#include <array>

#include <algorithm>
#include <cmath>
#include <thread>

namespace pyl {

// ------------------------------------------------------------------ mode plan
// Independent modes = one of each conjugate pair plus the self-conjugate modes,
// so per bin: N_indep = (N_all + N_self) / 2, where N_all counts the full
// dims^3 cube and N_self the modes with every component in {0, dims/2}.
// Every binned quantity depends only on (|kx|, |ky|, |kz|), so the cube is
// walked over absolute values with multiplicities.
void compute_mode_plan(int dims, int axis, ModePlan *mp) {
    const int m = dims / 2;
    const bool even = dims % 2 == 0;
    const int kmax_par = m;
    const int kmax_per = (int)std::sqrt((double)(m * m + m * m));
    const int kmax = (int)std::sqrt((double)(3 * m * m));
    const long long nb2 = (long long)(kmax_par + 1) * (kmax_per + 1);
    const long long nk2 = 3LL * m * m + 1;
    auto mult = [&](int a) -> long long { return (a == 0 || (even && a == m)) ? 1 : 2; };
    auto isq = [](long long v) -> int {
        long long s = (long long)std::sqrt((double)v);
        while (s * s > v) --s;
        while ((s + 1) * (s + 1) <= v) ++s;
        return (int)s;
    };
    std::vector<int> isqrt_tab(2LL * m * m + 1);
    for (long long v = 0; v < (long long)isqrt_tab.size(); ++v) isqrt_tab[v] = isq(v);

    unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    if (m < 32) nth = 1;
    struct Part {
        std::vector<long long> c3, c2, c1;
    };
    std::vector<Part> parts(nth);
    auto work = [&](unsigned t) {
        Part &P = parts[t];
        P.c3.assign(nk2, 0);
        P.c2.assign(nb2, 0);
        P.c1.assign(m + 1, 0);
        for (int a = (int)t; a <= m; a += (int)nth) {
            const long long ma = mult(a);
            for (int b = 0; b <= m; ++b) {
                const long long mab = ma * mult(b);
                for (int c = 0; c <= m; ++c) {
                    const long long w = mab * mult(c);
                    const long long k2 = (long long)a * a + (long long)b * b + (long long)c * c;
                    int kpar, kper;
                    if (axis == 0) { kpar = a; kper = isqrt_tab[b * b + c * c]; }
                    else if (axis == 1) { kpar = b; kper = isqrt_tab[a * a + c * c]; }
                    else { kpar = c; kper = isqrt_tab[a * a + b * b]; }
                    P.c3[k2] += w;
                    P.c2[(long long)(kmax_par + 1) * kper + kpar] += w;
                    if (k2 <= (long long)m * m) P.c1[kpar] += w;
                }
            }
        }
    };
    if (nth == 1) work(0);
    else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; ++t) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    std::vector<long long> c3(nk2, 0), c2(nb2, 0), c1(m + 1, 0);
    for (auto &P : parts) {
        for (long long i = 0; i < nk2; ++i) c3[i] += P.c3[i];
        for (long long i = 0; i < nb2; ++i) c2[i] += P.c2[i];
        for (int i = 0; i <= m; ++i) c1[i] += P.c1[i];
    }
    // self-conjugate modes: components in {0} (odd) or {0, m} (even)
    std::vector<int> vals = even ? std::vector<int>{0, m} : std::vector<int>{0};
    if (even && m == 0) vals = {0};
    for (int a : vals)
        for (int b : vals)
            for (int c : vals) {
                const long long k2 = (long long)a * a + (long long)b * b + (long long)c * c;
                int kpar, kper;
                if (axis == 0) { kpar = a; kper = isq((long long)b * b + (long long)c * c); }
                else if (axis == 1) { kpar = b; kper = isq((long long)a * a + (long long)c * c); }
                else { kpar = c; kper = isq((long long)a * a + (long long)b * b); }
                c3[k2] += 1;
                c2[(long long)(kmax_par + 1) * kper + kpar] += 1;
                if (k2 <= (long long)m * m) c1[kpar] += 1;
            }
    mp->dims = dims;
    mp->axis = axis;
    mp->N3D.assign(kmax + 1, 0.0);
    mp->k3D.assign(kmax + 1, 0.0);
    for (long long k2 = 0; k2 < nk2; ++k2) {
        long long n = c3[k2] / 2;  // (N_all + N_self) / 2, always an integer
        if (!n) continue;
        int b = isq(k2);
        mp->N3D[b] += (double)n;
        mp->k3D[b] += (double)n * std::sqrt((double)k2);
    }
    mp->N2D.resize(nb2);
    for (long long i = 0; i < nb2; ++i) mp->N2D[i] = (double)(c2[i] / 2);
    mp->N1D.resize(m + 1);
    mp->k1D.resize(m + 1);
    for (int i = 0; i <= m; ++i) {
        mp->N1D[i] = (double)(c1[i] / 2);
        mp->k1D[i] = (double)i * mp->N1D[i];
    }
}

const ModePlan *get_mode_plan(Ctx *, int dims, int axis) {
    SharedTables &st = shared_tables();
    std::lock_guard<std::mutex> lk(st.mu);
    const auto key = std::make_pair(dims, axis);
    auto it = st.mode.find(key);
    if (it != st.mode.end()) return &it->second;
    ModePlan mp;
    compute_mode_plan(dims, axis, &mp);
    return &st.mode.emplace(key, std::move(mp)).first->second;
}

namespace {

constexpr int kRun = 4;          // consecutive kz per lane
constexpr int kBinBlock = 512;

struct PkDims {
    int dims, m, zlen, zpad, even, kmax, kmax_par, kmax_per;
    long long nb2;
};

PkDims pk_dims(int dims) {
    PkDims p;
    p.dims = dims;
    p.m = dims / 2;
    p.zlen = p.m + 1;
    p.zpad = (p.zlen + kRun - 1) / kRun * kRun;
    p.even = dims % 2 == 0;
    p.kmax_par = p.m;
    p.kmax_per = (int)std::sqrt((double)(p.m * p.m + p.m * p.m));
    p.kmax = (int)std::sqrt((double)(3 * p.m * p.m));
    p.nb2 = (long long)(p.kmax_par + 1) * (p.kmax_per + 1);
    return p;
}

int row_key(int kxx, int kyy, int dims, int axis) {
    const int m = dims / 2;
    int kx = kxx > m ? kxx - dims : kxx;
    int ky = kyy > m ? kyy - dims : kyy;
    if (axis == 0) return std::abs(kx);
    if (axis == 1) return std::abs(ky);
    long long v = (long long)kx * kx + (long long)ky * ky;
    long long s = (long long)std::sqrt((double)v);
    while (s * s > v) --s;
    while ((s + 1) * (s + 1) <= v) ++s;
    return (int)s;
}

int get_bin_plan(Ctx *c, int dims, int axis, int nreal, BinPlan **out) {
    // rows per item: ~2048 items for one large grid; for a batch, ~16K items in total
    // so that the partials stay small
    const long long nrows_total = (long long)dims * dims;
    const long long target = nreal > 1 ? std::max(1LL, 16384LL / nreal) : 2048LL;
    int R = c->tune_bin_rows > 0 ? std::min(256, c->tune_bin_rows)
                                 : (int)std::max(4LL, std::min(256LL, nrows_total / target));
    // an exact bin sum's low word takes at most kFixMaxAdds additions: one per run of
    // kRun modes of the item's rows
    const int runs_per_row = (pk_dims(dims).zpad) / kRun;
    R = std::max(1, std::min(R, kFixMaxAdds / runs_per_row));
    SharedTables &st = shared_tables();
    std::lock_guard<std::mutex> lk(st.mu);
    const auto key = std::make_tuple(c->device, dims, axis, R);
    auto it = st.bin.find(key);
    if (it != st.bin.end()) {
        *out = &it->second;
        return PYL_OK;
    }
    PYL_CHECK(!c->capturing, PYL_EINVAL, "binning plan created while a graph is captured: run the "
              "same work once on the context before capturing it");
    PkDims p = pk_dims(dims);
    const int nkeys = (axis == 2) ? p.kmax_per + 1 : p.m + 1;
    std::vector<std::vector<int>> by_key(nkeys);
    for (int kxx = 0; kxx < dims; ++kxx)
        for (int kyy = 0; kyy < dims; ++kyy)
            by_key[row_key(kxx, kyy, dims, axis)].push_back(kxx * dims + kyy);
    std::vector<int> rows;
    rows.reserve(nrows_total);
    std::vector<int4> items;
    std::vector<int> key_items(nkeys + 1, 0);
    for (int k = 0; k < nkeys; ++k) {
        key_items[k] = (int)items.size();
        const auto &v = by_key[k];
        for (size_t s = 0; s < v.size(); s += R) {
            int n = (int)std::min<size_t>(R, v.size() - s);
            items.push_back(make_int4((int)rows.size() + (int)s, n, k, 0));
        }
        rows.insert(rows.end(), v.begin(), v.end());
    }
    key_items[nkeys] = (int)items.size();
    BinPlan bp;
    bp.dims = dims;
    bp.axis = axis;
    bp.R = R;
    bp.nitems = (int)items.size();
    bp.nkeys = nkeys;
    auto fail = [&](hipError_t e) {
        (void)hipFree(bp.d_rows);
        (void)hipFree(bp.d_items);
        (void)hipFree(bp.d_key_items);
        set_error("binning plan upload failed: %s", hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? PYL_ENOMEM : PYL_EHIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&bp.d_rows, rows.size() * sizeof(int))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&bp.d_items, items.size() * sizeof(int4))) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&bp.d_key_items, key_items.size() * sizeof(int))) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(bp.d_rows, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return fail(e);
    if ((e = hipMemcpy(bp.d_items, items.data(), items.size() * sizeof(int4),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e);
    if ((e = hipMemcpy(bp.d_key_items, key_items.data(), key_items.size() * sizeof(int),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e);
    *out = &st.bin.emplace(key, bp).first->second;
    return PYL_OK;
}

// ------------------------------------------------------------------ kernels
// Spectra computed by one launch: NFL = 1 -> {auto a}; NFL = 2 -> {auto a, auto b, cross ab}.
template <int NFL> struct Spec { static constexpr int S = NFL == 1 ? 1 : 3; };

struct BinArgs {
    const float2 *dk[2];     // half-complex fields
    const double *corr[2];   // MAS correction tables [dims] per field
    float cmax[2];           // largest product of three MAS corrections per field
    const unsigned *amax[2]; // per field [realisation][64]: largest |Re|, |Im| (float bits)
    int *escale;             // [realisation][2]: E of the power and the phase sums
    int nreal;
    const int *rows;
    const int4 *items;
    unsigned long long *plo; // [nreal][nitems][P] low words of the item sums
    int *phi;                // [nreal][nitems][P] high words
    long long fstride;       // elements between realisations in dk[] (blockIdx.y = realisation)
    long long zs;            // elements between rows (kx, ky) of dk[] (>= zlen)
    int nitems;
    int dims, m, zlen, zpad, even, kmax, L2, L1;
    int P, off2, off1;       // partial layout: acc3 at 0, line2 at off2, line1 at off1
    int rmax;                // max rows per item (LDS row table size)
};

// floor(sqrt(v)) for 0 <= v < 2^24: the hardware square root (within 1 ulp)
// lands within one of the answer, and the two integer tests correct it.
__device__ __forceinline__ int isqrt_i(int v) {
    int s = (int)__builtin_amdgcn_sqrtf((float)v);
    if (s * s > v) --s;
    if ((s + 1) * (s + 1) <= v) ++s;
    return s;
}

// atan2(a, b) on the half-plane b >= |a| (the phase of Pk_library.pyx:358,
// atan2(real, |delta|)) for b in [0.5, 2): the caller scales both components by a
// power of two so that the larger has magnitude in [0.5, 1).  t = a/b (hardware
// reciprocal, 1 ulp) lies in [-1, 1] and atan(t) is t*P(t^2), P of degree 7 fitted
// to atan(t)/t on [0, 1] by iteratively reweighted least squares on Chebyshev nodes
// (max error 1.7e-7 in float, 2.3e-7 relative).
__device__ __forceinline__ float atan_halfplane(float a, float b) {
    float t = a * __builtin_amdgcn_rcpf(b);
    t = fminf(fmaxf(t, -1.0f), 1.0f);       // |a| <= b: only rounding can leave [-1, 1]
    const float u = t * t;
    float p = -0.00478044f;
    p = p * u + 0.02455707f;
    p = p * u - 0.05990465f;
    p = p * u + 0.09942755f;
    p = p * u - 0.14029419f;
    p = p * u + 0.19971375f;
    p = p * u - 0.33332094f;
    p = p * u + 0.99999991f;
    return t * p;
}

template <int AXIS, int NFL, bool PHASE>
__global__ void __launch_bounds__(kBinBlock) pk_bin_kernel(BinArgs a) {
    constexpr int S = Spec<NFL>::S;
    constexpr int NA3 = 3 * S + (PHASE ? 1 : 0);   // per 3D bin: P0,P2,P4 per spectrum [+phase]
    extern __shared__ double lds[];
    // The P accumulators, fixed point: low words alo[P], high words ahi[P].  acc3
    // (indices [0, n3)) is [NA3][kmax+1] in LDS (consecutive bins in consecutive
    // banks: lanes flushing neighbouring bins do not collide), [kmax+1][NA3] in the
    // partials; the 2D line at off2, the 1D line at off1.
    unsigned long long *alo = (unsigned long long *)lds;
    int *ahi = (int *)(alo + a.P);
    const int K1 = a.kmax + 1;
    double *corrz = (double *)(ahi + ((a.P + 1) & ~1));  // [zlen*NFL]
    double *rowc = corrz + a.zlen * NFL;              // [rmax*NFL]
    int *rowi = (int *)(rowc + a.rmax * NFL);         // [rmax*3]: kx, ky, linear row
    __shared__ int sE[2];

    const int4 item = a.items[blockIdx.x];
    const int nrows = item.y;
    const int key = item.z;
    const int tid = threadIdx.x;
    const int n3 = (a.kmax + 1) * NA3;
    for (int i = tid; i < a.P; i += blockDim.x) {
        alo[i] = 0ull;
        ahi[i] = 0;
    }
    if (tid == 0) {
        // the realisation's scale: the same in every work item
        const double nm = (double)a.dims * a.dims * a.zlen;
        double bound = 0.0;
        for (int f = 0; f < NFL; ++f) {
            const unsigned *w = a.amax[f] + (long long)blockIdx.y * 64;
            float m = 0.0f;
            for (int j = 0; j < 64; ++j) m = fmaxf(m, __uint_as_float(w[j]));
            const double am = (double)m * (double)a.cmax[f];
            bound = fmax(bound, 2.0 * nm * am * am);
        }
        sE[0] = fix_exponent(bound * 1.01);
        sE[1] = fix_exponent(nm);
        if (blockIdx.x == 0) {
            a.escale[2 * blockIdx.y] = sE[0];
            a.escale[2 * blockIdx.y + 1] = sE[1];
        }
    }
    for (int i = tid; i < a.zlen * NFL; i += blockDim.x) {
        int f = i / a.zlen, kz = i - f * a.zlen;
        corrz[i] = a.corr[f][kz];
    }
    for (int r = tid; r < nrows; r += blockDim.x) {
        int lin = a.rows[item.x + r];
        int kxx = lin / a.dims, kyy = lin - kxx * a.dims;
        int kx = kxx > a.m ? kxx - a.dims : kxx;
        int ky = kyy > a.m ? kyy - a.dims : kyy;
        rowi[3 * r + 0] = kx;
        rowi[3 * r + 1] = ky;
        rowi[3 * r + 2] = lin;
        for (int f = 0; f < NFL; ++f) rowc[r * NFL + f] = a.corr[f][kxx] * a.corr[f][kyy];
    }
    __syncthreads();
    const int Ep = sE[0], Eph = sE[1];

    const int mm = a.m * a.m;
    // 1D accumulator (axes 0/1: k_par = key): each run's sum in fixed point, added
    // up per lane in integers, so the result does not depend on which lane took which
    // run (nor on how the rows were cut into work items)
    unsigned long long p1lo[S];
    long long p1hi[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        p1lo[s] = 0ull;
        p1hi[s] = 0;
    }

    // Lane runs start at flat index tid*kRun and advance by blockDim*kRun over the
    // item's rows x zpad (zpad is a multiple of kRun, so a run never crosses a row).
    const int step = blockDim.x * kRun;
    const int dr = step / a.zpad, dc = step - dr * a.zpad;
    int r = (tid * kRun) / a.zpad;
    int kz0 = tid * kRun - r * a.zpad;
    for (; r < nrows; r += dr, kz0 += dc) {
        if (kz0 >= a.zpad) {
            kz0 -= a.zpad;
            ++r;
            if (r >= nrows) break;
        }
        const int kx = rowi[3 * r + 0], ky = rowi[3 * r + 1];
        const long long rowoff = (long long)blockIdx.y * a.fstride + (long long)rowi[3 * r + 2] * a.zs;
        const int kxy2 = kx * kx + ky * ky;
        const bool kx_neg = kx < 0;
        const bool kx_edge = (kx == 0) || (a.even && kx == a.m);
        const bool ky_neg = ky < 0;
        double c01[NFL];
#pragma unroll
        for (int f = 0; f < NFL; ++f) c01[f] = rowc[r * NFL + f];

        float2 v[NFL][kRun];
#pragma unroll
        for (int f = 0; f < NFL; ++f)
#pragma unroll
            for (int s = 0; s < kRun; ++s)
                v[f][s] = (kz0 + s < a.zlen) ? a.dk[f][rowoff + kz0 + s] : make_float2(0.f, 0.f);

        int cur3 = -1, cur2 = -1;
        double r3[NA3], r2[S], p1[S];
#pragma unroll
        for (int i = 0; i < NA3; ++i) r3[i] = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            r2[s] = 0.0;
            p1[s] = 0.0;
        }

        // |k|^2 and k_per^2 of the run's first mode; along the run kz grows by one,
        // so each square grows by 2kz-1 <= 2*isqrt+1 and each integer root by at
        // most one: one test per mode instead of a square root
        int k2 = kxy2 + kz0 * kz0;
        int kidx = isqrt_i(k2);
        int p2 = (AXIS == 0 ? ky * ky : kx * kx) + kz0 * kz0;
        int kpr = (AXIS == 2) ? key : isqrt_i(p2);
#pragma unroll
        for (int s = 0; s < kRun; ++s) {
            const int kz = kz0 + s;
            if (s > 0) {
                k2 += 2 * kz - 1;
                kidx += ((kidx + 1) * (kidx + 1) <= k2) ? 1 : 0;
                if (AXIS != 2) {
                    p2 += 2 * kz - 1;
                    kpr += ((kpr + 1) * (kpr + 1) <= p2) ? 1 : 0;
                }
            }
            bool ok = kz < a.zlen;
            if (kz == 0 || (a.even && kz == a.m)) {     // Pk_library.pyx:324-327
                if (kx_neg) ok = false;
                else if (kx_edge && ky_neg) ok = false;
            }
            if (!ok) continue;
            int kpar;
            const int kper = kpr;
            if (AXIS == 0) kpar = kx < 0 ? -kx : kx;
            else if (AXIS == 1) kpar = ky < 0 ? -ky : ky;
            else kpar = kz;
            // mu^2 = k_par^2 / k^2 in float (hardware reciprocal, within 2 ulp)
            const float mu2 = (k2 == 0) ? 0.0f : (float)(kpar * kpar) * __builtin_amdgcn_rcpf((float)k2);
            const float L2 = 1.5f * mu2 - 0.5f;
            const float L4 = (35.0f * mu2 * mu2 - 30.0f * mu2 + 3.0f) * 0.125f;

            double re[NFL], im[NFL];
#pragma unroll
            for (int f = 0; f < NFL; ++f) {
                const float mf = (float)(c01[f] * corrz[f * a.zlen + kz]);
                const float xr = v[f][s].x * mf, xi = v[f][s].y * mf;   // complex64 * float
                re[f] = (double)xr;
                im[f] = (double)xi;
            }
            double d2[S];
            // the products of floats are exact in double: one rounding, as re*re + im*im
            d2[0] = fma(re[0], re[0], im[0] * im[0]);
            if (NFL == 2) {
                d2[1] = fma(re[1], re[1], im[1] * im[1]);
                d2[2] = fma(re[0], re[1], im[0] * im[1]);
            }
            // 3D: merge equal bins in registers, flush on change
            if (kidx != cur3) {
                if (cur3 >= 0) {
#pragma unroll
                    for (int i = 0; i < NA3; ++i)
                        fix_add(alo, ahi, i * K1 + cur3, r3[i], (PHASE && i == NA3 - 1) ? Eph : Ep);
                }
                cur3 = kidx;
#pragma unroll
                for (int i = 0; i < NA3; ++i) r3[i] = 0.0;
            }
#pragma unroll
            for (int sp = 0; sp < S; ++sp) {
                r3[3 * sp + 0] += d2[sp];
                r3[3 * sp + 1] += d2[sp] * (double)L2;
                r3[3 * sp + 2] += d2[sp] * (double)L4;
            }
            if (PHASE) {
                // atan2(Re, |delta|) from the float components scaled by a power of two
                // near 1/max(|Re|, |Im|) (exact): |delta|^2 formed in float would
                // overflow or underflow for |delta| beyond ~1e19 or below ~1e-19
                float ph = 0.0f;
                const float xr = (float)re[0], xi = (float)im[0];
                const float mx = fmaxf(fabsf(xr), fabsf(xi));
                if (mx > 0.0f) {
                    const int e = __builtin_amdgcn_frexp_expf(mx);
                    const float xs = __builtin_ldexpf(xr, -e), ys = __builtin_ldexpf(xi, -e);
                    ph = atan_halfplane(xs, __builtin_amdgcn_sqrtf(xs * xs + ys * ys));
                }
                r3[NA3 - 1] += (double)(ph * ph);
            }
            // 2D line: axes 0/1 index k_per (slowly varying), axis 2 index kz
            const int li = (AXIS == 2) ? kz : kper;
            if (AXIS == 2) {
#pragma unroll
                for (int sp = 0; sp < S; ++sp) fix_add(alo, ahi, a.off2 + li * S + sp, d2[sp], Ep);
            } else {
                if (li != cur2) {
                    if (cur2 >= 0) {
#pragma unroll
                        for (int sp = 0; sp < S; ++sp) fix_add(alo, ahi, a.off2 + cur2 * S + sp, r2[sp], Ep);
                    }
                    cur2 = li;
#pragma unroll
                    for (int sp = 0; sp < S; ++sp) r2[sp] = 0.0;
                }
#pragma unroll
                for (int sp = 0; sp < S; ++sp) r2[sp] += d2[sp];
            }
            // 1D: modes with k <= middle
            if (k2 <= mm) {
                if (AXIS == 2) {
#pragma unroll
                    for (int sp = 0; sp < S; ++sp) fix_add(alo, ahi, a.off1 + kz * S + sp, d2[sp], Ep);
                } else {
#pragma unroll
                    for (int sp = 0; sp < S; ++sp) p1[sp] += d2[sp];
                }
            }
        }
        if (cur3 >= 0) {
#pragma unroll
            for (int i = 0; i < NA3; ++i)
                fix_add(alo, ahi, i * K1 + cur3, r3[i], (PHASE && i == NA3 - 1) ? Eph : Ep);
        }
        if (AXIS != 2 && cur2 >= 0) {
#pragma unroll
            for (int sp = 0; sp < S; ++sp) fix_add(alo, ahi, a.off2 + cur2 * S + sp, r2[sp], Ep);
        }
        if (AXIS != 2) {
#pragma unroll
            for (int sp = 0; sp < S; ++sp) {
                if (p1[sp] == 0.0) continue;
                unsigned long long lo;
                int hi;
                fix_split(p1[sp], Ep, lo, hi);
                fix_acc(p1lo[sp], p1hi[sp], lo, hi);
            }
        }
    }
    if (AXIS != 2) {
#pragma unroll
        for (int sp = 0; sp < S; ++sp) {
            // exact integer sum over the wave, then one addition per wave
            unsigned long long lo = p1lo[sp];
            long long hi = p1hi[sp];
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned long long olo = (unsigned long long)__shfl_xor((long long)lo, off);
                const long long ohi = __shfl_xor(hi, off);
                fix_acc(lo, hi, olo, ohi);
            }
            if ((tid & 63) == 0) fix_add_parts(alo, ahi, a.off1 + sp, lo, (int)hi);
        }
    }
    __syncthreads();
    const long long ob = ((long long)blockIdx.y * a.nitems + blockIdx.x) * a.P;
    unsigned long long *olo = a.plo + ob;
    int *ohi = a.phi + ob;
    for (int i = tid; i < n3; i += blockDim.x) {
        const int j = (i % NA3) * K1 + i / NA3;
        olo[i] = alo[j];
        ohi[i] = ahi[j];
    }
    for (int i = n3 + tid; i < a.P; i += blockDim.x) {
        olo[i] = alo[i];
        ohi[i] = ahi[i];
    }
}

// Sum the item partials (exact integer sums, so the order does not matter) and
// convert each bin to double with its realisation's scale.  blockIdx.y =
// realisation.  Output layout per realisation: out3[(kmax+1)*NA3], out2[nb2*S]
// (reference 2D index), out1[(m+1)*S].
struct ReduceArgs {
    const unsigned long long *plo;
    const int *phi;
    const int *key_items;
    const int *escale;
    double *out3, *out2, *out1;
    int nitems, nkeys, n3, L2, L1, S, P, off2, off1, axis, kmax_par, nreal;
    long long nb2, n1;       // out2 / out1 elements per realisation (nb2*S, (m+1)*S)
};

// Columns summed over every item (3D bins; the 1D line for axis 2) are reduced
// in two stages: chunks of kRedChunk items, then the chunk sums.
constexpr int kRedChunk = 64;

// (tlo, thi)[r][ch][c] = sum over items of chunk ch (of nch) of the partials' column
// col0 + c, realisation r
__device__ __forceinline__ void reduce_chunk_col(const unsigned long long *__restrict__ plo,
                                                 const int *__restrict__ phi, int nitems, int P, int col0,
                                                 int ncol, unsigned long long *__restrict__ tlo,
                                                 long long *__restrict__ thi, int c, int ch, int nch, int r) {
    const long long base = (long long)r * nitems * P + col0 + c;
    const int i0 = ch * kRedChunk, i1 = min(nitems, i0 + kRedChunk);
    unsigned long long slo = 0;
    long long shi = 0;
#pragma unroll 8
    for (int it = i0; it < i1; ++it) fix_acc(slo, shi, plo[base + (long long)it * P], phi[base + (long long)it * P]);
    const long long o = ((long long)r * nch + ch) * ncol + c;
    tlo[o] = slo;
    thi[o] = shi;
}

__global__ void pk_reduce_chunks_kernel(const unsigned long long *__restrict__ plo,
                                        const int *__restrict__ phi, int nitems, int P,
                                        int col0, int ncol, unsigned long long *__restrict__ tlo,
                                        long long *__restrict__ thi) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    reduce_chunk_col(plo, phi, nitems, P, col0, ncol, tlo, thi, c, blockIdx.y, gridDim.y, blockIdx.z);
}

// out[r*ostride + c] = sum over chunks of (tlo, thi)[r][ch][c], as a double; with
// na3 > 0, column c of the 3D block is a phase sum when c % na3 == na3 - 1
__global__ void pk_reduce_final_kernel(const unsigned long long *__restrict__ tlo,
                                       const long long *__restrict__ thi, int nch, int ncol,
                                       double *__restrict__ out, long long ostride,
                                       const int *__restrict__ escale, int na3) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncol) return;
    const long long t0 = (long long)blockIdx.y * nch * ncol + c;
    unsigned long long slo = 0;
    long long shi = 0;
    for (int ch = 0; ch < nch; ++ch) fix_acc(slo, shi, tlo[t0 + (long long)ch * ncol], thi[t0 + (long long)ch * ncol]);
    const int E = escale[2 * blockIdx.y + ((na3 > 0 && c % na3 == na3 - 1) ? 1 : 0)];
    out[(long long)blockIdx.y * ostride + c] = fix_value(slo, shi, E);
}

// 2D line element i = (key, line index, spectrum) of realisation r: sums the key's items.
__device__ __forceinline__ void reduce2_elem(const ReduceArgs &a, long long i, int r) {
    const long long base = (long long)r * a.nitems * a.P;
    int sp = (int)(i % a.S);
    int li = (int)((i / a.S) % a.L2);
    int key = (int)(i / ((long long)a.S * a.L2));
    unsigned long long slo = 0;
    long long shi = 0;
    for (int it = a.key_items[key]; it < a.key_items[key + 1]; ++it) {
        const long long j = base + (long long)it * a.P + a.off2 + li * a.S + sp;
        fix_acc(slo, shi, a.plo[j], a.phi[j]);
    }
    // reference 2D index: (kmax_par+1)*k_per + k_par
    long long idx2 = (a.axis == 2) ? (long long)(a.kmax_par + 1) * key + li
                                   : (long long)(a.kmax_par + 1) * li + key;
    a.out2[(long long)r * a.nb2 + idx2 * a.S + sp] = fix_value(slo, shi, a.escale[2 * r]);
}

// Axes 0/1: 1D bin = key = k_par; element i = (key, spectrum) sums the key's items.
// (Axis 2 keeps one 1D line per item and goes through the chunked reduction.)
__device__ __forceinline__ void reduce1_elem(const ReduceArgs &a, int i, int r) {
    const long long base = (long long)r * a.nitems * a.P;
    double *out1 = a.out1 + (long long)r * a.n1;
    int key = i / a.S, sp = i % a.S;
    unsigned long long slo = 0;
    long long shi = 0;
    for (int it = a.key_items[key]; it < a.key_items[key + 1]; ++it) {
        const long long j = base + (long long)it * a.P + a.off1 + sp;
        fix_acc(slo, shi, a.plo[j], a.phi[j]);
    }
    out1[key * a.S + sp] = fix_value(slo, shi, a.escale[2 * r]);
}

// The independent first reductions in one launch of 256-thread blocks, by block
// range: the 3D columns' chunk sums (g3 = ceil(n3/256) * nch * nreal blocks), then
// the 2D line (g2 blocks), then for axes 0/1 the 1D line (g1 blocks).
__global__ void __launch_bounds__(256)
pk_reduce_first_kernel(ReduceArgs a, unsigned long long *__restrict__ tlo, long long *__restrict__ thi,
                       int nch, int g3, int g2) {
    const int b = blockIdx.x, t = threadIdx.x;
    if (b < g3) {
        const int xb = (a.n3 + 255) / 256;
        const int c = (b % xb) * 256 + t, rest = b / xb;
        if (c < a.n3) reduce_chunk_col(a.plo, a.phi, a.nitems, a.P, 0, a.n3, tlo, thi, c, rest % nch, nch, rest / nch);
    } else if (b < g3 + g2) {
        const int xb = g2 / a.nreal, bb = b - g3;
        const long long i = (long long)(bb % xb) * 256 + t;
        if (i < (long long)a.nkeys * a.L2 * a.S) reduce2_elem(a, i, bb / xb);
    } else {
        const int bb = b - g3 - g2, xb = (a.nkeys * a.S + 255) / 256;
        const int i = (bb % xb) * 256 + t;
        if (i < a.nkeys * a.S) reduce1_elem(a, i, bb / xb);
    }
}

// FFT prologue: out = (in / scale) - shift in float32 (the reference caller's
// "delta /= mean; delta -= 1.0", Pk_snapshot.py:88), or a plain copy.  blockIdx.y
// selects one of several contiguous cubes of n elements, each with its own
// scale/shift when the device arrays are given.
__global__ void prologue_kernel(const float *__restrict__ in, float *__restrict__ out,
                                long long n, const float *__restrict__ scales,
                                const float *__restrict__ shifts, float scale, float shift,
                                int apply, int sstride = 1) {
    const long long cube = blockIdx.y;
    in += cube * n;
    out += cube * n;
    if (scales) scale = scales[cube * sstride];
    if (shifts) shift = shifts[cube * sstride];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float x = in[i];
        if (apply) {
            x = x / scale;
            x = x - shift;
        }
        out[i] = x;
    }
}

__global__ void rsd_kernel(float *__restrict__ pos, const float *__restrict__ vel, long long n,
                           float BoxSize, float factor, int axis) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        pos[3 * i + axis] = rsd_shift(pos[3 * i + axis], vel[3 * i + axis], factor, BoxSize);
    }
}

template <typename K>
int set_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        PYL_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bytes));
    return PYL_OK;
}

// Run the binning of NFL fields (device half-complex arrays) -> reduced sums
// appended into host arrays of spectra [s_off, s_off+S).
struct HostAcc {
    int S_total;
    std::vector<double> P3;     // [(kmax+1)*3*S_total] l-major per bin: (bin*3+l)*S_total+s
    std::vector<double> phase;  // [kmax+1]
    std::vector<double> P2;     // [nb2*S_total]
    std::vector<double> P1;     // [(m+1)*S_total]
    // the 2D and 1D sums to finalize: P2/P1, or the pinned download itself when one
    // launch produced every spectrum in output order (no copy of the 2D bins)
    const double *P2v = nullptr, *P1v = nullptr;
};

using AccsPtr = std::shared_ptr<std::vector<HostAcc>>;

// Copies the downloaded sums of one binning launch (h3 | h2 | h1, per realisation)
// into the spectrum columns spec_out of the host accumulators.
template <int NFL, bool PHASE>
void bin_host_part(const PkDims &p, const double *hp, size_t out3, size_t out2, size_t out1,
                   int nreal, const int *spec_out, std::vector<HostAcc> &accs) {
    constexpr int S = Spec<NFL>::S;
    constexpr int NA3 = 3 * S + (PHASE ? 1 : 0);
    const double *h3 = hp, *h2 = h3 + out3 * nreal, *h1 = h2 + out2 * nreal;
    for (int rr = 0; rr < nreal; ++rr) {
        HostAcc &acc = accs[rr];
        const double *H3 = h3 + out3 * rr, *H2 = h2 + out2 * rr, *H1 = h1 + out1 * rr;
        const int ST = acc.S_total;
        for (int b = 0; b <= p.kmax; ++b) {
            for (int sp = 0; sp < S; ++sp) {
                int so = spec_out[sp];
                if (so < 0) continue;
                for (int l = 0; l < 3; ++l)
                    acc.P3[((size_t)b * 3 + l) * ST + so] = H3[(size_t)b * NA3 + 3 * sp + l];
            }
            if (PHASE) acc.phase[b] = H3[(size_t)b * NA3 + NA3 - 1];
        }
        bool identity = ST == S;
        for (int sp = 0; sp < S; ++sp) identity = identity && spec_out[sp] == sp;
        if (identity) {
            // the only launch of this call: finalize straight from the pinned buffer
            // (valid until the next download into it)
            acc.P2v = H2;
            acc.P1v = H1;
            continue;
        }
        for (long long i = 0; i < p.nb2; ++i)
            for (int sp = 0; sp < S; ++sp)
                if (spec_out[sp] >= 0) acc.P2[(size_t)i * ST + spec_out[sp]] = H2[(size_t)i * S + sp];
        for (int i = 0; i <= p.m; ++i)
            for (int sp = 0; sp < S; ++sp)
                if (spec_out[sp] >= 0) acc.P1[(size_t)i * ST + spec_out[sp]] = H1[(size_t)i * S + sp];
        acc.P2v = acc.P2.data();
        acc.P1v = acc.P1.data();
    }
}

template <int AXIS, int NFL, bool PHASE>
int launch_bin(Ctx *c, const BinPlan *bp, const PkDims &p, const float2 *const *dk,
               long long fstride, long long zs, int nreal, const double *const *corr,
               const float *cmax, const unsigned *const *amax, const AccsPtr &accs,
               const int *spec_out) {
    constexpr int S = Spec<NFL>::S;
    constexpr int NA3 = 3 * S + (PHASE ? 1 : 0);
    const int L2 = (AXIS == 2) ? p.zlen : p.kmax_per + 1;
    const int L1 = (AXIS == 2) ? p.zlen : 1;
    const int n3 = (p.kmax + 1) * NA3;
    const int P = n3 + (L2 + L1) * S;
    const int off2 = n3, off1 = n3 + L2 * S;
    const int rmax = bp->R;
    // LDS: P low words (8 B) and P high words (4 B, padded to an even count), the
    // kz corrections, the row corrections and the row table
    size_t lds = sizeof(unsigned long long) * (size_t)P + sizeof(int) * (size_t)((P + 1) & ~1) +
                 sizeof(double) * ((size_t)p.zlen * NFL + (size_t)rmax * NFL) +
                 sizeof(int) * 3 * (size_t)rmax;
    PYL_CHECK(lds <= 160 * 1024, PYL_EINVAL, "grid too large for the binning kernel (%zu B LDS)", lds);
    // scratch: partials (low words | high words), reduced outputs, chunk sums, scales
    const size_t part_elems = (size_t)bp->nitems * P * nreal;
    const size_t out3 = (size_t)n3, out2 = (size_t)p.nb2 * S, out1 = (size_t)(p.m + 1) * S;
    const int nch = ceil_div(bp->nitems, kRedChunk);
    const int ncol_all = (AXIS == 2) ? std::max(n3, L1 * S) : n3;
    const size_t tmp_elems = (size_t)nch * ncol_all * nreal;
    const size_t phi_bytes = (sizeof(int) * part_elems + 7) & ~(size_t)7;
    PYL_TRY(ensure(c, BUF_BINS, sizeof(unsigned long long) * part_elems + phi_bytes +
                                    sizeof(double) * (out3 + out2 + out1) * nreal +
                                    2 * sizeof(long long) * tmp_elems + sizeof(int) * 2 * (size_t)nreal));
    unsigned long long *plo = (unsigned long long *)c->bufs[BUF_BINS].ptr;
    int *phi = (int *)(plo + part_elems);
    double *d3 = (double *)((char *)phi + phi_bytes);
    double *d2 = d3 + out3 * nreal;
    double *d1 = d2 + out2 * nreal;
    unsigned long long *tlo = (unsigned long long *)(d1 + out1 * nreal);
    long long *thi = (long long *)(tlo + tmp_elems);
    int *escale = (int *)(thi + tmp_elems);

    BinArgs a{};
    for (int f = 0; f < NFL; ++f) {
        a.dk[f] = dk[f];
        a.corr[f] = corr[f];
        a.cmax[f] = cmax[f];
        a.amax[f] = amax[f];
    }
    a.escale = escale;
    a.nreal = nreal;
    a.rows = bp->d_rows;
    a.items = bp->d_items;
    a.plo = plo;
    a.phi = phi;
    a.fstride = fstride;
    a.zs = zs;
    a.nitems = bp->nitems;
    a.dims = p.dims; a.m = p.m; a.zlen = p.zlen; a.zpad = p.zpad; a.even = p.even;
    a.kmax = p.kmax; a.L2 = L2; a.L1 = L1; a.P = P; a.off2 = off2; a.off1 = off1; a.rmax = rmax;
    auto kern = pk_bin_kernel<AXIS, NFL, PHASE>;
    PYL_TRY(set_lds(kern, lds));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_BIN));
    kern<<<dim3(bp->nitems, nreal), kBinBlock, lds, c->stream>>>(a);
    PYL_HIP(hipGetLastError());
    ReduceArgs r{};
    r.plo = plo; r.phi = phi; r.escale = escale;
    r.key_items = bp->d_key_items; r.out3 = d3; r.out2 = d2; r.out1 = d1;
    r.nitems = bp->nitems; r.nkeys = bp->nkeys; r.n3 = n3; r.L2 = L2; r.L1 = L1; r.S = S;
    r.P = P; r.off2 = off2; r.off1 = off1; r.axis = AXIS; r.kmax_par = p.kmax_par;
    r.nb2 = (long long)out2; r.n1 = (long long)out1; r.nreal = nreal;
    // 3D bins (every item contributes: chunk sums, then their sum), the 2D line and,
    // for axes 0/1, the 1D line (each bin from its key's items): the first stages in
    // one launch
    const long long n2 = (long long)bp->nkeys * L2 * S;
    const long long g3 = (long long)ceil_div(n3, 256) * nch * nreal;
    const long long g2 = (n2 + 255) / 256 * nreal;
    const long long g1 = AXIS == 2 ? 0 : (long long)ceil_div(bp->nkeys * S, 256) * nreal;
    PYL_CHECK(g3 + g2 + g1 < (1ll << 31), PYL_EINVAL, "binning reduction too large");
    pk_reduce_first_kernel<<<(unsigned)(g3 + g2 + g1), 256, 0, c->stream>>>(r, tlo, thi, nch, (int)g3, (int)g2);
    pk_reduce_final_kernel<<<dim3(ceil_div(n3, 256), nreal), 256, 0, c->stream>>>(
        tlo, thi, nch, n3, d3, (long long)out3, escale, PHASE ? NA3 : 0);
    if (AXIS == 2) {   // 1D bins are k_par = kz: every item contributes
        pk_reduce_chunks_kernel<<<dim3(ceil_div(L1 * S, 256), nch, nreal), 256, 0, c->stream>>>(
            plo, phi, bp->nitems, P, off1, L1 * S, tlo, thi);
        pk_reduce_final_kernel<<<dim3(ceil_div(L1 * S, 256), nreal), 256, 0, c->stream>>>(
            tlo, thi, nch, L1 * S, d1, (long long)out1, escale, 0);
    }
    PYL_HIP(hipGetLastError());
    PYL_TRY(stage_end(c, PYL_STAGE_PK_BIN));
    // d3 | d2 | d1 are contiguous: one DMA into pinned memory
    const size_t nout = (out3 + out2 + out1) * nreal;
    void *hp = nullptr;
    if (c->capturing) {
        // a pinned buffer of the graph's own; the host part runs after each replay
        PYL_HIP(hipHostMalloc(&hp, sizeof(double) * nout, hipHostMallocDefault));
        c->cap_host.push_back(hp);
    } else {
        PYL_TRY(pinned(c, c->pin_down, sizeof(double) * nout, &hp));
    }
    PYL_HIP(hipMemcpyAsync(hp, d3, sizeof(double) * nout, hipMemcpyDeviceToHost, c->stream));
    if (c->capturing) {
        std::array<int, Spec<NFL>::S> so;
        for (int sp = 0; sp < Spec<NFL>::S; ++sp) so[sp] = spec_out[sp];
        const double *h = (const double *)hp;
        AccsPtr keep = accs;
        c->deferred.push_back([p, h, out3, out2, out1, nreal, so, keep]() {
            bin_host_part<NFL, PHASE>(p, h, out3, out2, out1, nreal, so.data(), *keep);
            return PYL_OK;
        });
        return PYL_OK;
    }
    PYL_HIP(hipStreamSynchronize(c->stream));
    bin_host_part<NFL, PHASE>(p, (const double *)hp, out3, out2, out1, nreal, spec_out, *accs);
    return PYL_OK;
}

template <int NFL, bool PHASE>
int launch_bin_axis(Ctx *c, int axis, const BinPlan *bp, const PkDims &p, const float2 *const *dk,
                    long long fstride, long long zs, int nreal, const double *const *corr,
                    const float *cmax, const unsigned *const *amax, const AccsPtr &accs,
                    const int *spec_out) {
    if (axis == 0)
        return launch_bin<0, NFL, PHASE>(c, bp, p, dk, fstride, zs, nreal, corr, cmax, amax, accs, spec_out);
    if (axis == 1)
        return launch_bin<1, NFL, PHASE>(c, bp, p, dk, fstride, zs, nreal, corr, cmax, amax, accs, spec_out);
    return launch_bin<2, NFL, PHASE>(c, bp, p, dk, fstride, zs, nreal, corr, cmax, amax, accs, spec_out);
}

// Normalisation, Pk_library.pyx:384-418 (and :735-791 for XPk).
// The outputs that depend only on the mode plan (k, bin centres, mode counts):
// filled while the GPU still transforms and bins.
int finalize_modes(const ModePlan *mp, const PkDims &p, double BoxSize, pyl_pk_out *out) {
    const double kF = 2.0 * M_PI / BoxSize;
    double sumN = 0.0;
    for (double v : mp->N3D) sumN += v;
    const long long own = (p.dims % 2 == 1) ? 1 : 8;
    const long long d3 = (long long)p.dims * p.dims * p.dims;
    const long long expected = (d3 - own) / 2 + own;
    PYL_CHECK((long long)sumN == expected, PYL_EMODES,
              "WARNING: Not all modes counted (counted %lld, expected %lld)", (long long)sumN,
              expected);
    for (int i = 1; i <= p.m; ++i) {
        if (out->k1D) out->k1D[i - 1] = (mp->k1D[i] / mp->N1D[i]) * kF;
        if (out->Nmodes1D) out->Nmodes1D[i - 1] = mp->N1D[i];
    }
    // bin centres (kpar fastest: contiguous writes)
    for (int kper = 0; kper <= p.kmax_per; ++kper)
        for (int kpar = 0; kpar <= p.kmax_par; ++kpar) {
            const long long idx = (long long)(p.kmax_par + 1) * kper + kpar;
            if (out->kpar) out->kpar[idx] = 0.5 * (kpar + kpar + 1) * kF;
            if (out->kper) out->kper[idx] = 0.5 * (kper + kper + 1) * kF;
        }
    if (out->Nmodes2D) std::memcpy(out->Nmodes2D, mp->N2D.data(), sizeof(double) * p.nb2);
    for (int b = 1; b <= p.kmax; ++b) {
        if (out->k3D) out->k3D[b - 1] = (mp->k3D[b] / mp->N3D[b]) * kF;
        if (out->Nmodes3D) out->Nmodes3D[b - 1] = mp->N3D[b];
    }
    return PYL_OK;
}

// The spectra (Pk_library.pyx:384-418): 1D drops the DC bin, 2D keeps every bin,
// 3D drops the DC bin with the multipoles times (2l+1).
void finalize_values(const ModePlan *mp, const PkDims &p, double BoxSize, const HostAcc &acc,
                     bool want_phase, pyl_pk_out *out) {
    const int S = acc.S_total;
    const double kF = 2.0 * M_PI / BoxSize;
    const double kN = p.m * kF;
    const double fact = std::pow(BoxSize / ((double)p.dims * (double)p.dims), 3.0);
    if (out->Pk1D)
        for (int i = 1; i <= p.m; ++i) {
            const double N = mp->N1D[i];
            const double k1 = (mp->k1D[i] / N) * kF;
            const double kmaxper = std::sqrt(kN * kN - k1 * k1);
            for (int s = 0; s < S; ++s) {
                double v = acc.P1v[(size_t)i * S + s] * fact;
                v = v * (M_PI * (kmaxper * kmaxper) / N) / ((2.0 * M_PI) * (2.0 * M_PI));
                out->Pk1D[(size_t)(i - 1) * S + s] = v;
            }
        }
    if (out->Pk2D) {
        // (the 2D bins are the bulk of the host work: a restrict-qualified loop that
        // the host compiler vectorizes for one spectrum)
        double *__restrict__ o2 = out->Pk2D;
        const double *__restrict__ s2 = acc.P2v;
        const double *__restrict__ n2 = mp->N2D.data();
        if (S == 1) {
            for (long long i = 0; i < p.nb2; ++i) o2[i] = s2[i] * fact / n2[i];
        } else {
            for (long long i = 0; i < p.nb2; ++i)
                for (int s = 0; s < S; ++s) o2[(size_t)i * S + s] = s2[(size_t)i * S + s] * fact / n2[i];
        }
    }
    for (int b = 1; b <= p.kmax; ++b) {
        const double N = mp->N3D[b];
        if (out->Pk3D)
            for (int s = 0; s < S; ++s) {
                const size_t o = (size_t)(b - 1) * 3 * S;
                out->Pk3D[o + 0 * S + s] = (acc.P3[((size_t)b * 3 + 0) * S + s] / N) * fact;
                out->Pk3D[o + 1 * S + s] = (acc.P3[((size_t)b * 3 + 1) * S + s] * 5.0 / N) * fact;
                out->Pk3D[o + 2 * S + s] = (acc.P3[((size_t)b * 3 + 2) * S + s] * 9.0 / N) * fact;
            }
        if (want_phase && out->Pkphase) out->Pkphase[b - 1] = (acc.phase[b] / N) * fact;
    }
}

}  // namespace

__global__ void zero_u32_kernel(unsigned *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// 64 zeroed words per cube for the transforms' largest |Re|, |Im| (a fill kernel,
// not a memset node, so that a captured call clears them on every replay).
int pk_amax_words(Ctx *c, int ncube, unsigned **out) {
    const int n = 64 * ncube;
    PYL_TRY(ensure(c, BUF_PK_AMAX, sizeof(unsigned) * (size_t)n));
    unsigned *a = (unsigned *)c->bufs[BUF_PK_AMAX].ptr;
    zero_u32_kernel<<<ceil_div(n, 256), 256, 0, c->stream>>>(a, n);
    PYL_HIP(hipGetLastError());
    *out = a;
    return PYL_OK;
}

int fft_device(Ctx *c, float *in, float2 *out, int dims, int batch, long long zs = 0,
               unsigned *amax = nullptr) {
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_FFT));
    PYL_TRY(fft_r2c_3d(c, in, out, dims, batch, zs, amax));
    PYL_TRY(stage_end(c, PYL_STAGE_PK_FFT));
    return PYL_OK;
}

// Binning of F device half-complex fields of nreal realisations -> finalized outputs.
// Field f of realisation r is dk[f] + r*fstride.
// Device-resident MAS correction table for (dims, p), built once per device with the
// host's libm.
int get_mas_tab(Ctx *c, int dims, int p, const double **out) {
    SharedTables &st = shared_tables();
    std::lock_guard<std::mutex> lk(st.mu);
    const auto key = std::make_tuple(c->device, dims, p);
    auto it = st.mas.find(key);
    if (it == st.mas.end()) {
        PYL_CHECK(!c->capturing, PYL_EINVAL, "MAS table created while a graph is captured");
        std::vector<double> tab(dims);
        mas_table(dims, p, tab.data());
        void *d = nullptr;
        PYL_HIP(hipMalloc(&d, sizeof(double) * dims));
        hipError_t e = hipMemcpy(d, tab.data(), sizeof(double) * dims, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d);
            set_error("MAS table upload failed: %s", hipGetErrorString(e));
            return PYL_EHIP;
        }
        it = st.mas.emplace(key, d).first;
    }
    *out = (const double *)it->second;
    return PYL_OK;
}

// Rows of each field are zs elements apart (zs = 0: dense, dims/2+1).  amax: per
// field, realisation r's 64 words of the transform's largest |Re|, |Im| at r*64
// (fft_r2c_3d, complex_absmax).
int pk_bin_device(Ctx *c, const float2 *const *dk, long long fstride, long long zs, int nreal, int F,
                  const int *mas, int dims, double BoxSize, int axis, int flags, pyl_pk_out *outs,
                  const unsigned *const *amax) {
    if (zs <= 0) zs = dims / 2 + 1;
    const PkDims p = pk_dims(dims);
    const ModePlan *mp = get_mode_plan(c, dims, axis);
    BinPlan *bp = nullptr;
    PYL_TRY(get_bin_plan(c, dims, axis, nreal, &bp));
    std::vector<const double *> tab(F);
    std::vector<float> cmax(F);
    for (int f = 0; f < F; ++f) {
        PYL_TRY(get_mas_tab(c, dims, mas[f], &tab[f]));
        // the largest product of three axis corrections, rounded up
        std::vector<double> h(dims);
        mas_table(dims, mas[f], h.data());
        double mx = 1.0;
        for (double v : h) mx = std::max(mx, v);
        cmax[f] = (float)(mx * mx * mx * 1.001);
    }
    // the mode-plan outputs are written while the FFT already queued runs
    for (int r = 0; r < nreal; ++r) PYL_TRY(finalize_modes(mp, p, BoxSize, &outs[r]));
    const int S = F + F * (F - 1) / 2;
    // host accumulators reused across calls on this thread (no fresh pages per call);
    // a captured call keeps its own, used by the host steps of every replay
    static thread_local AccsPtr tl_accs;
    AccsPtr accs_p;
    if (c->capturing) {
        accs_p = std::make_shared<std::vector<HostAcc>>();
    } else {
        if (!tl_accs) tl_accs = std::make_shared<std::vector<HostAcc>>();
        accs_p = tl_accs;
    }
    std::vector<HostAcc> &accs = *accs_p;
    accs.resize(nreal);
    for (auto &acc : accs) {      // every column is written by exactly one launch below
        acc.S_total = S;
        acc.P3.resize((size_t)(p.kmax + 1) * 3 * S);
        acc.phase.assign(p.kmax + 1, 0.0);
        acc.P2.resize((size_t)p.nb2 * S);
        acc.P1.resize((size_t)(p.m + 1) * S);
    }
    const bool phase = (F == 1) && !(flags & PYL_PK_NO_PHASE);
    if (F == 1) {
        const float2 *f0[1] = {dk[0]};
        const double *c0[1] = {tab[0]};
        const float m0[1] = {cmax[0]};
        const unsigned *a0[1] = {amax[0]};
        int so[1] = {0};
        if (phase)
            PYL_TRY((launch_bin_axis<1, true>(c, axis, bp, p, f0, fstride, zs, nreal, c0, m0, a0, accs_p, so)));
        else
            PYL_TRY((launch_bin_axis<1, false>(c, axis, bp, p, f0, fstride, zs, nreal, c0, m0, a0, accs_p, so)));
    } else {
        // pairs i<j in lexicographic order; each launch yields auto i, auto j, cross ij
        std::vector<bool> have(F, false);
        int xi = 0;
        for (int i = 0; i < F; ++i)
            for (int j = i + 1; j < F; ++j, ++xi) {
                const float2 *fp[2] = {dk[i], dk[j]};
                const double *cp[2] = {tab[i], tab[j]};
                const float mp2[2] = {cmax[i], cmax[j]};
                const unsigned *ap[2] = {amax[i], amax[j]};
                int so[3] = {have[i] ? -1 : i, have[j] ? -1 : j, F + xi};
                have[i] = have[j] = true;
                PYL_TRY((launch_bin_axis<2, false>(c, axis, bp, p, fp, fstride, zs, nreal, cp, mp2, ap,
                                                   accs_p, so)));
            }
    }
    if (c->capturing) {
        // the caller's result structs point at host arrays that outlive the graph
        std::vector<pyl_pk_out> o(outs, outs + nreal);
        c->deferred.push_back([mp, p, BoxSize, accs_p, phase, o, nreal]() mutable {
            for (int r = 0; r < nreal; ++r) finalize_values(mp, p, BoxSize, (*accs_p)[r], phase, &o[r]);
            return PYL_OK;
        });
        return PYL_OK;
    }
    for (int r = 0; r < nreal; ++r) finalize_values(mp, p, BoxSize, accs[r], phase, &outs[r]);
    if (nreal > 16) {                            // do not keep large batch accumulators
        accs.clear();
        accs.shrink_to_fit();
    }
    return PYL_OK;
}

}  // namespace pyl

using namespace pyl;

extern "C" {

double pyl_mas_correction(double x, int p) {
    return (x == 0.0) ? 1.0 : std::pow(x / std::sin(x), (double)p);
}

int pyl_pk_get_sizes(int dims, int nfields, pyl_pk_sizes *out) {
    PYL_CHECK(out && dims >= 2 && nfields >= 1, PYL_EINVAL, "bad arguments");
    PkDims p = pk_dims(dims);
    out->n1d = p.m;
    out->n2d = p.nb2;
    out->n3d = p.kmax;
    out->nspectra = nfields + nfields * (nfields - 1) / 2;
    return PYL_OK;
}

int pyl_mode_counts(int dims, int axis, double *N1D, double *N2D, double *N3D, double *k1Dsum,
                    double *k3Dsum) {
    PYL_CHECK(dims >= 2 && axis >= 0 && axis <= 2, PYL_EINVAL, "bad dims/axis");
    ModePlan mp;
    compute_mode_plan(dims, axis, &mp);
    if (N1D) std::copy(mp.N1D.begin(), mp.N1D.end(), N1D);
    if (N2D) std::copy(mp.N2D.begin(), mp.N2D.end(), N2D);
    if (N3D) std::copy(mp.N3D.begin(), mp.N3D.end(), N3D);
    if (k1Dsum) std::copy(mp.k1D.begin(), mp.k1D.end(), k1Dsum);
    if (k3Dsum) std::copy(mp.k3D.begin(), mp.k3D.end(), k3Dsum);
    return PYL_OK;
}

int pyl_fft_r2c(pyl_ctx *ctx, const float *delta, float *delta_k, int dims, int flags) {
    TraceRange trace_("pyl_fft_r2c");
    PYL_CHECK(ctx && delta && delta_k && dims >= 2, PYL_EINVAL, "bad arguments");
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    const long long n = (long long)dims * dims * dims;
    const long long nk = (long long)dims * dims * (dims / 2 + 1);
    PYL_TRY(ensure(c, BUF_FFT_IN, (size_t)n * sizeof(float)));
    float *in = (float *)c->bufs[BUF_FFT_IN].ptr;
    float2 *outp;
    if (flags & PYL_PTR_DEVICE) {
        if (flags & PYL_PK_DESTROY_INPUT) in = (float *)delta;
        else PYL_HIP(hipMemcpyAsync(in, delta, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        outp = (float2 *)delta_k;
    } else {
        PYL_TRY(ensure(c, BUF_FFT_OUT, (size_t)nk * sizeof(float2)));
        outp = (float2 *)c->bufs[BUF_FFT_OUT].ptr;
        PYL_HIP(hipMemcpyAsync(in, delta, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    PYL_TRY(fft_device(c, in, outp, dims, 1));
    if (!(flags & PYL_PTR_DEVICE))
        PYL_HIP(hipMemcpyAsync(delta_k, outp, nk * sizeof(float2), hipMemcpyDeviceToHost, c->stream));
    return finish(c, flags);
}

int pyl_pk(pyl_ctx *ctx, const float *const *deltas, int nfields, const int *mas, int dims,
           double BoxSize, int axis, float scale, float shift, int flags, pyl_pk_out *out) {
    TraceRange trace_("pyl_pk");
    PYL_CHECK(ctx && deltas && mas && out, PYL_EINVAL, "NULL argument");
    PYL_CHECK(nfields >= 1, PYL_EINVAL, "need at least one field");
    PYL_CHECK(dims >= 2, PYL_EINVAL, "dims must be >= 2");
    PYL_CHECK(axis >= 0 && axis <= 2, PYL_EINVAL, "axis must be 0, 1 or 2");
    PYL_CHECK(BoxSize > 0.0, PYL_EINVAL, "BoxSize must be positive");
    for (int f = 0; f < nfields; ++f) {
        PYL_CHECK(deltas[f] != nullptr, PYL_EINVAL, "field %d is NULL", f);
        PYL_CHECK(mas[f] >= 0 && mas[f] <= 4, PYL_EINVAL, "bad MAS id %d", mas[f]);
    }
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    const long long n = (long long)dims * dims * dims;
    // the library's own half-complex buffer: rows padded to 128 bytes (fft_zstride)
    const long long zs = fft_zstride(dims);
    const long long nk = (long long)dims * dims * zs;
    const bool apply = !(scale == 1.0f && shift == 0.0f);
    const bool dev = (flags & PYL_PTR_DEVICE) != 0;
    const bool destroy = dev && (flags & PYL_PK_DESTROY_INPUT);
    PYL_TRY(ensure(c, BUF_FFT_OUT, (size_t)nk * nfields * sizeof(float2)));
    float2 *dk = (float2 *)c->bufs[BUF_FFT_OUT].ptr;
    unsigned *amax = nullptr;
    PYL_TRY(pk_amax_words(c, nfields, &amax));
    std::vector<const unsigned *> amf(nfields);
    for (int f = 0; f < nfields; ++f) amf[f] = amax + 64 * f;
    std::vector<const float2 *> dks(nfields);
    const int pb = 256;
    const int nb = (int)std::min<long long>((n + pb - 1) / pb, 8192);
    if (destroy) {
        for (int f = 0; f < nfields; ++f) {
            float *x = (float *)deltas[f];
            if (apply) {
                PYL_TRY(stage_begin(c, PYL_STAGE_PK_PREP));
                prologue_kernel<<<nb, pb, 0, c->stream>>>(x, x, n, nullptr, nullptr, scale, shift, 1);
                PYL_HIP(hipGetLastError());
                PYL_TRY(stage_end(c, PYL_STAGE_PK_PREP));
            }
            PYL_TRY(fft_device(c, x, dk + (size_t)f * nk, dims, 1, zs, amax + 64 * f));
            dks[f] = dk + (size_t)f * nk;
        }
    } else {
        PYL_TRY(ensure(c, BUF_FFT_IN, (size_t)n * nfields * sizeof(float)));
        float *in = (float *)c->bufs[BUF_FFT_IN].ptr;
        PYL_TRY(stage_begin(c, PYL_STAGE_PK_PREP));
        for (int f = 0; f < nfields; ++f) {
            float *dst = in + (size_t)f * n;
            if (dev) {
                if (apply) {
                    prologue_kernel<<<nb, pb, 0, c->stream>>>(deltas[f], dst, n, nullptr, nullptr, scale,
                                                            shift, 1);
                    PYL_HIP(hipGetLastError());
                } else {
                    PYL_HIP(hipMemcpyAsync(dst, deltas[f], n * sizeof(float), hipMemcpyDeviceToDevice,
                                           c->stream));
                }
            } else {
                PYL_HIP(hipMemcpyAsync(dst, deltas[f], n * sizeof(float), hipMemcpyHostToDevice,
                                       c->stream));
                if (apply) {
                    prologue_kernel<<<nb, pb, 0, c->stream>>>(dst, dst, n, nullptr, nullptr, scale, shift, 1);
                    PYL_HIP(hipGetLastError());
                }
            }
        }
        PYL_TRY(stage_end(c, PYL_STAGE_PK_PREP));
        PYL_TRY(fft_device(c, in, dk, dims, nfields, zs, amax));
        for (int f = 0; f < nfields; ++f) dks[f] = dk + (size_t)f * nk;
    }
    PYL_TRY(pk_bin_device(c, dks.data(), 0, zs, 1, nfields, mas, dims, BoxSize, axis, flags, out,
                          amf.data()));
    return finish(c, flags);
}

int pyl_pk_batch(pyl_ctx *ctx, const float *const *deltas, int nreal, int nfields,
                 const int *mas, int dims, double BoxSize, int axis, const float *scales,
                 const float *shifts, int flags, pyl_pk_out *outs) {
    TraceRange trace_("pyl_pk_batch");
    PYL_CHECK(ctx && deltas && mas && outs, PYL_EINVAL, "NULL argument");
    PYL_CHECK(nreal >= 1 && nfields >= 1 && (long long)nreal * nfields <= 65535, PYL_EINVAL,
              "need 1 <= nreal*nfields <= 65535");
    PYL_CHECK(dims >= 2 && axis >= 0 && axis <= 2 && BoxSize > 0.0, PYL_EINVAL, "bad arguments");
    for (int f = 0; f < nfields; ++f) {
        PYL_CHECK(deltas[f] != nullptr, PYL_EINVAL, "field %d is NULL", f);
        PYL_CHECK(mas[f] >= 0 && mas[f] <= 4, PYL_EINVAL, "bad MAS id %d", mas[f]);
    }
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    const long long n = (long long)dims * dims * dims;
    const long long zs = fft_zstride(dims);              // padded rows (fft_zstride)
    const long long nk = (long long)dims * dims * zs;
    const int ncube = nreal * nfields;                 // cube (f, r) at index f*nreal + r
    const bool dev = (flags & PYL_PTR_DEVICE) != 0;
    const bool destroy = dev && (flags & PYL_PK_DESTROY_INPUT);
    const bool apply = scales != nullptr || shifts != nullptr;
    PYL_TRY(ensure(c, BUF_FFT_OUT, (size_t)nk * ncube * sizeof(float2)));
    float2 *dk = (float2 *)c->bufs[BUF_FFT_OUT].ptr;
    unsigned *amax = nullptr;                          // cube (f, r) at (f*nreal + r)*64
    PYL_TRY(pk_amax_words(c, ncube, &amax));
    std::vector<const unsigned *> amf(nfields);
    for (int f = 0; f < nfields; ++f) amf[f] = amax + (size_t)64 * nreal * f;
    float *in = nullptr;
    if (!destroy) {
        PYL_TRY(ensure(c, BUF_FFT_IN, (size_t)n * ncube * sizeof(float)));
        in = (float *)c->bufs[BUF_FFT_IN].ptr;
    }
    // Per-cube scale/shift: caller order [r][f].  Device arrays are read in place
    // (stride nfields from field f); host arrays are uploaded through pinned
    // staging in cube order [f][r] (no wait for the work already queued).
    const float *dsc = nullptr, *dsh = nullptr;
    int sstride = 1;
    const bool sdev = (flags & PYL_PK_SCALES_DEVICE) != 0;
    if (apply && sdev) {
        PYL_CHECK(scales != nullptr && shifts != nullptr, PYL_EINVAL,
                  "PYL_PK_SCALES_DEVICE needs both scales and shifts");
        dsc = scales;
        dsh = shifts;
        sstride = nfields;
    } else if (apply) {
        PYL_TRY(ensure(c, BUF_SCAN_TMP, sizeof(float) * 2 * (size_t)ncube));
        float *d = (float *)c->bufs[BUF_SCAN_TMP].ptr;
        void *stage = nullptr;
        PYL_TRY(pinned(c, c->pin_scales, sizeof(float) * 2 * (size_t)ncube, &stage));
        float *sc = (float *)stage, *sh = sc + ncube;
        for (int r = 0; r < nreal; ++r)
            for (int f = 0; f < nfields; ++f) {
                sc[(size_t)f * nreal + r] = scales ? scales[(size_t)r * nfields + f] : 1.0f;
                sh[(size_t)f * nreal + r] = shifts ? shifts[(size_t)r * nfields + f] : 0.0f;
            }
        PYL_HIP(hipMemcpyAsync(d, stage, sizeof(float) * 2 * ncube, hipMemcpyHostToDevice, c->stream));
        PYL_TRY(pinned_done(c, c->pin_scales));
        dsc = d;
        dsh = d + ncube;
    }
    const int nbx = (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 2048));
    PYL_TRY(stage_begin(c, PYL_STAGE_PK_PREP));
    for (int f = 0; f < nfields; ++f) {
        float *dst = destroy ? (float *)deltas[f] : in + (size_t)f * nreal * n;
        if (!destroy)
            PYL_HIP(hipMemcpyAsync(dst, deltas[f], (size_t)n * nreal * sizeof(float),
                                   dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
        if (apply) {
            const size_t off = sdev ? (size_t)f : (size_t)f * nreal;
            prologue_kernel<<<dim3(nbx, nreal), 256, 0, c->stream>>>(
                dst, dst, n, dsc + off, dsh + off, 1.0f, 0.0f, 1, sstride);
            PYL_HIP(hipGetLastError());
        }
    }
    PYL_TRY(stage_end(c, PYL_STAGE_PK_PREP));
    if (destroy) {
        for (int f = 0; f < nfields; ++f)
            PYL_TRY(fft_device(c, (float *)deltas[f], dk + (size_t)f * nreal * nk, dims, nreal, zs,
                               amax + (size_t)64 * nreal * f));
    } else {
        PYL_TRY(fft_device(c, in, dk, dims, ncube, zs, amax));
    }
    std::vector<const float2 *> dks(nfields);
    for (int f = 0; f < nfields; ++f) dks[f] = dk + (size_t)f * nreal * nk;
    PYL_TRY(pk_bin_device(c, dks.data(), nk, zs, nreal, nfields, mas, dims, BoxSize, axis, flags, outs,
                          amf.data()));
    return finish(c, flags);
}

int pyl_pk_from_fourier(pyl_ctx *ctx, const float *const *delta_k, int nfields, const int *mas,
                        int dims, double BoxSize, int axis, int flags, pyl_pk_out *out) {
    TraceRange trace_("pyl_pk_from_fourier");
    PYL_CHECK(ctx && delta_k && mas && out, PYL_EINVAL, "NULL argument");
    PYL_CHECK(nfields >= 1 && dims >= 2 && axis >= 0 && axis <= 2 && BoxSize > 0.0, PYL_EINVAL,
              "bad arguments");
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    reset_stages(c);
    const long long nk = (long long)dims * dims * (dims / 2 + 1);
    std::vector<const float2 *> dks(nfields);
    if (flags & PYL_PTR_DEVICE) {
        for (int f = 0; f < nfields; ++f) dks[f] = (const float2 *)delta_k[f];
    } else {
        PYL_TRY(ensure(c, BUF_FFT_OUT, (size_t)nk * nfields * sizeof(float2)));
        float2 *dk = (float2 *)c->bufs[BUF_FFT_OUT].ptr;
        for (int f = 0; f < nfields; ++f) {
            PYL_HIP(hipMemcpyAsync(dk + (size_t)f * nk, delta_k[f], nk * sizeof(float2),
                                   hipMemcpyHostToDevice, c->stream));
            dks[f] = dk + (size_t)f * nk;
        }
    }
    unsigned *amax = nullptr;
    PYL_TRY(pk_amax_words(c, nfields, &amax));
    std::vector<const unsigned *> amf(nfields);
    for (int f = 0; f < nfields; ++f) {
        PYL_TRY(complex_absmax(c, dks[f], dims, 0, 1, nk, amax + 64 * f));
        amf[f] = amax + 64 * f;
    }
    PYL_TRY(pk_bin_device(c, dks.data(), 0, 0, 1, nfields, mas, dims, BoxSize, axis, flags, out,
                          amf.data()));
    return finish(c, flags);
}

int pyl_pos_redshift_space(pyl_ctx *ctx, float *pos, const float *vel, int64_t n, float BoxSize,
                           float Hubble, float redshift, int axis, int flags) {
    PYL_CHECK(ctx && pos && vel, PYL_EINVAL, "NULL argument");
    PYL_CHECK(axis >= 0 && axis <= 2, PYL_EINVAL, "axis must be 0, 1 or 2");
    Ctx *c = &ctx->c;
    PYL_HIP(hipSetDevice(c->device));
    if (n <= 0) return PYL_OK;
    const float factor = (float)((1.0 + (double)redshift) / (double)Hubble);
    float *dpos = pos;
    const float *dvel = vel;
    if (!(flags & PYL_PTR_DEVICE)) {
        PYL_TRY(ensure(c, BUF_POS, (size_t)n * 3 * sizeof(float)));
        PYL_TRY(ensure(c, BUF_VEL, (size_t)n * 3 * sizeof(float)));
        dpos = (float *)c->bufs[BUF_POS].ptr;
        dvel = (const float *)c->bufs[BUF_VEL].ptr;
        PYL_HIP(hipMemcpyAsync(dpos, pos, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        PYL_HIP(hipMemcpyAsync((void *)dvel, vel, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice,
                               c->stream));
    }
    const int nb = (int)std::min<long long>((n + 255) / 256, 65536);
    rsd_kernel<<<nb, 256, 0, c->stream>>>(dpos, dvel, n, BoxSize, factor, axis);
    PYL_HIP(hipGetLastError());
    if (!(flags & PYL_PTR_DEVICE))
        PYL_HIP(hipMemcpyAsync(pos, dpos, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return finish(c, flags);
}

}  // extern "C"
