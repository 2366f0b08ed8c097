/*
 * ORACLE -- test infrastructure only.  Nothing in the product path may link,
 * load or call this file; only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py use it, and only as the checker.
 *
 * CPU restatement of the reference power-spectrum mode loops:
 *   library/Pk_library/Pk_library.pyx:83-84    MAS_correction
 *   library/Pk_library/Pk_library.pyx:311-378  Pk  mode loop (serial)
 *   library/Pk_library/Pk_library.pyx:623-732  XPk mode loop (serial)
 *   library/Pk_library/Pk_library.pyx:466-497  Pk_plane mode loop
 *   library/Pk_library/Pk_library.pyx:1155-1203 XPk_plane mode loop
 *   library/Pk_library/Pk_library.pyx:1910-1929 correct_MAS mode scaling
 * The loops take the half-complex FFT output (dims, dims, dims/2+1) and fill
 * the raw per-bin sums; normalisation (Pk_library.pyx:384-418, 735-791) is in
 * oracle/oracle.py.  The float32 variant keeps the reference's types: float
 * MAS_factor, complex64 scaling, double |delta_k|^2 and double accumulators.
 * The float64 variant keeps everything in double (the FP64 oracle).
 *
 * Pinned by tests/test_oracle_pk.py: mode counts against the closed form of
 * check_number_modes and brute-force Hermitian enumeration, single plane waves
 * (closed-form bin value), modes on the line of sight (Legendre factors 1),
 * the MAS window closed form, and XPk([d,d]) / XPk([d,-d]) identities.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static double mas_correction(double x, int p)
{
    return (x == 0.0) ? 1.0 : pow(x / sin(x), (double)p);
}

/* Shared per-mode geometry, Pk_library.pyx:313-348. Returns 0 if skipped. */
typedef struct {
    double k, mu2;
    int k_index, k_par, k_per;
} mode_geom;

static int geom(int kx, int ky, int kz, int dims, int middle, int axis, mode_geom *g)
{
    if (kz == 0 || (kz == middle && dims % 2 == 0)) {
        if (kx < 0) return 0;
        else if (kx == 0 || (kx == middle && dims % 2 == 0)) {
            if (ky < 0) return 0;
        }
    }
    g->k = sqrt((double)(kx * kx + ky * ky + kz * kz));
    g->k_index = (int)g->k;
    if (axis == 0) { g->k_par = kx; g->k_per = (int)sqrt((double)(ky * ky + kz * kz)); }
    else if (axis == 1) { g->k_par = ky; g->k_per = (int)sqrt((double)(kx * kx + kz * kz)); }
    else { g->k_par = kz; g->k_per = (int)sqrt((double)(kx * kx + ky * ky)); }
    {
        double mu = (g->k == 0.0) ? 0.0 : g->k_par / g->k;
        g->mu2 = mu * mu;
    }
    if (g->k_par < 0) g->k_par = -g->k_par;
    return 1;
}

/*
 * Pk mode loop.  dk: interleaved (re,im) of F complex fields, each
 * (dims,dims,middle+1); is_f64 selects complex128 input and double MAS factor.
 * Outputs for field f (auto) or pair p (cross, i<j lexicographic):
 *   Pk1D[(middle+1)*S], Pk2D[nb2*S], Pk3D[(kmax+1)*3*S], Pkphase[kmax+1] (F==1 only)
 * where S = F + F(F-1)/2 spectra (autos first, then crosses).
 * Counters (shared): k1D, N1D [middle+1], N2D [nb2], k3D, N3D [kmax+1].
 */
static void pk_loop_range(const void *dk, int is_f64, int F, int dims, int axis,
                          const int *mas_index, int kx0, int kx1,
                          double *k1D, double *Pk1D, double *N1D,
                          double *Pk2D, double *N2D,
                          double *k3D, double *Pk3D, double *Pkphase, double *N3D)
{
    int middle = dims / 2;
    int kmax_par = middle;
    int kmax_per = (int)sqrt((double)(middle * middle + middle * middle));
    int kmax = (int)sqrt((double)(3 * middle * middle));
    long nb2 = (long)(kmax_par + 1) * (kmax_per + 1);
    int S = F + F * (F - 1) / 2;
    long zlen = middle + 1;
    double prefact = M_PI / dims;
    double *corr = (double *)malloc(sizeof(double) * 3 * F);
    double *re = (double *)malloc(sizeof(double) * F);
    double *im = (double *)malloc(sizeof(double) * F);
    int kxx, kyy, kzz, f, i, j, s;
    (void)kmax;

    for (kxx = kx0; kxx < kx1; kxx++) {
        int kx = (kxx > middle) ? kxx - dims : kxx;
        for (f = 0; f < F; f++) corr[3 * f + 0] = mas_correction(prefact * kx, mas_index[f]);
        for (kyy = 0; kyy < dims; kyy++) {
            int ky = (kyy > middle) ? kyy - dims : kyy;
            for (f = 0; f < F; f++) corr[3 * f + 1] = mas_correction(prefact * ky, mas_index[f]);
            for (kzz = 0; kzz < middle + 1; kzz++) {
                int kz = kzz;
                mode_geom g;
                long index_2D;
                for (f = 0; f < F; f++) corr[3 * f + 2] = mas_correction(prefact * kz, mas_index[f]);
                if (!geom(kx, ky, kz, dims, middle, axis, &g)) continue;

                for (f = 0; f < F; f++) {
                    long e = (((long)kxx * dims + kyy) * zlen + kzz);
                    if (is_f64) {
                        const double *d = (const double *)dk + 2 * ((long)f * dims * dims * zlen + e);
                        double mf = corr[3 * f + 0] * corr[3 * f + 1] * corr[3 * f + 2];
                        re[f] = d[0] * mf; im[f] = d[1] * mf;
                    } else {
                        const float *d = (const float *)dk + 2 * ((long)f * dims * dims * zlen + e);
                        float mf = (float)(corr[3 * f + 0] * corr[3 * f + 1] * corr[3 * f + 2]);
                        float r = d[0] * mf, m = d[1] * mf;   /* complex64 * float */
                        re[f] = (double)r; im[f] = (double)m;
                    }
                }
                index_2D = (long)(kmax_par + 1) * g.k_per + g.k_par;
                if (g.k <= middle) { k1D[g.k_par] += g.k_par; N1D[g.k_par] += 1.0; }
                N2D[index_2D] += 1.0;
                k3D[g.k_index] += g.k;
                N3D[g.k_index] += 1.0;

                s = 0;
                for (i = 0; i < F; i++) {
                    for (j = i; j < F; j++) {
                        double d2;
                        int sp;
                        if (j == i) { d2 = re[i] * re[i] + im[i] * im[i]; sp = i; }
                        else { d2 = re[i] * re[j] + im[i] * im[j]; sp = F + s++; }
                        if (g.k <= middle) Pk1D[(long)g.k_par * S + sp] += d2;
                        Pk2D[index_2D * S + sp] += d2;
                        Pk3D[((long)g.k_index * 3 + 0) * S + sp] += d2;
                        Pk3D[((long)g.k_index * 3 + 1) * S + sp] += (d2 * (3.0 * g.mu2 - 1.0) / 2.0);
                        Pk3D[((long)g.k_index * 3 + 2) * S + sp] +=
                            (d2 * (35.0 * g.mu2 * g.mu2 - 30.0 * g.mu2 + 3.0) / 8.0);
                        if (F == 1 && Pkphase) {
                            double phase = atan2(re[0], sqrt(d2));
                            Pkphase[g.k_index] += phase * phase;
                        }
                    }
                }
            }
        }
    }
    (void)nb2;
    free(corr); free(re); free(im);
}


void ora_pk_loop(const void *dk, int is_f64, int F, int dims, int axis,
                 const int *mas_index,
                 double *k1D, double *Pk1D, double *N1D,
                 double *Pk2D, double *N2D,
                 double *k3D, double *Pk3D, double *Pkphase, double *N3D)
{
    pk_loop_range(dk, is_f64, F, dims, axis, mas_index, 0, dims, k1D, Pk1D, N1D, Pk2D, N2D,
                  k3D, Pk3D, Pkphase, N3D);
}

/*
 * The same loop over `threads` OpenMP threads: thread t takes a contiguous range
 * of kxx into private sums, which are then added into the outputs in thread
 * order.  Deterministic; per-bin double sums differ from the serial loop's only
 * by the order of their additions (1e-16 relative).  For grids of 512^3 and more.
 */
void ora_pk_loop_threads(const void *dk, int is_f64, int F, int dims, int axis,
                         const int *mas_index,
                         double *k1D, double *Pk1D, double *N1D,
                         double *Pk2D, double *N2D,
                         double *k3D, double *Pk3D, double *Pkphase, double *N3D, int threads)
{
    int middle = dims / 2;
    long n1 = middle + 1;
    long kmax_per = (long)sqrt((double)(middle * middle + middle * middle));
    long nb2 = (long)(middle + 1) * (kmax_per + 1);
    long n3 = (long)sqrt((double)(3 * middle * middle)) + 1;
    long S = F + F * (F - 1) / 2;
    /* per-thread block: k1D N1D | Pk1D | Pk2D | N2D | k3D N3D Pkphase | Pk3D */
    long blk = 2 * n1 + n1 * S + nb2 * S + nb2 + 3 * n3 + n3 * 3 * S;
    int t, T = threads < 1 ? 1 : (threads > dims ? dims : threads);
    long e;
    double *priv = (double *)calloc((size_t)blk * T, sizeof(double));
    if (!priv) return;
#pragma omp parallel for num_threads(T) schedule(static, 1)
    for (t = 0; t < T; t++) {
        double *b = priv + (long)t * blk;
        double *pk1 = b, *pn1 = b + n1, *pP1 = b + 2 * n1, *pP2 = pP1 + n1 * S, *pN2 = pP2 + nb2 * S;
        double *pk3 = pN2 + nb2, *pN3 = pk3 + n3, *pph = pN3 + n3, *pP3 = pph + n3;
        int kx0 = (int)((long)dims * t / T), kx1 = (int)((long)dims * (t + 1) / T);
        pk_loop_range(dk, is_f64, F, dims, axis, mas_index, kx0, kx1, pk1, pP1, pn1, pP2, pN2,
                      pk3, pP3, Pkphase ? pph : NULL, pN3);
    }
    for (t = 0; t < T; t++) {
        double *b = priv + (long)t * blk;
        double *pk1 = b, *pn1 = b + n1, *pP1 = b + 2 * n1, *pP2 = pP1 + n1 * S, *pN2 = pP2 + nb2 * S;
        double *pk3 = pN2 + nb2, *pN3 = pk3 + n3, *pph = pN3 + n3, *pP3 = pph + n3;
        for (e = 0; e < n1; e++) { k1D[e] += pk1[e]; N1D[e] += pn1[e]; }
        for (e = 0; e < n1 * S; e++) Pk1D[e] += pP1[e];
        for (e = 0; e < nb2 * S; e++) Pk2D[e] += pP2[e];
        for (e = 0; e < nb2; e++) N2D[e] += pN2[e];
        for (e = 0; e < n3; e++) { k3D[e] += pk3[e]; N3D[e] += pN3[e]; }
        if (Pkphase) for (e = 0; e < n3; e++) Pkphase[e] += pph[e];
        for (e = 0; e < n3 * 3 * S; e++) Pk3D[e] += pP3[e];
    }
    free(priv);
}

/* MAS_correction exported for the closed-form tests (Pk_library.pyx:83-84). */
double ora_mas_correction(double x, int p) { return mas_correction(x, p); }

/*
 * 2D mode loops of Pk_plane (F == 1, Pk_library.pyx:466-497) and XPk_plane
 * (F == 2, Pk_library.pyx:1155-1203).  dk: F half-complex planes
 * (grid, grid/2+1), interleaved (re,im).  Outputs per |k| bin (kmax+1 bins,
 * kmax = (int)sqrt(2*middle^2)): k2D, N, Pk[(kmax+1)*S] with S = 1 (F == 1) or
 * 3 (two autos, then the cross).  float32 variant: float MAS factor (the
 * double product of the two axis corrections, rounded), complex64 scaling,
 * double squares and sums.
 */
void ora_pk_plane_loop(const void *dk, int is_f64, int F, int grid, const int *mas_index,
                       double *k2D, double *Pk, double *N)
{
    int middle = grid / 2;
    int S = (F == 1) ? 1 : 3;
    long zlen = middle + 1;
    double prefact = M_PI / grid;
    int kxx, kyy, f;
    for (kxx = 0; kxx < grid; kxx++) {
        int kx = (kxx > middle) ? kxx - grid : kxx;
        double cx[2], cy[2];
        for (f = 0; f < F; f++) cx[f] = mas_correction(prefact * kx, mas_index[f]);
        for (kyy = 0; kyy < middle + 1; kyy++) {
            int ky = kyy;
            double k, re[2], im[2];
            int k_index;
            for (f = 0; f < F; f++) cy[f] = mas_correction(prefact * ky, mas_index[f]);
            if (ky == 0 || (ky == middle && grid % 2 == 0))
                if (kx < 0) continue;
            k = sqrt((double)(kx * kx + ky * ky));
            k_index = (int)k;
            for (f = 0; f < F; f++) {
                long e = (long)f * grid * zlen + (long)kxx * zlen + kyy;
                if (is_f64) {
                    const double *d = (const double *)dk + 2 * e;
                    double mf = cx[f] * cy[f];
                    re[f] = d[0] * mf; im[f] = d[1] * mf;
                } else {
                    const float *d = (const float *)dk + 2 * e;
                    float mf = (float)(cx[f] * cy[f]);
                    float r = d[0] * mf, m = d[1] * mf;
                    re[f] = (double)r; im[f] = (double)m;
                }
            }
            k2D[k_index] += k;
            N[k_index] += 1.0;
            if (F == 1) {
                Pk[k_index] += re[0] * re[0] + im[0] * im[0];
            } else {
                Pk[(long)k_index * 3 + 0] += re[0] * re[0] + im[0] * im[0];
                Pk[(long)k_index * 3 + 1] += re[1] * re[1] + im[1] * im[1];
                Pk[(long)k_index * 3 + 2] += re[0] * re[1] + im[0] * im[1];
            }
        }
    }
}

/*
 * MAS deconvolution of correct_MAS (Pk_library.pyx:1910-1929) applied in place
 * to a half-complex cube (dims, dims, dims/2+1): independent modes are scaled
 * by the float MAS factor in complex64 (or double for is_f64); the redundant
 * half of the kz = 0 and kz = middle planes is left as it is, as the reference
 * does.  The caller runs the unnormalised inverse transform.
 */
void ora_correct_mas_modes(void *dk, int is_f64, int dims, int mas_index)
{
    int middle = dims / 2;
    long zlen = middle + 1;
    double prefact = M_PI / dims;
    int kxx, kyy, kzz;
    for (kxx = 0; kxx < dims; kxx++) {
        int kx = (kxx > middle) ? kxx - dims : kxx;
        double c0 = mas_correction(prefact * kx, mas_index);
        for (kyy = 0; kyy < dims; kyy++) {
            int ky = (kyy > middle) ? kyy - dims : kyy;
            double c1 = mas_correction(prefact * ky, mas_index);
            for (kzz = 0; kzz < middle + 1; kzz++) {
                int kz = kzz;
                double c2 = mas_correction(prefact * kz, mas_index);
                long e = ((long)kxx * dims + kyy) * zlen + kzz;
                if (kz == 0 || (kz == middle && dims % 2 == 0)) {
                    if (kx < 0) continue;
                    else if (kx == 0 || (kx == middle && dims % 2 == 0)) {
                        if (ky < 0) continue;
                    }
                }
                if (is_f64) {
                    double *d = (double *)dk + 2 * e;
                    double mf = c0 * c1 * c2;
                    d[0] *= mf; d[1] *= mf;
                } else {
                    float *d = (float *)dk + 2 * e;
                    float mf = (float)(c0 * c1 * c2);
                    d[0] = d[0] * mf; d[1] = d[1] * mf;
                }
            }
        }
    }
}
