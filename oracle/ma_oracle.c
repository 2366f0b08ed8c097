/*
 * ORACLE -- test infrastructure only.  Nothing in the product path may link,
 * load or call this file; only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py use it, and only as the checker.
 *
 * CPU restatement of the reference mass-assignment (MA) arithmetic:
 *   library/MAS_library/MAS_library.pyx:57-112   MA() dispatch, 2D expand_dims path
 *   library/MAS_library/MAS_library.pyx:123-166  CIC     :179-216 CICW
 *   library/MAS_library/MAS_library.pyx:273-292  NGP     :305-325 NGPW
 *   library/MAS_library/MAS_library.pyx:369-404  TSC     :417-452 TSCW
 *   library/MAS_library/MAS_library.pyx:463-497  PCS     :510-545 PCSW
 *   library/MAS_library/MAS_c.c:8-266            native-2D C kernels
 *
 * Build with -ffp-contract=off (oracle/Makefile): every float expression below
 * is evaluated in exactly the C types the reference uses (float products,
 * double temporaries where a double literal appears), serially, in the
 * reference's particle order.
 *
 * Pinned by tests/test_oracle_ma.py: closed-form single-particle footprints,
 * partition of unity, periodic wrap, float64 brute force on tiny grids, and
 * bitwise equality with the reference's own MAS_c.c compiled by
 * oracle/Makefile into oracle/_ref/ (strict flags).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_NGP 1
#define ORA_CIC 2
#define ORA_TSC 3
#define ORA_PCS 4

static inline int wrap_index(long i, int dims)
{
    /* (i+dims)%dims for i in [-dims, 2*dims): the reference's own wrap
     * (MAS_library.pyx:396,489).  Positions far outside [0,BoxSize] are
     * undefined behaviour in the reference; here they are wrapped. */
    long r = i % dims;
    if (r < 0) r += dims;
    return (int)r;
}

/* Per-axis footprint in float, exactly as the reference computes it.
 * Returns the footprint width F and fills idx[0..F) and C[0..F).          */
static int axis_weights_f32(float x, float inv_cell_size, int dims, int mas,
                            int idx[4], float C[4])
{
    float dist, diff;
    int j, minimum;
    switch (mas) {
    case ORA_NGP: /* MAS_library.pyx:288-290: <int>(pos*inv + 0.5), 0.5 is a double */
        idx[0] = wrap_index((long)(int)((double)(x * inv_cell_size) + 0.5), dims);
        C[0] = 1.0f;
        return 1;
    case ORA_CIC: /* MAS_library.pyx:151-156 */
        dist = x * inv_cell_size;
        C[1] = dist - (float)(int)dist;          /* u */
        C[0] = (float)(1.0 - (double)C[1]);      /* d = 1.0 - u, double literal */
        idx[0] = wrap_index((long)(int)dist, dims);
        idx[1] = wrap_index((long)idx[0] + 1, dims);
        return 2;
    case ORA_TSC: /* MAS_library.pyx:392-399 */
        dist = x * inv_cell_size;
        minimum = (int)floor((double)dist - 1.5);
        for (j = 0; j < 3; j++) {
            idx[j] = wrap_index((long)minimum + j + 1 + dims, dims);
            diff = (float)fabs((double)((float)(minimum + j + 1) - dist));
            if (diff < 0.5f) {
                float d2 = diff * diff;                       /* float product */
                C[j] = (float)(0.75 - (double)d2);
            } else if (diff < 1.5f) {
                C[j] = (float)(0.5 * (1.5 - (double)diff) * (1.5 - (double)diff));
            } else {
                C[j] = 0.0f;
            }
        }
        return 3;
    case ORA_PCS: /* MAS_library.pyx:485-492 */
        dist = x * inv_cell_size;
        minimum = (int)floor((double)dist - 2.0);
        for (j = 0; j < 4; j++) {
            double dd;
            idx[j] = wrap_index((long)minimum + j + 1 + dims, dims);
            diff = (float)fabs((double)((float)(minimum + j + 1) - dist));
            dd = (double)diff;
            if (diff < 1.0f)
                C[j] = (float)((4.0 - 6.0 * dd * dd + 3.0 * dd * dd * dd) / 6.0);
            else if (diff < 2.0f)
                C[j] = (float)((2.0 - dd) * (2.0 - dd) * (2.0 - dd) / 6.0);
            else
                C[j] = 0.0f;
        }
        return 4;
    }
    return 0;
}

/*
 * Float32 mass assignment in the reference's serial order.
 *   coord      2 or 3 (columns of pos)
 *   number     flat grid: dims^3 (coord==3) or dims^2 (coord==2), += semantics
 *   cython2d   coord==2 only: 1 = MA()'s expand_dims path (the 3D kernel on a
 *              (d,d,1) grid, every (l,m) deposit repeated F times into the same
 *              cell, MAS_library.pyx:84-103; renormalisation is left to the
 *              caller), 0 = the native 2D C kernel (MAS_c.c:28-34, once).
 * inv_cell_size = dims/BoxSize in float (MAS_library.pyx:135).
 */
void ora_ma_f32(const float *pos, long n, int coord, const float *W,
                float *number, int dims, float BoxSize, int mas, int cython2d)
{
    long i;
    float inv_cell_size = (float)dims / BoxSize;
    int idx[3][4];
    float C[3][4];
    int F = 1, Fz, l, m, q, a, b;
    for (a = 0; a < 3; a++)
        for (b = 0; b < 4; b++) { idx[a][b] = 0; C[a][b] = 1.0f; }

    for (i = 0; i < n; i++) {
        int axis;
        for (axis = 0; axis < coord; axis++)
            F = axis_weights_f32(pos[(long)coord * i + axis], inv_cell_size, dims,
                                 mas, idx[axis], C[axis]);
        if (coord == 3) Fz = F;
        else Fz = cython2d ? F : 1;
        for (l = 0; l < F; l++)
            for (m = 0; m < F; m++)
                for (q = 0; q < Fz; q++) {
                    long cell;
                    float w = C[0][l] * C[1][m];
                    w = w * C[2][q];
                    if (W) w = w * W[i];
                    if (coord == 3)
                        cell = ((long)idx[0][l] * dims + idx[1][m]) * dims + idx[2][q];
                    else
                        cell = (long)idx[0][l] * dims + idx[1][m];
                    number[cell] += w;
                }
    }
}

/*
 * Float64 oracle: the same footprints with dist, weights and accumulation in
 * double (positions and weights are the float32 inputs, widened exactly).
 * 2D is the native single deposit.
 */
/* Float64 footprint of one coordinate: the B-spline definitions evaluated in
 * double from the float32 position (MAS_library.pyx:123-545 without roundings). */
static int f64_axis(double dist, int dims, int mas, int *idx, double *C)
{
    int mn, j;
    switch (mas) {
    case ORA_NGP:
        idx[0] = wrap_index((long)floor(dist + 0.5), dims);
        C[0] = 1.0;
        return 1;
    case ORA_CIC:
        mn = (int)floor(dist);
        C[1] = dist - mn; C[0] = 1.0 - C[1];
        idx[0] = wrap_index(mn, dims);
        idx[1] = wrap_index((long)mn + 1, dims);
        return 2;
    case ORA_TSC:
        mn = (int)floor(dist - 1.5);
        for (j = 0; j < 3; j++) {
            double diff = fabs(mn + j + 1 - dist);
            idx[j] = wrap_index((long)mn + j + 1, dims);
            C[j] = diff < 0.5 ? 0.75 - diff * diff
                 : diff < 1.5 ? 0.5 * (1.5 - diff) * (1.5 - diff) : 0.0;
        }
        return 3;
    case ORA_PCS:
        mn = (int)floor(dist - 2.0);
        for (j = 0; j < 4; j++) {
            double diff = fabs(mn + j + 1 - dist);
            idx[j] = wrap_index((long)mn + j + 1, dims);
            C[j] = diff < 1.0 ? (4.0 - 6.0 * diff * diff + 3.0 * diff * diff * diff) / 6.0
                 : diff < 2.0 ? (2.0 - diff) * (2.0 - diff) * (2.0 - diff) / 6.0 : 0.0;
        }
        return 4;
    default:
        return 0;
    }
}

/* Adds particle i's float64 footprint into number (atomically when `atomic`). */
static void f64_particle(const float *pos, long i, int coord, const float *W, double *number,
                         int dims, double inv, int mas, int atomic)
{
    int idx[3][4];
    double C[3][4];
    int F = 1, Fz, l, m, q, a, b;
    for (a = 0; a < 3; a++)
        for (b = 0; b < 4; b++) { idx[a][b] = 0; C[a][b] = 1.0; }
    for (a = 0; a < coord; a++)
        F = f64_axis((double)pos[(long)coord * i + a] * inv, dims, mas, idx[a], C[a]);
    Fz = (coord == 3) ? F : 1;
    for (l = 0; l < F; l++)
        for (m = 0; m < F; m++)
            for (q = 0; q < Fz; q++) {
                long cell;
                double w = C[0][l] * C[1][m] * C[2][q];
                if (W) w *= (double)W[i];
                if (coord == 3)
                    cell = ((long)idx[0][l] * dims + idx[1][m]) * dims + idx[2][q];
                else
                    cell = (long)idx[0][l] * dims + idx[1][m];
                if (atomic) {
#pragma omp atomic
                    number[cell] += w;
                } else {
                    number[cell] += w;
                }
            }
}

void ora_ma_f64(const float *pos, long n, int coord, const float *W,
                double *number, int dims, double BoxSize, int mas)
{
    long i;
    const double inv = (double)dims / BoxSize;
    if (mas < ORA_NGP || mas > ORA_PCS) return;
    for (i = 0; i < n; i++) f64_particle(pos, i, coord, W, number, dims, inv, mas, 0);
}

/* The same sums with `threads` OpenMP threads, deterministic and without atomics.
 * Particles are grouped by the x slab of their first footprint node (2T slabs of
 * at least F-1 planes, kept in input order within a slab); even slabs run in
 * parallel, then odd slabs.  A particle of slab s writes only planes of slabs s
 * and s+1 (periodically), so no two threads of a phase touch the same cell.  Per
 * cell, the additions come in a fixed order that differs from the serial one
 * only at the 1e-16 level of double rounding.  Falls back to the serial loop
 * when the grid is too thin for two slabs per thread. */
void ora_ma_f64_threads(const float *pos, long n, int coord, const float *W,
                        double *number, int dims, double BoxSize, int mas, int threads)
{
    const double inv = (double)dims / BoxSize;
    const int F = mas;                      /* footprint width: NGP 1 .. PCS 4 */
    int T = threads < 1 ? 1 : threads;
    int nslab, s;
    long i, *cnt, *start;
    unsigned *order;
    if (mas < ORA_NGP || mas > ORA_PCS) return;
    nslab = 2 * T;
    while (nslab > 2 && (dims / nslab < F - 1 || dims / nslab < 1)) nslab -= 2;
    if (T == 1 || nslab < 2 || dims / nslab < F - 1 || dims / nslab < 1 || n >= 4294967295L) {
        ora_ma_f64(pos, n, coord, W, number, dims, BoxSize, mas);
        return;
    }
    {
        unsigned char *slab_of = (unsigned char *)malloc((size_t)n);
        long *tcnt = (long *)calloc((size_t)T * nslab, sizeof(long));
        cnt = (long *)calloc((size_t)nslab + 1, sizeof(long));
        start = (long *)calloc((size_t)nslab + 1, sizeof(long));
        order = (unsigned *)malloc(sizeof(unsigned) * (size_t)(n > 0 ? n : 1));
        if (!slab_of || !tcnt || !cnt || !start || !order) {
            free(slab_of); free(tcnt); free(cnt); free(start); free(order);
            ora_ma_f64(pos, n, coord, W, number, dims, BoxSize, mas);
            return;
        }
        /* slab of each particle, per-thread counts over contiguous particle ranges */
#pragma omp parallel num_threads(T)
        {
            int t = 0, k;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            long a = n * t / T, b = n * (t + 1) / T, j;
            for (j = a; j < b; j++) {
                int idx[4];
                double C[4];
                f64_axis((double)pos[(long)coord * j] * inv, dims, mas, idx, C);
                k = (int)((long)idx[0] * nslab / dims);
                slab_of[j] = (unsigned char)k;
                tcnt[(long)t * nslab + k]++;
            }
        }
        /* stable order: slab, then thread range, then particle index */
        {
            long run = 0;
            int t;
            for (s = 0; s < nslab; s++) {
                start[s] = run;
                for (t = 0; t < T; t++) {
                    long c = tcnt[(long)t * nslab + s];
                    tcnt[(long)t * nslab + s] = run;
                    run += c;
                }
            }
            start[nslab] = run;
        }
#pragma omp parallel num_threads(T)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            long a = n * t / T, b = n * (t + 1) / T, j;
            for (j = a; j < b; j++) order[tcnt[(long)t * nslab + slab_of[j]]++] = (unsigned)j;
        }
        free(slab_of);
        free(tcnt);
    }
    for (int phase = 0; phase < 2; phase++) {
#pragma omp parallel for num_threads(T) schedule(dynamic, 1)
        for (s = phase; s < nslab; s += 2) {
            long j;
            for (j = start[s]; j < start[s + 1]; j++)
                f64_particle(pos, (long)order[j], coord, W, number, dims, inv, mas, 0);
        }
    }
    free(cnt); free(start); free(order);
}

/*
 * CIC interpolation of a 3D grid at particle positions, the gather adjoint of CIC:
 * library/MAS_library/MAS_library.pyx:558-599.  Float arithmetic, terms summed
 * left to right in the reference's order.
 */
void ora_cic_interp(const float *density, int dims, float BoxSize, const float *pos, long n,
                    float *den)
{
    long i;
    float inv_cell_size = (float)dims / BoxSize;
    for (i = 0; i < n; i++) {
        float u[3], d[3];
        int id[3], iu[3], axis;
        for (axis = 0; axis < 3; axis++) {
            float dist = pos[3 * i + axis] * inv_cell_size;
            u[axis] = dist - (float)(int)dist;
            d[axis] = (float)(1.0 - (double)u[axis]);
            id[axis] = wrap_index((long)(int)dist, dims);
            iu[axis] = wrap_index((long)id[axis] + 1, dims);
        }
#define G(a, b, c) density[((long)(a) * dims + (b)) * dims + (c)]
        den[i] = G(id[0], id[1], id[2]) * d[0] * d[1] * d[2] +
                 G(id[0], id[1], iu[2]) * d[0] * d[1] * u[2] +
                 G(id[0], iu[1], id[2]) * d[0] * u[1] * d[2] +
                 G(id[0], iu[1], iu[2]) * d[0] * u[1] * u[2] +
                 G(iu[0], id[1], id[2]) * u[0] * d[1] * d[2] +
                 G(iu[0], id[1], iu[2]) * u[0] * d[1] * u[2] +
                 G(iu[0], iu[1], id[2]) * u[0] * u[1] * d[2] +
                 G(iu[0], iu[1], iu[2]) * u[0] * u[1] * u[2];
#undef G
    }
}

/* Single-axis footprint, exposed so tests can pin the closed forms. */
int ora_axis_weights(float x, float inv_cell_size, int dims, int mas,
                     int *idx_out, float *C_out)
{
    return axis_weights_f32(x, inv_cell_size, dims, mas, idx_out, C_out);
}

/*
 * Redshift-space shift, library/redshift_space_library/redshift_space_library.pyx:29-46:
 *   factor = (1+z)/H ; x += v*factor ; while x<0: x+=L ; if x>L: x = fmod(x, L)
 */
void ora_pos_redshift_space(float *pos, const float *vel, long n, float BoxSize,
                            float Hubble, float redshift, int axis)
{
    long i;
    float factor = (float)((1.0 + (double)redshift) / (double)Hubble);
    for (i = 0; i < n; i++) {
        float *p = &pos[3 * i + axis];
        *p = *p + vel[3 * i + axis] * factor;
        while (*p < 0.0f) *p += BoxSize;
        if (*p > BoxSize) *p = fmodf(*p, BoxSize);
    }
}
