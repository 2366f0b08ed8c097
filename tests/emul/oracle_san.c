/*
 * Sanitizer driver for the oracle's C code (oracle/ma_oracle.c, oracle/pk_oracle.c),
 * built with AddressSanitizer and UndefinedBehaviorSanitizer by
 * tests/test_sanitizers.py and run on the CPU.  It calls every exported oracle
 * routine on grids from 1 to 33 nodes per axis, all schemes, 2D and 3D, with and
 * without weights, positions on the box edges, and checks what needs no other
 * reference: the deposited mass equals the particle mass (partition of unity) and
 * the mode counts add up to the closed form of check_number_modes
 * (Pk_library.pyx:87-99).  Any out-of-bounds access or undefined operation aborts.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/ma_oracle.c"
#include "../../oracle/pk_oracle.c"

static unsigned long long rng_state = 88172645463325252ULL;
static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int fail(const char *what, int dims, int mas, int coord)
{
    printf("FAIL %s dims=%d mas=%d coord=%d\n", what, dims, mas, coord);
    return 1;
}

static int check_ma(int dims, int mas, int coord, int weighted)
{
    const long n = 3000;
    const float box = 37.5f;
    float *pos = (float *)malloc(sizeof(float) * n * coord);
    float *W = weighted ? (float *)malloc(sizeof(float) * n) : NULL;
    long cells = (long)dims * dims * (coord == 3 ? dims : 1);
    float *g = (float *)calloc(cells, sizeof(float));
    double *g64 = (double *)calloc(cells, sizeof(double));
    double *g64t = (double *)calloc(cells, sizeof(double));
    double mass = 0.0, s32 = 0.0, s64 = 0.0, s64t = 0.0;
    long i, c;
    int bad = 0, F = mas;
    for (i = 0; i < n * coord; i++) pos[i] = (float)(urand() * box);
    /* edges: 0, the box length itself, just below it, exact node positions */
    for (i = 0; i < coord; i++) {
        pos[i] = 0.0f;
        pos[coord + i] = box;
        pos[2 * coord + i] = nextafterf(box, 0.0f);
        pos[3 * coord + i] = box / dims * (dims / 2);
    }
    for (i = 0; i < n; i++) {
        double w = weighted ? (W[i] = (float)(0.1 + 3.0 * urand())) : 1.0;
        mass += w;
    }
    ora_ma_f32(pos, n, coord, W, g, dims, box, mas, coord == 2 ? 1 : 0);
    ora_ma_f64(pos, n, coord, W, g64, dims, box, mas);
    ora_ma_f64_threads(pos, n, coord, W, g64t, dims, box, mas, 3);
    for (c = 0; c < cells; c++) {
        s32 += g[c];
        s64 += g64[c];
        s64t += g64t[c];
    }
    /* the 2D expand_dims path repeats each deposit F times */
    if (coord == 2) s32 /= F;
    /* float32 sums: on grids of a few nodes one cell takes thousands of additions */
    if (fabs(s32 - mass) > (dims >= 4 ? 1e-4 : 2e-2) * mass) bad |= fail("f32 mass", dims, mas, coord);
    if (fabs(s64 - mass) > 1e-9 * mass) bad |= fail("f64 mass", dims, mas, coord);
    if (fabs(s64t - mass) > 1e-9 * mass) bad |= fail("f64 threads mass", dims, mas, coord);
    if (coord == 3) {
        float *den = (float *)malloc(sizeof(float) * n);
        ora_cic_interp(g, dims, box, pos, n, den);
        free(den);
    }
    {
        float *vel = (float *)malloc(sizeof(float) * n * 3);
        float *p3 = (float *)malloc(sizeof(float) * n * 3);
        for (i = 0; i < n * 3; i++) {
            vel[i] = (float)((urand() - 0.5) * 4000.0);
            p3[i] = (float)(urand() * box);
        }
        ora_pos_redshift_space(p3, vel, n, box, 100.0f, 0.5f, (int)(dims % 3));
        for (i = 0; i < n * 3; i++)
            if (!(p3[i] >= 0.0f && p3[i] <= box)) { bad |= fail("rsd wrap", dims, mas, coord); break; }
        free(vel);
        free(p3);
    }
    free(pos);
    free(W);
    free(g);
    free(g64);
    free(g64t);
    return bad;
}

static int check_pk(int dims, int F, int is_f64, int axis)
{
    int middle = dims / 2;
    int kmax_per = (int)sqrt((double)(2 * middle * middle));
    int kmax = (int)sqrt((double)(3 * middle * middle));
    long nb2 = (long)(middle + 1) * (kmax_per + 1);
    int S = F + F * (F - 1) / 2, f, mas[3] = {2, 3, 4}, bad = 0;
    long nk = (long)dims * dims * (middle + 1), i;
    size_t esz = is_f64 ? sizeof(double) : sizeof(float);
    void *dk = malloc(esz * 2 * nk * F);
    double *k1 = calloc(middle + 1, 8), *P1 = calloc((middle + 1) * S, 8), *N1 = calloc(middle + 1, 8);
    double *P2 = calloc(nb2 * S, 8), *N2 = calloc(nb2, 8);
    double *k3 = calloc(kmax + 1, 8), *P3 = calloc((kmax + 1) * 3 * S, 8), *ph = calloc(kmax + 1, 8);
    double *N3 = calloc(kmax + 1, 8);
    double total = 0.0;
    long own = (dims % 2) ? 1 : 8;
    for (i = 0; i < 2 * nk * F; i++) {
        if (is_f64) ((double *)dk)[i] = urand() - 0.5;
        else ((float *)dk)[i] = (float)(urand() - 0.5);
    }
    ora_pk_loop(dk, is_f64, F, dims, axis, mas, k1, P1, N1, P2, N2, k3, P3, F == 1 ? ph : NULL, N3);
    for (i = 0; i <= kmax; i++) total += N3[i];
    if (total != (double)(((long)dims * dims * dims - own) / 2 + own)) bad |= fail("Nmodes", dims, F, axis);
    memset(N3, 0, (kmax + 1) * 8);
    ora_pk_loop_threads(dk, is_f64, F, dims, axis, mas, k1, P1, N1, P2, N2, k3, P3,
                        F == 1 ? ph : NULL, N3, 3);
    total = 0.0;
    for (i = 0; i <= kmax; i++) total += N3[i];
    if (total != (double)(((long)dims * dims * dims - own) / 2 + own)) bad |= fail("Nmodes threads", dims, F, axis);
    if (F <= 2) {
        /* the first plane of each field as a 2D half-complex map (grid = dims) */
        int kp = (int)sqrt((double)(2 * middle * middle));
        double *k2 = calloc(kp + 1, 8), *Pp = calloc((kp + 1) * 3, 8), *Np = calloc(kp + 1, 8);
        ora_pk_plane_loop(dk, is_f64, F, dims, mas, k2, Pp, Np);
        free(k2); free(Pp); free(Np);
    }
    ora_correct_mas_modes(dk, is_f64, dims, 3);
    free(dk); free(k1); free(P1); free(N1); free(P2); free(N2); free(k3); free(P3); free(ph); free(N3);
    return bad;
}

int main(void)
{
    int bad = 0, dims, mas, coord, w, F, axis, is64;
    for (dims = 1; dims <= 33; dims += (dims < 9 ? 1 : 8))
        for (mas = 1; mas <= 4; mas++)
            for (coord = 2; coord <= 3; coord++)
                for (w = 0; w <= 1; w++) bad |= check_ma(dims, mas, coord, w);
    for (dims = 2; dims <= 17; dims += 3)
        for (F = 1; F <= 3; F++)
            for (is64 = 0; is64 <= 1; is64++)
                for (axis = 0; axis <= 2; axis++) bad |= check_pk(dims, F, is64, axis);
    {
        int idx[4];
        float C[4];
        bad |= ora_axis_weights(3.25f, 1.0f, 8, 4, idx, C) != 4;
        bad |= ora_mas_correction(0.0, 3) != 1.0;
    }
    printf(bad ? "oracle sanitizer run: FAILED\n" : "oracle sanitizer run: ok\n");
    return bad;
}
