/*
 * pylians_hip.h -- C ABI of libpylians_hip.so, the MI355X (gfx950) density-field
 * and power-spectrum path.
 *
 * All entry points are extern "C", take plain pointers and sizes, return an int
 * status (PYL_OK = 0, negative on error; pyl_last_error() gives a thread-local
 * message), and never abort the process.  Host pointers are the default
 * (the library stages them through device memory); PYL_PTR_DEVICE marks
 * device-resident inputs/outputs.  A context (pyl_ctx) owns one HIP stream,
 * grow-only scratch memory within an HBM budget, and resident FFT and
 * mode-count plans.  One context per host thread.
 *
 * Reference interfaces replaced (library/ paths are in the reference tree):
 *   pyl_ma            MAS_library.MA                library/MAS_library/MAS_library.pyx:57-112
 *   pyl_ma_delta      MA + "delta /= mean; delta -= 1.0"  library/Pk_library/Pk_snapshot.py:80-88
 *   pyl_ma_rsd        pos_redshift_space + MA       redshift_space_library.pyx:29-46 + MAS_library.pyx:57-112
 *                     and its kernels NGP..PCSW      MAS_library.pyx:123-545
 *   NGP/CIC/TSC/PCS   the C kernels, same signature  library/MAS_library/MAS_c.h:3-10 (MAS_c.c:8-266)
 *   pyl_fft_r2c       FFT3Dr_f                       library/Pk_library/Pk_library.pyx:117-130
 *   pyl_pk            Pk (F=1) and XPk (F>=2)        Pk_library.pyx:263-420, :529-793
 *   pyl_pk_from_fourier  the mode loops alone        Pk_library.pyx:311-378, :623-732
 *   pyl_mode_counts   Nmodes/k sums of those loops and check_number_modes
 *                                                    Pk_library.pyx:87-99, :331-378
 *   pyl_mas_correction  MAS_correction               Pk_library.pyx:83-84
 *   pyl_pos_redshift_space  pos_redshift_space       library/redshift_space_library/
 *                                                    redshift_space_library.pyx:29-46
 *   pyl_pk_plane      Pk_plane (F=1), XPk_plane (F=2) Pk_library.pyx:435-511, :1093-1220
 *   pyl_correct_mas   correct_MAS                    Pk_library.pyx:1882-1939
 */
#ifndef PYLIANS_HIP_H
#define PYLIANS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PYL_OK        0
#define PYL_EINVAL   -1   /* bad argument (the reference's sys.exit()/ValueError cases) */
#define PYL_EHIP     -2   /* HIP runtime error */
#define PYL_ENOMEM   -3   /* device allocation failed */
#define PYL_EFFT     -4   /* rocFFT error */
#define PYL_EMODES   -5   /* check_number_modes failed */
#define PYL_EBUDGET  -6   /* request exceeds the context's HBM budget */

/* ---- mass-assignment schemes (Pk_library.pyx:72-78 MAS_function) ----- */
#define PYL_MAS_NONE 0
#define PYL_NGP      1
#define PYL_CIC      2
#define PYL_TSC      3
#define PYL_PCS      4

/* ---- flags ------------------------------------------------------------ */
#define PYL_PTR_DEVICE     0x0001  /* array arguments are device pointers */
#define PYL_MA_ATOMIC      0x0002  /* MA: plain per-particle global float atomics (the
                                      non-deterministic "plain GPU form") */
#define PYL_MA_CYTHON_2D   0x0004  /* MA, axes==2: MA()'s expand_dims semantics, i.e. each
                                      deposit counted F times (MAS_library.pyx:84-110) */
#define PYL_MA_RENORM_2D   0x0008  /* with PYL_MA_CYTHON_2D: divide the whole grid by F */
#define PYL_MA_OVERWRITE   0x0010  /* MA: grid = deposit instead of grid += deposit */
#define PYL_NO_SYNC        0x0020  /* device pointers: return without synchronising the stream */
#define PYL_PK_DESTROY_INPUT 0x0040 /* Pk, device delta: the FFT may overwrite delta */
#define PYL_PK_NO_PHASE    0x0080  /* Pk: skip the Pkphase accumulator */
#define PYL_PK_SCALES_DEVICE 0x0100 /* pyl_pk_batch: scales and shifts are device arrays */

/* ---- context ---------------------------------------------------------- */
typedef struct pyl_ctx pyl_ctx;

const char *pyl_last_error(void);
int pyl_version(void);
/* hbm_budget: bytes of device scratch the context may hold (0 = 90% of free memory). */
int pyl_ctx_create(pyl_ctx **out, int device, size_t hbm_budget);
int pyl_ctx_destroy(pyl_ctx *ctx);
/* Use an external HIP stream (hipStream_t) for all work; NULL restores the own stream.
 * When the stream changes, the new stream first waits for the work already queued on
 * the old one (the context's scratch is shared by its calls). */
int pyl_ctx_set_stream(pyl_ctx *ctx, void *hip_stream);
int pyl_ctx_synchronize(pyl_ctx *ctx);
/* Release scratch, FFT plans and mode-count plans. */
int pyl_ctx_trim(pyl_ctx *ctx);
/* Bytes of device memory the context currently holds. */
int pyl_ctx_memory(pyl_ctx *ctx, size_t *bytes_out);
/* Per-stage device times (ms, HIP events on the context stream) of the last call.
 * Stage ids are PYL_STAGE_*; enable=1 turns event recording on. */
#define PYL_STAGE_MA_COUNT   0
#define PYL_STAGE_MA_SCAN    1
#define PYL_STAGE_MA_SCATTER 2
#define PYL_STAGE_MA_DEPOSIT 3
#define PYL_STAGE_MA_ATOMIC  4
#define PYL_STAGE_PK_PREP    5
#define PYL_STAGE_PK_FFT     6
#define PYL_STAGE_PK_BIN     7
#define PYL_STAGE_MA_FINAL   8   /* halo finalize of the tiled MA */
#define PYL_STAGE_MA_BUCKET  9   /* first pass of the two-level sort (particles into buckets) */
#define PYL_STAGE_MA_SPILL   10  /* capacity sort: ordering of the particles its tile slots missed */
#define PYL_STAGE_COUNT      11
int pyl_ctx_enable_timing(pyl_ctx *ctx, int enable);   /* all stages; also clears the totals */
/* Times only the stages whose bit (1u << PYL_STAGE_*) is set: each timed stage adds
 * two HIP events to the stream, so a caller timing one kernel records only its own. */
int pyl_ctx_enable_stage_timing(pyl_ctx *ctx, unsigned stage_mask);
/* Stage times of the latest call (ms; 0 for stages it did not run). */
int pyl_ctx_stage_times(pyl_ctx *ctx, double *ms_out /* [PYL_STAGE_COUNT] */);
/* Summed stage times and call counts since timing was enabled; recording them
 * needs no synchronisation between calls (read back here, after the stream). */
int pyl_ctx_stage_totals(pyl_ctx *ctx, double *ms_out /* [PYL_STAGE_COUNT] */,
                         long long *count_out /* [PYL_STAGE_COUNT] */);

/* ---- mass assignment (MAS_library.pyx:57-545) --------------------------
 * pos     (n, axes) float32, C order (x,y[,z] per particle)
 * W       (n,) float32 or NULL
 * grid    (dims,dims[,dims]) float32, C order; grid += deposit (or = with OVERWRITE)
 * mas     PYL_NGP..PYL_PCS
 * Default algorithm: particles binned by home tile with a device counting sort,
 * each tile plus its halo summed in LDS with exact (order-independent)
 * accumulation, the halos added in a fixed order -- bitwise reproducible.
 * PYL_MA_ATOMIC selects per-particle global atomics.
 */
int pyl_ma(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
           float *grid, int dims, float BoxSize, int mas, int flags);

/* MA that allocates nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, algorithm (tiled) and
 * bits as pyl_ma with PYL_PTR_DEVICE, one grid, device pointers only; every piece
 * of scratch comes from one fixed per-device array of PYL_MA_NOALLOC_SCRATCH bytes
 * that the library holds from load until the process ends, and counters in it are
 * cleared by fill kernels (no hipMalloc, hipFree, hipMemsetAsync or staging upload).
 * A call that would need more scratch (about 12 B per particle, 16 with W, plus the
 * sort tables and tile partials) returns PYL_EBUDGET before launching anything.
 * All calls on one device share the array: calls that overlap on one GPU (other
 * contexts, other streams without ordering) race and must be serialised by the caller. */
#define PYL_MA_NOALLOC_SCRATCH ((size_t)128 << 20)
int pyl_ma_noalloc(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
                   float *grid, int dims, float BoxSize, int mas, int flags);

/* MA fused with the density contrast: delta = deposit/scale - shift in float32,
 * written once (overwrite).  The caller's two steps after MA in the reference
 * pipeline (library/Pk_library/Pk_snapshot.py:80-88, "delta /= mean; delta -= 1.0")
 * with scale = the mean and shift = 1; scale = 1, shift = 0 gives pyl_ma with
 * PYL_MA_OVERWRITE. */
int pyl_ma_delta(pyl_ctx *ctx, const float *pos, int64_t n, int axes, const float *W,
                 float *delta, int dims, float BoxSize, int mas, float scale, float shift,
                 int flags);

/* MA of redshift-space positions, the shift of pos_redshift_space
 * (library/redshift_space_library/redshift_space_library.pyx:29-46) fused into the
 * position load: each particle is deposited at pos[:,axis] + vel[:,axis]*(1+z)/H,
 * wrapped as the reference does; pos itself is not modified.  pos and vel are
 * (n,3) float32.  With PYL_MA_OVERWRITE the grid receives deposit/scale - shift
 * (as pyl_ma_delta); without it, grid += deposit and scale/shift must be 1/0. */
int pyl_ma_rsd(pyl_ctx *ctx, const float *pos, const float *vel, int64_t n, const float *W,
               float *grid, int dims, float BoxSize, int mas, float Hubble, float redshift,
               int axis, float scale, float shift, int flags);

/* Batched MA: nreal independent grids in one launch sequence.  Particles of grid r are
 * pos[offsets[r] .. offsets[r+1]) (host array, nreal+1 entries, offsets[0] = 0), W is
 * concatenated the same way (or NULL), grids holds nreal contiguous grids.  Same
 * semantics as nreal calls of pyl_ma (each grid's result is bitwise the same as alone).
 * Replaces a Python loop of MA calls (MAS_library.pyx:57-112) over realisations. */
int pyl_ma_batch(pyl_ctx *ctx, const float *pos, const int64_t *offsets, int nreal, int axes,
                 const float *W, float *grids, int dims, float BoxSize, int mas, int flags);

/* Batched MA with the particle ranges in device memory: realisation r holds particles
 * d_offsets[r] .. d_offsets[r+1] (int64, device, nreal+1 entries, d_offsets[0] = 0 and
 * d_offsets[nreal] <= n_cap).  Each realisation is cut into chunks_per_real sort
 * chunks on the device, so the launch sequence depends only on (nreal, n_cap,
 * chunks_per_real, dims, MAS) and not on the counts: the form a HIP graph can replay
 * with new particles.  Device pointers only, tiled algorithm; grids as pyl_ma_batch.
 * Replaces the same Python loop of MA calls as pyl_ma_batch (MAS_library.pyx:57-112). */
int pyl_ma_batch_dev(pyl_ctx *ctx, const float *pos, const int64_t *d_offsets, int nreal,
                     int64_t n_cap, int chunks_per_real, int axes, const float *W, float *grids,
                     int dims, float BoxSize, int mas, int flags);

/* Reference C kernels, same names and signatures as MAS_c.h:3-10 (host pointers,
 * threads ignored, += semantics, native 2D as in MAS_c.c:28-34).  They run on
 * a process-wide context on device 0.  On error they print pyl_last_error() to
 * stderr and leave number unchanged. */
void NGP(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads);
void CIC(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads);
void TSC(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads);
void PCS(float *pos, float *number, float *W, long particles, int dims, int axes,
         float BoxSize, int threads);

/* CIC interpolation of a (dims,dims,dims) float32 grid at n positions (n,3):
 * den[i] = sum of the 8 CIC-weighted neighbours (gather adjoint of CIC).
 * Replaces CIC_interp, MAS_library.pyx:558-599. */
int pyl_cic_interp(pyl_ctx *ctx, const float *density, int dims, float BoxSize,
                   const float *pos, int64_t n, float *den, int flags);

/* ---- FFT (Pk_library.pyx:117-130) -------------------------------------
 * Forward, unnormalised r2c of a (dims,dims,dims) float32 cube into a
 * (dims,dims,dims/2+1) complex64 array (interleaved re,im).  delta is not
 * modified (unless PYL_PK_DESTROY_INPUT with device pointers). */
int pyl_fft_r2c(pyl_ctx *ctx, const float *delta, float *delta_k, int dims, int flags);

/* ---- power spectra (Pk_library.pyx:263-420, :529-793) -------------------
 * Output lengths for a dims^3 grid (F fields, S = F + F(F-1)/2 spectra):
 *   n1d = dims/2            (1D bins, DC dropped)
 *   n2d = (dims/2+1)*(int(sqrt(2)*(dims/2))+1)   (2D bins, all kept)
 *   n3d = int(sqrt(3)*(dims/2))                    (3D bins, DC dropped)
 * Result arrays (float64, caller-allocated, any may be NULL to skip):
 *   k1D[n1d] Nmodes1D[n1d] Pk1D[n1d*S]            Pk1D[i*S+s]
 *   kpar[n2d] kper[n2d] Nmodes2D[n2d] Pk2D[n2d*S] Pk2D[i*S+s]
 *   k3D[n3d] Nmodes3D[n3d] Pk3D[n3d*3*S]          Pk3D[(i*3+l)*S+s], l = 0,2,4 multipoles
 *   Pkphase[n3d]                                   (F==1 only)
 * Spectrum s < F is the auto-spectrum of field s; s >= F are the cross spectra
 * of pairs i<j in lexicographic order (XPk's index_X).
 */
typedef struct {
    int64_t n1d, n2d, n3d;
    int nspectra;
} pyl_pk_sizes;
int pyl_pk_get_sizes(int dims, int nfields, pyl_pk_sizes *out);

typedef struct {
    double *k1D, *Nmodes1D, *Pk1D;
    double *kpar, *kper, *Nmodes2D, *Pk2D;
    double *k3D, *Nmodes3D, *Pk3D, *Pkphase;
} pyl_pk_out;

/* deltas: nfields pointers to (dims,dims,dims) float32 fields.  mas: nfields MAS
 * ids (PYL_MAS_NONE = no deconvolution).  Optional prologue, applied to every
 * field before the FFT exactly as "delta /= scale; delta -= shift" in float32:
 * pass scale=1, shift=0 for none.  axis: line of sight, 0..2. */
int pyl_pk(pyl_ctx *ctx, const float *const *deltas, int nfields, const int *mas,
           int dims, double BoxSize, int axis, float scale, float shift, int flags,
           pyl_pk_out *out);

/* Batched Pk/XPk: nreal realisations of nfields fields.  deltas[f] points to nreal
 * contiguous (dims,dims,dims) float32 cubes of field f (field-major, so each field can
 * live in its own array).  One batched rocFFT and one binning launch for the whole
 * batch.  scales/shifts: nreal*nfields values, [r*nfields + f], for the prologue
 * (NULL = none); host arrays, or device arrays with PYL_PK_SCALES_DEVICE (then both
 * are given and read on the device, so they may come from device work queued just
 * before the call).  outs: nreal result structs (each sized as for pyl_pk with nfields
 * fields).  Each realisation's result is bitwise the same as a single pyl_pk call on
 * it.  Replaces a Python loop of Pk/XPk calls over realisations
 * (Pk_library.pyx:263-420, :529-793). */
int pyl_pk_batch(pyl_ctx *ctx, const float *const *deltas, int nreal, int nfields,
                 const int *mas, int dims, double BoxSize, int axis, const float *scales,
                 const float *shifts, int flags, pyl_pk_out *outs);

/* Same binning pass on fields that are already in Fourier space:
 * delta_k[f] is (dims,dims,dims/2+1) complex64 (interleaved). */
int pyl_pk_from_fourier(pyl_ctx *ctx, const float *const *delta_k, int nfields,
                        const int *mas, int dims, double BoxSize, int axis, int flags,
                        pyl_pk_out *out);

/* Host-only: exact mode counts of the reference loop for (dims, axis), with the DC
 * bins kept: N1D[dims/2+1], N2D[n2d], N3D[n3d+1], and the k sums k1Dsum[dims/2+1]
 * (sum of |k_par|) and k3Dsum[n3d+1] (sum of |k|), in units of kF. */
int pyl_mode_counts(int dims, int axis, double *N1D, double *N2D, double *N3D,
                    double *k1Dsum, double *k3Dsum);

/* Host-only: MAS_correction(x, p) = 1 if x == 0 else (x/sin x)^p. */
double pyl_mas_correction(double x, int p);

/* ---- redshift-space shift (redshift_space_library.pyx:29-46) -----------
 * pos/vel (n,3) float32; pos[:,axis] += vel[:,axis]*(1+z)/H, then periodic wrap. */
int pyl_pos_redshift_space(pyl_ctx *ctx, float *pos, const float *vel, int64_t n,
                           float BoxSize, float Hubble, float redshift, int axis, int flags);

/* ---- 2D spectra of (grid,grid) maps ------------------------------------
 * Pk_plane (nfields = 1) and XPk_plane (nfields = 2): 2D r2c FFT, MAS
 * deconvolution (float factor, complex64 scaling), binning by int(|k|), then the
 * reference normalisation with the DC bin dropped.  Outputs have nbins =
 * pyl_pk_plane_nbins(grid) entries: k[nbins], Nmodes[nbins] and
 * Pk[nbins*S] with S = 1, or S = 3 for two fields (auto 1, auto 2, cross),
 * [i*S + s].  Deterministic: every bin is summed in a fixed order. */
int pyl_pk_plane_nbins(int grid);
int pyl_pk_plane(pyl_ctx *ctx, const float *const *deltas, int nfields, const int *mas,
                 int grid, double BoxSize, int flags, double *k, double *Nmodes, double *Pk);

/* ---- MAS deconvolution of a field (correct_MAS) ------------------------
 * out = unnormalised inverse r2c transform (FFTW_BACKWARD convention, so the
 * result is dims^3 times the deconvolved field) of the forward transform of delta
 * with its independent modes divided by the MAS window.  The redundant half of
 * the kz = 0 and kz = dims/2 planes is left unscaled, as the reference does, and
 * enters the real inverse transform through its Hermitian part.  delta and out
 * are (dims,dims,dims) float32 and may alias. */
int pyl_correct_mas(pyl_ctx *ctx, const float *delta, float *out, int dims, double BoxSize,
                    int mas, int flags);

/* ---- HIP graphs of a context's work --------------------------------------
 * pyl_graph_begin starts capturing the context's stream; the library calls that
 * follow (device pointers: pyl_ma_batch_dev, pyl_ma with PYL_PTR_DEVICE, pyl_pk,
 * pyl_pk_batch with PYL_PK_SCALES_DEVICE, ...) only enqueue their kernels and
 * copies, and pyl_graph_end turns them into a graph.  Host steps that need the
 * device's results (the spectra's normalisation into the caller's result arrays,
 * the sort's particle-count check) run after every replay, in pyl_graph_launch,
 * which returns when the results are in the arrays given at capture.  Run the same
 * work once on the context before capturing it: scratch, plans and tables may not
 * be created or grow while capturing (such a call fails).  A context that captured
 * a graph serves that graph: other calls on it may grow its scratch and invalidate
 * the graph.  Inputs are read, and outputs written, at the addresses given at capture.
 * The reference has no counterpart: replaces a Python loop of MA and Pk/XPk calls
 * over realisations (MAS_library.pyx:57-112, Pk_library.pyx:263-420, :529-793). */
typedef struct pyl_graph pyl_graph;
int pyl_graph_begin(pyl_ctx *ctx);
int pyl_graph_end(pyl_ctx *ctx, pyl_graph **out);
int pyl_graph_abort(pyl_ctx *ctx);          /* ends a failed capture, discarding it */
int pyl_graph_launch(pyl_graph *graph);     /* replay, wait, run the host steps */
int pyl_graph_destroy(pyl_graph *graph);

/* Launch parameters for a tuner (value 0 / -1: the built-in choice). */
#define PYL_TUNE_BIN_ROWS      1   /* binning rows per work item, 1..256 */
#define PYL_TUNE_DEPOSIT_ORDER 2   /* MA deposit in largest-tile-first order: 1 on, 0 off */
#define PYL_TUNE_SORT_WINDOW   3   /* tiled MA sort: 0 the one-pass sort with the grid's
                                      default window; 1 the wide window (every tile in LDS;
                                      3D grids of 16385..32768 tiles); 2 the two-level sort
                                      (buckets of 8x8x4 tiles, then tiles; one 3D grid of
                                      16 to 96 tiles of 16 nodes per axis, 256^3 to 1536^3);
                                      3 the capacity sort (one read of the particles into
                                      per-tile slots of 3x the mean occupancy, the rest
                                      sorted after; one 3D grid whose sort window is every
                                      tile or the 8x8-column box); 4 the two-level sort with
                                      the capacity form as its second pass; -1 the adaptive
                                      choice: per (dims, MAS), when the previous call found
                                      more than a quarter of its particles outside their
                                      sort chunk's 8x8-column box, the two-level sort, else
                                      the one-pass sort; either with the capacity form when
                                      the tiles hold 256 particles or more on average, except
                                      for the 16 such calls that follow a capacity call of the
                                      context on that grid which spilled more than 1/64 of its
                                      particles.  A form that does not apply falls back to the
                                      default. */
int pyl_ctx_set_tuning(pyl_ctx *ctx, int knob, int value);
/* Sort window of the context's latest tiled MA (-1 none yet). */
#define PYL_WINDOW_DIRECT 0        /* every tile of the realisation counted in LDS */
#define PYL_WINDOW_BOX    1        /* 8 x 8 tile columns around the chunk */
#define PYL_WINDOW_LINEAR 2        /* 16384 consecutive keys around the chunk */
#define PYL_WINDOW_WIDE   3        /* every tile (up to 32768) in 128 KB of LDS */
#define PYL_WINDOW_TWOLEVEL 4      /* buckets of 8x8x4 tiles, then the tiles of each bucket */
#define PYL_WINDOW_CAPACITY 5      /* per-tile slots filled in one pass, then the spill sorted */
#define PYL_WINDOW_TWOLEVEL_CAPACITY 6  /* buckets, then the capacity form within each bucket */
int pyl_ctx_last_sort_window(pyl_ctx *ctx, int *window);

#ifdef __cplusplus
}
#endif
#endif /* PYLIANS_HIP_H */
