#!/usr/bin/env python3
"""Benchmark of the density-field -> power-spectrum hot path on one MI355X.

Workload (clustered field, 512^3): a Gaussian random field with
P(k) = A k^-1.5 exp(-(k/0.5)^2) (sigma_8 = 0.8) in a 1000 Mpc/h box, a 512^3
particle lattice moved by its Zel'dovich displacement plus a redshift-space
shift along axis 0 (f = 0.5); 134,217,728 float32 positions resident in HBM.

One step = the reference caller's chain (Pk_snapshot.py:77-89):
    grid = MA(pos, TSC)            (grid written, not accumulated: it starts at zero)
    delta = grid / mean - 1        (fused into the FFT input pass)
    Pk(delta, axis=0, MAS='TSC')   (own r2c FFT passes + fused binning + finalisation)
value = particles / step time (Mparticles/s).  The plain GPU form (per-particle
global atomics) and a CPU baseline are timed in the same run, and so are the same
particles in random order (each MA sort form), a 1024^3 precision study, batched
realisations and a serving stream.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-plain]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--grid', type=int, default=512, help='grid and particle-lattice size per axis (default 512: 512^3 particles onto 512^3)')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-plain', action='store_true', help='skip the plain GPU (atomic) leg')
    ap.add_argument('--cpu-seconds', type=float, default=20.0, help='target CPU sample time')
    ap.add_argument('--no-batched', action='store_true', help='skip the batched-realisations leg')
    ap.add_argument('--no-serving', action='store_true', help='skip the serving leg')
    ap.add_argument('--no-random', action='store_true', help='skip the random-order sort leg')
    ap.add_argument('--no-precision', action='store_true',
                    help='skip the 1024^3 PCS precision study (tiled vs atomic vs FP64)')
    ap.add_argument('--nreal', type=int, default=256, help='realisations in the batched leg')
    ap.add_argument('--calib-copy', action='store_true',
                    help='one untimed 512 MiB device-to-device copy before the steps: the '
                         'known-size kernel that scales PMC FETCH_SIZE/WRITE_SIZE to bytes')
    return ap.parse_args()


class Pipeline:
    """Device-resident MA + Pk step through the C ABI."""

    def __init__(self, N, ctx, pos, dims, box, mas, axis, atomic=False):
        import torch
        self.N, self.ctx = N, ctx
        self.pos, self.dims, self.box, self.mas, self.axis = pos, dims, box, mas, axis
        self.np = pos.shape[0]
        self.grid = torch.empty((dims,) * 3, dtype=torch.float32, device=pos.device)
        self.mean = float(ctypes.c_float(float(self.np) / dims ** 3).value)
        self.out, self.arrs, _ = N.alloc_pk_out(dims, 1)
        self.flags_ma = N.PTR_DEVICE | N.MA_OVERWRITE | N.NO_SYNC | (N.MA_ATOMIC if atomic else 0)
        self.gp = ctypes.c_void_p(self.grid.data_ptr())
        self.ptrs = (ctypes.c_void_p * 1)(self.grid.data_ptr())
        self.masid = (ctypes.c_int * 1)(mas)
        self.totals = {}

    def step(self):
        # MA fused with delta = rho/mean - 1 (Pk_snapshot.py:80-88), then Pk on delta
        N, L = self.N, self.N.lib()
        N.check(L.pyl_ma_delta(self.ctx.handle, ctypes.c_void_p(self.pos.data_ptr()), self.np, 3,
                               None, self.gp, self.dims, self.box, self.mas, self.mean, 1.0,
                               self.flags_ma))
        N.check(L.pyl_pk(self.ctx.handle, self.ptrs, 1, self.masid, self.dims, self.box, self.axis,
                         1.0, 0.0, N.PTR_DEVICE | N.PK_DESTROY_INPUT, ctypes.byref(self.out)))

    def reset(self, stages=None):
        # clears the stage totals; the stage events of the timed steps are read
        # back only after the last step (HIP events on the library's stream)
        self.ctx.enable_timing(True, stages)
        self.totals = {}

    def collect(self):
        self.totals = self.ctx.stage_totals()

    def stage_avg(self):
        """Average ms per launch of every stage that ran in the timed steps."""
        return {k: ms / n for k, (ms, n) in self.totals.items() if n > 0 and ms > 0}


def sort_form(N, ctx):
    """Name of the tile-sort form the context's latest tiled MA took."""
    w = ctypes.c_int()
    N.check(N.lib().pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(w)))
    return {N.WINDOW_DIRECT: 'one_pass_direct', N.WINDOW_BOX: 'one_pass_box', N.WINDOW_LINEAR: 'one_pass_linear',
            N.WINDOW_WIDE: 'wide', N.WINDOW_TWOLEVEL: 'two_level', N.WINDOW_CAPACITY: 'capacity',
            N.WINDOW_TWOLEVEL_CAPACITY: 'two_level_capacity'}.get(w.value, w.value)


def algorithmic_bytes(n, dims):
    """Algorithmic HBM bytes per launch of each timed kernel of the MA + Pk step
    (n particles, dims^3 grid, TSC, unweighted): what the kernel must move at
    minimum, not what it moves."""
    nk = dims * dims * (dims // 2 + 1)
    return {
        'ma_count': 12.0 * n,                    # positions read once
        'ma_scatter': 12.0 * n + 12.0 * n,       # positions read, one 12 B record written
        'ma_deposit': 12.0 * n + 4.0 * dims ** 3,  # records read, grid written once
        'pk_fft': 4.0 * dims ** 3 + 8.0 * nk,    # real grid read, half-complex written once
        'pk_bin': 8.0 * nk,                      # half-complex read once
    }


def committed_traffic(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC FETCH_SIZE and
    WRITE_SIZE passes on this workload, from the committed summary
    profiles/<kernel>_512_tsc_traffic.txt (line 'traffic_bytes: <int>'), else None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                        '%s_512_tsc_traffic.txt' % kernel)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('traffic_bytes:'):
                    return int(line.split(':', 1)[1])
    except OSError:
        return None
    return None


def time_steps(pipe, steps, warmup, stages=None):
    """ms per step; HIP events around `stages` (None: every stage) of every step."""
    import torch
    for _ in range(warmup):
        pipe.step()
    torch.cuda.synchronize()
    pipe.reset(stages)
    pipe.ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.step()
    pipe.ctx.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    pipe.collect()
    return ms


def host_cores():
    """Cores this process may use: its affinity set, capped by a cgroup CPU quota."""
    n = len(os.sched_getaffinity(0))
    try:
        with open('/sys/fs/cgroup/cpu.max') as f:
            q, per = f.read().split()[:2]
        if q != 'max':
            n = max(1, min(n, int(-(-int(q) // int(per)))))
    except (OSError, ValueError):
        pass
    return n


def _runs(fn, warm=3, runs=20):
    """Median and p99 seconds of `runs` calls after `warm` untimed ones."""
    import numpy as np
    for _ in range(warm):
        fn()
    t = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return float(np.median(t)), float(np.percentile(t, 99))


def _ma_sample(run, pos_host, dims, target_s):
    """Seconds for all particles, scaled from a sample grown to ~target_s."""
    import numpy as np
    n_total = pos_host.shape[0]
    ns = min(n_total, 1 << 20)
    g = np.zeros((dims,) * 3, np.float32)
    while True:
        t0 = time.perf_counter()
        run(pos_host[:ns], g)
        dt = time.perf_counter() - t0
        if dt > target_s or ns == n_total:
            break
        ns = min(n_total, ns * 4)
    return dt * n_total / ns, ns, dt


def cpu_baseline(pos_host, grid_host, dims, box, mas_name, axis, target_s):
    """The CPU path on this host's cores (count from the affinity set and cgroup quota):
      MA     the reference's MAS_c.c built with OpenMP (oracle/_ref, all cores), and the
             reference's own Cython MA() (serial loops, MAS_library.pyx:57-545, -O3
             -ffast-math as shipped), each on a sample scaled to the full count;
      FFT    scipy.fft.rfftn float32 with all cores (FFTW is not installed here);
      bins   the reference mode loop (Pk_library.pyx:311-378) restated in C, serial as
             the reference runs it and split over all cores.
    FFT and the threaded loop: 3 warm-ups, 20 timed runs, median and p99.  `value`
    is particles / (MA OpenMP + FFT + threaded bins); `as_shipped_step_s` is what
    the reference's MA() + Pk() path takes (serial MA, threaded FFT, serial bins)."""
    import numpy as np
    from oracle import oracle as O
    cores = host_cores()
    n_total = pos_host.shape[0]
    out = {'cores': cores}
    if O.ref_available('omp'):
        kind = 'reference'
        t_ma, ns, dt = _ma_sample(lambda p, g: O.ma_reference(p, g, box, mas_name, None, flavour='omp',
                                                              threads=cores), pos_host, dims, target_s / 4)
        ma_desc = 'reference MAS_c.c (OpenMP, %d threads) on %d of %d particles (%.2f s)' % (cores, ns, n_total, dt)
    else:
        kind = 'port'
        t_ma, ns, dt = _ma_sample(lambda p, g: O.ma_f32(p, g, box, mas_name), pos_host, dims, target_s / 4)
        ma_desc = 'oracle restatement (serial) on %d of %d particles (%.2f s)' % (ns, n_total, dt)
    if O.cython_available('shipped'):
        t_ser, ns2, dt2 = _ma_sample(lambda p, g: O.ma_cython(p, g, box, mas_name, flavour='shipped'),
                                     pos_host, dims, target_s / 4)
        ser_desc = 'reference Cython MA() (serial) on %d particles (%.2f s)' % (ns2, dt2)
    else:
        t_ser, ns2, dt2 = _ma_sample(lambda p, g: O.ma_f32(p, g, box, mas_name), pos_host, dims,
                                     target_s / 4)
        ser_desc = 'oracle restatement of MA() (serial) on %d particles (%.2f s)' % (ns2, dt2)
    delta = grid_host / np.float32(grid_host.mean()) - np.float32(1.0)
    fft_med, fft_p99 = _runs(lambda: O.fft_f32(delta, workers=cores))
    dk = O.fft_f32(delta, workers=cores)
    bin_med, bin_p99 = _runs(lambda: O.raw_bins([dk], dims, axis, [mas_name], threads=cores))
    t0 = time.perf_counter()
    O.raw_bins([dk], dims, axis, [mas_name])
    t_bin_serial = time.perf_counter() - t0
    total = t_ma + fft_med + bin_med
    out.update({
        'value': round(n_total / total / 1e6, 3),
        'unit': 'Mparticles/s',
        'kind': kind,
        'sample': ('MA: %s, scaled linearly; MA() serial: %s; FFT: scipy.fft.rfftn float32, %d '
                   'workers, full %d^3, median of 20; bins: reference mode loop restated in C over '
                   '%d threads, full grid, median of 20; serial bins once' % (
                       ma_desc, ser_desc, cores, dims, cores)),
        'step_s': round(total, 3),
        'grid_cells_per_s': round(dims ** 3 / total, 1),
        'ma_omp_s': round(t_ma, 4),
        'ma_serial_s': round(t_ser, 3),
        'fft_s': {'median': round(fft_med, 4), 'p99': round(fft_p99, 4)},
        'bins_threads_s': {'median': round(bin_med, 4), 'p99': round(bin_p99, 4)},
        'bins_serial_s': round(t_bin_serial, 3),
        'as_shipped_step_s': round(t_ser + fft_med + t_bin_serial, 3),
    })
    return out


def _errs(x, ref, floor=None):
    import numpy as np
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    fl = np.mean(np.abs(ref)) if floor is None else np.abs(np.asarray(floor, np.float64)).ravel()
    den = np.maximum(np.abs(ref), fl)
    num = np.abs(x - ref)
    # bins that are zero in both (Pk1D at k_par = middle) carry no error
    rel = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))
    return {'max': float('%.3g' % rel.max()), 'rms': float('%.3g' % np.sqrt(np.mean(rel * rel)))}


def cpu_check_errors(pos_host, grid_gpu, pk_gpu, dims, box, mas_name, axis, mean, threads):
    """Errors of the GPU results against the FP64 oracle (MA, FFT and bins in double,
    on all particles) and against the float32 oracle on the GPU's own delta.
    pk_gpu: dict of the GPU's finalized arrays (Pk3D, Pk1D, Pk2D).  The oracle is
    the checker here, run on the host cores."""
    import numpy as np
    from oracle import oracle as O
    g64 = O.ma_f64(pos_host, dims, box, mas_name, threads=threads)
    d64 = g64 / g64.mean() - 1.0
    r64 = O.pk(d64, box, axis=axis, MAS=mas_name, f64=True, workers=threads, threads=threads)
    del d64
    delta = grid_gpu / np.float32(mean) - np.float32(1.0)
    r32 = O.pk(delta, box, axis=axis, MAS=mas_name, workers=threads, threads=threads)
    del delta
    P = pk_gpu['Pk3D'][:, :, 0]
    P1, P2 = pk_gpu['Pk1D'][:, 0], pk_gpu['Pk2D'][:, 0]

    def block(ref, grid_ref):
        b = {} if grid_ref is None else {'grid': _errs(grid_gpu, grid_ref)}
        P0 = ref.Pk[:, 0]
        b.update({'Pk0': _errs(P[:, 0], P0, P0), 'Pk2': _errs(P[:, 1], ref.Pk[:, 1], P0),
                  'Pk4': _errs(P[:, 2], ref.Pk[:, 2], P0), 'Pk1D': _errs(P1, ref.Pk1D, ref.Pk1D),
                  'Pk2D': _errs(P2, ref.Pk2D, np.median(np.abs(ref.Pk2D)))})
        return b

    g32 = np.zeros((dims,) * 3, np.float32)
    O.ma_f32(pos_host, g32, box, mas_name)       # the reference's serial float32 arithmetic
    out = {'vs_fp64': block(r64, g64), 'vs_f32_oracle_same_delta': block(r32, None),
           'f32_oracle_vs_fp64': {'grid': _errs(g32, g64),
                                  'Pk0': _errs(r32.Pk[:, 0], r64.Pk[:, 0], r64.Pk[:, 0])},
           'gpu_vs_f32_oracle_grid': _errs(grid_gpu, g32),
           'floors': 'grid: mean cell; Pk2, Pk4: |Pk0| of the bin; Pk2D: median |Pk2D|'}
    return out


def random_order_leg(N, pos, dims, box, mas, reps=5):
    """The timed step's particles in random order (a seeded permutation): one MA per
    sort form, forced on a context of its own -- the one-pass sort with the grid's
    default window, the wide window, the two-level sort, the two-level sort with the
    capacity form as its second pass -- then the adaptive choice (the first call takes
    one pass and tallies its box misses; the next calls take the two-level sort).  Median of `reps` timed MAs after one warm-up (HIP
    events around each stage), and whether every form gave the same bits."""
    import numpy as np
    import torch
    L = N.lib()
    g = torch.Generator(device='cuda').manual_seed(7)
    perm = torch.randperm(pos.shape[0], generator=g, device='cuda')
    sp = pos[perm].contiguous()
    del perm
    # the library's stream does not wait for torch's: the copy must be complete first
    torch.cuda.synchronize()
    n = sp.shape[0]
    grid = torch.empty((dims,) * 3, device='cuda')
    ref = None
    out = {'workload': 'the timed step\'s %d particles in random order, %s onto %d^3'
                       % (n, {N.NGP: 'NGP', N.CIC: 'CIC', N.TSC: 'TSC', N.PCS: 'PCS'}[mas], dims),
           'forms': {}}
    same = True
    for name, force in (('one_pass', N.SORT_ONEPASS), ('wide_window', N.SORT_WIDE),
                        ('two_level', N.SORT_TWOLEVEL),
                        ('two_level_capacity', N.SORT_TWOLEVEL_CAPACITY), ('adaptive', -1)):
        ctx = N.Context(0)
        N.check(L.pyl_ctx_set_tuning(ctx.handle, N.TUNE_SORT_WINDOW, force))

        def one():
            N.check(L.pyl_ma(ctx.handle, ctypes.c_void_p(sp.data_ptr()), n, 3, None,
                             ctypes.c_void_p(grid.data_ptr()), dims, box, mas,
                             N.PTR_DEVICE | N.MA_OVERWRITE))
        one()
        torch.cuda.synchronize()
        first_form = ctypes.c_int()
        N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(first_form)))
        ctx.enable_timing(True)
        t = []
        for _ in range(reps):
            t0 = time.perf_counter()
            one()
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        tot = ctx.stage_totals()
        form = ctypes.c_int()
        N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(form)))
        if ref is None:
            ref = grid.clone()
        else:
            same = same and bool(torch.equal(grid, ref))
        names = {N.WINDOW_DIRECT: 'direct', N.WINDOW_BOX: 'box', N.WINDOW_LINEAR: 'linear',
                 N.WINDOW_WIDE: 'wide', N.WINDOW_TWOLEVEL: 'two_level', N.WINDOW_CAPACITY: 'capacity',
                 N.WINDOW_TWOLEVEL_CAPACITY: 'two_level_capacity'}
        out['forms'][name] = {
            'ma_ms_median': round(1e3 * float(np.median(t)), 3),
            'form_taken': names.get(form.value, form.value),
            'first_call_form': names.get(first_form.value, first_form.value),
            'stages_ms': {k: round(v[0] / v[1], 3) for k, v in tot.items() if v[1]}}
        ctx.close()
    out['bitwise_equal_across_forms'] = same
    f = out['forms']
    out['two_level_speedup_vs_one_pass'] = round(f['one_pass']['ma_ms_median'] /
                                                 f['two_level']['ma_ms_median'], 2)
    out['two_level_speedup_vs_wide'] = round(f['wide_window']['ma_ms_median'] /
                                             f['two_level']['ma_ms_median'], 2)
    out['two_level_capacity_speedup_vs_two_level'] = round(f['two_level']['ma_ms_median'] /
                                                           f['two_level_capacity']['ma_ms_median'], 2)
    del sp, grid, ref
    torch.cuda.empty_cache()
    return out


def precision_leg(N, reps=10, steps=5):
    """Large single field (1024^3 Zel'dovich particles, box 2000 Mpc/h, seed 3, PCS onto
    1024^3, Pk axis 0): the tiled (sorted, exact integer sums) and the atomic float32
    scatters.  Times the full MA + delta + Pk step, checks `reps` repeated tiled MAs
    and 3 atomic ones for bitwise reproducibility, and keeps both grids and their
    spectra for the CPU leg's FP64 comparison."""
    import numpy as np
    import torch
    from pylians3_amd import workloads
    dims, box, mas = 1024, 2000.0, N.PCS
    pos = workloads.config_c3()
    torch.cuda.synchronize()            # generated on torch's stream; read on the library's
    n = pos.shape[0]
    ctx = N.Context(0)
    L = N.lib()
    out = {'workload': '1024^3 Zel\'dovich particles (seed 3, box 2000 Mpc/h), PCS onto 1024^3, '
                       'Pk axis 0', 'particles': n}
    pipe = Pipeline(N, ctx, pos, dims, box, mas, 0)
    ms = time_steps(pipe, steps, 2)
    st = pipe.stage_avg()
    out['ms_per_step'] = round(ms, 3)
    out['sort_form'] = sort_form(N, ctx)
    out['mparticles_per_s'] = round(n / ms / 1e3, 1)
    out['grid_cells_per_s'] = round(dims ** 3 / (ms * 1e-3), 1)
    out['kernels_ms'] = {k: round(v, 4) for k, v in st.items()}
    ctx.enable_timing(False)
    grid = pipe.grid
    keep = {}
    for name, flag, nrep in (('tiled', 0, reps), ('atomic', N.MA_ATOMIC, 3)):
        first = None
        differ = 0
        t = []
        for r in range(nrep):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            N.check(L.pyl_ma(ctx.handle, ctypes.c_void_p(pos.data_ptr()), n, 3, None,
                             ctypes.c_void_p(grid.data_ptr()), dims, box, mas,
                             N.PTR_DEVICE | N.MA_OVERWRITE | flag))
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
            if first is None:
                first = grid.clone()
            elif not torch.equal(grid, first):
                differ += 1
        arrs = N.alloc_pk_out(dims, 1)
        mean = float(np.float32(n / dims ** 3))
        N.check(L.pyl_pk(ctx.handle, (ctypes.c_void_p * 1)(first.data_ptr()), 1,
                         (ctypes.c_int * 1)(mas), dims, box, 0, mean, 1.0, N.PTR_DEVICE,
                         ctypes.byref(arrs[0])))
        keep[name] = (first.cpu().numpy(), arrs[1])
        del first
        out[name] = {'runs': nrep, 'runs_differing_from_first': differ,
                     'bitwise_reproducible': differ == 0,
                     'ma_ms_median': round(1e3 * float(np.median(t)), 3)}
    del pipe, grid
    ctx.close()
    pos_host = pos.cpu().numpy()
    del pos
    torch.cuda.empty_cache()
    return out, keep, pos_host


def cpu_precision_errors(pos_host, keep, threads):
    """FP64 oracle (MA, FFT, bins in double) of the 1024^3 field: per-cell and per-bin
    relative errors of the tiled and atomic GPU results."""
    import numpy as np
    from oracle import oracle as O
    dims, box = 1024, 2000.0
    g64 = O.ma_f64(pos_host, dims, box, 'PCS', threads=threads)
    d64 = g64 / g64.mean() - 1.0
    r64 = O.pk(d64, box, axis=0, MAS='PCS', f64=True, workers=threads, threads=threads)
    del d64
    res = {}
    P0 = r64.Pk[:, 0]
    for name, (g, a) in keep.items():
        P = a['Pk3D'][:, :, 0]
        rel0 = np.abs(P[:, 0] - P0) / np.abs(P0)
        res[name] = {'grid': _errs(g, g64), 'Pk0': _errs(P[:, 0], P0, P0),
                     'Pk2': _errs(P[:, 1], r64.Pk[:, 1], P0), 'Pk4': _errs(P[:, 2], r64.Pk[:, 2], P0),
                     'Pk1D': _errs(a['Pk1D'][:, 0], r64.Pk1D, r64.Pk1D),
                     'Pk0_rel_err_by_k': {'k_h_per_mpc': [round(float(r64.k3D[i]), 4) for i in (0, 9, 99, 499, 885)],
                                          'rel': [float('%.3g' % rel0[i]) for i in (0, 9, 99, 499, 885)]}}
    g_t, g_a = keep['tiled'][0], keep['atomic'][0]
    res['tiled_vs_atomic_grid'] = _errs(g_a, g_t)
    return res


def batched_leg(N, nreal, dims=128, box=500.0, seed0=1000):
    """Batched leg: nreal independent 128^3 realisations (box 500): a Zel'dovich matter
    field (128^3 particles) and Poisson(lambda) weighted tracers, lambda log-uniform
    in [1e5, 2e6], CIC MA of both, then Pk of both and their cross spectrum (XPk).
    Forms, all device-resident: batched (pyl_ma_batch x2 + pyl_pk_batch), the same
    work as one tuned HIP graph (batch.BatchGraph), and one call per realisation
    (pyl_ma x2 + pyl_pk) with the tiled and with the atomic scatter."""
    import numpy as np
    import torch
    from pylians3_amd import workloads
    wl = workloads.batched_realisations(nreal, dims, box, seed0)
    mpos, moffs, tpos, tw, toffs, counts = (wl['mpos'], wl['moffs'], wl['tpos'], wl['tw'],
                                            wl['toffs'], wl['counts'])
    del wl
    ctx = N.Context(0)
    grids = [torch.empty((nreal, dims, dims, dims), dtype=torch.float32, device='cuda')
             for _ in range(2)]
    allocs = [N.alloc_pk_out(dims, 2) for _ in range(nreal)]
    outs = (N.PkOut * nreal)(*[a[0] for a in allocs])
    mas = (ctypes.c_int * 2)(N.CIC, N.CIC)
    tot_w = np.array([tw[toffs[r]:toffs[r + 1]].double().sum().item() for r in range(nreal)])
    scales = np.stack([tot_w / dims ** 3, np.full(nreal, 1.0)], 1).astype(np.float32)
    shifts = np.ones((nreal, 2), np.float32)
    L = N.lib()
    flags_ma = N.PTR_DEVICE | N.MA_OVERWRITE | N.NO_SYNC

    def batched():
        N.check(L.pyl_ma_batch(ctx.handle, ctypes.c_void_p(tpos.data_ptr()), toffs.ctypes.data,
                               nreal, 3, ctypes.c_void_p(tw.data_ptr()),
                               ctypes.c_void_p(grids[0].data_ptr()), dims, box, N.CIC, flags_ma))
        N.check(L.pyl_ma_batch(ctx.handle, ctypes.c_void_p(mpos.data_ptr()), moffs.ctypes.data,
                               nreal, 3, None, ctypes.c_void_p(grids[1].data_ptr()), dims, box,
                               N.CIC, flags_ma))
        ptrs = (ctypes.c_void_p * 2)(grids[0].data_ptr(), grids[1].data_ptr())
        N.check(L.pyl_pk_batch(ctx.handle, ptrs, nreal, 2, mas, dims, box, 0,
                               scales.ctypes.data, shifts.ctypes.data,
                               N.PTR_DEVICE | N.PK_DESTROY_INPUT, outs))

    one_out, one_arrs, _ = N.alloc_pk_out(dims, 2)   # one_arrs keeps the result arrays alive

    def per_call(flag):
        def run():
            for r in range(nreal):
                g0 = grids[0][r]
                g1 = grids[1][r]
                N.check(L.pyl_ma(ctx.handle, ctypes.c_void_p(tpos[toffs[r]:].data_ptr()),
                                 int(counts[r]), 3, ctypes.c_void_p(tw[toffs[r]:].data_ptr()),
                                 ctypes.c_void_p(g0.data_ptr()), dims, box, N.CIC,
                                 N.PTR_DEVICE | N.MA_OVERWRITE | flag))
                N.check(L.pyl_ma(ctx.handle, ctypes.c_void_p(mpos[moffs[r]:].data_ptr()),
                                 dims ** 3, 3, None, ctypes.c_void_p(g1.data_ptr()), dims, box, N.CIC,
                                 N.PTR_DEVICE | N.MA_OVERWRITE | flag))
                ptrs = (ctypes.c_void_p * 2)(g0.data_ptr(), g1.data_ptr())
                N.check(L.pyl_pk(ctx.handle, ptrs, 2, mas, dims, box, 0, 1.0, 0.0, N.PTR_DEVICE,
                                 ctypes.byref(one_out)))
        return run

    res = {}
    for name, fn, reps in (('batched', batched, 3), ('per_call_tiled', per_call(0), 2),
                           ('per_call_atomic', per_call(N.MA_ATOMIC), 2)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t0) / reps
    ctx.close()
    del grids
    torch.cuda.empty_cache()
    # the same work as one captured HIP graph, launch parameters from the tuner
    from pylians3_amd import batch, tune
    fields = [dict(pos=tpos, offs=torch.from_numpy(toffs).cuda(), W=tw, mas='CIC'),
              dict(pos=mpos, offs=torch.from_numpy(moffs).cuda(), W=None, mas='CIC')]
    t0 = time.perf_counter()
    best = tune.tune_batch(fields, nreal, dims, box, 0, use_cache=False, reps=2,
                           candidates={'chunks_per_real': (4, 16, 64), 'bin_rows': (0, 32),
                                       'deposit_order': (-1,)})
    t_tune = time.perf_counter() - t0
    tuning = {k: best[k] for k in ('chunks_per_real', 'bin_rows', 'deposit_order')}
    g = batch.BatchGraph(fields, nreal, dims, box, 0, tuning=tuning)
    try:
        res['graph'] = tune.time_graph(g, reps=5)
    finally:
        g.close()
    npart = int(counts.sum()) + nreal * dims ** 3
    return {'workload': '%d x %d^3 realisations: ZA matter + Poisson lognormal-weighted tracers '
                        '(lambda log-uniform 1e5-2e6), CIC MA of both, Pk + XPk, axis 0'
                        % (nreal, dims),
            'particles': npart,
            'batched_s': round(res['batched'], 4),
            'graph_s': round(res['graph'], 4),
            'per_call_tiled_s': round(res['per_call_tiled'], 4),
            'per_call_atomic_s': round(res['per_call_atomic'], 4),
            'realisations_per_s': round(nreal / min(res['batched'], res['graph']), 2),
            'speedup_vs_per_call_tiled': round(res['per_call_tiled'] / min(res['batched'], res['graph']), 2),
            'speedup_vs_per_call_atomic': round(res['per_call_atomic'] / min(res['batched'], res['graph']), 2),
            'graph_tuning': tuning,
            'tuner_s': round(t_tune, 2),
            'tuner_timings_ms': [dict({k: t[k] for k in tuning}, ms=round(1e3 * t['seconds'], 2))
                                 for t in best.get('timings', [])],
            'forms': {'batched': 'pyl_ma_batch x2 + pyl_pk_batch (launch sequence per call)',
                      'graph': 'the batched work captured once as a HIP graph, replayed (host '
                               'normalisation after each replay), tuned launch parameters',
                      'per_call_tiled': 'pyl_ma x2 (tiled) + pyl_pk per realisation',
                      'per_call_atomic': 'pyl_ma x2 (per-particle atomics) + pyl_pk per realisation'}}


def serving_leg(N, nreq=120, seed=5):
    """Serving leg: mixed MA+Pk / MA+XPk requests (grid 64^3-512^3, N log-uniform
    1e5-1.3e8, MAS and axis drawn uniformly) against the continuous-batching server.
    Plain form: one request at a time, one slot, arrival order, no fusion; its
    service latency alone, then under Poisson arrivals at 0.9x its capacity.
    Served form (4 slots, admission in arrival order, fusion of small requests
    waiting together): the same arrival schedule, then a saturated burst.  Each
    server first runs the largest request once on every slot (Server.warm)."""
    import numpy as np
    import torch
    from pylians3_amd import workloads
    from pylians3_amd.serving import Request, Server
    rng = np.random.default_rng(seed)
    box = 1000.0
    psi = workloads.displacement_field(256, box, seed)
    nmax = 130_000_000
    pool, poolw = workloads.tracers(nmax, psi, box, seed + 1)
    del psi
    reqs = []
    for i in range(nreq):
        d = int(rng.choice([64, 128, 256, 512]))
        n = int(np.exp(rng.uniform(np.log(1e5), np.log(1.3e8))))
        o = int(rng.integers(0, nmax - n + 1))
        mas = str(rng.choice(['NGP', 'CIC', 'TSC', 'PCS']))
        axis = int(rng.integers(0, 3))
        if rng.random() < 0.5:
            reqs.append(Request(pool[o:o + n], d, box, mas, axis))
        else:
            n2 = max(100000, n // 4)
            o2 = int(rng.integers(0, nmax - n2 + 1))
            reqs.append(Request(pool[o:o + n], d, box, mas, axis, pos2=pool[o2:o2 + n2],
                                W2=poolw[o2:o2 + n2]))
    # every server is warmed with the largest request on each of its slots first
    biggest = max(reqs, key=lambda r: r.hbm_bytes())
    # plain forms: one request after the other, one slot, no fusion; the tiled
    # scatter, and the per-particle atomic scatter (the plain GPU form)
    with Server(slots=1, max_batch=1, policy='fifo') as srv:
        srv.warm(biggest)
        srv.submit(reqs[0]).result()
        t0 = time.perf_counter()
        plain_lat = np.array([srv.submit(r).result().latency_s for r in reqs])
        t_plain = time.perf_counter() - t0
    with Server(slots=1, max_batch=1, policy='fifo', atomic=True) as srv:
        srv.warm(biggest)
        srv.submit(reqs[0]).result()
        t0 = time.perf_counter()
        atomic_lat = np.array([srv.submit(r).result().latency_s for r in reqs])
        t_atomic = time.perf_counter() - t0
    cap = nreq / t_plain
    out = {'requests': nreq, 'plain_throughput_rps': round(cap, 2),
           'plain_p50_ms': round(1e3 * np.percentile(plain_lat, 50), 3),
           'plain_p99_ms': round(1e3 * np.percentile(plain_lat, 99), 3),
           'plain_atomic_throughput_rps': round(nreq / t_atomic, 2),
           'plain_atomic_p50_ms': round(1e3 * np.percentile(atomic_lat, 50), 3),
           'plain_atomic_p99_ms': round(1e3 * np.percentile(atomic_lat, 99), 3)}
    # Poisson arrivals at 0.9x the plain capacity, the same schedule for both forms
    gaps = rng.exponential(1.0 / (0.9 * cap), nreq)

    def loaded(srv):
        srv.submit(reqs[1]).result()
        srv.completed.clear()
        futs = []
        t0 = time.perf_counter()
        for r, gap in zip(reqs, gaps):
            time.sleep(gap)
            futs.append(srv.submit(r))
        return np.array([f.result().latency_s for f in futs]), time.perf_counter() - t0

    with Server(slots=1, max_batch=1, policy='fifo') as srv:
        srv.warm(biggest)
        lat_p, _ = loaded(srv)
    out.update({'plain_loaded_p50_ms': round(1e3 * np.percentile(lat_p, 50), 3),
                'plain_loaded_p99_ms': round(1e3 * np.percentile(lat_p, 99), 3)})
    with Server(slots=4, max_batch=32) as srv:
        srv.warm(biggest)
        lat, t_arr = loaded(srv)
        srv.completed.clear()
        t0 = time.perf_counter()
        futs = [srv.submit(r) for r in reqs]
        for f in futs:
            f.result()
        t_sat = time.perf_counter() - t0
    out.update({'arrival_rate_rps': round(0.9 * cap, 2),
                'p50_ms': round(1e3 * np.percentile(lat, 50), 3),
                'p99_ms': round(1e3 * np.percentile(lat, 99), 3),
                'served_window_s': round(t_arr, 3),
                'saturated_throughput_rps': round(nreq / t_sat, 2),
                'speedup_vs_plain': round((nreq / t_sat) / cap, 2),
                'speedup_vs_plain_atomic': round(t_atomic / t_sat, 2),
                'forms': {'plain': '1 slot, arrival order, no fusion, tiled scatter',
                          'plain_atomic': '1 slot, arrival order, no fusion, per-particle '
                                          'atomic scatter (the plain per-call GPU form)',
                          'served': '4 slots, admission in arrival order, fusion of '
                                    'small requests waiting together, tiled scatter'}})
    return out


def main():
    args = parse()
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print('bench.py: this build runs on one GPU only; no multi-GPU path exists '
              '(requested --gpus %d, WORLD_SIZE %d)' % (args.gpus, world), file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from pylians3_amd import _native as N, workloads

    dims, box, axis = args.grid, 1000.0, 0
    mas = N.TSC
    log('generating %d^3 Zel\'dovich + RSD particles on the GPU' % dims)
    pos = workloads.zeldovich(dims, box, 2, f_rsd=0.5, rsd_axis=0)
    torch.cuda.synchronize()
    n = pos.shape[0]
    ctx = N.Context(0)
    ctx.enable_timing(True)
    if args.calib_copy:
        src = torch.ones(1 << 27, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log('calibration copy: %d bytes read, %d bytes written' % (src.numel() * 4, dst.numel() * 4))
        del src, dst

    pipe = Pipeline(N, ctx, pos, dims, box, mas, axis)
    kernel_bytes = algorithmic_bytes(n, dims)
    # 1) per-kernel times: events around every stage (each pair adds a few us of
    #    stream serialisation, so these steps do not give `value`)
    ms_all = time_steps(pipe, max(3, min(args.steps, 10)), args.warmup)
    st = pipe.stage_avg()
    log('instrumented step %.3f ms  stages %s' % (ms_all, {k: round(v, 3) for k, v in st.items()}))
    dom = max((k for k in st if k in kernel_bytes), key=lambda k: st[k])
    # 2) timed steps: events only around the dominant kernel (the roofline)
    log('tiled MA + Pk: warmup %d, steps %d (events around %s)' % (args.warmup, args.steps, dom))
    ms = time_steps(pipe, args.steps, args.warmup, stages=[dom])
    dom_ms = pipe.stage_avg()[dom]
    log('step %.3f ms  %s %.3f ms' % (ms, dom, dom_ms))
    # the whole step (MA, FFT, binning) is the same bits on every run: hash every
    # output of the last timed step and of three more steps
    import hashlib

    def _outputs_hash():
        return hashlib.sha1(b''.join(np.ascontiguousarray(v).tobytes()
                                     for _, v in sorted(pipe.arrs.items()))).hexdigest()
    hashes = [_outputs_hash()]
    for _ in range(3):
        pipe.step()
        torch.cuda.synchronize()
        hashes.append(_outputs_hash())
    reproducible = {'steps_compared': len(hashes), 'bitwise_identical': len(set(hashes)) == 1,
                    'outputs_sha1': hashes[0][:16]}
    main_form = sort_form(N, ctx)
    grid_host = None
    if not args.no_cpu:
        # one more MA to keep the density grid for the CPU leg
        N.check(N.lib().pyl_ma(ctx.handle, ctypes.c_void_p(pos.data_ptr()), n, 3, None,
                               ctypes.c_void_p(pipe.grid.data_ptr()), dims, box, mas,
                               N.PTR_DEVICE | N.MA_OVERWRITE))
        grid_host = pipe.grid.cpu().numpy()

    plain = None
    if not args.no_plain:
        log('plain GPU form (atomic MA) + Pk')
        pp = Pipeline(N, ctx, pos, dims, box, mas, axis, atomic=True)
        ms_p = time_steps(pp, args.steps, max(1, args.warmup // 2))
        sp = pp.stage_avg()
        plain = {'ms_per_step': round(ms_p, 4), 'mparticles_per_s': round(n / ms_p / 1e3, 2),
                 'ma_ms': round(sp.get('ma_atomic', 0.0), 4)}
        del pp

    random_order = None
    if not args.no_random:
        log('random-order leg: the same particles shuffled, each sort form')
        random_order = random_order_leg(N, pos, dims, box, mas)

    ma_ms = sum(st.get(k, 0.0) for k in ('ma_bucket', 'ma_count', 'ma_scan', 'ma_scatter',
                                          'ma_spill', 'ma_deposit', 'ma_final'))
    pk_ms = sum(st.get(k, 0.0) for k in ('pk_prep', 'pk_fft', 'pk_bin'))
    nk = dims * dims * (dims // 2 + 1)
    B_ma = 12.0 * n + 4.0 * dims ** 3            # positions once + grid written once
    B_pk = 4.0 * dims ** 3 + 16.0 * nk           # real grid, half-complex written + read once
    stage_bytes = {'ma': (B_ma, ma_ms), 'pk': (B_pk, pk_ms)}
    stages = {k: {'ms': round(t, 4), 'algorithmic_GB': round(b / 1e9, 4),
                  'GBps': round(b / (t * 1e-3) / 1e9, 1) if t > 0 else None}
              for k, (b, t) in stage_bytes.items()}
    kernels = {k: round(v, 4) for k, v in st.items()}
    # dominant kernel: the longest stage; its duration is from the timed steps
    # (HIP events on the library's stream)
    ach = kernel_bytes[dom] / (dom_ms * 1e-3) / 1e9
    lds = None
    if dom == 'ma_deposit':
        # the deposit is LDS-atomic bound: 27 ds_add_u64 per TSC particle against the
        # measured ceiling of 64-bit LDS atomics with conflict-free lanes (6.2 clk per
        # wave instruction per CU: 64 lanes / 6.2 clk * 2.4 GHz * 256 CUs)
        atomics = 27.0 * n
        peak = 64.0 / 6.2 * 2.4e9 * 256
        lds = {'atomics_per_s': round(atomics / (dom_ms * 1e-3), -6), 'peak_per_s': round(peak, -9),
               'frac': round(atomics / (dom_ms * 1e-3) / peak, 4)}
    # The roofline is HBM bandwidth (no kernel here is a dense contraction).  What
    # limits the kernel below it: the deposit and the binning are bound by their LDS
    # atomics and VALU work, the sort and FFT passes by HBM.
    limiter = {'ma_deposit': 'lds_atomics+valu', 'pk_bin': 'lds_atomics+valu'}.get(dom, 'hbm')
    roofline = {'bound': 'hbm', 'limiter': limiter, 'kernel': dom, 'kernel_ms': round(dom_ms, 4), 'lds_atomics': lds,
                'achieved': round(ach, 1), 'peak': HBM_PEAK_GBS,
                'unit': 'GB/s', 'frac': round(ach / HBM_PEAK_GBS, 4),
                'algorithmic_bytes': int(kernel_bytes[dom]),
                'traffic': committed_traffic(dom) if dims == 512 else None}
    per_kernel_gbps = {k: round(kernel_bytes[k] / (t * 1e-3) / 1e9, 1)
                       for k, t in st.items() if k in kernel_bytes and t > 0}

    cpu = None
    errors = None
    if not args.no_cpu:
        log('CPU baseline leg')
        pos_host = pos.cpu().numpy()
        cpu = cpu_baseline(pos_host, grid_host, dims, box, 'TSC', axis, args.cpu_seconds)
        log('CPU leg: errors of the timed step against the FP64 oracle')
        errors = cpu_check_errors(pos_host, grid_host, pipe.arrs, dims, box, 'TSC', axis, pipe.mean,
                                  cpu['cores'])
        del pos_host
    del pipe
    ctx.close()
    del pos
    torch.cuda.empty_cache()

    precision = None
    if not args.no_precision and dims == 512:
        log('precision leg: 1024^3 PCS, tiled vs atomic')
        precision, keep, pos3 = precision_leg(N)
        if not args.no_cpu:
            log('CPU leg: 1024^3 FP64 oracle')
            precision['errors_vs_fp64'] = cpu_precision_errors(pos3, keep, cpu['cores'])
        del keep, pos3
        torch.cuda.empty_cache()

    batched = None
    if not args.no_batched:
        log('batched realisations leg (%d x 128^3)' % args.nreal)
        batched = batched_leg(N, args.nreal)
        torch.cuda.empty_cache()
    serving = None
    if not args.no_serving:
        log('serving leg')
        serving = serving_leg(N)
        torch.cuda.empty_cache()

    value = n / (ms * 1e-3) / 1e6
    line = {
        'metric': 'Mparticles/s (MA) + Pk ms at 512^3 TSC vs plain GPU; p50/p99 request latency',
        'value': round(value, 2),
        'unit': 'Mparticles/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 grid/FFT, f64 bin sums',
        'data': 'synthetic: Zel\'dovich + RSD (f=0.5, axis 0) lattice from a seeded Gaussian field',
        'config': {'workload': 'clustered field: %d^3 particles, TSC onto %d^3, Pk axis 0 '
                               '(MA + delta + Pk per step)' % (dims, dims),
                   'particles': n, 'grid': dims, 'box_mpc_h': box, 'mas': 'TSC', 'axis': axis},
        'ma_ms': round(ma_ms, 4),
        'sort_form': main_form,
        'pk_ms': round(pk_ms, 4),
        'kernel_GBps': per_kernel_gbps,
        'ma_mparticles_per_s': round(n / ma_ms / 1e3, 2) if ma_ms > 0 else None,
        'grid_cells_per_s': round(dims ** 3 / (ms * 1e-3), 1),
        'errors_vs_fp64': errors,
        'step_reproducible': reproducible,
        'precision_1024': precision,
        'kernels_ms': kernels,                 # from the instrumented steps (events on every stage)
        'instrumented_ms_per_step': round(ms_all, 4),
        'stages': stages,
        'plain_gpu': plain,
        'random_order': random_order,
        'batched': batched,
        'serving': serving,
        'roofline': roofline,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
