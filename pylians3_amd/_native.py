"""ctypes binding of libpylians_hip.so (include/pylians_hip.h).

The library is built in-tree by ``python -m pylians3_amd.build``.  There is no
CPU fallback: if the library is missing or no GPU is visible, every compute
call raises.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libpylians_hip.so')

PYL_OK, PYL_EINVAL, PYL_EHIP, PYL_ENOMEM, PYL_EFFT, PYL_EMODES, PYL_EBUDGET = 0, -1, -2, -3, -4, -5, -6
NGP, CIC, TSC, PCS = 1, 2, 3, 4
PTR_DEVICE = 0x0001
MA_ATOMIC = 0x0002
MA_CYTHON_2D = 0x0004
MA_RENORM_2D = 0x0008
MA_OVERWRITE = 0x0010
NO_SYNC = 0x0020
PK_DESTROY_INPUT = 0x0040
PK_NO_PHASE = 0x0080
PK_SCALES_DEVICE = 0x0100
TUNE_BIN_ROWS, TUNE_DEPOSIT_ORDER, TUNE_SORT_WINDOW = 1, 2, 3
WINDOW_DIRECT, WINDOW_BOX, WINDOW_LINEAR, WINDOW_WIDE, WINDOW_TWOLEVEL = 0, 1, 2, 3, 4
WINDOW_CAPACITY, WINDOW_TWOLEVEL_CAPACITY = 5, 6
# values of TUNE_SORT_WINDOW (-1 adaptive)
SORT_ONEPASS, SORT_WIDE, SORT_TWOLEVEL, SORT_CAPACITY, SORT_TWOLEVEL_CAPACITY = 0, 1, 2, 3, 4

STAGES = ['ma_count', 'ma_scan', 'ma_scatter', 'ma_deposit', 'ma_atomic', 'pk_prep', 'pk_fft',
          'pk_bin', 'ma_final', 'ma_bucket', 'ma_spill']

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class PkOut(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ('k1D', 'Nmodes1D', 'Pk1D', 'kpar', 'kper', 'Nmodes2D', 'Pk2D',
                                   'k3D', 'Nmodes3D', 'Pk3D', 'Pkphase')]


class PkSizes(ctypes.Structure):
    _fields_ = [('n1d', ctypes.c_int64), ('n2d', ctypes.c_int64), ('n3d', ctypes.c_int64),
                ('nspectra', ctypes.c_int)]


class NativeError(RuntimeError):
    """A libpylians_hip call failed; .status holds the PYL_E* code."""

    def __init__(self, status, msg):
        super().__init__('%s (status %d)' % (msg, status))
        self.status = status


class BudgetError(NativeError):
    pass


_LIB = None
_LOCK = threading.Lock()
EXPORTS = {
    'pyl_last_error': (ctypes.c_char_p, []),
    'pyl_version': (ctypes.c_int, []),
    'pyl_ctx_create': (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_size_t]),
    'pyl_ctx_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_ctx_set_stream': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    'pyl_ctx_synchronize': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_ctx_trim': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_ctx_memory': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    'pyl_ctx_enable_timing': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    'pyl_ctx_enable_stage_timing': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint]),
    'pyl_ctx_stage_times': (ctypes.c_int, [ctypes.c_void_p, _dp]),
    'pyl_ctx_stage_totals': (ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    'pyl_ma': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                              ctypes.c_int, ctypes.c_int]),
    'pyl_ma_noalloc': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_int, ctypes.c_int]),
    'pyl_ma_delta': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int]),
    'pyl_ma_rsd': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                  ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_int]),
    'pyl_ma_batch': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_int, ctypes.c_int]),
    'pyl_ma_batch_dev': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_int]),
    'pyl_graph_begin': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_graph_end': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    'pyl_graph_abort': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_graph_launch': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_graph_destroy': (ctypes.c_int, [ctypes.c_void_p]),
    'pyl_ctx_set_tuning': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    'pyl_ctx_last_sort_window': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    'NGP': (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                   ctypes.c_int, ctypes.c_float, ctypes.c_int]),
    'CIC': (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                   ctypes.c_int, ctypes.c_float, ctypes.c_int]),
    'TSC': (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                   ctypes.c_int, ctypes.c_float, ctypes.c_int]),
    'PCS': (None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_int,
                   ctypes.c_int, ctypes.c_float, ctypes.c_int]),
    'pyl_fft_r2c': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_int]),
    'pyl_cic_interp': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]),
    'pyl_pk_get_sizes': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(PkSizes)]),
    'pyl_pk': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                              ctypes.POINTER(PkOut)]),
    'pyl_pk_batch': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.POINTER(PkOut)]),
    'pyl_pk_from_fourier': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(PkOut)]),
    'pyl_mode_counts': (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]),
    'pyl_mas_correction': (ctypes.c_double, [ctypes.c_double, ctypes.c_int]),
    'pyl_pos_redshift_space': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int64, ctypes.c_float, ctypes.c_float,
                                              ctypes.c_float, ctypes.c_int, ctypes.c_int]),
    'pyl_pk_plane_nbins': (ctypes.c_int, [ctypes.c_int]),
    'pyl_pk_plane': (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _dp, _dp, _dp]),
    'pyl_correct_mas': (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int]),
}


def _share_hip_runtime():
    """PyTorch wheels bundle their own HIP runtime with the same soname
    (libamdhip64.so.7) but load it under another name; if this library were
    loaded first, a later ``import torch`` would bring a second HIP runtime
    into the process and lose the GPU.  Importing torch first makes the
    library bind to the runtime already in the process."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libpylians_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        with _LOCK:
            if _LIB is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError('libpylians_hip.so is not built (%s); run '
                                      '"python -m pylians3_amd.build"' % LIB_PATH)
                _share_hip_runtime()
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in EXPORTS.items():
                    f = getattr(L, name)
                    f.restype = res
                    f.argtypes = args
                _LIB = L
    return _LIB


def last_error():
    return lib().pyl_last_error().decode(errors='replace')


def check(status):
    if status != PYL_OK:
        msg = last_error()
        if status == PYL_EBUDGET:
            raise BudgetError(status, msg)
        raise NativeError(status, msg)


class Context:
    """A pyl_ctx: one HIP stream, scratch memory within an HBM budget, resident plans."""

    def __init__(self, device=0, hbm_budget=0):
        h = ctypes.c_void_p()
        check(lib().pyl_ctx_create(ctypes.byref(h), int(device), int(hbm_budget)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().pyl_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        check(lib().pyl_ctx_set_stream(self.handle, ctypes.c_void_p(stream_handle or 0)))

    def synchronize(self):
        check(lib().pyl_ctx_synchronize(self.handle))

    def trim(self):
        check(lib().pyl_ctx_trim(self.handle))

    def memory(self):
        v = ctypes.c_size_t()
        check(lib().pyl_ctx_memory(self.handle, ctypes.byref(v)))
        return v.value

    def enable_timing(self, on=True, stages=None):
        """Time every stage (stages=None) or only the named ones; clears the totals."""
        if stages is None:
            check(lib().pyl_ctx_enable_timing(self.handle, 1 if on else 0))
        else:
            mask = 0
            for s in stages:
                mask |= 1 << STAGES.index(s)
            check(lib().pyl_ctx_enable_stage_timing(self.handle, mask if on else 0))

    def stage_times(self):
        """Stage times (ms) of the latest call."""
        buf = (ctypes.c_double * len(STAGES))()
        check(lib().pyl_ctx_stage_times(self.handle, buf))
        return dict(zip(STAGES, list(buf)))

    def stage_totals(self):
        """{stage: (summed ms, calls)} since timing was enabled (no per-call sync)."""
        ms = (ctypes.c_double * len(STAGES))()
        cnt = (ctypes.c_longlong * len(STAGES))()
        check(lib().pyl_ctx_stage_totals(self.handle, ms, cnt))
        return {s: (ms[i], cnt[i]) for i, s in enumerate(STAGES)}


_tls = threading.local()


def default_context():
    """Per-thread default context on device 0 (the drop-in API uses it)."""
    ctx = getattr(_tls, 'ctx', None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx


def pk_sizes(dims, nfields):
    s = PkSizes()
    check(lib().pyl_pk_get_sizes(int(dims), int(nfields), ctypes.byref(s)))
    return s


def alloc_pk_out(dims, nfields):
    """Allocate float64 result arrays and the PkOut struct pointing at them."""
    s = pk_sizes(dims, nfields)
    S = s.nspectra
    arrs = dict(k1D=np.zeros(s.n1d), Nmodes1D=np.zeros(s.n1d), Pk1D=np.zeros((s.n1d, S)),
                kpar=np.zeros(s.n2d), kper=np.zeros(s.n2d), Nmodes2D=np.zeros(s.n2d),
                Pk2D=np.zeros((s.n2d, S)), k3D=np.zeros(s.n3d), Nmodes3D=np.zeros(s.n3d),
                Pk3D=np.zeros((s.n3d, 3, S)), Pkphase=np.zeros(s.n3d))
    out = PkOut(**{k: v.ctypes.data_as(_dp) for k, v in arrs.items()})
    out._keep = arrs          # the struct holds raw pointers into these arrays
    return out, arrs, S


def mode_counts(dims, axis):
    m = dims // 2
    s = pk_sizes(dims, 1)
    N1 = np.zeros(m + 1)
    N2 = np.zeros(s.n2d)
    N3 = np.zeros(s.n3d + 1)
    k1 = np.zeros(m + 1)
    k3 = np.zeros(s.n3d + 1)
    check(lib().pyl_mode_counts(int(dims), int(axis), N1.ctypes.data_as(_dp), N2.ctypes.data_as(_dp),
                                N3.ctypes.data_as(_dp), k1.ctypes.data_as(_dp), k3.ctypes.data_as(_dp)))
    return N1, N2, N3, k1, k3


def mas_correction(x, p):
    return lib().pyl_mas_correction(float(x), int(p))


# ---------------------------------------------------------------- array plumbing
def is_device_array(a):
    """True for a torch tensor on a GPU device."""
    return type(a).__module__.startswith('torch') and getattr(a, 'is_cuda', False)


def device_ptr(t):
    if not t.is_contiguous():
        raise ValueError('device arrays must be contiguous')
    return t.data_ptr()


class torch_stream:
    """Order library work after torch's queued work for the duration of a call.

    A side stream of torch is handed to the context, so the call is queued on it.
    Torch's default stream has the handle 0, which the context reads as "own
    stream" (a non-blocking stream, not ordered with it): then torch's queued work
    is waited for on entry, and the context's on exit, so that inputs written just
    before the call are complete and outputs are complete after it."""

    def __init__(self, ctx, active):
        self.ctx, self.active = ctx, active
        self.legacy = False

    def __enter__(self):
        if self.active:
            import torch
            s = torch.cuda.current_stream()
            self.legacy = not s.cuda_stream
            if self.legacy:
                s.synchronize()
            self.ctx.set_stream(s.cuda_stream)
        return self

    def __exit__(self, *exc):
        if self.active:
            self.ctx.set_stream(None)
            if self.legacy:
                if exc[0] is None:
                    self.ctx.synchronize()
                else:
                    # Kernels already queued may still write torch-owned outputs:
                    # finish them before the caller sees the error.  The error
                    # being raised is the one that matters.
                    try:
                        self.ctx.synchronize()
                    except Exception:
                        pass
        return False
