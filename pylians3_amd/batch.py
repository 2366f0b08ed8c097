"""Batched MA and Pk/XPk over many realisations (one launch sequence per batch).

    MA_batch(pos, offsets, grids, BoxSize, MAS='CIC', W=None, accumulate=True)
        particles of realisation r are pos[offsets[r]:offsets[r+1]]; grids is
        (nreal, d, d, d) float32; same result per grid as MAS_library.MA on it.
    Pk_batch(deltas, BoxSize, axis=2, MAS=('CIC',), scales=None, shifts=None)
        deltas is a list of nfields (nreal, d, d, d) float32 arrays (or one
        (nreal, nfields, d, d, d) array); returns one result per
        realisation with the attributes of Pk_library.Pk (nfields == 1) or
        Pk_library.XPk (nfields >= 2).  The optional prologue applies
        delta = delta/scale - shift per field in float32 before the FFT.

Arrays may be numpy (staged) or contiguous torch tensors on the GPU.  Each
realisation's result is bitwise identical to the corresponding single call:
the tiled scatter and the binning reduce in an order that does not depend on
the batch.
"""
import ctypes

import numpy as np

from . import _native as N
from .MAS_library import _MAS_IDS
from .Pk_library import MAS_function


class Result:
    """Attributes of Pk_library.Pk / XPk for one realisation."""


def MA_batch(pos, offsets, grids, BoxSize, MAS='CIC', W=None, accumulate=True, atomic=False,
             ctx=None):
    ctx = ctx or N.default_context()
    offs = np.ascontiguousarray(offsets, dtype=np.int64)
    nreal = len(offs) - 1
    dims = grids.shape[1]
    axes = pos.shape[1]
    if grids.shape[0] != nreal:
        raise ValueError('grids must have one grid per realisation')
    flags = 0 if accumulate else N.MA_OVERWRITE
    if atomic:
        flags |= N.MA_ATOMIC
    if N.is_device_array(pos):
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_ma_batch(ctx.handle, N.device_ptr(pos), offs.ctypes.data, nreal, axes,
                                         None if W is None else N.device_ptr(W),
                                         N.device_ptr(grids), dims, float(BoxSize), _MAS_IDS[MAS],
                                         flags | N.PTR_DEVICE))
        return grids
    p = np.ascontiguousarray(pos, dtype=np.float32)
    w = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    if not grids.flags['C_CONTIGUOUS'] or grids.dtype != np.float32:
        raise ValueError('grids must be a C-contiguous float32 array')
    N.check(N.lib().pyl_ma_batch(ctx.handle, p.ctypes.data, offs.ctypes.data, nreal, axes,
                                 None if w is None else w.ctypes.data, grids.ctypes.data, dims,
                                 float(BoxSize), _MAS_IDS[MAS], flags))
    return grids


def _fields(deltas):
    """List of F field arrays, each (nreal, d, d, d) contiguous float32.  Accepts such
    a list, or one (nreal, F, d, d, d) array (split into per-field copies)."""
    if isinstance(deltas, (list, tuple)):
        fields = list(deltas)
    else:
        fields = [deltas[:, f] for f in range(deltas.shape[1])]
    out = []
    for a in fields:
        if N.is_device_array(a):
            out.append(a.contiguous())
        else:
            out.append(np.ascontiguousarray(a, dtype=np.float32))
    return out


def Pk_batch(deltas, BoxSize, axis=2, MAS=('CIC',), scales=None, shifts=None, destroy=False,
             ctx=None):
    ctx = ctx or N.default_context()
    fields = _fields(deltas)
    F = len(fields)
    nreal, dims = fields[0].shape[0], fields[0].shape[1]
    if len(MAS) != F:
        raise ValueError('one MAS per field')
    for a in fields:
        if tuple(a.shape) != (nreal, dims, dims, dims):
            raise ValueError('every field must be (nreal, dims, dims, dims)')
    mas = (ctypes.c_int * F)(*[MAS_function(m) for m in MAS])
    allocs = [N.alloc_pk_out(dims, F) for _ in range(nreal)]
    outs = (N.PkOut * nreal)(*[a[0] for a in allocs])
    sc = None if scales is None else np.ascontiguousarray(scales, dtype=np.float32).reshape(-1)
    sh = None if shifts is None else np.ascontiguousarray(shifts, dtype=np.float32).reshape(-1)
    for a in (sc, sh):
        if a is not None and a.size != nreal * F:
            raise ValueError('scales/shifts need nreal*nfields values')
    dev = N.is_device_array(fields[0])
    ptrs = (ctypes.c_void_p * F)(*[(N.device_ptr(a) if dev else a.ctypes.data) for a in fields])
    flags = (N.PTR_DEVICE | (N.PK_DESTROY_INPUT if destroy else 0)) if dev else 0
    with N.torch_stream(ctx, dev):
        N.check(N.lib().pyl_pk_batch(ctx.handle, ptrs, nreal, F, mas, dims, float(BoxSize),
                                     int(axis), None if sc is None else sc.ctypes.data,
                                     None if sh is None else sh.ctypes.data, flags, outs))
    return _results(allocs, F)


class BatchGraph:
    """Batched MA of every field and Pk/XPk of nreal realisations, captured once as a
    HIP graph and replayed.

        g = BatchGraph(fields, nreal, dims, BoxSize, axis=0)
        res = g.run()          # list of nreal results (Pk / XPk attributes)

    fields: list of dicts, one per field, with device tensors bound at capture:
        pos   (n_cap, 3) float32 positions; realisation r holds pos[offs[r]:offs[r+1]]
        offs  (nreal+1,) int64 particle ranges (device), offs[0] = 0, offs[nreal] <= n_cap
        W     (n_cap,) float32 weights or None
        mas   'NGP' | 'CIC' | 'TSC' | 'PCS'
    Between replays the caller may rewrite pos, W and offs in place (new particles,
    other counts up to n_cap); the graph reads them where they are.  run() forms
    delta = grid/mean - 1 per field with mean = count/dims^3, or sum(W)/dims^3
    (float64 sums rounded to float32), unless scales are given.  The launch
    sequence is fixed at capture: it depends on n_cap and chunks_per_real, not on
    the counts (pyl_ma_batch_dev cuts the sort chunks on the device).

    tuning: {'chunks_per_real': int, 'bin_rows': int, 'deposit_order': -1|0|1}
    (see tune.py); the graph owns a private context, so nothing else can grow the
    scratch its kernels were captured with.
    """

    def __init__(self, fields, nreal, dims, BoxSize, axis=0, tuning=None, device=0):
        import torch
        self.fields = fields
        self.nreal, self.dims, self.box, self.axis = int(nreal), int(dims), float(BoxSize), int(axis)
        self.F = len(fields)
        t = dict(chunks_per_real=16, bin_rows=0, deposit_order=-1)
        t.update(tuning or {})
        self.tuning = t
        self.ctx = N.Context(device)
        L = N.lib()
        N.check(L.pyl_ctx_set_tuning(self.ctx.handle, N.TUNE_BIN_ROWS, int(t['bin_rows'])))
        N.check(L.pyl_ctx_set_tuning(self.ctx.handle, N.TUNE_DEPOSIT_ORDER, int(t['deposit_order'])))
        for f in fields:
            for k in ('pos', 'offs'):
                if not N.is_device_array(f[k]) or not f[k].is_contiguous():
                    raise ValueError('%s must be a contiguous GPU tensor' % k)
            if f['offs'].dtype != torch.int64 or f['offs'].numel() != self.nreal + 1:
                raise ValueError('offs must be an int64 tensor of nreal+1 entries')
            if f.get('W') is not None and (not N.is_device_array(f['W']) or f['W'].numel() < f['pos'].shape[0]):
                raise ValueError('W must be a GPU tensor with one weight per position')
        d = self.dims
        self.grids = [torch.empty((self.nreal, d, d, d), dtype=torch.float32, device='cuda')
                      for _ in range(self.F)]
        self.scales = torch.ones((self.nreal, self.F), dtype=torch.float32, device='cuda')
        self.shifts = torch.ones((self.nreal, self.F), dtype=torch.float32, device='cuda')
        self.allocs = [N.alloc_pk_out(d, self.F) for _ in range(self.nreal)]
        self.outs = (N.PkOut * self.nreal)(*[a[0] for a in self.allocs])
        self.mas = (ctypes.c_int * self.F)(*[MAS_function(f['mas']) for f in fields])
        self.ptrs = (ctypes.c_void_p * self.F)(*[g.data_ptr() for g in self.grids])
        torch.cuda.synchronize()
        self._sequence()                     # once uncaptured: sizes scratch, builds plans
        N.check(self.ctx_sync())
        N.check(L.pyl_graph_begin(self.ctx.handle))
        try:
            self._sequence()
        except Exception:
            L.pyl_graph_abort(self.ctx.handle)
            raise
        g = ctypes.c_void_p()
        N.check(L.pyl_graph_end(self.ctx.handle, ctypes.byref(g)))
        self.graph = g

    def ctx_sync(self):
        return N.lib().pyl_ctx_synchronize(self.ctx.handle)

    def _sequence(self):
        L = N.lib()
        cpr = int(self.tuning['chunks_per_real'])
        for f, grid in zip(self.fields, self.grids):
            W = f.get('W')
            N.check(L.pyl_ma_batch_dev(self.ctx.handle, N.device_ptr(f['pos']), N.device_ptr(f['offs']),
                                       self.nreal, int(f['pos'].shape[0]), cpr, 3,
                                       None if W is None else N.device_ptr(W), N.device_ptr(grid),
                                       self.dims, self.box, _MAS_IDS[f['mas']],
                                       N.PTR_DEVICE | N.MA_OVERWRITE | N.NO_SYNC))
        N.check(L.pyl_pk_batch(self.ctx.handle, self.ptrs, self.nreal, self.F, self.mas, self.dims,
                               self.box, self.axis, N.device_ptr(self.scales),
                               N.device_ptr(self.shifts),
                               N.PTR_DEVICE | N.PK_DESTROY_INPUT | N.PK_SCALES_DEVICE, self.outs))

    def means(self):
        """Mean density of every (realisation, field) from the current inputs, on
        the device: count/dims^3, or sum(W)/dims^3 in float64, rounded to float32
        (a zero mean becomes 1)."""
        import torch
        cells = float(self.dims) ** 3
        m = torch.empty((self.nreal, self.F), dtype=torch.float64, device='cuda')
        for i, f in enumerate(self.fields):
            offs = f['offs']
            W = f.get('W')
            if W is None:
                m[:, i] = (offs[1:] - offs[:-1]).double() / cells
            else:
                c = torch.zeros(W.numel() + 1, dtype=torch.float64, device='cuda')
                torch.cumsum(W.double(), 0, out=c[1:])
                m[:, i] = (c[offs[1:]] - c[offs[:-1]]) / cells
        m = m.float()
        return torch.where(m == 0, torch.ones_like(m), m)

    def run(self, scales=None):
        """Replay: MA of every field, Pk/XPk of every realisation.  Returns one result
        per realisation (arrays owned by the graph, overwritten by the next run)."""
        import torch
        self.scales.copy_(self.means() if scales is None else
                          torch.as_tensor(scales, dtype=torch.float32, device='cuda').view(self.nreal, self.F))
        torch.cuda.current_stream().synchronize()     # inputs and scales written
        N.check(N.lib().pyl_graph_launch(self.graph))
        return _results(self.allocs, self.F)

    def close(self):
        if getattr(self, 'graph', None):
            N.lib().pyl_graph_destroy(self.graph)
            self.graph = None
        if getattr(self, 'ctx', None):
            self.ctx.close()
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _results(allocs, F):
    X = F * (F - 1) // 2
    results = []
    for _, arrs, S in allocs:
        r = Result()
        r.k1D, r.Nmodes1D = arrs['k1D'], arrs['Nmodes1D']
        r.kpar, r.kper, r.Nmodes2D = arrs['kpar'], arrs['kper'], arrs['Nmodes2D']
        r.k3D, r.Nmodes3D = arrs['k3D'], arrs['Nmodes3D']
        if F == 1:
            r.Pk = arrs['Pk3D'][:, :, 0]
            r.Pk1D = arrs['Pk1D'][:, 0]
            r.Pk2D = arrs['Pk2D'][:, 0]
            r.Pkphase = arrs['Pkphase']
        else:
            r.Pk, r.XPk = arrs['Pk3D'][:, :, :F], arrs['Pk3D'][:, :, F:F + X]
            r.Pk1D, r.PkX1D = arrs['Pk1D'][:, :F], arrs['Pk1D'][:, F:F + X]
            r.Pk2D, r.PkX2D = arrs['Pk2D'][:, :F], arrs['Pk2D'][:, F:F + X]
        results.append(r)
    return results
