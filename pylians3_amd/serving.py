"""Continuous-batching request loop for MA + Pk / XPk requests on one GPU.

A request is a density-field -> power-spectrum job:
    Request(pos, dims, BoxSize, MAS, axis, W=None, pos2=None, W2=None)
        MA(pos[, W]) onto a dims^3 grid, delta = grid/mean - 1, then Pk(delta)
        (or, with pos2, XPk of the two fields).  pos/W are contiguous torch
        tensors on the GPU (or numpy arrays).
The server keeps `slots` worker contexts, each with its own HIP stream, its
resident FFT plans and its scratch memory.  A dispatcher admits queued requests
while a slot is free and the estimated HBM need of all in-flight work stays
within the budget; requests that arrive while others run are admitted as
slots and memory free up, and their kernels overlap on the device (one stream
per slot).  Requests are admitted in arrival order (policy='fifo', the default:
at moderate load it gives the lowest tail latency, since a large request then
starts as soon as a slot frees instead of waiting behind smaller ones); 'sjf'
takes the shortest waiting job first (work = particles + grid nodes) unless one
has waited longer than max_wait_s, and 'hrrn' the highest response ratio
(wait + service) / service.  Small requests with the same (dims, MAS, axis,
kind) that are waiting together are fused into one batched launch sequence
(pyl_ma_batch + pyl_pk_batch).  Each result is bitwise the same whether the
request runs alone, in a batch or alongside other requests: the scatter and
the binning are order-independent/deterministic.

    with Server(hbm_budget=64 << 30, slots=4) as srv:
        fut = srv.submit(Request(pos, 256, 1000.0, 'CIC', 0))
        res = fut.result()          # .Pk, .k3D, ... ; .latency_s
"""
import collections
import ctypes
import queue
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _native as N
from .MAS_library import _MAS_IDS
from .Pk_library import MAS_function


class Request:
    def __init__(self, pos, dims, BoxSize, MAS='CIC', axis=0, W=None, pos2=None, W2=None):
        self.pos, self.dims, self.BoxSize, self.MAS, self.axis = pos, int(dims), float(BoxSize), MAS, int(axis)
        self.W, self.pos2, self.W2 = W, pos2, W2
        self.t_submit = None
        self.t_start = None
        self.ready = None      # event on the submitter's stream after which the inputs are complete

    @property
    def nfields(self):
        return 1 if self.pos2 is None else 2

    def key(self):
        return (self.dims, self.MAS, self.axis, self.nfields, self.W is not None,
                self.W2 is not None, self.BoxSize)

    def particles(self):
        n = self.pos.shape[0]
        return n + (0 if self.pos2 is None else self.pos2.shape[0])

    def work(self):
        """Relative cost: particles deposited plus grid nodes transformed and binned."""
        return self.particles() + 4 * self.nfields * self.dims ** 3

    def service_s(self):
        """Rough GPU seconds alone: a fixed cost per request, about 30 ps per particle
        deposited and 10 ps per grid node transformed and binned (MI355X)."""
        return 3e-4 + 3e-11 * self.particles() + 1e-11 * self.nfields * self.dims ** 3

    def hbm_bytes(self):
        """Estimated device bytes while running (scratch + FFT buffers): per particle
        the sorted record (16 B with weights), the bucketed or spill copy (16 B) and the
        capacity sort's tile slots (3 records)."""
        d = self.dims
        n = self.particles()
        cells = d ** 3
        return int(80 * n + self.nfields * (4 * cells + 8 * d * d * (d // 2 + 1)) * 2 + 4 * cells
                   + (64 << 20))


class Result:
    pass


def _is_small(req):
    return req.dims <= 128 and req.particles() <= 4_000_000


class Server:
    def __init__(self, hbm_budget=None, slots=4, max_batch=32, device=0, policy='fifo',
                 max_wait_s=2.0, atomic=False):
        import torch
        self.device = device
        # atomic=True: per-particle atomic scatter (the plain GPU form, for comparison)
        self.ma_flags = N.MA_ATOMIC if atomic else 0
        if policy not in ('sjf', 'fifo', 'hrrn'):
            raise ValueError('policy must be "sjf", "hrrn" or "fifo"')
        self.policy = policy
        self.max_wait_s = float(max_wait_s)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(free * 0.6)
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.in_flight_bytes = 0
        self.free_slots = list(range(slots))
        self.ctxs = [N.Context(device) for _ in range(slots)]
        # one default-priority stream per slot: torch glue and library kernels share it
        self.streams = [torch.cuda.Stream(device=device) for _ in range(slots)]
        for c, s in zip(self.ctxs, self.streams):
            c.set_stream(s.cuda_stream)
        self.stop = False
        self.completed = []
        # one resident worker thread per slot, fed by the dispatcher
        self.work = [queue.SimpleQueue() for _ in range(slots)]
        self.workers = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                        for i in range(slots)]
        for w in self.workers:
            w.start()
        self.dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self.dispatcher.start()

    # ------------------------------------------------------------- client side
    def submit(self, req):
        if req.hbm_bytes() > self.budget:
            raise N.BudgetError(N.PYL_EBUDGET, 'request needs %d bytes, budget is %d'
                                % (req.hbm_bytes(), self.budget))
        fut = Future()
        if any(N.is_device_array(a) for a in (req.pos, req.W, req.pos2, req.W2)):
            # the slot stream must not read the inputs before the work that wrote
            # them on the submitter's stream is done
            import torch
            req.ready = torch.cuda.Event()
            req.ready.record(torch.cuda.current_stream())
        req.t_submit = time.perf_counter()
        with self.cv:
            self.queue.append((req, fut))
            self.cv.notify_all()
        return fut

    def warm(self, req):
        """Runs `req` once on every slot, one slot after the other, so that each slot's
        context has grown its scratch buffers and built its plans for requests up to
        that size before serving starts (a slot's first large request otherwise pays
        the device allocations).  Call before submitting requests: it uses the slot
        contexts directly and waits for each run."""
        with self.cv:
            busy = len(self.free_slots) != len(self.ctxs) or bool(self.queue)
        if busy:
            raise RuntimeError('warm() needs an idle server')
        for slot in range(len(self.ctxs)):
            self._run_batch(self.ctxs[slot], self.streams[slot], [req])

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        self.dispatcher.join()
        for q in self.work:
            q.put(None)
        for w in self.workers:
            w.join()
        for c in self.ctxs:
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def latencies(self):
        return np.array([r.latency_s for r in self.completed])

    # ------------------------------------------------------------- scheduling
    def _pick(self):
        """Index of the next request: the oldest ('fifo'); the highest response ratio
        ('hrrn'); or the shortest job (work ~ particles + grid nodes, 'sjf'), except
        that a request waiting longer than `max_wait_s` goes first (oldest first), so
        that large requests are not starved."""
        now = time.perf_counter()
        if self.policy == 'fifo':
            return 0
        if self.policy == 'hrrn':
            # highest response ratio (wait + service) / service first
            return max(range(len(self.queue)),
                       key=lambda i: (now - self.queue[i][0].t_submit) / self.queue[i][0].service_s())
        oldest = min(range(len(self.queue)), key=lambda i: self.queue[i][0].t_submit)
        if now - self.queue[oldest][0].t_submit > self.max_wait_s:
            return oldest
        return min(range(len(self.queue)), key=lambda i: self.queue[i][0].work())

    def _dispatch(self):
        while True:
            with self.cv:
                while not self.stop and (not self.queue or not self.free_slots):
                    self.cv.wait(0.05)
                if self.stop and not self.queue:
                    return
                if not self.queue or not self.free_slots:
                    continue
                i = self._pick()
                head = self.queue[i][0]
                if head.hbm_bytes() + self.in_flight_bytes > self.budget:
                    self.cv.wait(0.05)              # wait for memory to free up
                    continue
                group = [self.queue[i]]
                del self.queue[i]
                need = head.hbm_bytes()
                if _is_small(head):
                    # fuse waiting requests of the same shape into one batched launch
                    k = head.key()
                    keep = collections.deque()
                    while self.queue and len(group) < self.max_batch:
                        item = self.queue.popleft()
                        if item[0].key() == k and need + item[0].hbm_bytes() + \
                                self.in_flight_bytes <= self.budget:
                            group.append(item)
                            need += item[0].hbm_bytes()
                        else:
                            keep.append(item)
                    keep.extend(self.queue)
                    self.queue = keep
                slot = self.free_slots.pop()
                self.in_flight_bytes += need
            self.work[slot].put((group, need))

    def _worker(self, slot):
        while True:
            item = self.work[slot].get()
            if item is None:
                return
            self._run(slot, *item)

    def _run(self, slot, group, need):
        ctx = self.ctxs[slot]
        t0 = time.perf_counter()
        for r, _ in group:
            r.t_start = t0
        try:
            results = self._run_batch(ctx, self.streams[slot], [g[0] for g in group])
            t1 = time.perf_counter()
            for (req, fut), res in zip(group, results):
                res.latency_s = t1 - req.t_submit
                res.queue_s = t0 - req.t_submit
                res.batch = len(group)
                self.completed.append(res)
                fut.set_result(res)
        except Exception as e:  # noqa: BLE001 - forwarded to the caller's future
            for _, fut in group:
                if not fut.done():
                    fut.set_exception(e)
        finally:
            with self.cv:
                self.free_slots.append(slot)
                self.in_flight_bytes -= need
                self.cv.notify_all()

    # ------------------------------------------------------------- execution
    @staticmethod
    def _ptr(a):
        return N.device_ptr(a) if N.is_device_array(a) else a.ctypes.data

    def _grid_fields(self, ctx, reqs, grids):
        """MA of field f of every request into grids[f] (nreal, d, d, d), overwrite.
        Returns the mean density of every (request, field) as a device float32
        (nreal, F) tensor, as the reference caller computes it (Pk_snapshot.py:77-83):
        len(pos)/dims**3, or sum(W)/dims**3 (float64 sum), rounded to float32; a
        zero mean is replaced by 1.  Nothing here waits for the device."""
        import torch
        d = reqs[0].dims
        F = reqs[0].nfields
        mas = _MAS_IDS[reqs[0].MAS]
        means = torch.empty((len(reqs), F), dtype=torch.float64, device='cuda')
        for f in range(F):
            pos = [r.pos if f == 0 else r.pos2 for r in reqs]
            Ws = [r.W if f == 0 else r.W2 for r in reqs]
            counts = [p.shape[0] for p in pos]
            offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
            P = torch.cat([torch.as_tensor(p, device='cuda') for p in pos]) if len(reqs) > 1 \
                else torch.as_tensor(pos[0], device='cuda')
            P = P.contiguous()
            Wc = None
            if Ws[0] is not None:
                Wd = [torch.as_tensor(w, device='cuda') for w in Ws]
                Wc = torch.cat(Wd) if len(reqs) > 1 else Wd[0]
                Wc = Wc.contiguous()
                for i, w in enumerate(Wd):
                    means[i, f] = w.double().sum() / float(d) ** 3
            else:
                for i, c in enumerate(counts):
                    means[i, f] = float(c) / float(d) ** 3
            N.check(N.lib().pyl_ma_batch(ctx.handle, N.device_ptr(P), offs.ctypes.data, len(reqs), 3,
                                         None if Wc is None else N.device_ptr(Wc),
                                         N.device_ptr(grids[f]), d, reqs[0].BoxSize, mas,
                                         N.PTR_DEVICE | N.MA_OVERWRITE | N.NO_SYNC | self.ma_flags))
        means = means.float()
        return torch.where(means == 0, torch.ones_like(means), means).contiguous()

    def _run_batch(self, ctx, stream, reqs):
        import torch
        d = reqs[0].dims
        F = reqs[0].nfields
        with torch.cuda.stream(stream):
            for r in reqs:
                if r.ready is not None:
                    stream.wait_event(r.ready)
                    for a in (r.pos, r.W, r.pos2, r.W2):
                        if N.is_device_array(a):
                            a.record_stream(stream)
            grids = [torch.empty((len(reqs), d, d, d), dtype=torch.float32, device='cuda')
                     for _ in range(F)]
            scales = self._grid_fields(ctx, reqs, grids)
            shifts = torch.ones_like(scales)
            mas = (ctypes.c_int * F)(*([MAS_function(reqs[0].MAS)] * F))
            allocs = [N.alloc_pk_out(d, F) for _ in reqs]
            outs = (N.PkOut * len(reqs))(*[a[0] for a in allocs])
            ptrs = (ctypes.c_void_p * F)(*[N.device_ptr(g) for g in grids])
            # returns once the spectra are on the host (the call waits for its own work)
            N.check(N.lib().pyl_pk_batch(ctx.handle, ptrs, len(reqs), F, mas, d,
                                         reqs[0].BoxSize, reqs[0].axis, N.device_ptr(scales),
                                         N.device_ptr(shifts),
                                         N.PTR_DEVICE | N.PK_DESTROY_INPUT | N.PK_SCALES_DEVICE,
                                         outs))
            del grids, scales, shifts
        results = []
        X = F * (F - 1) // 2
        for _, arrs, S in allocs:
            r = Result()
            r.k3D, r.Nmodes3D, r.k1D, r.kpar, r.kper = (arrs['k3D'], arrs['Nmodes3D'], arrs['k1D'],
                                                        arrs['kpar'], arrs['kper'])
            r.Pk = arrs['Pk3D'][:, :, 0] if F == 1 else arrs['Pk3D'][:, :, :F]
            r.Pk2D = arrs['Pk2D'][:, 0] if F == 1 else arrs['Pk2D'][:, :F]
            r.Pk1D = arrs['Pk1D'][:, 0] if F == 1 else arrs['Pk1D'][:, :F]
            if F > 1:
                r.XPk = arrs['Pk3D'][:, :, F:F + X]
                r.PkX2D = arrs['Pk2D'][:, F:F + X]
            results.append(r)
        return results
