"""Synthetic inputs for the benchmark configurations (generated on the GPU with torch).

Recipes:
  uniform(n, box, seed)            positions uniform in [0, box)^3
  zeldovich(d, box, seed, f_rsd)   Gaussian random field on a d^3 grid with
                                   P(k) = A k^-1.5 exp(-(k/0.5)^2) (k in h/Mpc),
                                   A set by sigma(R=8 Mpc/h) = 0.8 (top-hat);
                                   Zel'dovich displacement psi_k = i k/k^2 delta_k
                                   of a d^3 lattice q = (i+1/2) box/d; optional
                                   redshift-space shift x_0 += f psi_0; positions
                                   wrapped into [0, box).
The field is delta_k = rfftn(w) sqrt(P(k) d^3 / box^3) for white noise w ~ N(0,1),
so that the power-spectrum estimator of the linear field has expectation P(k).
These only build inputs; nothing here is part of the measured path.
"""
import math

import numpy as np

_A_CACHE = {}


def _sigma8_amplitude(R=8.0, sigma=0.8):
    """A such that sigma(R) = sigma for P(k) = A k^-1.5 exp(-(k/0.5)^2)."""
    key = (R, sigma)
    if key not in _A_CACHE:
        from scipy import integrate

        def integrand(lnk):
            k = math.exp(lnk)
            x = k * R
            W = 3.0 * (math.sin(x) - x * math.cos(x)) / x ** 3 if x > 1e-4 else 1.0 - x * x / 10.0
            return k ** 3 * k ** -1.5 * math.exp(-(k / 0.5) ** 2) * W * W / (2 * math.pi ** 2)

        val, _ = integrate.quad(integrand, math.log(1e-5), math.log(20.0), limit=400)
        _A_CACHE[key] = sigma ** 2 / val
    return _A_CACHE[key]


def power_law_cutoff(k, A=None):
    A = _sigma8_amplitude() if A is None else A
    return A * k ** -1.5 * np.exp(-(k / 0.5) ** 2)


def uniform(n, box, seed, device='cuda'):
    import torch
    g = torch.Generator(device=device).manual_seed(int(seed))
    return torch.rand((n, 3), generator=g, device=device, dtype=torch.float32) * float(box)


def _kgrid(d, box, device):
    import torch
    kF = 2 * math.pi / box
    k1 = torch.fft.fftfreq(d, d=1.0 / d, device=device) * kF
    kz = torch.fft.rfftfreq(d, d=1.0 / d, device=device) * kF
    return k1, kz


def zeldovich(d, box, seed, f_rsd=0.0, rsd_axis=0, device='cuda', chunk_planes=None):
    """Zel'dovich-displaced d^3 lattice (positions float32, shape (d^3, 3), lattice order)."""
    import torch
    g = torch.Generator(device=device).manual_seed(int(seed))
    w = torch.randn((d, d, d), generator=g, device=device, dtype=torch.float32)
    wk = torch.fft.rfftn(w)
    del w
    k1, kz = _kgrid(d, box, device)
    A = _sigma8_amplitude()
    kx = k1.view(d, 1, 1)
    ky = k1.view(1, d, 1)
    kzz = kz.view(1, 1, d // 2 + 1)
    k2 = kx * kx + ky * ky + kzz * kzz
    k = torch.sqrt(k2)
    k[0, 0, 0] = 1.0
    amp = torch.sqrt(A * k ** -1.5 * torch.exp(-(k / 0.5) ** 2) * (d ** 3) / box ** 3)
    amp[0, 0, 0] = 0.0
    dk = wk * amp
    del wk, amp, k
    k2[0, 0, 0] = 1.0
    pos = torch.empty((d, d, d, 3), device=device, dtype=torch.float32)
    q = (torch.arange(d, device=device, dtype=torch.float32) + 0.5) * (box / d)
    for ax, kv in enumerate((kx, ky, kzz)):
        psik = dk * (1j * kv / k2)
        psik[0, 0, 0] = 0.0
        psi = torch.fft.irfftn(psik, s=(d, d, d))
        del psik
        shape = [1, 1, 1]
        shape[ax] = d
        x = q.view(*shape) + psi
        if f_rsd and ax == rsd_axis:
            x = x + f_rsd * psi
        pos[..., ax] = torch.remainder(x, box)
        del psi, x
    pos = pos.view(-1, 3)
    # remainder can return box for tiny negative inputs rounded up; keep [0, box)
    pos[pos >= box] = 0.0
    return pos


def displacement_field(d, box, seed, device='cuda'):
    """Zel'dovich displacement psi on a d^3 grid: (3, d, d, d) float32 (Mpc/h)."""
    import torch
    g = torch.Generator(device=device).manual_seed(int(seed))
    w = torch.randn((d, d, d), generator=g, device=device, dtype=torch.float32)
    wk = torch.fft.rfftn(w)
    del w
    k1, kz = _kgrid(d, box, device)
    A = _sigma8_amplitude()
    kx, ky, kzz = k1.view(d, 1, 1), k1.view(1, d, 1), kz.view(1, 1, d // 2 + 1)
    k2 = kx * kx + ky * ky + kzz * kzz
    k = torch.sqrt(k2)
    k[0, 0, 0] = 1.0
    amp = torch.sqrt(A * k ** -1.5 * torch.exp(-(k / 0.5) ** 2) * (d ** 3) / box ** 3)
    amp[0, 0, 0] = 0.0
    dk = wk * amp
    k2[0, 0, 0] = 1.0
    psi = torch.empty((3, d, d, d), device=device, dtype=torch.float32)
    for ax, kv in enumerate((kx, ky, kzz)):
        pk = dk * (1j * kv / k2)
        pk[0, 0, 0] = 0.0
        psi[ax] = torch.fft.irfftn(pk, s=(d, d, d))
    return psi


def tracers(n, psi, box, seed, weights=True, device='cuda'):
    """n tracers at q ~ U[0, box)^3 moved by psi CIC-interpolated at q (the batched
    and serving recipe), with lognormal(0, 1) weights.  Returns (pos, W)."""
    import torch
    from . import MAS_library as MASL
    g = torch.Generator(device=device).manual_seed(int(seed))
    q = torch.rand((n, 3), generator=g, device=device, dtype=torch.float32) * float(box)
    disp = torch.empty((n,), device=device, dtype=torch.float32)
    pos = torch.empty_like(q)
    for ax in range(3):
        MASL.CIC_interp(psi[ax].contiguous(), box, q, disp)
        pos[:, ax] = torch.remainder(q[:, ax] + disp, box)
    pos[pos >= box] = 0.0
    W = None
    if weights:
        W = torch.exp(torch.randn((n,), generator=g, device=device, dtype=torch.float32))
    return pos.contiguous(), W


def config_c2(device='cuda'):
    """Clustered field: 512^3 ZA particles, box 1000 Mpc/h, RSD along axis 0 (f=0.5), seed 2."""
    return zeldovich(512, 1000.0, 2, f_rsd=0.5, rsd_axis=0, device=device)


def batched_realisations(nreal=256, dims=128, box=500.0, seed0=1000, device='cuda'):
    """Batched workload: nreal independent realisations (box 500 Mpc/h, 128^3).
    Realisation r (seed seed0 + r): a Zel'dovich-displaced dims^3 matter lattice
    and N_r ~ Poisson(lambda_r) tracers (lambda_r log-uniform in [1e5, 2e6], N_r
    clipped to that range) at q ~ U[0, box)^3 moved by the CIC-interpolated
    displacement, with lognormal(0, 1) weights.  Returns a dict with the matter
    positions and per-realisation offsets (mpos, moffs), the tracer positions,
    weights and offsets (tpos, tw, toffs) and the tracer counts."""
    import torch
    rng = np.random.default_rng(seed0)
    lam = np.exp(rng.uniform(np.log(1e5), np.log(2e6), nreal))
    counts = np.clip(rng.poisson(lam), 100000, 2000000).astype(np.int64)
    matter, tr_pos, tr_w = [], [], []
    q = (torch.arange(dims, device=device, dtype=torch.float32) + 0.5) * (box / dims)
    lat = torch.stack(torch.meshgrid(q, q, q, indexing='ij'), -1).view(-1, 3)
    for r in range(nreal):
        psi = displacement_field(dims, box, seed0 + r, device=device)
        matter.append(torch.remainder(lat + psi.view(3, -1).t(), box))
        p, w = tracers(int(counts[r]), psi, box, seed0 + 5000 + r, device=device)
        tr_pos.append(p)
        tr_w.append(w)
        del psi
    mpos = torch.cat(matter).contiguous()
    mpos[mpos >= box] = 0.0
    tpos = torch.cat(tr_pos).contiguous()
    tw = torch.cat(tr_w).contiguous()
    return dict(mpos=mpos, moffs=np.arange(nreal + 1, dtype=np.int64) * dims ** 3,
                tpos=tpos, tw=tw, toffs=np.concatenate([[0], np.cumsum(counts)]).astype(np.int64),
                counts=counts, nreal=nreal, dims=dims, box=box)


def config_c3(device='cuda'):
    """Large single field: 1024^3 ZA particles, box 2000 Mpc/h, no RSD, seed 3."""
    return zeldovich(1024, 2000.0, 3, device=device)


def config_c1(device='cuda'):
    """128^3 uniform particles in a 1000 Mpc/h box, seed 1."""
    return uniform(128 ** 3, 1000.0, 1, device=device)
