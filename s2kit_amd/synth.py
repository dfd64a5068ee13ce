"""Synthetic workloads on the GPU (torch RNG), laid out as the transforms
expect: coefficients [F, bw*bw] in IndexOfHarmonicCoeff order
(util.c:42-49), grids [F, 2bw, 2bw]."""
import numpy as np


def degrees(bw):
    """Degree l of each coefficient slot: orders 0..bw-1, then -(bw-1)..-1."""
    pos = [np.arange(m, bw) for m in range(bw)]
    neg = [np.arange(m, bw) for m in range(bw - 1, 0, -1)]
    return np.concatenate(pos + neg)


def orders(bw):
    pos = [np.full(bw - m, m) for m in range(bw)]
    neg = [np.full(bw - m, -m) for m in range(bw - 1, 0, -1)]
    return np.concatenate(pos + neg)


def power_law_coeffs(bw, nfields, seed, device):
    """Complex Gaussian coefficients with E|c|^2 = (l+1)^-2, split evenly over
    the real and imaginary parts."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    sd = torch.tensor(np.sqrt(0.5) / (degrees(bw) + 1.0), dtype=torch.float64, device=device)
    re = torch.randn((nfields, bw * bw), generator=g, dtype=torch.float64, device=device) * sd
    im = torch.randn((nfields, bw * bw), generator=g, dtype=torch.float64, device=device) * sd
    return re, im


def uniform_coeffs(bw, nfields, seed, device):
    """Coefficients with real and imaginary parts i.i.d. uniform[-1, 1]."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    re = torch.rand((nfields, bw * bw), generator=g, dtype=torch.float64, device=device) * 2 - 1
    im = torch.rand((nfields, bw * bw), generator=g, dtype=torch.float64, device=device) * 2 - 1
    return re, im


CONV_BWS = (32, 64, 128, 256, 512)
CONV_P = (0.3, 0.3, 0.2, 0.15, 0.05)


def conv_uneven_requests(nreq, seed, device, outputs=("o64", "omx")):
    """The uneven-bandwidth convolution workload: nreq requests with bw drawn
    from {32, 64, 128, 256, 512} (p = .3, .3, .2, .15, .05) by numpy's
    default_rng(seed), each a signal grid synthesised (FP64 inverse
    transform) from uniform[-1, 1] coefficients and a real zonal filter grid
    with h_l = exp(-l(l+1)/(2 sigma^2)), sigma ~ U[5, bw/2].

    Returns (bws, data): bws the request bandwidths in request order; data
    maps bw to a dict of [count, 2bw, 2bw] float64 device tensors, the
    requests of that bandwidth back to back in request order: "sre", "sim",
    "hre", "him", and for each name in `outputs` an uninitialised output
    pair name+"r", name+"i"."""
    import torch
    import s2kit_amd as S
    rng = np.random.default_rng(seed)
    bws = rng.choice(list(CONV_BWS), size=nreq, p=list(CONV_P))
    g = torch.Generator(device=device)
    g.manual_seed(seed + 5)
    data = {}
    for bw in sorted(set(bws.tolist())):
        cnt = int((bws == bw).sum())
        n = 2 * bw
        names = ["sre", "sim", "hre", "him"] + [o + s for o in outputs for s in ("r", "i")]
        t = {k: torch.empty((cnt, n, n), dtype=torch.float64, device=device) for k in names}
        cr = torch.rand((cnt, bw * bw), generator=g, dtype=torch.float64, device=device) * 2 - 1
        ci = torch.rand((cnt, bw * bw), generator=g, dtype=torch.float64, device=device) * 2 - 1
        p = S.Plan(bw, cnt, precision=S.FP64)
        p.inverse(cr, ci, t["sre"], t["sim"])
        sig = torch.tensor(rng.uniform(5, bw / 2, cnt), dtype=torch.float64, device=device)
        l = torch.arange(bw, dtype=torch.float64, device=device)
        hc = torch.zeros((cnt, bw * bw), dtype=torch.float64, device=device)
        hc[:, :bw] = torch.exp(-l[None, :] * (l[None, :] + 1) / (2 * sig[:, None] ** 2))
        p.inverse(hc, torch.zeros_like(hc), t["hre"], t["him"])
        torch.cuda.synchronize()
        p.close()
        data[bw] = t
    return bws, data


def conv_requests_of(bws, data, out):
    """Request tuples (bw, sre, sim, hre, him, ore, oim) in request order,
    outputs into data[bw][out + "r"], data[bw][out + "i"]."""
    slot = {bw: 0 for bw in data}
    reqs = []
    for bw in bws.tolist():
        j = slot[bw]
        slot[bw] += 1
        t = data[bw]
        reqs.append((bw, t["sre"][j], t["sim"][j], t["hre"][j], t["him"][j], t[out + "r"][j], t[out + "i"][j]))
    return reqs
