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
