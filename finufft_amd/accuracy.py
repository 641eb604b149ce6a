"""FP64 direct NUDFT on the GPU at sampled modes / targets (torch).

Measures the error of a transform on problems far too large for a CPU direct
sum (M = 1e8 points, 256^3 modes): the definition is that of the reference's
test/utils/dirft3d.hpp (modes from -N/2, array axis d paired with coordinate
d), evaluated in complex128 by chunked products on the GPU. Type 2 uses the
separable structure of the sum (one matrix product per axis).
"""

from __future__ import annotations

import torch


def _freqs(n, device):
    return torch.arange(-(n // 2), n - n // 2, device=device, dtype=torch.float64)


def type1_at(pts, c, isign, n_modes, flat_idx, chunk=1 << 19):
    """f at the flat (C-order) mode indices flat_idx; pts, c on the GPU."""
    dev = c.device
    idx = torch.as_tensor(flat_idx, device=dev, dtype=torch.int64)
    ks = []
    rem = idx.clone()
    for n in reversed(n_modes):
        ks.append(_freqs(n, dev)[rem % n])
        rem //= n
    ks = ks[::-1]
    out = torch.zeros(len(idx), dtype=torch.complex128, device=dev)
    M = c.shape[-1]
    for j0 in range(0, M, chunk):
        j1 = min(M, j0 + chunk)
        ph = sum(pts[d][j0:j1].to(torch.float64)[:, None] * ks[d][None, :] for d in range(len(pts)))
        e = torch.polar(torch.ones_like(ph), isign * ph)
        out += (e * c[j0:j1].to(torch.complex128)[:, None]).sum(dim=0)
    return out


def type2_at(pts, f, isign, pt_idx):
    """c at the points pt_idx; f has shape n_modes (array order), on the GPU."""
    dev = f.device
    idx = torch.as_tensor(pt_idx, device=dev, dtype=torch.int64)
    dim = len(pts)
    P = [pts[d][idx].to(torch.float64) for d in range(dim)]
    g = f.to(torch.complex128)
    n_modes = f.shape
    S = len(idx)
    # contract the last axis first: g[..., j]
    k = _freqs(n_modes[-1], dev)
    e = torch.polar(torch.ones(S, len(k), device=dev, dtype=torch.float64),
                    isign * P[-1][:, None] * k[None, :])  # (S, N_last)
    g = (g.reshape(-1, n_modes[-1]) @ e.T).reshape(*n_modes[:-1], S)
    for d in range(dim - 2, -1, -1):
        k = _freqs(n_modes[d], dev)
        e = torch.polar(torch.ones(S, len(k), device=dev, dtype=torch.float64),
                        isign * P[d][:, None] * k[None, :])  # (S, N_d)
        shape = g.shape  # (*n_modes[:d+1], S)
        g = g.reshape(-1, n_modes[d], S)
        g = (g * e.T[None, :, :]).sum(dim=1).reshape(*shape[:-2], S)
    return g.reshape(S)


def relerr(a, b) -> float:
    a = a.to(torch.complex128).reshape(-1)
    b = b.to(torch.complex128).reshape(-1)
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
