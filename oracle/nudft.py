"""FP64 direct nonuniform DFTs (the oracle of every accuracy test).

Restates test/utils/dirft1d.hpp, dirft2d.hpp, dirft3d.hpp: for type 1

    f[k] = sum_j c[j] exp(i * isign * k . x_j),   k_d in -(N_d//2) .. (N_d-1)//2

(CMCL order, modeord=0; modeord=1 is FFT order), type 2 is the transpose
``c[j] = sum_k f[k] exp(i * isign * k . x_j)``, and type 3 is
``f[k] = sum_j c[j] exp(i * isign * s_k . x_j)`` at arbitrary frequencies.
The relative l2 error is test/utils/norms.hpp:21-37.

Convention: array axis d pairs with coordinate d (``pts[d]``), exactly as the
Python ``Plan`` interface; the C layout is the reversed view.

Everything is computed in complex128 by chunked matrix products, so sizes up
to M * N ~ 1e8 finish in seconds.
"""

from __future__ import annotations

import numpy as np

CHUNK = 2048


def freqs(n: int, modeord: int = 0) -> np.ndarray:
    """Integer frequencies of the n modes of one axis, in storage order."""
    k = np.arange(-(n // 2), n - n // 2)
    if modeord == 1:
        k = np.fft.ifftshift(k)
    return k


def _phase(x: np.ndarray, k: np.ndarray, isign: int) -> np.ndarray:
    return np.exp(1j * isign * np.outer(np.asarray(x, np.float64), k.astype(np.float64)))


def type1(pts, c, isign, n_modes, modeord=0):
    """All modes. pts: list of dim arrays (M,), c: (M,) or (ntr, M)."""
    c = np.asarray(c, np.complex128)
    batched = c.ndim == 2
    cb = c if batched else c[None]
    dim = len(pts)
    ks = [freqs(n, modeord) for n in n_modes]
    out = np.zeros((cb.shape[0],) + tuple(n_modes), np.complex128)
    M = cb.shape[1]
    for j0 in range(0, M, CHUNK):
        sl = slice(j0, min(M, j0 + CHUNK))
        e = [_phase(pts[d][sl], ks[d], isign) for d in range(dim)]
        for t in range(cb.shape[0]):
            ec = e[0] * cb[t, sl, None]
            if dim == 1:
                out[t] += ec.sum(axis=0)
            elif dim == 2:
                out[t] += ec.T @ e[1]
            else:
                e12 = (e[1][:, :, None] * e[2][:, None, :]).reshape(len(e[1]), -1)
                out[t] += (ec.T @ e12).reshape(n_modes)
    return out if batched else out[0]


def type1_at(pts, c, isign, n_modes, flat_idx, modeord=0):
    """Type 1 at selected flat (C-order) mode indices; c: (M,) or (ntr, M)."""
    c = np.asarray(c, np.complex128)
    batched = c.ndim == 2
    cb = c if batched else c[None]
    sub = np.unravel_index(np.asarray(flat_idx), tuple(n_modes))
    kk = [freqs(n, modeord)[s] for n, s in zip(n_modes, sub)]
    out = np.zeros((cb.shape[0], len(kk[0])), np.complex128)
    M = cb.shape[1]
    for j0 in range(0, M, CHUNK):
        sl = slice(j0, min(M, j0 + CHUNK))
        ph = sum(np.outer(kk[d].astype(np.float64), np.asarray(pts[d][sl], np.float64))
                 for d in range(len(pts)))
        out += (np.exp(1j * isign * ph) @ cb[:, sl].T).T
    return out if batched else out[0]


def type2(pts, f, isign, modeord=0, idx=None):
    """Values at all points (or at the points idx). f: n_modes or (ntr, *n_modes)."""
    dim = len(pts)
    f = np.asarray(f, np.complex128)
    batched = f.ndim == dim + 1
    fb = f if batched else f[None]
    n_modes = fb.shape[1:]
    ks = [freqs(n, modeord) for n in n_modes]
    P = [np.asarray(p, np.float64) if idx is None else np.asarray(p, np.float64)[idx]
         for p in pts]
    M = len(P[0])
    out = np.zeros((fb.shape[0], M), np.complex128)
    for j0 in range(0, M, CHUNK):
        sl = slice(j0, min(M, j0 + CHUNK))
        e = [_phase(P[d][sl], ks[d], isign) for d in range(dim)]
        for t in range(fb.shape[0]):
            if dim == 1:
                out[t, sl] = e[0] @ fb[t]
            elif dim == 2:
                out[t, sl] = np.einsum("ja,jb,ab->j", e[0], e[1], fb[t], optimize=True)
            else:
                e12 = (e[1][:, :, None] * e[2][:, None, :]).reshape(len(e[1]), -1)
                g = e12 @ fb[t].reshape(n_modes[0], -1).T  # (J, N0)
                out[t, sl] = (g * e[0]).sum(axis=1)
    return out if batched else out[0]


def type2_grad(pts, f, isign, modeord=0):
    """Derivatives of the type-2 values with respect to the point coordinates:
    out[d] = sum_k f_k (i isign k_d) exp(i isign k . x_j), shape (dim, [ntr,] M)
    (the closed-form derivative of the direct sum)."""
    dim = len(pts)
    f = np.asarray(f, np.complex128)
    n_modes = f.shape[-dim:]
    out = []
    for d in range(dim):
        shape = [1] * f.ndim
        shape[f.ndim - dim + d] = n_modes[d]
        kd = freqs(n_modes[d], modeord).astype(np.float64).reshape(shape)
        out.append(type2(pts, f * (1j * isign * kd), isign, modeord))
    return np.stack(out)


def type3(pts, c, isign, targets, idx=None):
    """f at targets (or targets[idx]); c: (M,) or (ntr, M)."""
    c = np.asarray(c, np.complex128)
    batched = c.ndim == 2
    cb = c if batched else c[None]
    T = [np.asarray(t, np.float64) if idx is None else np.asarray(t, np.float64)[idx]
         for t in targets]
    X = [np.asarray(p, np.float64) for p in pts]
    K = len(T[0])
    out = np.zeros((cb.shape[0], K), np.complex128)
    for k0 in range(0, K, CHUNK):
        sl = slice(k0, min(K, k0 + CHUNK))
        for j0 in range(0, len(X[0]), CHUNK):
            sj = slice(j0, min(len(X[0]), j0 + CHUNK))
            ph = sum(np.outer(T[d][sl], X[d][sj]) for d in range(len(X)))
            out[:, sl] += (np.exp(1j * isign * ph) @ cb[:, sj].T).T
    return out if batched else out[0]


def relerr(a, b) -> float:
    """||a - b||_2 / ||b||_2 with b the reference (norms.hpp:21-37 uses ||a||)."""
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
