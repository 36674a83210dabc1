"""Synthetic test matrices for the RBRP column ID (host-side NumPy generators).

These replace the reference's generators (gpu/driver.cpp:23-111,
matlab/Matrix_Gaussian_exp.m:1-15, RBRP.ipynb c5:49-159), which draw from
unseeded global RNGs.  Every generator here takes an explicit seed.

* ``haar(m, n, rng)``: ``Q`` of the QR of an ``m x n`` N(0,1) matrix with its
  columns multiplied by ``sign(diag(R))`` (Haar on the Stiefel manifold).
* ``spectrum_matrix``: ``A = U diag(s) V^T`` with Haar ``U`` and ``V``.
* ``gmm_rows``: rows from a Gaussian mixture plus a scaled outlier component.
* ``kernel_blocks``: Gaussian-kernel blocks between two separated clusters.
"""

from __future__ import annotations

import numpy as np


def haar(m: int, n: int, rng: np.random.Generator) -> np.ndarray:
    G = rng.standard_normal((m, n))
    Q, R = np.linalg.qr(G)
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return np.asfortranarray(Q * d)


def exp_spectrum(n: int, rate: float) -> np.ndarray:
    """``s_i = exp(-i / rate)`` for ``i = 1..n``."""
    return np.exp(-np.arange(1, n + 1, dtype=np.float64) / rate)


def poly_spectrum(n: int, power: float = 1.0) -> np.ndarray:
    """``s_i = i^-power`` for ``i = 1..n``."""
    return np.arange(1, n + 1, dtype=np.float64) ** (-power)


def spectrum_matrix(m: int, n: int, s: np.ndarray, seed: int) -> np.ndarray:
    """``A = U diag(s) V^T`` (m x n, Fortran order) with Haar ``U`` (m x r), ``V`` (n x r)."""
    r = min(m, n, len(s))
    rng = np.random.default_rng(seed)
    U = haar(m, r, rng)
    V = haar(n, r, rng)
    return np.asfortranarray((U * s[:r]) @ V.T)


def haar_torch(m: int, n: int, gen, device="cuda"):
    """Device version of :func:`haar` (Q of a seeded N(0,1) matrix, sign-fixed)."""
    import torch
    G = torch.randn((m, n), generator=gen, dtype=torch.float64, device=device)
    Q, R = torch.linalg.qr(G)
    d = torch.sign(torch.diagonal(R))
    d[d == 0] = 1.0
    return Q * d


def spectrum_matrix_torch(m: int, n: int, s: np.ndarray, seed: int, device="cuda"):
    """``A = U diag(s) V^T`` generated on the device; returns a column-major (m, n) view."""
    import torch
    r = min(m, n, len(s))
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    U = haar_torch(m, r, gen, device)
    V = haar_torch(n, r, gen, device)
    st = torch.as_tensor(np.asarray(s[:r], dtype=np.float64), device=device)
    At = V @ (U * st).t()            # (n, m) row-major == A column-major
    del U
    return At.t()


def poly_outlier_torch(m: int, n: int, seed: int, power: float = 1.0, n_outlier=None,
                       device="cuda"):
    """FP32-stored ``A = U diag(s) V^T`` with ``s_i = i^-power`` (Haar U, V, built in
    FP64 on the device), then ``n_outlier`` rows (default 0.05% of m) replaced by
    i.i.d. standard-Cauchy rows.  Returns a column-major (m, n) float32 view."""
    import torch
    if n_outlier is None:
        n_outlier = int(round(5e-4 * m))
    r = min(m, n)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    V = haar_torch(n, r, gen, device)
    U = haar_torch(m, r, gen, device)
    st = torch.as_tensor(poly_spectrum(r, power), device=device)
    At = torch.empty((n, m), dtype=torch.float32, device=device)   # column-major (m, n)
    step = 1 << 17
    for r0 in range(0, m, step):                                   # bounded FP64 temporaries
        r1 = min(m, r0 + step)
        At[:, r0:r1] = (V @ (U[r0:r1] * st).t()).float()
    del U
    rows = torch.randperm(m, generator=gen, device=device)[:n_outlier]
    out = torch.empty((n_outlier, n), dtype=torch.float32, device=device)
    out.cauchy_(generator=gen)
    At[:, rows] = out.t()
    return At.t()


def optimal_error(s: np.ndarray, k: int) -> float:
    """Best rank-k relative Frobenius error ``sqrt(sum_{i>k} s_i^2) / ||s||``."""
    s = np.asarray(s, dtype=np.float64)
    return float(np.sqrt(np.sum(s[k:] ** 2)) / np.sqrt(np.sum(s ** 2)))


def gmm_rows(m: int, d: int, seed: int, components: int = 20, var: float = 0.01,
             outlier_frac: float = 1e-3, outlier_scale: float = 50.0) -> np.ndarray:
    """``m x d`` matrix whose rows come from a ``components``-way Gaussian mixture.

    Means ~ N(0, I), covariance ``var * I``, weights ~ Dirichlet(1).  A further
    ``outlier_frac`` of the rows comes from one extra component scaled by
    ``outlier_scale`` (drawn from the same recipe, then multiplied).
    """
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal((components + 1, d))
    wts = rng.dirichlet(np.ones(components))
    n_out = int(round(m * outlier_frac))
    n_in = m - n_out
    lab = rng.choice(components, size=n_in, p=wts)
    X = np.empty((m, d), dtype=np.float64, order="F")
    X[:n_in] = mu[lab] + np.sqrt(var) * rng.standard_normal((n_in, d))
    X[n_in:] = outlier_scale * (mu[components] + np.sqrt(var) * rng.standard_normal((n_out, d)))
    order = rng.permutation(m)
    return np.asfortranarray(X[order])


def kernel_block(m: int, n: int, rng: np.random.Generator, sigma: float = 0.5,
                 shift: float = 4.0) -> np.ndarray:
    """Gaussian-kernel block ``K_ij = exp(-||x_i - y_j||^2 / (2 sigma^2))``.

    ``x`` uniform in the unit cube, ``y`` uniform in the unit cube shifted by
    ``(shift, 0, 0)``.
    """
    x = rng.random((m, 3))
    y = rng.random((n, 3))
    y[:, 0] += shift
    d2 = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return np.asfortranarray(np.exp(-d2 / (2.0 * sigma * sigma)))


def kernel_blocks(count: int, seed: int, m_range=(64, 1024), n_range=(64, 512)):
    """``count`` kernel blocks with ``m ~ U{m_range}``, ``n ~ U{n_range}``."""
    rng = np.random.default_rng(seed)
    ms = rng.integers(m_range[0], m_range[1] + 1, size=count)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    return [kernel_block(int(a), int(c), rng) for a, c in zip(ms, ns)]


def kernel_block_torch(m: int, n: int, gen, sigma: float = 0.5, shift: float = 4.0,
                       device="cuda"):
    """``kernel_block`` generated on the device from the torch generator ``gen``;
    returns a column-major (m, n) view."""
    import torch
    x = torch.rand((m, 3), generator=gen, dtype=torch.float64, device=device)
    y = torch.rand((n, 3), generator=gen, dtype=torch.float64, device=device)
    y[:, 0] += shift
    d2 = ((y[:, None, :] - x[None, :, :]) ** 2).sum(-1)      # (n, m) row-major == (m, n) col-major
    return torch.exp(-d2 / (2.0 * sigma * sigma)).t()


def kernel_blocks_torch(count: int, seed: int, m_range=(64, 1024), n_range=(64, 512),
                        device="cuda"):
    """``count`` device kernel blocks with ``m ~ U{m_range}``, ``n ~ U{n_range}`` (seeded)."""
    import torch
    rng = np.random.default_rng(seed)
    ms = rng.integers(m_range[0], m_range[1] + 1, size=count)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    return [kernel_block_torch(int(a), int(c), gen, device=device) for a, c in zip(ms, ns)]


def gmm_rows_torch(m: int, d: int, seed: int, components: int = 20, var: float = 0.01,
                   outlier_frac: float = 1e-3, outlier_scale: float = 50.0, device="cuda"):
    """``gmm_rows`` drawn on the device (same recipe; the mixture parameters come
    from NumPy, the rows from a seeded torch generator); column-major (m, d) view."""
    import torch
    rng = np.random.default_rng(seed)
    mu = torch.as_tensor(rng.standard_normal((components + 1, d)), device=device)
    wts = rng.dirichlet(np.ones(components))
    n_out = int(round(m * outlier_frac))
    n_in = m - n_out
    lab = torch.as_tensor(rng.choice(components, size=n_in, p=wts), device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    Xt = torch.empty((d, m), dtype=torch.float64, device=device)   # column-major (m, d)
    Xt[:, :n_in] = (mu[lab] + float(np.sqrt(var)) *
                    torch.randn((n_in, d), generator=gen, dtype=torch.float64, device=device)).t()
    Xt[:, n_in:] = (outlier_scale * (mu[components] + float(np.sqrt(var)) * torch.randn(
        (n_out, d), generator=gen, dtype=torch.float64, device=device))).t()
    order = torch.as_tensor(rng.permutation(m), device=device)
    return Xt[:, order].t()

