"""Seeded synthetic workloads of the benchmark configurations.

* gmm_integer: a Gaussian mixture rounded to integers and clamped to [0, 255] (the 10k x 128 and
  1M x 128 L2 bases). Centers, per-component sigmas and component weights come from
  `mixture_seed`; base vectors and queries are drawn from the same mixture with their own seeds.
* anisotropic_cosine: Gaussian vectors with eigen-spectrum lambda_i ~ i^-alpha under a seeded
  random rotation, unit-normalised (the 1.2M x 100 cosine base).
* integer_matrix: vecdex's test data (tests/hnsw_parallel_test.cpp:41-48): components in
  [-128, 127] from the top 8 bits of mt19937, which numpy's RandomState is.

All generation is vectorised numpy, chunked so peak memory stays near the output size.
"""

from __future__ import annotations

import numpy as np


def integer_matrix(rows: int, dim: int, seed: int) -> np.ndarray:
    raw = np.random.RandomState(seed).randint(0, 2**32, size=rows * dim, dtype=np.uint32)
    return ((raw >> 24).astype(np.int32) - 128).astype(np.float32).reshape(rows, dim)


def random_vectors(n: int, dim: int, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal((n, dim), dtype=np.float32)


class GaussianMixture:
    """Isotropic (sigma_low == sigma_high) or per-dimension log-uniform sigmas."""

    def __init__(self, dim: int, components: int, center_low: float, center_high: float,
                 sigma_low: float, sigma_high: float, mixture_seed: int):
        rng = np.random.default_rng(mixture_seed)
        self.dim = dim
        self.centers = rng.uniform(center_low, center_high, size=(components, dim)).astype(np.float32)
        if sigma_low == sigma_high:
            self.sigmas = np.full((components, dim), sigma_low, dtype=np.float32)
        else:
            self.sigmas = np.exp(rng.uniform(np.log(sigma_low), np.log(sigma_high),
                                             size=(components, dim))).astype(np.float32)

    def sample(self, n: int, seed: int, chunk: int = 1 << 18) -> np.ndarray:
        rng = np.random.default_rng(seed)
        out = np.empty((n, self.dim), dtype=np.float32)
        for start in range(0, n, chunk):
            m = min(chunk, n - start)
            comp = rng.integers(0, self.centers.shape[0], size=m)
            noise = rng.standard_normal((m, self.dim), dtype=np.float32)
            block = self.centers[comp] + noise * self.sigmas[comp]
            np.rint(block, out=block)
            np.clip(block, 0.0, 255.0, out=block)
            out[start:start + m] = block
        return out


def gmm64() -> GaussianMixture:
    """10,000 x 128 L2: 64 components, centers U[0,100], sigma 10."""
    return GaussianMixture(128, 64, 0.0, 100.0, 10.0, 10.0, mixture_seed=0)


def gmm1024(components: int = 1024) -> GaussianMixture:
    """1,000,000 x 128 L2: 1,024 components, centers U[0,128], per-dim sigma log-uniform [4,32]."""
    return GaussianMixture(128, components, 0.0, 128.0, 4.0, 32.0, mixture_seed=0)


def gmm_base_and_queries(mixture: GaussianMixture, n: int, nq: int, base_seed: int = 1,
                         query_seed: int = 2):
    return mixture.sample(n, base_seed), mixture.sample(nq, query_seed)


class AnisotropicGaussian:
    """lambda_i ~ i^-alpha, rotated by a seeded random orthogonal matrix (QR of a Gaussian)."""

    def __init__(self, dim: int = 100, alpha: float = 0.8, seed: int = 0):
        rng = np.random.default_rng(seed)
        self.dim = dim
        q, r = np.linalg.qr(rng.standard_normal((dim, dim)))
        q *= np.sign(np.diag(r))[None, :]
        scale = np.arange(1, dim + 1, dtype=np.float64) ** (-alpha / 2.0)
        self.transform = (scale[:, None] * q.T).astype(np.float32)  # x = z * sqrt(lambda) @ Q^T

    def sample(self, n: int, seed: int, chunk: int = 1 << 18) -> np.ndarray:
        rng = np.random.default_rng(seed)
        out = np.empty((n, self.dim), dtype=np.float32)
        for start in range(0, n, chunk):
            m = min(chunk, n - start)
            z = rng.standard_normal((m, self.dim), dtype=np.float32)
            x = z @ self.transform
            x /= np.linalg.norm(x, axis=1, keepdims=True)
            out[start:start + m] = x
        return out
