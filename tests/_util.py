"""Helpers shared by the GPU parity tests (inputs, transfers, error norms)."""

import math

import numpy as np
import torch

from fluxoptics import inputs


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def c64(u: np.ndarray) -> np.ndarray:
    """Round to complex64 and promote back: the oracle sees exactly the device input."""
    return np.asarray(u, dtype=np.complex64).astype(np.complex128)


def to_dev(u: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(np.asarray(u, dtype=np.complex64))).cuda()


def to_host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.complex128)


def speckle_batch(batch, nx, ny, dx, dy, w0, seed):
    """Gaussian envelope times exp(i*phi), phi ~ U[0, 2pi) per pixel, per field."""
    g = inputs.gaussian(nx, ny, dx, dy, w0)
    out = np.empty((batch, ny, nx), np.complex128)
    for b in range(batch):
        phi = inputs.uniform(seed + b, (ny, nx), 0.0, 2 * math.pi)
        out[b] = g * np.exp(1j * phi)
    return out


def random_field(shape, seed):
    return inputs.normal(seed, shape) + 1j * inputs.normal(seed + 7919, shape)
