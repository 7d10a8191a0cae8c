"""Shared helpers of the test suite (seeded inputs, error metrics)."""
import numpy as np


def rel_l2(actual, expected) -> float:
    a = np.asarray(actual, dtype=np.complex128).reshape(-1)
    e = np.asarray(expected, dtype=np.complex128).reshape(-1)
    den = np.linalg.norm(e)
    num = np.linalg.norm(a - e)
    if den == 0:
        return float(num)
    return float(num / den)


def tol(dtype) -> float:
    """Relative L2 bound of the north star: 2e-6 for f32, 1e-12 for f64."""
    return 2e-6 if np.dtype(dtype) in (np.float32, np.complex64) else 1e-12


def uniform_real(rng, shape, dtype):
    return rng.uniform(-1.0, 1.0, size=shape).astype(dtype)


def uniform_complex(rng, shape, dtype):
    c = rng.uniform(-1.0, 1.0, size=shape) + 1j * rng.uniform(-1.0, 1.0, size=shape)
    return c.astype(dtype)


def real_of(dtype):
    return np.float32 if np.dtype(dtype) in (np.float32, np.complex64) else np.float64


def complex_of(dtype):
    return np.complex64 if np.dtype(dtype) in (np.float32, np.complex64) else np.complex128
