"""CPU pin of the host gradient reference used by the GPU gradient tests:
the oracle-built reverse sweep of the 20-step pseudo-spectral map agrees
with central finite differences of its own forward sweep."""
import numpy as np

from gradient_ref import oracle_gradient


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    n = 8
    u0 = rng.standard_normal(n ** 3) * 0.5
    loss, g = oracle_gradient(u0, n)
    v = rng.standard_normal(n ** 3)
    v /= np.linalg.norm(v)
    eps = 1e-5
    lp, _ = oracle_gradient(u0 + eps * v, n)
    lm, _ = oracle_gradient(u0 - eps * v, n)
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - g @ v) <= 1e-6 * abs(g @ v), (fd, g @ v)


def test_a_wrong_adjoint_is_caught():
    """Replacing the C2R adjoint by the plain R2C forward in the reverse sweep
    (a plausible mistake: it drops the Hermitian weights and the 1/n^3)
    breaks the finite-difference agreement."""
    import oracle as O

    rng = np.random.default_rng(2)
    n = 8
    u0 = rng.standard_normal(n ** 3) * 0.5
    _, g = oracle_gradient(u0, n, adj_inverse_op=O.FORWARD)
    v = rng.standard_normal(n ** 3)
    v /= np.linalg.norm(v)
    eps = 1e-5
    lp, _ = oracle_gradient(u0 + eps * v, n)
    lm, _ = oracle_gradient(u0 - eps * v, n)
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - g @ v) > 1e-2 * abs(fd)
