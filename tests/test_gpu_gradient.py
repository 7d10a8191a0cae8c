"""Gradient of a pseudo-spectral run with respect to its initial field
(real 256^3, f64 and f32; 20 steps of R2C -> exp(-nu |k|^2 dt) -> C2R ->
u + 0.1 dt u^2, nu = 0.01, dt = 1e-3; loss ||u_20||^2; checkpoints every 5
steps), computed with the adjoint R2C / C2R operations on the device.

Checks: the dot-product identity <J du, g> = <du, J^T g> of the whole
20-step map (f64, 1e-12); central finite differences with eps = 1e-4 (f64);
checkpointing every 5 steps gives the same bits as storing every step; and,
at a size the oracle runs quickly, the gradient assembled entirely from the
oracle's FP64 transforms and adjoints (the reference's route restated on the
CPU) against the device gradient (1e-12).
"""
import numpy as np
import pytest
import torch

import pencil_fft_hip as pf
from gradient_ref import oracle_gradient
from pseudo_spectral import PseudoSpectralModel

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return float(torch.sum(a.double() * b.double()).item())


@pytest.mark.slow
def test_dot_identity_of_the_20_step_map_f64(gpu):
    m = PseudoSpectralModel(256, dtype=torch.float64)
    u0 = m.initial_field(5)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    du = torch.randn(256 ** 3, generator=g, device="cuda", dtype=torch.float64)
    gbar = torch.randn(256 ** 3, generator=g, device="cuda", dtype=torch.float64)
    _, jdu = m.tangent(u0, du)
    jtg = m.vjp(u0, gbar)
    lhs, rhs = _dot(jdu, gbar), _dot(du, jtg)
    rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"dot identity: {lhs:.15e} vs {rhs:.15e} (rel {rel:.2e})", flush=True)
    assert rel <= 1e-12


@pytest.mark.slow
def test_central_finite_differences_f64(gpu):
    m = PseudoSpectralModel(256, dtype=torch.float64)
    u0 = m.initial_field(5)
    loss, grad = m.gradient(u0)
    g = torch.Generator(device="cuda")
    g.manual_seed(12)
    v = torch.randn(256 ** 3, generator=g, device="cuda", dtype=torch.float64)
    v /= torch.linalg.vector_norm(v)
    eps = 1e-4
    lp, _, _ = m.forward(u0 + eps * v)
    lm, _, _ = m.forward(u0 - eps * v)
    fd = (lp - lm) / (2 * eps)
    an = _dot(grad, v)
    rel = abs(fd - an) / abs(an)
    print(f"loss {loss:.6e}; FD {fd:.12e} vs adjoint {an:.12e} (rel {rel:.2e})", flush=True)
    assert rel <= 1e-6


@pytest.mark.slow
def test_checkpointing_matches_storing_every_step(gpu):
    a = PseudoSpectralModel(256, dtype=torch.float32, checkpoint_every=5)
    b = PseudoSpectralModel(256, dtype=torch.float32, checkpoint_every=1)
    u0 = a.initial_field(5)
    la, ga = a.gradient(u0)
    lb, gb = b.gradient(u0)
    assert la == lb
    assert torch.equal(ga, gb)
    # f32 against f64
    c = PseudoSpectralModel(256, dtype=torch.float64)
    _, g64 = c.gradient(u0.double())
    rel = float(torch.linalg.vector_norm(ga.double() - g64) / torch.linalg.vector_norm(g64))
    print(f"f32 gradient vs f64 gradient rel-L2 {rel:.2e}", flush=True)
    assert rel <= 1e-4


def test_gradient_matches_oracle_built_adjoint_f64(gpu):
    n = 24
    m = PseudoSpectralModel(n, dtype=torch.float64)
    u0 = m.initial_field(5)
    loss, grad = m.gradient(u0)
    lref, gref = oracle_gradient(u0.cpu().numpy(), n)
    rel = np.linalg.norm(grad.cpu().numpy() - gref) / np.linalg.norm(gref)
    print(f"loss {loss:.15e} vs oracle {lref:.15e}; gradient rel-L2 {rel:.2e}", flush=True)
    assert abs(loss - lref) <= 1e-12 * abs(lref)
    assert rel <= 1e-12


def test_pointwise_kernels(gpu):
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    for dt in (torch.float32, torch.float64):
        w = torch.randn(100003, generator=g, device="cuda", dtype=torch.float64).to(dt)
        gg = torch.randn(100003, generator=g, device="cuda", dtype=torch.float64).to(dt)
        a = 0.37
        rtol = 1e-6 if dt == torch.float32 else 1e-14
        # bounds relative to the terms' magnitudes (the sums cancel near their roots)
        u = pf.quadratic(w, a)
        assert bool(torch.all((u - (w + a * w * w)).abs() <= rtol * (w.abs() + a * w * w)))
        out = pf.quadratic_adjoint(w, gg, a)
        assert bool(torch.all((out - gg * (1 + 2 * a * w)).abs() <= rtol * gg.abs() * (1 + 2 * a * w.abs())))
        s = pf.sum_squares(w).item()
        ref = float(torch.sum(w.double() ** 2))
        assert abs(s - ref) <= 1e-12 * ref
        c = torch.complex(w, gg)
        sc = pf.sum_squares(c).item()
        assert abs(sc - float(torch.sum(w.double() ** 2 + gg.double() ** 2))) <= 1e-12 * sc
    # separable scaling on a spectrum, against the same product built in torch
    n = 12
    plan = pf.R2cPlan((n, n, n), dtype=torch.float64)
    x = torch.randn(n ** 3, generator=g, device="cuda", dtype=torch.float64)
    spec = plan.forward(x)
    f = [torch.arange(1, n + 1, device="cuda", dtype=torch.float64) * (a + 1) * 0.1 for a in range(2)]
    f.append(torch.arange(1, n // 2 + 2, device="cuda", dtype=torch.float64) * 0.3)
    ref = spec.reshape(n // 2 + 1, n, n) * (f[2][:, None, None] * f[1][None, :, None] * f[0][None, None, :])
    got = pf.scale_axes(plan, 1, spec.clone(), torch.cat(f))
    assert torch.allclose(got.reshape(ref.shape), ref, rtol=1e-14, atol=0)



def test_spectral_derivative_with_complex_axis_factors(gpu):
    """d/dx of u = sin(2 x) cos(3 y) (1 + cos z) on a 32 x 24 x 16 grid
    (x = 2 pi i / n): R2C, multiply by i k_x on axis 0 (complex per-axis
    factors; 0 at the Nyquist wavenumber, which keeps the derivative real),
    normalized C2R, against 2 cos(2 x) cos(3 y) (1 + cos z); and the factor
    product against torch's complex arithmetic on the spectrum. (With a field
    whose kz = 0 plane is zero the product would leave only rounding noise in
    that plane, with imaginary parts as large as its real parts, and the C2R
    endpoint check would rightly reject it.)"""
    import math

    n0, n1, n2 = 32, 24, 16
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-6)):
        plan = pf.R2cPlan((n0, n1, n2), dtype=dt)
        x = torch.arange(n0, dtype=torch.float64) * 2 * math.pi / n0
        y = torch.arange(n1, dtype=torch.float64) * 2 * math.pi / n1
        z = torch.arange(n2, dtype=torch.float64) * 2 * math.pi / n2
        X, Y, Z = torch.meshgrid(x, y, z, indexing="ij")
        u = (torch.sin(2 * X) * torch.cos(3 * Y) * (1 + torch.cos(Z))).to(dt).cuda().reshape(-1)
        du = (2 * torch.cos(2 * X) * torch.cos(3 * Y) * (1 + torch.cos(Z))).cuda().reshape(-1)
        spec = plan.forward(u)
        kx = [k if k < n0 // 2 else (0 if k == n0 // 2 else k - n0) for k in range(n0)]
        cdt = torch.complex64 if dt == torch.float32 else torch.complex128
        f = torch.cat([torch.tensor([1j * k for k in kx], dtype=torch.complex128),
                       torch.ones(n1, dtype=torch.complex128), torch.ones(n2 // 2 + 1, dtype=torch.complex128)])
        f = f.to(cdt).cuda()
        ref = spec.reshape(n2 // 2 + 1, n1, n0) * f[:n0].reshape(1, 1, n0)
        pf.scale_axes(plan, 1, spec, f)
        assert torch.allclose(spec.reshape(ref.shape), ref, rtol=tol, atol=tol * float(ref.abs().max()))
        got = plan.inverse(spec)
        err = float((got.double() - du).abs().max())
        assert err <= 10 * tol * 2, err
    with pytest.raises(pf.PfhError):
        pf.scale_axes(plan, 0, u, f)  # complex factors on the real side
