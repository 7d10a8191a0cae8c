"""Pseudo-spectral time stepping on the device path, and its gradient with
respect to the initial field through the adjoint transforms.

Model (real periodic field u on an n x n x n grid, integer wavenumbers
k_a in {0, 1, ..., n/2, -n/2 + 1, ..., -1}, the R2C axis keeping 0..n/2):

    w_j     = C2R( E * R2C(u_j) )       normalized inverse (R2cPlan.inverse)
    u_{j+1} = w_j + gamma * w_j^2       gamma = 0.1 * dt
    E(k)    = exp(-nu |k|^2 dt)          nu = 0.01, dt = 1e-3
    L       = sum_x u_N(x)^2             N = 20 steps

The gradient dL/du_0 is the reverse sweep

    g_N = 2 u_N
    g_w = g_{j+1} * (1 + 2 gamma w_j)
    g_j = R2C^T( E * C2R^T(g_w) )        ADJOINT_INVERSE, then ADJOINT_FORWARD

with u_j checkpointed every `checkpoint_every` steps: each segment is
recomputed forward from its checkpoint (keeping its w_j) before its reverse
sweep. The transforms and the pointwise steps run on the device through the
C ABI (pencil_fft_hip); there is no CPU fallback.
"""
from __future__ import annotations

from typing import List, Optional, Tuple

import torch

import pencil_fft_hip as pf


def wavenumbers(n: int, half: bool = False) -> List[int]:
    """Integer wavenumbers of a length-n periodic axis in FFT order (the
    first n/2 + 1 of them for the R2C axis)."""
    k = [i if i <= n // 2 else i - n for i in range(n)]
    return k[: n // 2 + 1] if half else k


class PseudoSpectralModel:
    def __init__(self, n: int, dtype=torch.float64, nu: float = 0.01, dt: float = 1e-3,
                 gamma: Optional[float] = None, steps: int = 20, checkpoint_every: int = 5):
        if dtype not in (torch.float32, torch.float64):
            raise TypeError("real field dtype must be float32 or float64")
        self.n, self.dtype, self.nu, self.dt = n, dtype, nu, dt
        self.gamma = 0.1 * dt if gamma is None else gamma
        self.steps, self.every = steps, checkpoint_every
        self.plan = pf.R2cPlan((n, n, n), dtype=dtype)
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        self.ws = self.plan.allocate_workspace()
        self.spec = torch.empty(self.plan.output_elements(), dtype=self.cdtype, device="cuda")
        self.sq_ws = torch.empty(8192, dtype=torch.uint8, device="cuda")
        # per-axis factors of E (the spectrum's logical shape is n, n, n/2+1)
        self.decay = self.axis_factors(lambda k: -nu * dt * k * k)

    # -- helpers ------------------------------------------------------------
    def axis_factors(self, fn) -> torch.Tensor:
        """exp(fn(k_a)) per axis, concatenated in logical axis order, in f64
        on the host and rounded to the field's real dtype."""
        import math

        vals = []
        for a in range(3):
            vals += [math.exp(fn(k)) for k in wavenumbers(self.n, half=(a == 2))]
        return torch.tensor(vals, dtype=torch.float64).to(self.dtype).cuda()

    def _filtered(self, u: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """out = C2R(E * R2C(u))."""
        self.plan.forward(u, self.spec, self.ws)
        pf.scale_axes(self.plan, 1, self.spec, self.decay)
        self.plan.inverse(self.spec, out, self.ws, flags=pf.EXEC_DEFER_STATUS)
        return out

    def step(self, u: torch.Tensor, w: torch.Tensor, u_next: torch.Tensor):
        self._filtered(u, w)
        pf.quadratic(w, self.gamma, out=u_next)

    def loss(self, u: torch.Tensor) -> torch.Tensor:
        return pf.sum_squares(u, self.sq_ws)

    def initial_field(self, seed: int = 5) -> torch.Tensor:
        """White N(0,1) noise (seed) whose spectrum is multiplied by
        exp(-|k|^2 / 64), back to real space."""
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        xi = torch.randn(self.n ** 3, generator=g, device="cuda", dtype=torch.float64).to(self.dtype)
        smooth = self.axis_factors(lambda k: -k * k / 64.0)
        u0 = torch.empty_like(xi)
        self.plan.forward(xi, self.spec, self.ws)
        pf.scale_axes(self.plan, 1, self.spec, smooth)
        self.plan.inverse(self.spec, u0, self.ws, flags=pf.EXEC_DEFER_STATUS)
        self.plan.take_status(self.ws)
        return u0

    # -- forward / reverse ---------------------------------------------------
    def forward(self, u0: torch.Tensor) -> Tuple[float, torch.Tensor, List[torch.Tensor]]:
        """Runs the N steps; returns (loss, u_N, checkpoints u_0, u_e, u_2e, ...)."""
        u = u0.clone()
        w = torch.empty_like(u0)
        nxt = torch.empty_like(u0)
        ckpt = []
        for j in range(self.steps):
            if j % self.every == 0:
                ckpt.append(u.clone())
            self.step(u, w, nxt)
            u, nxt = nxt, u
        self.plan.take_status(self.ws)
        return float(self.loss(u).item()), u, ckpt

    def _reverse(self, ckpt: List[torch.Tensor], gbar: torch.Tensor) -> torch.Tensor:
        """J^T gbar for u_0 -> u_N: segment by segment from the last, each
        recomputed forward from its checkpoint (keeping its w_j), then swept
        backwards through the adjoint step."""
        g = gbar.clone()
        gw = torch.empty_like(g)
        spec = torch.empty_like(self.spec)
        for s in range(len(ckpt) - 1, -1, -1):
            j0 = s * self.every
            j1 = min(j0 + self.every, self.steps)
            ws_seg = []
            u = ckpt[s].clone()
            nxt = torch.empty_like(u)
            for _ in range(j0, j1):
                w = torch.empty_like(u)
                self.step(u, w, nxt)
                ws_seg.append(w)
                u, nxt = nxt, u
            for j in range(j1 - 1, j0 - 1, -1):
                pf.quadratic_adjoint(ws_seg[j - j0], g, self.gamma, out=gw)
                self.plan.adjoint_inverse(gw, spec, self.ws)
                pf.scale_axes(self.plan, 1, spec, self.decay)
                self.plan.adjoint_forward(spec, g, self.ws)
        self.plan.take_status(self.ws)
        return g

    def gradient(self, u0: torch.Tensor) -> Tuple[float, torch.Tensor]:
        """(L, dL/du_0): the forward sweep with checkpoints, then the reverse
        sweep seeded with dL/du_N = 2 u_N."""
        loss, u_n, ckpt = self.forward(u0)
        return loss, self._reverse(ckpt, 2.0 * u_n)

    def tangent(self, u0: torch.Tensor, du0: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """(u_N, du_N): the forward sweep with its linearization applied to du0
        (for the dot test <J du, gbar> = <du, J^T gbar>)."""
        u, du = u0.clone(), du0.clone()
        w = torch.empty_like(u0)
        dw = torch.empty_like(u0)
        nxt = torch.empty_like(u0)
        for _ in range(self.steps):
            self._filtered(u, w)
            self._filtered(du, dw)
            # du_{j+1} = dw (1 + 2 gamma w): the same map as the adjoint step (diagonal)
            pf.quadratic_adjoint(w, dw, self.gamma, out=du)
            pf.quadratic(w, self.gamma, out=nxt)
            u, nxt = nxt, u
        self.plan.take_status(self.ws)
        return u, du

    def vjp(self, u0: torch.Tensor, gbar: torch.Tensor) -> torch.Tensor:
        """J^T gbar for the map u_0 -> u_N."""
        _, _, ckpt = self.forward(u0)
        return self._reverse(ckpt, gbar)
