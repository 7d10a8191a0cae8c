"""Full-size parity and adjoint checks on the GPU.

512^3 R2C (f32 and f64) against the FP64 oracle on the same seeded N(0,1)
input (rel-L2 <= 2e-6 for f32, <= 1e-12 for f64), the normalized round
trip, the raw backward (= n^3 x), and dot-product tests of every adjoint
operation (<A x, y> = <x, A* y>, f64, relative difference <= 1e-12); the
largest size, 1024^3 f32, through a closed-form spectrum and the round trip.
"""
import numpy as np
import pytest
import torch

import oracle as O
import pencil_fft_hip as pf
from util import rel_l2

pytestmark = pytest.mark.gpu


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_512cubed_r2c_against_fp64_oracle(gpu):
    n = 512
    shape = (n, n, n)
    rng = np.random.default_rng(2)
    x32 = rng.standard_normal(shape, dtype=np.float32)
    print("oracle f64 forward of the f32 input ...", flush=True)
    ref = O.transform(O.KIND_R2C, x32.astype(np.float64), O.FORWARD, shape=shape, threads=16)
    for dt, bound in ((torch.float32, 2e-6), (torch.float64, 1e-12)):
        plan = pf.R2cPlan(shape, dtype=dt)
        xd = torch.from_numpy(x32).cuda().to(dt).reshape(-1)
        spec = plan.forward(xd)
        err = rel_l2(_host(spec), ref)
        print(f"{dt}: forward rel-L2 vs FP64 oracle {err:.3e}", flush=True)
        assert err <= bound, (dt, err)
        back = plan.inverse(spec)
        rt = float(torch.linalg.norm((back - xd).double()) / torch.linalg.norm(xd.double()))
        print(f"{dt}: round trip rel-L2 {rt:.3e}", flush=True)
        assert rt <= bound
        raw = plan.backward(spec)
        rb = float(torch.linalg.norm((raw - xd * float(n ** 3)).double()) / (torch.linalg.norm(xd.double()) * n ** 3))
        print(f"{dt}: raw backward rel-L2 vs n^3 x {rb:.3e}", flush=True)
        assert rb <= bound
        del plan, spec, back, raw, xd
        torch.cuda.empty_cache()


def _dot(a, b):
    a = torch.view_as_real(a) if a.is_complex() else a
    b = torch.view_as_real(b) if b.is_complex() else b
    return float(torch.sum(a.double() * b.double()))


def _rand(numel, complex_, g):
    if complex_:
        return torch.complex(torch.randn(numel, generator=g, device="cuda", dtype=torch.float64),
                             torch.randn(numel, generator=g, device="cuda", dtype=torch.float64))
    return torch.randn(numel, generator=g, device="cuda", dtype=torch.float64)


CASES = [
    ("c2c", (12, 10, 16), {}),
    ("c2c", (6, 5, 7, 4), {}),
    ("r2c", (12, 10, 16), {}),
    ("r2c", (9, 10, 15), {}),
    ("r2c", (8, 6, 10), {"selection": [0, 1]}),
    ("r2r", (12, 10, 16), {"kinds": [pf.R2rKind.DctII, pf.R2rKind.DstI, pf.R2rKind.DctIV]}),
    ("r2r", (9, 8, 7), {"kinds": [pf.R2rKind.DctI, pf.R2rKind.DstII, pf.R2rKind.DstIII]}),
    ("r2r", (8, 9, 6), {"kinds": [pf.R2rKind.DctIII, pf.R2rKind.DstIV, pf.R2rKind.Dht]}),
]


@pytest.mark.parametrize("kind,shape,kw", CASES)
def test_adjoint_dot_products_f64(gpu, kind, shape, kw):
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    if kind == "c2c":
        plan = pf.C2cPlan(shape, dtype=torch.complex128, **kw)
    elif kind == "r2c":
        plan = pf.R2cPlan(shape, dtype=torch.float64, **kw)
    else:
        plan = pf.R2rPlan(shape, dtype=torch.float64, **kw)
    _, _, n_in, b_in = plan.layout(0)
    _, _, n_out, b_out = plan.layout(1)
    for op, adj in ((pf.FORWARD, pf.ADJOINT_FORWARD), (pf.INVERSE, pf.ADJOINT_INVERSE),
                    (pf.BACKWARD, pf.ADJOINT_BACKWARD)):
        if op == pf.FORWARD:
            x = _rand(n_in, b_in == 16, g)
            y = _rand(n_out, b_out == 16, g)
        else:
            # INVERSE / BACKWARD map the output layout back; feed a valid spectrum
            x = plan.forward(_rand(n_in, b_in == 16, g))
            y = _rand(n_in, b_in == 16, g)
        ax = plan.execute(op, x)
        aty = plan.execute(adj, y)
        lhs, rhs = _dot(ax, y), _dot(x, aty)
        rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
        assert rel <= 1e-12, (kind, shape, op, lhs, rhs, rel)


def _axis_lines(t, axis, fn):
    """Applies fn to the lines of a logical 3D tensor along `axis` (made the
    contiguous last axis, transformed, moved back)."""
    nd = t.dim()
    order = [a for a in range(nd) if a != axis] + [axis]
    moved = t.permute(order).contiguous()
    out = fn(moved.reshape(-1)).reshape(moved.shape)
    inv = [order.index(a) for a in range(nd)]
    return out.permute(inv).contiguous()


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_mixed_r2r_and_dht_256x384x240(gpu):
    """256x384x240 real U[-1,1) (seed 3): R2rPlan with (DCT-II, DST-I,
    DCT-IV) per axis and a DhtPlan over all axes, f32 and f64, against the
    FP64 oracle (which follows the reference's embedding algorithm,
    r2r.rs:697-796), against per-axis LocalR2rPlan / LocalDhtPlan results,
    and the normalized inverse (DCT-III, DST-I, DCT-IV) round trip."""
    shape = (256, 384, 240)
    kinds = [pf.R2rKind.DctII, pf.R2rKind.DstI, pf.R2rKind.DctIV]
    rng = np.random.default_rng(3)
    x32 = rng.uniform(-1.0, 1.0, size=shape).astype(np.float32)
    x64 = x32.astype(np.float64)
    print("oracle f64 R2R and DHT of the f32-rounded input ...", flush=True)
    ref_r2r = O.transform(O.KIND_R2R, x64, O.FORWARD, shape=shape, r2r=[int(k) for k in kinds], threads=16)
    ref_dht = O.transform(O.KIND_R2R, x64, O.FORWARD, shape=shape, r2r=[O.DHT] * 3, threads=16)
    for dt, bound in ((torch.float32, 2e-6), (torch.float64, 1e-12)):
        xd = torch.from_numpy(x32).cuda().to(dt)
        for name, plan, ref, ks in (("r2r", pf.R2rPlan(shape, kinds, dtype=dt), ref_r2r, kinds),
                                    ("dht", pf.DhtPlan(shape, dtype=dt), ref_dht, [pf.R2rKind.Dht] * 3)):
            y = plan.forward(xd.reshape(-1))
            err = rel_l2(_host(y), ref)
            print(f"{dt} {name}: forward rel-L2 vs FP64 oracle {err:.3e}", flush=True)
            assert err <= bound, (dt, name, err)
            # the route against the local line plans applied axis by axis
            loc = xd
            for axis in (2, 1, 0):
                n, k = shape[axis], ks[axis]
                lp = pf.LocalDhtPlan(n, dtype=dt) if k == pf.R2rKind.Dht else pf.LocalR2rPlan(n, k, dtype=dt)
                loc = _axis_lines(loc, axis, lp.forward)
            # route output memory order is the reversed axes
            y_log = y.reshape(shape[::-1]).permute(2, 1, 0)
            lerr = float(torch.linalg.norm((y_log - loc).double()) / torch.linalg.norm(loc.double()))
            print(f"{dt} {name}: route vs local plans rel-L2 {lerr:.3e}", flush=True)
            assert lerr <= bound, (dt, name, lerr)
            back = plan.inverse(y)
            rt = float(torch.linalg.norm((back - xd.reshape(-1)).double()) / torch.linalg.norm(xd.double()))
            print(f"{dt} {name}: inverse round trip rel-L2 {rt:.3e}", flush=True)
            assert rt <= bound, (dt, name, rt)
            del plan, y, back, loc
        torch.cuda.empty_cache()


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_1024cubed_r2c_closed_form_and_round_trip(gpu):
    """The largest size exercised: 1024^3 f32 (4 GiB real, 4.3 GB half
    spectrum). A single cosine mode cos(2 pi (a i + b j + c k) / n) has two
    non-zero bins, n^3 / 2 at (a, b, c) and at its mirror, which the half
    spectrum (kz >= 0) does not store; every other stored bin must be at
    rounding level. Then a random N(0,1) field's
    normalized round trip (size independent: rel-L2 <= 2e-6)."""
    n = 1024
    plan = pf.R2cPlan((n, n, n), dtype=torch.float32)
    ws = plan.allocate_workspace()
    a, b, c = 3, 5, 7
    i = torch.arange(n, device="cuda", dtype=torch.float64)
    phase = (a * i).reshape(n, 1, 1) + (b * i).reshape(1, n, 1) + (c * i).reshape(1, 1, n)
    x = torch.cos(2 * torch.pi * torch.remainder(phase, n) / n).to(torch.float32).reshape(-1)
    del phase
    spec = plan.forward(x, workspace=ws)
    torch.cuda.synchronize()
    # memory order [kz][ky][kx], kz = 0 .. n/2
    s = spec.reshape(n // 2 + 1, n, n)
    peak = s[c, b, a]
    assert abs(peak.real.item() - n ** 3 / 2) <= 1e-5 * n ** 3 and abs(peak.imag.item()) <= 1e-5 * n ** 3
    s[c, b, a] = 0
    assert float(s.abs().max()) <= 1e-5 * n ** 3
    del s, spec
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x.normal_(generator=g)
    spec = plan.forward(x, workspace=ws)
    back = plan.inverse(spec, workspace=ws)
    err = float(torch.linalg.vector_norm((back - x).double()) / torch.linalg.vector_norm(x.double()))
    print(f"1024^3 f32 round trip rel-L2 {err:.2e}", flush=True)
    assert err <= 2e-6
