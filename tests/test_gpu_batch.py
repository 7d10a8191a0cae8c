"""GPU parity of the batched uneven-size C2C path.

The reference runs a collection of plans one after another
(crates/pencil-fft/src/distributed/collections.rs:229-311); the batched
path must give each problem the result of its own C2cPlan. Small batches
are checked against the oracle; the full workload (96 problems, edges from
{48,60,64,84,96,112,128,160}, extra dim [3]) is checked problem by problem
against the per-problem plans on the device, plus a round trip.
"""
import numpy as np
import pytest
import torch

import oracle as O
import pencil_fft_hip as pf
from util import rel_l2, tol

pytestmark = pytest.mark.gpu

EDGES = (48, 60, 64, 84, 96, 112, 128, 160)


def _tdev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _shapes(rng, count, edges):
    return [tuple(int(e) for e in rng.choice(edges, size=3)) for _ in range(count)]


def _inputs(rng, shapes, extra, cd):
    out = []
    for s in shapes:
        n = int(np.prod(s)) * extra
        out.append((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(cd))
    return out


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("autotune", [False, True])
def test_small_batch_matches_oracle(dt, autotune):
    rng = np.random.default_rng(41 + int(autotune))
    shapes = [(6, 10, 8), (12, 5, 7), (16, 16, 16), (3, 9, 20), (1, 4, 13), (8, 1, 1)]
    xs = _inputs(rng, shapes, 3, dt)
    torch_dt = torch.complex64 if dt == np.complex64 else torch.complex128
    plan = pf.BatchC2cPlan(shapes, dtype=torch_dt, autotune=autotune, extra_shapes=[(3,)] * len(shapes))
    if autotune:
        block, maxr, col, one, persistent, plane, streams = plan.tuning()
        assert block in (128, 256, 512) and maxr in (8, 16) and col in (0, 32) and one in (True, False)
        assert persistent in (True, False) and not plane  # no length here has a plane plan
        assert streams in (1, 2)
    srcs = [_tdev(x) for x in xs]
    for op in (O.FORWARD, O.INVERSE, O.BACKWARD):
        for use_graph in (False, True):
            outs = plan.execute(op, srcs, use_graph=use_graph)
            for k, (s, x) in enumerate(zip(shapes, xs)):
                ref = O.transform(O.KIND_C2C, x.astype(np.complex128), op, shape=s, extra=3)
                err = rel_l2(_host(outs[k]), ref)
                assert err <= tol(dt), (op, use_graph, k, s, err)


def test_tuning_cache_reuses_the_measured_tuning(tmp_path, monkeypatch):
    """An autotuned batch is measured once: the same problems on the same
    device reuse the tuning in the process and, after the process map is
    cleared (as in a new process), from the cache file; the file-tuned plan
    gives the oracle's results."""
    path = tmp_path / "batch_tuning.txt"
    monkeypatch.setenv("PFH_TUNE_CACHE", str(path))
    pf.clear_tuning_cache()
    shapes = [(48, 60, 64), (12, 7, 9), (84, 16, 20)]
    ex = [(3,)] * len(shapes)
    try:
        a = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=ex)
        assert a.tuning_source == "measured"
        b = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=ex)
        assert b.tuning_source == "process" and b.tuning() == a.tuning()
        pf.clear_tuning_cache()
        c = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=ex)
        assert c.tuning_source == "file" and c.tuning() == a.tuning()
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and lines[0].startswith("pfhbt3;arch=gfx950"), lines
        assert pf.BatchC2cPlan(shapes, autotune=False, extra_shapes=ex).tuning_source == "default"
        rng = np.random.default_rng(5)
        xs = _inputs(rng, shapes, 3, np.complex64)
        outs = c.forward([_tdev(x) for x in xs], use_graph=True)
        for k, (s, x) in enumerate(zip(shapes, xs)):
            ref = O.transform(O.KIND_C2C, x.astype(np.complex128), O.FORWARD, shape=s, extra=3)
            assert rel_l2(_host(outs[k]), ref) <= tol(np.complex64), (k, s)
    finally:
        pf.clear_tuning_cache()


@pytest.mark.parametrize("persistent", [1, 0])
def test_two_streams_match_oracle(tmp_path, monkeypatch, persistent):
    """Problems whose first pass is a plane stage run on a second stream,
    concurrently with the others (a tuning record with two streams in the
    cache file forces it): every operation, with and without a graph, against
    the oracle and the per-problem plans, twice in a row."""
    path = tmp_path / "batch_tuning.txt"
    monkeypatch.setenv("PFH_TUNE_CACHE", str(path))
    pf.clear_tuning_cache()
    shapes = [(48, 60, 64), (128, 112, 20), (96, 60, 7), (160, 84, 5), (112, 96, 48), (160, 128, 48), (12, 7, 9),
              (84, 16, 20)]
    ex = [(3,)] * len(shapes)
    try:
        pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=ex)
        key = path.read_text().splitlines()[0].split(" ")[0]
        with open(path, "a") as f:
            f.write(f"{key} 256 8 0 1 {persistent} 1 2 0.0\n")
        pf.clear_tuning_cache()
        plan = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=ex)
        assert plan.tuning_source == "file"
        t = plan.tuning()
        assert t[5] and t[6] == 2 and t[4] == bool(persistent), t
        rng = np.random.default_rng(90 + persistent)
        xs = _inputs(rng, shapes, 3, np.complex64)
        srcs = [_tdev(x) for x in xs]
        for op in (O.FORWARD, O.INVERSE, O.BACKWARD):
            for use_graph in (False, True, True):
                outs = plan.execute(op, srcs, use_graph=use_graph)
                for k, (s, x) in enumerate(zip(shapes, xs)):
                    got = _host(outs[k])
                    ref = O.transform(O.KIND_C2C, x.astype(np.complex128), op, shape=s, extra=3)
                    assert rel_l2(got, ref) <= tol(np.complex64), (op, use_graph, k, s)
    finally:
        pf.clear_tuning_cache()


def test_graph_replays_match_eager():
    """Graph replays, with eager work and synchronizations between them,
    give the same bits as the direct launches."""
    rng = np.random.default_rng(7)
    shapes = _shapes(rng, 8, (12, 20, 28, 32))
    xs = _inputs(rng, shapes, 3, np.complex64)
    plan = pf.BatchC2cPlan(shapes, autotune=False, extra_shapes=[(3,)] * len(shapes))
    srcs = [_tdev(x) for x in xs]
    ref = [_host(t) for t in plan.forward(srcs, use_graph=False)]
    dsts = [torch.empty_like(t) for t in srcs]
    ws = plan.allocate_workspace()
    for rep in range(4):
        for d in dsts:
            d.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.forward(srcs, dsts, workspace=ws, use_graph=True)
        got = [_host(t) for t in dsts]
        for k in range(len(shapes)):
            assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), (rep, k)
        _ = (srcs[0] * 2).sum().item()  # unrelated eager work between replays


def test_new_buffers_after_graph_capture():
    rng = np.random.default_rng(8)
    shapes = [(8, 12, 10), (20, 6, 4)]
    plan = pf.BatchC2cPlan(shapes, autotune=False)
    for trial in range(3):
        xs = _inputs(rng, shapes, 1, np.complex64)
        outs = plan.forward([_tdev(x) for x in xs], use_graph=True)
        for k, (s, x) in enumerate(zip(shapes, xs)):
            ref = O.transform(O.KIND_C2C, x.astype(np.complex128), O.FORWARD, shape=s)
            assert rel_l2(_host(outs[k]), ref) <= tol(np.complex64), (trial, k)


def test_rejects_short_workspace():
    plan = pf.BatchC2cPlan([(4, 4, 4)], autotune=False)
    x = torch.zeros(64, dtype=torch.complex64, device="cuda")
    with pytest.raises(pf.PfhError):
        plan.forward([x], workspace=torch.zeros(0, dtype=torch.uint8, device="cuda"))


@pytest.mark.parametrize("extra", [1, 3])
def test_plane_stages_match_oracle_and_per_problem_plans(extra):
    """Problems with plane plans on two consecutive route axes run those two
    stages as one fused plane stage: the inner pair (forward: axes 1 then 0;
    inverse / backward: axes 0 then 1; e.g. (48, 60, 64), (160, 84, 5)) or
    the outer pair where its plane is smaller ((112, 96, 48),
    (160, 128, 48), whose (x, y) plane does not fit in LDS): every
    operation against the oracle and against the problem's own three-stage
    C2cPlan, with and without a graph; the other problems keep their line
    stages in the same launches."""
    rng = np.random.default_rng(60 + extra)
    shapes = [(48, 60, 64), (128, 112, 20), (96, 60, 7), (160, 84, 5), (64, 160, 3), (112, 96, 48),
              (160, 128, 48), (12, 7, 9), (84, 16, 20), (60, 48, 1)]
    xs = _inputs(rng, shapes, extra, np.complex64)
    ex = [(extra,) if extra > 1 else ()] * len(shapes)
    plan = pf.BatchC2cPlan(shapes, autotune=False, extra_shapes=ex)
    assert plan.tuning()[5], "plane stages are expected for these shapes"
    srcs = [_tdev(x) for x in xs]
    for op in (O.FORWARD, O.INVERSE, O.BACKWARD):
        for use_graph in (False, True):
            outs = plan.execute(op, srcs, use_graph=use_graph)
            for k, (s, x) in enumerate(zip(shapes, xs)):
                got = _host(outs[k])
                ref = O.transform(O.KIND_C2C, x.astype(np.complex128), op, shape=s, extra=extra)
                assert rel_l2(got, ref) <= tol(np.complex64), (op, use_graph, k, s)
                single = pf.C2cPlan(s, extra_shape=ex[k]).execute(op, srcs[k])
                assert rel_l2(got, _host(single)) <= 2 * tol(np.complex64), (op, use_graph, k, s)


@pytest.mark.slow
def test_full_workload_matches_per_problem_plans():
    """96 problems, f32, extra [3], autotuned, under a graph: every problem
    within the f32 bound of its own C2cPlan result and of the FP64 oracle on a
    sample, and the normalized inverse returns the input."""
    rng = np.random.default_rng(20261202)
    shapes = _shapes(rng, 96, EDGES)
    plan = pf.BatchC2cPlan(shapes, autotune=True, extra_shapes=[(3,)] * 96)
    srcs = []
    for s in shapes:
        n = int(np.prod(s)) * 3
        srcs.append(torch.randn(n, dtype=torch.complex64, device="cuda"))
    outs = plan.forward(srcs, use_graph=True)
    torch.cuda.synchronize()
    for k, s in enumerate(shapes):
        single = pf.C2cPlan(s, extra_shape=(3,)).forward(srcs[k])
        num = torch.linalg.vector_norm((outs[k] - single).to(torch.complex128)).item()
        den = torch.linalg.vector_norm(single.to(torch.complex128)).item()
        assert num / den <= 2 * tol(np.complex64), (k, s, num / den)
    for k in (0, 17, 95):
        x = srcs[k].cpu().numpy().astype(np.complex128)
        ref = O.transform(O.KIND_C2C, x, O.FORWARD, shape=shapes[k], extra=3)
        assert rel_l2(_host(outs[k]), ref) <= tol(np.complex64), k
    back = plan.inverse(outs, use_graph=True)
    torch.cuda.synchronize()
    for k in range(96):
        num = torch.linalg.vector_norm((back[k] - srcs[k]).to(torch.complex128)).item()
        den = torch.linalg.vector_norm(srcs[k].to(torch.complex128)).item()
        assert num / den <= tol(np.complex64), (k, num / den)
