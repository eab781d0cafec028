"""The serving loop on the GPU: every result is the same alone and in the mix, requests
overlap on the device, and the admitted footprint bounds what a request allocates."""
import copy

import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _small_stream(n, seed, rate=50.0):
    from adfwi_amd.serving import make_stream
    reqs = make_stream(n, rate, seed, nz_range=(128, 160), nx_range=(256, 352), nt_range=(60, 140),
                       shots_range=(1, 4))
    for r in reqs:
        r.keep = True
    return reqs


def test_results_alone_equal_results_in_the_mix():
    from adfwi_amd.serving import ServingLoop, run_alone
    reqs = _small_stream(8, seed=7)
    assert {r.kind for r in reqs} == {"acoustic", "elastic"} and {r.mode for r in reqs} == {"forward", "gradient"}
    rep = ServingLoop(DEV, max_slots=3).run(copy.deepcopy(reqs), realtime=True)
    assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
    assert rep.max_concurrent >= 2
    alone = run_alone(reqs, DEV)
    for a, b in zip(rep.outcomes, alone):
        assert a.rid == b.rid and a.result.keys() == b.result.keys()
        for k in b.result:
            x, y = a.result[k].detach().cpu().numpy(), b.result[k].detach().cpu().numpy()
            assert rel_l2(x, y) <= 1e-6, (a.rid, k, rel_l2(x, y))
    s = rep.summary()
    assert s["failed"] == 0 and s["latency_p99_s"] >= s["latency_p50_s"] > 0


def test_footprint_bounds_device_allocation():
    from adfwi_amd.serving import execute, footprint_bytes
    for r in _small_stream(6, seed=11):
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        out, _ = execute(r, DEV)
        torch.cuda.synchronize()
        peak = torch.cuda.max_memory_allocated() - base
        del out
        assert peak <= footprint_bytes(r), (r.kind, r.mode, peak, footprint_bytes(r))


def test_packed_requests_equal_their_results_alone():
    """Requests against one shared model are packed into shared launches (modelling of both
    kinds, acoustic gradients with members of odd shot counts, which the pack pads with silent
    shots); every member's records are bit-identical to its run alone, its gradient and loss
    equal within 1e-6 (the gradient planes of a pack may group the shots differently)."""
    from adfwi_amd.serving import Request, ServingLoop, run_alone
    spec = [("acoustic", "forward", 1), ("acoustic", "gradient", 3), ("acoustic", "forward", 2),
            ("elastic", "forward", 2), ("acoustic", "gradient", 1), ("acoustic", "gradient", 2),
            ("elastic", "forward", 1), ("acoustic", "forward", 3), ("elastic", "gradient", 1)]
    reqs = []
    for i, (kind, mode, ns) in enumerate(spec):
        model = 101 if kind == "acoustic" else 202
        reqs.append(Request(i, kind, mode, 128, 288, 90, ns, 1000 + i, 0.0, 20, keep=True, model=model))
    rep = ServingLoop(DEV, max_slots=2).run(copy.deepcopy(reqs), realtime=False)
    assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
    sizes = {tuple(p) for p in rep.packs}
    assert (0, 2, 7) in sizes and (1, 4, 5) in sizes and (3, 6) in sizes and (8,) in sizes, sizes
    alone = run_alone(reqs, DEV)
    for a, b in zip(rep.outcomes, alone):
        assert a.rid == b.rid and a.result.keys() == b.result.keys(), (a.rid, a.result.keys(), b.result.keys())
        for k in b.result:
            x, y = a.result[k].detach().cpu().numpy(), b.result[k].detach().cpu().numpy()
            if reqs[a.rid].mode == "forward":
                assert np.array_equal(x, y), (a.rid, k, rel_l2(x, y))
            else:
                assert rel_l2(x, y) <= 1e-6, (a.rid, k, rel_l2(x, y))


def test_requests_of_different_models_share_launches():
    """Acoustic modelling requests of one grid and nt but each with its own model are packed
    into shared launches: every member's model is one coefficient set, indexed per shot (members
    padded to the coefficient group with silent shots).  Every member's records are
    bit-identical to its run alone; elastic modelling requests with their own models stay alone.
    Acoustic gradient requests of different models pack the same way when the loop packs them
    across models: every member's gradient and loss equal its run alone within 1e-6 (the pack's
    gradient planes group shots differently)."""
    from adfwi_amd.serving import Request, ServingLoop, run_alone
    spec = [("acoustic", 1), ("acoustic", 3), ("acoustic", 2), ("elastic", 2), ("acoustic", 5), ("elastic", 1)]
    reqs = [Request(i, kind, "forward", 128, 288, 90, ns, 2000 + i, 0.0, 20, keep=True)
            for i, (kind, ns) in enumerate(spec)]
    rep = ServingLoop(DEV, max_slots=2, graph_steps=0).run(copy.deepcopy(reqs), realtime=False)
    assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
    sizes = {tuple(p) for p in rep.packs}
    assert (0, 1, 2, 4) in sizes and (3,) in sizes and (5,) in sizes, sizes
    alone = run_alone(reqs, DEV)
    for a, b in zip(rep.outcomes, alone):
        assert a.rid == b.rid and a.result.keys() == b.result.keys()
        for k in b.result:
            x, y = a.result[k].detach().cpu().numpy(), b.result[k].detach().cpu().numpy()
            assert np.array_equal(x, y), (a.rid, k, rel_l2(x, y))

    spec = [("acoustic", 3), ("acoustic", 1), ("elastic", 1), ("acoustic", 2)]
    reqs = [Request(i, kind, "gradient", 96, 224, 80, ns, 3000 + i, 0.0, 20, keep=True)
            for i, (kind, ns) in enumerate(spec)]
    rep = ServingLoop(DEV, max_slots=2, graph_steps=0, pack_across_models=True).run(copy.deepcopy(reqs),
                                                                                  realtime=False)
    assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
    sizes = {tuple(p) for p in rep.packs}
    assert (0, 1, 3) in sizes and (2,) in sizes, sizes
    alone = run_alone(reqs, DEV)
    for a, b in zip(rep.outcomes, alone):
        assert a.rid == b.rid and a.result.keys() == b.result.keys()
        for k in b.result:
            x, y = a.result[k].detach().cpu().numpy(), b.result[k].detach().cpu().numpy()
            assert np.abs(y).max() > 0 and rel_l2(x, y) <= 1e-6, (a.rid, k, rel_l2(x, y))


def test_large_requests_run_in_shot_chunks_within_their_footprint(monkeypatch):
    """With the caps lowered, gradient requests run in shot chunks (their parameters' gradients
    summed over the chunks) and acoustic ones recompute their segments instead of keeping every
    stencil image: gradients and losses equal the single-launch resident runs within 1e-6, and
    the peak allocation stays within the (chunk's) footprint."""
    from adfwi_amd.serving import Request, execute, footprint_bytes
    from adfwi_amd.serving import requests as rq
    reqs = [Request(0, "acoustic", "gradient", 128, 288, 120, 7, 41, keep=True),
            Request(1, "elastic", "gradient", 128, 256, 100, 5, 42, keep=True)]
    whole = [execute(r, DEV)[0] for r in reqs]
    fp_whole = [footprint_bytes(r) for r in reqs]
    monkeypatch.setattr(rq, "RESIDENT_IMAGE_CAP_BYTES", 0)
    monkeypatch.setattr(rq, "REQUEST_CAP_BYTES", min(fp_whole) // 2)
    for r, w in zip(reqs, whole):
        n = rq.shot_chunk(r)
        assert 1 <= n < r.nshots and not rq.images_resident(r), (r.kind, n)
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        out, _ = execute(r, DEV)
        torch.cuda.synchronize()
        peak = torch.cuda.max_memory_allocated() - base
        assert peak <= footprint_bytes(r), (r.kind, peak, footprint_bytes(r))
        assert out.keys() == w.keys()
        for k in w:
            x, y = out[k].detach().cpu().numpy(), w[k].detach().cpu().numpy()
            assert np.abs(y).max() > 0 and rel_l2(x, y) <= 1e-6, (r.kind, k, rel_l2(x, y))


def test_exclusive_requests_run_alone_on_the_device():
    """With exclusive_s = 0 every request is charged the whole budget: the loop runs them one
    at a time, and every result equals its run alone."""
    from adfwi_amd.serving import ServingLoop, run_alone
    reqs = _small_stream(5, seed=13)
    rep = ServingLoop(DEV, max_slots=3, exclusive_s=0.0).run(copy.deepcopy(reqs), realtime=False)
    assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
    assert rep.max_concurrent == 1 and not rep.rejected
    alone = run_alone(reqs, DEV)
    for a, b in zip(rep.outcomes, alone):
        assert a.rid == b.rid and a.result.keys() == b.result.keys()
        for k in b.result:
            x, y = a.result[k].detach().cpu().numpy(), b.result[k].detach().cpu().numpy()
            assert rel_l2(x, y) <= 1e-6, (a.rid, k, rel_l2(x, y))


def test_packed_gradients_with_four_shots_per_wavefront_and_odd_segments(monkeypatch):
    """Four shots per forward wavefront and two per two-step adjoint wavefront, with nt 80 so
    that every checkpoint segment (k = 9) has an odd length and its first reversed step runs
    the one-step adjoint kernel: both adjoint kernels must add a shot's image terms into the
    plane of its gradient group, so every pack member's gradient equals its run alone (members
    of 1-3 shots, padded to the group, against one shared model and against models of their
    own)."""
    from adfwi_amd.serving import Request, ServingLoop, run_alone
    monkeypatch.setenv("ADFWI_SHOTS_PER_BLOCK", "4")
    for across, base in ((False, 4000), (True, 4100)):
        spec = [3, 1, 2, 3]
        reqs = [Request(i, "acoustic", "gradient", 96, 224, 80, ns, base + i, 0.0, 20, keep=True,
                        model=None if across else 303) for i, ns in enumerate(spec)]
        rep = ServingLoop(DEV, max_slots=1, graph_steps=0, pack_across_models=across).run(copy.deepcopy(reqs),
                                                                                        realtime=False)
        assert all(o.error is None for o in rep.outcomes), [o.error for o in rep.outcomes]
        assert any(len(p) >= 3 for p in rep.packs), rep.packs
        alone = run_alone(reqs, DEV)
        for a, b in zip(rep.outcomes, alone):
            assert a.rid == b.rid and a.result.keys() == b.result.keys()
            for k in b.result:
                x, y = a.result[k].detach().cpu().numpy(), b.result[k].detach().cpu().numpy()
                assert np.abs(y).max() > 0 and rel_l2(x, y) <= 1e-6, (across, a.rid, k, rel_l2(x, y))
