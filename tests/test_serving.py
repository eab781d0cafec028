"""Host-side tests of the serving loop: the request stream, the footprint model and the
admission policy (no device needed)."""
import numpy as np
import pytest

from adfwi_amd.serving import Request, Scheduler, footprint_bytes, make_stream, pack_key


def _req(rid, nbytes_tag=0, arrival=0.0):
    r = Request(rid, "acoustic", "forward", 128, 256, 100, 1, rid, arrival)
    r.extra["tag"] = nbytes_tag
    return r


def test_stream_is_seeded_and_follows_the_mix():
    a = make_stream(2000, 2.0, seed=4)
    b = make_stream(2000, 2.0, seed=4)
    assert [(r.kind, r.mode, r.nz, r.nx, r.nt, r.nshots, r.seed) for r in a] == \
           [(r.kind, r.mode, r.nz, r.nx, r.nt, r.nshots, r.seed) for r in b]
    kinds = np.mean([r.kind == "acoustic" for r in a])
    grads = np.mean([r.mode == "gradient" for r in a])
    assert abs(kinds - 0.6) < 0.04 and abs(grads - 0.5) < 0.04
    assert all(r.nz % 32 == 0 and 128 <= r.nz <= 512 for r in a)
    assert all(r.nx % 32 == 0 and 256 <= r.nx <= 2048 for r in a)
    assert all(1000 <= r.nt <= 6000 and 1 <= r.nshots <= 64 for r in a)
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 0.5) < 0.05          # Poisson process at 2 requests/s


def test_footprint_grows_with_every_dimension():
    base = Request(0, "acoustic", "forward", 128, 256, 1000, 4, 0)
    f0 = footprint_bytes(base)
    for kw in ({"nz": 256}, {"nx": 512}, {"nt": 4000}, {"nshots": 8}, {"mode": "gradient"}, {"kind": "elastic"}):
        r = Request(**{**base.__dict__, **kw})
        assert footprint_bytes(r) > f0, kw
    # gradient: checkpoints grow with sqrt(nt), records with nt
    g1 = Request(0, "elastic", "gradient", 256, 512, 1000, 8, 0)
    g4 = Request(0, "elastic", "gradient", 256, 512, 4000, 8, 0)
    assert footprint_bytes(g4) < 4 * footprint_bytes(g1)


def test_admission_respects_budget_and_slots():
    s = Scheduler(budget_bytes=100, max_slots=2)
    sizes = {0: 60, 1: 30, 2: 30, 3: 10}
    for rid, n in sizes.items():
        s.submit(_req(rid), n)
    got = [r.rid for r in s.admit()]
    assert got == [0, 1]                     # 2 and 3 fit the budget but no slot is left
    assert s.used == 90
    s.finish(0)
    got = [r.rid for r in s.admit()]
    assert got == [2]                        # 2 fits in 70; one slot free
    s.finish(1)
    assert [r.rid for r in s.admit()] == [3]
    s.finish(2)
    s.finish(3)
    assert s.idle()


def test_large_request_is_not_starved():
    s = Scheduler(budget_bytes=100, max_slots=8, max_bypass=2)
    s.submit(_req(0), 50)
    assert [r.rid for r in s.admit()] == [0]
    s.submit(_req(1), 85)                    # does not fit while 0 runs
    for rid in (2, 3, 4, 5):
        s.submit(_req(rid), 10)
    got = [r.rid for r in s.admit()]
    assert got == [2, 3]                     # 1 is passed twice, then blocks the queue
    s.finish(0)
    assert [r.rid for r in s.admit()] == []  # 85 does not fit beside 2 and 3 (20 used); 4, 5 wait behind it
    s.finish(2)
    s.finish(3)
    assert [r.rid for r in s.admit()][0] == 1


def test_oversized_request_is_rejected():
    s = Scheduler(budget_bytes=100, max_slots=2)
    assert not s.submit(_req(0), 101)
    assert s.rejected and "exceeds" in s.rejected[0][1]
    assert s.idle()


def test_full_stream_admission_never_exceeds_budget():
    reqs = make_stream(300, 5.0, seed=1)
    budget = 200 * 2**30
    s = Scheduler(budget, max_slots=6)
    rng = np.random.default_rng(0)
    for r in reqs:
        s.submit(r, footprint_bytes(r))
    done = 0
    while not s.idle():
        for r in s.admit():
            pass
        assert s.used <= budget and len(s.running) <= 6
        # finish a random running request
        if s.running:
            rid = list(s.running)[rng.integers(len(s.running))]
            s.finish(rid)
            done += 1
    assert done + len(s.rejected) == len(reqs)


def test_model_pool_stream_and_pack_keys():
    reqs = make_stream(400, 5.0, seed=9, model_pool=3)
    models = {r.model for r in reqs}
    assert len(models) == 3 and None not in models
    for m in models:   # kind, grid and nt are the model's
        assert len({(r.kind, r.nz, r.nx, r.nt) for r in reqs if r.model == m}) == 1
    for r in reqs:
        k = pack_key(r)
        if r.mode == "gradient" and r.kind == "elastic":
            assert k is None
        elif r.kind == "acoustic" and r.mode == "forward":    # any model: one coefficient set each
            assert k == (r.kind, r.mode, r.nz, r.nx, r.nt, r.nabc)
        else:
            assert k == (r.kind, r.mode, r.nz, r.nx, r.nt, r.nabc, r.model)
        if r.kind == "acoustic":  # gradients too when packed across models
            assert pack_key(r, True) == (r.kind, r.mode, r.nz, r.nx, r.nt, r.nabc)
    # own models: acoustic modelling requests pack by shape, gradients only across models
    for r in make_stream(50, 5.0, seed=9):
        assert (pack_key(r) is None) == (r.kind != "acoustic" or r.mode != "forward")
        assert (pack_key(r, True) is None) == (r.kind != "acoustic")


def test_full_size_gradient_requests_fit_one_gpu():
    """Acoustic gradient requests keep all their stencil images only under the cap, and gradient
    requests whose footprint exceeds the request cap run in shot chunks: the largest requests of
    the full-size mix (nz 512, nx 2048, nt 6000, 64 shots) fall back to checkpoints and
    recompute, elastic gradients of that size run in chunks of 16 shots (one launch of all 64
    would need 489 GiB), and every request's footprint fits the 288 GB of one MI355X."""
    from adfwi_amd.serving.requests import (REQUEST_CAP_BYTES, RESIDENT_IMAGE_CAP_BYTES, _footprint,
                                            images_resident, shot_chunk)
    small = Request(0, "acoustic", "gradient", 128, 256, 200, 3, 1)
    big = Request(1, "acoustic", "gradient", 512, 2048, 6000, 64, 2)
    assert images_resident(small) and not images_resident(big)
    Z, X = big.padded
    assert big.nt * 64 * Z * X * 4 > RESIDENT_IMAGE_CAP_BYTES
    assert shot_chunk(small) == 3 and pack_key(small, True) is not None
    el = Request(2, "elastic", "gradient", 512, 2048, 6000, 64, 3)
    assert _footprint(el, 64) > 400 * 2**30 and shot_chunk(el) == 16 and pack_key(el, True) is None
    for kind in ("acoustic", "elastic"):
        for mode in ("forward", "gradient"):
            r = Request(2, kind, mode, 512, 2048, 6000, 64, 3)
            assert footprint_bytes(r) <= max(REQUEST_CAP_BYTES, _footprint(r, 1)), (kind, mode)
            assert (shot_chunk(r) < 64) == (mode == "gradient"), (kind, mode)
    for r in make_stream(300, 1.0, seed=4):
        assert footprint_bytes(r) <= REQUEST_CAP_BYTES, r


def test_shape_pool_stream():
    reqs = make_stream(300, 5.0, seed=11, shape_pool=4)
    assert all(r.model is None for r in reqs)
    shapes = {(r.kind, r.nz, r.nx, r.nt) for r in reqs}
    assert len(shapes) <= 4
    seeds = {r.model_seed for r in reqs}
    assert len(seeds) == len(reqs)                           # every request its own model
    for mode in ("forward", "gradient"):
        ac = [r for r in reqs if r.kind == "acoustic" and r.mode == mode]
        assert len({pack_key(r, True) for r in ac}) <= 4


def test_packs_share_a_slot_and_respect_shots_and_budget():
    def req(rid, key, shots, nbytes):
        r = Request(rid, "acoustic", "forward", 128, 256, 100, shots, rid, 0.0)
        r.extra.update(key=key, nbytes=nbytes)
        return r
    s = Scheduler(budget_bytes=100, max_slots=2, pack_key=lambda r: r.extra["key"], max_pack_shots=8)
    reqs = [req(0, "a", 4, 10), req(1, "b", 2, 10), req(2, "a", 3, 10), req(3, "a", 2, 10),
            req(4, None, 1, 10), req(5, "a", 2, 75)]
    for r in reqs:
        assert s.submit(r, r.extra["nbytes"])
    packs = s.admit_packs()
    # 0 takes 2 along (4 + 3 shots); 3 and 5 would exceed 8 shots
    assert [[r.rid for r in p] for p in packs] == [[0, 2], [1]]
    assert s.slots_used == 2 and s.used == 30
    assert [w.req.rid for w in s.waiting] == [3, 4, 5]
    for rid in (0, 2):
        s.finish(rid)
    assert s.slots_used == 1 and s.used == 10
    packs = s.admit_packs()   # 3 leads a new pack with 5 (2 + 2 shots, 85 B fits in 90 B); no slot for 4
    assert [[r.rid for r in p] for p in packs] == [[3, 5]]
    assert s.used == 95 and [w.req.rid for w in s.waiting] == [4]


def test_exclusive_requests_run_alone():
    """A request charged the whole budget (the loop's exclusive policy for long requests) starts
    only when nothing runs and blocks admission while it runs; smaller ones pass it at most
    max_bypass times."""
    from adfwi_amd.serving.requests import service_estimate_s
    s = Scheduler(budget_bytes=100, max_slots=4, max_bypass=2)
    s.submit(_req(0), 10)
    assert [r.rid for r in s.admit()] == [0]
    s.submit(_req(1), 100)                   # exclusive
    for rid in (2, 3, 4):
        s.submit(_req(rid), 10)
    assert [r.rid for r in s.admit()] == [2, 3]   # 1 passed twice, then it blocks the queue
    s.finish(0)
    s.finish(2)
    assert s.admit() == []
    s.finish(3)
    assert [r.rid for r in s.admit()] == [1]
    assert s.admit() == [] and s.used == 100
    s.finish(1)
    assert [r.rid for r in s.admit()] == [4]
    small = Request(0, "acoustic", "forward", 128, 256, 600, 4, 0)
    big = Request(1, "elastic", "gradient", 512, 2048, 6000, 64, 0)
    assert service_estimate_s(small) < 0.01 < 10 < service_estimate_s(big)


def test_pack_members_taken_past_a_waiting_request_count_as_bypasses():
    """A pack lead takes later members of its key past an earlier waiting request that does not
    fit: each such pack passes that request once, and once it has been passed max_bypass times
    no member is taken past it any more, so packing cannot starve it."""
    def req(rid, key, nbytes):
        r = Request(rid, "acoustic", "forward", 128, 256, 100, 1, rid, 0.0)
        r.extra.update(key=key, nbytes=nbytes)
        return r
    s = Scheduler(budget_bytes=100, max_slots=8, max_bypass=2, pack_key=lambda r: r.extra["key"])
    s.submit(req(0, None, 40), 40)
    assert [r.rid for r in s.admit()] == [0]
    # 1 leads a pack; 2 (70 B) does not fit beside 0 and 1; 3 and 4 share 1's key
    for r in (req(1, "a", 10), req(2, None, 70), req(3, "a", 10)):
        s.submit(r, r.extra["nbytes"])
    assert [[r.rid for r in p] for p in s.admit_packs()] == [[1, 3]]
    assert [(w.req.rid, w.bypassed) for w in s.waiting] == [(2, 1)]
    for r in (req(4, "a", 5), req(5, "a", 5)):
        s.submit(r, r.extra["nbytes"])
    # 4 leads from behind 2 and takes 5 along: the pack passes 2 once more (its second time)
    assert [[r.rid for r in p] for p in s.admit_packs()] == [[4, 5]]
    assert [(w.req.rid, w.bypassed) for w in s.waiting] == [(2, 2)]
    s.submit(req(6, "a", 5), 5)
    assert s.admit_packs() == []          # 2 blocks the queue until it fits
    for rid in (0, 1, 3, 4, 5):
        s.finish(rid)
    assert [[r.rid for r in p] for p in s.admit_packs()] == [[2], [6]]

    # a lead waiting ahead of a request that has reached max_bypass takes no member past it
    s = Scheduler(budget_bytes=100, max_slots=4, max_bypass=1, pack_key=lambda r: r.extra["key"])
    s.submit(req(0, None, 40), 40)
    assert [r.rid for r in s.admit()] == [0]
    for r in (req(1, "a", 65), req(2, None, 70), req(3, None, 10)):
        s.submit(r, r.extra["nbytes"])
    assert [r.rid for r in s.admit()] == [3]            # passes 1 and 2 once each
    s.submit(req(4, "a", 10), 10)
    s.finish(0)
    assert [[r.rid for r in p] for p in s.admit_packs()] == [[1]]   # 4 stays behind 2
    assert [(w.req.rid, w.bypassed) for w in s.waiting] == [(2, 1), (4, 0)]


def test_weighted_slots_limit_long_requests_and_do_not_starve_them():
    """A request submitted with weight 2 takes two slots: with four slots at most two such
    requests run together, short ones fill the rest; a weight-2 request passed max_bypass
    times by short ones blocks the queue until two slots are free."""
    s = Scheduler(budget_bytes=100, max_slots=4, max_bypass=2)
    for rid in (0, 1, 2):
        s.submit(_req(rid), 10, weight=2)
    s.submit(_req(3), 10)
    assert [r.rid for r in s.admit()] == [0, 1]          # two long requests fill the four slots
    assert s.slots_used == 4
    s.finish(0)
    assert [r.rid for r in s.admit()] == [2]             # the next long one takes the two slots freed
    s.finish(1)
    assert [r.rid for r in s.admit()] == [3]             # a short one takes one of the two freed
    assert s.slots_used == 3
    s.submit(_req(4), 10, weight=2)                      # one slot free: it must wait
    for rid in (5, 6, 7):
        s.submit(_req(rid), 10)
    assert [r.rid for r in s.admit()] == [5]             # 5 passes 4 (once) and takes the last slot
    s.finish(5)
    assert [r.rid for r in s.admit()] == [6]             # passed twice: 4 now blocks the queue
    s.finish(6)
    assert [r.rid for r in s.admit()] == []              # 7 may not pass it; one slot is not enough
    s.finish(3)
    assert [r.rid for r in s.admit()] == [4]             # two slots free
    with pytest.raises(ValueError):
        s.submit(_req(8), 10, weight=5)
