"""HIP-graph replay of the acoustic forward (modelling without gradient): records and energy
maps bit-identical to direct launches on every replay, with other work and synchronisations
between the replays, for two models of one shape sharing the cached graphs."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from helpers import acoustic_inputs

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _problem(name, vp_scale=1.0, src_shift=0):
    from adfwi_amd.propagator.acoustic_kernels import AcousticGeometry, coefficient_fields
    g, m = acoustic_inputs(name)
    nabc = m["nabc"]
    geom = AcousticGeometry(m["nz"] + 2 * nabc, m["nx"] + 2 * nabc, m["nt"], m["fs"], nabc,
                            g["src_z"] + nabc, g["src_x"] + nabc + src_shift, g["rcv_z"] + nabc,
                            g["rcv_x"] + nabc, DEV)
    vp = torch.tensor(g["vp"], device=DEV) * vp_scale
    a1, k1, a2, k2, k3 = coefficient_fields(vp, torch.tensor(g["rho"], device=DEV),
                                            torch.tensor(g["damp"], dtype=torch.float32, device=DEV),
                                            m["nz"], m["nx"], nabc, m["dt"], m["dx"], m["fs"], DEV)
    coef = geom.pack_coef(1.0 - k1, a1, 1.0 - k2, a2, 1.0 - k3)
    src = (m["dt"] * torch.tensor(g["wavelet"], device=DEV)).contiguous()
    return geom, coef, src, m


def _same(res, ref):
    for i in range(3):
        x, y = res["rec"][i].cpu().numpy(), ref["rec"][i].cpu().numpy()
        assert np.array_equal(x, y), (i, rel_l2(x, y))
    assert torch.equal(res["energy"], ref["energy"])


@pytest.mark.parametrize("name", ["acoustic_nofs", "acoustic_fs"])
@pytest.mark.parametrize("window", [16, 32])
def test_graph_replays_equal_direct_launches(name, window):
    from adfwi_amd.propagator import graphs
    from adfwi_amd.propagator.acoustic_kernels import run_forward
    graphs.clear_cache()
    geom, coef_a, src, m = _problem(name)
    _, coef_b, _, _ = _problem(name, vp_scale=1.07)
    eu = m["nt"] // 3                      # the u energy map starts inside the run
    assert m["nt"] % window and m["nt"] // window >= 2   # full windows and a shorter last one
    ref_a = run_forward(geom, coef_a, src, energy_u_from=eu)
    ref_b = run_forward(geom, coef_b, src, energy_u_from=eu)
    for trial, (coef, ref) in enumerate([(coef_a, ref_a), (coef_b, ref_b), (coef_a, ref_a), (coef_b, ref_b)]):
        res = graphed_forward_checked(geom, coef, src, eu, window)
        _same(res, ref)
        run_forward(geom, coef_b if trial % 2 == 0 else coef_a, src, energy_u_from=eu)   # work between replays
        torch.cuda.synchronize()
    # new geometry objects with the same content replay the cached graphs, also with other
    # source positions (read from device memory); the first geometry keeps its own sources
    assert len(graphs._LOCAL.cache) == 1
    geom2, _, _, _ = _problem(name)
    _same(graphed_forward_checked(geom2, coef_a, src, eu, window), ref_a)
    geom3, _, _, _ = _problem(name, src_shift=3)
    ref_3 = run_forward(geom3, coef_a, src, energy_u_from=eu)
    assert not np.array_equal(ref_3["rec"][0].cpu().numpy(), ref_a["rec"][0].cpu().numpy())
    _same(graphed_forward_checked(geom3, coef_a, src, eu, window), ref_3)
    assert len(graphs._LOCAL.cache) == 1
    _same(run_forward(geom, coef_a, src, energy_u_from=eu), ref_a)
    graphs.clear_cache()


def graphed_forward_checked(geom, coef, src, eu, window):
    from adfwi_amd.propagator.graphs import graphed_forward
    res = graphed_forward(geom, coef, src, eu, window)
    torch.cuda.synchronize()
    return res


def test_propagator_forward_with_graphs_equals_direct():
    """The drop-in propagator with graph_steps set (modelling under no_grad) returns the same
    records and energy maps as with direct launches."""
    from adfwi_amd.model import AcousticModel
    from adfwi_amd.propagator import AcousticPropagator, graphs
    from adfwi_amd.survey import Receiver, Source, Survey
    rng = np.random.default_rng(3)
    nz, nx, nt = 40, 90, 150
    vp = (2000 + 800 * rng.random((nz, nx))).astype(np.float32)
    src = Source(nt, 1e-3, 15.0)
    t = np.arange(nt) * 1e-3 - 0.08
    src.add_sources([10, 45, 80], [2, 2, 2], (1 - 2 * (np.pi * 15 * t) ** 2) * np.exp(-(np.pi * 15 * t) ** 2))
    rcv = Receiver(nt, 1e-3)
    rcv.add_receivers(np.arange(nx), np.full(nx, 3), "pr")
    outs = []
    for steps in (0, 32, 32, 32):   # direct, then the first graph run (direct), capture, replay
        model = AcousticModel(0, 0, nx, nz, 10.0, 10.0, vp=vp, rho=np.full((nz, nx), 2000.0, np.float32),
                              vp_grad=False, nabc=10, free_surface=False, device=DEV)
        prop = AcousticPropagator(model, Survey(src, rcv), device=DEV)
        prop.graph_steps = steps
        with torch.no_grad():
            outs.append(prop.forward())
        torch.cuda.synchronize()
    for o in outs[1:]:
        for k in ("p", "u", "w", "forward_wavefield_p", "forward_wavefield_u", "forward_wavefield_w"):
            assert torch.equal(o[k], outs[0][k]), k
    graphs.clear_cache()
