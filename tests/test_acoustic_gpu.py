"""Parity of the HIP acoustic path (libadfwi_amd.so through the C ABI) with the oracle and
with the reference's own vectors.  Tolerances: forward records rel-L2 <= 1e-5 (bit-exact when
the kernel is fed the oracle's coefficient fields), gradients rel-L2 <= 1e-4."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from helpers import acoustic_inputs, run_acoustic_oracle

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _geom_from_golden(g, m, **kw):
    from adfwi_amd.propagator.acoustic_kernels import AcousticGeometry
    nabc = m["nabc"]
    return AcousticGeometry(m["nz"] + 2 * nabc, m["nx"] + 2 * nabc, m["nt"], m["fs"], nabc,
                            g["src_z"] + nabc, g["src_x"] + nabc, g["rcv_z"] + nabc, g["rcv_x"] + nabc, DEV, **kw)


@pytest.mark.parametrize("name", ["acoustic_nofs", "acoustic_fs"])
@pytest.mark.parametrize("spb", [0, 1])
def test_kernel_bit_exact_given_oracle_coefficients(name, spb):
    """Same coefficient bits in -> same record bits out: the fused step keeps the reference's
    fp32 association (no contraction), stage order, windows and free-surface rows."""
    from adfwi_amd.propagator.acoustic_kernels import run_forward
    from oracle import acoustic as oa
    g, m = acoustic_inputs(name)
    omk1, a1, omk2, a2, omk3 = oa.coefficients(torch.tensor(g["vp"]), torch.tensor(g["rho"]),
                                               torch.tensor(g["damp"], dtype=torch.float32), m["nz"], m["nx"],
                                               m["nabc"], m["dt"], m["dx"], m["fs"])
    src = m["dt"] * torch.tensor(g["wavelet"])
    geom = _geom_from_golden(g, m, shots_per_block=spb)
    coef = geom.pack_coef(*[t.to(DEV) for t in (omk1, a1, omk2, a2, omk3)])
    res = run_forward(geom, coef, src.to(DEV), energy_u_from=0)
    torch.cuda.synchronize()
    for i, k in enumerate("puw"):
        got = res["rec"][i].cpu().numpy()
        assert np.array_equal(got, g[f"seg1_rec_{k}"]), (k, rel_l2(got, g[f"seg1_rec_{k}"]))
    nz, nx, nabc = m["nz"], m["nx"], m["nabc"]
    ep = res["energy"][0].cpu().numpy()[nabc:nabc + nz, nabc:nabc + nx]
    assert rel_l2(ep, g["seg1_fw_p"]) < 1e-6


@pytest.mark.parametrize("name", ["acoustic_nofs", "acoustic_fs"])
@pytest.mark.parametrize("seg", [1, 3])
def test_forward_kernel_dropin_matches_reference_vectors(name, seg):
    from adfwi_amd.propagator import acoustic_kernels as ak
    g, m = acoustic_inputs(name)
    v = torch.tensor(g["vp"], device=DEV, requires_grad=True)
    r = torch.tensor(g["rho"], device=DEV, requires_grad=True)
    w = torch.tensor(g["wavelet"], device=DEV, requires_grad=True)
    res = ak.forward_kernel(m["nx"], m["nz"], m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                            torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), 2, w,
                            torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]), len(g["rcv_x"]),
                            torch.tensor(g["damp"], dtype=torch.float32, device=DEV), v, r,
                            checkpoint_segments=seg, device=DEV, ckpt_every=7)
    loss = sum((res[k] * torch.tensor(g["weights"][i], device=DEV)).sum() for i, k in enumerate("puw"))
    loss.backward()
    f64 = run_acoustic_oracle(g, m, seg, dtype=torch.float64, want_grad=True)
    for k in "puw":
        got = res[k].detach().cpu().numpy()
        e32 = rel_l2(got, g[f"seg{seg}_rec_{k}"])
        print(f"{name} seg{seg} rec_{k}: rel-L2 vs reference fp32 {e32:.2e}, vs fp64 oracle "
              f"{rel_l2(got, f64[k]):.2e} (reference fp32 vs fp64: {rel_l2(g[f'seg{seg}_rec_{k}'], f64[k]):.2e})")
        assert e32 <= 1e-5
        assert rel_l2(res["forward_wavefield_" + k].cpu().numpy(), g[f"seg{seg}_fw_{k}"]) <= 1e-5
    for k, t in (("vp", v), ("rho", r), ("wavelet", w)):
        got = t.grad.cpu().numpy()
        e = rel_l2(got, g[f"seg{seg}_grad_{k}"])
        print(f"{name} seg{seg} grad_{k}: rel-L2 vs reference autograd {e:.2e}, vs fp64 oracle "
              f"{rel_l2(got, f64['grad_' + k]):.2e}")
        assert e <= 1e-4, k


def test_adjoint_is_the_transpose_of_the_source_to_record_map():
    """The wavelet -> records map is linear, so <F s, d> = <s, F^T d> up to fp32 rounding.
    Exercises every transposed stage except the A1/A2 imaging."""
    from adfwi_amd.propagator.acoustic_kernels import AcousticWave
    from oracle import acoustic as oa
    g, m = acoustic_inputs("acoustic_fs")
    rng = np.random.default_rng(9)
    omk1, a1, omk2, a2, omk3 = [t.to(DEV) for t in oa.coefficients(
        torch.tensor(g["vp"]), torch.tensor(g["rho"]), torch.tensor(g["damp"], dtype=torch.float32),
        m["nz"], m["nx"], m["nabc"], m["dt"], m["dx"], m["fs"])]
    geom = _geom_from_golden(g, m)
    s = torch.tensor(rng.standard_normal((2, m["nt"])), dtype=torch.float32, device=DEV, requires_grad=True)
    dd = [torch.tensor(rng.standard_normal((2, m["nt"], len(g["rcv_x"]))), dtype=torch.float32, device=DEV)
          for _ in range(3)]
    out = AcousticWave.apply(a1, a2, s, omk1, omk2, omk3, geom, 0, 5)
    lhs = sum((out[i].double() * dd[i].double()).sum() for i in range(3))
    torch.autograd.backward(out[:3], dd)
    rhs = (s.detach().double() * s.grad.double()).sum()
    assert abs(float(lhs - rhs)) <= 1e-5 * abs(float(lhs)), (float(lhs), float(rhs))


def test_shot_batch_equals_single_shots():
    from adfwi_amd.propagator import acoustic_kernels as ak
    g, m = acoustic_inputs("acoustic_nofs")
    args = lambda sl: (m["nx"], m["nz"], m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                       torch.tensor(g["src_x"][sl]), torch.tensor(g["src_z"][sl]), len(g["src_x"][sl]),
                       torch.tensor(g["wavelet"][sl], device=DEV), torch.tensor(g["rcv_x"]),
                       torch.tensor(g["rcv_z"]), len(g["rcv_x"]),
                       torch.tensor(g["damp"], dtype=torch.float32, device=DEV),
                       torch.tensor(g["vp"], device=DEV), torch.tensor(g["rho"], device=DEV))
    both = ak.forward_kernel(*args(slice(0, 2)), device=DEV)
    for s in range(2):
        one = ak.forward_kernel(*args(slice(s, s + 1)), device=DEV)
        for k in "puw":
            assert torch.equal(one[k][0], both[k][s])


@pytest.mark.parametrize("free_surface", [False, True])
def test_multi_tile_grid_sources_and_receivers_on_tile_edges(free_surface):
    """A grid spanning several tiles in x and z and several shot groups; sources and receivers
    sit on tile boundaries; compared with the oracle, gradient included."""
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.propagator.boundary_condition import bc_pml
    from oracle import acoustic as oa
    rng = np.random.default_rng(17)
    nz, nx, nabc, nt, dt, dx = 45, 150, 10, 150, 1e-3, 10.0
    vp = rng.uniform(1500, 3200, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2400, (nz, nx)).astype(np.float32)
    damp = torch.tensor(bc_pml(nx, nz, dx, dx, nabc, float(vp.max()), False), dtype=torch.float32)
    src_x = np.array([0, 53, 54, 117, 118, 149, 63]); src_z = np.array([0, 3, 13, 14, 21, 44, 1])
    rcv_x = np.array([0, 53, 54, 117, 118, 149, 64, 64]); rcv_z = np.array([0, 1, 5, 14, 22, 44, 6, 6])
    S = len(src_x)
    wav = np.stack([np.sin(np.arange(nt) * 0.07 * (1 + 0.1 * s)) * np.exp(-((np.arange(nt) - 30) / 15.0) ** 2)
                    for s in range(S)]).astype(np.float32)
    wts = rng.standard_normal((3, S, nt, len(rcv_x))).astype(np.float32)

    def run(mod, dev):
        v = torch.tensor(vp, device=dev, requires_grad=True)
        r = torch.tensor(rho, device=dev, requires_grad=True)
        w = torch.tensor(wav, device=dev, requires_grad=True)
        kw = dict(device=dev) if dev != "cpu" else {}
        res = mod.forward_kernel(nx, nz, dx, dx, nt, dt, nabc, free_surface, torch.tensor(src_x),
                                 torch.tensor(src_z), S, w, torch.tensor(rcv_x), torch.tensor(rcv_z), len(rcv_x),
                                 damp.to(dev), v, r, checkpoint_segments=2, **kw)
        loss = sum((res[k] * torch.tensor(wts[i], device=dev)).sum() for i, k in enumerate("puw"))
        loss.backward()
        return ({k: res[k].detach().cpu().numpy() for k in res},
                {"vp": v.grad.cpu().numpy(), "rho": r.grad.cpu().numpy(), "wav": w.grad.cpu().numpy()})
    rec_h, grad_h = run(ak, DEV)
    rec_o, grad_o = run(oa, "cpu")
    for k in rec_o:
        assert rel_l2(rec_h[k], rec_o[k]) <= 1e-5, (k, rel_l2(rec_h[k], rec_o[k]))
    for k in grad_o:
        assert rel_l2(grad_h[k], grad_o[k]) <= 1e-4, (k, rel_l2(grad_h[k], grad_o[k]))


def test_propagator_class_end_to_end():
    """AcousticModel + Survey + AcousticPropagator drive the HIP path like the reference classes;
    vp.grad through model.forward() (Gardner rho from detached vp) matches the oracle."""
    from adfwi_amd.model import AcousticModel
    from adfwi_amd.propagator import AcousticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    from adfwi_amd.utils import wavelet
    from oracle import acoustic as oa
    nz, nx, nabc, nt, dt = 30, 50, 10, 200, 1e-3
    rng = np.random.default_rng(2)
    vp = (2000 + 500 * rng.random((nz, nx))).astype(np.float32)
    model = AcousticModel(0, 0, nx, nz, 10.0, 10.0, vp=vp, rho=np.ones_like(vp), vp_grad=True, nabc=nabc,
                          free_surface=True, device=DEV)
    src = Source(nt, dt, 15.0)
    _, wv = wavelet(nt, dt, 15.0)
    for x in (10, 25, 40):
        src.add_source(x, 1, wv)
    rcv = Receiver(nt, dt)
    rcv.add_receivers(np.arange(nx), np.ones(nx, dtype=int), "pr")
    F = AcousticPropagator(model, Survey(src, rcv), device=DEV)
    out = F.forward(shot_index=np.array([0, 2]))
    out["p"].square().sum().backward()
    g_h = model.vp.grad.cpu().numpy().copy()
    v = torch.tensor(model.vp.detach().cpu().numpy(), requires_grad=True)
    rho = torch.tensor(model.rho.detach().cpu().numpy())
    ref = oa.forward_kernel(nx, nz, 10.0, 10.0, nt, dt, nabc, True, torch.tensor([10, 40]), torch.tensor([1, 1]), 2,
                            F.wavelet[[0, 2]].cpu(), torch.arange(nx), torch.ones(nx, dtype=torch.long), nx,
                            F.damp.cpu(), v, rho)
    ref["p"].square().sum().backward()
    assert rel_l2(out["p"].detach().cpu().numpy(), ref["p"].detach().numpy()) <= 1e-5
    assert rel_l2(g_h, v.grad.numpy()) <= 1e-4
    assert F.src_n == 3 and out["forward_wavefield_p"].shape == (nz, nx)


@pytest.mark.parametrize("free_surface", [False, True])
@pytest.mark.parametrize("nt,spb,tz", [(121, 4, 0), (150, 2, 8), (97, 1, 4)])
@pytest.mark.parametrize("needs", ["all", "a1_p", "a1_all"])
def test_two_steps_per_launch_equal_one_step_per_launch(monkeypatch, free_surface, nt, spb, tz, needs):
    """The fused two-step kernels (intermediate state in registers, the receivers of both steps
    written / their receiver sources added in the kernel) reproduce the one-step kernels (separate
    gather and inject launches): records and the wavelet gradient bit for bit; energy maps and the
    A1/A2 gradients to rounding (two steps' terms are added before one read-modify-write of the
    accumulator planes).  needs: the gradients asked for, which pick the adjoint variant: "all"
    (A1, A2, wavelet, the three records' adjoints: the variant that leaves the first reversed
    step's record adjoints to a separate launch), "a1_p" (A1 from the pressure records only: the
    lean variant), "a1_all" (A1 and the wavelet from all three records).  Both forms of the
    two-step kernels are checked: stores issued on every lane with the accesses of lanes that own
    no point dropped (the default below 2 GiB per two fields), and stores in branches
    (ADFWI_PAIR_DROP=0, the form for larger fields); the 7 shots leave partial shot groups."""
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.propagator.boundary_condition import bc_pml
    rng = np.random.default_rng(nt + spb)
    nz, nx, nabc, dt, dx = 70, 170, 10, 1e-3, 10.0
    vp = rng.uniform(1500, 3200, (nz, nx)).astype(np.float32)
    rho = rng.uniform(1800, 2400, (nz, nx)).astype(np.float32)
    damp = torch.tensor(bc_pml(nx, nz, dx, dx, nabc, float(vp.max()), False), dtype=torch.float32, device=DEV)
    S = 7
    src_x = np.linspace(1, nx - 2, S).astype(int) + nabc
    src_z = (np.arange(S) * 9 % (nz - 2) + (nabc if free_surface else 1)) + (0 if free_surface else nabc - 1)
    rcv_x = np.arange(0, nx, 3) + nabc
    rcv_z = np.full(len(rcv_x), nabc + 2)
    rcv_z[::5] = nabc + 30                                  # a second receiver row
    wav = torch.tensor(rng.standard_normal((S, nt)), dtype=torch.float32, device=DEV)
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    a1, k1, a2, k2, k3 = ak.coefficient_fields(torch.tensor(vp, device=DEV), torch.tensor(rho, device=DEV), damp,
                                               nz, nx, nabc, dt, dx, free_surface, DEV)
    outs = []
    for fuse, drop in (("0", "1"), ("1", "1"), ("1", "0")):
        monkeypatch.setenv("ADFWI_FUSE_STEPS", fuse)
        monkeypatch.setenv("ADFWI_PAIR_DROP", drop)
        if tz:
            monkeypatch.setenv("ADFWI_TILE_Z", str(tz))
        geom = ak.AcousticGeometry(Z, X, nt, free_surface, nabc, src_z, src_x, rcv_z, rcv_x, DEV, shots_per_block=spb)
        assert (geom.rcv_map is not None) == (fuse == "1")
        a1_ = a1.detach().clone().requires_grad_(True)
        a2_ = a2.detach().clone().requires_grad_(needs == "all")
        w_ = wav.clone().requires_grad_(needs != "a1_p")
        res = ak.AcousticWave.apply(a1_, a2_, dt * w_, 1.0 - k1, 1.0 - k2, 1.0 - k3, geom, nt // 3, 9)
        nrec = 1 if needs == "a1_p" else 3
        g = [torch.randn(r.shape, generator=torch.Generator().manual_seed(i)).to(DEV) for i, r in enumerate(res[:nrec])]
        torch.autograd.backward(res[:nrec], g)
        torch.cuda.synchronize()
        outs.append(([r.detach().cpu() for r in res], a1_.grad.cpu(), None if a2_.grad is None else a2_.grad.cpu(),
                     None if w_.grad is None else w_.grad.cpu()))
    (r0, ga0, gb0, gw0) = outs[0]
    for form, (r1, ga1, gb1, gw1) in zip(("dropped", "branches"), outs[1:]):
        for i in range(3):
            assert torch.equal(r0[i], r1[i]), (form, "record", i, rel_l2(r1[i].numpy(), r0[i].numpy()))
        for i in range(3, 6):
            assert rel_l2(r1[i].numpy(), r0[i].numpy()) <= 1e-6, (form, "energy", i)
        assert (gw0 is None) == (gw1 is None) and (gw0 is None or torch.equal(gw0, gw1)), form
        assert rel_l2(ga1.numpy(), ga0.numpy()) <= 1e-6, form
        assert (gb0 is None) == (gb1 is None) and (gb0 is None or rel_l2(gb1.numpy(), gb0.numpy()) <= 1e-6), form


@pytest.mark.parametrize("abc", ["gerjan", "sincos"])
def test_propagator_other_absorbing_boundaries(abc):
    """abc_type 'gerjan' / anything else builds the damping map with bc_gerjan / bc_sincos
    (acoustic_propagator.py:102-118); the step kernel is the same, so records and the vp
    gradient match the oracle fed the same map."""
    from adfwi_amd.model import AcousticModel
    from adfwi_amd.propagator import AcousticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    from adfwi_amd.utils import wavelet
    from oracle import acoustic as oa
    from oracle import boundary as ob
    nz, nx, nabc, nt, dt = 30, 60, 12, 220, 1e-3
    rng = np.random.default_rng(4)
    vp = (2200 + 400 * rng.random((nz, nx))).astype(np.float32)
    model = AcousticModel(0, 0, nx, nz, 10.0, 10.0, vp=vp, rho=np.ones_like(vp), vp_grad=True, nabc=nabc,
                          free_surface=False, abc_type=abc, device=DEV)
    src = Source(nt, dt, 15.0)
    _, wv = wavelet(nt, dt, 15.0)
    for x in (12, 47):
        src.add_source(x, 2, wv)
    rcv = Receiver(nt, dt)
    rcv.add_receivers(np.arange(nx), np.full(nx, 2), "pr")
    F = AcousticPropagator(model, Survey(src, rcv), device=DEV)
    damp = (ob.bc_gerjan(nx, nz, 10.0, 10.0, nabc, alpha=model.abc_jerjan_alpha, free_surface=False)
            if abc == "gerjan" else ob.bc_sincos(nx, nz, 10.0, 10.0, nabc, free_surface=False))
    assert np.array_equal(F.damp.cpu().numpy(), np.asarray(damp, dtype=np.float32))
    out = F.forward()
    out["p"].square().sum().backward()
    v = torch.tensor(model.vp.detach().cpu().numpy(), requires_grad=True)
    ref = oa.forward_kernel(nx, nz, 10.0, 10.0, nt, dt, nabc, False, torch.tensor([12, 47]), torch.tensor([2, 2]), 2,
                            F.wavelet.cpu(), torch.arange(nx), torch.full((nx,), 2), nx,
                            torch.tensor(damp, dtype=torch.float32), v, torch.tensor(model.rho.detach().cpu().numpy()))
    ref["p"].square().sum().backward()
    assert rel_l2(out["p"].detach().cpu().numpy(), ref["p"].detach().numpy()) <= 1e-5
    assert rel_l2(model.vp.grad.cpu().numpy(), v.grad.numpy()) <= 1e-4


@pytest.mark.parametrize("rho_grad", [False, True])
def test_resident_images_equal_recomputed_segments(monkeypatch, rho_grad):
    """Keeping the last segments' stencil (and pressure) images from the forward pass instead of
    recomputing those segments in the backward changes nothing: records and energy maps are
    bit-identical with none, some and all segments resident (nt 61 in 8-step segments, the last
    one odd), and so are the vp / rho / wavelet gradients (the padding's backward is a
    deterministic sum)."""
    from adfwi_amd.propagator import acoustic_kernels as ak
    g, m = acoustic_inputs("acoustic_fs")
    nt = 61
    out = {}
    monkeypatch.setenv("ADFWI_POISON_RESIDENT", "1")     # unwritten image points would be NaN
    for resident in ("0", "3", "100", None):
        if resident is None:
            monkeypatch.delenv("ADFWI_RESIDENT_IMAGES", raising=False)
        else:
            monkeypatch.setenv("ADFWI_RESIDENT_IMAGES", resident)
        v = torch.tensor(g["vp"], device=DEV, requires_grad=True)
        r = torch.tensor(g["rho"], device=DEV, requires_grad=rho_grad)
        w = torch.tensor(g["wavelet"][:, :nt], device=DEV, requires_grad=True)
        res = ak.forward_kernel(m["nx"], m["nz"], m["dx"], m["dx"], nt, m["dt"], m["nabc"], m["fs"],
                                torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), 2, w,
                                torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]), len(g["rcv_x"]),
                                torch.tensor(g["damp"], dtype=torch.float32, device=DEV), v, r,
                                device=DEV, ckpt_every=8)
        loss = sum((res[k] * torch.tensor(g["weights"][i][:, :nt], device=DEV)).sum() for i, k in enumerate("puw"))
        loss.backward()
        o = {k: res[k].detach().cpu().numpy() for k in ("p", "u", "w", "forward_wavefield_p")}
        o.update(vp=v.grad.cpu().numpy(), wavelet=w.grad.cpu().numpy())
        if rho_grad:
            o["rho"] = r.grad.cpu().numpy()
        out[resident] = o
    for resident in ("3", "100", None):
        for k, a in out["0"].items():
            assert np.array_equal(out[resident][k], a), (resident, k, rel_l2(out[resident][k], a))


def test_resident_budget_peak_stays_within_the_free_memory_it_assumed():
    """The resident-image budget (80 % of the free memory less 8 GiB, less the checkpoints and the
    backward's buffers) against the real peak allocation of a gradient: a blocker tensor leaves
    about 24 GiB free so that the budget binds (some but not all segments resident), then one
    32-shot 464 x 1232 gradient (nt 1000, 32-step segments of 2.4 GB of images) must allocate at
    most 80 % of the free memory it was given, and its gradient must equal the run with every
    segment recomputed."""
    from adfwi_amd.propagator import acoustic_kernels as ak
    from adfwi_amd.propagator.boundary_condition import bc_pml
    nz, nx, nabc, nt, S = 384, 1152, 40, 1000, 32
    rng = np.random.default_rng(8)
    vp = (2500 + 300 * rng.random((nz, nx))).astype(np.float32)
    rho = (310 * vp ** 0.25).astype(np.float32)
    damp = torch.tensor(bc_pml(nx, nz, 12.5, 12.5, nabc, float(vp.max()), False), dtype=torch.float32, device=DEV)
    Z, X = nz + 2 * nabc, nx + 2 * nabc
    sx, sz = np.linspace(5, nx - 6, S).astype(int), np.full(S, 2)
    rx, rz = np.arange(nx), np.full(nx, 2)
    wav = torch.tensor(rng.standard_normal((S, nt)), dtype=torch.float32, device=DEV)
    k = ak.checkpoint_interval(nt)
    nseg = -(-nt // k)

    def gradient(resident):
        geom = ak.AcousticGeometry(Z, X, nt, False, nabc, sz + nabc, sx + nabc, rz + nabc, rx + nabc, DEV,
                                   resident_images=resident)
        v = torch.tensor(vp, device=DEV, requires_grad=True)
        res = ak.forward_kernel(nx, nz, 12.5, 12.5, nt, 1e-3, nabc, False, torch.tensor(sx), torch.tensor(sz), S,
                                wav, torch.tensor(rx), torch.tensor(rz), nx, damp, v, torch.tensor(rho, device=DEV),
                                device=DEV, geometry=geom)
        (res["p"].double() ** 2).sum().backward()
        return v.grad.detach(), geom
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0, _ = torch.cuda.mem_get_info(DEV)
    blocker = torch.empty(max(0, free0 - (40 << 30)) // 4, dtype=torch.float32, device=DEV)
    try:
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated(DEV)
        free, _ = torch.cuda.mem_get_info(DEV)
        free += torch.cuda.memory_reserved(DEV) - base
        torch.cuda.reset_peak_memory_stats(DEV)
        probe = ak.AcousticGeometry(Z, X, nt, False, nabc, sz + nabc, sx + nabc, rz + nabc, rx + nabc, DEV)
        nres = ak.resident_segments(probe, k, True, False)
        g_res, _ = gradient(None)
        torch.cuda.synchronize()
        peak = torch.cuda.max_memory_allocated(DEV) - base
        print(f"free {free / 2**30:.1f} GiB, resident segments {nres} of {nseg}, peak {peak / 2**30:.1f} GiB",
              flush=True)
        assert 0 < nres < nseg, (nres, nseg)
        assert peak <= 0.8 * free, (peak / 2**30, free / 2**30)
    finally:
        del blocker
        torch.cuda.empty_cache()
    g_none, _ = gradient(0)
    assert torch.equal(g_res, g_none)
