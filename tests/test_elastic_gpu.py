"""Parity of the HIP elastic path (libadfwi_amd.so through the C ABI) with the oracle and with
the reference's own vectors.  Tolerances: forward records rel-L2 <= 1e-5 (bit-exact when the
kernel is fed the oracle's padded coefficient fields), gradients rel-L2 <= 1e-4."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from helpers import (ELASTIC_NAMES, elastic_abc, elastic_cc_general, elastic_damp, elastic_inputs,
                     elastic_random_case)

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _oracle_setup(g, m, *, shots_per_wave=0, tile_z=0, fdc=None):
    """Kernel inputs built on the CPU by the oracle (padded fields, PML factors, scalars).
    fdc: FD coefficients to use (default: the oracle's float32 solve on this host)."""
    from adfwi_amd.propagator.elastic_kernels import ElasticGeometry, pml_factors
    from oracle import elastic as oe
    nz, nx, nabc, fs = m["nz"], m["nx"], m["nabc"], m["fs"]
    nn = m["fd"] // 2
    fdc = oe.diff_coef(nn) if fdc is None else fdc
    cc = elastic_cc_general(g, nz, nx)
    f = oe.padded_fields(torch.tensor(g["bx"]), torch.tensor(g["bz"]), cc,
                         torch.tensor(g["bcx"], dtype=torch.float32), torch.tensor(g["bcz"], dtype=torch.float32),
                         nabc, fs, m["fd"])
    Z = nz + (nabc + nn if fs else 2 * nabc + nn)
    X = nx + 2 * nabc
    zoff = nn if fs else nabc + nn
    mt = torch.tensor(g["mt"])
    damping = elastic_abc(g) != "PML"
    if damping:
        src_m = torch.stack([(-1 / 3) * mt[:, 0, 0], (-1 / 3) * mt[:, 2, 2], (-1 / 3) * mt[:, 0, 2]], dim=1)
    else:
        half = mt / 2
        src_m = torch.stack([-half[:, 0, 0], -half[:, 2, 2], -half[:, 0, 2]], dim=1)
    aniso = "c15" in g
    geom = ElasticGeometry(Z, X, m["nt"], fs, g["src_z"] + zoff, g["src_x"] + nabc, src_m,
                           g["rcv_z"] + zoff, g["rcv_x"] + nabc, DEV, aniso=aniso,
                           shots_per_wave=shots_per_wave, tile_z=tile_z, damping=damping, nn=nn,
                           c2=float(fdc[2]) if nn > 2 else 0.0)
    fields = [f[k] for k in ("c11", "c13", "c33", "c55", "bx", "bz")]
    fields += [f["c15"], f["c35"]] if aniso else [None, None]
    coef = geom.pack_coef([None if t is None else t.to(DEV) for t in fields])
    if damping:
        px, pz = geom.pack_plane(oe.pad_elastic(elastic_damp(g), 0, nn, fs).to(DEV)), None
    else:
        px, pz = pml_factors(f["bcx"], f["bcz"], m["dt"], geom.pitch)
    f32 = lambda v: float(np.float32(v))
    scalars = (f32(m["dt"]), f32(m["dt"] / m["dx"]), f32(m["dt"] / m["dx"]), f32(m["dx"]), f32(m["dx"]),
               float(fdc[0]), float(fdc[1]))
    src = torch.tensor(g["wavelet"]).to(DEV)
    return geom, coef, px.to(DEV), None if pz is None else pz.to(DEV), src, scalars


def _run(geom, coef, px, pz, src, scalars, **kw):
    """run_forward with the absorbing-boundary tables of either scheme."""
    from adfwi_amd.propagator.elastic_kernels import run_forward
    if geom.damping:
        return run_forward(geom, coef, None, None, src, scalars, damp=px, **kw)
    return run_forward(geom, coef, px, pz, src, scalars, **kw)


LEAVES = ("c11", "c13", "c33", "c55", "bx", "bz", "c15", "c35")
CC_INDEX = {"c11": 0, "c13": 2, "c15": 4, "c33": 11, "c35": 13, "c55": 18}


def _leaves(g, device, want_grad):
    leaves = {k: torch.tensor(g[k], device=device, requires_grad=want_grad) for k in LEAVES if k in g}
    cc = [t.to(device) for t in elastic_cc_general(g, g["c11"].shape[0], g["c11"].shape[1])]
    for k, i in CC_INDEX.items():
        if k in leaves:
            cc[i] = leaves[k]
    return leaves, cc


def _run_oracle(g, m, seg=1, want_grad=False):
    from oracle import elastic as oe
    nz, nx = m["nz"], m["nx"]
    leaves, cc = _leaves(g, "cpu", want_grad)
    w = torch.tensor(g["wavelet"], requires_grad=want_grad)
    z = torch.zeros((nz, nx))
    res = oe.forward_kernel(nx, nz, m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                            torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), len(g["src_x"]), w,
                            torch.tensor(g["mt"]), torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]),
                            len(g["rcv_x"]), elastic_abc(g), torch.tensor(g["bcx"], dtype=torch.float32),
                            torch.tensor(g["bcz"], dtype=torch.float32), elastic_damp(g), z, z,
                            leaves["bx"], leaves["bz"], cc, n_segments=seg, fd_order=m["fd"])
    out = {k: res[k].detach().numpy() for k in ELASTIC_NAMES}
    out.update({"fw_" + k: res["forward_wavefield_" + k].numpy() for k in ELASTIC_NAMES})
    if want_grad:
        loss = sum((res[k] * torch.tensor(g["weights"][i])).sum() for i, k in enumerate(ELASTIC_NAMES))
        loss.backward()
        out.update({"grad_" + k: v.grad.numpy() for k, v in leaves.items()})
        out["grad_wavelet"] = w.grad.numpy()
    return out


def _reference_coefficients(fd):
    """The reference's DiffCoef(fd // 2, 's') as its host solved it (a float32 inverse of an
    ill-conditioned Vandermonde matrix, whose last bits depend on the host's LAPACK: at fd_order
    10 another host's solve moves the records by ~1e-5)."""
    from conftest import golden
    return torch.tensor(golden("diffcoef")["fdc"] if fd == 4 else golden("diffcoef6")[f"fdc{fd}"])


def _dropin(g, m, seg=1, want_grad=False, wavelet=None, fdc=None, **leaf_kw):
    """The drop-in forward_kernel on the GPU with the case's inputs; fdc: FD coefficients to
    feed it in place of its own DiffCoef solve (the reference's, for comparisons with the
    reference's vectors; the product's DiffCoef is checked against them on the CPU)."""
    from unittest import mock
    from adfwi_amd.propagator import elastic_kernels as ek
    nz, nx = m["nz"], m["nx"]
    leaves, cc = _leaves(g, DEV, want_grad)
    w = torch.tensor(g["wavelet"] if wavelet is None else wavelet, device=DEV, requires_grad=want_grad)
    z = torch.zeros((nz, nx), device=DEV)
    own = ek.DiffCoef
    with mock.patch.object(ek, "DiffCoef", lambda order, scheme: own(order, scheme) if fdc is None else fdc.clone()):
        res = ek.forward_kernel(nx, nz, m["dx"], m["dx"], m["nt"], m["dt"], m["nabc"], m["fs"],
                                torch.tensor(g["src_x"]), torch.tensor(g["src_z"]), len(g["src_x"]), w,
                                torch.tensor(g["mt"]), torch.tensor(g["rcv_x"]), torch.tensor(g["rcv_z"]),
                                len(g["rcv_x"]), elastic_abc(g), torch.tensor(g["bcx"], dtype=torch.float32),
                                torch.tensor(g["bcz"], dtype=torch.float32), elastic_damp(g), z, z, leaves["bx"],
                                leaves["bz"], cc, fd_order=m["fd"], n_segments=seg, device=DEV, **leaf_kw)
    return res, leaves, w


ALL_CASES = ["elastic_iso_fs", "elastic_vti_nofs", "elastic_abl_gerjan_fs", "elastic_abl_sincos_nofs",
             "elastic_fd6_iso_fs", "elastic_fd6_vti_nofs", "elastic_fd8_iso_fs", "elastic_fd10_vti_nofs",
             "elastic_fd6_abl_gerjan_fs", "elastic_fd8_abl_sincos_nofs"]


@pytest.mark.parametrize("name", ALL_CASES)
def test_kernel_bit_exact_given_oracle_coefficients(name):
    """Same coefficient bits in -> same record bits out: the fused step keeps the reference's
    fp32 association, stage order, split-field PML and free-surface rows.  The FD coefficients
    are the reference's own (a float32 matrix solve whose last bit can depend on the host's
    LAPACK), taken from the fixture."""
    g, m = elastic_inputs(name)
    fdc = _reference_coefficients(m["fd"])
    geom, coef, px, pz, src, scalars = _oracle_setup(g, m, fdc=fdc)
    res = _run(geom, coef, px, pz, src, scalars, record=True)
    torch.cuda.synchronize()
    for i, k in enumerate(ELASTIC_NAMES):
        got = res["rec"][i].cpu().numpy()
        assert np.array_equal(got, g[f"seg1_rec_{k}"]), (k, rel_l2(got, g[f"seg1_rec_{k}"]))


@pytest.mark.parametrize("name", ALL_CASES)
@pytest.mark.parametrize("seg", [1, 3])
def test_dropin_forward_matches_reference_vectors(name, seg):
    g, m = elastic_inputs(name)
    res, _, _ = _dropin(g, m, seg=seg, fdc=_reference_coefficients(m["fd"]))
    torch.cuda.synchronize()
    for k in ELASTIC_NAMES:
        got = res[k].detach().cpu().numpy()
        assert got.shape == g[f"seg{seg}_rec_{k}"].shape
        assert rel_l2(got, g[f"seg{seg}_rec_{k}"]) <= 1e-5, k
        fw = res["forward_wavefield_" + k].cpu().numpy()
        assert fw.shape == g[f"seg{seg}_fw_{k}"].shape
        assert rel_l2(fw, g[f"seg{seg}_fw_{k}"]) <= 1e-5, ("energy", k)


@pytest.mark.parametrize("fs,aniso,spw,tz,abc", [(True, False, 1, 4, "PML"), (False, False, 2, 8, "PML"),
                                                  (True, True, 1, 32, "PML"), (False, False, 2, 24, "PML"),
                                                  (True, True, 1, 16, "PML"), (False, True, 0, 0, "PML"),
                                                  (True, False, 1, 4, "gerjan"), (False, True, 2, 8, "sincos"),
                                                  (True, True, 0, 0, "gerjan"),
                                                  (True, False, 1, 8, "PML6"), (False, True, 2, 16, "PML6"),
                                                  (True, True, 0, 0, "PML8"), (False, False, 1, 8, "PML10"),
                                                  (True, False, 1, 8, "gerjan6"), (False, True, 2, 16, "sincos8"),
                                                  (True, True, 0, 0, "sincos10")])
def test_multi_strip_grid_matches_oracle(fs, aniso, spw, tz, abc):
    """A grid of several 55-column (51 at orders 6-10) strips and row tiles, several shots per
    wavefront: strip edges, tile edges, PML corners, both absorbing schemes, the three-term
    scheme of orders 6, 8, 10 and the anisotropic C15/C35 terms against the oracle."""
    abc, fd = _split_abc(abc)
    g, m = elastic_random_case(44, 150, 8, 50, 9, fs, seed=11 + int(fs) + 2 * int(aniso), aniso=aniso,
                               abc=abc, fd_order=fd)
    ref = _run_oracle(g, m)
    geom, coef, px, pz, src, scalars = _oracle_setup(g, m, shots_per_wave=spw, tile_z=tz)
    assert geom.layout.ntiles_x >= 3 and geom.layout.ntiles_z >= 2 or tz == 0
    res = _run(geom, coef, px, pz, src, scalars, record=True)
    torch.cuda.synchronize()
    for i, k in enumerate(ELASTIC_NAMES):
        got = res["rec"][i].cpu().numpy()
        assert np.array_equal(got, ref[k]), (k, rel_l2(got, ref[k]))


def _split_abc(spec):
    """'PML6' -> ('PML', 6), 'gerjan' -> ('gerjan', 4): absorbing scheme and FD order of a case."""
    name = spec.rstrip("0123456789")
    return name, int(spec[len(name):] or 4)


def _loss(res, weights):
    return sum((res[k] * torch.as_tensor(weights[i], device=res[k].device)).sum() for i, k in enumerate(ELASTIC_NAMES))


@pytest.mark.parametrize("name", ALL_CASES)
@pytest.mark.parametrize("seg", [1, 3])
def test_dropin_gradients_match_reference_vectors(name, seg):
    """Hand-written adjoint vs the reference's autograd gradients of a seeded linear functional
    of all five records (C11, C13, C33, C55, bx, bz on their own staggered shapes, wavelet)."""
    g, m = elastic_inputs(name)
    res, leaves, w = _dropin(g, m, seg=seg, want_grad=True, fdc=_reference_coefficients(m["fd"]))
    _loss(res, g["weights"]).backward()
    torch.cuda.synchronize()
    for k in ("c11", "c13", "c33", "c55", "bx", "bz"):
        got = leaves[k].grad.cpu().numpy()
        ref = g[f"seg{seg}_grad_{k}"]
        assert got.shape == ref.shape
        assert rel_l2(got, ref) <= 1e-4, (k, rel_l2(got, ref))
    assert rel_l2(w.grad.cpu().numpy(), g[f"seg{seg}_grad_wavelet"]) <= 1e-4


@pytest.mark.parametrize("fs,aniso,abc", [(True, False, "PML"), (False, True, "PML"), (True, True, "PML"),
                                          (True, False, "gerjan"), (False, True, "sincos"), (True, True, "sincos"),
                                          (True, False, "PML6"), (False, True, "PML6"), (True, True, "PML8"),
                                          (True, False, "PML10"), (True, False, "gerjan6"), (False, True, "sincos8"),
                                          (True, True, "gerjan10")])
def test_gradients_multi_strip_match_oracle(fs, aniso, abc):
    """Several strips and row tiles, several shot groups, anisotropic C15/C35 gradients too, both
    absorbing schemes and the three-term scheme of orders 6, 8, 10, against the oracle's autograd
    gradients."""
    abc, fd = _split_abc(abc)
    g, m = elastic_random_case(44, 150, 8, 60, 9, fs, seed=21 + int(fs) + 2 * int(aniso), aniso=aniso,
                               abc=abc, fd_order=fd)
    ref = _run_oracle(g, m, want_grad=True)
    res, leaves, w = _dropin(g, m, want_grad=True)
    _loss(res, g["weights"]).backward()
    torch.cuda.synchronize()
    for k, v in leaves.items():
        got = v.grad.cpu().numpy()
        assert rel_l2(got, ref["grad_" + k]) <= 1e-4, (k, rel_l2(got, ref["grad_" + k]))
    assert rel_l2(w.grad.cpu().numpy(), ref["grad_wavelet"]) <= 1e-4


@pytest.mark.parametrize("fs,aniso,repeat", [(True, False, False), (False, True, False), (True, True, False),
                                             (True, False, True)])
def test_fused_receivers_equal_separate_launches(monkeypatch, fs, aniso, repeat):
    """Records taken, and record adjoints added, inside the 4th-order PML step kernels (through
    the receiver map) against the separate gather and inject launches: a surface line on every
    column (with a free surface the free-surface row, whose tzz reads zero) and a well line down
    one column across several row strips of both kernels.  Records and gradients bit-identical,
    records within 1e-5 of the oracle.  repeat: one receiver point twice, where the geometry
    builds no map and the separate launches take over; their atomic adds of the two record
    adjoints onto one point come in either order, so gradients then agree to rel-L2 1e-6."""
    from adfwi_amd.propagator.elastic_kernels import ElasticGeometry
    g, m = elastic_random_case(44, 150, 8, 40, 9, fs, seed=31 + int(fs) + 2 * int(aniso), aniso=aniso)
    nz, nx = m["nz"], m["nx"]
    rx = [np.arange(nx), np.full(nz - 2, 37)] + ([np.array([37])] if repeat else [])
    rz = [np.zeros(nx), np.arange(1, nz - 1)] + ([np.array([5])] if repeat else [])
    g["rcv_x"], g["rcv_z"] = np.concatenate(rx).astype(np.int64), np.concatenate(rz).astype(np.int64)
    g["weights"] = np.random.default_rng(5).standard_normal((5, 9, m["nt"], len(g["rcv_x"]))).astype(np.float32)
    monkeypatch.setenv("ADFWI_EL_TILE_Z", "8")        # forward strips of 8 rows
    monkeypatch.setenv("ADFWI_ADJ_TILE_Z", "16")      # adjoint strips of 16 rows
    built = []
    own_init = ElasticGeometry.__init__

    def spy(self, *a, **kw):
        own_init(self, *a, **kw)
        built.append(self.rcv_map is not None)
    monkeypatch.setattr(ElasticGeometry, "__init__", spy)
    out = {}
    for fused in (True, False):
        monkeypatch.setenv("ADFWI_EL_FUSED_RECORDS", "1" if fused else "0")
        res, leaves, w = _dropin(g, m, want_grad=True)
        _loss(res, g["weights"]).backward()
        torch.cuda.synchronize()
        out[fused] = ({k: res[k].detach().cpu().numpy() for k in ELASTIC_NAMES},
                      {k: v.grad.cpu().numpy() for k, v in leaves.items()}, w.grad.cpu().numpy())
    assert built == [not repeat, False], built
    (rec_f, grad_f, w_f), (rec_s, grad_s, w_s) = out[True], out[False]
    for k in ELASTIC_NAMES:
        assert np.array_equal(rec_f[k], rec_s[k]), (k, rel_l2(rec_f[k], rec_s[k]))
        assert np.abs(rec_f[k]).max() > 0, k
    for k in grad_s:
        if repeat:
            assert rel_l2(grad_f[k], grad_s[k]) <= 1e-6, (k, rel_l2(grad_f[k], grad_s[k]))
        else:
            assert np.array_equal(grad_f[k], grad_s[k]), (k, rel_l2(grad_f[k], grad_s[k]))
    assert np.array_equal(w_f, w_s) if not repeat else rel_l2(w_f, w_s) <= 1e-6
    if fs:
        assert not np.any(rec_f["tzz"][:, :, :nx]), "the free-surface row's tzz records are zero"
    ref = _run_oracle(g, m)
    for k in ELASTIC_NAMES:
        assert rel_l2(rec_f[k], ref[k]) <= 1e-5, (k, rel_l2(rec_f[k], ref[k]))


@pytest.mark.parametrize("abc", ["PML", "gerjan", "PML6", "gerjan6"])
def test_adjoint_dot_product_in_the_wavelet(abc):
    """Records are linear in the wavelet: <J dw, dd> = <dw, J^T dd> with J^T dd the wavelet
    gradient of the hand-written adjoint (free surface and anisotropy on, several tiles)."""
    abc, fd = _split_abc(abc)
    g, m = elastic_random_case(40, 130, 8, 80, 5, True, seed=5, aniso=True, abc=abc, fd_order=fd)
    rng = np.random.default_rng(3)
    dw = rng.standard_normal(g["wavelet"].shape).astype(np.float32)
    dd = rng.standard_normal(g["weights"].shape).astype(np.float32)
    res, _, w = _dropin(g, m, want_grad=True, wavelet=dw)
    loss = _loss(res, dd)
    loss.backward()
    lhs = float(loss.detach().double())
    rhs = float((w.grad.double() * torch.tensor(dw, device=DEV).double()).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_checkpoint_interval_does_not_change_gradients():
    g, m = elastic_random_case(30, 90, 6, 45, 3, True, seed=8)
    out = []
    for k in (1, 7, 45):
        res, leaves, w = _dropin(g, m, want_grad=True, ckpt_every=k)
        _loss(res, g["weights"]).backward()
        out.append([leaves[n].grad.cpu().numpy() for n in ("c11", "c33", "c55", "bx")] + [w.grad.cpu().numpy()])
    for other in out[1:]:
        for a, b in zip(out[0], other):
            assert rel_l2(b, a) <= 1e-6


@pytest.mark.parametrize("name", ["elastic_iso_fs", "elastic_vti_nofs", "elastic_fd6_iso_fs", "elastic_fd6_vti_nofs"])
def test_elastic_propagator_class_matches_reference_vectors(name):
    """ElasticPropagator(model, survey).forward() with the reference's model parameterisation
    (IsotropicElasticModel / AnisotropicElasticModel from vp, vs, rho, eps, delta) reproduces the
    reference's records of the same case."""
    from adfwi_amd.model import AnisotropicElasticModel, IsotropicElasticModel
    from adfwi_amd.propagator import ElasticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    g, m = elastic_inputs(name)
    nz, nx = m["nz"], m["nx"]
    src = Source(m["nt"], m["dt"], 60.0)
    for s in range(len(g["src_x"])):
        src.add_source(int(g["src_x"][s]), int(g["src_z"][s]), g["wavelet"][s], "mt", g["mt"][s])
    rcv = Receiver(m["nt"], m["dt"])
    rcv.add_receivers(g["rcv_x"], g["rcv_z"], "vz")
    f32 = lambda k: g[k].astype(np.float32)
    common = dict(free_surface=m["fs"], nabc=m["nabc"], auto_update_rho=False, device=DEV)
    if np.any(g["eps"] != 0):
        model = AnisotropicElasticModel(0, 0, nx, nz, m["dx"], m["dx"], vp=f32("vp"), vs=f32("vs"), rho=f32("rho"),
                                        eps=f32("eps"), gamma=np.zeros((nz, nx), np.float32), delta=f32("delta"),
                                        **common)
    else:
        model = IsotropicElasticModel(0, 0, nx, nz, m["dx"], m["dx"], vp=f32("vp"), vs=f32("vs"), rho=f32("rho"),
                                      **common)
    prop = ElasticPropagator(model, Survey(src, rcv), device=DEV)
    with torch.no_grad():
        res = prop.forward(fd_order=m["fd"])
    for k in ELASTIC_NAMES:
        got = res[k].cpu().numpy()
        assert rel_l2(got, g[f"seg1_rec_{k}"]) <= 1e-5, (k, rel_l2(got, g[f"seg1_rec_{k}"]))


@pytest.mark.parametrize("name", ["elastic_abl_gerjan_fs", "elastic_fd6_abl_gerjan_fs"])
def test_elastic_propagator_damping_scheme_matches_reference_vectors(name):
    """ElasticPropagator with abc_type 'gerjan' (Cerjan taper from the model's alpha, the
    damping scheme of the reference) reproduces the reference's records of that case, at 4th
    and 6th order."""
    from adfwi_amd.model import IsotropicElasticModel
    from adfwi_amd.propagator import ElasticPropagator
    from adfwi_amd.survey import Receiver, Source, Survey
    g, m = elastic_inputs(name)
    nz, nx = m["nz"], m["nx"]
    src = Source(m["nt"], m["dt"], 60.0)
    for s_ in range(len(g["src_x"])):
        src.add_source(int(g["src_x"][s_]), int(g["src_z"][s_]), g["wavelet"][s_], "mt", g["mt"][s_])
    rcv = Receiver(m["nt"], m["dt"])
    rcv.add_receivers(g["rcv_x"], g["rcv_z"], "vz")
    f32 = lambda k: g[k].astype(np.float32)
    model = IsotropicElasticModel(0, 0, nx, nz, m["dx"], m["dx"], vp=f32("vp"), vs=f32("vs"), rho=f32("rho"),
                                  free_surface=m["fs"], nabc=m["nabc"], auto_update_rho=False, abc_type="gerjan",
                                  device=DEV)
    prop = ElasticPropagator(model, Survey(src, rcv), device=DEV)
    assert np.allclose(prop.damp.cpu().numpy(), g["damp"].astype(np.float32))
    with torch.no_grad():
        res = prop.forward(fd_order=m["fd"])
    for k in ELASTIC_NAMES:
        got = res[k].cpu().numpy()
        assert rel_l2(got, g[f"seg1_rec_{k}"]) <= 1e-5, (k, rel_l2(got, g[f"seg1_rec_{k}"]))


@pytest.mark.parametrize("abc,fs", [("PML", True), ("gerjan", True), ("PML6", False), ("gerjan6", True),
                                    ("sincos", False)])
def test_resident_traces_equal_recomputed_segments(monkeypatch, abc, fs):
    """Keeping the last segments' per-step states from the forward pass instead of recomputing
    them in the backward changes nothing (the same kernel writes them): records and energy maps
    are bit-identical with none, some and all segments resident, also when the energy-map chunks
    of checkpoint_segments=3 split a resident segment, and so are the gradients (the padding's
    backward is a deterministic sum)."""
    abc, fd = _split_abc(abc)
    g, m = elastic_random_case(30, 90, 6, 45, 3, fs, seed=8, aniso=True, abc=abc, fd_order=fd)
    out = {}
    monkeypatch.setenv("ADFWI_POISON_RESIDENT", "1")     # unwritten trace points would be NaN
    for resident in ("0", "2", "100", None):
        if resident is None:
            monkeypatch.delenv("ADFWI_RESIDENT_IMAGES", raising=False)
        else:
            monkeypatch.setenv("ADFWI_RESIDENT_IMAGES", resident)
        res, leaves, w = _dropin(g, m, seg=3, want_grad=True, ckpt_every=7)
        _loss(res, g["weights"]).backward()
        o = {k: res[k].detach().cpu().numpy() for k in ELASTIC_NAMES}
        o.update({"fw_" + k: res["forward_wavefield_" + k].cpu().numpy() for k in ELASTIC_NAMES})
        o.update({"grad_" + k: v.grad.cpu().numpy() for k, v in leaves.items()})
        o["grad_wavelet"] = w.grad.cpu().numpy()
        out[resident] = o
    for resident in ("2", "100", None):
        for k, a in out["0"].items():
            assert np.array_equal(out[resident][k], a), (resident, k, rel_l2(out[resident][k], a))


def test_energy_chunk_ending_on_a_checkpoint_step(monkeypatch):
    """checkpoint_segments=3 over nt 45 ends its chunks after steps 14 and 29, whose states are
    checkpoints when the checkpoint interval is 5: the next chunk must start from that
    checkpoint slot.  Records and energy maps against the oracle."""
    monkeypatch.setenv("ADFWI_RESIDENT_IMAGES", "0")
    g, m = elastic_random_case(30, 90, 6, 45, 3, True, seed=9)
    ref = _run_oracle(g, m, seg=3)
    res, _, _ = _dropin(g, m, seg=3, want_grad=True, ckpt_every=5)
    for k in ELASTIC_NAMES:
        assert rel_l2(res[k].detach().cpu().numpy(), ref[k]) <= 1e-5, k
        assert rel_l2(res["forward_wavefield_" + k].cpu().numpy(), ref["fw_" + k]) <= 1e-5, ("energy", k)
