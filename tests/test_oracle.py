"""The oracle is pinned: against the reference's own outputs (golden vectors), against an
explicit-loop restatement of one step, and against invariants of the discrete system."""
import numpy as np
import pytest
import torch

from conftest import golden, rel_l2
from helpers import (ELASTIC_LEAVES, ELASTIC_NAMES, acoustic_inputs, elastic_inputs, run_acoustic_oracle,
                     run_elastic_oracle)


@pytest.mark.parametrize("name", ["acoustic_nofs", "acoustic_fs"])
@pytest.mark.parametrize("seg", [1, 3])
def test_acoustic_oracle_matches_reference_vectors(name, seg):
    g, m = acoustic_inputs(name)
    out = run_acoustic_oracle(g, m, seg)
    for k in "puw":
        assert np.array_equal(out[k], g[f"seg{seg}_rec_{k}"]), k
        assert np.array_equal(out["fw_" + k], g[f"seg{seg}_fw_{k}"]), k
    for k in ("vp", "rho", "wavelet"):
        assert rel_l2(out["grad_" + k], g[f"seg{seg}_grad_{k}"]) < 1e-6, k


@pytest.mark.parametrize("name", ["elastic_iso_fs", "elastic_vti_nofs", "elastic_abl_gerjan_fs",
                                  "elastic_abl_sincos_nofs", "elastic_fd6_iso_fs", "elastic_fd6_vti_nofs",
                                  "elastic_fd8_iso_fs", "elastic_fd10_vti_nofs", "elastic_fd6_abl_gerjan_fs",
                                  "elastic_fd8_abl_sincos_nofs"])
@pytest.mark.parametrize("seg", [1, 3])
def test_elastic_oracle_matches_reference_vectors(name, seg):
    g, m = elastic_inputs(name)
    out = run_elastic_oracle(g, m, seg)
    for k in ELASTIC_NAMES:
        assert np.array_equal(out[k], g[f"seg{seg}_rec_{k}"]), k
        assert np.array_equal(out["fw_" + k], g[f"seg{seg}_fw_{k}"]), k
    for k in ELASTIC_LEAVES + ("wavelet",):
        assert rel_l2(out["grad_" + k], g[f"seg{seg}_grad_{k}"]) < 2e-6, k


def test_fd_coefficients_closed_form():
    from oracle.elastic import diff_coef
    c = diff_coef(2).numpy()
    assert np.array_equal(c, golden("diffcoef")["fdc"])
    assert abs(float(c[0]) - 9 / 8) < 1e-7 and abs(float(c[1]) + 1 / 24) < 1e-7
    # 6th order: 75/64, -25/384, 3/640 (staggered-grid Taylor coefficients); the reference
    # solves the Vandermonde system in float32, so the closed form holds to a few 1e-6
    c6 = diff_coef(3).numpy()
    assert np.array_equal(c6, golden("diffcoef6")["fdc6"])
    assert np.allclose(c6, [75 / 64, -25 / 384, 3 / 640], rtol=3e-5, atol=0)
    # 8th order: 1225/1024, -245/3072, 49/5120, -5/7168
    c8 = diff_coef(4).numpy()
    assert np.array_equal(c8, golden("diffcoef6")["fdc8"])
    assert np.allclose(c8, [1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168], rtol=3e-4, atol=0)
    assert np.array_equal(diff_coef(5).numpy(), golden("diffcoef6")["fdc10"])


def test_oracle_boundary_profiles_match_reference_vectors():
    from oracle import boundary as ob
    g = golden("boundary")
    assert np.array_equal(ob.bc_pml(24, 16, 10.0, 10.0, 6, 3000.0, False), g["pml_nofs"])
    assert np.array_equal(ob.bc_pml(24, 16, 10.0, 10.0, 6, 3000.0, True), g["pml_fs"])
    for fs in (0, 1):
        bx, bz = ob.bc_pml_xz(24, 16, 5.0, 5.0, 6, 3000.0, bool(fs))
        assert np.array_equal(bx, g[f"pmlxz_fs{fs}_x"]) and np.array_equal(bz, g[f"pmlxz_fs{fs}_z"])
        assert np.array_equal(ob.bc_sincos(24, 16, 10.0, 10.0, 6, bool(fs)), g[f"sincos_fs{fs}"])
        assert np.array_equal(ob.bc_gerjan(24, 16, 10.0, 10.0, 6, 0.0053, bool(fs)), g[f"gerjan_fs{fs}"])


def test_pml_profile_closed_form():
    """d(k) = kappa (k dx / a)^2, kappa = -3 vmax ln(1e-6) / (2a), a = (pml-1) dx; max(dx-, dz-profile)."""
    from oracle import boundary as ob
    pml, dx, vmax = 6, 10.0, 3000.0
    d = ob.bc_pml(24, 16, dx, dx, pml, vmax, False)
    a = (pml - 1) * dx
    prof = -3 * vmax * np.log(1e-6) / (2 * a) * (np.arange(pml) * dx / a) ** 2
    Z, X = d.shape
    dist_x = np.minimum(np.arange(X), X - 1 - np.arange(X))   # cells from the left/right edge
    dist_z = np.minimum(np.arange(Z), Z - 1 - np.arange(Z))
    px = np.where(dist_x < pml, prof[np.clip(pml - 1 - dist_x, 0, pml - 1)], 0.0)
    pz = np.where(dist_z < pml, prof[np.clip(pml - 1 - dist_z, 0, pml - 1)], 0.0)
    np.testing.assert_allclose(d, np.maximum(px[None, :], pz[:, None]), rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("free_surface", [False, True])
def test_vectorised_step_equals_loop_step(free_surface):
    """One vectorised oracle step equals the explicit-loop step (window bounds, FS rows)."""
    from oracle import acoustic as oa
    from oracle import loops
    rng = np.random.default_rng(5)
    Z, X, nabc = 14, 17, 4
    fs = nabc if free_surface else 1
    p = rng.standard_normal((Z, X)); u = rng.standard_normal((Z, X - 1)); w = rng.standard_normal((Z - 1, X))
    coef = [rng.uniform(0.5, 1.0, (Z, X)) for _ in range(5)]
    src = (fs + 2, 6, 0.37)
    rcv = [(fs + 1, 3), (Z - 3, X - 3), (fs, 8)]
    pl, ul, wl, recl = loops.acoustic_step(p, u, w, *coef, fs, free_surface, src, rcv)
    T = lambda a: torch.tensor(a, dtype=torch.float64)
    out = oa._steps(X - 2 * nabc, Z - 2 * nabc, 1.0, nabc, free_surface,
                    torch.tensor([src[1]]), torch.tensor([src[0]]), torch.tensor([[src[2]]], dtype=torch.float64),
                    torch.tensor([r[1] for r in rcv]), torch.tensor([r[0] for r in rcv]),
                    *[T(c) for c in coef], oa.C1, oa.C2, T(p)[None], T(u)[None], T(w)[None])
    np.testing.assert_allclose(out[0][0].numpy(), pl, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(out[1][0].numpy(), ul, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(out[2][0].numpy(), wl, rtol=1e-13, atol=1e-13)
    recs = np.stack([out[3][0, 0].numpy(), out[4][0, 0].numpy(), out[5][0, 0].numpy()], 1)
    np.testing.assert_allclose(recs, np.array(recl), rtol=1e-13, atol=1e-13)


def test_acoustic_oracle_linear_in_wavelet_and_shot_independent():
    g, m = acoustic_inputs("acoustic_nofs")
    base = run_acoustic_oracle(g, m, 1, dtype=torch.float64, want_grad=False)
    g2 = dict(g)
    g2["wavelet"] = 2.5 * g["wavelet"].astype(np.float64)
    scaled = run_acoustic_oracle(g2, m, 1, dtype=torch.float64, want_grad=False)
    assert rel_l2(scaled["p"], 2.5 * base["p"]) < 1e-12
    # shot 1 alone equals shot 1 of the batch
    g3 = {k: g[k] for k in g.files}
    g3["src_x"], g3["src_z"], g3["wavelet"] = g["src_x"][1:], g["src_z"][1:], g["wavelet"][1:]
    one = run_acoustic_oracle(g3, m, 1, dtype=torch.float64, want_grad=False)
    assert np.array_equal(one["p"][0], base["p"][1])


def test_acoustic_oracle_gradient_matches_central_differences():
    """Autograd of the oracle agrees with central differences on a tiny grid (free surface on)."""
    from oracle import acoustic as oa
    rng = np.random.default_rng(3)
    nz, nx, nabc, nt = 6, 7, 3, 12
    damp = torch.tensor(rng.uniform(0, 20, (nz + 2 * nabc, nx + 2 * nabc)))
    wav = torch.tensor(rng.standard_normal((1, nt)))

    def f(vp):
        res = oa.forward_kernel(nx, nz, 10.0, 10.0, nt, 1e-3, nabc, True, torch.tensor([3]), torch.tensor([1]),
                                1, wav, torch.tensor([1, 5]), torch.tensor([1, 2]), 2, damp, vp,
                                torch.full((nz, nx), 2000.0, dtype=torch.float64), dtype=torch.float64)
        return res["p"].sum() + res["w"].sum()
    vp = torch.tensor(rng.uniform(1500, 2500, (nz, nx)), requires_grad=True)
    f(vp).backward()
    d = torch.tensor(rng.standard_normal((nz, nx)))
    h = 0.5  # the padded velocity is float32 by design (reference quirk), so h spans many ulps
    with torch.no_grad():
        fd = (f(vp + h * d) - f(vp - h * d)) / (2 * h)
    assert abs(float(fd) - float((vp.grad * d).sum())) <= 1e-3 * abs(float(fd))
