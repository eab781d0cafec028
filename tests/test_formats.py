"""On-disk formats (no GPU): SeismicData records and model files round-trip with plain arrays
only, and a file in the reference's pickled layout is refused, never unpickled."""
import numpy as np
import pytest
import torch

from adfwi_amd.model import AcousticModel, IsotropicElasticModel, load_model_data
from adfwi_amd.survey import Receiver, SeismicData, Source, Survey


def _survey(nt=50):
    src = Source(nt, 1e-3, 10.0)
    src.add_sources([3, 9, 15], [2, 2, 2], np.sin(np.arange(nt) * 0.3))
    rcv = Receiver(nt, 1e-3)
    rcv.add_receivers(np.arange(20), np.full(20, 1), "pr")
    return Survey(src, rcv)


def test_seismic_data_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    recs = {k: torch.tensor(rng.standard_normal((3, 50, 20)), dtype=torch.float32) for k in "puw"}
    recs["forward_wavefield_p"] = torch.tensor(rng.random((12, 20)), dtype=torch.float32)
    d = SeismicData(_survey())
    d.record_data(dict(recs))
    assert all(isinstance(v, np.ndarray) for v in d.data.values())
    d.save(str(tmp_path / "shots.npz"))
    e = SeismicData(_survey())
    e.load(str(tmp_path / "shots.npz"))
    assert set(e.data) == set(recs)
    for k, v in recs.items():
        assert np.array_equal(e.data[k], v.numpy()) and e.data[k].dtype == np.float32
    assert (e.nt, e.dt, e.src_num, e.rcv_num) == (50, 1e-3, 3, 20)
    assert np.array_equal(e.src_loc, d.src_loc) and np.array_equal(e.rcv_loc, d.rcv_loc)
    assert list(e.rcv_type) == ["pr"] * 20 and list(e.src_type) == ["mt"] * 3
    p, u, w = e.parse_acoustic_data(normalize=True)
    assert np.allclose(np.abs(p).max(axis=1), 1.0) and np.allclose(p * np.abs(recs["p"].numpy()).max(axis=1,
                       keepdims=True), recs["p"].numpy())


def test_elastic_parse_gives_negative_stress_sum_as_pressure():
    rng = np.random.default_rng(1)
    d = SeismicData(_survey())
    d.record_data({k: rng.standard_normal((3, 50, 20)) for k in ("txx", "tzz", "txz", "vx", "vz")})
    pressure, txz, vx, vz = d.parse_elastic_data()
    assert np.array_equal(pressure, -(d.data["txx"] + d.data["tzz"])) and vz is d.data["vz"]


def test_reference_layout_is_refused_not_unpickled(tmp_path):
    path = tmp_path / "ref_layout.npz"
    np.savez(path, data={"p": np.zeros((1, 2, 3))}, src_loc=np.zeros((1, 2)), nt=2, dt=0.1)
    with pytest.raises(ValueError, match="pickled"):
        SeismicData(_survey()).load(str(path))


@pytest.mark.parametrize("bounded", [False, True])
def test_model_file_round_trip(tmp_path, bounded):
    rng = np.random.default_rng(2)
    nz, nx = 12, 18
    vp = (2000 + 1000 * rng.random((nz, nx))).astype(np.float32)
    kw = dict(vp_bound=(1500.0, 4000.0)) if bounded else {}
    m = AcousticModel(0.0, 0.0, nx, nz, 10.0, 10.0, vp=vp, rho=np.full((nz, nx), 2000.0, np.float32),
                      vp_grad=True, nabc=8, free_surface=True, **kw)
    m.save(str(tmp_path / "model.npz"))
    with np.load(str(tmp_path / "model.npz"), allow_pickle=False) as f:
        assert all(f[k].dtype != object for k in f.files)
    got = load_model_data(str(tmp_path / "model.npz"))
    assert got["vp_bound"] == ([1500.0, 4000.0] if bounded else [None, None])
    assert got["vp_grad"] is True and got["rho_grad"] is False and got["nabc"] == 8 and got["free_surface"] is True
    m2 = AcousticModel(**got)
    assert np.array_equal(m2.get_model("vp"), m.get_model("vp")) and m2.get_bound("vp") == m.get_bound("vp")
    e = IsotropicElasticModel(0.0, 0.0, nx, nz, 5.0, 5.0, vp=vp, vs=vp / 1.7, rho=np.full((nz, nx), 2100.0),
                              nabc=8, auto_update_rho=False)
    e.save(str(tmp_path / "el.npz"))
    ge = load_model_data(str(tmp_path / "el.npz"))
    assert np.array_equal(ge["vs"], e.get_model("vs")) and ge["vs_bound"] == [None, None]
