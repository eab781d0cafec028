"""On-disk formats (no GPU): SeismicData records and model files round-trip with plain arrays
only; a file in the reference's pickled layout is read by a decoder that admits nothing but
containers, strings, numbers and plain arrays."""
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


def _reference_layout_file(path, recs, numpy1_names=False):
    """An .npz in the layout the reference's SeismicData.save writes (survey/data.py:63-75): the
    record dictionary as one object entry, the other fields as np.savez stores them.  With
    numpy1_names, the pickled entry names numpy.core.multiarray as numpy 1.x (the reference's
    pinned 1.24.4) writes it."""
    np.savez(path, data=dict(recs), src_loc=np.array([[3, 2], [9, 2]]), rcv_loc=np.array([[0, 1], [1, 1]]),
             src_num=2, rcv_num=2, rcv_type=["pr", "pr"], src_type=["mt", "mt"], t=np.arange(4) * 1e-3,
             nt=4, dt=1e-3)
    if numpy1_names:
        import zipfile
        with zipfile.ZipFile(path) as z:
            entries = {n: z.read(n) for n in z.namelist()}
        raw = entries["data.npy"]
        n = raw.count(b"\x8c\x16numpy._core.multiarray")
        assert n >= 1
        raw = raw.replace(b"\x8c\x16numpy._core.multiarray", b"\x8c\x15numpy.core.multiarray")
        f = raw.index(b"\x80\x04\x95") + 3       # protocol 4: one frame, its length follows
        size = int.from_bytes(raw[f:f + 8], "little") - n
        entries["data.npy"] = raw[:f] + size.to_bytes(8, "little") + raw[f + 8:]
        with zipfile.ZipFile(path, "w") as z:
            for n, b in entries.items():
                z.writestr(n, b)


@pytest.mark.parametrize("numpy1_names", [False, True])
def test_reference_layout_is_read_without_unpickling_code(tmp_path, numpy1_names):
    rng = np.random.default_rng(5)
    recs = {"p": rng.standard_normal((2, 4, 2)).astype(np.float32), "u": np.zeros((2, 4, 2), np.float32),
            "forward_wavefield_p": rng.random((3, 5)), "n": np.arange(3, dtype=">i8")}
    path = tmp_path / "ref_layout.npz"
    _reference_layout_file(path, recs, numpy1_names)
    d = SeismicData(_survey())
    d.load(str(path))
    assert set(d.data) == set(recs)
    for k, v in recs.items():
        assert d.data[k].dtype == v.dtype and np.array_equal(d.data[k], v), k
    assert (d.nt, d.dt, d.src_num, d.rcv_num) == (4, 1e-3, 2, 2)
    assert list(d.rcv_type) == ["pr", "pr"] and list(d.src_type) == ["mt", "mt"]
    assert np.array_equal(d.src_loc, [[3, 2], [9, 2]]) and np.array_equal(d.t, np.arange(4) * 1e-3)
    p, u, w = d.parse_acoustic_data() if "w" in d.data else (d.data["p"], d.data["u"], None)
    assert np.array_equal(p, recs["p"])


class _Trap:
    """Pickles as a call of a function the decoder must refuse (nothing is ever called here:
    the test asserts the refusal happens before any global is resolved to a callable)."""

    def __init__(self, target):
        self.target = target

    def __reduce__(self):
        return (self.target, ("unexpected",))


@pytest.mark.parametrize("payload", ["builtin", "numpy_other", "structured"])
def test_reference_layout_decoder_refuses_other_globals(tmp_path, payload):
    import pickle
    from adfwi_amd.survey.data import read_reference_npz
    if payload == "builtin":
        bad = {"p": _Trap(print)}
    elif payload == "numpy_other":
        bad = {"p": _Trap(np.load)}
    else:
        bad = {"p": np.zeros(2, dtype=[("a", "f4"), ("b", "i4")])}
    path = tmp_path / "bad.npz"
    np.savez(path, data=np.array(bad, dtype=object), nt=1)
    with pytest.raises(pickle.UnpicklingError):
        read_reference_npz(str(path))
    with pytest.raises(pickle.UnpicklingError):
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
