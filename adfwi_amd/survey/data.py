"""Shot-gather records on disk (ADFWI/survey/data.py:22-132).

`SeismicData` keeps the reference's constructor, fields (src_num, rcv_num, src_loc, rcv_loc,
src_type, rcv_type, nt, dt, t, data), `record_data`, `save`, `load` and `parse_*_data`.  The
file is an .npz whose entries are all plain arrays: each record component `k` is stored as
`data_k`, the component names in `data_keys`, the type lists as string arrays and the scalars
as 0-d arrays, so `load` reads it with `allow_pickle=False`.  The reference stores its record
dictionary as one pickled object (`np.savez(..., data=dict)`, data.py:63-75); such a file is
refused with an explanation rather than unpickled.  `to_device` puts the records on the GPU,
optionally normalised per trace, as the FWI drivers use the observed data
(fwi/acoustic_fwi.py:66-69).
"""
from typing import Dict, Optional

import numpy as np
import torch

FORMAT = "adfwi_amd.seismic_data.v1"


def _to_numpy(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy().copy()
    return np.array(v, copy=True)


class SeismicData:
    def __init__(self, survey) -> None:
        self.survey = survey
        self.src_num = survey.source.num
        self.rcv_num = survey.receiver.num
        self.src_loc = survey.source.get_loc()
        self.rcv_loc = survey.receiver.get_loc()
        self.src_type = survey.source.get_type()
        self.rcv_type = survey.receiver.get_type()
        self.nt = survey.receiver.nt
        self.dt = survey.receiver.dt
        self.t = np.arange(self.nt) * self.dt
        self.data: Optional[Dict[str, np.ndarray]] = None

    def __repr__(self):
        return (f"Seismic Data:\n  Source number : {self.src_num}\n  Receiver number : {self.rcv_num}\n"
                f"  Time samples : {self.nt} samples at {self.dt * 1000:.2f} ms\n")

    def record_data(self, data: dict) -> None:
        """Stores host copies of every entry of a propagator's result dictionary (data.py:52-61)."""
        for key, value in data.items():
            data[key] = _to_numpy(value)
        self.data = data

    def save(self, path: str) -> None:
        if self.data is None:
            raise ValueError("no records: call record_data first")
        keys = list(self.data)
        out = {f"data_{k}": np.asarray(self.data[k]) for k in keys}
        out.update(format=np.array(FORMAT), data_keys=np.array(keys, dtype=str),
                   src_loc=np.asarray(self.src_loc), rcv_loc=np.asarray(self.rcv_loc),
                   src_num=np.array(self.src_num), rcv_num=np.array(self.rcv_num),
                   rcv_type=np.array(list(self.rcv_type), dtype=str), src_type=np.array(list(self.src_type), dtype=str),
                   t=np.asarray(self.t), nt=np.array(self.nt), dt=np.array(self.dt))
        np.savez(path, **out)

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as f:
            names = set(f.files)
            if "data_keys" not in names:
                if "data" in names:
                    raise ValueError(f"{path} stores its records as one pickled object (the reference's layout); "
                                     "this loader never unpickles files: write the records with "
                                     "SeismicData.save, which keeps every component as a plain array")
                raise ValueError(f"{path} is not a seismic-data file (no data_keys entry)")
            self.data = {str(k): f[f"data_{k}"] for k in f["data_keys"]}
            self.src_loc, self.rcv_loc = f["src_loc"], f["rcv_loc"]
            self.src_num, self.rcv_num = int(f["src_num"]), int(f["rcv_num"])
            self.rcv_type, self.src_type = f["rcv_type"], f["src_type"]
            self.t, self.nt, self.dt = f["t"], int(f["nt"]), float(f["dt"])

    @staticmethod
    def _norm(a):
        return a / np.max(np.abs(a), axis=1, keepdims=True)

    def parse_elastic_data(self, normalize=False):
        """(pressure = -(txx + tzz), txz, vx, vz), each optionally normalised per trace (data.py:102-114)."""
        d = self.data
        pressure, txz, vx, vz = -(d["txx"] + d["tzz"]), d["txz"], d["vx"], d["vz"]
        if normalize:
            pressure, txz, vx, vz = (self._norm(a) for a in (pressure, txz, vx, vz))
        return pressure, txz, vx, vz

    def parse_acoustic_data(self, normalize=False):
        """(p, u, w), each optionally normalised per trace (data.py:116-124)."""
        d = self.data
        p, u, w = d["p"], d["u"], d["w"]
        if normalize:
            p, u, w = (self._norm(a) for a in (p, u, w))
        return p, u, w

    def to_device(self, device, keys=None, normalize: bool = False, dtype=torch.float32) -> Dict[str, torch.Tensor]:
        """The record components (all (shots, nt, receivers) ones by default) as device tensors,
        optionally divided per trace by their maximum |value| (fwi/acoustic_fwi.py:66-69)."""
        from ..fwi.misfit import normalize_traces
        keys = [k for k, v in self.data.items() if np.ndim(v) == 3] if keys is None else keys
        out = {}
        for k in keys:
            t = torch.as_tensor(np.asarray(self.data[k]), device=device).to(dtype)
            out[k] = normalize_traces(t) if normalize else t
        return out
