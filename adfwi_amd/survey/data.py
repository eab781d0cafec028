"""Shot-gather records on disk (ADFWI/survey/data.py:22-132).

`SeismicData` keeps the reference's constructor, fields (src_num, rcv_num, src_loc, rcv_loc,
src_type, rcv_type, nt, dt, t, data), `record_data`, `save`, `load` and `parse_*_data`.  The
file is an .npz whose entries are all plain arrays: each record component `k` is stored as
`data_k`, the component names in `data_keys`, the type lists as string arrays and the scalars
as 0-d arrays, so `load` reads it with `allow_pickle=False`.  The reference stores its record
dictionary as one pickled object (`np.savez(..., data=dict)`, data.py:63-75); `load` reads such
files too, through `read_reference_npz`, whose decoder admits only containers, strings, numbers
and numpy arrays, dtypes and scalars rebuilt by this module's own code: no class or function
named in the file is imported or called.  `to_device` puts the records on the GPU,
optionally normalised per trace, as the FWI drivers use the observed data
(fwi/acoustic_fwi.py:66-69).
"""
import io
import pickle
import zipfile
from typing import Dict, Optional

import numpy as np
import torch

FORMAT = "adfwi_amd.seismic_data.v1"

# -- the reference's pickled layout, decoded without running anything from the file -------------
# numpy writes an object array (the record dictionary) as a pickle of the array; its arrays,
# dtypes and scalars are rebuilt through numpy.core.multiarray._reconstruct (numpy._core in
# numpy 2), numpy.dtype and numpy.core.multiarray.scalar.  The decoder maps exactly those names
# to the builders below and refuses every other global, so the pickle can only produce
# containers, strings, numbers and plain numeric / string arrays.
_MULTIARRAY = ("numpy.core.multiarray", "numpy._core.multiarray")
_MAX_ENTRY_BYTES = 1 << 36


class _ArrayTag:
    """Stands for numpy.ndarray in the stream (only used as _reconstruct's first argument)."""


class _DtypeState:
    def __init__(self, name, align=False, copy=True):
        if not isinstance(name, str):
            raise pickle.UnpicklingError("dtype name must be a string")
        self.dtype = np.dtype(name)
        if self.dtype.fields is not None or self.dtype.subdtype is not None:
            raise pickle.UnpicklingError(f"structured dtype {name!r} is not admitted")

    def __setstate__(self, state):
        if not isinstance(state, tuple) or len(state) < 2 or state[1] not in ("<", ">", "|", "="):
            raise pickle.UnpicklingError("malformed dtype state")
        if len(state) >= 5 and (state[3] is not None or state[4] is not None):
            raise pickle.UnpicklingError("structured dtypes are not admitted")
        kind = self.dtype.kind
        if kind not in "biufcUSO":
            raise pickle.UnpicklingError(f"dtype kind {kind!r} is not admitted")
        if kind in "US" and len(state) >= 6 and isinstance(state[5], int) and state[5] > 0:
            # flexible types carry their item size in bytes here (4 bytes per unicode character)
            self.dtype = np.dtype(f"{kind}{state[5] // 4 if kind == 'U' else state[5]}")
        if state[1] in "<>" and self.dtype.byteorder != "|":
            self.dtype = self.dtype.newbyteorder(state[1])


class _ArrayState:
    def __init__(self):
        self.value = None

    def __setstate__(self, state):
        if not isinstance(state, tuple) or len(state) != 5:
            raise pickle.UnpicklingError("malformed array state")
        _, shape, dt, fortran, raw = state
        if not isinstance(dt, _DtypeState) or not isinstance(shape, tuple) or \
                not all(isinstance(n, int) and n >= 0 for n in shape):
            raise pickle.UnpicklingError("malformed array state")
        count = int(np.prod(shape, dtype=np.int64))
        if dt.dtype.kind == "O":
            if not isinstance(raw, list) or len(raw) != count:
                raise pickle.UnpicklingError("malformed object array")
            out = np.empty(count, dtype=object)
            for i, v in enumerate(raw):
                out[i] = v
            self.value = out.reshape(shape, order="F" if fortran else "C")
            return
        if not isinstance(raw, (bytes, bytearray)) or len(raw) != count * dt.dtype.itemsize:
            raise pickle.UnpicklingError("array data does not match its shape and dtype")
        self.value = np.frombuffer(bytes(raw), dtype=dt.dtype).reshape(shape, order="F" if fortran else "C").copy()


def _reconstruct(tag, shape, typecode):
    if tag is not _ArrayTag:
        raise pickle.UnpicklingError("only numpy arrays may be reconstructed")
    return _ArrayState()


def _scalar(dt, raw=None):
    if not isinstance(dt, _DtypeState) or dt.dtype.kind == "O" or not isinstance(raw, (bytes, bytearray)) \
            or len(raw) != dt.dtype.itemsize:
        raise pickle.UnpicklingError("malformed numpy scalar")
    return np.frombuffer(bytes(raw), dtype=dt.dtype)[0]


class _RecordUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if module in _MULTIARRAY and name == "_reconstruct":
            return _reconstruct
        if module in _MULTIARRAY and name == "scalar":
            return _scalar
        if module == "numpy" and name == "ndarray":
            return _ArrayTag
        if module == "numpy" and name == "dtype":
            return _DtypeState
        raise pickle.UnpicklingError(f"{module}.{name} is not admitted in a seismic-data file")


def _finish(v):
    """Replaces the decoder's array stand-ins by the arrays they built, recursively."""
    if isinstance(v, _ArrayState):
        if v.value is None:
            raise pickle.UnpicklingError("array without state")
        a = v.value
        if a.dtype == object:
            flat = [_finish(x) for x in a.reshape(-1)]
            if a.shape == ():
                return flat[0]
            out = np.empty(len(flat), dtype=object)
            out[:] = flat
            return out.reshape(a.shape)
        return a
    if isinstance(v, (_DtypeState, _ArrayTag)) or v is _ArrayTag:
        raise pickle.UnpicklingError("dtype or array type outside an array")
    if isinstance(v, dict):
        return {_finish(k): _finish(x) for k, x in v.items()}
    if isinstance(v, list):
        return [_finish(x) for x in v]
    if isinstance(v, tuple):
        return tuple(_finish(x) for x in v)
    if v is None or isinstance(v, (str, bytes, bool, int, float, complex, np.generic)):
        return v
    raise pickle.UnpicklingError(f"{type(v).__name__} is not admitted in a seismic-data file")


def read_reference_npz(path: str) -> Dict[str, object]:
    """Every entry of an .npz written by np.savez, as the reference's SeismicData.save writes it
    (data.py:63-75): plain arrays are read with allow_pickle=False; an object array (the record
    dictionary) is decoded by the restricted decoder above.  Raises pickle.UnpicklingError for
    anything but containers, strings, numbers and numeric / string arrays."""
    out = {}
    with zipfile.ZipFile(path) as z:
        for info in z.infolist():
            if not info.filename.endswith(".npy"):
                continue
            if info.file_size > _MAX_ENTRY_BYTES:
                raise ValueError(f"{info.filename}: entry of {info.file_size} bytes is too large")
            f = io.BytesIO(z.read(info))
            version = np.lib.format.read_magic(f)
            if version == (1, 0):
                shape, fortran, dtype = np.lib.format.read_array_header_1_0(f)
            else:
                shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
            key = info.filename[:-4]
            if dtype.hasobject:
                out[key] = _finish(_RecordUnpickler(f).load())
            else:
                f.seek(0)
                out[key] = np.lib.format.read_array(f, allow_pickle=False)
    return out


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
                    self._load_reference_layout(path)
                    return
                raise ValueError(f"{path} is not a seismic-data file (no data_keys entry)")
            self.data = {str(k): f[f"data_{k}"] for k in f["data_keys"]}
            self.src_loc, self.rcv_loc = f["src_loc"], f["rcv_loc"]
            self.src_num, self.rcv_num = int(f["src_num"]), int(f["rcv_num"])
            self.rcv_type, self.src_type = f["rcv_type"], f["src_type"]
            self.t, self.nt, self.dt = f["t"], int(f["nt"]), float(f["dt"])

    def _load_reference_layout(self, path: str) -> None:
        """The reference's file (data.py:63-99): the record dictionary pickled as `data`, the
        other fields as arrays."""
        f = read_reference_npz(path)
        data = f["data"]
        if not isinstance(data, dict) or not all(isinstance(k, str) and isinstance(v, np.ndarray)
                                                 for k, v in data.items()):
            raise ValueError(f"{path}: `data` is not a dictionary of arrays")
        self.data = data
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
