"""Acquisition geometry read by the propagators (ADFWI/survey/{source,receiver,survey}.py).

Integer grid indices, per-shot wavelets (S, nt) and moment tensors (S, 3, 3);
receivers are shared by all shots.  Same constructors, methods and errors as
the reference (source.py:58-144, receiver.py:70-112, survey.py:20-56).
"""
from typing import Optional

import numpy as np

from ..utils import list2numpy, numpy2list

_IDENTITY_MT = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]])


class Source:
    """Seismic sources: locations (grid indices), wavelets and moment tensors."""

    def __init__(self, nt: int, dt: float, f0: float) -> None:
        self.nt, self.dt, self.f0 = nt, dt, f0
        self.t = np.arange(nt) * dt
        self.loc_x, self.loc_z, self.loc = [], [], []
        self.type, self.wavelet, self.moment_tensor = [], [], []
        self.num = 0

    def __repr__(self):
        if not self.num:
            return "Seismic Source:\n  empty\n"
        x, z = np.array(self.loc_x), np.array(self.loc_z)
        return (f"Seismic Source:\n  Source wavelet: {self.nt} samples at {self.dt * 1000:.2f} ms\n"
                f"  Source number : {self.num}\n  Source types  : {self.get_type(unique=True)}\n"
                f"  Source x range: {x.min()} - {x.max()} (grids)\n"
                f"  Source z range: {z.min()} - {z.max()} (grids)\n")

    def _validate(self, src_wavelet, src_type, src_mt):
        # "vf" (vertical force, elastic propagator only) has no counterpart in the reference,
        # which accepts moment-tensor sources alone
        if src_type.lower() not in ["mt", "vf"]:
            raise ValueError("Source type must be either mt or vf")
        if src_type.lower() == "vf":
            return
        if src_wavelet.shape[0] != self.nt:
            raise ValueError("Source wavelet must have the same length as the number of time samples")
        if src_mt is None:
            raise ValueError("Moment tensor must be provided for mt source")
        if np.shape(src_mt) != (3, 3):
            raise ValueError("Moment tensor must be a 3x3 matrix")

    def add_sources(self, src_x: np.ndarray, src_z: np.ndarray, src_wavelet: np.ndarray,
                    src_type: Optional[str] = "mt", src_mt: Optional[np.ndarray] = _IDENTITY_MT) -> None:
        """Adds several sources sharing one wavelet and moment tensor."""
        if np.shape(src_x) != np.shape(src_z):
            raise ValueError("Source location along x and z direction must have the same shape")
        self._validate(src_wavelet, src_type, src_mt)
        n = len(src_x)
        self.loc_x.extend(numpy2list(np.asarray(src_x).reshape(-1)))
        self.loc_z.extend(numpy2list(np.asarray(src_z).reshape(-1)))
        self.type.extend([src_type] * n)
        self.wavelet.extend(np.ones((n, self.nt)) * src_wavelet)
        mt = np.zeros((3, 3)) if src_type.lower() == "vf" else src_mt
        self.moment_tensor.extend(np.ones((n, 3, 3)) * mt)
        self.num += n

    def add_source(self, src_x: int, src_z: int, src_wavelet: np.ndarray,
                   src_type: Optional[str] = "mt", src_mt: Optional[np.ndarray] = _IDENTITY_MT) -> None:
        """Appends one source."""
        self._validate(src_wavelet, src_type, src_mt)
        self.loc_x.append(src_x)
        self.loc_z.append(src_z)
        self.type.append(src_type)
        self.wavelet.append(src_wavelet)
        self.moment_tensor.append(np.zeros((3, 3)) if src_type.lower() == "vf" else src_mt)
        self.num += 1

    def get_loc(self):
        loc = np.hstack((list2numpy(self.loc_x).reshape(-1, 1), list2numpy(self.loc_z).reshape(-1, 1)))
        self.loc = loc.copy()
        return loc

    def get_wavelet(self):
        return list2numpy(self.wavelet)

    def get_moment_tensor(self):
        return list2numpy(self.moment_tensor)

    def get_type(self, unique=False):
        return list2numpy(list(set(self.type))) if unique else list2numpy(self.type)


class Receiver:
    """Receivers shared by every shot (grid indices)."""

    def __init__(self, nt: int, dt: float) -> None:
        self.nt, self.dt = nt, dt
        self.loc_x, self.loc_z, self.locs, self.type = [], [], [], []
        self.num = 0

    def __repr__(self):
        if not self.num:
            return "Seismic Receiver:\n  empty\n"
        x, z = np.array(self.loc_x), np.array(self.loc_z)
        return (f"Seismic Receiver:\n  Receiver data   : {self.nt} samples at {self.dt * 1000:.2f} ms\n"
                f"  Receiver number : {self.num}\n  Receiver types  : {self.get_type(unique=True)}\n"
                f"  Receiver x range: {x.min()} - {x.max()} (grids)\n"
                f"  Receiver z range: {z.min()} - {z.max()} (grids)\n")

    def add_receivers(self, rcv_x: np.ndarray, rcv_z: np.ndarray, rcv_type: str) -> None:
        if np.shape(rcv_x) != np.shape(rcv_z):
            raise ValueError("Receiver Error: Inconsistant number of receiver in X and Z directions")
        if rcv_type.lower() not in ["pr", "vx", "vz"]:
            raise ValueError("Receiver type must be either pr, vx, vz")
        n = len(rcv_x)
        self.loc_x.extend(numpy2list(np.asarray(rcv_x).reshape(-1)))
        self.loc_z.extend(numpy2list(np.asarray(rcv_z).reshape(-1)))
        self.type.extend([rcv_type] * n)
        self.num += n

    def add_receiver(self, rcv_x: int, rcv_z: int, rcv_type: str) -> None:
        if rcv_type.lower() not in ["pr", "vx", "vz"]:
            raise ValueError("Receiver type must be either pr, vx, vz")
        self.loc_x.append(rcv_x)
        self.loc_z.append(rcv_z)
        self.type.append(rcv_type)
        self.num += 1

    def get_loc(self):
        loc = np.hstack((list2numpy(self.loc_x).reshape(-1, 1), list2numpy(self.loc_z).reshape(-1, 1)))
        self.loc = loc.copy()
        return loc

    def get_type(self, unique=False):
        return list2numpy(list(set(self.type))) if unique else list2numpy(self.type)


class Survey:
    """Sources plus shared receivers."""

    def __init__(self, source: Source, receiver: Receiver) -> None:
        self.source = source
        self.receiver = receiver

    def __repr__(self):
        return "Survey Information:\n" + repr(self.source) + "\n" + repr(self.receiver)


from .data import SeismicData  # noqa: E402  (uses Survey above)
