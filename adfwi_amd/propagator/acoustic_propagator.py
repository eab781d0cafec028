"""AcousticPropagator: drop-in for ADFWI/propagator/acoustic_propagator.py:21-157.

Same constructor, attributes (src_n, damp, wavelet, rcv_*, ...), forward()
signature and record dictionary; the time stepping runs in libadfwi_amd.so.
Models and surveys are duck-typed: the reference's classes, DIP models and the
classes of ``adfwi_amd.model`` / ``adfwi_amd.survey`` all work.
"""
from collections import OrderedDict
from typing import Dict, Optional

import numpy as np
import torch
from torch import Tensor

from .. import _native as nat
from ..utils import numpy2tensor
from .acoustic_kernels import AcousticGeometry, forward_kernel
from .boundary_condition import bc_gerjan, bc_pml, bc_sincos

_MODEL_ATTRS = ("ox", "oz", "dx", "dz", "nx", "nz", "abc_type", "nabc", "free_surface", "vp", "rho", "forward")


def _check_duck(model, survey, attrs):
    missing = [a for a in attrs if not hasattr(model, a)]
    if missing:
        raise ValueError(f"model is not an instance of AbstractModel (missing {missing})")
    if not (hasattr(survey, "source") and hasattr(survey, "receiver")):
        raise ValueError("survey is not an instance of Survey")


def shot_subset(index, n):
    """Normalises shot_index (None, int, list, ndarray, tensor) to a sorted-as-given int64 array."""
    if index is None:
        return np.arange(n)
    if torch.is_tensor(index):
        index = index.cpu().numpy()
    return np.atleast_1d(np.asarray(index, dtype=np.int64))


class AcousticPropagator(torch.nn.Module):
    """Isotropic acoustic (p-u-w) staggered-grid propagator on the MI355X.

    Parameters: model, survey, device (a ROCm GPU, e.g. 'cuda:0'), cpu_num and
    gpu_num (kept for signature compatibility, unused as in the reference),
    dtype (torch.float32).
    """

    def __init__(self, model, survey, device: Optional[str] = "cpu", cpu_num: Optional[int] = 1,
                 gpu_num: Optional[int] = 1, dtype: torch.dtype = torch.float32):
        super().__init__()
        _check_duck(model, survey, _MODEL_ATTRS)
        nat.require_gpu(device)
        if dtype != torch.float32:
            raise NotImplementedError("the HIP propagator computes in float32")
        self.model, self.survey = model, survey
        self.device, self.dtype = device, dtype
        self.cpu_num, self.gpu_num = cpu_num, gpu_num
        self.ox, self.oz = model.ox, model.oz
        self.dx, self.dz = model.dx, model.dz
        self.nx, self.nz = model.nx, model.nz
        self.nt = survey.source.nt
        self.dt = survey.source.dt
        self.f0 = survey.source.f0
        self.abc_type = model.abc_type
        self.nabc = model.nabc
        self.free_surface = model.free_surface
        self.bcx, self.bcz, self.damp = None, None, None
        self.boundary_condition()
        self.source = survey.source
        self.src_loc = self.source.get_loc()
        self.src_x = numpy2tensor(self.src_loc[:, 0], torch.long).to(self.device)
        self.src_z = numpy2tensor(self.src_loc[:, 1], torch.long).to(self.device)
        self.src_n = self.source.num
        self.wavelet = numpy2tensor(self.source.get_wavelet(), self.dtype).to(self.device)
        self.moment_tensor = numpy2tensor(self.source.get_moment_tensor(), self.dtype).to(self.device)
        self.receiver = survey.receiver
        self.rcv_loc = self.receiver.get_loc()
        self.rcv_x = numpy2tensor(self.rcv_loc[:, 0], torch.long).to(self.device)
        self.rcv_z = numpy2tensor(self.rcv_loc[:, 1], torch.long).to(self.device)
        self.rcv_n = self.receiver.num
        self._geoms = OrderedDict()
        self.ckpt_every = 0  # 0 = ceil(sqrt(nt)) wavefield checkpoint spacing
        # forward runs without gradient replay HIP graphs of this many steps (None: the
        # ADFWI_GRAPH_STEPS environment variable, default 0 = direct launches)
        self.graph_steps = None
        # checkpoint segments whose stencil images the forward keeps for the backward (None: as
        # many as the free device memory allows; 0: recompute every segment)
        self.resident_images = None

    def boundary_condition(self, vmax=None):
        """Damping profile (acoustic_propagator.py:102-118); PML vmax frozen at construction."""
        kind = self.abc_type.lower()
        if kind == "pml":
            if vmax is None:
                vmax = self.model.vp.cpu().detach().numpy().max()
            damp = bc_pml(self.nx, self.nz, self.dx, self.dz, pml=self.nabc, vmax=vmax, free_surface=False)
        elif kind == "gerjan":
            damp = bc_gerjan(self.nx, self.nz, self.dx, self.dz, pml=self.nabc,
                             alpha=self.model.abc_jerjan_alpha, free_surface=False)
        else:
            damp = bc_sincos(self.nx, self.nz, self.dx, self.dz, pml=self.nabc, free_surface=False)
        self.damp = numpy2tensor(damp, self.dtype).to(self.device)

    def _geometry(self, shots: np.ndarray) -> AcousticGeometry:
        key = tuple(int(s) for s in shots)
        g = self._geoms.get(key)
        if g is None:
            sl = self.src_loc[shots]
            nabc = self.nabc
            g = AcousticGeometry(self.nz + 2 * nabc, self.nx + 2 * nabc, self.nt, self.free_surface, nabc,
                                 sl[:, 1] + nabc, sl[:, 0] + nabc, self.rcv_loc[:, 1] + nabc,
                                 self.rcv_loc[:, 0] + nabc, self.device, graph_steps=self.graph_steps,
                                 resident_images=self.resident_images)
            self._geoms[key] = g
            while len(self._geoms) > 16:
                self._geoms.popitem(last=False)
        return g

    def forward(self, model=None, shot_index: Optional[int] = None,
                checkpoint_segments: int = 1) -> Dict[str, Tensor]:
        """Forward simulation of the selected shots (acoustic_propagator.py:120-157)."""
        model = self.model if model is None else model
        model.forward()
        shots = shot_subset(shot_index, self.src_n)
        sel = torch.as_tensor(shots, device=self.src_x.device)
        src_x, src_z = self.src_x[sel], self.src_z[sel]
        wavelet = self.wavelet[sel]
        return forward_kernel(self.nx, self.nz, self.dx, self.dz, self.nt, self.dt, self.nabc,
                              self.free_surface, src_x, src_z, len(shots), wavelet,
                              self.rcv_x, self.rcv_z, self.rcv_n, self.damp, model.vp, model.rho,
                              checkpoint_segments=checkpoint_segments, device=self.device, dtype=self.dtype,
                              geometry=self._geometry(shots), ckpt_every=self.ckpt_every)
