"""ElasticPropagator: drop-in for ADFWI/propagator/elastic_propagator.py:152-261.

Same constructor, attributes (src_n, bcx, bcz, damp, wavelet, moment_tensor, rcv_*),
forward(model, shot_index, fd_order, checkpoint_segments) and record dictionary; the
time stepping and its adjoint run in libadfwi_amd.so.  Models (isotropic or
anisotropic, the reference's or adfwi_amd's) are duck-typed.
"""
from collections import OrderedDict
from typing import Optional

import numpy as np
import torch

from .. import _native as nat
from ..utils import numpy2tensor
from .acoustic_propagator import _check_duck, shot_subset
from .boundary_condition import bc_gerjan, bc_pml_xz, bc_sincos
from .elastic_kernels import DiffCoef, ElasticGeometry, forward_kernel

_MODEL_ATTRS = ("ox", "oz", "dx", "dz", "nx", "nz", "abc_type", "nabc", "free_surface", "vp", "forward")


class ElasticPropagator(torch.nn.Module):
    """Elastic velocity-stress staggered-grid propagator (4th order, split-field PML) on the MI355X."""

    def __init__(self, model, survey, device: Optional[str] = "cpu", cpu_num: Optional[int] = 1,
                 gpu_num: Optional[int] = 1, dtype=torch.float32):
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
        # vertical-force sources ("vf", an extension of the reference's moment-tensor sources)
        types = np.array([str(t).lower() for t in self.source.get_type()])
        self.force_weight = (torch.tensor((types == "vf").astype(np.float32), device=self.device)
                             if np.any(types == "vf") else None)
        self.receiver = survey.receiver
        self.rcv_loc = self.receiver.get_loc()
        self.rcv_x = numpy2tensor(self.rcv_loc[:, 0], torch.long).to(self.device)
        self.rcv_z = numpy2tensor(self.rcv_loc[:, 1], torch.long).to(self.device)
        self.rcv_n = self.receiver.num
        self._geoms = OrderedDict()
        self.ckpt_every = 0
        # checkpoint segments whose per-step states the forward keeps for the backward (None: as
        # many as the free device memory allows; 0: recompute every segment)
        self.resident_images = None

    def boundary_condition(self):
        """Damping profiles (elastic_propagator.py:99-114); 'PML' is case-sensitive as there."""
        if self.abc_type == "PML":
            bcx, bcz = bc_pml_xz(self.nx, self.nz, self.dx, self.dz, pml=self.nabc,
                                 vmax=self.model.vp.cpu().detach().numpy().max(), free_surface=self.free_surface)
            self.bcx = numpy2tensor(bcx, self.dtype).to(self.device)
            self.bcz = numpy2tensor(bcz, self.dtype).to(self.device)
        elif self.abc_type == "gerjan":
            damp = bc_gerjan(self.nx, self.nz, self.dx, self.dz, pml=self.nabc, alpha=self.model.abc_jerjan_alpha,
                             free_surface=self.free_surface)
            self.damp = numpy2tensor(damp, self.dtype).to(self.device)
        else:
            damp = bc_sincos(self.nx, self.nz, self.dx, self.dz, pml=self.nabc, free_surface=self.free_surface)
            self.damp = numpy2tensor(damp, self.dtype).to(self.device)

    def _geometry(self, shots, aniso, damping, fd_order=4):
        key = (tuple(int(s) for s in shots), bool(aniso), bool(damping), int(fd_order))
        g = self._geoms.get(key)
        if g is None:
            fso = fd_order // 2
            Z = self.nz + (self.nabc + fso if self.free_surface else 2 * self.nabc + fso)
            X = self.nx + 2 * self.nabc
            zoff = fso if self.free_surface else self.nabc + fso
            sl = self.src_loc[shots]
            mt = self.moment_tensor[torch.as_tensor(shots, device=self.moment_tensor.device)]
            if damping:   # -MT/3 (elastic_kernels.py:703)
                src_m = torch.stack([(-1 / 3) * mt[:, 0, 0], (-1 / 3) * mt[:, 2, 2], (-1 / 3) * mt[:, 0, 2]], dim=1)
            else:         # -MT/2 into each split (elastic_kernels.py:358-372)
                half = mt / 2
                src_m = torch.stack([-half[:, 0, 0], -half[:, 2, 2], -half[:, 0, 2]], dim=1)
            fz = None if self.force_weight is None else self.force_weight[torch.as_tensor(shots, device=self.device)]
            g = ElasticGeometry(Z, X, self.nt, self.free_surface, sl[:, 1] + zoff, sl[:, 0] + self.nabc, src_m,
                                self.rcv_loc[:, 1] + zoff, self.rcv_loc[:, 0] + self.nabc, self.device, aniso=aniso,
                                damping=damping, src_fz=fz, nn=fso,
                                c2=float(DiffCoef(fso, "s")[2]) if fso > 2 else 0.0,
                                resident_traces=self.resident_images)
            self._geoms[key] = g
            while len(self._geoms) > 16:
                self._geoms.popitem(last=False)
        return g

    def forward(self, model=None, shot_index=None, fd_order=4, checkpoint_segments=1):
        """Forward simulation of the selected shots (elastic_propagator.py:116-142)."""
        model = self.model if model is None else model
        model.forward()
        shots = shot_subset(shot_index, self.src_n)
        sel = torch.as_tensor(shots, device=self.src_x.device)
        src_x, src_z = self.src_x[sel], self.src_z[sel]
        wavelet = self.wavelet[sel]
        mt = self.moment_tensor[sel]
        CC = model.CC
        aniso = bool((CC[4] != 0).any() or (CC[13] != 0).any())
        geom = None
        if (self.abc_type == "PML") == (self.abc_type.lower() == "pml") and fd_order in (4, 6, 8, 10):
            # the reference's propagator builds bcx/bcz only for 'PML' exactly (elastic_propagator.py
            # :219) while its driver takes the PML branch for any case of 'pml' (elastic_kernels.py:948)
            geom = self._geometry(shots, aniso, damping=self.abc_type != "PML", fd_order=fd_order)
        return forward_kernel(self.nx, self.nz, self.dx, self.dz, self.nt, self.dt, self.nabc, self.free_surface,
                              src_x, src_z, len(shots), wavelet, mt, self.rcv_x, self.rcv_z, self.rcv_n,
                              self.abc_type, self.bcx, self.bcz, self.damp, model.lamu, model.lam,
                              model.bx, model.bz, CC, fd_order=fd_order, n_segments=checkpoint_segments,
                              device=self.device, dtype=self.dtype, geometry=geom, ckpt_every=self.ckpt_every,
                              src_fz=None if self.force_weight is None else self.force_weight[sel])
