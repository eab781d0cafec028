"""Velocity models feeding the propagators (ADFWI/model/{base,acoustic_model,elastic_model}.py).

``forward()`` re-derives the coefficient tensors exactly as the reference does
(clamping in place, Gardner density from *detached* vp, Lame/Thomsen moduli and
staggered averages), so autograd carries kernel-coefficient gradients back to
the model parameters.  The propagators duck-type any object with these
attributes, so the reference's own model classes (and DIP models) work too.
"""
from typing import Optional, Tuple, Union

import numpy as np
import torch
from torch import Tensor

from ..utils import gpu2cpu, numpy2tensor
from .parameters import (elastic_moduli_for_isotropic, elastic_moduli_for_TI, elastic_moduli_init,
                         parameter_staggered_grid, thomsen_to_elastic_moduli, vs_vp_to_Lame)

_EPS = 1e-7


class AbstractModel(torch.nn.Module):
    """Grid, boundary settings and per-parameter bounds / gradient flags (base.py:35-349)."""

    def __init__(self, ox: float, oz: float, nx: int, nz: int, dx: float, dz: float,
                 free_surface: Optional[bool] = False, abc_type: Optional[str] = "PML",
                 abc_jerjan_alpha: Optional[float] = 0.0053, nabc: Optional[int] = 20,
                 device="cpu", dtype=torch.float32) -> None:
        super().__init__()
        self.ox, self.oz, self.dx, self.dz, self.nx, self.nz = ox, oz, dx, dz, nx, nz
        self.free_surface, self.abc_type, self.abc_jerjan_alpha = free_surface, abc_type, abc_jerjan_alpha
        self.nabc, self.device, self.dtype = nabc, device, dtype
        assert self.dx == self.dz, "Model grid size dx and dz must be the same"
        self.x = np.arange(nx) * self.dx + self.ox
        self.z = np.arange(nz) * self.dz + self.oz
        self.pars, self.requires_grad, self.lower_bound, self.upper_bound = [], {}, {}, {}

    def forward(self, *args, **kwargs):
        raise NotImplementedError("Forward method must be implemented by the subclass")

    def _set_bounds(self, bounds):
        for par, b in bounds.items():
            self.lower_bound[par] = b[0] if b is not None else None
            self.upper_bound[par] = b[1] if b is not None else None

    def _check_bounds(self):
        for par in self.pars:
            lo, hi = self.lower_bound[par], self.upper_bound[par]
            if lo is not None and hi is not None:
                assert lo < hi, "Lower bound must be smaller than upper bound"
            if lo is not None and lo + _EPS > self.get_model(par).min():
                self.lower_bound[par] = self.get_model(par).min() - _EPS
            if hi is not None and hi - _EPS < self.get_model(par).max():
                self.upper_bound[par] = self.get_model(par).max() + _EPS

    def check_dims(self) -> None:
        for par in self.pars:
            assert self.get_model(par).shape == (self.nz, self.nx), "Model dimensions must be (nz, nx)"

    def get_model(self, par: str) -> np.ndarray:
        if par not in self.pars:
            raise ValueError(f"Parameter {par} not in model")
        return gpu2cpu(getattr(self, par).clone())

    def set_model(self, par: str, model) -> None:
        if par not in self.pars:
            raise ValueError(f"Parameter {par} not in model")
        if model.shape != (self.nz, self.nx):
            raise ValueError("Model dimensions must be (nz, nx)")
        m = numpy2tensor(gpu2cpu(model))
        setattr(self, par, torch.nn.Parameter(m.to(torch.float32), requires_grad=self.requires_grad[par]))

    def get_bound(self, par: str):
        if par not in self.pars:
            raise ValueError(f"Parameter {par} not in model")
        lo, hi = self.lower_bound[par], self.upper_bound[par]
        return [None, None] if lo is None or hi is None else [lo, hi]

    def get_requires_grad(self, par: str) -> bool:
        if par not in self.pars:
            raise ValueError(f"Parameter {par} not in model")
        return self.requires_grad[par]

    def get_grad(self, par: str) -> np.ndarray:
        if par not in self.pars:
            raise ValueError(f"Parameter {par} not in model")
        m = getattr(self, par)
        return np.zeros(m.shape) if m.grad is None else m.grad.cpu().detach().numpy()

    def get_clone_data(self):
        kw = {}
        for par in self.pars:
            kw[par] = self.get_model(par)
            kw[par + "_bound"] = self.get_bound(par)
            kw[par + "_grad"] = self.get_requires_grad(par)
        kw.update(ox=self.ox, oz=self.oz, dx=self.dx, dz=self.dz, nx=self.nx, nz=self.nz,
                  free_surface=self.free_surface, nabc=self.nabc)
        return kw

    def save(self, filename: str) -> None:
        """The clone data as an .npz (base.py:284-296) with plain arrays only: a bound is a
        float pair, NaN for "no bound", so `load_model_data` reads it without unpickling."""
        kw = self.get_clone_data()
        for par in self.pars:
            lo, hi = kw[par + "_bound"]
            kw[par + "_bound"] = np.array([np.nan if lo is None else lo, np.nan if hi is None else hi],
                                          dtype=np.float64)
        np.savez(filename, **kw)

    def clip_params(self) -> None:
        """In-place clamp of every bounded parameter (base.py:298-308)."""
        for par in self.pars:
            lo, hi = self.lower_bound[par], self.upper_bound[par]
            if lo is not None and hi is not None:
                getattr(self, par).data.clamp_(lo, hi)

    def constrain_range(self, value, min_value, max_value) -> Tensor:
        """Clamped copy; raises ValueError on inf/negative/NaN (base.py:310-350)."""
        if min_value is None or max_value is None:
            return value
        if torch.isnan(value).any():
            value[torch.isnan(value)] = max_value
        value = torch.clamp(value, min_value, max_value)
        if torch.isinf(value).any():
            raise ValueError("Value is inf")
        if (value < 0.0).any():
            raise ValueError("Value is negative")
        if torch.isnan(value).any():
            raise ValueError("Value is nan")
        return value

    def _as_parameter(self, a, flag):
        return torch.nn.Parameter(numpy2tensor(a, self.dtype).to(self.device), requires_grad=flag)

    def _gardner_rho(self):
        """rho = 310 vp^0.25 from detached vp (acoustic_model.py:111-118)."""
        rho = np.power(self.vp.cpu().detach().numpy(), 0.25) * 310
        self.rho = torch.nn.Parameter(numpy2tensor(rho, self.dtype).to(self.device), requires_grad=self.rho_grad)


class AcousticModel(AbstractModel):
    """vp / rho model (acoustic_model.py:18-134).  forward() clamps vp and resets rho by Gardner."""

    def __init__(self, ox: float, oz: float, nx: int, nz: int, dx: float, dz: float,
                 vp=None, rho=None, vp_bound: Optional[Tuple[float, float]] = None,
                 rho_bound: Optional[Tuple[float, float]] = None, vp_grad: Optional[bool] = False,
                 rho_grad: Optional[bool] = False, free_surface: Optional[bool] = False,
                 abc_type: Optional[str] = "PML", abc_jerjan_alpha: Optional[float] = 0.0053,
                 nabc: Optional[int] = 20, device="cpu", dtype=torch.float32) -> None:
        super().__init__(ox, oz, nx, nz, dx, dz, free_surface, abc_type, abc_jerjan_alpha, nabc, device, dtype)
        self.pars = ["vp", "rho"]
        self.vp_grad, self.rho_grad = vp_grad, rho_grad
        self.vp = self._as_parameter(vp.copy(), vp_grad)
        self.rho = self._as_parameter(rho.copy(), rho_grad)
        self._set_bounds({"vp": vp_bound, "rho": rho_bound})
        self.requires_grad["vp"], self.requires_grad["rho"] = vp_grad, rho_grad
        self._check_bounds()
        self.check_dims()

    def set_rho_using_empirical_function(self):
        self._gardner_rho()

    def forward(self):
        self.clip_params()
        self.constrain_range(self.vp, self.lower_bound["vp"], self.upper_bound["vp"])
        self.set_rho_using_empirical_function()
        self.constrain_range(self.rho, self.lower_bound["rho"], self.upper_bound["rho"])


class _ElasticBase(AbstractModel):
    _thomsen_pars = ("vp", "vs", "rho", "eps", "gamma", "delta")

    def _init_elastic(self, vals, grads, bounds, auto_update_rho, auto_update_vp):
        for par in self._thomsen_pars:
            setattr(self, par + "_grad", grads.get(par, False))
            setattr(self, par, self._as_parameter(np.array(vals[par], copy=True), grads.get(par, False)))
        self.mu = self.lamu = self.lam = self.muxz = self.b = self.bx = self.bz = None
        self._parameterization_Lame()
        self.CC = []
        self._parameterization_elastic_moduli()
        self._set_bounds({p: bounds.get(p) for p in self.pars})
        for p in self.pars:
            self.requires_grad[p] = grads.get(p, False)
        self._check_bounds()
        self.check_dims()
        self.auto_update_rho, self.auto_update_vp = auto_update_rho, auto_update_vp

    def _parameterization_Lame(self):
        self.mu, self.lamu, self.lam, self.b = vs_vp_to_Lame(self.vp, self.vs, self.rho)

    def _fill(self, CC):
        raise NotImplementedError

    def _parameterization_elastic_moduli(self):
        CC = elastic_moduli_init(self.nz, self.nx, self.device, self.dtype)
        C11, C13, C33, C44, C66 = thomsen_to_elastic_moduli(self.vp, self.vs, self.rho, self.eps, self.delta,
                                                            self.gamma)
        CC[0], CC[2], CC[11], CC[15], CC[20] = C11, C13, C33, C44, C66
        CC = self._fill(CC)
        bx, bz, muxz, C44s, C55s, C66s = parameter_staggered_grid(self.mu, self.b, CC[15], CC[18], CC[20],
                                                                  self.nx, self.nz)
        self.bx, self.bz, self.muxz = bx, bz, muxz
        CC[15], CC[18], CC[20] = C44s, C55s, C66s
        self.CC = CC

    def set_rho_using_empirical_function(self):
        self._gardner_rho()

    def set_vp_using_empirical_function(self):
        vp = self.vs.cpu().detach().numpy() * np.sqrt(3)
        self.vp = torch.nn.Parameter(numpy2tensor(vp, self.dtype).to(self.device), requires_grad=self.vp_grad)

    def forward(self):
        """elastic_model.py:217-234 / :459-476."""
        if self.auto_update_rho:
            self.set_rho_using_empirical_function()
        if self.auto_update_vp:
            self.set_vp_using_empirical_function()
        self.clip_params()
        self._parameterization_Lame()
        self._parameterization_elastic_moduli()


class IsotropicElasticModel(_ElasticBase):
    """vp / vs / rho isotropic model (elastic_model.py:22-234); Gardner rho by default."""

    def __init__(self, ox: float, oz: float, nx: int, nz: int, dx: float, dz: float,
                 vp=None, vs=None, rho=None, vp_bound=None, vs_bound=None, rho_bound=None,
                 vp_grad=False, vs_grad=False, rho_grad=False, free_surface=False, abc_type="PML",
                 abc_jerjan_alpha=0.0053, nabc=20, auto_update_rho=True, auto_update_vp=False,
                 device="cpu", dtype=torch.float32) -> None:
        super().__init__(ox, oz, nx, nz, dx, dz, free_surface, abc_type, abc_jerjan_alpha, nabc, device, dtype)
        self.pars = ["vp", "vs", "rho"]
        z = np.zeros((nz, nx))
        self._init_elastic(dict(vp=vp, vs=vs, rho=rho, eps=z, gamma=z, delta=z),
                           dict(vp=vp_grad, vs=vs_grad, rho=rho_grad),
                           dict(vp=vp_bound, vs=vs_bound, rho=rho_bound), auto_update_rho, auto_update_vp)

    def _fill(self, CC):
        return elastic_moduli_for_isotropic(CC)


class AnisotropicElasticModel(_ElasticBase):
    """Thomsen-parameterised TI model (elastic_model.py:237-476); no Gardner rho by default."""

    def __init__(self, ox: float, oz: float, nx: int, nz: int, dx: float, dz: float,
                 vp=None, vs=None, rho=None, eps=None, gamma=None, delta=None,
                 vp_bound=None, vs_bound=None, rho_bound=None, eps_bound=None, gamma_bound=None,
                 delta_bound=None, vp_grad=False, vs_grad=False, rho_grad=False, eps_grad=False,
                 gamma_grad=False, delta_grad=False, free_surface=False, auto_update_rho=False,
                 auto_update_vp=False, anisotropic_type="vti", abc_type="PML", abc_jerjan_alpha=0.0053,
                 nabc=20, device="cpu", dtype=torch.float32) -> None:
        super().__init__(ox, oz, nx, nz, dx, dz, free_surface, abc_type, abc_jerjan_alpha, nabc, device, dtype)
        self.pars = ["vp", "vs", "rho", "eps", "gamma", "delta"]
        self.anisotropic_type = anisotropic_type
        self._init_elastic(dict(vp=vp, vs=vs, rho=rho, eps=eps, gamma=gamma, delta=delta),
                           dict(vp=vp_grad, vs=vs_grad, rho=rho_grad, eps=eps_grad, gamma=gamma_grad,
                                delta=delta_grad),
                           dict(vp=vp_bound, vs=vs_bound, rho=rho_bound, eps=eps_bound, gamma=gamma_bound,
                                delta=delta_bound), auto_update_rho, auto_update_vp)

    def _fill(self, CC):
        return elastic_moduli_for_TI(CC, anisotropic_type=self.anisotropic_type)


__all__ = ["AbstractModel", "AcousticModel", "IsotropicElasticModel", "AnisotropicElasticModel", "load_model_data"]


def load_model_data(filename: str) -> dict:
    """Keyword arguments of a model saved by `save` (read with allow_pickle=False): every
    parameter array, its bound ([None, None] when unbounded) and grad flag, and the grid."""
    out = {}
    with np.load(filename, allow_pickle=False) as f:
        for k in f.files:
            v = f[k]
            if k.endswith("_bound"):
                lo, hi = (float(x) for x in v)
                out[k] = [None, None] if np.isnan(lo) or np.isnan(hi) else [lo, hi]
            elif v.ndim == 0:
                out[k] = v.item()
            else:
                out[k] = v
    return out
