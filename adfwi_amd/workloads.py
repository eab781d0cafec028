"""Seeded synthetic workloads used by the benchmark, the serving loop and the tests.

Each recipe returns plain numpy arrays (models, geometry, wavelets); nothing
here touches the GPU.  Grids are model grids (the propagator pads them).

acoustic_small  100 x 200, dx 10 m, 20-cell PML, dt 1 ms, nt 1000, 10 shots (10 Hz),
                200 receivers: five flat layers 1500->4000 m/s + one +-300 m/s Gaussian
                anomaly (sigma 8 cells); initial model = true model smoothed (sigma 10).
acoustic_large  384 x 1152, dx 12.5 m, 40-cell PML, dt 1 ms, nt 4000, 64 shots (8 Hz,
                spacing 18 cells), 1152 receivers: 1500->4500 m/s gradient + Gaussian random
                field (corr. 20 cells, std 300) + two dipping faulted interfaces, clipped to
                [1500, 4700]; initial model smoothed (sigma 10).
elastic_iso     200 x 600, dx 5 m, free surface, 40-cell PML, dt 0.5 ms, nt 4000, 32 shots.
elastic_vti     320 x 960, dx 5 m, free surface, 40-cell PML, dt 0.4 ms, nt 5000, 48 shots.
"""
from dataclasses import dataclass, field
from typing import Dict

import numpy as np
from scipy.ndimage import gaussian_filter


def ricker(nt, dt, f0, t0=None):
    t = np.arange(nt) * dt
    t0 = 1.2 / f0 if t0 is None else t0
    tau = (np.pi * f0) ** 2
    return (1 - 2 * tau * (t - t0) ** 2) * np.exp(-tau * (t - t0) ** 2)


def gardner(vp):
    return (np.power(vp.astype(np.float32), 0.25) * 310).astype(np.float32)


@dataclass
class Workload:
    name: str
    nz: int
    nx: int
    dx: float
    nabc: int
    dt: float
    nt: int
    f0: float
    free_surface: bool
    src_x: np.ndarray
    src_z: np.ndarray
    rcv_x: np.ndarray
    rcv_z: np.ndarray
    models: Dict[str, np.ndarray] = field(default_factory=dict)

    @property
    def nshots(self):
        return len(self.src_x)

    @property
    def padded(self):
        """Padded acoustic grid (Z, X)."""
        return self.nz + 2 * self.nabc, self.nx + 2 * self.nabc

    @property
    def padded_elastic(self):
        """Padded 4th-order elastic grid (Z, X): two free-surface rows on top."""
        return self.padded_elastic_order(4)

    def padded_elastic_order(self, fd_order: int):
        """Padded elastic grid (Z, X) of an fd_order: fd_order // 2 free-surface rows on top."""
        nn = fd_order // 2
        top = self.nabc + nn if self.free_surface else 2 * self.nabc + nn
        return self.nz + top, self.nx + 2 * self.nabc

    def wavelets(self):
        return np.tile(ricker(self.nt, self.dt, self.f0)[None, :], (self.nshots, 1)).astype(np.float32)


def _grf(rng, nz, nx, corr, std):
    a = gaussian_filter(rng.standard_normal((nz, nx)), corr, mode="reflect")
    return a / a.std() * std


def acoustic_small(seed: int = 0, nt: int = 1000) -> Workload:
    rng = np.random.default_rng(seed)
    nz, nx = 100, 200
    depths = np.sort(rng.choice(np.arange(10, 90), size=4, replace=False))
    speeds = np.linspace(1500, 4000, 5)
    vp = np.empty((nz, nx))
    edges = np.concatenate([[0], depths, [nz]])
    for k in range(5):
        vp[edges[k]:edges[k + 1], :] = speeds[k]
    cz, cx = rng.integers(30, 70), rng.integers(50, 150)
    amp = 300.0 * rng.choice([-1.0, 1.0])
    zz, xx = np.mgrid[0:nz, 0:nx]
    vp = vp + amp * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 8.0 ** 2))
    vp_true = vp.astype(np.float32)
    vp_init = gaussian_filter(vp_true, 10, mode="reflect").astype(np.float32)
    w = Workload("acoustic_small", nz, nx, 10.0, 20, 1e-3, nt, 10.0, False,
                 src_x=10 + 20 * np.arange(10), src_z=np.full(10, 2),
                 rcv_x=np.arange(nx), rcv_z=np.full(nx, 2))
    w.models = dict(vp_true=vp_true, vp_init=vp_init, rho_true=gardner(vp_true), rho_init=gardner(vp_init))
    return w


def acoustic_large(seed: int = 1, nt: int = 4000, nshots: int = 64) -> Workload:
    rng = np.random.default_rng(seed)
    nz, nx = 384, 1152
    zz, xx = np.mgrid[0:nz, 0:nx]
    vp = 1500.0 + 3000.0 * zz / (nz - 1)
    vp = vp + _grf(rng, nz, nx, 20, 300.0)
    for _ in range(2):
        dip = np.deg2rad(rng.uniform(5, 25))
        z0 = rng.uniform(0.3, 0.7) * nz
        xf = rng.uniform(0.3, 0.7) * nx
        throw = rng.uniform(5, 20)
        surf = z0 + np.tan(dip) * (xx - nx / 2) + np.where(xx > xf, throw, 0.0)
        vp = vp + np.where(zz > surf, 250.0, 0.0)
    vp_true = np.clip(vp, 1500, 4700).astype(np.float32)
    vp_init = gaussian_filter(vp_true, 10, mode="reflect").astype(np.float32)
    w = Workload("acoustic_large", nz, nx, 12.5, 40, 1e-3, nt, 8.0, False,
                 src_x=9 + 18 * np.arange(nshots), src_z=np.full(nshots, 2),
                 rcv_x=np.arange(nx), rcv_z=np.full(nx, 2))
    w.models = dict(vp_true=vp_true, vp_init=vp_init, rho_true=gardner(vp_true), rho_init=gardner(vp_init))
    return w


def elastic_layers(rng, nz, nx, nlayers, lo, hi, grf_std):
    depths = np.sort(rng.choice(np.arange(nz // 10, nz - nz // 10), size=nlayers - 1, replace=False))
    edges = np.concatenate([[0], depths, [nz]])
    vp = np.empty((nz, nx))
    for k, v in enumerate(np.linspace(lo, hi, nlayers)):
        vp[edges[k]:edges[k + 1], :] = v
    vp = vp + _grf(rng, nz, nx, 20, grf_std)
    vs = vp / 1.73 * (0.95 + 0.1 * (lambda a: (a - a.min()) / (a.max() - a.min()))(
        gaussian_filter(rng.standard_normal((nz, nx)), 10)))
    return vp.astype(np.float32), vs.astype(np.float32)


def elastic_iso(seed: int = 2, nt: int = 4000, nshots: int = 32) -> Workload:
    rng = np.random.default_rng(seed)
    nz, nx = 200, 600
    vp, vs = elastic_layers(rng, nz, nx, 6, 1800, 4500, 200.0)
    w = Workload("elastic_iso", nz, nx, 5.0, 40, 5e-4, nt, 8.0, True,
                 src_x=np.linspace(10, nx - 10, nshots).astype(int), src_z=np.full(nshots, 4),
                 rcv_x=np.arange(nx), rcv_z=np.full(nx, 2))
    w.models = dict(vp=vp, vs=vs, rho=gardner(vp))
    return w


def elastic_vti(seed: int = 3, nt: int = 5000, nshots: int = 48) -> Workload:
    rng = np.random.default_rng(seed)
    nz, nx = 320, 960
    vp, vs = elastic_layers(rng, nz, nx, 6, 1800, 4500, 200.0)
    unit = lambda a: (a - a.min()) / (a.max() - a.min())
    eps = 0.25 * unit(gaussian_filter(rng.standard_normal((nz, nx)), 30))
    delta = -0.05 + 0.2 * unit(gaussian_filter(rng.standard_normal((nz, nx)), 30))
    w = Workload("elastic_vti", nz, nx, 5.0, 40, 4e-4, nt, 10.0, True,
                 src_x=np.linspace(10, nx - 10, nshots).astype(int), src_z=np.full(nshots, 4),
                 rcv_x=np.arange(nx), rcv_z=np.full(nx, 2))
    w.models = dict(vp=vp, vs=vs, rho=gardner(vp), eps=eps.astype(np.float32), delta=delta.astype(np.float32))
    return w
