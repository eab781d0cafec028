"""Pins of the oracle's record and gradient post-processing (no GPU): brute-force loops on
tiny inputs, known answers worked by hand, invariances that a dropped term would break, and
finite differences of the autograd gradient."""
import numpy as np
import pytest
import torch

from oracle import gradient_process as ogp
from oracle import misfit as om


def _loops_misfit(rec, obs, dt, normalize):
    S, nt, R = rec.shape
    total = 0.0
    for s in range(S):
        for r in range(R):
            peak = max(abs(float(rec[s, t, r])) for t in range(nt)) if normalize else 1.0
            acc = 0.0
            for t in range(nt):
                d = float(rec[s, t, r]) / peak - float(obs[s, t, r])
                acc += d * d * dt
            total += acc ** 0.5
    return total


@pytest.mark.parametrize("normalize", [False, True])
def test_misfit_equals_loops(normalize):
    rng = np.random.default_rng(0)
    rec = rng.standard_normal((2, 9, 3))
    obs = rng.standard_normal((2, 9, 3)) * 0.3
    got = float(om.driver_misfit(torch.tensor(rec), torch.tensor(obs), 0.004, normalize))
    assert abs(got - _loops_misfit(rec, obs, 0.004, normalize)) <= 1e-12 * abs(got)


def test_least_squares_misfit_equals_loops_and_half_the_squared_roots():
    """waveform_l2_squared: brute-force loops; and on every trace, half the square of
    waveform_l2's per-trace root (so a dropped 1/2 or dt, or a sum over the wrong axis, fails)."""
    rng = np.random.default_rng(3)
    rec = rng.standard_normal((2, 7, 4))
    obs = rng.standard_normal((2, 7, 4)) * 0.5
    dt = 0.003
    got = float(om.waveform_l2_squared(torch.tensor(obs), torch.tensor(rec), dt))
    loops = sum(0.5 * sum((obs[s, t, r] - rec[s, t, r]) ** 2 * dt for t in range(7))
                for s in range(2) for r in range(4))
    assert abs(got - loops) <= 1e-12 * abs(loops)
    roots = [float(om.waveform_l2(torch.tensor(obs[s:s + 1, :, r:r + 1]), torch.tensor(rec[s:s + 1, :, r:r + 1]), dt))
             for s in range(2) for r in range(4)]
    assert abs(got - 0.5 * sum(x * x for x in roots)) <= 1e-12 * abs(got)
    # gradient: d/d syn = -(obs - syn) dt (known answer)
    syn = torch.tensor(rec, requires_grad=True)
    om.waveform_l2_squared(torch.tensor(obs), syn, dt).backward()
    assert np.allclose(syn.grad.numpy(), -(obs - rec) * dt, rtol=1e-12, atol=0)


def test_misfit_known_answer():
    # one trace [1, -2, 0.5] normalised by 2 against zeros: sum of squares 0.25 + 1 + 0.0625
    rec = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64).view(1, 3, 1)
    got = float(om.driver_misfit(rec, torch.zeros_like(rec), 0.5, True))
    assert abs(got - (1.3125 * 0.5) ** 0.5) < 1e-15


def test_normalised_misfit_is_scale_invariant_and_its_gradient_orthogonal_to_each_trace():
    """loss(c x) = loss(x) for c > 0 per trace, so d/dc at c = 1 vanishes: sum_t g[t] x[t] = 0 for
    every trace; dropping the normalisation's max term breaks this."""
    rng = np.random.default_rng(1)
    rec = torch.tensor(rng.standard_normal((3, 40, 5)), requires_grad=True)
    obs = om.normalize_traces(torch.tensor(rng.standard_normal((3, 40, 5))))
    loss = om.driver_misfit(rec, obs, 0.002, True)
    loss.backward()
    scale = torch.tensor(rng.uniform(0.5, 3.0, (3, 1, 5)))
    assert abs(float(om.driver_misfit(rec.detach() * scale, obs, 0.002, True)) - loss.item()) < 1e-12
    proj = (rec.grad * rec.detach()).sum(dim=1)
    assert float(proj.abs().max()) < 1e-12 * float(rec.grad.abs().max())


@pytest.mark.parametrize("normalize", [False, True])
def test_misfit_gradient_matches_central_differences(normalize):
    rng = np.random.default_rng(2)
    rec = torch.tensor(rng.standard_normal((2, 12, 3)), requires_grad=True)
    obs = torch.tensor(rng.standard_normal((2, 12, 3)) * 0.2)
    om.driver_misfit(rec, obs, 0.01, normalize).backward()
    d = torch.tensor(rng.standard_normal((2, 12, 3)))
    h = 1e-6
    fd = (float(om.driver_misfit(rec.detach() + h * d, obs, 0.01, normalize))
          - float(om.driver_misfit(rec.detach() - h * d, obs, 0.01, normalize))) / (2 * h)
    assert abs(fd - float((rec.grad * d).sum())) <= 1e-6 * abs(fd)


def _loops_smooth(z, span):
    f = ogp.gauss_kernel(span)
    k = f.shape[0] // 2
    nz, nx = z.shape
    num = np.zeros((nz, nx))
    den = np.zeros((nz, nx))
    for i in range(nz):
        for j in range(nx):
            for a in range(-k, k + 1):
                for b in range(-k, k + 1):
                    ii, jj = i - a, j - b
                    if 0 <= ii < nz and 0 <= jj < nx:
                        num[i, j] += f[a + k, b + k] * z[ii, jj]
                        den[i, j] += f[a + k, b + k]
    return num / den


def test_smooth2d_equals_loops_and_keeps_constants():
    rng = np.random.default_rng(3)
    z = rng.standard_normal((7, 9))
    assert np.allclose(ogp.smooth2d(z, 2), _loops_smooth(z, 2), rtol=0, atol=1e-13)
    assert np.allclose(ogp.smooth2d(np.full((11, 13), 3.5), 4), 3.5, rtol=0, atol=1e-13)
    f = ogp.gauss_kernel(3)   # separable: rank one, symmetric, unit sum
    assert np.linalg.matrix_rank(f, tol=1e-14) == 1 and np.allclose(f, f.T) and abs(f.sum() - 1) < 1e-14


def test_grad_taper_marine_rows_and_land_columns():
    from scipy.signal import windows
    assert np.allclose(windows.hamming(4), [0.08, 0.77, 0.77, 0.08])
    m = ogp.grad_taper(30, 40, tapersize=6, marine_or_land="Marine")
    assert np.all(m[:6] == 0) and np.all(m[6:] == 1)
    land = ogp.grad_taper(30, 40, tapersize=6, thred=0.001, marine_or_land="land")
    assert np.allclose(land, land[0:1, :], atol=1e-12)          # every row alike: the columns are tapered
    assert land[0, 0] < 0.01 and abs(land[0, -1] - 1.0) < 1e-6 and np.all(np.diff(land[0]) >= -1e-12)
    assert np.allclose(ogp.grad_taper(30, 40, 6, 0.0, "marine"), ogp.grad_taper(30, 40, 6, 0.0, "land"))


def test_grad_processor_known_answers():
    nz, nx = 30, 50
    gp = ogp.GradProcessor(grad_mute=0, grad_smooth=0, norm_grad=False, forw_illumination=True)
    g = gp.forward(nx, nz, 3000.0, np.ones((nz, nx)), forw=np.ones((nz, nx)))
    assert np.allclose(g, (1.0 + 1e-5) ** 2, rtol=1e-13)        # precondition 1 / (1 + 1e-5) everywhere
    g = ogp.GradProcessor(norm_grad=True, forw_illumination=False).forward(
        nx, nz, 2500.0, np.random.default_rng(4).standard_normal((nz, nx)))
    assert abs(np.abs(g).max() - 2500.0) < 1e-9
