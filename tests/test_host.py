"""CPU tests of the product's host side: damping profiles, model parameterisation, the C ABI
library (loads, exports every declared entry point, host-only layout / binning / errors)."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import torch

from conftest import ROOT, golden

LIB = os.path.join(ROOT, "adfwi_amd", "libadfwi_amd.so")
HEADER = os.path.join(ROOT, "include", "adfwi_amd.h")


# ---------------------------------------------------------------- profiles
@pytest.mark.parametrize("fs", [False, True])
@pytest.mark.parametrize("shape", [(24, 16, 6), (37, 21, 9), (200, 100, 20)])
def test_product_profiles_equal_oracle_loops(fs, shape):
    from adfwi_amd.propagator import boundary_condition as bc
    from oracle import boundary as ob
    nx, nz, pml = shape
    assert np.array_equal(bc.bc_pml(nx, nz, 10.0, 10.0, pml, 3512.5, fs), ob.bc_pml(nx, nz, 10.0, 10.0, pml, 3512.5, fs))
    for a, b in zip(bc.bc_pml_xz(nx, nz, 5.0, 5.0, pml, 4100.0, fs), ob.bc_pml_xz(nx, nz, 5.0, 5.0, pml, 4100.0, fs)):
        assert np.array_equal(a, b)
    assert np.array_equal(bc.bc_sincos(nx, nz, 10.0, 10.0, pml, fs), ob.bc_sincos(nx, nz, 10.0, 10.0, pml, fs))
    assert np.array_equal(bc.bc_gerjan(nx, nz, 10.0, 10.0, pml, 0.0053, fs),
                          ob.bc_gerjan(nx, nz, 10.0, 10.0, pml, 0.0053, fs))


def test_product_profiles_equal_reference_vectors():
    from adfwi_amd.propagator import boundary_condition as bc
    g = golden("boundary")
    assert np.array_equal(bc.bc_pml(24, 16, 10.0, 10.0, 6, 3000.0, False), g["pml_nofs"])
    assert np.array_equal(bc.bc_pml_xz(24, 16, 5.0, 5.0, 6, 3000.0, True)[0], g["pmlxz_fs1_x"])


# ------------------------------------------------------------------ models
def test_fd_coefficients_equal_the_reference_solve():
    """The product's DiffCoef(n, 's') reproduces the reference's float32 solve on this host
    (n = 2..5, the stencils of fd_order 4..10) and the staggered-grid closed forms."""
    from adfwi_amd.propagator.elastic_kernels import DiffCoef
    assert np.array_equal(DiffCoef(2, "s").numpy(), golden("diffcoef")["fdc"])
    for n in (3, 4, 5):
        assert np.array_equal(DiffCoef(n, "s").numpy(), golden("diffcoef6")[f"fdc{2 * n}"]), n
    assert np.allclose(DiffCoef(2, "s").numpy(), [9 / 8, -1 / 24], rtol=1e-6)
    assert np.allclose(DiffCoef(3, "s").numpy(), [75 / 64, -25 / 384, 3 / 640], rtol=3e-5)


def test_wavelet_closed_form():
    from adfwi_amd.utils import wavelet
    t, w = wavelet(500, 1e-3, 10.0)
    t0, tau = 0.12, (np.pi * 10.0) ** 2
    np.testing.assert_allclose(w, (1 - 2 * tau * (t - t0) ** 2) * np.exp(-tau * (t - t0) ** 2), rtol=1e-14)
    assert np.argmax(w) == 120 and abs(w.max() - 1.0) < 1e-12


def test_acoustic_model_forward_gardner_and_clip():
    from adfwi_amd.model import AcousticModel
    rng = np.random.default_rng(0)
    vp = rng.uniform(1700, 3800, (10, 12)).astype(np.float32)
    m = AcousticModel(0, 0, 12, 10, 10.0, 10.0, vp=vp, rho=np.ones_like(vp), vp_bound=(1600.0, 3900.0),
                      vp_grad=True)
    with torch.no_grad():                      # an optimiser step leaves the bounds
        m.vp[0, 0] = 1000.0
        m.vp[1, 1] = 5000.0
    m.forward()
    v = m.vp.detach().numpy()
    assert v.min() >= 1600.0 - 1e-3 and v.max() <= 3900.0 + 1e-3       # clamped in place
    np.testing.assert_array_equal(m.rho.detach().numpy(), (np.power(v, 0.25) * 310).astype(np.float32))
    assert not m.rho.requires_grad and m.vp.requires_grad


def test_elastic_parameterisation_matches_oracle_and_closed_forms():
    from adfwi_amd.model import AnisotropicElasticModel, IsotropicElasticModel
    from oracle import model as om
    rng = np.random.default_rng(1)
    nz, nx = 9, 11
    vp = rng.uniform(2000, 3000, (nz, nx)); vs = vp / 1.73; rho = rng.uniform(1800, 2400, (nz, nx))
    eps = rng.uniform(0, 0.25, (nz, nx)); delta = rng.uniform(-0.05, 0.15, (nz, nx)); gamma = np.zeros((nz, nx))
    m = AnisotropicElasticModel(0, 0, nx, nz, 5.0, 5.0, vp=vp, vs=vs, rho=rho, eps=eps, gamma=gamma, delta=delta)
    m.forward()
    T = lambda a: torch.tensor(a, dtype=torch.float32)
    CC, bx, bz, lamu, lam = om.elastic_cc(T(vp), T(vs), T(rho), T(eps), T(delta), T(gamma), nx, nz)
    for k in (0, 2, 11, 18):
        assert torch.equal(m.CC[k], CC[k]), k
    assert torch.equal(m.bx, bx) and torch.equal(m.bz, bz)
    # closed forms: C33 = vp^2 rho, C11 = C33 (1 + 2 eps), C13 from delta
    c33 = vp.astype(np.float32) ** 2 * rho.astype(np.float32)
    np.testing.assert_allclose(m.CC[11].numpy(), c33, rtol=1e-6)
    np.testing.assert_allclose(m.CC[0].numpy(), c33 * (1 + 2 * eps), rtol=1e-5)
    c44 = vs ** 2 * rho
    np.testing.assert_allclose(m.CC[2].numpy(), np.sqrt(2 * c33 * (c33 - c44) * delta + (c33 - c44) ** 2) - c44,
                               rtol=1e-4)
    iso = IsotropicElasticModel(0, 0, nx, nz, 5.0, 5.0, vp=vp, vs=vs, rho=rho, auto_update_rho=False)
    iso.forward()
    mu = (vs ** 2 * rho)
    five = 0.2 * (mu[1:-1, 1:-1] + 2 * mu[2:, 1:-1] + mu[1:-1, 2:] + mu[2:, 2:])
    np.testing.assert_allclose(iso.CC[18].numpy(), five, rtol=1e-5)
    np.testing.assert_allclose(iso.CC[2].numpy(), vp ** 2 * rho - 2 * mu, rtol=1e-4)


# --------------------------------------------------------------------- ABI
def _declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(adfwi_\w+)\s*\(", text, re.M)))


def test_library_is_built_and_exports_every_declared_symbol():
    assert os.path.exists(LIB), "run python -m adfwi_amd.build first"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (adfwi_\w+)", out))
    declared = _declared_entry_points()
    assert declared and set(declared) <= exported, set(declared) - exported
    from adfwi_amd import _native as nat
    assert set(declared) == set(nat.exported_symbols())


def test_library_loads_without_gpu_and_reports_version():
    from adfwi_amd import _native as nat
    lib = nat.load()
    assert lib.adfwi_abi_version() == nat.ABI_VERSION
    assert lib.adfwi_strerror(-2) == b"invalid argument"


def _desc(S=3, Z=60, X=150, fs=1, free_surface=0, nt=10, nrec=5):
    from adfwi_amd import _native as nat
    d = nat.AcousticDesc()
    d.abi_version, d.struct_size = nat.ABI_VERSION, ctypes.sizeof(nat.AcousticDesc)
    d.nshots, d.nz, d.nx, d.pitch = S, Z, X, -(-X // 64) * 64
    d.fs, d.free_surface, d.nt, d.nrec = fs, free_surface, nt, nrec
    return d


def test_layout_host_side():
    from adfwi_amd import _native as nat
    lib = nat.load()
    d = _desc(free_surface=1, fs=20)
    L = nat.Layout()
    assert lib.adfwi_acoustic_get_layout(ctypes.byref(d), ctypes.byref(L)) == 0
    assert L.zbase == 19 and L.ntiles == L.ntiles_x * L.ntiles_z
    assert L.ntiles_x * L.tile_x >= 150 and L.zbase + L.ntiles_z * L.tile_z >= 60
    assert L.ngroups * L.shots_per_block >= 3 and (L.ngroups - 1) * L.shots_per_block < 3
    assert L.state_floats == 3 * 3 * 60 * d.pitch
    d.shots_per_block = 2
    assert lib.adfwi_acoustic_get_layout(ctypes.byref(d), ctypes.byref(L)) == 0 and L.ngroups == 2


def test_abi_rejects_bad_arguments_with_messages():
    from adfwi_amd import _native as nat
    lib = nat.load()
    L = nat.Layout()
    d = _desc()
    d.struct_size = 8
    assert lib.adfwi_acoustic_get_layout(ctypes.byref(d), ctypes.byref(L)) == -1
    assert b"size" in lib.adfwi_last_error()
    d = _desc(S=0)
    assert lib.adfwi_acoustic_get_layout(ctypes.byref(d), ctypes.byref(L)) == -2
    d = _desc()
    a = nat.AcousticFwdArgs()   # null buffers: rejected before any device access
    assert lib.adfwi_acoustic_forward(ctypes.byref(d), ctypes.byref(a), None) == -2
    with pytest.raises(ValueError):
        nat.check(-2, "test")


def test_noalloc_adjoint_rejects_before_any_device_access():
    """adfwi_acoustic_adjoint_noalloc checks its arguments and the size of the gradient planes
    against its scratch array before it touches a device (these calls run without a GPU)."""
    from adfwi_amd import _native as nat
    lib = nat.load()
    fake = ctypes.c_void_p(0x1000)
    # planes of one gradient: (shot groups) x Z x pitch floats; 64 single-shot groups of
    # 2000 x 2048 exceed the scratch array, one group of 60 x 192 fits
    d = _desc(S=64, Z=2000, X=2000)
    d.shots_per_block, d.adj_shots_per_block = 1, 1
    a = nat.AcousticAdjArgs()
    a.image = fake.value
    assert lib.adfwi_acoustic_adjoint_noalloc(ctypes.byref(d), ctypes.byref(a), fake, None, None) == -2
    assert b"scratch" in lib.adfwi_last_error() and b"maximum" in lib.adfwi_last_error()
    L = nat.Layout()
    assert lib.adfwi_acoustic_get_layout(ctypes.byref(d), ctypes.byref(L)) == 0
    assert L.accum_floats > nat.NOALLOC_SCRATCH_FLOATS
    a.grad_a1 = fake.value          # the planes are the library's: caller planes are refused
    assert lib.adfwi_acoustic_adjoint_noalloc(ctypes.byref(_desc()), ctypes.byref(a), None, None, None) == -2
    assert b"must be NULL" in lib.adfwi_last_error()
    a.grad_a1, a.image = None, None  # a gradient sum without the stencil image is refused
    assert lib.adfwi_acoustic_adjoint_noalloc(ctypes.byref(_desc()), ctypes.byref(a), fake, None, None) == -2
    assert b"stencil image" in lib.adfwi_last_error()
    assert "#define ADFWI_NOALLOC_SCRATCH_FLOATS (32ll << 20)" in open(HEADER).read()
    assert nat.NOALLOC_SCRATCH_FLOATS == 32 << 20


def test_fault_injection_surfaces_as_error():
    code = ("import ctypes; from adfwi_amd import _native as nat; lib = nat.load(); "
            "print(lib.adfwi_acoustic_forward(None, None, None))")
    env = dict(os.environ, ADFWI_FAULT_INJECT="adfwi_acoustic_forward")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=ROOT)
    assert out.stdout.strip() == "-4", out.stderr


def test_product_refuses_cpu_device():
    from adfwi_amd import _native as nat
    with pytest.raises(nat.NativeLibraryError):
        nat.require_gpu("cpu")


def test_ctypes_structures_match_the_c_header(tmp_path):
    """Every ctypes mirror in adfwi_amd/_native.py has the size and field offsets of its struct
    in include/adfwi_amd.h (compiled with the host C compiler; field names must match)."""
    import os
    import shutil
    import subprocess
    from adfwi_amd import _native as nat
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    pairs = {"AcousticDesc": "adfwi_acoustic_desc", "Layout": "adfwi_layout",
             "AcousticFwdArgs": "adfwi_acoustic_fwd_args", "AcousticAdjArgs": "adfwi_acoustic_adj_args",
             "ElasticDesc": "adfwi_elastic_desc", "ElasticFwdArgs": "adfwi_elastic_fwd_args",
             "ElasticAdjArgs": "adfwi_elastic_adj_args"}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "adfwi_amd.h"', "int main(void) {"]
    expect = []
    for py, c in pairs.items():
        cls = getattr(nat, py)
        lines.append(f'printf("{py} sizeof %zu\\n", sizeof({c}));')
        expect.append(f"{py} sizeof {ctypes.sizeof(cls)}")
        for f in cls._fields_:
            lines.append(f'printf("{py} {f[0]} %zu\\n", offsetof({c}, {f[0]}));')
            expect.append(f"{py} {f[0]} {getattr(cls, f[0]).offset}")
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    assert [g for g in got if g] == expect


def test_resident_segment_counts(monkeypatch):
    """How many checkpoint segments the forward keeps resident: capped by the geometry's setting,
    by the number of segments and by the memory budget (here set through the environment), and
    none when no image is needed."""
    from types import SimpleNamespace
    from adfwi_amd.propagator.acoustic_kernels import resident_segments
    from adfwi_amd.propagator.elastic_kernels import resident_trace_segments
    k, S, Z, pitch, nt = 8, 2, 10, 64, 61                    # 8 segments, the last 5 steps
    seg = k * S * Z * pitch * 4
    monkeypatch.setenv("ADFWI_IMAGE_BUDGET_BYTES", str(3 * seg + 5))
    g = SimpleNamespace(nt=nt, S=S, Z=Z, pitch=pitch, resident_images=None, device="cpu")
    assert resident_segments(g, k, True, False) == 3
    assert resident_segments(g, k, True, True) == 1          # two planes per step
    assert resident_segments(g, k, False, False) == 0
    g.resident_images = 2
    assert resident_segments(g, k, True, False) == 2
    g.resident_images = 0
    assert resident_segments(g, k, True, False) == 0
    g.resident_images = None
    monkeypatch.setenv("ADFWI_IMAGE_BUDGET_BYTES", str(100 * seg))
    assert resident_segments(g, k, True, False) == 8         # every segment
    e = SimpleNamespace(nt=nt, resident_traces=None, device="cpu", layout=SimpleNamespace(state_floats=1000))
    monkeypatch.setenv("ADFWI_IMAGE_BUDGET_BYTES", str(2 * k * 1000 * 4))
    assert resident_trace_segments(e, k) == 2
    e.resident_traces = 1
    assert resident_trace_segments(e, k) == 1


def test_forward_state_lookup_finds_work_checkpoint_and_trace_slots():
    """The elastic forward driver maps the state pointer a call ends on back to its tensor: a
    work buffer, a checkpoint slot or a trace slot (a call can end on a checkpoint step)."""
    from adfwi_amd.propagator.elastic_kernels import _state_at
    work = torch.zeros((2, 3, 4, 5))
    ckpt = torch.zeros((4, 3, 4, 5))
    trace = torch.zeros((7, 3, 4, 5))
    assert _state_at(work[1].data_ptr(), work, [ckpt, trace]).data_ptr() == work[1].data_ptr()
    assert _state_at(ckpt[2].data_ptr(), work, [ckpt, trace]).data_ptr() == ckpt[2].data_ptr()
    assert _state_at(trace[6].data_ptr(), work, [ckpt, trace]).data_ptr() == trace[6].data_ptr()
    with pytest.raises(RuntimeError):
        _state_at(trace[6].data_ptr() + 4, work, [ckpt, trace])


def test_pending_backward_reserves_are_counted_until_released():
    """A forward pass that keeps resident segments sets its backward's reserve aside for later
    budgets until its resident set is released (after its backward, or when the graph is freed)."""
    import gc
    from adfwi_amd.propagator.acoustic_kernels import ResidentSet
    base = ResidentSet.pending_bytes
    a = ResidentSet({0: None}).hold(1000)
    b = ResidentSet().hold(24)
    assert ResidentSet.pending_bytes == base + 1024
    del a
    gc.collect()
    assert ResidentSet.pending_bytes == base + 24
    b.pop(0, None)
    del b
    gc.collect()
    assert ResidentSet.pending_bytes == base


def test_integration_stub_matches_the_binding():
    """The ctypes stub a maintainer copies from the integration notes declares the same descriptor
    fields and ABI version as the binding the package uses (and so as include/adfwi_amd.h)."""
    import ctypes as ct
    from adfwi_amd import _native as nat
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"class AcousticDesc\(ctypes\.Structure\):\s*_fields_ = (\[.*?\])\n", text, re.S)
    assert m, "stub not found"
    fields = eval(m.group(1), {"_i32": ct.c_int32, "_vp": ct.c_void_p})
    assert [f[0] for f in fields] == [f[0] for f in nat.AcousticDesc._fields_]
    assert [f[1] for f in fields] == [f[1] for f in nat.AcousticDesc._fields_]
    m = re.search(r"class ElasticDesc\(ctypes\.Structure\):\s*_fields_ = (\[.*?\])\n", text, re.S)
    assert m, "elastic stub not found"
    fields = eval(m.group(1), {"_i32": ct.c_int32, "_vp": ct.c_void_p, "_f32": ct.c_float})
    assert [f[0] for f in fields] == [f[0] for f in nat.ElasticDesc._fields_]
    assert [f[1] for f in fields] == [f[1] for f in nat.ElasticDesc._fields_]
    v = re.search(r"AcousticDesc\(abi_version=(\d+)", text)
    assert v and int(v.group(1)) == nat.ABI_VERSION
    hdr = open(os.path.join(ROOT, "include", "adfwi_amd.h")).read()
    assert int(re.search(r"#define ADFWI_ABI_VERSION (\d+)", hdr).group(1)) == nat.ABI_VERSION


def test_receiver_maps_list_rows_then_indices_and_refuse_repeated_points(monkeypatch):
    """The receiver map both geometries hand to the step kernels: [Z] receivers per row, then
    [Z][pitch] receiver index (-1 elsewhere); no map when a point repeats (its records are then
    taken by separate launches), when there are no receivers, or, elastic, when switched off."""
    from adfwi_amd.propagator.acoustic_kernels import AcousticGeometry
    from adfwi_amd.propagator.elastic_kernels import ElasticGeometry
    Z, X, pitch = 12, 20, 32
    rz = np.array([3, 3, 7, 9], dtype=np.int64)
    rx = np.array([2, 5, 5, 17], dtype=np.int64)
    for cls in (AcousticGeometry, ElasticGeometry):
        g = object.__new__(cls)          # the map builder only, without a device
        g.Z, g.X, g.pitch, g.R, g.device = Z, X, pitch, len(rz), torch.device("cpu")
        g.free_surface, g.fs = False, 2
        m = g._receiver_map(rz, rx).numpy()
        assert m.shape == (Z + Z * pitch,)
        assert list(m[:Z]) == [0, 0, 0, 2, 0, 0, 0, 1, 0, 1, 0, 0]
        idx = m[Z:].reshape(Z, pitch)
        for r, (z, x) in enumerate(zip(rz, rx)):
            assert idx[z, x] == r
        assert (idx >= 0).sum() == len(rz)
        assert g._receiver_map(np.array([3, 3]), np.array([5, 5])) is None     # a repeated point
        g.R = 0
        assert g._receiver_map(rz[:0], rx[:0]) is None
    g = object.__new__(ElasticGeometry)
    g.Z, g.X, g.pitch, g.R, g.device = Z, X, pitch, len(rz), torch.device("cpu")
    monkeypatch.setenv("ADFWI_EL_FUSED_RECORDS", "0")
    assert g._receiver_map(rz, rx) is None


def test_elastic_layout_refuses_gradient_planes_over_2_gib():
    """The elastic adjoint addresses a workgroup's gradient planes through one buffer descriptor
    whose lanes without a point of their own aim past the planes, so one set of planes must stay
    under 2 GiB: a single-shot 8200 x 8200 grid (its state still under the descriptor's 4 GiB)
    is refused with a message, a 4000 x 4000 one accepted (host-side, no GPU)."""
    from adfwi_amd import _native as nat
    lib = nat.load()

    def desc(nz, nx):
        d = nat.ElasticDesc()
        d.abi_version, d.struct_size = nat.ABI_VERSION, ctypes.sizeof(nat.ElasticDesc)
        d.nshots, d.nz, d.nx, d.pitch = 1, nz, nx, -(-nx // 64) * 64
        d.nt, d.nrec = 10, 0
        return d
    L = nat.Layout()
    assert lib.adfwi_elastic_get_layout(ctypes.byref(desc(4000, 4000)), ctypes.byref(L)) == 0
    assert lib.adfwi_elastic_get_layout(ctypes.byref(desc(8200, 8200)), ctypes.byref(L)) == -2
    assert b"gradient planes" in lib.adfwi_last_error()
