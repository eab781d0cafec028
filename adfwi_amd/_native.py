"""ctypes binding of libadfwi_amd.so (declared in include/adfwi_amd.h).

The library is loaded from this package directory; there is no CPU fallback.
When it is missing or cannot be loaded, every propagator call raises
``NativeLibraryError`` with the reason.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadfwi_amd.so")
ABI_VERSION = 3
# floats of the library's per-GPU scratch array (ADFWI_NOALLOC_SCRATCH_FLOATS, 128 MiB)
NOALLOC_SCRATCH_FLOATS = 32 << 20


class NativeLibraryError(RuntimeError):
    """libadfwi_amd.so is missing, unloadable, or returned an error code."""


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class AcousticDesc(ctypes.Structure):
    _fields_ = [("abi_version", _i32), ("struct_size", _i32), ("nshots", _i32), ("nz", _i32),
                ("nx", _i32), ("pitch", _i32), ("fs", _i32), ("free_surface", _i32), ("nt", _i32),
                ("nrec", _i32), ("shots_per_block", _i32), ("row_unroll", _i32), ("tile_z", _i32), ("adj_shots_per_block", _i32),
                ("coef", _vp), ("src", _vp), ("src_z", _vp), ("src_x", _vp), ("rcv_z", _vp),
                ("rcv_x", _vp), ("coef_set", _vp)]


class Layout(ctypes.Structure):
    _fields_ = [("tile_z", _i32), ("tile_x", _i32), ("zbase", _i32), ("ntiles_z", _i32),
                ("ntiles_x", _i32), ("ntiles", _i32), ("shots_per_block", _i32), ("ngroups", _i32),
                ("nfields", _i32), ("reserved", _i32), ("field_floats", _i64), ("state_floats", _i64),
                ("accum_floats", _i64)]


class AcousticFwdArgs(ctypes.Structure):
    _fields_ = [("state_in", _vp), ("work", _vp * 2), ("ckpt", _vp), ("ckpt_every", _i32),
                ("t0", _i32), ("nsteps", _i32), ("energy_u_from", _i32),
                ("rec_p", _vp), ("rec_u", _vp), ("rec_w", _vp), ("energy", _vp),
                ("image", _vp), ("image_p", _vp), ("state_out", _vp), ("rcv_map", _vp),
                ("fuse_steps", _i32), ("reserved2", _i32), ("t_base", _vp)]


class AcousticAdjArgs(ctypes.Structure):
    _fields_ = [("lam_in", _vp), ("work", _vp * 2), ("t_hi", _i32), ("nsteps", _i32),
                ("image_t0", _i32), ("reserved", _i32), ("image", _vp), ("image_p", _vp),
                ("grad_rec_p", _vp), ("grad_rec_u", _vp), ("grad_rec_w", _vp),
                ("grad_a1", _vp), ("grad_a2", _vp), ("grad_src", _vp), ("lam_out", _vp),
                ("rcv_map", _vp), ("fuse_steps", _i32), ("reserved2", _i32)]


_f32 = ctypes.c_float


class ElasticDesc(ctypes.Structure):
    _fields_ = [("abi_version", _i32), ("struct_size", _i32), ("nshots", _i32), ("nz", _i32), ("nx", _i32),
                ("pitch", _i32), ("free_surface", _i32), ("nt", _i32), ("nrec", _i32), ("aniso", _i32),
                ("shots_per_wave", _i32), ("tile_z", _i32), ("dt", _f32), ("dt_dx", _f32), ("dt_dz", _f32),
                ("dx", _f32), ("dz", _f32), ("c0", _f32), ("c1", _f32), ("damping", _i32),
                ("coef", _vp), ("pml_x", _vp), ("pml_z", _vp), ("src", _vp), ("src_m", _vp),
                ("src_z", _vp), ("src_x", _vp), ("rcv_z", _vp), ("rcv_x", _vp), ("damp", _vp), ("src_fz", _vp),
                ("nn", _i32), ("c2", _f32), ("rcv_map", _vp)]


class ElasticFwdArgs(ctypes.Structure):
    _fields_ = [("state_in", _vp), ("work", _vp * 2), ("ckpt", _vp), ("ckpt_every", _i32), ("t0", _i32),
                ("nsteps", _i32), ("reserved", _i32), ("rec", _vp * 5), ("trace", _vp), ("state_out", _vp)]


class ElasticAdjArgs(ctypes.Structure):
    _fields_ = [("lam_in", _vp), ("work", _vp * 2), ("t_hi", _i32), ("nsteps", _i32), ("fwd0", _vp),
                ("trace", _vp), ("grad_rec", _vp * 5), ("grad_coef", _vp), ("grad_src", _vp), ("lam_out", _vp)]


# function name -> (restype, argtypes)
_SIGNATURES = {
    "adfwi_abi_version": (ctypes.c_int, []),
    "adfwi_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "adfwi_last_error": (ctypes.c_char_p, []),
    "adfwi_acoustic_get_layout": (ctypes.c_int, [ctypes.POINTER(AcousticDesc), ctypes.POINTER(Layout)]),
    "adfwi_acoustic_forward": (ctypes.c_int, [ctypes.POINTER(AcousticDesc),
                                              ctypes.POINTER(AcousticFwdArgs), _vp]),
    "adfwi_acoustic_adjoint": (ctypes.c_int, [ctypes.POINTER(AcousticDesc),
                                              ctypes.POINTER(AcousticAdjArgs), _vp]),
    "adfwi_acoustic_adjoint_noalloc": (ctypes.c_int, [ctypes.POINTER(AcousticDesc),
                                                      ctypes.POINTER(AcousticAdjArgs), _vp, _vp, _vp]),
    "adfwi_elastic_get_layout": (ctypes.c_int, [ctypes.POINTER(ElasticDesc), ctypes.POINTER(Layout)]),
    "adfwi_elastic_forward": (ctypes.c_int, [ctypes.POINTER(ElasticDesc), ctypes.POINTER(ElasticFwdArgs), _vp]),
    "adfwi_elastic_adjoint": (ctypes.c_int, [ctypes.POINTER(ElasticDesc), ctypes.POINTER(ElasticAdjArgs), _vp]),
    "adfwi_trace_l2": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, ctypes.c_float, _i32, _vp, _vp, _vp, _vp]),
}

_lock = threading.Lock()
_lib = None
_load_error = None


def load(path: str = LIB_PATH):
    """Loads (once) and returns the library handle; raises NativeLibraryError if unavailable."""
    global _lib, _load_error
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise NativeLibraryError(
                f"{path} is not built; run `python -m adfwi_amd.build` (hipcc, gfx950)")
        try:
            lib = ctypes.CDLL(path)
        except OSError as e:
            _load_error = str(e)
            raise NativeLibraryError(f"cannot load {path}: {e}") from e
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.adfwi_abi_version() != ABI_VERSION:
            raise NativeLibraryError(f"{path} has ABI {lib.adfwi_abi_version()}, expected {ABI_VERSION}")
        _lib = lib
        return lib


def exported_symbols():
    """Names this binding resolves (every entry point declared in include/adfwi_amd.h)."""
    return list(_SIGNATURES)


def check(rc: int, what: str):
    if rc != 0:
        lib = load()
        msg = lib.adfwi_last_error().decode(errors="replace")
        kind = lib.adfwi_strerror(rc).decode(errors="replace")
        if rc == -2:
            raise ValueError(f"{what}: {kind}: {msg}")
        raise NativeLibraryError(f"{what} failed ({rc}, {kind}): {msg}")


def ptr(t):
    """Device/host address of a tensor, or None."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu(device):
    """The HIP library is the only compute path: refuse non-GPU tensors loudly."""
    dev = torch.device(device)
    if dev.type != "cuda":
        raise NativeLibraryError(
            f"adfwi_amd propagators run on the MI355X through libadfwi_amd.so; device {dev} is not a GPU "
            "(there is no CPU fallback)")
    if not torch.cuda.is_available():
        raise NativeLibraryError("no ROCm GPU is visible to torch; adfwi_amd has no CPU fallback")
    load()
