"""ctypes binding of libkwavehip.so (the C ABI declared in include/kwave_hip.h).

The library is built in-tree (`make -C k-wave-swift_amd` or `__graft_entry__.build()`); importing
this module fails loudly if it is missing -- there is no CPU fallback for the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os

ABI_VERSION = 2
BATCH_GRAPH = 1    # kwh_batch_run mode flags (KWH_BATCH_GRAPH / KWH_BATCH_FUSED / KWH_BATCH_PLANE)
BATCH_FUSED = 2
BATCH_PLANE = 4

OK, EINVAL, ENOMEM, EDEVICE, ENONFINITE, EUNSUPPORTED, ESTATE = 0, -1, -2, -3, -4, -5, -6

(OUT_P, OUT_P_MAX, OUT_P_MIN, OUT_P_RMS, OUT_P_FINAL, OUT_UX, OUT_UY, OUT_UZ, OUT_UX_MAX, OUT_UY_MAX,
 OUT_UZ_MAX, OUT_UX_RMS, OUT_UY_RMS, OUT_UZ_RMS, OUT_UX_FINAL, OUT_UY_FINAL, OUT_UZ_FINAL, OUT_IX_AVG,
 OUT_IY_AVG, OUT_IZ_AVG) = range(20)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkwavehip.so")

# Every symbol include/kwave_hip.h declares.
EXPORTED = [
    "kwh_abi_version", "kwh_last_error", "kwh_sim_validate", "kwh_ctx_create", "kwh_ctx_destroy", "kwh_sim_create",
    "kwh_sim_create_axisymmetric",
    "kwh_sim_run", "kwh_sim_launch", "kwh_sim_sync", "kwh_sim_output_shape", "kwh_sim_read_output",
    "kwh_sim_steps_done", "kwh_sim_state_bytes", "kwh_sim_save_state", "kwh_sim_load_state",
    "kwh_sim_set_kernel_timing", "kwh_sim_set_kernel_timing_only", "kwh_sim_kernel_stats", "kwh_sim_set_graph",
    "kwh_sim_set_plane",
    "kwh_sim_element_traces", "kwh_sim_autotune", "kwh_sim_set_checkpointing", "kwh_sim_vjp", "kwh_sim_destroy",
    "kwh_batch_create", "kwh_batch_run", "kwh_batch_sim", "kwh_batch_read_output", "kwh_batch_destroy",
    "kwh_time_reversal", "kwh_sim_set_nonfinite_check", "kwh_sim_read_output_noalloc",
    "kwh_h5_available", "kwh_h5_write_input", "kwh_h5_read_input", "kwh_h5_input_desc", "kwh_h5_input_destroy",
    "kwh_h5_write_output", "kwh_h5_read_dataset", "kwh_h5_read_root_attribute",
]


class Field(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("count", C.c_int64)]


class SourceField(C.Structure):
    _fields_ = [("index", C.POINTER(C.c_int64)), ("n_points", C.c_int64), ("signal", C.POINTER(C.c_float)),
                ("rows", C.c_int32), ("len", C.c_int32), ("mode", C.c_int32),
                ("delay", C.POINTER(C.c_int32)), ("scale", C.POINTER(C.c_float))]


class SimDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32),
        ("dim", C.c_int32), ("n", C.c_int32 * 3), ("d", C.c_double * 3), ("dt", C.c_double),
        ("nt", C.c_int32), ("dtype", C.c_int32),
        ("pml_size", C.c_int32), ("pml_alpha", C.c_double),
        ("smooth_p0", C.c_int32), ("smooth_c0", C.c_int32), ("smooth_rho0", C.c_int32),
        ("sound_speed", Field), ("density", Field), ("alpha_coeff", Field), ("alpha_power", C.c_double),
        ("alpha_mode", C.c_int32), ("bona", Field),
        ("p0", Field), ("p_source", SourceField), ("u_source", SourceField * 3),
        ("sensor_index", C.POINTER(C.c_int64)), ("n_sensor", C.c_int64),
        ("sensor_cartesian", C.POINTER(C.c_float)), ("n_cartesian", C.c_int64),
        ("record", C.c_uint32), ("directivity_angle", C.POINTER(C.c_float)),
        ("directivity_size", C.c_double), ("directivity_pattern", C.c_int32),
    ]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes_per_launch", C.c_double)]


STEP_CB = C.CFUNCTYPE(None, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p)

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the library (once). Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libkwavehip.so not built at {path}; run `make -C k-wave-swift_amd`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.kwh_abi_version.restype = C.c_int
    lib.kwh_last_error.restype = C.c_char_p
    lib.kwh_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.kwh_ctx_destroy.argtypes = [vp]
    lib.kwh_ctx_destroy.restype = None
    lib.kwh_sim_validate.argtypes = [C.POINTER(SimDesc)]
    lib.kwh_sim_create.argtypes = [vp, C.POINTER(SimDesc), C.POINTER(vp)]
    lib.kwh_sim_create_axisymmetric.argtypes = [vp, C.POINTER(SimDesc), C.POINTER(vp)]
    lib.kwh_sim_run.argtypes = [vp, C.c_int32, C.c_int32, STEP_CB, C.c_int32, vp]
    lib.kwh_sim_launch.argtypes = [vp, C.c_int32, C.c_int32]
    lib.kwh_sim_sync.argtypes = [vp]
    lib.kwh_sim_output_shape.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
    lib.kwh_sim_read_output.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    lib.kwh_sim_read_output_noalloc.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    lib.kwh_sim_steps_done.argtypes = [vp, C.POINTER(C.c_int32)]
    lib.kwh_sim_state_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.kwh_sim_save_state.argtypes = [vp, vp, C.c_size_t]
    lib.kwh_sim_load_state.argtypes = [vp, vp, C.c_size_t]
    lib.kwh_sim_set_kernel_timing.argtypes = [vp, C.c_int]
    lib.kwh_sim_set_kernel_timing_only.argtypes = [vp, C.c_char_p]
    lib.kwh_sim_kernel_stats.argtypes = [vp, C.POINTER(KernelStat), C.c_int32, C.POINTER(C.c_int32)]
    lib.kwh_sim_set_graph.argtypes = [vp, C.c_int]
    lib.kwh_sim_set_plane.argtypes = [vp, C.c_int]
    lib.kwh_sim_element_traces.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int32,
                                           C.c_double, vp, C.c_size_t]
    lib.kwh_sim_autotune.argtypes = [vp, C.c_char_p, C.c_size_t]
    lib.kwh_sim_set_checkpointing.argtypes = [vp, C.c_int32]
    lib.kwh_sim_set_nonfinite_check.argtypes = [vp, C.c_int32]
    lib.kwh_sim_vjp.argtypes = [vp, vp, C.c_int32, vp, vp]
    lib.kwh_sim_destroy.argtypes = [vp]
    lib.kwh_sim_destroy.restype = None
    lib.kwh_time_reversal.argtypes = [vp, C.POINTER(SimDesc), C.POINTER(C.c_float), C.c_int32, C.c_double, vp]
    lib.kwh_batch_create.argtypes = [vp, C.POINTER(SimDesc), C.c_int32, C.POINTER(vp)]
    lib.kwh_batch_run.argtypes = [vp, C.c_int32]
    lib.kwh_batch_sim.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    lib.kwh_batch_read_output.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_size_t]
    lib.kwh_batch_destroy.argtypes = [vp]
    lib.kwh_batch_destroy.restype = None
    lib.kwh_h5_available.restype = C.c_int
    lib.kwh_h5_write_input.argtypes = [C.c_char_p, C.POINTER(SimDesc)]
    lib.kwh_h5_read_input.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.kwh_h5_input_desc.argtypes = [vp]
    lib.kwh_h5_input_desc.restype = C.POINTER(SimDesc)
    lib.kwh_h5_input_destroy.argtypes = [vp]
    lib.kwh_h5_input_destroy.restype = None
    lib.kwh_h5_write_output.argtypes = [C.c_char_p, vp, vp]
    lib.kwh_h5_read_dataset.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double), C.c_size_t]
    lib.kwh_h5_read_root_attribute.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
    if lib.kwh_abi_version() != ABI_VERSION:
        raise ImportError("libkwavehip ABI version mismatch")
    _lib = lib
    return lib


class KWaveHIPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class KWaveNonFiniteError(KWaveHIPError, FloatingPointError):
    """KWH_ENONFINITE: the optional blow-up check found a NaN or an infinity in the pressure."""


def check(code: int) -> None:
    """Map a kwh_status to a Python exception (the reference traps on invalid input)."""
    if code == OK:
        return
    msg = load().kwh_last_error().decode("utf-8", "replace")
    if code == EINVAL:
        raise ValueError(msg)
    if code == EUNSUPPORTED:
        raise NotImplementedError(msg)
    if code == ENOMEM:
        raise MemoryError(msg)
    if code == ENONFINITE:
        raise KWaveNonFiniteError(code, msg)
    raise KWaveHIPError(code, msg)
