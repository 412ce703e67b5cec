"""ctypes binding of ``libpypty_amd.so`` (the C ABI declared in include/pypty_amd.h).

The library is the only compute path of this package: if it is missing or
cannot be loaded every entry point raises ``RuntimeError``; there is no CPU
fallback.  Pointers handed to the library are ``torch.Tensor.data_ptr()`` of
device tensors; the stream is torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpypty_amd.so")

PTY_OK = 0
PTY_FFT_FWD, PTY_FFT_INV, PTY_FFT_FWD_SHIFT, PTY_FFT_INV_ISHIFT = 0, 1, 2, 3
PTY_LOSS_LSQ_SQRT, PTY_LOSS_ML, PTY_LOSS_LSQ, PTY_LOSS_LSQ_SQRT_2 = 0, 1, 2, 3
PTY_LOSS_LSQ_COMPRESSED, PTY_LOSS_GAUSS_COMPRESSED, PTY_LOSS_POISSON_COMPRESSED = 4, 5, 6
PTY_MEASURED_BY_POSITION, PTY_MEASURED_BY_SEQUENCE = 0, 1
PTY_LAYOUT_REFERENCE, PTY_LAYOUT_PLANAR = 0, 1
PTY_ERR_WORKSPACE = -5
PTY_MAX_ABERRATIONS = 64  # coefficients per aberration row
PTY_STATIC_WORKSPACE_BYTES = 64 << 20  # the workspace array of pty_loss_and_direction_static

EXPORTED = (
    "pty_status_string", "pty_version", "pty_supported_window", "pty_fft2_c64", "pty_bandlimit_c64",
    "pty_scatteradd_c64", "pty_workspace_bytes", "pty_loss_and_direction", "pty_tables_bytes",
    "pty_ms_forward", "pty_ms_adjoint", "pty_timing_enable", "pty_timing_read", "pty_timing_reset",
    "pty_fourier_clean_c64", "pty_plane_fft2_c64", "pty_multidot_c64", "pty_axpby_c64", "pty_axpby_f32",
    "pty_phase_grad_c64", "pty_charge_flip_c64", "pty_orthogonalize_modes_c64", "pty_phase_step_c64",
    "pty_l1_constraint_c64", "pty_plan_check", "pty_loss_and_direction_static", "pty_center_of_mass",
    "pty_gemm_bf16", "pty_cnn_im2col", "pty_cnn_col2im", "pty_cnn_bias_act", "pty_cnn_act_grad",
    "pty_cnn_bias_grad", "pty_cnn_to_bf16", "pty_cnn_mse", "pty_cnn_adam",
)


class PtyProblem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("n_slices", ctypes.c_int), ("n_obj_modes", ctypes.c_int),
        ("n_probe_modes", ctypes.c_int), ("obj_ny", ctypes.c_int), ("obj_nx", ctypes.c_int),
        ("n_positions", ctypes.c_int), ("compute_batch", ctypes.c_int), ("algorithm", ctypes.c_int),
        ("obj_layout", ctypes.c_int), ("step_obj", ctypes.c_int), ("step_probe", ctypes.c_int),
        ("step_pos", ctypes.c_int), ("shift_active", ctypes.c_int), ("loss_weight", ctypes.c_float),
        ("slice_distance", ctypes.c_float), ("wavelength", ctypes.c_float),
        ("pixel_size_x", ctypes.c_float), ("pixel_size_y", ctypes.c_float),
        ("tan_x", ctypes.c_float), ("tan_y", ctypes.c_float), ("damping_cutoff", ctypes.c_float),
        ("edge_multiplier", ctypes.c_float), ("n_masks", ctypes.c_int), ("raw_n", ctypes.c_int),
        ("data_shift_y", ctypes.c_int), ("data_shift_x", ctypes.c_int), ("data_bin", ctypes.c_int),
        ("data_pad", ctypes.c_int), ("data_multiplier", ctypes.c_float),
        ("checkpoint_slices", ctypes.c_int), ("lanes", ctypes.c_int),
        ("step_static_background", ctypes.c_int), ("step_beam_current", ctypes.c_int),
        ("n_aberrations", ctypes.c_int), ("n_markers", ctypes.c_int), ("step_aberrations", ctypes.c_int),
        ("n_distances", ctypes.c_int), ("tilt_mode", ctypes.c_int), ("n_tilts", ctypes.c_int),
        ("step_tilts", ctypes.c_int),
    ]


class PtyBuffers(ctypes.Structure):
    _fields_ = [
        ("obj", ctypes.c_void_p), ("probe", ctypes.c_void_p), ("positions", ctypes.c_void_p),
        ("corrections", ctypes.c_void_p), ("measured", ctypes.c_void_p), ("sequence", ctypes.c_void_p),
        ("n_seq", ctypes.c_int), ("loss_sse", ctypes.c_void_p), ("obj_grad", ctypes.c_void_p),
        ("probe_grad", ctypes.c_void_p), ("pos_grad", ctypes.c_void_p), ("sim_patterns", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t), ("masks", ctypes.c_void_p),
        ("measured_order", ctypes.c_int), ("accumulate", ctypes.c_int),
        ("static_background", ctypes.c_void_p), ("static_background_grad", ctypes.c_void_p),
        ("beam_current", ctypes.c_void_p), ("beam_current_grad", ctypes.c_void_p),
        ("aberrations", ctypes.c_void_p), ("aberration_marker", ctypes.c_void_p),
        ("aberrations_grad", ctypes.c_void_p), ("distances", ctypes.c_void_p), ("tilts", ctypes.c_void_p),
        ("tilts_grad", ctypes.c_void_p),
    ]


_lib = None
_load_error = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Load (once) and return the library, with argument types declared."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"pypty_amd: native library not built: {p} (run __graft_entry__.build())")
    try:
        lib = ctypes.CDLL(p)
    except OSError as exc:  # pragma: no cover - depends on the host
        _load_error = exc
        raise RuntimeError(f"pypty_amd: cannot load {p}: {exc}") from exc
    vp, i, i64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_size_t
    P = ctypes.POINTER
    lib.pty_status_string.restype = ctypes.c_char_p
    lib.pty_status_string.argtypes = [i]
    lib.pty_version.restype = i
    lib.pty_supported_window.restype = i
    lib.pty_supported_window.argtypes = [i]
    lib.pty_fft2_c64.restype = i
    lib.pty_fft2_c64.argtypes = [vp, i64, i, i, vp]
    lib.pty_bandlimit_c64.restype = i
    lib.pty_bandlimit_c64.argtypes = [vp, i64, i, vp, vp]
    lib.pty_scatteradd_c64.restype = i
    lib.pty_scatteradd_c64.argtypes = [vp, i, i, i, i, vp, i, i, vp, vp]
    lib.pty_workspace_bytes.restype = i
    lib.pty_workspace_bytes.argtypes = [P(PtyProblem), P(sz)]
    lib.pty_loss_and_direction.restype = i
    lib.pty_loss_and_direction.argtypes = [P(PtyProblem), P(PtyBuffers), vp]
    lib.pty_tables_bytes.restype = i
    lib.pty_tables_bytes.argtypes = [i, P(sz)]
    lib.pty_ms_forward.restype = i
    lib.pty_ms_forward.argtypes = [P(PtyProblem), i, vp, vp, vp, vp, vp, vp, vp]
    lib.pty_ms_adjoint.restype = i
    lib.pty_ms_adjoint.argtypes = [P(PtyProblem), i, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.pty_fourier_clean_c64.restype = i
    lib.pty_fourier_clean_c64.argtypes = [vp, i64, i, i, ctypes.c_double, ctypes.c_double, vp]
    lib.pty_plane_fft2_c64.restype = i
    lib.pty_plane_fft2_c64.argtypes = [vp, vp, i64, i, i, i, vp]
    lib.pty_multidot_c64.restype = i
    lib.pty_multidot_c64.argtypes = [P(vp), P(vp), i, i64, i, vp, vp, vp]
    lib.pty_axpby_c64.restype = i
    lib.pty_axpby_c64.argtypes = [vp, vp, i64, ctypes.c_float, ctypes.c_float, vp]
    lib.pty_axpby_f32.restype = i
    lib.pty_axpby_f32.argtypes = [vp, vp, i64, ctypes.c_float, ctypes.c_float, vp]
    lib.pty_phase_grad_c64.restype = i
    lib.pty_phase_grad_c64.argtypes = [vp, vp, vp, i64, vp]
    f = ctypes.c_float
    lib.pty_charge_flip_c64.restype = i
    lib.pty_charge_flip_c64.argtypes = [vp, i64, i, i, f, f, f, f, vp, vp]
    lib.pty_orthogonalize_modes_c64.restype = i
    lib.pty_orthogonalize_modes_c64.argtypes = [vp, i64, i, vp]
    lib.pty_phase_step_c64.restype = i
    lib.pty_phase_step_c64.argtypes = [vp, vp, f, vp, i64, vp]
    lib.pty_center_of_mass.restype = i
    lib.pty_center_of_mass.argtypes = [vp, i64, i, i, i, vp, vp, vp]
    lib.pty_l1_constraint_c64.restype = i
    lib.pty_l1_constraint_c64.argtypes = [vp, i64, ctypes.c_float, ctypes.c_float, vp, vp, vp, vp]
    lib.pty_loss_and_direction_static.restype = i
    lib.pty_loss_and_direction_static.argtypes = [P(PtyProblem), P(PtyBuffers), vp]
    lib.pty_plan_check.restype = i
    lib.pty_plan_check.argtypes = [P(PtyProblem), i, i, i, i, i, i, ctypes.c_char_p, sz]
    lib.pty_timing_enable.restype = None
    lib.pty_timing_enable.argtypes = [i]
    lib.pty_timing_read.restype = i
    lib.pty_timing_read.argtypes = [P(ctypes.c_char_p), P(ctypes.c_double), P(ctypes.c_int64), i]
    lib.pty_timing_reset.restype = None
    lib.pty_timing_reset.argtypes = []
    f = ctypes.c_float
    for name, args in (
            ("pty_gemm_bf16", [i, i, i, i, i, vp, i, vp, i, vp, i, vp, sz, vp]),
            ("pty_cnn_im2col", [vp, i, i, i, i, i, i, i, vp, vp]),
            ("pty_cnn_col2im", [vp, i, i, i, i, i, i, i, vp, vp]),
            ("pty_cnn_bias_act", [vp, i64, i, vp, i, vp, vp, vp]),
            ("pty_cnn_act_grad", [vp, vp, i64, i, vp, vp, vp]),
            ("pty_cnn_bias_grad", [vp, i64, i, vp, vp, sz, vp]),
            ("pty_cnn_to_bf16", [vp, i64, vp, vp]),
            ("pty_cnn_mse", [vp, vp, i64, vp, vp, vp, sz, vp]),
            ("pty_cnn_adam", [vp, vp, vp, vp, vp, i64, f, f, f, f, i, vp])):
        fn = getattr(lib, name)
        fn.restype = i
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    """Raise RuntimeError on a nonzero library status."""
    if rc != PTY_OK:
        msg = load().pty_status_string(rc).decode()
        raise RuntimeError(f"pypty_amd: {what} failed with status {rc} ({msg})")


def stream_ptr(device=None) -> int:
    import torch

    return torch.cuda.current_stream(device).cuda_stream


def require_device_tensor(t, name: str):
    import torch

    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"pypty_amd: {name} must be a torch tensor on a HIP device")
    if not t.is_contiguous():
        raise ValueError(f"pypty_amd: {name} must be contiguous")
    return t


def timing_enable(on: bool = True) -> None:
    load().pty_timing_enable(1 if on else 0)


def timing_reset() -> None:
    load().pty_timing_reset()


def timing_read() -> dict:
    """Per-kernel totals {name: (total_ms, launches)} since the last reset."""
    lib = load()
    cap = 32
    names = (ctypes.c_char_p * cap)()
    ms = (ctypes.c_double * cap)()
    cnt = (ctypes.c_int64 * cap)()
    n = lib.pty_timing_read(names, ms, cnt, cap)
    return {names[i].decode(): (ms[i], int(cnt[i])) for i in range(min(n, cap))}


def plan_check(problem: "PtyProblem", n_seq: int, lanes: int, calls: int = 1, accumulate: bool = False,
               verbose: bool = False, omit: int = 0):
    """Host-side check of the launch plan of pty_loss_and_direction (no GPU needed):
    returns (violations, report text).  ``omit`` plants faults in the recorded plan only."""
    buf = ctypes.create_string_buffer(1 << 22)
    n = load().pty_plan_check(ctypes.byref(problem), int(n_seq), int(lanes), int(calls), int(bool(accumulate)),
                              int(bool(verbose)), int(omit), buf, len(buf))
    if n < 0:
        check(n, "plan_check")
    return n, buf.value.decode()
