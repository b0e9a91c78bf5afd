"""ctypes binding of libpgpe_hip.so (declared in include/pgpe.h).

The library is the only compute path of this package: there is no CPU
fallback.  ``lib()`` raises ``RuntimeError`` when the shared library is
missing, and every wrapper raises ``RuntimeError`` with the library's own
message when a call fails.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpgpe_hip.so"

PGPE_RES_NONE, PGPE_RES_SINGLE, PGPE_RES_DOUBLE, PGPE_RES_QDOUBLE = 0, 1, 2, 3
PGPE_LAP_FIVE, PGPE_LAP_NINE = 0, 1
LAPLACIAN_KIND = {"five-point": PGPE_LAP_FIVE, "isotropic-9pt": PGPE_LAP_NINE}
PGPE_BC_PERIODIC, PGPE_BC_CLOSED = 0, 1
BOUNDARY_KIND = {"periodic": PGPE_BC_PERIODIC, "closed-interval": PGPE_BC_CLOSED}
RESERVOIR_KIND = {"none": PGPE_RES_NONE, "single": PGPE_RES_SINGLE, "double": PGPE_RES_DOUBLE,
                  "quadratic-double": PGPE_RES_QDOUBLE}
RESERVOIR_FIELDS = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}

vp = C.c_void_p


class Member(C.Structure):
    """Mirror of ``pgpe_member``."""
    _fields_ = [
        ("off", C.c_int64), ("pump_off", C.c_int64),
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("tiles_x", C.c_int32), ("tiles_y", C.c_int32),
        ("tile_begin", C.c_int32), ("components", C.c_int32),
        ("tile_h", C.c_int32), ("res_components", C.c_int32),
        ("inv_dx2", C.c_double), ("inv_dy2", C.c_double),
        ("hbar_over_2m", C.c_double), ("inv_hbar", C.c_double), ("kin", C.c_double),
        ("hbar", C.c_double), ("m_eff", C.c_double),
        ("g_C", C.c_double), ("g_R", C.c_double), ("g_I", C.c_double), ("R", C.c_double),
        ("gamma_C", C.c_double), ("eta", C.c_double),
        ("res_a", C.c_double), ("res_b", C.c_double), ("res_c", C.c_double), ("res_d", C.c_double),
        ("alpha_2", C.c_double), ("zeeman", C.c_double), ("te_tm", C.c_double), ("inv_4dxdy", C.c_double),
        ("noise_gamma", C.c_double), ("noise_R", C.c_double), ("noise_strength", C.c_double),
        ("noise_dt_over_4v", C.c_double),
        ("seed", C.c_uint64),
        ("pump_w_plus", C.c_double), ("pump_w_minus", C.c_double),
        ("pot_off", C.c_int64), ("laplacian", C.c_int32), ("boundary", C.c_int32),
    ]


class Rk4Args(C.Structure):
    """Mirror of ``pgpe_rk4_args``."""
    _fields_ = [
        ("precision", C.c_int32), ("reservoir", C.c_int32), ("stage", C.c_int32),
        ("num_tiles", C.c_int32), ("tile_y", C.c_int32), ("noise", C.c_int32),
        ("members", C.c_int32), ("components", C.c_int32), ("res_components", C.c_int32),
        ("five_point", C.c_int32),
        ("dt", C.c_double),
        ("table", vp), ("tile_map", vp),
        ("psi0", vp), ("kprev", vp), ("k1", vp), ("k2", vp),
        ("kout", vp), ("psi_out", vp),
        ("res0", vp * 2), ("kres_prev", vp * 2), ("kres1", vp * 2), ("kres2", vp * 2),
        ("kres_out", vp * 2), ("res_out", vp * 2),
        ("pump", vp), ("coef", vp),
        ("step", C.c_uint64), ("step_base", vp),
        ("potential", vp), ("active_tiles", vp), ("strips", C.c_int32), ("no_eta", C.c_int32),
    ]


class SplitArgs(C.Structure):
    """Mirror of ``pgpe_split_args``."""
    _fields_ = [
        ("precision", C.c_int32), ("reservoir", C.c_int32), ("ordering", C.c_int32), ("noise", C.c_int32),
        ("members", C.c_int32), ("components", C.c_int32),
        ("dt", C.c_double),
        ("table", vp), ("plan", vp), ("psi", vp), ("tmp", vp), ("res", vp * 2),
        ("pump", vp), ("coef", vp), ("step", C.c_uint64), ("step_base", vp),
        ("res_components", C.c_int32), ("chain", C.c_int32),
    ]


CHAIN_IN, CHAIN_OUT = 1, 2   # pgpe_split_chain


MAX_STAGES = 12


class LawsonArgs(C.Structure):
    """Mirror of ``pgpe_lawson_args``."""
    _fields_ = [
        ("precision", C.c_int32), ("reservoir", C.c_int32), ("op", C.c_int32), ("components", C.c_int32),
        ("res_components", C.c_int32), ("members", C.c_int32),
        ("nterms", C.c_int32), ("lap", C.c_int32), ("nerr", C.c_int32), ("stage_real", C.c_int32),
        ("noise", C.c_int32), ("pad0", C.c_int32),
        ("h", C.c_double), ("frac0", C.c_double), ("atol", C.c_double), ("rtol", C.c_double),
        ("hc", C.c_double * MAX_STAGES), ("fr", C.c_double * MAX_STAGES), ("rc", C.c_double * MAX_STAGES),
        ("he5", C.c_double * MAX_STAGES), ("he3", C.c_double * MAX_STAGES),
        ("table", vp), ("plan", vp), ("psi0k", vp), ("f", vp * MAX_STAGES), ("dres", (vp * 2) * MAX_STAGES),
        ("work", vp), ("aux", vp), ("f_out", vp), ("psi_real", vp), ("res0", vp * 2), ("dres_out", vp * 2),
        ("pump", vp), ("coef", vp), ("potential", vp), ("psi_out", vp), ("res_out", vp * 2), ("psik_out", vp),
        ("psi_start", vp), ("partials", vp), ("error_sums", vp), ("step", C.c_uint64), ("step_base", vp),
        ("h_dev", vp), ("skip_dev", vp),
    ]


COMMIT_MAX = 8
ADAPTIVE_UNDERFLOW = 1


class AdaptiveState(C.Structure):
    """Mirror of ``pgpe_adaptive_state`` (the device-resident trial controller)."""
    _fields_ = [
        ("t", C.c_double), ("target", C.c_double), ("h_next", C.c_double), ("err_prev", C.c_double),
        ("dt_max", C.c_double), ("k", C.c_double), ("h", C.c_double), ("prop", C.c_double),
        ("error", C.c_double), ("dt_min", C.c_double), ("dt_max_used", C.c_double), ("dt_last", C.c_double),
        ("count", C.c_double), ("members", C.c_int32), ("blend", C.c_int32),
        ("clamped", C.c_int32), ("no_growth", C.c_int32), ("accepted", C.c_int32), ("rejected", C.c_int32),
        ("done", C.c_int32), ("commit", C.c_int32), ("status", C.c_int32), ("pad0", C.c_int32),
    ]


class CommitList(C.Structure):
    """Mirror of ``pgpe_commit_list``."""
    _fields_ = [("dst", vp * COMMIT_MAX), ("src", vp * COMMIT_MAX), ("bytes", C.c_uint64 * COMMIT_MAX),
                ("count", C.c_int32), ("pad0", C.c_int32)]


class GemmArgs(C.Structure):
    """Mirror of ``pgpe_gemm_args``."""
    _fields_ = [
        ("A", vp), ("B", vp), ("C", vp), ("bias", vp), ("residual", vp),
        ("M", C.c_int32), ("N", C.c_int32), ("K", C.c_int32), ("batch", C.c_int32),
        ("sam", C.c_int64), ("sak", C.c_int64), ("sab", C.c_int64),
        ("sbk", C.c_int64), ("sbn", C.c_int64), ("sbb", C.c_int64),
        ("scm", C.c_int64), ("scn", C.c_int64), ("scb", C.c_int64), ("srb", C.c_int64),
        ("alpha", C.c_float), ("beta", C.c_float),
        ("bias_mode", C.c_int32), ("gelu", C.c_int32), ("out_bf16", C.c_int32), ("pad0", C.c_int32),
    ]


# name -> (restype, argtypes); every function returns an int status except pgpe_last_error.
_SIGNATURES = {
    "pgpe_last_error": (C.c_char_p, []),
    "pgpe_version": (C.c_int, []),
    "pgpe_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pgpe_device_init": (C.c_int, [C.c_int]),
    "pgpe_device_name": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t]),
    "pgpe_device_cu_count": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "pgpe_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "pgpe_free": (C.c_int, [vp]),
    "pgpe_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "pgpe_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "pgpe_memcpy_d2h_async": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "pgpe_host_alloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "pgpe_host_free": (C.c_int, [vp]),
    "pgpe_memcpy_d2d": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "pgpe_fill_zero": (C.c_int, [vp, C.c_size_t, vp]),
    "pgpe_stream_create": (C.c_int, [C.POINTER(vp)]),
    "pgpe_stream_destroy": (C.c_int, [vp]),
    "pgpe_stream_sync": (C.c_int, [vp]),
    "pgpe_device_sync": (C.c_int, []),
    "pgpe_event_create": (C.c_int, [C.POINTER(vp)]),
    "pgpe_event_destroy": (C.c_int, [vp]),
    "pgpe_event_record": (C.c_int, [vp, vp]),
    "pgpe_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
    "pgpe_event_sync": (C.c_int, [vp]),
    "pgpe_stream_wait_event": (C.c_int, [vp, vp]),
    "pgpe_graph_begin": (C.c_int, [vp]),
    "pgpe_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
    "pgpe_graph_launch": (C.c_int, [vp, vp]),
    "pgpe_graph_destroy": (C.c_int, [vp]),
    "pgpe_philox_words": (C.c_int, [vp, vp, C.c_uint64, C.c_int64, vp]),
    "pgpe_normal_pairs": (C.c_int, [vp, C.c_int, C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64, vp]),
    "pgpe_wigner_add": (C.c_int, [vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_uint64, vp]),
    "pgpe_rk4_stage": (C.c_int, [C.POINTER(Rk4Args), vp]),
    "pgpe_rk4_step": (C.c_int, [C.POINTER(Rk4Args), vp]),
    "pgpe_fft_plan_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(vp)]),
    "pgpe_fft_plan_destroy": (C.c_int, [vp]),
    "pgpe_fft2": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp]),
    "pgpe_fft_columns": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, vp]),
    "pgpe_split_step": (C.c_int, [C.POINTER(SplitArgs), vp]),
    "pgpe_lawson": (C.c_int, [C.POINTER(LawsonArgs), vp]),
    "pgpe_adaptive_propose": (C.c_int, [vp, vp]),
    "pgpe_adaptive_decide": (C.c_int, [vp, vp, vp]),
    "pgpe_adaptive_commit": (C.c_int, [vp, C.POINTER(CommitList), vp]),
    "pgpe_gemm_bf16": (C.c_int, [C.POINTER(GemmArgs), vp]),
    "pgpe_f32_to_bf16": (C.c_int, [vp, vp, C.c_int64, vp]),
    "pgpe_gelu_fwd": (C.c_int, [vp, vp, vp, C.c_int64, vp]),
    "pgpe_gelu_bwd": (C.c_int, [vp, vp, vp, vp, C.c_int64, vp]),
    "pgpe_mode_scale": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "pgpe_mode_scale_grad": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    "pgpe_colsum": (C.c_int, [vp, vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, vp]),
    "pgpe_density_delta": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, vp]),
    "pgpe_density_delta_noalloc": (C.c_int, [vp, vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp,
                                             C.c_int, vp]),
    "pgpe_accumulate_power": (C.c_int, [vp, vp, C.c_int64, C.c_int, vp]),
    "pgpe_accumulate_moments": (C.c_int, [vp, vp, vp, C.c_int64, C.c_int, vp]),
    "pgpe_spectrum_line": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, vp]),
    "pgpe_spectrum_block": (C.c_int, [vp, vp, vp, C.c_int, C.c_int64, C.c_int, C.c_int, vp]),
    "pgpe_spectrum_power": (C.c_int, [vp, vp, C.c_int, C.c_int64, C.c_int, vp]),
    "pgpe_frame_fields": (C.c_int, [vp, C.c_int, C.c_int, C.c_int64, C.c_int64, vp, vp, vp, vp]),
}

_lib = None


def declared_symbols() -> list[str]:
    """Every entry point this binding declares (the same set as include/pgpe.h)."""
    return sorted(_SIGNATURES)


def lib():
    """Load libpgpe_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("PGPE_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m polarism_amd.build_native` "
            "(there is no CPU fallback for the HIP solvers)")
    handle = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def last_error() -> str:
    msg = lib().pgpe_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Raise RuntimeError for a nonzero status."""
    if rc != 0:
        raise RuntimeError(f"{what or 'pgpe call'} failed (status {rc}): {last_error()}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
