"""ctypes binding of the C ABI in `include/typecv_ba.h` (`libtypecv_ba.so`).

Loading fails loudly when the library is missing: the engine has no CPU fallback.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libtypecv_ba.so"

FP64 = 0
MIXED = 1
FLAG_TIMING = 1
FLAG_IMPLICIT = 2
FLAG_SCHUR_CACHE = 4
FLAG_DIRECT = 8

STATUS = {
    0: "ok",
    1: "argument",
    2: "index",
    3: "device",
    4: "memory",
    5: "no-device",
    6: "solve",
}

# The symbols include/typecv_ba.h declares, in declaration order.
EXPORTS = (
    "tcv_ba_output_size",
    "tcv_ba_scene_size",
    "tcv_ba_index_size",
    "tcv_ba_status_string",
    "tcv_ba_last_error",
    "tcv_ba_create",
    "tcv_ba_destroy",
    "tcv_ba_device_bytes",
    "tcv_ba_release",
    "tcv_ba_set_working_bytes",
    "tcv_ba_fail_reads",
    "tcv_ba_solve",
    "tcv_ba_solve_device",
    "tcv_ba_solve_batch",
    "tcv_ba_schur_probe",
    "tcv_ba_jacobian",
    "tcv_ba_polish",
    "tcv_ba_backward",
    "tcv_obs_evaluate",
    "tcv_obs_evaluate_noalloc",
)


class Control(ctypes.Structure):
    """`tcv_ba_control`."""

    _fields_ = [
        ("cameras", ctypes.c_uint32),
        ("points", ctypes.c_uint32),
        ("directional_points", ctypes.c_uint32),
        ("observations", ctypes.c_uint32),
        ("anchor", ctypes.c_uint32),
        ("scale_view", ctypes.c_uint32),
        ("normalize_scale", ctypes.c_uint32),
        ("shared_focal", ctypes.c_uint32),
        ("iterations", ctypes.c_uint32),
        ("max_cg", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("huber", ctypes.c_double),
        ("relative_tolerance", ctypes.c_double),
        ("stall_tolerance", ctypes.c_double),
        ("refinement_tolerance", ctypes.c_double),
    ]


class Report(ctypes.Structure):
    """`tcv_ba_report`."""

    _fields_ = [
        ("solve_ms", ctypes.c_double),
        ("setup_ms", ctypes.c_double),
        ("schur_ms", ctypes.c_double),
        ("schur_point_ms", ctypes.c_double),
        ("schur_camera_ms", ctypes.c_double),
        ("refinement_steps", ctypes.c_double),
        ("lm_iterations", ctypes.c_uint64),
        ("cg_iterations", ctypes.c_uint64),
        ("matvecs", ctypes.c_uint64),
        ("matvec_launches", ctypes.c_uint64),
        ("flag_reads", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("point_tiles", ctypes.c_uint64),
        ("camera_tasks", ctypes.c_uint64),
        ("direct_solves", ctypes.c_uint64),
        ("timed_matvecs", ctypes.c_uint64),
        ("direct_corrections", ctypes.c_uint64),
        ("camera_sweep_cap", ctypes.c_uint64),
        ("fused_product", ctypes.c_uint64),
        ("streamed", ctypes.c_uint64),
        ("held_points", ctypes.c_uint64),
        ("onepass_rows", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ObsControl(ctypes.Structure):
    """`tcv_obs_control` (include/typecv_ba.h)."""
    _fields_ = [
        ("views", ctypes.c_uint32),
        ("points", ctypes.c_uint32),
        ("observations", ctypes.c_uint32),
        ("tracks", ctypes.c_uint32),
        ("max_error_px", ctypes.c_double),
        ("max_relative_depth_std_dev", ctypes.c_double),
        ("pixel_noise_std_dev_px", ctypes.c_double),
    ]


_lib = None


class EngineMissing(RuntimeError):
    pass


def library() -> ctypes.CDLL:
    """The loaded engine library; raises EngineMissing when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBRARY.exists():
        raise EngineMissing(
            f"{LIBRARY} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(str(LIBRARY))
    P = ctypes.POINTER
    c_control = P(Control)
    lib.tcv_ba_output_size.argtypes = [c_control]
    lib.tcv_ba_output_size.restype = ctypes.c_size_t
    lib.tcv_ba_scene_size.argtypes = [c_control]
    lib.tcv_ba_scene_size.restype = ctypes.c_size_t
    lib.tcv_ba_index_size.argtypes = [c_control]
    lib.tcv_ba_index_size.restype = ctypes.c_size_t
    lib.tcv_ba_status_string.argtypes = [ctypes.c_int]
    lib.tcv_ba_status_string.restype = ctypes.c_char_p
    lib.tcv_ba_last_error.argtypes = []
    lib.tcv_ba_last_error.restype = ctypes.c_char_p
    lib.tcv_ba_create.argtypes = [ctypes.c_int, P(ctypes.c_void_p)]
    lib.tcv_ba_create.restype = ctypes.c_int
    lib.tcv_ba_destroy.argtypes = [ctypes.c_void_p]
    lib.tcv_ba_destroy.restype = None
    lib.tcv_ba_device_bytes.argtypes = [ctypes.c_void_p]
    lib.tcv_ba_device_bytes.restype = ctypes.c_uint64
    lib.tcv_ba_release.argtypes = [ctypes.c_void_p]
    lib.tcv_ba_release.restype = ctypes.c_int
    lib.tcv_ba_set_working_bytes.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    lib.tcv_ba_set_working_bytes.restype = ctypes.c_int
    lib.tcv_ba_fail_reads.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    lib.tcv_ba_fail_reads.restype = ctypes.c_int
    lib.tcv_ba_solve.argtypes = [ctypes.c_void_p, c_control, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, P(Report)]
    lib.tcv_ba_solve.restype = ctypes.c_int
    lib.tcv_ba_solve_device.argtypes = [ctypes.c_void_p, c_control, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, P(Report),
                                        ctypes.c_void_p]
    lib.tcv_ba_solve_device.restype = ctypes.c_int
    lib.tcv_ba_solve_batch.argtypes = [ctypes.c_void_p, ctypes.c_uint32, c_control,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       P(Report), ctypes.c_void_p]
    lib.tcv_ba_solve_batch.restype = ctypes.c_int
    lib.tcv_ba_schur_probe.argtypes = [ctypes.c_void_p, c_control, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                       P(Report)]
    lib.tcv_ba_schur_probe.restype = ctypes.c_int
    vp = ctypes.c_void_p
    lib.tcv_ba_jacobian.argtypes = [vp, c_control, vp, vp, ctypes.c_double, vp, P(Report)]
    lib.tcv_ba_jacobian.restype = ctypes.c_int
    lib.tcv_ba_polish.argtypes = [vp, c_control, vp, vp, ctypes.c_double, ctypes.c_uint32,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, vp, P(Report)]
    lib.tcv_ba_polish.restype = ctypes.c_int
    lib.tcv_ba_backward.argtypes = [vp, c_control, vp, vp, ctypes.c_double, vp, ctypes.c_uint32,
                                    ctypes.c_double, ctypes.c_uint32, vp, vp, vp, P(Report)]
    lib.tcv_ba_backward.restype = ctypes.c_int
    lib.tcv_obs_evaluate.argtypes = [vp, P(ObsControl), vp, vp, vp, vp, vp, vp, vp]
    lib.tcv_obs_evaluate.restype = ctypes.c_int
    lib.tcv_obs_evaluate_noalloc.argtypes = [P(ObsControl), vp, vp, vp, vp, vp, vp, vp, vp]
    lib.tcv_obs_evaluate_noalloc.restype = ctypes.c_int
    lib.TCV_OBS_NOALLOC_MAX_OBSERVATIONS = 2097152
    _lib = lib
    return lib


class EngineError(RuntimeError):
    def __init__(self, status: int, where: str):
        lib = library()
        text = lib.tcv_ba_status_string(status).decode()
        detail = lib.tcv_ba_last_error().decode()
        super().__init__(f"{where}: {text}" + (f" ({detail})" if detail else ""))
        self.status = status


def check(status: int, where: str) -> None:
    if status != 0:
        raise EngineError(status, where)
