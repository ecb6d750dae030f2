"""Loads libopenpstd_amd.so (built in-tree by openpstd_amd.build).

There is no fallback: a missing library is an error.
"""
import ctypes as C
import os

from . import abi_types as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libopenpstd_amd.so")
_lib = None

EXPORTED = [
    "pstd_ctx_create", "pstd_ctx_upload_plan", "pstd_ctx_set_field", "pstd_ctx_get_field", "pstd_ctx_advance",
    "pstd_ctx_advance_async", "pstd_ctx_advance_static", "pstd_ctx_synchronize", "pstd_ctx_read_frame", "pstd_ctx_steps_done",
    "pstd_ctx_profile_stages", "pstd_ctx_profile_kernels", "pstd_ctx_cells_per_stage", "pstd_ctx_upload_plans", "pstd_ctx_set_graph_steps", "pstd_ctx_capture_graph",
    "pstd_ctx_set_stage_overlap",
    "pstd_ctx_set_tile_order", "pstd_ctx_set_kernel_flavour", "pstd_ctx_set_tile_threads",
    "pstd_ctx_get_tile_threads", "pstd_ctx_get_tuning", "pstd_ctx_get_autotune", "pstd_batch_prepare",
    "pstd_derivative_matrix", "pstd_shift_operator_row", "pstd_ctx_set_form", "pstd_ctx_get_form", "pstd_kernel_set_form", "pstd_ctx_gemm_work",
    "pstd_ctx_last_error", "pstd_ctx_destroy", "pstd_kernel_create", "pstd_kernel_initialize", "pstd_kernel_run",
    "pstd_kernel_get_metadata", "pstd_kernel_domain_info", "pstd_kernel_read_field", "pstd_kernel_read_pml",
    "pstd_kernel_prepare", "pstd_kernel_segment_count", "pstd_kernel_get_segment", "pstd_kernel_last_error", "pstd_kernel_destroy",
    "pstd_file_open", "pstd_file_new", "pstd_file_get_scene", "pstd_file_set_scene", "pstd_file_frame_count",
    "pstd_file_get_frame", "pstd_file_get_receiver", "pstd_file_commit", "pstd_file_run", "pstd_file_last_error",
    "pstd_file_close", "pstd_file_get_raw", "pstd_file_set_raw", "pstd_file_delete_raw", "pstd_file_record_count",
    "pstd_file_record_key",
]


class PSTDError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"openpstd_amd native library not built: {LIB_PATH} "
                           "(run `python -m openpstd_amd.build`)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    L.pstd_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pstd_ctx_upload_plan.argtypes = [vp, vp]
    L.pstd_ctx_set_field.argtypes = [vp, C.c_int, C.c_int, fp, C.c_size_t]
    L.pstd_ctx_get_field.argtypes = [vp, C.c_int, C.c_int, fp, C.c_size_t]
    L.pstd_ctx_advance.argtypes = [vp, C.c_int, fp]
    L.pstd_ctx_advance_async.argtypes = [vp, C.c_int]
    L.pstd_ctx_advance_static.argtypes = [vp, C.c_int]
    L.pstd_ctx_synchronize.argtypes = [vp]
    L.pstd_ctx_read_frame.argtypes = [vp, C.c_int, fp, C.c_size_t]
    L.pstd_ctx_steps_done.argtypes = [vp]
    L.pstd_ctx_profile_stages.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.pstd_ctx_profile_kernels.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.pstd_ctx_upload_plans.argtypes = [vp, vp, C.c_int]
    L.pstd_ctx_set_graph_steps.argtypes = [vp, C.c_int]
    L.pstd_ctx_capture_graph.argtypes = [vp]
    L.pstd_ctx_set_stage_overlap.argtypes = [vp, C.c_int]
    L.pstd_ctx_set_tile_order.argtypes = [vp, C.c_int]
    L.pstd_ctx_set_kernel_flavour.argtypes = [vp, C.c_int]
    L.pstd_ctx_set_tile_threads.argtypes = [vp, C.c_int]
    L.pstd_ctx_get_tile_threads.argtypes = [vp]
    L.pstd_ctx_get_tuning.argtypes = [vp, ip, ip, C.POINTER(C.c_double)]
    L.pstd_ctx_get_autotune.argtypes = [vp, ip, ip]
    L.pstd_batch_prepare.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp)]
    L.pstd_derivative_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_size_t]
    L.pstd_derivative_matrix.restype = C.c_longlong
    L.pstd_shift_operator_row.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                          C.POINTER(C.c_double), C.c_double, C.c_int, C.POINTER(C.c_double),
                                          C.c_size_t]
    L.pstd_shift_operator_row.restype = C.c_longlong
    fp = C.POINTER(C.c_float)
    L.pstd_file_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.pstd_file_new.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.pstd_file_get_scene.argtypes = [vp, C.c_int, C.POINTER(T.pstd_config)]
    L.pstd_file_set_scene.argtypes = [vp, C.POINTER(T.pstd_config)]
    L.pstd_file_frame_count.argtypes = [vp, C.c_int]
    L.pstd_file_get_frame.argtypes = [vp, C.c_int, C.c_int, fp, C.c_size_t]
    L.pstd_file_get_frame.restype = C.c_longlong
    L.pstd_file_get_receiver.argtypes = [vp, C.c_int, fp, C.c_size_t]
    L.pstd_file_get_receiver.restype = C.c_longlong
    L.pstd_file_commit.argtypes = [vp]
    L.pstd_file_run.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(T.pstd_callbacks)]
    L.pstd_file_last_error.argtypes = [vp]
    L.pstd_file_last_error.restype = C.c_char_p
    L.pstd_file_close.argtypes = [vp]
    L.pstd_file_get_raw.argtypes = [vp, C.c_char_p, C.c_int, C.c_void_p, C.c_size_t]
    L.pstd_file_get_raw.restype = C.c_longlong
    L.pstd_file_set_raw.argtypes = [vp, C.c_char_p, C.c_int, C.c_char_p, C.c_longlong, C.c_int]
    L.pstd_file_delete_raw.argtypes = [vp, C.c_char_p, C.c_int]
    L.pstd_file_record_count.argtypes = [vp]
    L.pstd_file_record_key.argtypes = [vp, C.c_int, C.c_void_p, C.c_int]
    L.pstd_ctx_set_form.argtypes = [vp, C.c_int]
    L.pstd_ctx_get_form.argtypes = [vp]
    L.pstd_ctx_gemm_work.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    L.pstd_kernel_set_form.argtypes = [vp, C.c_int]
    L.pstd_ctx_cells_per_stage.argtypes = [vp]
    L.pstd_ctx_cells_per_stage.restype = C.c_longlong
    L.pstd_ctx_last_error.argtypes = [vp]
    L.pstd_ctx_last_error.restype = C.c_char_p
    L.pstd_ctx_destroy.argtypes = [vp]
    L.pstd_ctx_destroy.restype = None
    L.pstd_kernel_create.argtypes = [C.c_int, C.c_int, C.POINTER(vp)]
    L.pstd_kernel_initialize.argtypes = [vp, C.POINTER(T.pstd_config), C.POINTER(T.pstd_callbacks)]
    L.pstd_kernel_run.argtypes = [vp, C.POINTER(T.pstd_callbacks)]
    L.pstd_kernel_get_metadata.argtypes = [vp, ip, ip, C.c_int32, ip]
    L.pstd_kernel_domain_info.argtypes = [vp, C.c_int, ip, fp]
    L.pstd_kernel_read_field.argtypes = [vp, C.c_int, C.c_int, fp, C.c_size_t]
    L.pstd_kernel_read_pml.argtypes = [vp, C.c_int, C.c_int, fp, C.c_size_t]
    L.pstd_kernel_prepare.argtypes = [vp, C.POINTER(vp)]
    L.pstd_kernel_segment_count.argtypes = [vp]
    L.pstd_kernel_get_segment.argtypes = [vp, C.c_int, C.POINTER(T.pstd_plan_segment)]
    L.pstd_kernel_last_error.argtypes = [vp]
    L.pstd_kernel_last_error.restype = C.c_char_p
    L.pstd_kernel_destroy.argtypes = [vp]
    L.pstd_kernel_destroy.restype = None
    _lib = L
    return L
