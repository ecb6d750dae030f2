"""ctypes mirrors of the C structures declared in include/openpstd_amd.h.

Types only: importing this module loads no library.
"""
import ctypes as C

PSTD_OK = 0
PSTD_ERR_INVALID = -1
PSTD_ERR_NOT_CONFIGURED = -2
PSTD_ERR_DEVICE = -3
PSTD_ERR_UNSUPPORTED = -4
PSTD_ERR_ALLOC = -5

PSTD_FIELD_P0, PSTD_FIELD_PX, PSTD_FIELD_PY, PSTD_FIELD_VX, PSTD_FIELD_VY = range(5)
FIELD_IDS = {"p0": 0, "px0": 1, "py0": 2, "vx0": 3, "vy0": 4}

PSTD_RECEIVER_PHYSICAL = 0
PSTD_RECEIVER_REFERENCE = 1
PSTD_RECEIVER_SPECTRAL = 2

PSTD_FORM_FFT = 0
PSTD_FORM_GEMM_F16 = 1
PSTD_FORM_GEMM_BF16 = 2
PSTD_FORM_AUTO = 3
FORMS = {"fft": PSTD_FORM_FFT, "gemm": PSTD_FORM_GEMM_F16, "gemm_f16": PSTD_FORM_GEMM_F16,
         "gemm_bf16": PSTD_FORM_GEMM_BF16, "auto": PSTD_FORM_AUTO}


class pstd_settings(C.Structure):
    _fields_ = [
        ("calctime", C.c_float),
        ("c1", C.c_float),
        ("ampmax", C.c_float),
        ("freqmax", C.c_float),
        ("rho", C.c_float),
        ("patcherror", C.c_float),
        ("tfactRK", C.c_float),
        ("grid_spacing", C.c_float),
        ("band_width", C.c_float),
        ("wave_length", C.c_float),
        ("rk_coefficients", C.c_float * 6),
        ("spectral_interpolation", C.c_int32),
        ("pml_cells", C.c_int32),
        ("save_nth", C.c_int32),
    ]


class pstd_edge(C.Structure):
    _fields_ = [("absorption", C.c_float), ("lr", C.c_int32)]


class pstd_domain_conf(C.Structure):
    _fields_ = [
        ("top_left", C.c_float * 2),
        ("size", C.c_float * 2),
        ("T", pstd_edge),
        ("L", pstd_edge),
        ("B", pstd_edge),
        ("R", pstd_edge),
    ]


class pstd_config(C.Structure):
    _fields_ = [
        ("settings", pstd_settings),
        ("n_speakers", C.c_int32),
        ("speakers", C.POINTER(C.c_float)),
        ("n_receivers", C.c_int32),
        ("receivers", C.POINTER(C.c_float)),
        ("n_domains", C.c_int32),
        ("domains", C.POINTER(pstd_domain_conf)),
        ("receiver_mode", C.c_int32),
    ]


class pstd_plan_segment(C.Structure):
    _fields_ = [
        ("domain", C.c_int32),
        ("direction", C.c_int32),
        ("field", C.c_int32),
        ("line_begin", C.c_int32),
        ("line_end", C.c_int32),
        ("side1", C.c_int32),
        ("side2", C.c_int32),
        ("side1_offset", C.c_int32),
        ("side2_offset", C.c_int32),
        ("wlen", C.c_int32),
        ("table_len", C.c_int32),
        ("c_side1", C.c_float),
        ("c_main1", C.c_float),
        ("c_side2", C.c_float),
        ("c_main2", C.c_float),
        ("interp_only", C.c_int32),
    ]


FRAME_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_size_t)
SAMPLE_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_size_t)
LOG_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_char_p)


class pstd_callbacks(C.Structure):
    _fields_ = [
        ("user", C.c_void_p),
        ("write_frame", FRAME_CB),
        ("write_sample", SAMPLE_CB),
        ("log", LOG_CB),
    ]
