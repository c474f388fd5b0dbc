"""ctypes binding of libfftwpp_amd.so (the C ABI of include/fftwpp_amd.h and
include/cfftw++.h).  Loading fails loudly when the library has not been built:
there is no CPU fallback."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfftwpp_amd.so")

P = ctypes.c_void_p
Z = ctypes.c_size_t
I = ctypes.c_int
D = ctypes.c_double
ZP = ctypes.POINTER(ctypes.c_size_t)

# name -> (restype, argtypes)
SIGNATURES = {
    # kernel layer
    "fftwpp_amd_last_error": (ctypes.c_char_p, []),
    "fftwpp_amd_dim_create": (I, [ctypes.POINTER(P), Z, Z, Z, I]),
    "fftwpp_amd_dim_destroy": (I, [P]),
    "fftwpp_amd_max_m": (Z, [I]),
    "fftwpp_amd_internal_engine": (I, [P, I]),
    "fftwpp_amd_internal_rows_fit": (I, [P, I]),
    "fftwpp_amd_forward_many": (I, [P, I, P, Z, Z, Z, Z, P, Z, Z, Z, Z, P]),
    "fftwpp_amd_backward_many": (I, [P, I, P, Z, Z, P, Z, Z, Z, Z, Z, Z, I, D, P]),
    "fftwpp_amd_real_planes": (Z, [P]),
    "fftwpp_amd_forward_real_many": (I, [P, I, P, Z, Z, Z, Z, P, Z, Z, P]),
    "fftwpp_amd_backward_real_many": (I, [P, I, P, Z, Z, P, Z, Z, Z, Z, D, P]),
    "fftwpp_amd_conv_rows": (I, [P, I, I, P, P, Z, Z, P, Z, D, P]),
    "fftwpp_amd_dim_kernel_args": (P, [P, I]),
    "fftwpp_amd_plane_work": (Z, [I]),
    "fftwpp_amd_plane_supported": (I, [P, P]),
    "fftwpp_amd_plane_conv": (I, [P, P, I, I, P, P, P, Z, Z, Z, P, P]),
    "fftwpp_amd_plane_status": (I, [P, P]),
    "fftwpp_amd_hconv_rows_long_part": (Z, [P, Z]),
    "fftwpp_amd_hconv_rows_long": (I, [P, I, P, P, Z, Z, P, Z, P, D, P]),
    "fftwpp_amd_batch_rows_lds": (Z, [P, I, Z, Z]),
    "fftwpp_amd_batch_rows_tps": (Z, [P, I]),
    "fftwpp_amd_batch_rows": (I, [P, I, I, P, Z, P, P, P, Z, Z, Z, Z, Z, P, Z, D, P]),
    "fftwpp_amd_batch_rows_noalloc": (I, [P, I, I, P, Z, P, P, P, Z, Z, Z, Z, Z, Z, D, P]),
    "fftwpp_amd_real_balance_bytes": (Z, []),
    "fftwpp_amd_real_balance": (I, [I, P, P, Z, P, P]),
    "fftwpp_amd_real_pack": (I, [I, P, P, P, Z, P, P]),
    "fftwpp_amd_real_unpack": (I, [I, P, P, Z, D, P, P]),
    "fftwpp_amd_multiply": (I, [I, I, P, P, Z, P]),
    "fftwpp_amd_scale": (I, [I, P, Z, D, P]),
    "fftwpp_amd_symmetrize_x": (I, [Z, Z, Z, P, Z, P]),
    "fftwpp_amd_symmetrize_xy": (I, [Z, Z, Z, Z, Z, P, Z, Z, P]),
    "fftwpp_amd_hermitian_plane": (I, [I, I, Z, Z, Z, Z, P, Z, Z, P]),
    "fftwpp_amd_malloc": (I, [ctypes.POINTER(P), Z]),
    "fftwpp_amd_free": (I, [P]),
    "fftwpp_amd_memcpy": (I, [P, P, Z, P]),
    "fftwpp_amd_memset": (I, [P, Z, P]),
    "fftwpp_amd_synchronize": (I, [P]),
    "fftwpp_amd_is_device_pointer": (I, [P]),
    "fftwpp_amd_device_count": (I, []),
    "fftwpp_amd_event_create": (I, [ctypes.POINTER(P)]),
    "fftwpp_amd_event_record": (I, [P, P]),
    "fftwpp_amd_event_elapsed": (I, [ctypes.POINTER(ctypes.c_float), P, P]),
    "fftwpp_amd_event_destroy": (I, [P]),
    # convolution layer
    "fftwpp_amd_conv_create": (I, [ctypes.POINTER(P), I, I, ZP, ZP, ZP, I]),
    "fftwpp_amd_conv_create_tuned": (I, [ctypes.POINTER(P), I, I, ZP, ZP, I, I, ZP]),
    "fftwpp_amd_conv_destroy": (I, [P]),
    "fftwpp_amd_conv_execute_fused": (I, [P, P, P, P, P, P, P]),
    "fftwpp_amd_conv_set_memory_limit": (I, [P, Z]),
    "fftwpp_amd_conv_work_bytes": (I, [P, ZP, ctypes.POINTER(ctypes.c_int)]),
    "fftwpp_amd_conv_execute": (I, [P, P, P, P]),
    "fftwpp_amd_conv_adjoint": (I, [P, P, P, P, P, P, P]),
    "fftwpp_amd_conv_geometry": (I, [P, ZP, ZP, ZP]),
    "fftwpp_amd_conv_execute_out": (I, [P, P, P, P, P]),
    "fftwpp_amd_conv_execute_timed": (I, [P, P, P, P, P, ctypes.POINTER(ctypes.c_float), I,
                                          ctypes.POINTER(I)]),
    "fftwpp_amd_conv_stage_name": (ctypes.c_char_p, [P, I]),
    "fftwpp_amd_fft_create": (I, [ctypes.POINTER(P), I, I, ZP, I, I, Z, Z, Z, Z, Z]),
    "fftwpp_amd_fft_destroy": (I, [P]),
    "fftwpp_amd_fft_execute": (I, [P, P, P, D, P]),
    "fftwpp_amd_fft_normalization": (Z, [P]),
    "fftwpp_amd_fft_normalize": (I, [P, P, P]),
    "fftwpp_amd_transpose": (I, [Z, Z, Z, P, P, P]),
    "fftwpp_amd_fft_shift": (I, [I, P, Z, Z, Z, Z, P]),
    "fftwpp_amd_conv_create_general": (I, [ctypes.POINTER(P), I, I, ZP, ZP, ZP, I, Z, Z, P, P]),
    "fftwpp_amd_conv_execute_general": (I, [P, ctypes.POINTER(P), P]),
    # batches of 1D convolutions
    "fftwpp_amd_batch_create": (I, [ctypes.POINTER(P), Z, ZP, ZP, I, I, I]),
    "fftwpp_amd_batch_destroy": (I, [P]),
    "fftwpp_amd_batch_size": (Z, [P]),
    "fftwpp_amd_batch_offsets": (I, [P, ZP]),
    "fftwpp_amd_batch_geometry": (I, [P, Z, ZP, ZP, ZP]),
    "fftwpp_amd_batch_describe": (I, [P, ctypes.c_char_p, Z]),
    "fftwpp_amd_batch_execute": (I, [P, P, P, P, P]),
    # cfftw++.h drop-in
    "create_doubleAlign": (P, [Z]),
    "delete_doubleAlign": (None, [P]),
    "create_complexAlign": (P, [Z]),
    "delete_complexAlign": (None, [P]),
    "get_fftwpp_maxthreads": (Z, []),
    "set_fftwpp_maxthreads": (None, [Z]),
    "fftwpp_create_conv1d": (P, [Z]),
    "fftwpp_conv1d_convolve": (None, [P, P, P]),
    "fftwpp_conv1d_delete": (None, [P]),
    "fftwpp_create_hconv1d": (P, [Z]),
    "fftwpp_hconv1d_convolve": (None, [P, P, P]),
    "fftwpp_hconv1d_delete": (None, [P]),
    "fftwpp_create_conv2d": (P, [Z, Z]),
    "fftwpp_conv2d_convolve": (None, [P, P, P]),
    "fftwpp_conv2d_delete": (None, [P]),
    "fftwpp_create_hconv2d": (P, [Z, Z]),
    "fftwpp_hconv2d_convolve": (None, [P, P, P]),
    "fftwpp_hconv2d_delete": (None, [P]),
    "fftwpp_create_conv3d": (P, [Z, Z, Z]),
    "fftwpp_conv3d_convolve": (None, [P, P, P]),
    "fftwpp_conv3d_delete": (None, [P]),
    "fftwpp_create_hconv3d": (P, [Z, Z, Z]),
    "fftwpp_hconv3d_convolve": (None, [P, P, P]),
    "fftwpp_hconv3d_delete": (None, [P]),
    "fftwpp_conv1d_convolve_device": (I, [P, P, P, P]),
    "fftwpp_hconv1d_convolve_device": (I, [P, P, P, P]),
    "fftwpp_conv2d_convolve_device": (I, [P, P, P, P]),
    "fftwpp_hconv2d_convolve_device": (I, [P, P, P, P]),
    "fftwpp_conv3d_convolve_device": (I, [P, P, P, P]),
    "fftwpp_hconv3d_convolve_device": (I, [P, P, P, P]),
    "fftwpp_HermitianSymmetrize": (None, [P]),
    "fftwpp_HermitianSymmetrizeX": (None, [Z, Z, Z, P]),
    "fftwpp_HermitianSymmetrizeXY": (None, [Z, Z, Z, Z, Z, P]),
}

F64, F32 = 0, 1
KIND_COMPLEX, KIND_CENTERED, KIND_HERMITIAN = 0, 1, 2
MULT_NONE, MULT_BINARY, MULT_REAL_BINARY, MULT_CORRELATION = 0, 1, 2, 3

_lib = None


class FFTWPPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fftwpp_amd error {code}: {msg}")
        self.code = code


def lib():
    """Load the in-tree library (raises when it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built; run `python -m fftwpp_amd.build`")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().fftwpp_amd_last_error()
        raise FFTWPPError(rc, msg.decode() if msg else "")


def sizes(vals):
    arr = (ctypes.c_size_t * 3)(*([int(v) for v in vals] + [0] * (3 - len(vals))))
    return arr
