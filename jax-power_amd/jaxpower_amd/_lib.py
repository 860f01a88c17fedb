"""ctypes binding of libjaxpower_amd.so (declared in include/jaxpower_amd.h).

torch is imported first so that the library binds to the HIP runtime torch already loaded
(both carry the soname libamdhip64.so.7): device pointers and streams are then shared.
The product path has no fallback: if the library or a device is missing, calls raise.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library must share)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, 'lib', 'libjaxpower_amd.so')

c_int, c_int64, c_size_t, c_double, c_float, c_void_p, c_char_p = (ctypes.c_int, ctypes.c_int64, ctypes.c_size_t,
                                                                   ctypes.c_double, ctypes.c_float, ctypes.c_void_p,
                                                                   ctypes.c_char_p)
P_I32 = ctypes.POINTER(ctypes.c_int32)
P_I64 = ctypes.POINTER(ctypes.c_int64)
P_F64 = ctypes.POINTER(ctypes.c_double)

# name: (restype, argtypes)
SIGNATURES = {
    'mp_last_error': (c_char_p, []),
    'mp_version': (c_int, []),
    'mp_timing_enable': (c_int, [c_int]),
    'mp_timing_collect': (c_int, []),
    'mp_timing_reset': (c_int, []),
    'mp_timing_report': (c_int, [c_char_p, c_int64]),
    'mp_paint_workspace_bytes': (c_int64, [P_I32, c_int64, c_int, c_int]),
    'mp_paint': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, P_I32, P_F64, P_F64, c_int, c_int, c_void_p,
                         c_void_p, c_size_t]),
    'mp_paint_atomic': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, P_I32, P_F64, P_F64, c_int, c_int, c_void_p]),
    'mp_paint_spectrum_workspace_bytes': (c_int64, [P_I32, c_int64, c_int, c_int]),
    'mp_paint_spectrum': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, P_I32, P_F64, P_F64, c_int, c_int, c_int,
                                  c_void_p, c_int64, c_void_p, c_size_t]),
    'mp_paint_partial_workspace_bytes': (c_int64, [P_I32, c_int64, c_int, c_int]),
    'mp_paint_partial': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, P_I32, P_F64, P_F64, c_int, c_int, c_void_p,
                                 c_int64, c_void_p, c_size_t]),
    'mp_combine_partial_workspace_bytes': (c_int64, [P_I32, c_int]),
    'mp_combine_partial': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_int, c_int, c_int, c_int, c_void_p, c_int64, c_void_p,
                                   c_size_t]),
    'mp_combine_partial_power': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                         c_int64, c_void_p, c_size_t]),
    'mp_bin_partial_workspace_bytes': (c_int64, [P_I32, c_int, c_int, c_int]),
    'mp_bin_poles_partial': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_int, P_F64, c_int, P_I32, P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_paint_partial_batch_workspace_bytes': (c_int64, [P_I32, c_int64, c_int, c_int, c_int]),
    'mp_paint_partial_batch': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int64, c_void_p, c_int, P_I32, P_F64, P_F64, c_int,
                                       c_int, c_void_p, c_int64, c_void_p, c_size_t]),
    'mp_bin_partial_batch_workspace_bytes': (c_int64, [P_I32, c_int, c_int, c_int, c_int]),
    'mp_bin_poles_partial_batch': (c_int, [c_void_p, c_void_p, c_int64, c_int, P_I32, c_int, c_int, c_int, c_int, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_int, P_F64, c_int, P_I32, P_I32, c_void_p, c_void_p,
                                           c_size_t]),
    'mp_catalogue_sums_workspace_bytes': (c_int64, [c_int]),
    'mp_catalogue_sums': (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size_t]),
    'mp_noalloc_scratch_bytes': (c_int64, []),
    'mp_bin_poles_partial_noalloc': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                             c_void_p, c_void_p, c_int, P_F64, c_int, P_I32, P_I32, c_void_p]),
    'mp_bin_adjoint_workspace_bytes': (c_int64, [P_I32, c_int]),
    'mp_bin_adjoint': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_int, P_F64, c_int, P_I32, P_I32, c_void_p, c_void_p, c_void_p, c_size_t]),
    'mp_read_grad': (c_int, [c_void_p, c_void_p, c_int, P_I32, P_F64, P_F64, c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p,
                             c_void_p]),
    'mp_read': (c_int, [c_void_p, c_void_p, P_I32, P_F64, P_F64, c_void_p, c_int, c_int64, c_int, c_void_p]),
    'mp_spectrum_pitch': (c_int64, [ctypes.c_int32]),
    'mp_rfft3d': (c_int, [c_void_p, c_void_p, c_void_p, P_I32, c_int64]),
    'mp_irfft3d_workspace_bytes': (c_int64, [P_I32]),
    'mp_irfft3d': (c_int, [c_void_p, c_void_p, c_int64, c_void_p, P_I32, c_void_p, c_size_t]),
    'mp_bin_modes': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, P_F64, c_void_p, c_void_p]),
    'mp_bin_workspace_bytes': (c_int64, [c_int, c_int]),
    'mp_bin_poles': (c_int, [c_void_p, c_int, ctypes.POINTER(c_void_p), P_I64, P_I32, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_int, P_F64, P_F64, c_int, P_I32, P_I32, P_I32, P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_bin_real_modes': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, P_F64, c_void_p, c_void_p]),
    'mp_bin_real_workspace_bytes': (c_int64, [P_I32, c_int, c_int]),
    'mp_bin_real_poles': (c_int, [c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, P_F64, P_F64, c_int, P_I32,
                                  P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_bispec_shells_workspace_bytes': (c_size_t, [P_I32]),
    'mp_bessel_shells_ylm': (c_int, [c_void_p, c_void_p, c_int64, P_I32, P_F64, P_F64, c_int, c_int, c_int, c_int, c_double, c_double,
                                     c_void_p, c_void_p, c_void_p, c_size_t]),
    'mp_hankel_pairs_workspace_bytes': (c_size_t, [c_int]),
    'mp_hankel_pairs': (c_int, [c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_int,
                                c_void_p, c_int, c_void_p, c_void_p, c_size_t]),
    'mp_bispec_shells_ylm': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                     c_int, c_int, c_void_p, c_void_p, c_size_t]),
    'mp_bispec_shells': (c_int, [c_void_p, c_void_p, c_int64, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, P_F64,
                                 c_int, c_void_p, c_void_p, c_size_t]),
    'mp_bispec_triangles_workspace_bytes': (c_size_t, [c_int]),
    'mp_bispec_triangles': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_int, c_void_p,
                                    c_void_p, c_size_t]),
    'mp_irfft3d_bin_real_workspace_bytes': (c_int64, [P_I32, c_int, c_int]),
    'mp_irfft3d_bin_real': (c_int, [c_void_p, c_void_p, c_int64, c_int, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, P_F64,
                                    c_int, P_I32, P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_spectrum_product': (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_int64, P_I32, c_void_p, c_int64]),
    'mp_mul_ylm': (c_int, [c_void_p, c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int]),
    'mp_ylm_workspace_bytes': (c_int64, [P_I32]),
    'mp_ylm_accumulate': (c_int, [c_void_p, c_void_p, P_I32, P_F64, P_F64, c_int, c_void_p, c_int64, c_void_p, c_size_t]),
    'mp_ylm_bin_workspace_bytes': (c_int64, [P_I32, c_int, c_int]),
    'mp_ylm_bin': (c_int, [c_void_p, c_void_p, c_void_p, c_int64, P_I32, P_F64, P_F64, c_int, c_int, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_int, c_void_p, c_void_p, c_size_t]),
    'mp_window_shell': (c_int, [c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, P_F64, c_float,
                                c_void_p, c_int64]),
    'mp_window_bin_workspace_bytes': (c_int64, [P_I32, c_int, c_int]),
    'mp_window_bin': (c_int, [c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, P_F64, P_F64, P_I32,
                              P_F64, c_float, c_int, P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_tophat_shell_xi': (c_double, [c_int, c_double, c_double, c_double]),
    'mp_window_tophat_workspace_bytes': (c_int64, [P_I32, c_int, c_int]),
    'mp_window_bin_tophat': (c_int, [c_void_p, c_void_p, P_I32, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, P_F64,
                                     P_F64, P_I32, P_F64, c_double, c_int, P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_tophat_shell_xi_array': (c_int, [c_int, c_double, c_double, c_void_p, c_int64, c_void_p]),
    'mp_spherical_jn_array': (c_int, [c_int, c_void_p, c_int64, c_void_p]),
    'mp_window_bin_tophat_local': (c_int, [c_void_p, c_void_p, P_I32, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                           P_F64, P_F64, c_int, c_int, c_int, c_int, c_double, c_void_p, c_void_p, c_size_t]),
    'mp_mul_acc': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int]),
    'mp_window_shell_ylm_workspace_bytes': (c_int64, [P_I32]),
    'mp_window_shell_ylm': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, c_int, c_float,
                                    c_void_p, c_void_p, c_size_t]),
    'mp_bin_ylm_real_workspace_bytes': (c_int64, [P_I32, c_int]),
    'mp_bin_ylm_real': (c_int, [c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                                c_void_p, c_void_p, c_size_t]),
    'mp_spectrum_mean_bin': (c_int, [c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                                     c_void_p, c_int, P_F64, c_float, c_int, P_I32, c_void_p, c_void_p, c_size_t]),
    'mp_theory_ylm_real': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                                   c_float, c_void_p, c_void_p, c_size_t]),
    'mp_real_gaunt': (c_double, [c_int, c_int, c_int, c_int, c_int, c_int]),
    'mp_axpy': (c_int, [c_void_p, c_float, c_void_p, c_void_p, c_int64]),
    'mp_window_shell_ylm2': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, c_int, c_int, c_int,
                                     c_float, c_void_p, c_void_p, c_size_t]),
    'mp_mul_radial_power': (c_int, [c_void_p, c_void_p, c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_double]),
    'mp_theory_ylm2_real': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                                    c_int, c_int, c_float, c_void_p, c_void_p, c_size_t]),
    'mp_ylm_legendre_spectrum': (c_int, [c_void_p, P_I32, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, P_F64, c_void_p, c_int64]),
    'mp_reduce_workspace_bytes': (c_int64, []),
    'mp_sum_f64': (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p, c_void_p, c_size_t]),
    'mp_fill_f32': (c_int, [c_void_p, c_void_p, c_int64, c_float]),
    'mp_set_paint_path': (c_int, [c_int]),
    'mp_set_paint_tuning': (c_int, [c_int]),
    'mp_paint_tuning_report': (c_int, [c_char_p, c_int64]),
    'mp_tune_export': (c_int, [c_char_p, c_int64]),
    'mp_tune_set': (c_int, [c_char_p, c_int, c_int]),
    'mp_tune_force': (c_int, [c_int, c_int]),
    'mp_tune_enable': (c_int, [c_int]),
    'mp_tune_clear': (c_int, []),
}


class LibraryError(RuntimeError):
    pass


_lib = None


def load(path=LIBPATH):
    """Load and type the shared library (once)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LibraryError(f'{path} not found: build it with `python jax-power_amd/build_lib.py`')
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def check(rc, name=''):
    if rc != 0:
        msg = load().mp_last_error().decode(errors='replace')
        exc = NotImplementedError if rc == -2 else LibraryError
        raise exc(f'{name} failed ({rc}): {msg}')


def call(name, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc


def i32x3(values):
    return (ctypes.c_int32 * 3)(*[int(v) for v in values])


def f64x3(values):
    return (ctypes.c_double * 3)(*[float(v) for v in values])


def ptr(tensor):
    return ctypes.c_void_p(tensor.data_ptr()) if tensor is not None else ctypes.c_void_p(None)


def stream_handle(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_device(tensor, name='array'):
    if not (isinstance(tensor, torch.Tensor) and tensor.is_cuda):
        raise ValueError(f'{name} must be a torch tensor on the GPU, got {type(tensor)} on '
                         f'{getattr(tensor, "device", None)}')


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
