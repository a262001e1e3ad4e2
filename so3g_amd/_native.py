"""ctypes binding of libso3g_hip.so (the C ABI in include/so3g_hip.h).

The HIP library is the only compute path of this package: if it is not
built, importing the engine raises at first use with instructions, and
there is no CPU fallback.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'lib', 'libso3g_hip.so')

# Enumerations mirrored from include/so3g_hip.h.
PROJ = {'Flat': 0, 'Quat': 1, 'CAR': 2, 'CEA': 3, 'ARC': 4, 'SIN': 5, 'TAN': 6, 'ZEA': 7}
PIX_FLAT, PIX_HEALPIX = 0, 1
OK, ERR_VALUE, ERR_RUNTIME, ERR_TYPE, ERR_HIP = 0, 1, 2, 3, 4
DOMAIN_STRIPE, DOMAIN_MAP, DOMAIN_TILE = 0, 1, 2
PREC_REF, PREC_F32, PREC_F64 = 0, 1, 2

c_int, c_i32, c_i64, c_vp = ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)


class Geom(ctypes.Structure):
    """so3g_geom: engine template parameters + pixelizor arguments."""
    _fields_ = [
        ('proj', c_i32), ('pixelizor', c_i32), ('tiled', c_i32), ('interp', c_i32),
        ('ncomp', c_i32), ('naxis', c_i32 * 2), ('cdelt', ctypes.c_double * 2),
        ('crpix', ctypes.c_double * 2), ('tile_shape', c_i32 * 2), ('nside', c_i32),
        ('nest', c_i32), ('nside_tile', c_i32), ('allow_nside_max', c_i32),
        ('n_tile', c_i32), ('index_count', c_i32), ('npix', c_i64),
    ]


class Obs(ctypes.Structure):
    """so3g_obs: one observation of a batch (device pointers)."""
    _fields_ = [
        ('bore', c_vp), ('n_t', c_i64), ('dets', c_vp), ('n_det', c_i64), ('resp', c_vp),
        ('signal', c_vp), ('signal_out', c_vp), ('det_weights', c_vp),
    ]


BATCH_EAGER, BATCH_GRAPH, BATCH_PER_OBS = 0, 1, 2

# name -> (restype, argtypes)
_SIGS = {
    'so3g_init': (c_int, [c_int]),
    'so3g_last_error': (ctypes.c_char_p, []),
    'so3g_abi_version': (c_int, []),
    'so3g_get_tables': (c_int, [c_vp, c_vp]),
    'so3g_geom_init': (c_int, [ctypes.POINTER(Geom)]),
    'so3g_coords': (c_int, [ctypes.POINTER(Geom), c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    'so3g_pixels': (c_int, [ctypes.POINTER(Geom), c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    'so3g_pointing_matrix': (c_int, [ctypes.POINTER(Geom), c_vp, c_i64, c_vp, c_i64, c_vp,
                                     c_vp, c_vp, c_vp]),
    'so3g_from_map': (c_int, [ctypes.POINTER(Geom), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp,
                              c_i64, c_vp, c_vp, c_int, c_vp]),
    'so3g_to_map': (c_int, [ctypes.POINTER(Geom), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64,
                            c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    'so3g_to_weights': (c_int, [ctypes.POINTER(Geom), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp,
                                c_i64, c_vp, c_vp, c_vp, c_i64, c_vp]),
    'so3g_to_map_prec': (c_int, [ctypes.POINTER(Geom), c_vp, c_vp, c_vp, c_vp, c_i64, c_vp,
                                 c_i64, c_vp, c_vp, c_vp, c_int, c_vp]),
    'so3g_to_weights_prec': (c_int, [ctypes.POINTER(Geom), c_vp, c_vp, c_vp, c_vp, c_i64,
                                     c_vp, c_i64, c_vp, c_vp, c_int, c_vp]),
    'so3g_tile_hits': (c_int, [ctypes.POINTER(Geom), c_vp, c_i64, c_vp, c_i64, c_vp, c_vp]),
    'so3g_sample_domains': (c_int, [ctypes.POINTER(Geom), c_vp, c_i64, c_vp, c_i64, c_int,
                                    c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_create': (c_int, [c_vp, c_i32, c_vp, c_i32, c_i64, c_i64, c_vp, c_i64,
                                    ctypes.POINTER(c_vp), c_vp]),
    'so3g_pc_plan_to_map': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_to_map_set': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_to_map_f64': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_from_map': (c_int, [c_vp, c_vp, c_vp, c_int, c_vp]),
    'so3g_pc_plan_stats': (c_int, [c_vp, c_vp]),
    'so3g_pc_plan_create_tiled': (c_int, [c_vp, c_vp, c_i32, c_i64, c_i64, c_vp, c_i32, c_i64,
                                          c_vp, c_vp]),
    'so3g_pc_plan_to_map_tiles': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_to_weights': (c_int, [c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_to_weights_tiles': (c_int, [c_vp, c_vp, c_vp, c_vp]),
    'so3g_pc_plan_from_map_tiles': (c_int, [c_vp, c_vp, c_vp, c_int, c_vp]),
    'so3g_precomp_from_map_f64': (c_int, [c_vp, c_vp, c_vp, c_vp, c_i32, c_i64, c_i64, c_vp,
                                          c_int, c_vp]),
    'so3g_pc_plan_info': (c_int, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    'so3g_pc_plan_destroy': (c_int, [c_vp]),
    'so3g_batch_create': (c_int, [ctypes.POINTER(Geom), ctypes.POINTER(Obs), c_i64,
                                  ctypes.POINTER(c_vp)]),
    'so3g_batch_to_map': (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_vp]),
    'so3g_batch_from_map': (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp]),
    'so3g_batch_set_shape': (c_int, [c_vp, c_int, c_int, c_int]),
    'so3g_batch_set_table': (c_int, [c_vp, c_int]),
    'so3g_batch_tune': (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_vp, c_int,
                                ctypes.POINTER(c_int), c_vp]),
    'so3g_batch_info': (c_int, [c_vp, c_vp]),
    'so3g_batch_destroy': (c_int, [c_vp]),
    'so3g_nmat_detvecs_apply': (c_int, [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_vp, c_i64,
                                        ctypes.c_float, ctypes.c_float, c_int, c_vp]),
    'so3g_boresight': (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, ctypes.c_double, c_i64, c_vp,
                               c_vp, c_vp, c_vp]),
    'so3g_cuts_measure': (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_vp]),
    'so3g_flac_encode': (c_int, [c_vp, c_i64, c_int, c_int, c_vp, c_i64, c_vp]),
    'so3g_g3s_split': (c_int, [c_vp, c_int, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp]),
    'so3g_g3s_decode': (c_int, [c_vp, c_i64, c_vp, c_i64, c_int, c_i64, c_i64, c_i64, c_vp,
                                c_i64, c_vp, c_i64, c_i64, c_vp]),
    'so3g_g3s_stats': (c_int, [c_vp]),
    'so3g_process_cuts': (c_int, [c_vp, c_i64, c_int, c_int, c_int, c_int, c_vp, c_i64, c_i64,
                                  c_vp, c_i64, c_int, c_vp]),
    'so3g_translate_cuts': (c_int, [c_vp, c_vp, c_i64, c_int, c_int, c_int, c_i64, c_i64, c_vp,
                                    c_i64, c_vp, c_i64, c_int, c_vp]),
    'so3g_nn_conv3x3': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_vp, c_vp, c_int, c_int,
                                c_vp, c_vp]),
    'so3g_nn_conv3x3_wgrad': (c_int, [c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp,
                                      c_vp, c_vp]),
    'so3g_nn_weight_flip': (c_int, [c_vp, c_int, c_int, c_vp, c_vp]),
    'so3g_nn_relu_mask': (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    'so3g_nn_add': (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    'so3g_nn_pool': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_vp, c_int,
                             c_vp]),
    'so3g_nn_up': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_vp, c_int,
                           c_vp]),
    'so3g_nn_from_planes': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                    c_vp, c_vp]),
    'so3g_nn_to_planes': (c_int, [c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                  c_vp, c_vp]),
    'so3g_nn_adam': (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, ctypes.c_float, c_int, c_vp, c_vp]),
    'so3g_nn_to_bf16': (c_int, [c_vp, c_i64, c_vp, c_vp]),
    'so3g_set_option': (c_int, [ctypes.c_char_p, c_i64]),
    'so3g_get_option': (c_int, [ctypes.c_char_p, ctypes.POINTER(c_i64)]),
    'so3g_precomp_from_map': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32, c_i64,
                                      c_i64, c_vp, c_int, c_vp]),
    'so3g_precomp_from_map_noalloc': (c_int, [c_vp, c_vp, c_vp, c_vp, c_i32, c_i64, c_i64,
                                              c_vp, c_int, c_vp]),
    'so3g_precomp_to_map': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32, c_i64,
                                    c_i64, c_vp, c_vp, c_vp, c_i64, c_vp]),
    'so3g_precomp_to_weights': (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_vp, c_i32,
                                        c_i64, c_i64, c_vp, c_vp, c_i64, c_vp]),
}

_lib = None
_lock = threading.Lock()


class NativeMissing(RuntimeError):
    pass


def load(path=LIB_PATH):
    """Load (once) and return the ctypes library handle."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise NativeMissing(
                f'{path} is not built; run `python -m so3g_amd.build` '
                '(the HIP library is the only compute path).')
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def exported_symbols():
    return sorted(_SIGS)


def check(rc):
    """Raise the Python exception class the reference would raise."""
    if rc == OK:
        return
    msg = load().so3g_last_error().decode(errors='replace')
    if rc == ERR_VALUE:
        raise ValueError(msg)
    if rc == ERR_TYPE:
        raise TypeError(msg)
    raise RuntimeError(msg)


_initialised = set()


def init(device=None):
    """Initialise the library on the current torch device (tables, flags)."""
    import torch
    if device is None:
        device = torch.cuda.current_device()
    if device in _initialised:
        return
    check(load().so3g_init(int(device)))
    _initialised.add(device)


def get_tables():
    import numpy as np
    a = np.zeros(1 << 14)
    b = np.zeros(1 << 14)
    check(load().so3g_get_tables(a.ctypes.data, b.ctypes.data))
    return a, b


def set_option(name, value):
    check(load().so3g_set_option(name.encode(), int(value)))


def get_option(name):
    v = c_i64(0)
    check(load().so3g_get_option(name.encode(), ctypes.byref(v)))
    return int(v.value)
