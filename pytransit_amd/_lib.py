"""ctypes binding of ``librr_hip.so`` (the C-ABI declared in ``include/rr_hip.h``).

The library is built in-tree by :func:`pytransit_amd.build.build_library`.  It must
share torch's HIP runtime (torch bundles its own ``libamdhip64``): torch is imported
first, the library is loaded after it, and :func:`load` checks that exactly one
``libamdhip64`` is mapped into the process.  There is no fallback: when the
library or a GPU is missing, :func:`load` raises.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_size_t, c_void_p

import torch

LIB_NAME = 'librr_hip.so'
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

RR_OK = 0
RR_FP64 = 0
RR_MIXED = 1

# Limb-darkening law ids (RR_LD_* in include/rr_hip.h), keyed by the reference's
# RoadRunnerModel.ldmodels names (rrmodel.py:48-58); value: (id, minimum coefficients).
LD_LAW_IDS = {
    'uniform': (0, 0), 'linear': (1, 1), 'quadratic': (2, 2), 'quadratic-tri': (3, 2),
    'nonlinear': (4, 4), 'general': (5, 0), 'square_root': (6, 2), 'logarithmic': (7, 2),
    'exponential': (8, 2), 'power-2': (9, 2), 'power-2-pm': (10, 2),
}


RR_LD_TABULATED = -1


class RRTable(ctypes.Structure):
    _fields_ = [('nk', c_int32), ('ng', c_int32), ('nz', c_int32), ('pad_', c_int32),
                ('kmin', c_double), ('kmax', c_double), ('dk', c_double), ('dg', c_double),
                ('weights', c_void_p), ('weights32', c_void_p), ('ze', c_void_p), ('mu', c_void_p)]


class RRLd(ctypes.Structure):
    _fields_ = [('law', c_int32), ('nldc', c_int32), ('ldc', c_void_p), ('ldp', c_void_p), ('istar', c_void_p)]


class RRLcData(ctypes.Structure):
    _fields_ = [('npt', c_int32), ('nlc', c_int32), ('npb', c_int32), ('nep', c_int32),
                ('sorted', c_int32), ('uniform_nsamples', c_int32),
                ('times', c_void_p), ('lcids', c_void_p), ('pbids', c_void_p), ('epids', c_void_p),
                ('nsamples', c_void_p), ('exptimes', c_void_p)]


class RRParams(ctypes.Structure):
    _fields_ = [('npv', c_int32), ('kcols', c_int32), ('t0cols', c_int32), ('pad_', c_int32),
                ('k', c_void_p), ('t0', c_void_p), ('p', c_void_p), ('a', c_void_p), ('inc', c_void_p),
                ('e', c_void_p), ('w', c_void_p)]


class RRRequest(ctypes.Structure):
    _fields_ = [('lc', RRLcData), ('pv', RRParams), ('ld', RRLd), ('workspace', c_void_p),
                ('workspace_bytes', c_size_t), ('flux', c_void_p)]


class RRBatchShape(ctypes.Structure):
    _fields_ = [('setup_blocks', c_int32), ('setup_multi', c_int32), ('setup_lds', c_int32),
                ('flux_blocks', c_int32 * 2), ('flux_lds', c_int32 * 2)]


# name -> (restype, argtypes)
_SIGNATURES = {
    'rr_last_error': (ctypes.c_char_p, []),
    'rr_abi_version': (c_int, []),
    'rr_z_grid': (c_int, [c_double, c_int, c_int, POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    'rr_weights_3d': (c_int, [c_int, c_double, c_double, c_void_p, c_int, c_int, c_void_p, c_void_p,
                              POINTER(c_double), POINTER(c_double), c_void_p]),
    'rr_ld_profiles': (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    'rr_integrate_profiles': (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    'rr_ldtk_interpolate': (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_double, c_double, c_double,
                                    c_double, c_double, c_double, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    'rr_orbit': (c_int, [c_int] + [c_void_p] * 9),
    'rr_workspace_bytes': (c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    'rr_setup': (c_int, [c_int, POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), POINTER(RRLd), c_void_p,
                         c_size_t, c_void_p]),
    'rr_flux': (c_int, [c_int, POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), c_void_p, c_size_t,
                        c_void_p, c_void_p]),
    'rr_evaluate': (c_int, [c_int, POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), POINTER(RRLd),
                            c_void_p, c_size_t, c_void_p, c_void_p]),
    'ts_workspace_bytes': (c_size_t, [c_int, c_int, c_int, c_int, c_int]),
    'rr_lnlike': (c_int, [c_int, POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), POINTER(RRLd), c_void_p,
                          c_void_p, c_void_p, c_void_p, c_void_p]),
    'ts_evaluate': (c_int, [c_int, POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), POINTER(RRLd),
                            c_void_p, c_size_t, c_void_p, c_void_p]),
    'rr_max_batch': (c_int, []),
    'ma_workspace_bytes': (c_size_t, [c_int]),
    'ma_evaluate': (c_int, [POINTER(RRLcData), POINTER(RRParams), c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    'rr_evaluate_batch': (c_int, [c_int, POINTER(RRTable), c_int, POINTER(RRRequest), c_void_p]),
    'rr_batch_desc_bytes': (c_size_t, []),
    'rr_batch_plan': (c_int, [c_int, POINTER(RRTable), c_int, POINTER(RRRequest), c_void_p,
                              POINTER(RRBatchShape)]),
    'rr_batch_launch': (c_int, [c_int, c_void_p, POINTER(RRBatchShape), c_void_p]),
    'ec_workspace_bytes': (c_size_t, [c_int]),
    'eclipse_evaluate': (c_int, [POINTER(RRLcData), POINTER(RRParams), c_void_p, c_void_p, c_size_t, c_void_p,
                                 c_void_p]),
    'es_evaluate': (c_int, [POINTER(RRLcData), POINTER(RRParams), c_void_p, c_void_p, c_int, c_void_p, c_size_t,
                            c_void_p, c_void_p]),
    'op_evaluate': (c_int, [POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), POINTER(RRLd), c_void_p,
                            c_void_p, c_int, c_void_p, c_size_t, c_void_p, c_void_p]),
    'rr_evaluate_noalloc': (c_int, [c_int, POINTER(RRTable), POINTER(RRLcData), POINTER(RRParams), POINTER(RRLd),
                                    c_void_p, c_void_p]),
}
RR_NOALLOC_SCRATCH_BYTES = 32 << 20

_lib = None


class RRError(RuntimeError):
    """A C-ABI call returned an error code."""


def _hip_runtimes_mapped():
    try:
        with open('/proc/self/maps') as f:
            return sorted({line.split()[-1] for line in f if 'libamdhip64' in line})
    except OSError:
        return []


def load(require_gpu: bool = True):
    """Load the library (once) and return the ctypes handle.

    Raises RuntimeError if the built library is missing, if it would bind a second
    HIP runtime next to torch's, or (``require_gpu``) if no GPU is visible.
    """
    global _lib
    if require_gpu and not torch.cuda.is_available():
        raise RuntimeError('pytransit_amd needs a ROCm GPU; torch.cuda.is_available() is False')
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'{LIB_PATH} is missing: build it with pytransit_amd.build.build_library() '
                           f'(or __graft_entry__.build()) before using the GPU model')
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    runtimes = _hip_runtimes_mapped()
    if len(runtimes) > 1:
        raise RuntimeError(f'two HIP runtimes are mapped ({runtimes}); librr_hip.so must bind torch\'s')
    _lib = lib
    return lib


def check(rc: int):
    if rc != RR_OK:
        msg = _lib.rr_last_error().decode() if _lib is not None else ''
        raise RRError(f'librr_hip error {rc}: {msg}')


def to_host(t: torch.Tensor):
    """A device tensor copied to a new numpy array through page-locked memory.

    The copy goes at the link's speed (a pageable destination takes the driver's
    bounce-buffer path, several times slower for flux arrays of hundreds of MB);
    torch's caching host allocator hands the page-locked block back for reuse
    once the returned array is released."""
    if t.device.type != 'cuda':
        return t.numpy()
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h.numpy()


def ptr(t) -> int:
    """Device (or host) address of a tensor, 0 for None."""
    return 0 if t is None else t.data_ptr()


def stream_handle(device) -> int:
    """hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def z_grid(zcut: float, nin: int, nedge: int):
    """Host call of rr_z_grid; returns (ze, zm, mu) as float64 numpy arrays."""
    import numpy as np
    lib = load(require_gpu=False)
    n = nin + nedge
    ze = np.zeros(n)
    zm = np.zeros(n)
    mu = np.zeros(n)
    dp = ctypes.POINTER(c_double)
    check(lib.rr_z_grid(zcut, nin, nedge, ze.ctypes.data_as(dp), zm.ctypes.data_as(dp), mu.ctypes.data_as(dp)))
    return ze, zm, mu
