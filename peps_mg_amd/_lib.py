"""ctypes binding of libpepsmg.so (declarations in include/pepsmg.h).

The product path has no fallback: if the HIP library is missing this
module raises instead of computing anything on the CPU.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpepsmg.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

OK = 0
ERR_ARG, ERR_LEVELS, ERR_GHOST, ERR_MATRIX, ERR_HIP, ERR_NOMEM, ERR_NOTCONV, ERR_DIVERGED = -1, -2, -3, -4, -5, -6, -7, -8


class PepsmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpepsmg error {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dxm", _dp), ("dxg", _dp), ("dym", _dp), ("dyg", _dp), ("dzm", _dp), ("dzg", _dp)]


class Options(C.Structure):
    _fields_ = [("smoother", C.c_int32), ("precision", C.c_int32), ("coarse_switch", C.c_int32),
                ("inner_cycles", C.c_int32), ("max_refine", C.c_int32), ("fused_smoother", C.c_int32),
                ("tile_y", C.c_int32), ("verbose", C.c_int32), ("tile_x", C.c_int32),
                ("inner_nu", C.c_int32), ("inner_omega", C.c_double),
                ("inner_nu_post", C.c_int32)]


def option_value(name, v):
    """Value of option `name` in its field's type (inner_omega is a double)."""
    return float(v) if name == "inner_omega" else int(v)


class Stats(C.Structure):
    _fields_ = [("cycles", C.c_int32), ("refine_steps", C.c_int32), ("converged", C.c_int32),
                ("history_cap", C.c_int32), ("res0", C.c_double), ("rho", C.c_double),
                ("history", _dp), ("coarse_iters", _ip), ("device_ms", C.c_double)]


# name -> (restype, argtypes); every function include/pepsmg.h declares
SIGNATURES = {
    "pepsmg_last_error": (C.c_char_p, []),
    "pepsmg_default_options": (None, [C.POINTER(Options)]),
    "pepsmg_multigrid_create": (C.c_int, [C.POINTER(Grid), C.c_int32, C.c_int32, C.c_int32, C.c_int32, _ip,
                                          C.POINTER(C.c_void_p)]),
    "pepsmg_set_options": (C.c_int, [C.c_void_p, C.POINTER(Options)]),
    "pepsmg_multigrid_solve_vcycle": (C.c_int, [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int32, C.c_double,
                                                C.c_double, C.POINTER(Stats)]),
    "pepsmg_multigrid_solve_vcycle_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                       C.c_int32, C.c_double, C.c_double, C.c_void_p,
                                                       C.POINTER(Stats)]),
    "pepsmg_multigrid_destroy": (None, [C.c_void_p]),
    "pepsmg_rbgs_iterator": (C.c_int, [C.POINTER(Grid), _dp, _dp, _dp, C.c_int32, C.c_double]),
    "pepsmg_rbgs_iterator_fp32": (C.c_int, [C.POINTER(Grid), _dp, _dp, C.c_int32, C.c_double]),
    "pepsmg_rbgs_solver": (C.c_int, [C.POINTER(Grid), _dp, _dp, _dp, C.c_int32, C.c_double, C.c_double,
                                     _ip, _dp]),
    "pepsmg_cg_iterator": (C.c_int, [C.POINTER(Grid), _dp, _dp, _dp, C.c_int32]),
    "pepsmg_cg_solver": (C.c_int, [C.POINTER(Grid), _dp, _dp, _dp, C.c_int32, C.c_double, _ip, _dp]),
    "pepsmg_mv_mul": (C.c_int, [C.POINTER(Grid), _dp, _dp, _dp]),
    "pepsmg_vv_dot": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp]),
    "pepsmg_vv_dot_device_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p]),
    "pepsmg_bench_smoother": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, _dp, _dp,
                                        C.POINTER(C.c_char_p)]),
    "pepsmg_bench_stream": (C.c_int, [C.c_int64, C.c_int32, _dp, _dp]),
    "pepsmg_kernel_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "pepsmg_kernel_timing_read": (C.c_int, [C.c_void_p, _dp, _ip, C.POINTER(C.c_char_p)]),
    "pepsmg_fill_rhs_device": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_uint64, C.c_double,
                                         C.c_double, _dp, _dp, _dp, C.c_void_p]),
    "pepsmg_fill_rhs_host": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _dp, C.c_uint64, C.c_double, C.c_double,
                                       _dp, _dp, _dp]),
    "pepsmg_batch_create": (C.c_int, [C.c_int32, C.POINTER(Grid), _ip, _ip, C.c_int32, C.POINTER(C.c_void_p)]),
    "pepsmg_batch_set_options": (C.c_int, [C.c_void_p, C.POINTER(Options)]),
    "pepsmg_batch_size": (C.c_int32, [C.c_void_p]),
    "pepsmg_batch_set_tiles": (C.c_int, [C.c_void_p, _ip, _ip]),
    "pepsmg_batch_solve_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32,
                                            C.c_double, C.c_double, C.c_int32, C.c_void_p, C.POINTER(Stats)]),
    "pepsmg_batch_destroy": (None, [C.c_void_p]),
}

_lib = None


def lib():
    """Load libpepsmg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m peps_mg_amd.build`")
        # torch ships its own HIP runtime; loading it first makes the library
        # bind to that runtime, so torch device pointers and streams passed
        # through the C-ABI belong to the same runtime instance.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH, mode=os.RTLD_NOW)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().pepsmg_last_error().decode()


def check(rc):
    if rc != OK:
        raise PepsmgError(rc, last_error())
    return rc


def dptr(a, name="array"):
    """Pointer to a C-contiguous float64 numpy array (None passes through).
    Raises PepsmgError(ERR_ARG) for anything else: the C entry points read
    and write whole padded grids through the pointer."""
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise PepsmgError(ERR_ARG, f"{name} must be a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_dp)


def padded_ptr(a, nx, ny, nz, name):
    """Pointer to a padded grid array a(0:nx+1,0:ny+1,0:nz+1), numpy shape
    (nz+2, ny+2, nx+2); a wrong shape raises PepsmgError(ERR_ARG) (the C
    entry points copy (nx+2)(ny+2)(nz+2) values in and out)."""
    if a is None:
        raise PepsmgError(ERR_ARG, f"{name} is None")
    want = (nz + 2, ny + 2, nx + 2)
    if getattr(a, "shape", None) != want:
        raise PepsmgError(ERR_ARG, f"{name} has shape {getattr(a, 'shape', None)}, expected {want} "
                                   f"(padded (0:nx+1,0:ny+1,0:nz+1) in C order)")
    return dptr(a, name)


def coeff_ptr(a_poisson, nx, ny, nz):
    """Pointer to a matrix's AoS coefficients coeff(0:6,nx,ny,nz), numpy
    shape (nz, ny, nx, 7), or None when no matrix is given."""
    coeff = None if a_poisson is None else a_poisson.coeff
    if coeff is None:
        return None
    want = (nz, ny, nx, 7)
    if getattr(coeff, "shape", None) != want:
        raise PepsmgError(ERR_ARG, f"coeff has shape {getattr(coeff, 'shape', None)}, expected {want}")
    return dptr(coeff, "coeff")


def make_grid(nx, ny, nz, dxm, dxg, dym, dyg, dzm, dzg):
    """Grid descriptor; keeps references to the metric arrays alive."""
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (dxm, dxg, dym, dyg, dzm, dzg)]
    g = Grid(nx, ny, nz, *[dptr(a) for a in arrs])
    g._keep = arrs
    return g


def default_options():
    o = Options()
    lib().pepsmg_default_options(C.byref(o))
    return o
