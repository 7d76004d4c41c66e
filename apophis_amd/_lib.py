"""ctypes binding of the C ABI declared in include/apophis_hydro.h.

The shared library is built in-tree (apophis_amd/_build/libapophis_hydro.so) by
``apophis_amd.build``.  There is no fallback: if the library is missing or a call
fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libapophis_hydro.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "apophis_hydro.h")

APH_OK = 0
ERRORS = {-1: "APH_E_INVALID", -2: "APH_E_HIP", -3: "APH_E_NOMEM", -4: "APH_E_STATE"}

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_d = ctypes.c_double
_pp = ctypes.POINTER(ctypes.c_void_p)


class HydroError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Tile(ctypes.Structure):
    _fields_ = [("n", _i32), ("tx", _i32), ("ty", _i32), ("tz", _i32)]


# name -> (restype, argtypes)
_SIGS = {
    "aph_abi_version": (_i32, []),
    "aph_context_create": (_i32, [_vp, _pp]),
    "aph_context_destroy": (_i32, [_vp]),
    "aph_last_error": (ctypes.c_char_p, [_vp]),
    "aph_context_synchronize": (_i32, [_vp]),
    "aph_blockset_create": (_i32, [_vp, _vp, _i32, _i32, _i32, _pp]),
    "aph_blockset_destroy": (_i32, [_vp, _vp]),
    "aph_blockset_ncells": (_i64, [_vp]),
    "aph_cons_to_prim": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp]),
    "aph_calculate_fluxes": (_i32, [_vp, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    "aph_update_with_flux_divergence": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                                               _d, _d, _d]),
    "aph_apply_gravity": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _d]),
    "aph_fill_ghosts": (_i32, [_vp, _vp, _i32, _i32, _vp]),
    "aph_estimate_timestep": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _d, _vp]),
    "aph_global_sum": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "aph_global_sum_noalloc": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "aph_amr_plan_create": (_i32, [_vp, _vp, _vp, _i64, _vp, _i64, _vp, _i64, _pp]),
    "aph_amr_plan_destroy": (_i32, [_vp, _vp]),
    "aph_amr_build": (_i32, [_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp, _vp, ctypes.c_char_p, ctypes.c_size_t]),
    "aph_amr_regrid_tree": (_i32, [_i32, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i32,
                                   ctypes.c_char_p, ctypes.c_size_t]),
    "aph_amr_block_table": (_i32, [_i32, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _i32, _vp, _vp,
                                   _vp, ctypes.c_char_p, ctypes.c_size_t]),
    "aph_amr_remap": (_i32, [_i32, _vp, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp, _vp,
                             _vp, ctypes.c_char_p, ctypes.c_size_t]),
    "aph_amr_fill": (_i32, [_vp, _vp, _vp]),
    "aph_amr_flux_correct": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "aph_plan_set_amr": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64]),
    "aph_refinement_tag": (_i32, [_vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "aph_plan_create": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _pp]),
    "aph_plan_destroy": (_i32, [_vp, _vp]),
    "aph_plan_set_tile": (_i32, [_vp, _vp, ctypes.POINTER(Tile)]),
    "aph_plan_set_integrator": (_i32, [_vp, _vp, _i32]),
    "aph_plan_status": (_i32, [_vp, _vp]),
    "aph_plan_set_wait_limit": (_i32, [_vp, _vp, _i32]),
    "aph_plan_autotune": (_i32, [_vp, _vp, _vp, _vp, _vp, _i32, ctypes.c_char_p,
                                 ctypes.c_size_t]),
    "aph_plan_get_tile": (_i32, [_vp, _vp, _i32, ctypes.POINTER(Tile)]),
    "aph_plan_get_launch": (_i32, [_vp, _vp, _i32, ctypes.POINTER(ctypes.c_int32)]),
    "aph_plan_set_workgroup": (_i32, [_vp, _vp, _i32]),
    "aph_plan_set_lds_reserve": (_i32, [_vp, _vp, _i64]),
    "aph_apply_gravity_field": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _vp, _d]),
    "aph_monopole_create": (_i32, [_vp, _i32, _i32, _d, _d, ctypes.POINTER(_vp)]),
    "aph_monopole_destroy": (_i32, [_vp, _vp]),
    "aph_monopole_zero_radial_bins": (_i32, [_vp, _vp, _vp, _vp]),
    "aph_monopole_accumulate_mass": (_i32, [_vp, _vp, _vp, _i32, _i32, _vp]),
    "aph_monopole_build_radial_tables": (_i32, [_vp, _vp]),
    "aph_monopole_gravity_field": (_i32, [_vp, _vp, _vp, _i32, _i32, _vp, _vp]),
    "aph_monopole_get_tables": (_i32, [_vp, _vp, _vp, _vp]),
    "aph_plan_set_monopole": (_i32, [_vp, _vp, _vp]),
    "aph_poisson_create": (_i32, [_vp, _vp, _d, _d, _d, _i32, ctypes.POINTER(_vp)]),
    "aph_poisson_destroy": (_i32, [_vp, _vp]),
    "aph_poisson_create_tree": (_i32, [_vp, _vp, _vp, _vp, _vp, _i32, _i32, _d, _d, _d, _i32,
                                       ctypes.POINTER(_vp)]),
    "aph_poisson_tree_matrix": (_i32, [_vp, _i32, _vp, _vp, _vp, _i32, _i32, _d, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _vp, _vp, ctypes.c_char_p,
                                       ctypes.c_size_t]),
    "aph_poisson_solve": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "aph_poisson_get_info": (_i32, [_vp, _vp, _vp, ctypes.POINTER(_i32),
                                    ctypes.POINTER(_d)]),
    "aph_plan_set_poisson": (_i32, [_vp, _vp, _vp]),
    "aph_fused_stage": (_i32, [_vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "aph_plan_step": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i32]),
    "aph_plan_graph_capture": (_i32, [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _d, _d, _i32]),
    "aph_plan_graph_launch": (_i32, [_vp, _vp]),
    "aph_plan_estimate_timestep": (_i32, [_vp, _vp, _vp, _d, _vp]),
    "aph_adjoint_stage": (_i32, [_vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "aph_adjoint_step": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "aph_adjoint_step_stored": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


def header_symbols():
    """Function names declared in include/apophis_hydro.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(aph_[a-z0-9_]+)\s*\(", txt)))


def available():
    return os.path.exists(LIB_PATH)


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"apophis_amd native library not found at {LIB_PATH}; "
                "build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        if L.aph_abi_version() != 1:
            raise RuntimeError("apophis_amd: ABI version mismatch")
        _lib = L
    return _lib


class Context:
    """An aph_context bound to a HIP stream (int handle; 0 = default stream)."""

    def __init__(self, stream=0):
        self.h = _vp()
        rc = lib().aph_context_create(_vp(stream or None), ctypes.byref(self.h))
        if rc != APH_OK:
            raise HydroError(rc, "aph_context_create failed (no HIP device?)")
        self.stream = stream

    def check(self, rc):
        if rc != APH_OK:
            raise HydroError(rc, lib().aph_last_error(self.h).decode())
        return rc

    def synchronize(self):
        self.check(lib().aph_context_synchronize(self.h))

    def close(self):
        if self.h:
            lib().aph_context_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
