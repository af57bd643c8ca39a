"""Loader of the gfx950 library ``libdclaw_hip.so`` (ctypes, C-ABI only).

torch is imported first on purpose: torch bundles its own HIP runtime whose
SONAME is ``libamdhip64.so.7``, the same as the library's dependency, so the
library binds to the runtime torch already loaded and device pointers from
torch tensors are valid for it.  There is no fallback: if the library is
missing the import fails with instructions to build it.
"""
from __future__ import annotations

import ctypes
import os

try:  # noqa: SIM105 - torch must be loaded before the HIP library
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

from .params import CParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DCLAW_HIP_LIB", os.path.join(HERE, "_build", "libdclaw_hip.so"))

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double


class CAmrRegion(ctypes.Structure):
    """Mirror of ``dclaw_amr_region`` in include/dclaw_hip.h."""

    _fields_ = [
        ("q", _vp), ("aux", _vp), ("flag", _vp),
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("maux", ctypes.c_int),
        ("r0", ctypes.c_int), ("c0", ctypes.c_int),
        ("ilo", ctypes.c_int), ("ihi", ctypes.c_int), ("jlo", ctypes.c_int), ("jhi", ctypes.c_int),
    ]


class CAdjPatch(ctypes.Structure):
    """Mirror of ``dclaw_adj_patch`` in include/dclaw_hip.h."""

    _fields_ = [
        ("q_in", _vp), ("aux", _vp), ("lam_out", _vp), ("lam_in", _vp), ("q_out", _vp),
        ("work", _vp), ("grad", _vp),
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("maux", ctypes.c_int),
        ("member", ctypes.c_int), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
    ]


class DClawError(RuntimeError):
    pass


class CPatch(ctypes.Structure):
    """Mirror of ``dclaw_patch`` in include/dclaw_hip.h."""

    _fields_ = [
        ("q_in", _vp), ("q_out", _vp), ("aux", _vp),
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("maux", ctypes.c_int),
        ("member", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
    ]


# name -> (restype, argtypes); every symbol declared in include/dclaw_hip.h
SIGNATURES = {
    "dclaw_last_error": (ctypes.c_char_p, []),
    "dclaw_version": (ctypes.c_int, []),
    "dclaw_strip_width": (ctypes.c_int, []),
    "dclaw_set_seg_rows": (ctypes.c_int, [ctypes.c_int]),
    "dclaw_bc_fill": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _ip, _vp]),
    "dclaw_stepgrid_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          _vp, _vp, _vp, _vp]),
    "dclaw_stepgrid_bc_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             _ip, _vp, _vp, _vp, _vp]),
    "dclaw_worklist_size": (ctypes.c_int, [_ip, _ip, ctypes.c_int, ctypes.c_int]),
    "dclaw_worklist_build": (ctypes.c_int, [_ip, _ip, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int32)]),
    "dclaw_step_batched": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                          _ip, _vp, _vp, _vp, _vp, _vp]),
    "dclaw_step_batched_ws": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                          _ip, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "dclaw_tile_count": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "dclaw_split_workspace_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "dclaw_split_resident_size": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "dclaw_set_dry_rows": (ctypes.c_int, [ctypes.c_int]),
    "dclaw_set_split_mode": (ctypes.c_int, [ctypes.c_int]),
    "dclaw_gauges_tl_dev": (ctypes.c_int, [_vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _vp, _i,
                                           _vp, _vp, _vp]),
    "dclaw_step_batched_tl": (ctypes.c_int, [_vp, _vp, _i, _vp, _i, _i, _ip, _vp, _vp, _vp, _vp,
                                             _vp]),
    "dclaw_stepgrid_tl_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _ip, _vp,
                                             _vp, _vp, _vp, _vp]),
    "dclaw_adjoint_workspace_doubles": (ctypes.c_size_t, [_i, _i]),
    "dclaw_adjoint_step_dev": (ctypes.c_int, [_vp, _vp, _i, _i, _i, _d, _d, _ip, _vp, _vp, _vp, _i,
                                              _vp, _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "dclaw_adjoint_batch_doubles": (ctypes.c_size_t, [_vp, _i]),
    "dclaw_adjoint_batched": (ctypes.c_int, [_vp, _vp, _i, _ip, _vp, _vp, _vp, _i, _vp,
                                             ctypes.c_size_t, _vp]),
    "dclaw_gauges_adj_dev": (ctypes.c_int, [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _vp, _i, _vp,
                                            _vp, _vp]),
    "dclaw_amr_average_dev": (ctypes.c_int, [_vp, _vp, _i, _i, _i, _vp, _vp, _i, _i, _i, _i, _i, _i,
                                             _i, _i, _i, _i, _i, _d, _vp]),
    "dclaw_amr_region_init_dev": (ctypes.c_int, [_vp, _i, _i, _vp]),
    "dclaw_amr_copy_dev": (ctypes.c_int, [_vp, _vp, _vp, _d, _vp, _i, _i, _i, _i, _i, _i, _vp]),
    "dclaw_amr_unset_box_dev": (ctypes.c_int, [_vp, _vp, _vp]),
    "dclaw_amr_coarse_box": (None, [_i, _i, _i, _i, _i, _i, _vp]),
    "dclaw_amr_filpatch_dev": (ctypes.c_int, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i]
                               + [_d] * 10 + [_vp]),
    "dclaw_amr_region_bc_dev": (ctypes.c_int, [_vp, _i, _i, _vp, _vp]),
    "dclaw_amr_setaux_dev": (ctypes.c_int, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _i, _i, _d, _d,
                                            _vp]),
    "dclaw_amr_interp_dev": (ctypes.c_int, [_vp, _vp, _i, _i, _i, _vp, _vp, _i, _i, _i, _vp, _i,
                                            _i, _d, _d, _d, _d, _vp]),
    "dclaw_stepgrid_split_dev": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                _ip, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "dclaw_stepgrid_split_static_dev": (ctypes.c_int, [_vp, _vp, _vp, _i, _i, _i, _d, _d, _ip,
                                                       _vp, _vp, _vp, _vp]),
    "dclaw_static_workspace_ints": (ctypes.c_int, []),
    "dclaw_batched_resident_size": (ctypes.c_int, [_i, _i]),
    "dclaw_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "dclaw_destroy": (ctypes.c_int, [_vp]),
    "dclaw_handle_set_seg_rows": (ctypes.c_int, [_vp, _i]),
    "dclaw_handle_set_split_mode": (ctypes.c_int, [_vp, _i]),
    "dclaw_handle_set_dry_rows": (ctypes.c_int, [_vp, _i]),
    "dclaw_handle_tile_count": (ctypes.c_int, [_vp, _i, _i]),
    "dclaw_handle_stepgrid_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i, _i, _i, _d, _d, _ip,
                                                 _vp, _vp, _vp, _vp]),
    "dclaw_handle_state_written": (ctypes.c_int, [_vp, _vp]),
    "dclaw_curvature_workspace_doubles": (ctypes.c_size_t, [_i, _i]),
    "dclaw_stepgrid_curvature_dev": (ctypes.c_int, [_vp, _vp, _vp, _i, _i, _i, _d, _d, _ip, _vp,
                                                    _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "dclaw_src2_curvature_dev": (ctypes.c_int, [_vp, _vp, _i, _i, _i, _d, _d, _vp, _vp, _vp,
                                                ctypes.c_size_t, _vp]),
    "dclaw_dt_update": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double, _vp]),
    "dclaw_set_params": (ctypes.c_int, [ctypes.POINTER(CParams)]),
    "stepgrid_dclaw_": (None, [_dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp]),
    "stepgrid_fluxes_dclaw_": (None, [_dp, _dp, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, _dp]),
    "dclaw_dt_update_out": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_double, _vp]),
    "dclaw_gauges_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int, _vp,
                                        _vp]),
    "dclaw_check4nans_dev": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "dclaw_b4step2_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]),
    "dclaw_step2_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _vp]),
    "dclaw_src2_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                      _vp, _vp, _vp]),
    "b4step2_": (None, [_ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _ip]),
    "step2_": (None, [_ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                      _dp, _dp, _vp, _vp]),
    "src2_": (None, [_ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp]),
}

_lib = None


def lib():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DClawError(
                f"{LIB_PATH} not found: build it with `python -m dclaw_amd.build` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().dclaw_last_error()
        raise DClawError(f"dclaw error {rc}: {msg.decode() if msg else ''}")
