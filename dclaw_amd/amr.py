"""Device transfers between AMR levels (the numerical part of the coupling an
AMR driver performs around the per-patch step): fine -> coarse averaging
(update.f90:112-203) and coarse -> fine interpolation (filval.f90:190-381).

Tensors are device float64 in the library's layout: q as (ny, nx, 7) and aux
as (ny, nx, maux) -- the bytes of Fortran q(7, nx, ny) -- and needs_set as a
(ny, nx) uint8 tensor.  Indices are 0-based over the whole arrays (ghost
cells included)."""
from __future__ import annotations

import torch

from . import _lib


def _chk(t: torch.Tensor, dtype, ndim: int, name: str) -> None:
    if not (t.is_cuda and t.dtype == dtype and t.dim() == ndim and t.is_contiguous()):
        raise _lib.DClawError(f"{name} must be a contiguous {ndim}-d {dtype} device tensor")


def average_to_coarse(qc: torch.Tensor, auxc: torch.Tensor, qf: torch.Tensor,
                      auxf: torch.Tensor, ci0: int, cj0: int, ni: int, nj: int, fi0: int,
                      fj0: int, rx: int, ry: int, dry_tolerance: float) -> None:
    """Overwrite coarse cells (ci0 + a, cj0 + b), a < ni, b < nj, of qc with
    the wet-surface average of their rx x ry fine blocks of qf starting at
    (fi0 + a rx, fj0 + b ry)."""
    for t, n in ((qc, "qc"), (auxc, "auxc"), (qf, "qf"), (auxf, "auxf")):
        _chk(t, torch.float64, 3, n)
    if qc.shape[2] != 7 or qf.shape[2] != 7:
        raise _lib.DClawError("q tensors must have 7 components")
    _lib.check(_lib.lib().dclaw_amr_average_dev(
        qc.data_ptr(), auxc.data_ptr(), auxc.shape[2], qc.shape[1], qc.shape[0], qf.data_ptr(),
        auxf.data_ptr(), auxf.shape[2], qf.shape[1], qf.shape[0], ci0, cj0, ni, nj, fi0, fj0,
        rx, ry, float(dry_tolerance), torch.cuda.current_stream().cuda_stream))


def interpolate_to_fine(valc: torch.Tensor, auxc: torch.Tensor, val: torch.Tensor,
                        auxf: torch.Tensor, needs_set: torch.Tensor, rx: int, ry: int,
                        dry_tolerance: float, sea_level: float, rho_f: float,
                        grav: float) -> None:
    """Fill the fine patch val (2 ghost cells) from the coarse patch valc
    enlarged by one coarse cell all round: depth where needs_set, the other
    components over every interior coarse cell's fine block."""
    for t, n in ((valc, "valc"), (auxc, "auxc"), (val, "val"), (auxf, "auxf")):
        _chk(t, torch.float64, 3, n)
    _chk(needs_set, torch.uint8, 2, "needs_set")
    if tuple(needs_set.shape) != tuple(val.shape[:2]):
        raise _lib.DClawError("needs_set must have the fine patch's (ny, nx) shape")
    _lib.check(_lib.lib().dclaw_amr_interp_dev(
        valc.data_ptr(), auxc.data_ptr(), auxc.shape[2], valc.shape[1], valc.shape[0],
        val.data_ptr(), auxf.data_ptr(), auxf.shape[2], val.shape[1], val.shape[0],
        needs_set.data_ptr(), rx, ry, float(dry_tolerance), float(sea_level), float(rho_f),
        float(grav), torch.cuda.current_stream().cuda_stream))
