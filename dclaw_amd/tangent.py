"""Tangent-linear (forward-mode) derivatives of the fused step on the GPU.

One tangent-linear step (``dclaw_stepgrid_tl_dev``) advances the state q and
its tangent dq together: q_{n+1} is the step's own result, bit for bit, and
dq_{n+1} = J_q dq_n + J_P dP, the derivative of q_{n+1} along a direction
(dq_0, dP) in (initial state, parameters).  Chaining steps gives the
derivative of the trajectory; the dt sequence is the one the values choose
(its own derivative is not carried, as for a fixed time grid).  At a branch
of the scheme the tangent is the derivative of the active branch.

Directions over parameters use the ``dclaw_params`` record's own units:
``phi`` in radians, ``kref`` in m^2, and so on (see ``param_direction``).
The gradient of a scalar function of the final state with respect to k
parameters takes k runs."""
from __future__ import annotations

import ctypes
from typing import Iterable, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .params import CParams, DigParams, to_c
from .step import (BC_EXTRAP, MBC, _ptr, _stream, aux_planes, dt_update, params_tensor,
                   to_device_q, to_host_q)

# floating fields of the parameter record (the ones a direction can seed)
PARAM_FIELDS = tuple(n for n, t in CParams._fields_ if t is ctypes.c_double)


def param_direction(**seeds: float) -> CParams:
    """A tangent record: field name -> derivative along the direction (e.g.
    ``param_direction(phi=1.0)`` differentiates with respect to phi in
    radians); unnamed fields are 0."""
    d = CParams()
    for name, v in seeds.items():
        if name not in PARAM_FIELDS:
            raise ValueError(f"{name} is not a floating field of dclaw_params")
        setattr(d, name, float(v))
    return d


def _record_tensor(rec: CParams, device="cuda") -> torch.Tensor:
    return torch.frombuffer(bytearray(bytes(rec)), dtype=torch.uint8).to(device)


def stepgrid_tl(q_in: torch.Tensor, dq_in: torch.Tensor, q_out: torch.Tensor,
                dq_out: torch.Tensor, aux: torch.Tensor, maux: int, mx: int, my: int, dx: float,
                dy: float, bc, dt_dev: torch.Tensor, params_dev: torch.Tensor,
                dparams_dev: torch.Tensor, cfl_dev: torch.Tensor) -> None:
    """One tangent-linear fused step (bc None: stored ghosts)."""
    arr = None if bc is None else (ctypes.c_int * 4)(*bc)
    _lib.check(_lib.lib().dclaw_stepgrid_tl_dev(
        _ptr(q_in), _ptr(dq_in), _ptr(q_out), _ptr(dq_out), _ptr(aux), maux, mx, my, dx, dy, arr,
        _ptr(dt_dev), _ptr(params_dev), _ptr(dparams_dev), _ptr(cfl_dev), _stream()))


class TangentRunner:
    """Device-resident run of the fused step with its tangent along one
    direction: dq0 (initial-state tangent, Fortran (7, mx+4, my+4) array,
    default 0) and dparams (a ``param_direction`` record, default 0)."""

    def __init__(self, q: np.ndarray, aux: np.ndarray, params: DigParams, dx: float, dy: float,
                 dparams: Optional[CParams] = None, dq0: Optional[np.ndarray] = None,
                 bc=(BC_EXTRAP,) * 4, device="cuda"):
        self.mx = q.shape[1] - 2 * MBC
        self.my = q.shape[2] - 2 * MBC
        self.dx, self.dy = float(dx), float(dy)
        self.bc = tuple(bc)
        self.qa = to_device_q(q, device)
        self.qb = torch.empty_like(self.qa)
        self.dqa = (torch.zeros_like(self.qa) if dq0 is None else to_device_q(dq0, device))
        self.dqb = torch.empty_like(self.qa)
        self.aux, self.maux = aux_planes(aux, params, device)
        self.params = params_tensor([params], device)
        self.dparams = _record_tensor(dparams if dparams is not None else CParams(), device)
        self.dt = torch.zeros(1, dtype=torch.float64, device=device)
        self.cfl = torch.zeros(1, dtype=torch.float64, device=device)
        self.dts: List[float] = []

    def set_dt(self, dt: float) -> None:
        self.dt.fill_(float(dt))

    def step(self) -> None:
        stepgrid_tl(self.qa, self.dqa, self.qb, self.dqb, self.aux, self.maux, self.mx, self.my,
                    self.dx, self.dy, self.bc, self.dt, self.params, self.dparams, self.cfl)
        self.qa, self.qb = self.qb, self.qa
        self.dqa, self.dqb = self.dqb, self.dqa

    def run(self, nsteps: int, dt0: float, cfl_desired: float, dt_max: float,
            record_dt: bool = False) -> None:
        """nsteps steps with the device dt rule (as PatchRunner.run); with
        record_dt the dt of every step is kept in self.dts."""
        self.set_dt(dt0)
        self.cfl.zero_()
        for _ in range(nsteps):
            if record_dt:
                self.dts.append(float(self.dt.item()))
            self.step()
            dt_update(self.dt, self.cfl, cfl_desired, dt_max)

    def run_dts(self, dts: Iterable[float]) -> None:
        """One step per given dt (a prescribed time grid)."""
        for dt in dts:
            self.set_dt(dt)
            self.cfl.zero_()
            self.step()

    def q(self) -> np.ndarray:
        torch.cuda.synchronize()
        return to_host_q(self.qa)

    def dq(self) -> np.ndarray:
        torch.cuda.synchronize()
        return to_host_q(self.dqa)


def directional_derivative(q: np.ndarray, aux: np.ndarray, params: DigParams, dx: float,
                           dy: float, dts: Sequence[float], dparams: Optional[CParams] = None,
                           dq0: Optional[np.ndarray] = None, bc=(BC_EXTRAP,) * 4):
    """(q_N, dq_N) after len(dts) steps of the given time grid."""
    r = TangentRunner(q, aux, params, dx, dy, dparams=dparams, dq0=dq0, bc=bc)
    r.run_dts(dts)
    return r.q(), r.dq()


def gauges_tl(q: torch.Tensor, dq: torch.Tensor, aux: torch.Tensor, maux: int, mx: int, my: int,
              xlower: float, ylower: float, dx: float, dy: float, dry_tolerance: float,
              xy_dev: torch.Tensor, out: torch.Tensor, dout: torch.Tensor) -> None:
    """Gauge values (ngauges, 8) into out and their tangents into dout."""
    _lib.check(_lib.lib().dclaw_gauges_tl_dev(
        _ptr(q), _ptr(dq), _ptr(aux), maux, mx, my, xlower, ylower, dx, dy, dry_tolerance,
        _ptr(xy_dev), xy_dev.shape[0], _ptr(out), _ptr(dout), _stream()))


def gauge_loss_gradient(q: np.ndarray, aux: np.ndarray, params: DigParams, dx: float, dy: float,
                        xlower: float, ylower: float, gauges_xy, targets: np.ndarray,
                        directions: dict, dt0: float = 1e-3, cfl_desired: float = 0.9,
                        dt_max: float = 1e9, component: int = 0, bc=(BC_EXTRAP,) * 4):
    """Loss = sum over steps n = 1..N and gauges g of (gauge_g(q_n)[component]
    - targets[n-1, g])^2, N = len(targets), with the device dt rule, and its
    derivative along each direction by one tangent-linear run:
    ``directions = {name: (dparams or None, dq0 or None)}``.  Returns
    (loss, {name: dloss/dname}, dts) -- dts is the time grid the values chose
    (the same in every run)."""
    targets = np.asarray(targets, dtype=np.float64)
    nsteps, ng = targets.shape
    xy = torch.tensor(np.asarray(gauges_xy, dtype=np.float64).reshape(-1, 2), device="cuda")
    if xy.shape[0] != ng:
        raise ValueError("targets must have one column per gauge")
    tgt = torch.tensor(targets, device="cuda")
    out = torch.zeros((ng, 8), dtype=torch.float64, device="cuda")
    dout = torch.zeros_like(out)
    dts = torch.zeros(nsteps, dtype=torch.float64, device="cuda")
    loss, grads = None, {}
    for name, (dparams, dq0) in directions.items():
        r = TangentRunner(q, aux, params, dx, dy, dparams=dparams, dq0=dq0, bc=bc)
        acc = torch.zeros(2, dtype=torch.float64, device="cuda")
        r.set_dt(dt0)
        r.cfl.zero_()
        for n in range(nsteps):
            dts[n].copy_(r.dt[0])
            r.step()
            dt_update(r.dt, r.cfl, cfl_desired, dt_max)
            gauges_tl(r.qa, r.dqa, r.aux, r.maux, r.mx, r.my, xlower, ylower, r.dx, r.dy,
                      params.dry_tolerance, xy, out, dout)
            res = out[:, component] - tgt[n]
            acc[0] += (res * res).sum()
            acc[1] += (2.0 * res * dout[:, component]).sum()
        torch.cuda.synchronize()
        if loss is None:
            loss = float(acc[0].item())
        grads[name] = float(acc[1].item())
    return loss, grads, dts.cpu().numpy()
