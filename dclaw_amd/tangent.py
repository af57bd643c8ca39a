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


class TangentEnsemble:
    """Members of uneven patch sizes advanced with their tangents by one
    launch per step (``dclaw_step_batched_tl``): member k carries the
    direction directions[k] (a ``param_direction`` record; default 0) and the
    initial tangent dq0[k] (default 0).  Values equal the forward ensemble's
    (src2 fused) bit for bit."""

    def __init__(self, members: List, directions: Optional[Sequence[CParams]] = None,
                 dq0: Optional[Sequence[Optional[np.ndarray]]] = None, bc=(BC_EXTRAP,) * 4,
                 seg_rows: int = 16, device="cuda"):
        from .step import Ensemble
        self.ens = Ensemble(members, bc=bc, seg_rows=seg_rows, device=device, split_src2=False)
        e = self.ens
        self.n = e.n
        self.dqa = [to_device_q(dq0[k], device) if dq0 is not None and dq0[k] is not None
                    else torch.zeros_like(e.qa[k]) for k in range(e.n)]
        self.dqb = [torch.empty_like(t) for t in self.dqa]
        recs = list(directions) if directions is not None else [CParams() for _ in range(e.n)]
        if len(recs) != e.n:
            raise ValueError("one direction record per member")
        self.dparams = torch.frombuffer(bytearray(b"".join(bytes(r) for r in recs)),
                                        dtype=torch.uint8).to(device)
        self._tpatches = [self._tangent_table(False, device), self._tangent_table(True, device)]

    def _tangent_table(self, flipped: bool, device) -> torch.Tensor:
        e = self.ens
        recs = []
        for k in range(e.n):
            di, do = (self.dqb[k], self.dqa[k]) if flipped else (self.dqa[k], self.dqb[k])
            recs.append(_lib.CPatch(_ptr(di), _ptr(do), _ptr(e.bed[k]), e.mx[k], e.my[k],
                                    e.maux[k], k, e.dx[k], e.dy[k]))
        raw = b"".join(bytes(r) for r in recs)
        return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)

    @property
    def dt(self) -> torch.Tensor:
        return self.ens.dt

    @property
    def cfl(self) -> torch.Tensor:
        return self.ens.cfl

    def step(self) -> None:
        """One launch for all members and their tangents."""
        e = self.ens
        bc = (ctypes.c_int * 4)(*e.bc)
        f = 1 if e.flip else 0
        _lib.check(_lib.lib().dclaw_step_batched_tl(
            _ptr(e._patches[f]), _ptr(self._tpatches[f]), e.n, _ptr(e.tiles), e.ntiles,
            e.seg_rows, bc, _ptr(e.dt), _ptr(e.params), _ptr(self.dparams), _ptr(e.cfl),
            _stream()))
        e.flip = not e.flip

    def step_plain(self) -> None:
        """The per-member loop of single-patch tangent-linear launches."""
        e = self.ens
        rec = ctypes.sizeof(CParams)
        for k in range(e.n):
            qi = e.current(k)
            qo = e.qa[k] if e.flip else e.qb[k]
            di = self.dqb[k] if e.flip else self.dqa[k]
            do = self.dqa[k] if e.flip else self.dqb[k]
            stepgrid_tl(qi, di, qo, do, e.bed[k], e.maux[k], e.mx[k], e.my[k], e.dx[k], e.dy[k],
                        e.bc, e.dt[k:k + 1], e.params[k * rec:(k + 1) * rec],
                        self.dparams[k * rec:(k + 1) * rec], e.cfl[k:k + 1])
        e.flip = not e.flip

    def run(self, nsteps: int, dt0, cfl_desired: float, dt_max: float,
            plain: bool = False) -> None:
        e = self.ens
        if np.isscalar(dt0):
            e.dt.fill_(float(dt0))
        else:
            e.dt.copy_(torch.as_tensor(dt0, dtype=torch.float64))
        e.cfl.zero_()
        for _ in range(nsteps):
            if plain:
                self.step_plain()
            else:
                self.step()
            dt_update(e.dt, e.cfl, cfl_desired, dt_max)

    def q(self, k: int) -> np.ndarray:
        torch.cuda.synchronize()
        return self.ens.q(k)

    def dq(self, k: int) -> np.ndarray:
        torch.cuda.synchronize()
        return to_host_q(self.dqb[k] if self.ens.flip else self.dqa[k])
