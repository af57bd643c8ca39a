"""Reverse-mode (adjoint) derivatives of the fused step on the GPU.

``dclaw_adjoint_step_dev`` maps lambda_{n+1} (the gradient of a loss with
respect to q_{n+1}) to lambda_n = J_q^T lambda_{n+1} and gives the gradient
contribution of up to four parameter directions, where J is the Jacobian of
the fused step (b4step2, step2 with rpn2/rpt2/flux2, update, src2) at q_n for
the step's dt (the dt sequence is the one the values choose and is not
differentiated, as for a fixed time grid).  J is the tangent-linear step's
(``tangent.py``): the adjoint is its transpose, and one reverse sweep gives
the gradient of a scalar loss with respect to the whole initial state and
the parameters together.

A run of N steps is reversed with two levels of checkpoints: the forward
pass keeps the state every ``checkpoint_every`` steps; the reverse pass
recomputes, from a block's checkpoint, the states at the block's
sub-checkpoints (every ``sub_every`` steps, by default about the square root
of the block length), then each sub-block's states from its sub-checkpoint
(bit for bit: the step is deterministic), and runs the adjoint steps back
through them.  At most N / checkpoint_every + checkpoint_every / sub_every +
sub_every + 1 states are held at once, for about two extra forward steps per
step.

``rollout`` is the ``torch.autograd.Function`` form: gauge records and the
final state of an N-step run as functions of the initial state and of named
parameters, differentiable by ``torch.autograd``."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .params import CParams, DigParams, to_c
from .step import (BC_EXTRAP, _ptr, _stream, aux_planes, dt_update, gauges, stepgrid_bc,
                   to_device_q, to_host_q)
from .tangent import PARAM_FIELDS, _record_tensor, param_direction

MAX_DIRECTIONS = 4


def workspace_doubles(mx: int, my: int) -> int:
    """Doubles of device scratch one adjoint step of an mx x my patch needs."""
    return int(_lib.lib().dclaw_adjoint_workspace_doubles(mx, my))


def new_workspace(mx: int, my: int, device="cuda") -> torch.Tensor:
    """A zero-filled adjoint workspace for mx x my patches."""
    return torch.zeros(workspace_doubles(mx, my), dtype=torch.float64, device=device)


def adjoint_step(q_in: torch.Tensor, aux: torch.Tensor, maux: int, mx: int, my: int, dx: float,
                 dy: float, bc, dt_dev: torch.Tensor, params_dev: torch.Tensor,
                 dparams_dev: Optional[torch.Tensor], np_dirs: int, lam_out: torch.Tensor,
                 lam_in: torch.Tensor, q_out: Optional[torch.Tensor],
                 grad_dev: Optional[torch.Tensor], work: torch.Tensor) -> None:
    """One adjoint step on device buffers (see include/dclaw_hip.h).  work
    must be zero-filled before its first use (``new_workspace``) and passed
    unchanged between calls for the same mx, my."""
    arr = None if bc is None else (ctypes.c_int * 4)(*bc)
    _lib.check(_lib.lib().dclaw_adjoint_step_dev(
        _ptr(q_in), _ptr(aux), maux, mx, my, dx, dy, arr, _ptr(dt_dev), _ptr(params_dev),
        None if dparams_dev is None else _ptr(dparams_dev), np_dirs, _ptr(lam_out), _ptr(lam_in),
        None if q_out is None else _ptr(q_out), None if grad_dev is None else _ptr(grad_dev),
        _ptr(work), work.numel(),
        _stream()))


def gauges_adj(q: torch.Tensor, aux: torch.Tensor, maux: int, mx: int, my: int, xlower: float,
               ylower: float, dx: float, dy: float, dry_tolerance: float, xy_dev: torch.Tensor,
               lam_gauges: torch.Tensor, lam_q: torch.Tensor) -> None:
    """lam_q += (d gauges / dq)^T lam_gauges (ngauges x 8)."""
    _lib.check(_lib.lib().dclaw_gauges_adj_dev(
        _ptr(q), _ptr(aux), maux, mx, my, xlower, ylower, dx, dy, dry_tolerance, _ptr(xy_dev),
        xy_dev.shape[0], _ptr(lam_gauges), _ptr(lam_q), _stream()))


def _directions(names: Sequence[str]) -> List[CParams]:
    if len(names) > MAX_DIRECTIONS:
        raise ValueError(f"at most {MAX_DIRECTIONS} parameters per run")
    for n in names:
        if n not in PARAM_FIELDS:
            raise ValueError(f"{n} is not a floating field of dclaw_params")
    return [param_direction(**{n: 1.0}) for n in names]


@dataclass
class RolloutSpec:
    """A fixed-patch run: the bed (aux), grid, parameter record, boundary
    conditions, dt rule (dt0, cfl_desired, dt_max), step count and gauges
    (x, y pairs; none for a run without gauge records).  ``parameters`` names
    the dclaw_params fields the run is differentiated with respect to (their
    values come from ``theta`` in ``rollout``)."""
    aux: np.ndarray
    params: DigParams
    dx: float
    dy: float
    nsteps: int
    dt0: float = 1e-3
    cfl_desired: float = 0.9
    dt_max: float = 1e9
    bc: Tuple[int, int, int, int] = (BC_EXTRAP,) * 4
    gauges_xy: Optional[np.ndarray] = None
    xlower: float = 0.0
    ylower: float = 0.0
    parameters: Tuple[str, ...] = ()
    checkpoint_every: int = 100
    sub_every: int = 0
    device: str = "cuda"


def sub_interval(every: int, sub: int) -> int:
    """Sub-checkpoint interval inside a block of `every` steps (sub <= 0: the
    integer nearest sqrt(every), which minimises the states held)."""
    if sub > 0:
        return min(sub, every)
    return max(1, int(round(every ** 0.5)))


def reverse_blocks(nsteps: int, every: int, sub: int, ckpt: Dict, advance, reverse_step,
                   load=None) -> None:
    """Drive the reverse sweep with two checkpoint levels.

    ckpt[b0] is the state at each block start b0 (multiples of `every`);
    advance(x, n) returns a new state q_{n+1} from q_n = x; reverse_step(n,
    q_n, q_{n+1}) runs the adjoint of step n+1.  Steps are reversed from the
    last to the first; within a block the states at the sub-checkpoints are
    recomputed first, then each sub-block's states, which are dropped once
    it has been reversed.  load(ckpt[b0]) gives the device state of a block
    start (checkpoints may be kept in host memory)."""
    s = sub_interval(every, sub)
    for b0 in sorted(ckpt, reverse=True):
        b1 = min(b0 + every, nsteps)
        starts = list(range(b0, b1, s))
        x = ckpt[b0] if load is None else load(ckpt[b0])
        sc = {b0: x}
        for n in range(b0, starts[-1]):
            x = advance(x, n)
            if n + 1 in starts:
                sc[n + 1] = x
        for c0 in reversed(starts):
            c1 = min(c0 + s, b1)
            states = [sc.pop(c0)]
            for n in range(c0, c1):
                states.append(advance(states[-1], n))
            for n in range(c1 - 1, c0 - 1, -1):
                reverse_step(n, states[n - c0], states[n + 1 - c0])
            del states


class Rollout:
    """N steps of the fused step with gauge records, and their reverse pass.

    forward(q0, record) runs the device loop (boundary conditions on load,
    device dt rule), keeping a checkpoint every ``checkpoint_every`` steps and
    the dt of every step; it returns the gauge records (nsteps, ngauges, 8)
    and the final state.  backward(g_series, g_final) runs the adjoint back
    through the run: returns lambda_0 (the gradient with respect to q0, layout
    of q; ghost entries 0) and the gradient with respect to each named
    parameter."""

    def __init__(self, spec: RolloutSpec):
        if spec.nsteps < 1:
            raise ValueError("nsteps must be >= 1")
        if spec.checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.spec = spec
        dev = spec.device
        aux = np.asfortranarray(spec.aux)
        self.mx, self.my = aux.shape[1] - 4, aux.shape[2] - 4
        self.aux, self.maux = aux_planes(aux, spec.params, dev)
        self.dirs = _directions(spec.parameters)
        self.np = len(self.dirs)
        self.dparams = (torch.cat([_record_tensor(d, dev) for d in self.dirs])
                        if self.dirs else None)
        # zero-filled before the first call; the library keeps it so
        self.work = new_workspace(self.mx, self.my, dev)
        self.xy = None
        self.ng = 0
        if spec.gauges_xy is not None:
            xy = np.asarray(spec.gauges_xy, dtype=np.float64).reshape(-1, 2)
            self.xy = torch.from_numpy(np.ascontiguousarray(xy)).to(dev)
            self.ng = xy.shape[0]

    # parameter record with the named fields set to theta
    def record(self, theta: Optional[Sequence[float]] = None) -> torch.Tensor:
        c = to_c(self.spec.params)
        if theta is not None:
            if len(theta) != self.np:
                raise ValueError("one value per named parameter")
            for n, v in zip(self.spec.parameters, theta):
                setattr(c, n, float(v))
        return _record_tensor(c, self.spec.device)

    def forward(self, q0: torch.Tensor, params_dev: torch.Tensor):
        s = self.spec
        self.params_dev = params_dev
        dev = q0.device
        qa = q0.detach().clone()
        qb = torch.empty_like(qa)
        dt = torch.full((1,), float(s.dt0), dtype=torch.float64, device=dev)
        cfl = torch.zeros(1, dtype=torch.float64, device=dev)
        self.dts = torch.zeros(s.nsteps, dtype=torch.float64, device=dev)
        self.ckpt: Dict[int, torch.Tensor] = {0: qa.clone()}
        series = torch.zeros((s.nsteps, max(self.ng, 1), 8), dtype=torch.float64, device=dev)
        for n in range(s.nsteps):
            self.dts[n].copy_(dt[0])
            stepgrid_bc(qa, qb, self.aux, self.maux, self.mx, self.my, s.dx, s.dy, s.bc, dt,
                        params_dev, cfl)
            qa, qb = qb, qa
            dt_update(dt, cfl, s.cfl_desired, s.dt_max)
            if self.ng:
                gauges(qa, self.aux, self.maux, self.mx, self.my, s.xlower, s.ylower, s.dx, s.dy,
                       s.params.dry_tolerance, self.xy, series[n])
            if (n + 1) % s.checkpoint_every == 0 and n + 1 < s.nsteps:
                self.ckpt[n + 1] = qa.clone()
        return series[:, :self.ng], qa

    def backward(self, g_series: Optional[torch.Tensor], g_final: Optional[torch.Tensor]):
        s = self.spec
        N = s.nsteps
        dev = self.dts.device
        q0 = self.ckpt[0]
        lam = (g_final.detach().clone().contiguous() if g_final is not None
               else torch.zeros_like(q0))
        lam_b = torch.empty_like(lam)
        grad = torch.zeros(max(self.np, 1), dtype=torch.float64, device=dev)
        gstep = torch.zeros_like(grad)
        cfl = torch.zeros(1, dtype=torch.float64, device=dev)
        dt = torch.zeros(1, dtype=torch.float64, device=dev)
        lams = [lam, lam_b]

        def advance(x, n):
            dt.copy_(self.dts[n:n + 1])
            y = torch.empty_like(q0)
            stepgrid_bc(x, y, self.aux, self.maux, self.mx, self.my, s.dx, s.dy, s.bc, dt,
                        self.params_dev, cfl)
            return y

        def reverse_step(n, qn, qn1):
            # step n+1 maps q_n -> q_{n+1}; its gauge record's adjoint joins first
            lo, li = lams
            if self.ng and g_series is not None:
                gauges_adj(qn1, self.aux, self.maux, self.mx, self.my, s.xlower, s.ylower, s.dx,
                           s.dy, s.params.dry_tolerance, self.xy,
                           g_series[n].detach().contiguous(), lo)
            dt.copy_(self.dts[n:n + 1])
            adjoint_step(qn, self.aux, self.maux, self.mx, self.my, s.dx, s.dy, s.bc, dt,
                         self.params_dev, self.dparams, self.np, lo, li, None,
                         gstep if self.np else None, self.work)
            if self.np:
                grad.add_(gstep)
            lams.reverse()

        reverse_blocks(N, s.checkpoint_every, s.sub_every, self.ckpt, advance, reverse_step)
        return lams[0], grad[:self.np]


def _adj_records(rows) -> tuple:
    """(host array, device copy) of dclaw_adj_patch records."""
    arr = (_lib.CAdjPatch * len(rows))(*rows)
    dev = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to("cuda")
    return arr, dev


class EnsembleAdjoint:
    """The adjoint step of every member of an ``step.Ensemble`` in one set of
    launches (``dclaw_adjoint_batched``): member k's lambda_{n+1} at its
    state q_n -> lambda_n, and d/d(parameters) for the named parameters (the
    same names for every member, each member's own values), with the members'
    own dt.  Each member keeps its own workspace.  ``step_plain`` is the
    per-member loop of single-patch adjoint steps, for comparison (on the
    same workspaces)."""

    def __init__(self, ens, parameters: Sequence[str] = ("phi", "kref")):
        self.ens = ens
        self.dirs = _directions(parameters)
        self.np = len(self.dirs)
        n = ens.n
        dev = ens.dt.device
        # one workspace per member, valid for the batched and the single-patch
        # call alike (same layout and contents between calls)
        self.work = [new_workspace(ens.mx[k], ens.my[k], dev) for k in range(n)]
        self.grad = torch.zeros((n, max(self.np, 1)), dtype=torch.float64, device=dev)
        # member k's directions at dparams[k * np .. k * np + np - 1]
        self.dparams = (torch.cat([_record_tensor(d, dev) for _ in range(n) for d in self.dirs])
                        if self.np else None)
        self._rec = ctypes.sizeof(CParams)
        self.batch = None

    def step(self, q: Sequence[torch.Tensor], lam_out: Sequence[torch.Tensor],
             lam_in: Sequence[torch.Tensor]) -> torch.Tensor:
        """lam_in[k] = J_k^T lam_out[k] at q[k] for every member (one launch per
        pass); returns the (n, np) parameter gradients (device)."""
        e = self.ens
        rows = []
        for k in range(e.n):
            rows.append(_lib.CAdjPatch(
                _ptr(q[k]), _ptr(e.bed[k]), _ptr(lam_out[k]), _ptr(lam_in[k]), None,
                _ptr(self.work[k]), _ptr(self.grad[k]) if self.np else None, e.mx[k], e.my[k],
                e.maux[k], k, e.dx[k], e.dy[k]))
        host, dev = _adj_records(rows)
        L = _lib.lib()
        nb = int(L.dclaw_adjoint_batch_doubles(host, e.n))
        if self.batch is None or self.batch.numel() < nb:
            self.batch = torch.empty(nb, dtype=torch.float64, device=e.dt.device)
        bc = (ctypes.c_int * 4)(*e.bc)
        _lib.check(L.dclaw_adjoint_batched(
            _ptr(dev), host, e.n, bc, _ptr(e.dt), _ptr(e.params),
            None if self.dparams is None else _ptr(self.dparams), self.np, _ptr(self.batch),
            self.batch.numel(), _stream()))
        self._keep = dev     # the record copy stays alive until the launches are queued
        return self.grad[:, :self.np]

    def step_plain(self, q: Sequence[torch.Tensor], lam_out: Sequence[torch.Tensor],
                   lam_in: Sequence[torch.Tensor]) -> torch.Tensor:
        """The same by one single-patch adjoint step per member."""
        e = self.ens
        for k in range(e.n):
            rec = e.params[k * self._rec:(k + 1) * self._rec]
            dpar = (self.dparams[k * self.np * self._rec:(k + 1) * self.np * self._rec]
                    if self.np else None)
            adjoint_step(q[k], e.bed[k], e.maux[k], e.mx[k], e.my[k], e.dx[k], e.dy[k], e.bc,
                         e.dt[k:k + 1], rec, dpar, self.np, lam_out[k], lam_in[k], None,
                         self.grad[k] if self.np else None, self.work[k])
        return self.grad[:, :self.np]


class EnsembleRollout:
    """N steps of an ensemble (one batched launch per step, each member with
    its own parameters and dt rule) with per-member gauge records, and the
    reverse pass of all members together (one batched adjoint launch per
    pass per step, checkpointed recompute): the particle-filter form of
    ``Rollout``.  Member k's results equal a ``Rollout`` of that member alone
    bit for bit (the batched step and adjoint equal the single-patch ones).

    gauges_xy: one (ng, 2) array of gauge positions per member (each member's
    patch has xlower = ylower = 0)."""

    def __init__(self, members: List, nsteps: int, gauges_xy: Sequence[np.ndarray],
                 parameters: Sequence[str] = ("phi", "kref"), checkpoint_every: int = 100,
                 dt0: float = 1e-3, cfl_desired: float = 0.9, dt_max: float = 1e9,
                 bc=(BC_EXTRAP,) * 4, seg_rows: int = 16, sub_every: int = 0,
                 host_checkpoints=False, hbm_reserve: Optional[int] = None):
        from .step import Ensemble
        if nsteps < 1 or checkpoint_every < 1:
            raise ValueError("nsteps and checkpoint_every must be >= 1")
        if len(gauges_xy) != len(members):
            raise ValueError("one gauge array per member")
        self.members = members
        self.nsteps = nsteps
        self.every = checkpoint_every
        self.sub_every = sub_every
        # block checkpoints: on the device (False), in host memory (True,
        # copied back when a block is reversed: HBM then holds only the
        # sub-checkpoints and one sub-block's states besides the step's own
        # buffers and workspaces), or "auto": on the device while it keeps
        # hbm_reserve bytes free after the copy (default: what the reverse
        # sweep allocates -- the sub-checkpoints of a block, one sub-block's
        # states, two adjoint states and two step temporaries -- plus 32 GB),
        # in host memory after that
        if host_checkpoints not in (False, True, "auto"):
            raise ValueError("host_checkpoints must be False, True or 'auto'")
        self.host_checkpoints = host_checkpoints
        self.hbm_reserve = hbm_reserve
        self.dt0, self.cfl_desired, self.dt_max = dt0, cfl_desired, dt_max
        self.ens = Ensemble(members, bc=bc, seg_rows=seg_rows)
        self.adj = EnsembleAdjoint(self.ens, parameters)
        dev = self.ens.dt.device
        self.xy = [torch.from_numpy(np.ascontiguousarray(
            np.asarray(g, dtype=np.float64).reshape(-1, 2))).to(dev) for g in gauges_xy]
        self.drytol = [m.patch.params.dry_tolerance for m in members]

    def _states(self) -> List[torch.Tensor]:
        return [self.ens.current(k) for k in range(self.ens.n)]

    def _checkpoint(self) -> List[torch.Tensor]:
        states = self._states()
        to_host = self.host_checkpoints is True
        if self.host_checkpoints == "auto":
            nbytes = sum(x.numel() * x.element_size() for x in states)
            reserve = self.hbm_reserve
            if reserve is None:
                s = sub_interval(self.every, self.sub_every)
                reserve = (-(-self.every // s) + s + 4) * nbytes + (32 << 30)
            free, _ = torch.cuda.mem_get_info(self.ens.dt.device)
            to_host = free - nbytes < reserve
        if not to_host:
            return [x.clone() for x in states]
        # pageable host memory: page-locked allocations of this size are not
        # granted everywhere, and a block checkpoint is taken only every
        # checkpoint_every steps
        return [x.to("cpu") for x in states]

    def _load(self, ck: List[torch.Tensor]) -> List[torch.Tensor]:
        dev = self.ens.dt.device
        return [h if h.device == dev else h.to(dev) for h in ck]

    def _gauges(self, series: List[torch.Tensor], n: int) -> None:
        e = self.ens
        for k in range(e.n):
            gauges(e.current(k), e.bed[k], e.maux[k], e.mx[k], e.my[k], 0.0, 0.0, e.dx[k],
                   e.dy[k], self.drytol[k], self.xy[k], series[k][n])

    def forward(self):
        """Runs the ensemble from its members' initial states; returns the
        gauge records (one (nsteps, ng_k, 8) tensor per member) and the final
        states."""
        e = self.ens
        e.dt.fill_(float(self.dt0))
        e.cfl.zero_()
        dev = e.dt.device
        self.dts = torch.zeros((self.nsteps, e.n), dtype=torch.float64, device=dev)
        self.ckpt = {0: self._checkpoint()}
        series = [torch.zeros((self.nsteps, xy.shape[0], 8), dtype=torch.float64, device=dev)
                  for xy in self.xy]
        for n in range(self.nsteps):
            self.dts[n].copy_(e.dt)
            e.step()
            dt_update(e.dt, e.cfl, self.cfl_desired, self.dt_max)
            self._gauges(series, n)
            if (n + 1) % self.every == 0 and n + 1 < self.nsteps:
                self.ckpt[n + 1] = self._checkpoint()
        return series, [x.clone() for x in self._states()]

    def backward(self, g_series: Optional[Sequence[torch.Tensor]] = None,
                 g_final: Optional[Sequence[torch.Tensor]] = None):
        """Adjoint of a loss given its gradients with respect to the gauge
        records and the final states: returns (lambda_0 per member, the (n, np)
        parameter gradients)."""
        e = self.ens
        n_m = e.n
        lam = [g.detach().clone().contiguous() if g is not None
               else torch.zeros_like(x, device=e.dt.device)
               for g, x in zip(g_final or [None] * n_m, self.ckpt[0])]
        lam_b = [torch.empty_like(x) for x in lam]
        grad = torch.zeros((n_m, self.adj.np), dtype=torch.float64, device=e.dt.device)
        lams = [lam, lam_b]

        last = [None]

        def advance(x, n):
            # one batched step from the states x (the ensemble's own buffers
            # serve as scratch); returns copies of q_{n+1}.  When x is the
            # previous call's result the buffers still hold it: no copy, and
            # the batched step goes on skipping the tiles that stayed dry
            if x is not last[0]:
                for k in range(n_m):
                    e.current(k).copy_(x[k])
                e.state_written()
            e.dt.copy_(self.dts[n])
            e.cfl.zero_()
            e.step()
            last[0] = [y.clone() for y in self._states()]
            return last[0]

        def reverse_step(n, qn, qn1):
            lo, li = lams
            if g_series is not None:
                for k in range(n_m):
                    gauges_adj(qn1[k], e.bed[k], e.maux[k], e.mx[k], e.my[k], 0.0, 0.0,
                               e.dx[k], e.dy[k], self.drytol[k], self.xy[k],
                               g_series[k][n].detach().contiguous(), lo[k])
            e.dt.copy_(self.dts[n])
            grad.add_(self.adj.step(qn, lo, li))
            lams.reverse()

        reverse_blocks(self.nsteps, self.every, self.sub_every, self.ckpt, advance, reverse_step,
                       load=self._load)
        return lams[0], grad


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, theta, spec):
        r = Rollout(spec)
        params_dev = r.record(theta.detach().cpu().tolist() if theta is not None else None)
        series, qf = r.forward(q0, params_dev)
        ctx.r = r
        ctx.theta_device = theta.device if theta is not None else None
        return series, qf

    @staticmethod
    def backward(ctx, g_series, g_final):
        lam0, grad = ctx.r.backward(g_series, g_final)
        ctx.r = None
        gtheta = grad.to(ctx.theta_device) if ctx.theta_device is not None else None
        return lam0, gtheta, None


def rollout(q0: torch.Tensor, theta: Optional[torch.Tensor], spec: RolloutSpec):
    """Differentiable N-step run: (gauge records (nsteps, ngauges, 8), final
    state) as functions of q0 (device tensor in the layout of
    ``step.to_device_q``) and theta (float64 tensor, one value per name in
    ``spec.parameters``, in the record's units: phi in radians, kref in m^2).
    The backward pass is the checkpointed adjoint."""
    if theta is not None and theta.numel() != len(spec.parameters):
        raise ValueError("theta needs one value per named parameter")
    return _RolloutFn.apply(q0, theta, spec)


def gauge_loss_gradient_adjoint(q: np.ndarray, aux: np.ndarray, params: DigParams, dx: float,
                                dy: float, xlower: float, ylower: float, gauges_xy,
                                targets: np.ndarray, parameters: Sequence[str] = ("phi", "kref"),
                                dt0: float = 1e-3, cfl_desired: float = 0.9, dt_max: float = 1e9,
                                component: int = 0, bc=(BC_EXTRAP,) * 4,
                                checkpoint_every: int = 100, sub_every: int = 0):
    """The gauge misfit of ``tangent.gauge_loss_gradient`` (sum over steps and
    gauges of (gauge value[component] - target)^2) and its gradient by one
    reverse sweep: returns (loss, {parameter: dloss/dparameter}, dloss/dq0 as
    a host array in q's layout (ghost entries 0), dts)."""
    targets = np.asarray(targets, dtype=np.float64)
    nsteps, ng = targets.shape
    spec = RolloutSpec(aux=aux, params=params, dx=dx, dy=dy, nsteps=nsteps, dt0=dt0,
                       cfl_desired=cfl_desired, dt_max=dt_max, bc=tuple(bc),
                       gauges_xy=np.asarray(gauges_xy, dtype=np.float64).reshape(-1, 2),
                       xlower=xlower, ylower=ylower, parameters=tuple(parameters),
                       checkpoint_every=checkpoint_every, sub_every=sub_every)
    if spec.gauges_xy.shape[0] != ng:
        raise ValueError("targets must have one column per gauge")
    r = Rollout(spec)
    series, _ = r.forward(to_device_q(q), r.record())
    tgt = torch.tensor(targets, device=series.device)
    res = series[:, :, component] - tgt
    loss = float((res * res).sum().item())
    g = torch.zeros_like(series)
    g[:, :, component] = 2.0 * res
    lam0, grad = r.backward(g, None)
    torch.cuda.synchronize()
    grads = {n: float(v) for n, v in zip(parameters, grad.cpu().tolist())}
    return loss, grads, to_host_q(lam0), r.dts.cpu().numpy()
