"""Host-side driver of the gfx950 step: the calls GeoClaw's time loop makes
for one patch (bound -> stepgrid -> dtnew), on device-resident buffers.

torch tensors are used only as device buffers and for the stream; every
computation happens in the HIP kernels behind the C-ABI (dclaw_amd/_lib.py).

Reference call sequence mirrored here (Clawpack 5 GeoClaw, restated):
``bound`` fills ghost cells (bc2amr), ``stepgrid`` runs b4step2
(b4step2.f90:7), step2 (flux2fw.f:8, rpn2_dclaw.f:24, rpt2_dclaw.f:16), the
conservative update and src2 (src2.f90:27), and returns cflgrid; the next
dt is dt*cfl_desired/cflgrid capped by dt_max (tick.f:484-531).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .params import BC_EXTRAP, CParams, DigParams, to_c

MEQN = 7
MBC = 2


def _ptr(t: torch.Tensor) -> int:
    return int(t.data_ptr())


def _stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def params_tensor(params: Sequence[DigParams], device="cuda") -> torch.Tensor:
    """Device array of dclaw_params records (one per member)."""
    recs = [to_c(p) for p in params]
    raw = b"".join(bytes(r) for r in recs)
    host = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
    return host.to(device)


def to_device_q(q: np.ndarray, device="cuda") -> torch.Tensor:
    """Fortran-ordered (7, mx+4, my+4) numpy array -> contiguous device tensor
    of shape (my+4, mx+4, 7): the same bytes, component fastest."""
    qf = np.asfortranarray(q, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(qf.transpose(2, 1, 0))).to(device)


def to_host_q(t: torch.Tensor) -> np.ndarray:
    a = t.detach().cpu().numpy()
    return np.asfortranarray(a.transpose(2, 1, 0))


def bed_tensor(aux: np.ndarray, device="cuda") -> torch.Tensor:
    """aux(1) (the bed) as a dense (my+4, mx+4) device array (maux = 1)."""
    b = np.asfortranarray(aux)[0]
    return torch.from_numpy(np.ascontiguousarray(b.T)).to(device)


def aux_planes(aux: np.ndarray, params, device="cuda"):
    """(device aux, maux) the step reads: the bed alone (maux = 1), or the
    whole aux array when entrainment is on (aux(7) holds the erodible depth
    h_e, digclaw_module.f90:110)."""
    if getattr(params, "entrainment", 0):
        if aux.shape[0] < 7:
            raise ValueError("entrainment needs maux >= 7 (aux(7) = h_e)")
        return aux_tensor(aux, device), aux.shape[0]
    if getattr(params, "curvature", 0):
        if aux.shape[0] < 3:
            raise ValueError("curvature needs maux >= 3 (aux(3) = theta)")
        return aux_tensor(aux, device), aux.shape[0]
    return bed_tensor(aux, device), 1


def bc_fill(q: torch.Tensor, mx: int, my: int, bc=(BC_EXTRAP,) * 4) -> None:
    """Ghost-cell fill in place (bc2amr extrap / wall)."""
    arr = (ctypes.c_int * 4)(*bc)
    _lib.check(_lib.lib().dclaw_bc_fill(_ptr(q), mx, my, arr, _stream()))


def stepgrid(q_in: torch.Tensor, q_out: torch.Tensor, aux: torch.Tensor, maux: int, mx: int,
             my: int, dx: float, dy: float, dt_dev: torch.Tensor, params_dev: torch.Tensor,
             cfl_dev: torch.Tensor) -> None:
    """One fused stepgrid on device buffers; cfl_dev receives max(cfl_dev, cflgrid)."""
    _lib.check(_lib.lib().dclaw_stepgrid_dev(
        _ptr(q_in), _ptr(q_out), _ptr(aux), maux, mx, my, dx, dy, _ptr(dt_dev),
        _ptr(params_dev), _ptr(cfl_dev), _stream()))


def stepgrid_bc(q_in: torch.Tensor, q_out: torch.Tensor, aux: torch.Tensor, maux: int, mx: int,
                my: int, dx: float, dy: float, bc, dt_dev: torch.Tensor,
                params_dev: torch.Tensor, cfl_dev: torch.Tensor) -> None:
    """stepgrid with the boundary conditions applied on load (no bound pass)."""
    arr = (ctypes.c_int * 4)(*bc)
    _lib.check(_lib.lib().dclaw_stepgrid_bc_dev(
        _ptr(q_in), _ptr(q_out), _ptr(aux), maux, mx, my, dx, dy, arr, _ptr(dt_dev),
        _ptr(params_dev), _ptr(cfl_dev), _stream()))


def aux_tensor(aux: np.ndarray, device="cuda") -> torch.Tensor:
    """The whole aux(maux, mx+4, my+4) array as a device tensor of shape
    (my+4, mx+4, maux) (the same bytes, component fastest)."""
    af = np.asfortranarray(aux, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(af.transpose(2, 1, 0))).to(device)


def to_host_aux(t: torch.Tensor) -> np.ndarray:
    return np.asfortranarray(t.detach().cpu().numpy().transpose(2, 1, 0))


NO_NAN = 0x7FFFFFFF


def nan_flag(device="cuda") -> torch.Tensor:
    """A device int32 for check4nans results, initialised to 'none'."""
    return torch.full((1,), NO_NAN, dtype=torch.int32, device=device)


def check4nans(q: torch.Tensor, mx: int, my: int, flag: torch.Tensor) -> None:
    _lib.check(_lib.lib().dclaw_check4nans_dev(_ptr(q), mx, my, _ptr(flag), _stream()))


def b4step2(q: torch.Tensor, aux: torch.Tensor, mx: int, my: int, dx: float, dy: float,
            params_dev: torch.Tensor, flag: Optional[torch.Tensor] = None) -> None:
    """b4step2 in place on q and aux(4..6) (aux: full (my+4, mx+4, maux) tensor)."""
    _lib.check(_lib.lib().dclaw_b4step2_dev(
        _ptr(q), _ptr(aux), aux.shape[2], mx, my, dx, dy, _ptr(params_dev),
        None if flag is None else _ptr(flag), _stream()))


def step2(q: torch.Tensor, aux: torch.Tensor, mx: int, my: int, dx: float, dy: float,
          dt_dev: torch.Tensor, params_dev: torch.Tensor, cfl_dev: torch.Tensor):
    """GeoClaw step2 on device buffers: returns (fm, fp, gm, gp) tensors shaped
    like q; cfl_dev receives max(cfl_dev, cflgrid)."""
    outs = [torch.empty_like(q) for _ in range(4)]
    _lib.check(_lib.lib().dclaw_step2_dev(
        _ptr(q), _ptr(aux), aux.shape[2], mx, my, dx, dy, _ptr(dt_dev), _ptr(params_dev),
        *[_ptr(o) for o in outs], _ptr(cfl_dev), _stream()))
    return outs


def src2(q: torch.Tensor, aux: torch.Tensor, mx: int, my: int, dt_dev: torch.Tensor,
         params_dev: torch.Tensor, flag: Optional[torch.Tensor] = None) -> None:
    """src2 in place on the interior cells of q."""
    _lib.check(_lib.lib().dclaw_src2_dev(
        _ptr(q), _ptr(aux), aux.shape[2], mx, my, _ptr(dt_dev), _ptr(params_dev),
        None if flag is None else _ptr(flag), _stream()))


SPLIT_MODES = ("single", "ordered", "dry", "hybrid")


def stepgrid_split(q_in: torch.Tensor, q_out: torch.Tensor, aux: torch.Tensor, maux: int, mx: int,
                   my: int, dx: float, dy: float, bc, dt_dev: torch.Tensor,
                   params_dev: torch.Tensor, cfl_dev: torch.Tensor, flags: torch.Tensor,
                   mode: str = "ordered") -> None:
    """The step with src2 in a second pass over the wet tiles (bc None: stored
    ghosts); flags is an int32 workspace (kept between steps for "ordered").
    mode "single": one strip launch in tile order (dclaw_tile_count ints
    suffice); "ordered": one strip launch with the tiles the previous step
    flagged first; "dry": a light dry pass writes the dry tiles and the strip
    kernel runs over the others; "hybrid": the previous step's wet segments
    singly and first, the others in tall groups of 4 segments (these three
    need dclaw_split_workspace_size ints).  All give the same result."""
    if mode not in SPLIT_MODES:
        raise ValueError(f"mode must be one of {SPLIT_MODES}")
    arr = None if bc is None else (ctypes.c_int * 4)(*bc)
    L = _lib.lib()
    n = flags.numel()
    if mode == "single":
        tc = L.dclaw_tile_count(mx, my)
        _lib.check(min(tc, 0))
        n = min(n, tc)
    else:
        _lib.check(min(0, L.dclaw_set_split_mode({"ordered": 1, "dry": 2, "hybrid": 3}[mode])))
    _lib.check(L.dclaw_stepgrid_split_dev(
        _ptr(q_in), _ptr(q_out), _ptr(aux), maux, mx, my, dx, dy, arr, _ptr(dt_dev),
        _ptr(params_dev), _ptr(cfl_dev), _ptr(flags), n, _stream()))


class Handle:
    """A library handle (dclaw_create): its own segment height, split mode and
    dry-pass rows instead of the library-wide ones, and the resident split
    step's tile workspace, so time loops with different settings run side by
    side.  One handle serves one time loop (one thread, one stream)."""

    def __init__(self, seg_rows: Optional[int] = None, split_mode: Optional[str] = None,
                 dry_rows: Optional[int] = None):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.dclaw_create(ctypes.byref(h)))
        self._h = h
        if seg_rows is not None:
            self.set_seg_rows(seg_rows)
        if split_mode is not None:
            self.set_split_mode(split_mode)
        if dry_rows is not None:
            _lib.check(min(0, L.dclaw_handle_set_dry_rows(self._h, int(dry_rows))))

    def set_seg_rows(self, rows: int) -> int:
        """Segment height of this handle's launches; returns the previous one."""
        prev = _lib.lib().dclaw_handle_set_seg_rows(self._h, int(rows))
        _lib.check(min(0, prev))
        return prev

    def set_split_mode(self, mode: str) -> None:
        """"single" (tile order), "ordered" (wet-first, dry-tile skipping),
        "dry" (a dry pass first) or "hybrid" (as stepgrid_split)."""
        if mode not in SPLIT_MODES:
            raise ValueError(f"mode must be one of {SPLIT_MODES}")
        _lib.check(min(0, _lib.lib().dclaw_handle_set_split_mode(self._h,
                                                                  SPLIT_MODES.index(mode))))

    def tile_count(self, mx: int, my: int) -> int:
        n = _lib.lib().dclaw_handle_tile_count(self._h, mx, my)
        _lib.check(min(0, n))
        return n

    def step(self, q_in: torch.Tensor, q_out: torch.Tensor, aux: torch.Tensor, maux: int,
             mx: int, my: int, dx: float, dy: float, bc, dt_dev: torch.Tensor,
             params_dev: torch.Tensor, cfl_dev: torch.Tensor) -> None:
        """The split step with this handle's settings and workspace (bc None:
        stored ghosts)."""
        arr = None if bc is None else (ctypes.c_int * 4)(*bc)
        _lib.check(_lib.lib().dclaw_handle_stepgrid_dev(
            self._h, _ptr(q_in), _ptr(q_out), _ptr(aux), maux, mx, my, dx, dy, arr,
            _ptr(dt_dev), _ptr(params_dev), _ptr(cfl_dev), _stream()))

    def state_written(self) -> None:
        """The loop's state was written outside its steps."""
        _lib.check(_lib.lib().dclaw_handle_state_written(self._h, _stream()))

    def close(self) -> None:
        h, self._h = self._h, None
        if h is not None and h.value:
            _lib.check(_lib.lib().dclaw_destroy(h))

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def stepgrid_curvature(q_in: torch.Tensor, q_out: torch.Tensor, aux: torch.Tensor, maux: int,
                       mx: int, my: int, dx: float, dy: float, bc, dt_dev: torch.Tensor,
                       params_dev: torch.Tensor, cfl_dev: torch.Tensor,
                       work: torch.Tensor) -> None:
    """The step with src2's curvature term (each cell's gravity includes the
    curvature accelerations of the cells before it in src2's sweep);
    work: float64 tensor of at least curvature_workspace_doubles(mx, my)."""
    arr = None if bc is None else (ctypes.c_int * 4)(*bc)
    _lib.check(_lib.lib().dclaw_stepgrid_curvature_dev(
        _ptr(q_in), _ptr(q_out), _ptr(aux), maux, mx, my, dx, dy, arr, _ptr(dt_dev),
        _ptr(params_dev), _ptr(cfl_dev), _ptr(work), work.numel(), _stream()))


def curvature_workspace_doubles(mx: int, my: int) -> int:
    return int(_lib.lib().dclaw_curvature_workspace_doubles(mx, my))


def dt_update(dt_dev: torch.Tensor, cfl_dev: torch.Tensor, cfl_desired: float,
              dt_max: float) -> None:
    _lib.check(_lib.lib().dclaw_dt_update(_ptr(dt_dev), _ptr(cfl_dev), dt_dev.numel(),
                                          cfl_desired, dt_max, _stream()))


def dt_update_out(dt_dev: torch.Tensor, cfl_dev: torch.Tensor, t_dev: torch.Tensor,
                  tout_dev: torch.Tensor, hit_dev: torch.Tensor, cfl_desired: float,
                  dt_max: float) -> None:
    """The dt rule with the output-time hit (tick.f:135-144); t advances by the
    dt just used."""
    _lib.check(_lib.lib().dclaw_dt_update_out(_ptr(dt_dev), _ptr(cfl_dev), _ptr(t_dev),
                                              _ptr(tout_dev), _ptr(hit_dev), dt_dev.numel(),
                                              cfl_desired, dt_max, _stream()))


def gauges(q: torch.Tensor, aux: torch.Tensor, maux: int, mx: int, my: int, xlower: float,
           ylower: float, dx: float, dy: float, dry_tolerance: float, xy_dev: torch.Tensor,
           out: torch.Tensor) -> None:
    """Gauge values (ngauges, 8) into `out` (a device view)."""
    _lib.check(_lib.lib().dclaw_gauges_dev(_ptr(q), _ptr(aux), maux, mx, my, xlower, ylower, dx,
                                           dy, dry_tolerance, _ptr(xy_dev), xy_dev.shape[0],
                                           _ptr(out), _stream()))


class RunResult:
    """What PatchRunner.run_until returns: final time, step count, the output
    times reached, the time of every step and the gauge records (nsteps,
    ngauges, 8: h hu hv hm pb hchi bdif eta after each step)."""

    def __init__(self, t, nsteps, outputs, times, gauge_values):
        self.t = t
        self.nsteps = nsteps
        self.outputs = outputs
        self.times = times
        self.gauge_values = gauge_values


class PatchRunner:
    """Device-resident fixed-patch time loop (one member)."""

    def __init__(self, q: np.ndarray, aux: np.ndarray, params: DigParams, dx: float, dy: float,
                 bc=(BC_EXTRAP,) * 4, device="cuda", split_src2: bool = True,
                 split_mode: str = "ordered", handle: Optional[Handle] = None):
        """handle: step through that library handle (its own settings and
        workspace; split_mode is then the handle's) instead of the library-wide
        settings and this runner's workspace."""
        self.split_src2 = split_src2
        self.handle = handle
        if split_mode not in SPLIT_MODES:
            raise ValueError(f"split_mode must be one of {SPLIT_MODES}")
        self.split_mode = split_mode
        self.mx = q.shape[1] - 2 * MBC
        self.my = q.shape[2] - 2 * MBC
        self.dx, self.dy = float(dx), float(dy)
        self.bc = tuple(bc)
        self.qa = to_device_q(q, device)
        self.qb = torch.empty_like(self.qa)
        self.bed, self.maux = aux_planes(aux, params, device)
        self.bed_host = np.array(np.asfortranarray(aux)[0:1], order="F", copy=True)
        self.params = params_tensor([params], device)
        self.params_host = params
        self.dt = torch.zeros(1, dtype=torch.float64, device=device)
        self.cfl = torch.zeros(1, dtype=torch.float64, device=device)
        # split-step workspace for any segment height (one row per tile at
        # most): tile flags, the wet list or tile order, counters, and the
        # dry-tile ages of the resident wet-first step (this loop swaps qa and
        # qb every step and writes neither in between)
        nstrips = -(-self.mx // _lib.lib().dclaw_strip_width())
        self.flags = torch.zeros(4 * nstrips * self.my + 12, dtype=torch.int32, device=device)
        # curvature: the step with src2's running gravity (its own workspace)
        self.curv_work = (torch.zeros(curvature_workspace_doubles(self.mx, self.my),
                                      dtype=torch.float64, device=device)
                          if getattr(params, "curvature", 0) else None)

    def set_dt(self, dt: float) -> None:
        self.dt.fill_(float(dt))

    def state_written(self) -> None:
        """Call after writing qa or qb directly: the split step's dry-tile
        skipping (which assumes the state changes only through step) starts
        over."""
        self.flags.zero_()
        if self.handle is not None:
            self.handle.state_written()

    def step(self, fused_bc: bool = True) -> None:
        """bound + stepgrid (boundary conditions applied on load when fused_bc);
        the result is left in self.qa and cflgrid in self.cfl."""
        if self.curv_work is not None:
            if not fused_bc:
                bc_fill(self.qa, self.mx, self.my, self.bc)
            stepgrid_curvature(self.qa, self.qb, self.bed, self.maux, self.mx, self.my, self.dx,
                               self.dy, self.bc if fused_bc else None, self.dt, self.params,
                               self.cfl, self.curv_work)
        elif fused_bc and self.split_src2 and self.handle is not None:
            self.handle.step(self.qa, self.qb, self.bed, self.maux, self.mx, self.my, self.dx,
                             self.dy, self.bc, self.dt, self.params, self.cfl)
        elif fused_bc and self.split_src2:
            stepgrid_split(self.qa, self.qb, self.bed, self.maux, self.mx, self.my, self.dx, self.dy,
                           self.bc, self.dt, self.params, self.cfl, self.flags, self.split_mode)
        elif fused_bc:
            stepgrid_bc(self.qa, self.qb, self.bed, self.maux, self.mx, self.my, self.dx, self.dy,
                        self.bc, self.dt, self.params, self.cfl)
        else:
            bc_fill(self.qa, self.mx, self.my, self.bc)
            stepgrid(self.qa, self.qb, self.bed, self.maux, self.mx, self.my, self.dx, self.dy,
                     self.dt,
                     self.params, self.cfl)
        self.qa, self.qb = self.qb, self.qa

    def run(self, nsteps: int, dt0: float, cfl_desired: float, dt_max: float,
            cfl_hist: Optional[torch.Tensor] = None, fused_bc: bool = True) -> None:
        self.set_dt(dt0)
        self.cfl.zero_()
        for n in range(nsteps):
            self.step(fused_bc)
            if cfl_hist is not None:
                cfl_hist[n].copy_(self.cfl[0])
            dt_update(self.dt, self.cfl, cfl_desired, dt_max)

    def run_until(self, t_outputs: Sequence[float], dt0: float, cfl_desired: float,
                  dt_max: float, gauges_xy: Optional[Sequence] = None, xlower: float = 0.0,
                  ylower: float = 0.0, check_nans: bool = False, on_output=None,
                  max_steps: int = 10 ** 7) -> RunResult:
        """Device-resident loop to the output times (the last one is the end
        time), as AMRClaw's tick does: bound + stepgrid + the dt rule with
        dt_max and the output-time hit; after every step the gauges are
        sampled on the device; at each output time on_output(self, t) is
        called.  check_nans runs check4nans on the state before every step and
        raises DClawError (with the first NaN's q(m,i,j)) at the next output."""
        outs = sorted(float(x) for x in t_outputs)
        if not outs or outs[0] <= 0.0:
            raise ValueError("output times must be > 0")
        dev = self.qa.device
        t = torch.zeros(1, dtype=torch.float64, device=dev)
        tout = torch.tensor([outs[0]], dtype=torch.float64, device=dev)
        hit = torch.zeros(1, dtype=torch.int32, device=dev)
        flag = nan_flag(dev) if check_nans else None
        ng = 0
        xy_dev = None
        if gauges_xy is not None:
            xy = np.asarray(gauges_xy, dtype=np.float64).reshape(-1, 2)
            for gx, gy in xy:
                # the four cells around the gauge must be interior:
                # x in [xlower + dx/2, xlower + (mx - 1/2) dx), same in y
                if not (xlower + 0.5 * self.dx <= gx < xlower + (self.mx - 0.5) * self.dx and
                        ylower + 0.5 * self.dy <= gy < ylower + (self.my - 0.5) * self.dy):
                    raise ValueError(f"gauge ({gx}, {gy}) is not half a cell inside the patch")
            ng = xy.shape[0]
            xy_dev = torch.from_numpy(np.ascontiguousarray(xy)).to(dev)
        cap = 1024
        gbuf = torch.zeros((cap, max(ng, 1), 8), dtype=torch.float64, device=dev)
        tbuf = torch.zeros(cap, dtype=torch.float64, device=dev)
        g_host, t_host = [], []
        self.cfl.zero_()
        # tick.f:135-144 before the first step: shorten dt0 to land on outs[0]
        target, landing = 0, None
        if 1.001 * dt0 >= outs[0]:
            dt0, landing, target = outs[0], 0, 1
        self.set_dt(dt0)
        tout.fill_(outs[target] if target < len(outs) else float("inf"))
        reached, n = [], 0
        while n < max_steps:
            if flag is not None:
                check4nans(self.qa, self.mx, self.my, flag)
            self.step()
            dt_update_out(self.dt, self.cfl, t, tout, hit, cfl_desired, dt_max)
            slot = n % cap
            tbuf[slot].copy_(t[0])
            if ng:
                gauges(self.qa, self.bed, self.maux, self.mx, self.my, xlower, ylower, self.dx,
                       self.dy,
                       self.params_host.dry_tolerance, xy_dev, gbuf[slot])
            n += 1
            if slot == cap - 1:
                t_host.append(tbuf.cpu().numpy().copy())
                g_host.append(gbuf.cpu().numpy().copy())
            if landing is not None:
                # the step just taken was shortened to land on outs[landing]
                if flag is not None and int(flag.item()) != NO_NAN:
                    raise _lib.DClawError(self._nan_message(int(flag.item())))
                reached.append(float(t.item()))
                if on_output is not None:
                    on_output(self, outs[landing])
                if landing == len(outs) - 1:
                    break
                landing = None
            if bool(hit.item()):
                # the next step lands on outs[target]; later checks use the one after
                landing, target = target, target + 1
                tout.fill_(outs[target] if target < len(outs) else float("inf"))
        rem = n % cap
        if rem:
            t_host.append(tbuf[:rem].cpu().numpy().copy())
            g_host.append(gbuf[:rem].cpu().numpy().copy())
        times = np.concatenate(t_host) if t_host else np.zeros(0)
        gv = np.concatenate(g_host)[:, :ng] if g_host else np.zeros((0, ng, 8))
        return RunResult(float(t.item()), n, reached, times, gv)

    def write_frame(self, outdir: str, frame: int, time: float,
                    output_format: str = "binary64", xlower: float = 0.0,
                    ylower: float = 0.0):
        """valout's frame of the current state (fort.q/fort.t, fort.b for the
        binary formats); ghost cells are filled on the device first, as valout
        calls bound() (valout.f90:254-257)."""
        from . import io as dio
        q = self.qa.clone()
        bc_fill(q, self.mx, self.my, self.bc)
        grid = dio.Grid(q=to_host_q(q), aux=self.bed_host, xlow=xlower, ylow=ylower, dx=self.dx,
                        dy=self.dy)
        return dio.write_frame(outdir, frame, time, [grid], output_format)

    def _nan_message(self, first: int) -> str:
        m, cell = first % 7, first // 7
        i, j = cell % (self.mx + 4) - 1, cell // (self.mx + 4) - 1
        return (f"SOLUTION ERROR --- ABORTING CALCULATION: NaN in q(m,i,j) with m,i,j = "
                f"{m + 1} {i} {j} (check4nans)")

    def q(self) -> np.ndarray:
        return to_host_q(self.qa)


class Ensemble:
    """Many members of uneven patch sizes advanced by one launch per step over
    a (patch, strip, segment) work list, each with its own parameters and dt."""

    def __init__(self, members: List, bc=(BC_EXTRAP,) * 4, seg_rows: int = 32, device="cuda",
                 split_src2: bool = True, ordered: bool = True):
        self.split_src2 = split_src2
        self.ordered = ordered
        self.n = len(members)
        if any(getattr(m.patch.params, "curvature", 0) for m in members):
            raise ValueError("the batched step has no curvature path (src2's running gravity "
                             "is a sweep over each patch); step such members with PatchRunner")
        self.seg_rows = seg_rows
        self.bc = tuple(bc)
        self.mx = [m.patch.mx for m in members]
        self.my = [m.patch.my for m in members]
        self.qa = [to_device_q(m.patch.q, device) for m in members]
        self.qb = [torch.empty_like(t) for t in self.qa]
        planes = [aux_planes(m.patch.aux, m.patch.params, device) for m in members]
        self.bed = [pl[0] for pl in planes]
        self.maux = [pl[1] for pl in planes]
        self.params = params_tensor([m.patch.params for m in members], device)
        self._prec = ctypes.sizeof(CParams)
        self.dt = torch.zeros(self.n, dtype=torch.float64, device=device)
        self.cfl = torch.zeros(self.n, dtype=torch.float64, device=device)
        L = _lib.lib()
        mxa = (ctypes.c_int * self.n)(*self.mx)
        mya = (ctypes.c_int * self.n)(*self.my)
        nt = L.dclaw_worklist_size(mxa, mya, self.n, seg_rows)
        if nt < 0:
            _lib.check(nt)
        tiles = (ctypes.c_int32 * (4 * nt))()
        _lib.check(min(0, L.dclaw_worklist_build(mxa, mya, self.n, seg_rows, tiles)))
        self.ntiles = nt
        self.tiles = torch.tensor(list(tiles), dtype=torch.int32, device=device)
        # src2 flags, plus (ordered) the tile order built from the previous
        # step's flags, its counters and the dry-tile bookkeeping of the
        # resident step (this class swaps each member's buffers every step
        # and writes neither in between; state_written() otherwise)
        nws = L.dclaw_batched_resident_size(nt, self.n) if ordered else nt
        _lib.check(min(0, nws))
        self.flags = torch.zeros(nws, dtype=torch.int32, device=device)
        self.dx = [m.patch.dx for m in members]
        self.dy = [m.patch.dy for m in members]
        self._patches = [self._patch_table(False, device), self._patch_table(True, device)]
        self.flip = False
        self._plain_flags = None
        self._plain_dirty = False

    def _patch_table(self, flipped: bool, device) -> torch.Tensor:
        recs = []
        for k in range(self.n):
            qi, qo = (self.qb[k], self.qa[k]) if flipped else (self.qa[k], self.qb[k])
            recs.append(_lib.CPatch(_ptr(qi), _ptr(qo), _ptr(self.bed[k]), self.mx[k],
                                    self.my[k], self.maux[k], k, self.dx[k], self.dy[k]))
        raw = b"".join(bytes(r) for r in recs)
        return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)

    def current(self, k: int) -> torch.Tensor:
        return self.qb[k] if self.flip else self.qa[k]

    def state_written(self) -> None:
        """Call after writing a member's state directly: the batched step's
        dry-tile skipping (which assumes the states change only through step)
        starts over (the four ints after the order's counters say whether
        the bookkeeping describes the last call; the previous flags still
        give the next step its wet-first order)."""
        if self.flags.numel() > 2 * self.ntiles + 4:
            self.flags[2 * self.ntiles + 4:2 * self.ntiles + 8].zero_()
        self._plain_dirty = True

    def step(self) -> None:
        """One launch for all members; boundary conditions applied on load."""
        L = _lib.lib()
        bc = (ctypes.c_int * 4)(*self.bc)
        _lib.check(L.dclaw_step_batched_ws(
            _ptr(self._patches[1 if self.flip else 0]), self.n, _ptr(self.tiles), self.ntiles,
            self.seg_rows, bc, _ptr(self.dt), _ptr(self.params), _ptr(self.cfl),
            _ptr(self.flags) if self.split_src2 else None, self.flags.numel(), _stream()))
        self.flip = not self.flip
        # the per-member loop's dry-tile ages no longer describe the state
        self._plain_dirty = True

    def step_plain(self) -> None:
        """The plain per-member loop, for comparison: one single-patch step per
        member, the reference's one stepgrid call per patch (AMRClaw advanc).
        Each call is the step PatchRunner takes: boundary conditions on load,
        src2 in a second pass over the wet tiles, the previous step's wet
        tiles dispatched first (split mode "ordered", each member with its own
        tile-flag workspace), at the library's single-patch segment height
        (dclaw_set_seg_rows; the caller tunes it for this loop)."""
        if self._plain_flags is None:
            w = _lib.lib().dclaw_strip_width()
            self._plain_flags = [
                torch.zeros(4 * (-(-self.mx[k] // w)) * self.my[k] + 12, dtype=torch.int32,
                            device=self.dt.device) for k in range(self.n)]
        elif self._plain_dirty:
            for f in self._plain_flags:
                f.zero_()
        self._plain_dirty = False
        for k in range(self.n):
            qi = self.current(k)
            qo = self.qa[k] if self.flip else self.qb[k]
            stepgrid_split(qi, qo, self.bed[k], self.maux[k], self.mx[k], self.my[k], self.dx[k],
                           self.dy[k], self.bc, self.dt[k:k + 1],
                           self.params[k * self._prec:(k + 1) * self._prec], self.cfl[k:k + 1],
                           self._plain_flags[k], "ordered")
        self.flip = not self.flip

    def run(self, nsteps: int, dt0, cfl_desired: float, dt_max: float,
            plain: bool = False) -> None:
        if np.isscalar(dt0):
            self.dt.fill_(float(dt0))
        else:
            self.dt.copy_(torch.as_tensor(dt0, dtype=torch.float64))
        self.cfl.zero_()
        for _ in range(nsteps):
            if plain:
                self.step_plain()
            else:
                self.step()
            dt_update(self.dt, self.cfl, cfl_desired, dt_max)

    def q(self, k: int) -> np.ndarray:
        return to_host_q(self.current(k))


class CapturedSteps:
    """k steps of a runner (PatchRunner or Ensemble: its step() followed by
    the device dt rule), captured once as a HIP graph; replay() advances the
    runner by k steps with one graph launch.  Everything a step touches is
    preallocated device memory read through pointers (state buffers, dt, cfl,
    tile flags, work lists) and no step records a memset, so a replay is the
    same work as k eager steps, bit for bit.  k must be even: the runners
    alternate between two state buffers, and an even count leaves them where
    the capture found them."""

    def __init__(self, runner, k: int, cfl_desired: float, dt_max: float):
        if k < 2 or k % 2:
            raise ValueError("k must be an even number >= 2")
        self.runner = runner
        self.k = k
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(k):
                runner.step()
                dt_update(runner.dt, runner.cfl, cfl_desired, dt_max)

    def replay(self) -> None:
        self.graph.replay()
