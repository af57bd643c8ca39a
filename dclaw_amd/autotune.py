"""Tile-shape autotuner of the fused step.

The only free tile parameter is the segment height (rows per wavefront
tile; the strip width is fixed at 58 interior columns by the 64-lane
wavefront).  Short segments shorten the critical path of wet regions and
spread wet work over more waves; long segments amortise the 6 halo rows each
segment recomputes.  Results are bit-identical for every height, so the
tuner only times: each candidate runs a few real steps of the caller's
workload on copies of its state, timed with HIP events on the current
stream, and the fastest height is kept.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Dict, Iterable, Tuple

import torch

from . import _lib

# 1 and 2 pay on narrow patches (few strips): a segment's march is its rows
# plus the halo rows, one after another
CANDIDATES = (1, 2, 4, 6, 8, 12, 16, 20, 24, 28, 32, 48)


def _time(fn: Callable[[], None], reps: int) -> float:
    fn()                                # warm (code objects, caches)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_patch(runner, candidates: Iterable[int] = CANDIDATES, reps: int = 5,
               split_modes: Iterable[str] = ("single", "ordered", "dry", "hybrid")
               ) -> Tuple[int, Dict]:
    """Pick the segment height for a PatchRunner's single-patch launches and,
    for the split step, its mode (one strip launch in tile order, with the
    previous step's wet tiles first, after a separate dry pass, or with the
    wet segments single and first and the dry ones in tall groups -- which
    wins depends on the wet fraction).  Every candidate starts from the same
    state (restored outside the timed region) and times `reps` consecutive
    steps at a fixed dt; the runner's state and dt are restored afterwards.
    Returns the height and the times in ms keyed by (height, mode)."""
    L = _lib.lib()
    handle = getattr(runner, "handle", None)

    def use(h, mode):
        # the runner's own handle if it has one, else the library-wide setting
        runner.split_mode = mode
        if handle is not None:
            handle.set_seg_rows(h)
            handle.set_split_mode(mode)
        else:
            _lib.check(min(0, L.dclaw_set_seg_rows(h)))

    qa, qb = runner.qa.clone(), runner.qb.clone()
    dt, cfl = runner.dt.clone(), runner.cfl.clone()
    modes = tuple(split_modes) if runner.split_src2 else (runner.split_mode,)
    times = {}
    for mode in modes:
        for h in candidates:
            use(h, mode)
            runner.qa.copy_(qa)
            runner.state_written()

            def one():
                runner.step()
                runner.cfl.zero_()
            times[(h, mode)] = _time(one, reps)
    best, mode = min(times, key=times.get)
    use(best, mode)
    runner.qa.copy_(qa)
    runner.qb.copy_(qb)
    runner.state_written()
    runner.dt.copy_(dt)
    runner.cfl.copy_(cfl)
    return best, times


def tune_ensemble(members, make, candidates: Iterable[int] = CANDIDATES,
                  reps: int = 3) -> Tuple[int, Dict]:
    """Pick the segment height of the batched launch: `make(seg_rows)` builds
    an Ensemble from `members`; each candidate times a few steps."""
    times = {}
    for h in candidates:
        ens = make(h)
        ens.dt.fill_(1e-3)

        def one():
            ens.step()
            ens.cfl.zero_()
        times[h] = _time(one, reps)
        del ens
    best = min(times, key=times.get)
    return best, times
