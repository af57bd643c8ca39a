"""On-disk formats of a D-Claw run: the dclaw.data parameter file and the
valout output frames (fort.qNNNN / fort.bNNNN / fort.tNNNN), so runs of the
GPU step can be set up from and compared with an xdclaw run's files.

* dclaw.data: read in the order of set_dig (digclaw_module.f90:129-166): one
  value per line, first field of the line (list-directed read); clawutil's
  header and comment lines start with '#'.  Written in clawutil's
  "value =: name  description" form (data.py:353-420 writes the same
  sequence).
* Frames: valout.f90:114-300 and :470-490.  fort.q holds one header per grid
  (header_format, valout.f90:74-81) followed, for ASCII output, by the
  interior cells row by row (meqn values and eta = h + aux(1) - bdif per
  line, format 50e26.16, a blank line after each row); for binary output the
  data goes to fort.b as the whole grid including ghost cells, (meqn+1)
  values per cell, component fastest, binary64 or binary32.  fort.t
  (t_file_format, valout.f90:83-89) records the time, meqn+1 (eta is a
  column), the number of grids, naux, ndim = 2, nghost and the format name.
  Values below 1e-90 in magnitude are written as 0 (interior cells), and an
  ASCII eta below 1e-99 as 0, as the reference does.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .params import DigParams

# set_dig's read sequence (digclaw_module.f90:129-166)
DCLAW_DATA_FIELDS = (
    "rho_s", "rho_f", "m_crit", "m0", "mref", "kref", "phi", "delta", "mu", "alpha_c", "c1",
    "sigma_0", "dd_manning", "manning_max", "src2method", "alphamethod", "rarecorrectortest",
    "bed_normal", "theta_input", "entrainment", "entrainment_method", "entrainment_rate", "me",
    "segregation", "beta_seg", "chi0", "chie", "delta_kr_order", "mom_autostop", "momlevel",
    "curvature",
)
_INT_FIELDS = {"src2method", "alphamethod", "bed_normal", "entrainment", "entrainment_method",
               "segregation", "momlevel", "curvature"}
_BOOL_FIELDS = {"dd_manning", "rarecorrectortest", "mom_autostop"}


def _parse_value(tok: str, name: str):
    t = tok.strip()
    if name in _BOOL_FIELDS:              # Fortran logical: T/F, .true./.false.
        u = t.upper().lstrip(".")
        if u.startswith("T"):
            return 1
        if u.startswith("F"):
            return 0
        return int(fortran_float(t))
    if name in _INT_FIELDS:
        return int(fortran_float(t))
    return fortran_float(t)


def read_dclaw_data(path: str) -> Dict[str, float]:
    """The 31 values of dclaw.data by name (phi and theta_input in degrees,
    as written)."""
    vals: List[str] = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            vals.append(s.split("=:")[0].split()[0])
    if len(vals) < len(DCLAW_DATA_FIELDS):
        raise ValueError(f"{path}: {len(vals)} values, set_dig reads {len(DCLAW_DATA_FIELDS)}")
    return {n: _parse_value(v, n) for n, v in zip(DCLAW_DATA_FIELDS, vals)}


def dig_params_from_data(path: str, base: Optional[DigParams] = None) -> Tuple[DigParams, float]:
    """A DigParams with the dclaw.data values (the GeoClaw ones -- grav,
    dry_tolerance, speed_limit, Manning -- and the method switches stay as in
    `base`), and m0 (the initial solid fraction, used by qinit)."""
    d = read_dclaw_data(path)
    p = dataclasses.replace(base) if base is not None else DigParams()
    for name in ("rho_s", "rho_f", "m_crit", "mref", "kref", "delta", "mu", "alpha_c", "c1",
                 "sigma_0", "manning_max", "beta_seg", "delta_kr_order"):
        setattr(p, name, float(d[name]))
    p.phi_deg = float(d["phi"])
    for name in ("src2method", "alphamethod", "rarecorrectortest", "segregation", "dd_manning",
                 "bed_normal", "entrainment", "curvature"):
        setattr(p, name, int(d[name]))
    return p, float(d["m0"])


def write_dclaw_data(path: str, p: DigParams, m0: float, extra: Optional[Dict] = None) -> None:
    """dclaw.data in clawutil's text form (the values of set_dig's sequence;
    `extra` overrides the ones DigParams does not hold)."""
    vals = {
        "rho_s": p.rho_s, "rho_f": p.rho_f, "m_crit": p.m_crit, "m0": m0, "mref": p.mref,
        "kref": p.kref, "phi": p.phi_deg, "delta": p.delta, "mu": p.mu, "alpha_c": p.alpha_c,
        "c1": p.c1, "sigma_0": p.sigma_0, "dd_manning": bool(p.dd_manning),
        "manning_max": p.manning_max, "src2method": p.src2method,
        "alphamethod": p.alphamethod, "rarecorrectortest": bool(p.rarecorrectortest),
        "bed_normal": p.bed_normal, "theta_input": 0.0, "entrainment": p.entrainment,
        "entrainment_method": 0, "entrainment_rate": 0.0, "me": 0.65,
        "segregation": p.segregation, "beta_seg": p.beta_seg, "chi0": 0.5, "chie": 0.5,
        "delta_kr_order": p.delta_kr_order, "mom_autostop": False, "momlevel": 1,
        "curvature": p.curvature,
    }
    vals.update(extra or {})
    with open(path, "w") as f:
        f.write("########################################################\n")
        f.write("### DO NOT EDIT THIS FILE:  GENERATED AUTOMATICALLY ####\n")
        f.write("### To modify data, edit  setrun.py                ####\n")
        f.write("########################################################\n\n")
        for name in DCLAW_DATA_FIELDS:
            v = vals[name]
            if isinstance(v, bool):
                s = "T" if v else "F"
            elif isinstance(v, int):
                s = str(v)
            else:
                s = repr(float(v))
            f.write(f"{s:<26} =: {name}\n")


# ---------------------------------------------------------------- Fortran edit descriptors
def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d output (gfortran): 0.ddd...E+xx with d digits, right-justified
    in w columns; exponents beyond two digits drop the 'E'."""
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("-Infinity" if x < 0 else "Infinity").rjust(w)
    sign = "-" if (x < 0 or (x == 0.0 and math.copysign(1.0, x) < 0)) else ""
    ax = abs(x)
    if ax == 0.0:
        digits, e = "0" * d, 0
    else:
        s = f"{ax:.{d - 1}e}"
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1
    if -99 <= e <= 99:
        es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    return f"{sign}0.{digits}{es}".rjust(w)


def fortran_float(s: str) -> float:
    """A real as Fortran writes it: D or E exponent, or a bare signed
    exponent for three-digit exponents (0.1234-100)."""
    t = s.strip().replace("D", "E").replace("d", "E")
    if "E" not in t.upper():
        k = max(t.rfind("+"), t.rfind("-"))
        if k > 0:
            t = t[:k] + "E" + t[k:]
    return float(t)


def fortran_i(n: int, w: int) -> str:
    return str(int(n)).rjust(w)


def fortran_a(s: str, w: int) -> str:
    return s.rjust(w) if len(s) <= w else s[:w]


# ---------------------------------------------------------------- frames
@dataclasses.dataclass
class Grid:
    """One patch of a frame: q(meqn, mx+2g, my+2g) and aux(naux, ...) in the
    Clawpack layout (Fortran order, ghosts included), lower corner, spacing."""
    q: np.ndarray
    aux: np.ndarray
    xlow: float
    ylow: float
    dx: float
    dy: float
    grid_number: int = 1
    level: int = 1
    nghost: int = 2


def _header(g: Grid, mx: int, my: int) -> str:
    return (f"{fortran_i(g.grid_number, 6)}                 grid_number\n"
            f"{fortran_i(g.level, 6)}                 AMR_level\n"
            f"{fortran_i(mx, 6)}                 mx\n"
            f"{fortran_i(my, 6)}                 my\n"
            f"{fortran_e(g.xlow, 26, 16)}    xlow\n"
            f"{fortran_e(g.ylow, 26, 16)}    ylow\n"
            f"{fortran_e(g.dx, 26, 16)}    dx\n"
            f"{fortran_e(g.dy, 26, 16)}    dy\n")


def _round_interior(q: np.ndarray, ng: int) -> np.ndarray:
    q = np.array(q, order="F", copy=True)
    inner = q[:, ng:q.shape[1] - ng, ng:q.shape[2] - ng]
    inner[np.abs(inner) < 1e-90] = 0.0
    return q


def write_frame(outdir: str, frame: int, time: float, grids: Sequence[Grid],
                output_format: str = "binary64") -> List[str]:
    """fort.qNNNN, fort.tNNNN and (binary formats) fort.bNNNN of one frame.
    For binary output the ghost cells are written as given (the reference
    fills them with bound() first)."""
    if output_format not in ("ascii", "binary32", "binary64"):
        raise ValueError("output_format must be ascii, binary32 or binary64")
    os.makedirs(outdir, exist_ok=True)
    tag = f"{frame:04d}"
    paths = [os.path.join(outdir, f"fort.q{tag}"), os.path.join(outdir, f"fort.t{tag}")]
    naux = grids[0].aux.shape[0] if grids else 0
    meqn = grids[0].q.shape[0] if grids else 7
    bfile = None
    if output_format != "ascii":
        paths.append(os.path.join(outdir, f"fort.b{tag}"))
        bfile = open(paths[2], "wb")
    with open(paths[0], "w") as fq:
        for g in grids:
            ng = g.nghost
            mx, my = g.q.shape[1] - 2 * ng, g.q.shape[2] - 2 * ng
            fq.write(_header(g, mx, my))
            fq.write("\n")
            q = _round_interior(g.q, ng)
            eta = q[0] + g.aux[0] - q[meqn - 1]
            if output_format == "ascii":
                for j in range(ng, my + ng):
                    for i in range(ng, mx + ng):
                        e = eta[i, j]
                        if abs(e) < 1e-99:
                            e = 0.0
                        fq.write("".join(fortran_e(v, 26, 16) for v in q[:, i, j]) +
                                 fortran_e(e, 26, 16) + "\n")
                    fq.write(" \n")
            else:
                qeta = np.empty((meqn + 1,) + q.shape[1:], order="F")
                qeta[:meqn] = q
                qeta[meqn] = eta
                raw = qeta.ravel(order="F")
                if output_format == "binary32":
                    raw = raw.astype(np.float32)
                bfile.write(np.ascontiguousarray(raw).tobytes())
    if bfile is not None:
        bfile.close()
    nghost = grids[0].nghost if grids else 2
    with open(paths[1], "w") as ft:
        ft.write(f"{fortran_e(time, 18, 8)}    time\n"
                 f"{fortran_i(meqn + 1, 6)}                 meqn\n"
                 f"{fortran_i(len(grids), 6)}                 ngrids\n"
                 f"{fortran_i(naux, 6)}                 naux\n"
                 f"{fortran_i(2, 6)}                 ndim\n"
                 f"{fortran_i(nghost, 6)}                 nghost\n"
                 f"{fortran_a(output_format, 10)}             format\n\n")
    return paths


def read_frame(outdir: str, frame: int) -> Tuple[float, List[Dict]]:
    """(time, grids) of a frame written by valout or write_frame: each grid is
    a dict of the header fields and 'qeta' ((meqn+1, mx, my) interior for
    ASCII; (meqn+1, mx+2g, my+2g) with ghosts for binary)."""
    tag = f"{frame:04d}"
    with open(os.path.join(outdir, f"fort.t{tag}")) as ft:
        tl = [ln.split() for ln in ft if ln.strip()]
    time = fortran_float(tl[0][0])
    meqn, ngrids, naux, ndim, nghost = (int(tl[k][0]) for k in range(1, 6))
    fmt = tl[6][0]
    grids = []
    with open(os.path.join(outdir, f"fort.q{tag}")) as fq:
        lines = [ln for ln in fq]
    pos = 0
    bdata = None
    if fmt in ("binary64", "binary32"):
        dtype = np.float64 if fmt == "binary64" else np.float32
        bdata = np.fromfile(os.path.join(outdir, f"fort.b{tag}"), dtype=dtype)
    boff = 0

    def next_nonblank():
        nonlocal pos
        while not lines[pos].strip():
            pos += 1
        s = lines[pos]
        pos += 1
        return s

    for _ in range(ngrids):
        hdr = [next_nonblank().split()[0] for _ in range(8)]
        g = {"grid_number": int(hdr[0]), "level": int(hdr[1]), "mx": int(hdr[2]),
             "my": int(hdr[3]), "xlow": fortran_float(hdr[4]),
             "ylow": fortran_float(hdr[5]), "dx": fortran_float(hdr[6]),
             "dy": fortran_float(hdr[7])}
        mx, my = g["mx"], g["my"]
        if bdata is None:
            qe = np.zeros((meqn, mx, my))
            for j in range(my):
                for i in range(mx):
                    s = next_nonblank().rstrip("\n")
                    qe[:, i, j] = [fortran_float(s[26 * k:26 * k + 26]) for k in range(meqn)]
        else:
            n = meqn * (mx + 2 * nghost) * (my + 2 * nghost)
            qe = bdata[boff:boff + n].astype(np.float64).reshape(
                (meqn, mx + 2 * nghost, my + 2 * nghost), order="F")
            boff += n
        g["qeta"] = qe
        grids.append(g)
    return time, grids
