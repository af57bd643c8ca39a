"""On-disk formats: Fortran edit descriptors of valout's formats, the
dclaw.data reader (set_dig's sequence) and frame round trips."""
import numpy as np
import pytest

from dclaw_amd import io as DIO
from dclaw_amd import workloads as W


def test_fortran_e_descriptor():
    # gfortran Ew.d: 0.ddd...E+xx right-justified; 3-digit exponents drop the E
    assert DIO.fortran_e(1.0, 26, 16) == "    0.1000000000000000E+01"
    assert DIO.fortran_e(0.0, 26, 16) == "    0.0000000000000000E+00"
    assert DIO.fortran_e(-2.5e-100, 26, 16) == "   -0.2500000000000000E-99"
    assert DIO.fortran_e(-2.5e-101, 26, 16) == "   -0.2500000000000000-100"
    big = DIO.fortran_e(3.0e150, 26, 16)          # three-digit exponent: no E
    assert len(big) == 26 and "E" not in big and big.endswith("+151")
    assert DIO.fortran_e(123456.789, 18, 8) == "    0.12345679E+06"
    assert DIO.fortran_e(-0.5, 26, 16) == "   -0.5000000000000000E+00"
    for x in (np.pi, -1e-80, 3.5e200, 2.0 ** -1074):
        assert DIO.fortran_float(DIO.fortran_e(x, 26, 16)) == float(f"{x:.15e}")


def test_dclaw_data_reader_follows_set_dig(tmp_path):
    # clawutil's layout: comment header, then "value =: name  description"
    vals = ["2700.0", "1000.0", "0.64", "0.62", "0.6", "1.d-10", "40.0", "0.01", "0.005", "0.01",
            "1.0", "1000.0", "F", "0.06", "0", "0", "F", "0", "0.0", "0", "0", "0.2", "0.65",
            "0", "0.0", "0.5", "0.5", "2.0", "T", "2", "0"]
    lines = ["########################", "### DO NOT EDIT ###", ""]
    lines += [f"{v:<26} =: {n}  some description" for v, n in zip(vals, DIO.DCLAW_DATA_FIELDS)]
    path = tmp_path / "dclaw.data"
    path.write_text("\n".join(lines) + "\n")
    d = DIO.read_dclaw_data(str(path))
    assert d["kref"] == 1e-10 and d["phi"] == 40.0 and d["m0"] == 0.62
    assert d["mom_autostop"] == 1 and d["dd_manning"] == 0 and d["momlevel"] == 2
    p, m0 = DIO.dig_params_from_data(str(path))
    assert m0 == 0.62 and p.m_crit == 0.64 and p.kref == 1e-10
    assert abs(p.phi - np.deg2rad(40.0)) < 1e-15


def test_dclaw_data_round_trip(tmp_path):
    p = W.config0_params()
    p.kref = 3.3e-11
    p.src2method = 1
    path = str(tmp_path / "dclaw.data")
    DIO.write_dclaw_data(path, p, 0.61)
    q, m0 = DIO.dig_params_from_data(path)
    assert m0 == 0.61
    for name in ("rho_s", "rho_f", "m_crit", "mref", "kref", "phi_deg", "delta", "mu",
                 "alpha_c", "c1", "sigma_0", "src2method", "alphamethod"):
        assert getattr(q, name) == getattr(p, name), name


def _grid(seed=0):
    from helpers import slope_patch
    pt = slope_patch(9, 5, seed=seed)
    q = np.array(pt.q, order="F", copy=True)
    q[3, 4, 3] = 1e-95                     # rounded to 0 on output (interior)
    return DIO.Grid(q=q, aux=pt.aux, xlow=-3.0, ylow=2.5, dx=0.5, dy=0.25, grid_number=7), pt


@pytest.mark.parametrize("fmt", ["ascii", "binary64", "binary32"])
def test_frame_round_trip(tmp_path, fmt):
    g, pt = _grid()
    DIO.write_frame(str(tmp_path), 3, 1.25, [g], fmt)
    t, grids = DIO.read_frame(str(tmp_path), 3)
    assert t == 1.25 and len(grids) == 1
    r = grids[0]
    assert (r["grid_number"], r["mx"], r["my"]) == (7, 9, 5)
    assert (r["xlow"], r["ylow"], r["dx"], r["dy"]) == (-3.0, 2.5, 0.5, 0.25)
    q = np.array(g.q, order="F", copy=True)
    q[3, 4, 3] = 0.0
    eta = q[0] + g.aux[0] - q[6]
    full = np.concatenate([q, eta[None]], axis=0)
    if fmt == "ascii":
        np.testing.assert_allclose(r["qeta"], full[:, 2:-2, 2:-2], rtol=1e-15, atol=0)
    elif fmt == "binary64":
        assert np.array_equal(r["qeta"], full)
    else:
        assert np.array_equal(r["qeta"], full.astype(np.float32).astype(np.float64))


def test_frame_headers_match_valout_formats(tmp_path):
    g, _ = _grid()
    DIO.write_frame(str(tmp_path), 0, 0.5, [g], "binary64")
    q = (tmp_path / "fort.q0000").read_text().splitlines()
    assert q[0] == "     7                 grid_number"
    assert q[1] == "     1                 AMR_level"
    assert q[2] == "     9                 mx"
    assert q[4] == "   -0.3000000000000000E+01    xlow"
    assert q[7] == "    0.2500000000000000E+00    dy"
    t = (tmp_path / "fort.t0000").read_text().splitlines()
    assert t[0] == "    0.50000000E+00    time"
    assert t[1] == "     8                 meqn"
    assert t[2] == "     1                 ngrids"
    assert t[4] == "     2                 ndim"
    assert t[6] == "  binary64             format"
