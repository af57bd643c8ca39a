"""AMR level transfers: the oracle's restatements of update.f90 (fine ->
coarse) and filval.f90's interpolation loop (coarse -> fine) pinned by
closed forms and invariants, and the product's cell bodies (dclaw_amr.h, run
on the host) against the oracle bit for bit on random wet/dry patches."""
import ctypes

import numpy as np
import pytest

DRY = 1e-3
RHO_F, GRAV = 1000.0, 9.81
_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


# ----------------------------------------------------------------- update.f90
def _average_case(rx, ry, h, b, mom, bc):
    """One coarse cell (index 2, 2 of a 5 x 5 coarse array) from one fine
    block (at 2, 2 of a fine array); h, b: (rx, ry); mom: (6, rx, ry)."""
    qc = np.full((7, 5, 5), 7.0, order="F")
    auxc = np.zeros((1, 5, 5), order="F")
    auxc[0, 2, 2] = bc
    qf = np.zeros((7, rx + 4, ry + 4), order="F")
    auxf = np.zeros((1, rx + 4, ry + 4), order="F")
    qf[0, 2:2 + rx, 2:2 + ry] = h
    qf[1:, 2:2 + rx, 2:2 + ry] = mom
    auxf[0, 2:2 + rx, 2:2 + ry] = b
    return qc, auxc, qf, auxf


@pytest.mark.parametrize("rx,ry", [(2, 2), (4, 4), (4, 2)])
def test_average_constant_state_is_reproduced(oracle_mod, rx, ry):
    """A uniform wet block on the coarse cell's own bed comes back exactly
    (dyadic values, power-of-two ratios)."""
    O = oracle_mod
    vals = np.array([0.5, -0.25, 1.25, 8000.0, 0.125, 0.375])
    mom = np.broadcast_to(vals[:, None, None], (6, rx, ry)).copy()
    qc, auxc, qf, auxf = _average_case(rx, ry, np.full((rx, ry), 2.0), np.full((rx, ry), 3.0),
                                       mom, 3.0)
    out = O.amr_average(qc, auxc, qf, auxf, 2, 2, 1, 1, 2, 2, rx, ry, DRY)
    assert np.array_equal(out[:, 2, 2], np.concatenate([[2.0], vals]))
    out[:, 2, 2] = 7.0
    assert (out == 7.0).all()                        # nothing else touched


def test_average_conserves_mass_and_momentum_on_a_flat_bed(oracle_mod):
    """All wet on a flat bed: the coarse depth and momenta are the block
    means, so coarse value x area = sum of the fine values."""
    rng = np.random.default_rng(1)
    rx, ry = 3, 4
    h = rng.uniform(0.5, 3.0, (rx, ry))
    mom = rng.normal(size=(6, rx, ry))
    qc, auxc, qf, auxf = _average_case(rx, ry, h, np.full((rx, ry), 1.5), mom, 1.5)
    out = oracle_mod.amr_average(qc, auxc, qf, auxf, 2, 2, 1, 1, 2, 2, rx, ry, DRY)
    assert out[0, 2, 2] * rx * ry == pytest.approx(h.sum(), rel=1e-14)
    for m in range(6):
        assert out[1 + m, 2, 2] * rx * ry == pytest.approx(mom[m].sum(), rel=1e-13, abs=1e-14)


def test_average_lake_at_rest(oracle_mod):
    """Flat surface over a varying fine bed, coarse bed = mean fine bed: the
    coarse depth is the surface minus the coarse bed and the water stays at
    rest."""
    rng = np.random.default_rng(2)
    rx = ry = 4
    b = rng.uniform(0.0, 2.0, (rx, ry))
    eta = 5.0
    qc, auxc, qf, auxf = _average_case(rx, ry, eta - b, b, np.zeros((6, rx, ry)), b.mean())
    out = oracle_mod.amr_average(qc, auxc, qf, auxf, 2, 2, 1, 1, 2, 2, rx, ry, DRY)
    assert out[0, 2, 2] == pytest.approx(eta - b.mean(), rel=1e-14)
    assert (out[1:, 2, 2] == 0.0).all()


def test_average_wet_cells_only_and_dry_depth_in_hsum(oracle_mod):
    """One wet fine cell among three with depth below dry_tolerance: the mean
    surface and mean depth divide by nwet = 1, the momentum factor divides by
    the depth of all four (update.f90:151-164).  Surface 3, coarse bed 2 ->
    depth 1; hsum = 2 + 3 * 4e-4."""
    h = np.array([[2.0, 4e-4], [4e-4, 4e-4]])
    b = np.array([[1.0, 9.0], [9.0, 9.0]])
    mom = np.zeros((6, 2, 2))
    mom[:, 0, 0] = [0.6, -0.2, 1.2, 9000.0, 0.0, 0.05]
    mom[:, 1, 1] = 5.0                                # a dry cell's momenta are dropped
    qc, auxc, qf, auxf = _average_case(2, 2, h, b, mom, 2.0)
    out = oracle_mod.amr_average(qc, auxc, qf, auxf, 2, 2, 1, 1, 2, 2, 2, 2, DRY)
    assert out[0, 2, 2] == 1.0
    hsum = 2.0 + 3 * 4e-4
    np.testing.assert_allclose(out[1:, 2, 2], mom[:, 0, 0] / hsum, rtol=1e-15)


def test_average_dry_block_and_bed_above_surface(oracle_mod):
    O = oracle_mod
    # all dry (bdif nonzero): zeros in every component
    mom = np.full((6, 2, 2), 0.3)
    qc, auxc, qf, auxf = _average_case(2, 2, np.full((2, 2), 5e-4), np.zeros((2, 2)), mom, 0.0)
    assert (O.amr_average(qc, auxc, qf, auxf, 2, 2, 1, 1, 2, 2, 2, 2, DRY)[:, 2, 2] == 0.0).all()
    # wet, but the coarse bed is above the mean surface: depth 0, momenta 0
    qc, auxc, qf, auxf = _average_case(2, 2, np.full((2, 2), 1.0), np.zeros((2, 2)), mom, 4.0)
    assert (O.amr_average(qc, auxc, qf, auxf, 2, 2, 1, 1, 2, 2, 2, 2, DRY)[:, 2, 2] == 0.0).all()


# ----------------------------------------------------------------- filval.f90
def _interp_setup(mic, mjc, rx, ry):
    valc = np.zeros((7, mic, mjc), order="F")
    auxc = np.zeros((1, mic, mjc), order="F")
    mit, mjt = (mic - 2) * rx + 4, (mjc - 2) * ry + 4
    val = np.full((7, mit, mjt), -1.0, order="F")
    auxf = np.zeros((1, mit, mjt), order="F")
    ns = np.ones((mit, mjt), dtype=bool)
    return valc, auxc, val, auxf, ns


def _interp(O, valc, auxc, val, auxf, ns, rx, ry, sea=0.0):
    return O.amr_interp(valc, auxc, val, auxf, ns, rx, ry, DRY, sea, RHO_F, GRAV)


@pytest.mark.parametrize("rx,ry", [(2, 2), (4, 2), (3, 3)])
def test_interp_constant_state(oracle_mod, rx, ry):
    """A uniform wet coarse state on a flat bed fills every fine cell with the
    same state, exactly; ghost cells of the fine patch are not touched."""
    valc, auxc, val, auxf, ns = _interp_setup(5, 4, rx, ry)
    state = np.array([2.0, 0.5, -0.25, 1.25, 19620.0 * 0.5, 0.0, 0.125])
    valc[:] = state[:, None, None]
    auxc[:] = 1.0
    auxf[:] = 1.0
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    inner = out[:, 2:-2, 2:-2]
    assert np.array_equal(inner, np.broadcast_to(state[:, None, None], inner.shape))
    ring = np.ones(out.shape[1:], bool)
    ring[2:-2, 2:-2] = False
    assert (out[:, ring] == -1.0).all()


def test_interp_linear_surface_conserves_mass(oracle_mod):
    """A planar wet surface on a flat bed: the fine depths are the plane at
    the fine centres (unlimited slopes), summing to the coarse depth x area."""
    rx, ry = 4, 2
    valc, auxc, val, auxf, ns = _interp_setup(5, 5, rx, ry)
    I, J = np.meshgrid(np.arange(5), np.arange(5), indexing="ij")
    valc[0] = 3.0 + 0.25 * I - 0.125 * J
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    for i in range(1, 4):
        for j in range(1, 4):
            blk = out[0, (i - 1) * rx + 2:i * rx + 2, (j - 1) * ry + 2:j * ry + 2]
            assert blk.sum() == pytest.approx(rx * ry * valc[0, i, j], rel=1e-15)
            xo = (np.arange(1, rx + 1) - 0.5) / rx - 0.5
            yo = (np.arange(1, ry + 1) - 0.5) / ry - 0.5
            np.testing.assert_allclose(blk, valc[0, i, j] + 0.25 * xo[:, None] - 0.125 * yo[None, :],
                                       rtol=1e-15)


def test_interp_lake_at_rest_over_fine_bed(oracle_mod):
    """Flat surface, water at rest, a rough fine bed: fine depth = surface -
    fine bed (0 where the bed is above it), momenta stay 0."""
    rng = np.random.default_rng(4)
    rx = ry = 2
    valc, auxc, val, auxf, ns = _interp_setup(6, 5, rx, ry)
    auxc[0] = rng.uniform(0.0, 1.0, auxc.shape[1:])
    valc[0] = 2.0 - auxc[0]
    auxf[0] = rng.uniform(0.0, 2.5, auxf.shape[1:])
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    inner = (slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(out[0][inner], np.maximum(0.0, 2.0 - auxf[0][inner]),
                               rtol=0, atol=1e-14)
    wet = out[0][inner] > DRY
    for m in (1, 2, 3, 5, 6):
        assert (out[m][inner][wet] == 0.0).all()


def test_interp_dry_neighbour_surface_is_sea_level(oracle_mod):
    """A dry neighbour enters the surface slope at sea_level, not at its own
    bed: centre surface 5, left dry (-> sea level 0), right 6 gives the
    x slope min(1, 5) = 1; a dry neighbour's own bed (here 100) would give
    an extremum and slope 0."""
    rx, ry = 2, 2
    valc, auxc, val, auxf, ns = _interp_setup(3, 3, rx, ry)
    valc[0, :, :] = 5.0
    valc[0, 0, 1] = 0.0
    auxc[0, 0, 1] = 100.0
    valc[0, 2, 1] = 6.0
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    np.testing.assert_array_equal(out[0, 2:4, 2:4], [[4.75, 4.75], [5.25, 5.25]])


def test_interp_velocity_bound_falls_back_to_depth_weighting(oracle_mod):
    """A bump in the fine bed makes hvf / h leave the coarse velocity range:
    the momentum of the block is then spread in proportion to the fine depth
    and its sum equals the coarse momentum x area (filval.f90:362-381)."""
    rx = ry = 2
    valc, auxc, val, auxf, ns = _interp_setup(3, 3, rx, ry)
    valc[0] = 2.0
    valc[1] = 2.0
    valc[1, 0, 1] = 1.6                               # hu slope 0.4, velocities in [0.8, 1.2]
    valc[1, 2, 1] = 2.4
    auxf[0, 2, 2] = 1.5                               # a shallow fine cell: 1.9 / 0.5 > 1.2
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    h = out[0, 2:4, 2:4]
    hu = out[1, 2:4, 2:4]
    np.testing.assert_allclose(hu, 4.0 * 2.0 / h.sum() * h, rtol=1e-15)
    assert hu.sum() == pytest.approx(4.0 * 2.0, rel=1e-15)
    # without the bump the limited-linear values stay (velocities in range)
    auxf[0, 2, 2] = 0.0
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    np.testing.assert_allclose(out[1, 2:4, 2:4], [[1.9, 1.9], [2.1, 2.1]], rtol=1e-15)


def test_interp_dry_centre_gets_hydrostatic_pore_pressure(oracle_mod):
    """A dry coarse cell between wet hydrostatic neighbours: the pore
    pressure's range starts at rho_f g (filval.f90:296-305), so the new fine
    cells get pb = rho_f g h; momenta come out 0."""
    rx = ry = 2
    valc, auxc, val, auxf, ns = _interp_setup(3, 3, rx, ry)
    valc[0] = 2.0
    valc[4] = RHO_F * GRAV * 2.0
    valc[:, 1, 1] = 0.0                               # dry centre, bed 0 -> surface sea level 1.0
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry, sea=1.0)
    h = out[0, 2:4, 2:4]
    assert (h > DRY).all()
    np.testing.assert_allclose(out[4, 2:4, 2:4], RHO_F * GRAV * h, rtol=1e-15)
    assert (out[1:4, 2:4, 2:4] == 0.0).all()


def test_interp_unset_cells_keep_depth_but_not_momentum(oracle_mod):
    """Fine cells not flagged needs_set keep their depth; the momentum pass
    still rewrites them, as the reference's second loop has no flag test."""
    rx = ry = 2
    valc, auxc, val, auxf, ns = _interp_setup(3, 3, rx, ry)
    valc[0] = 2.0
    valc[1] = 1.0
    valc[1, 0, 1] = 0.5                               # hu slope 0.5, velocities in [0.25, 0.75]
    valc[1, 2, 1] = 1.5
    val[0, 2, 2] = 1.7
    ns[2, 2] = False
    out = _interp(oracle_mod, valc, auxc, val, auxf, ns, rx, ry)
    assert out[0, 2, 2] == 1.7
    assert (out[0, 2:4, 2:4][ns[2:4, 2:4]] == 2.0).all()
    np.testing.assert_array_equal(out[1, 2:4, 2:4], [[0.875, 0.875], [1.125, 1.125]])


# ------------------------------------------------- product bodies vs the oracle
def _random_levels(rng, mic, mjc, rx, ry, maux=3):
    valc = np.zeros((7, mic, mjc), order="F")
    auxc = np.asfortranarray(rng.uniform(-1.0, 3.0, (maux, mic, mjc)))
    h = rng.uniform(0.0, 4.0, (mic, mjc))
    h[rng.random((mic, mjc)) < 0.3] = 0.0
    h[rng.random((mic, mjc)) < 0.1] = DRY
    valc[0] = h
    valc[1:3] = rng.normal(size=(2, mic, mjc)) * h
    valc[3] = rng.uniform(0.5, 0.65, (mic, mjc)) * h
    valc[4] = rng.uniform(0.3, 1.1, (mic, mjc)) * RHO_F * GRAV * h
    valc[5] = rng.uniform(0.0, 1.0, (mic, mjc)) * h
    valc[6] = rng.uniform(-0.2, 0.4, (mic, mjc))
    mit, mjt = (mic - 2) * rx + 4, (mjc - 2) * ry + 4
    val = np.asfortranarray(rng.normal(size=(7, mit, mjt)))
    val[0] = np.abs(val[0]) * 2.0
    val[0][rng.random((mit, mjt)) < 0.2] = 0.0
    auxf = np.asfortranarray(rng.uniform(-1.0, 3.0, (maux, mit, mjt)))
    ns = rng.random((mit, mjt)) < 0.8
    return valc, auxc, val, auxf, ns


@pytest.mark.parametrize("seed,rx,ry", [(0, 2, 2), (1, 4, 4), (2, 4, 2), (3, 3, 5), (4, 1, 1)])
def test_product_bodies_bitwise(oracle_mod, emu_lib, seed, rx, ry):
    O = oracle_mod
    rng = np.random.default_rng(seed)
    mic, mjc = 9, 7
    valc, auxc, val, auxf, ns = _random_levels(rng, mic, mjc, rx, ry)
    for sea in (0.0, 1.5):
        ref = O.amr_interp(valc, auxc, val, auxf, ns, rx, ry, DRY, sea, RHO_F, GRAV)
        got = np.array(val, order="F", copy=True)
        nsb = np.asfortranarray(ns, dtype=np.uint8)
        emu_lib.emu_amr_interp(_p(valc), _p(auxc), auxc.shape[0], mic, mjc, _p(got), _p(auxf),
                               auxf.shape[0], got.shape[1], nsb.ctypes.data_as(ctypes.c_void_p),
                               rx, ry, ctypes.c_double(DRY), ctypes.c_double(sea),
                               ctypes.c_double(RHO_F), ctypes.c_double(GRAV))
        assert np.array_equal(got, ref, equal_nan=True), f"{(got != ref).sum()} values differ"
    # fine -> coarse over a block of the coarse patch
    qf = np.array(ref, order="F", copy=True)
    ni, nj, ci0, cj0 = mic - 3, mjc - 3, 1, 2
    fi0, fj0 = 2, 2
    assert fi0 + ni * rx <= qf.shape[1] and fj0 + nj * ry <= qf.shape[2]
    refc = O.amr_average(valc, auxc, qf, auxf, ci0, cj0, ni, nj, fi0, fj0, rx, ry, DRY)
    gotc = np.array(valc, order="F", copy=True)
    emu_lib.emu_amr_average(_p(gotc), _p(auxc), auxc.shape[0], mic, _p(qf), _p(auxf),
                            auxf.shape[0], qf.shape[1], ci0, cj0, ni, nj, fi0, fj0, rx, ry,
                            ctypes.c_double(DRY))
    assert np.array_equal(gotc, refc), f"{(gotc != refc).sum()} values differ"
    assert not np.array_equal(refc, valc)
