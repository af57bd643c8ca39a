"""AMR region fills on the GPU (dclaw_amd.amr.AmrHierarchy) against the
oracle's composition (tests/amr_fill_ref.py): the ghost fill of every patch
of a 3-level hierarchy (same-level copies, filpatch's interpolation from the
coarser level, the boundary rules), the same between two stored times of
the coarser level, the regrid of the finest level (filval.f90 with setaux),
and the fine -> coarse average -- all bit for bit."""
import numpy as np
import pytest
import torch

from amr_fill_ref import RefHierarchy, RefLevel, RefPatch, from_device

pytestmark = pytest.mark.gpu

DRY, SEA, RHO_F, GRAV = 1e-3, 0.0, 1000.0, 9.81
PHI, THETA = 0.6, 0.0
MAUX = 6


def _to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def _bed(x, y):
    return 3.0 + 0.05 * x - 0.03 * y + 0.4 * np.sin(0.3 * x) * np.cos(0.2 * y)


def _state(rng, mx, my, x0, y0, dx, dy):
    """A patch state with 2 ghost cells: a mound of depth over the bed with
    dry cells around it, random momenta, m in [0.5, 0.65], pb near
    hydrostatic; the ghost cells hold junk the fill must overwrite."""
    x = x0 + (np.arange(-2, mx + 2) + 0.5) * dx
    y = y0 + (np.arange(-2, my + 2) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    b = _bed(X, Y)
    eta = 5.2 - 0.04 * (X - 20.0) ** 2 / 10.0 - 0.03 * (Y - 18.0) ** 2 / 10.0
    h = np.maximum(eta - b, 0.0)
    h[h < 0.01] = 0.0
    q = np.zeros((7,) + h.shape, order="F")
    q[0] = h
    q[1] = h * rng.normal(0.0, 1.0, h.shape)
    q[2] = h * rng.normal(0.0, 1.0, h.shape)
    m = rng.uniform(0.5, 0.65, h.shape)
    q[3] = h * m
    q[4] = RHO_F * GRAV * h * rng.uniform(0.6, 1.1, h.shape)
    q[5] = h * rng.uniform(0.0, 1.0, h.shape)
    q[6] = rng.uniform(0.0, 0.01, h.shape) * (h > 0)
    aux = np.zeros((MAUX, mx + 4, my + 4), order="F")
    aux[0] = b
    aux[1] = PHI
    # junk in the ghost cells
    for a in (q, aux):
        a[:, :2, :] = 777.0
        a[:, -2:, :] = 777.0
        a[:, :, :2] = 777.0
        a[:, :, -2:] = 777.0
    return q, aux


# level boxes (ilo, ihi, jlo, jhi) in each level's index space
L0 = dict(dx=2.0, dy=2.0, n=(24, 20), boxes=[(0, 23, 0, 19)])
L1 = dict(dx=1.0, dy=1.0, n=(48, 40), boxes=[(4, 19, 6, 21), (20, 35, 6, 17), (0, 7, 24, 39)])
L2 = dict(dx=0.5, dy=0.5, n=(96, 80), boxes=[(12, 27, 16, 31), (28, 39, 16, 27)])


def _build(bc, seed=1, time_levels=False):
    from dclaw_amd import amr as A
    rng = np.random.default_rng(seed)
    levs, rlevs = [], []
    for k, L in enumerate((L0, L1, L2)):
        lev = A.AmrLevel(L["dx"], L["dy"], L["n"][0], L["n"][1], 2 if k else 1, 2 if k else 1)
        rlev = RefLevel(L["dx"], L["dy"], L["n"][0], L["n"][1], 2 if k else 1, 2 if k else 1)
        for (ilo, ihi, jlo, jhi) in L["boxes"]:
            mx, my = ihi - ilo + 1, jhi - jlo + 1
            q, aux = _state(rng, mx, my, ilo * L["dx"], jlo * L["dy"], L["dx"], L["dy"])
            if k == 0:
                q[:, :2, :] = q[:, 2:3, :]; q[:, -2:, :] = q[:, -3:-2, :]
                q[:, :, :2] = q[:, :, 2:3]; q[:, :, -2:] = q[:, :, -3:-2]
                aux[:, :2, :] = aux[:, 2:3, :]; aux[:, -2:, :] = aux[:, -3:-2, :]
                aux[:, :, :2] = aux[:, :, 2:3]; aux[:, :, -2:] = aux[:, :, -3:-2]
            p = A.AmrPatch(_to_dev(q), _to_dev(aux), ilo, jlo, t=1.0)
            rp = RefPatch(q.copy(order="F"), aux.copy(order="F"), ilo, jlo, t=1.0)
            if time_levels and k == 1:
                qo, _ = _state(rng, mx, my, ilo * L["dx"], jlo * L["dy"], L["dx"], L["dy"])
                p.q_old, p.t_old = _to_dev(qo), 0.0
                rp.q_old, rp.t_old = qo, 0.0
            lev.patches.append(p)
            rlev.patches.append(rp)
        levs.append(lev)
        rlevs.append(rlev)
    H = A.AmrHierarchy(levs, 0.0, 0.0, bc, DRY, SEA, RHO_F, GRAV, PHI, THETA)
    return H, levs, rlevs


def _ref(oracle_mod, rlevs, bc):
    return RefHierarchy(oracle_mod, rlevs, 0.0, 0.0, bc, DRY, SEA, RHO_F, GRAV, PHI, THETA)


def _same(p, rp, what):
    q, aux = from_device(p.q), from_device(p.aux)
    assert np.array_equal(q, rp.q), f"{what}: q differs at {np.argwhere(q != rp.q)[:4]}"
    assert np.array_equal(aux, rp.aux), f"{what}: aux differs"
    assert not (q == 777.0).any(), f"{what}: a ghost cell was not filled"


@pytest.mark.parametrize("bc", [(1, 1, 1, 1), (3, 1, 1, 3)])
def test_ghost_fill_equals_oracle(oracle_mod, bc):
    H, levs, rlevs = _build(bc)
    R = _ref(oracle_mod, rlevs, bc)
    for level in (1, 2):
        for k in range(len(levs[level].patches)):
            print(f"[ghost fill] bc {bc} level {level} patch {k}", flush=True)
            H.fill_ghosts(level, k, 1.0)
            R.fill_ghosts(level, k, 1.0)
            _same(levs[level].patches[k], rlevs[level].patches[k], f"level {level} patch {k}")


def test_ghost_fill_between_coarse_times_equals_oracle(oracle_mod):
    bc = (1, 3, 1, 1)
    H, levs, rlevs = _build(bc, seed=2, time_levels=True)
    R = _ref(oracle_mod, rlevs, bc)
    for t in (0.25, 0.0):
        for k in range(len(levs[2].patches)):
            print(f"[ghost fill] t {t} level 2 patch {k}", flush=True)
            H.fill_ghosts(2, k, t)
            R.fill_ghosts(2, k, t)
            _same(levs[2].patches[k], rlevs[2].patches[k], f"t {t} patch {k}")


def test_regrid_equals_oracle(oracle_mod):
    bc = (1, 1, 3, 1)
    H, levs, rlevs = _build(bc, seed=3)
    R = _ref(oracle_mod, rlevs, bc)

    def bed_np(level, ilo, jlo, mx, my):
        dx = (L0, L1, L2)[level]["dx"]
        x = (ilo + np.arange(-2, mx + 2) + 0.5) * dx
        y = (jlo + np.arange(-2, my + 2) + 0.5) * dx
        X, Y = np.meshgrid(x, y, indexing="ij")
        return np.asfortranarray(_bed(X, Y) + 0.01)

    def bed_dev(level, ilo, jlo, mx, my):
        return torch.from_numpy(np.ascontiguousarray(bed_np(level, ilo, jlo, mx, my).T)).cuda()

    boxes = [(16, 31, 20, 35), (32, 43, 14, 21)]
    print("[regrid] level 2", flush=True)
    H.regrid(2, boxes, bed_dev, 1.0)
    R.regrid(2, boxes, bed_np, 1.0)
    for k in range(len(boxes)):
        p, rp = levs[2].patches[k], rlevs[2].patches[k]
        assert (p.ilo, p.jlo, p.mx, p.my) == (rp.ilo, rp.jlo, rp.mx, rp.my)
        q, aux = from_device(p.q), from_device(p.aux)
        assert np.array_equal(q, rp.q), f"regrid patch {k}: q"
        assert np.array_equal(aux, rp.aux), f"regrid patch {k}: aux"
        # cells the old patches covered keep their values
    old0 = _build(bc, seed=3)[2][2].patches[0]
    a0, b0 = 16, 20
    ov = (slice(None), slice(2, 2 + 12), slice(2, 2 + 12))     # fine cells 16..27 x 20..31
    newq = from_device(levs[2].patches[0].q)
    assert np.array_equal(newq[ov], old0.q[:, 2 + a0 - 12:2 + a0 - 12 + 12, 2 + b0 - 16:2 + b0 - 16 + 12])


def test_average_down_equals_oracle(oracle_mod):
    bc = (1, 1, 1, 1)
    H, levs, rlevs = _build(bc, seed=4)
    R = _ref(oracle_mod, rlevs, bc)
    for level in (2, 1):
        print(f"[average] level {level}", flush=True)
        H.average_down(level)
        R.average_down(level)
        for k, (p, rp) in enumerate(zip(levs[level - 1].patches, rlevs[level - 1].patches)):
            assert np.array_equal(from_device(p.q), rp.q), f"average into level {level - 1} patch {k}"


def test_filpatch_kernel_rejects_an_uncovered_region():
    """A region the coarser level does not cover is an error, not a guess."""
    from dclaw_amd import _lib
    H, levs, rlevs = _build((1, 1, 1, 1))
    q = torch.zeros((10, 10, 7), dtype=torch.float64, device="cuda")
    aux = torch.zeros((10, 10, MAUX), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.DClawError):
        H.fill_region(1, (40, 49, 30, 39), q, aux, 0, 0, 1.0, True)
