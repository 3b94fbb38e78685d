"""Pins how the oracle splits a reaction increment over the seven D3Q7 populations
(update_rxnLattices, src/complab3d_processors_part1.hh:345-415, and the speciation's
increment in run_equilibrium_biotic, src/complab3d_processors_part4_eqsolver.hh:1355-1494):
population i of the reacting voxel gains t_i * dC, t_0 = 1/4 and t_i = 1/8 for the six
moving populations, before streaming.

One reacting voxel in a field at rest (u = 0, every voxel at its collision equilibrium)
makes the split visible after one step: the voxel keeps its rest population, dC / 4, and
each face neighbour receives one moving population, dC / 8; no other voxel changes. Putting
all of dC in the rest population, spreading it evenly, or sending it to the wrong
neighbours fails these checks."""
import numpy as np

from conftest import d3q19_kind_omega
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact

NX = NY = NZ = 9
CX, CY, CZ = 4, 4, 4


def box():
    """A 9^3 box of pore voxels (code 2), bounce-back planes at x = 0 and x = 8, periodic in
    y and z."""
    m = np.full((NZ, NY, NX), 2, dtype=np.uint8)
    m[:, :, 0] = 1
    m[:, :, NX - 1] = 1
    return m.ravel()


def idx(x, y, z):
    return x + NX * (y + NY * z)


CENTER = idx(CX, CY, CZ)
FACES = [idx(CX + 1, CY, CZ), idx(CX - 1, CY, CZ), idx(CX, CY + 1, CZ), idx(CX, CY - 1, CZ),
         idx(CX, CY, CZ + 1), idx(CX, CY, CZ - 1)]


def test_kinetics_increment_split_single_voxel(oracle):
    m = box()
    kind, _ = d3q19_kind_omega()
    om = np.full((4, 256), 1.0 / 0.8)
    D, A, B = 0.2, 0.05, 7.0
    b = np.zeros(m.size)
    b[CENTER] = B
    run = oracle.AdeRun(NX, NY, NZ, m, kind, om, [1] * 8, [0.0] * 8, np.zeros((m.size, 3)), [D, A, 0.0, 0.0],
                        kinetics=True, dt=1.0, biomass=b)
    run.run(1)
    c = run.conc()
    # the dual-Monod rate of the voxel (defineKinetics.hh's closed form, below its cap)
    r = 1e-3 * B * D / (1e-2 + D) * A / (1e-2 + A)
    dC = np.array([-r, -r, r, r])
    c0 = np.array([D, A, 0.0, 0.0])
    fluid = m == 2
    for s in range(4):
        exp = np.where(fluid, c0[s], 0.0)
        exp[CENTER] += dC[s] / 4
        exp[FACES] += dC[s] / 8
        assert np.max(np.abs(c[s] - exp)) < 1e-15, s
    # the biomass grows only in the voxel, by Y r
    db = run.bmass - b
    assert abs(db[CENTER] - 0.4 * r) < 1e-15
    db[CENTER] = 0.0
    assert not db.any()


def test_speciation_increment_split_single_voxel(oracle):
    """Every pore voxel is at the carbonate equilibrium except the centre, whose complexes are
    moved off it. Only the centre's solve gives an increment. Streaming alone would leave the
    centre at c/4 + 3 c0/4 and a face neighbour at 7 c0/8 + c/8; the increment adds dC/4 and
    dC/8 to them, whatever dC is."""
    m = box()
    kind, _ = d3q19_kind_omega()
    om = np.full((4, 256), 1.0 / 0.8)
    c0 = carbonate_exact(1e-3, 1e-4)
    run = oracle.AdeRun(NX, NY, NZ, m, kind, om, [1] * 8, [0.0] * 8, np.zeros((m.size, 3)), c0,
                        tableau=(CARB_S, CARB_LOGK, CARB_COMP), eq_tol=1e-10, eq_max_iter=200)
    cc = c0.copy()
    cc[2] *= 1.05
    cc[3] *= 0.96
    for s in (2, 3):
        f = np.full(m.size, c0[s])
        f[CENTER] = cc[s]
        run.set_field(s, f)
    run.run(1)
    c = run.conc()
    fluid = m == 2
    others = fluid.copy()
    others[CENTER] = False
    others[FACES] = False
    dcen = np.zeros(4)
    for s in range(4):
        dcen[s] = c[s][CENTER] - (cc[s] / 4 + 3 * c0[s] / 4)
        dnb = c[s][FACES] - (7 * c0[s] / 8 + cc[s] / 8)
        # the populations are stored as deviations from t_i * 1: a few ulps of 1 absolute
        assert np.max(np.abs(dnb - dcen[s] / 2)) < 1e-9 * abs(dcen[s]) + 1e-15, s
        assert np.max(np.abs(c[s][others] - c0[s])) < 1e-15, s
    # the complexes were off equilibrium: the increment is not negligible
    assert np.all(np.abs(dcen) > 1e-10)
    # and it is the one the solve gives: 4 dcen is the voxel's limited increment, whose
    # totals (the components' mass) are those the voxel had
    dC = 4 * dcen
    T = CARB_S.T @ cc
    T1 = CARB_S.T @ (cc + dC)
    assert np.max(np.abs(T1 - T) / np.abs(T)) < 1e-9
