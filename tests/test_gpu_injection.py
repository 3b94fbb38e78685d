"""The split of a reaction increment over the seven D3Q7 populations on the GPU, through the
C ABI: population i of the reacting voxel gains t_i * dC (t_0 = 1/4, t_i = 1/8) before
streaming (update_rxnLattices, src/complab3d_processors_part1.hh:345-415; the speciation's
increment, src/complab3d_processors_part4_eqsolver.hh:1355-1494).

One reacting voxel in a field at rest: after one step it holds dC / 4 more and each face
neighbour dC / 8 more; nothing else changes. The kinetics case has dC in closed form (the
dual-Monod rate); the speciation case checks the split whatever dC is, in both speciation
modes (the fused one defers the increments to the next load, and the download folds them
in) and both layouts."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega
from test_oracle_injection import CENTER, FACES, NX, NY, NZ, box
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact

# FP64: a few ulps; FP32 store: populations hold the concentrations themselves (no
# reference subtracted), so a few FP32 ulps (1e-6) of the largest one
TOL = {0: 1e-15, 1: 1e-6}


def lattice(clb, prec, layout):
    m = box()
    kind, om = d3q19_kind_omega()
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(m)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [1] * 8, [0.0] * 8)
    lat.set_velocity(np.zeros(3 * m.size))
    return lat, m


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_kinetics_increment_split(clb, prec, layout):
    print(f"case: kinetics split prec={prec} layout={layout}", flush=True)
    lat, m = lattice(clb, prec, layout)
    D, A, B = 0.2, 0.05, 7.0
    b = np.zeros(m.size)
    b[CENTER] = B
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.solutes_init([D, A, 0.0, 0.0])
    lat.step(1)
    c = lat.conc()
    bm = lat.biomass()[0]
    lat.close()
    r = 1e-3 * B * D / (1e-2 + D) * A / (1e-2 + A)
    dC = np.array([-r, -r, r, r])
    c0 = np.array([D, A, 0.0, 0.0])
    tol = TOL[prec] * max(D, 1.0 if prec == 0 else D)
    for s in range(4):
        exp = np.where(m == 2, c0[s], 0.0)
        exp[CENTER] += dC[s] / 4
        exp[FACES] += dC[s] / 8
        assert np.max(np.abs(c[s] - exp)) < tol, s
    db = bm - b
    assert abs(db[CENTER] - 0.4 * r) < TOL[prec] * B
    db[CENTER] = 0.0
    assert not db.any()


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_speciation_increment_split(clb, mode, layout):
    print(f"case: speciation split mode={mode} layout={layout}", flush=True)
    lat, m = lattice(clb, 0, layout)
    lat.set_speciation_mode(mode)
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    c0 = carbonate_exact(1e-3, 1e-4)
    lat.solutes_init(c0)
    cc = c0.copy()
    cc[2] *= 1.05
    cc[3] *= 0.96
    for s in (2, 3):
        f = np.where(m == 2, c0[s], 0.0)
        f[CENTER] = cc[s]
        lat.set_solute_field(s, f)
    lat.step(1)
    c = lat.conc()
    lat.close()
    others = m == 2
    others[CENTER] = False
    others[FACES] = False
    dcen = np.zeros(4)
    for s in range(4):
        dcen[s] = c[s][CENTER] - (cc[s] / 4 + 3 * c0[s] / 4)
        dnb = c[s][FACES] - (7 * c0[s] / 8 + cc[s] / 8)
        assert np.max(np.abs(dnb - dcen[s] / 2)) < 1e-9 * abs(dcen[s]) + 1e-15, s
        assert np.max(np.abs(c[s][others] - c0[s])) < 1e-15, s
    assert np.all(np.abs(dcen) > 1e-10)
    T = CARB_S.T @ cc
    assert np.max(np.abs(CARB_S.T @ (cc + 4 * dcen) - T) / np.abs(T)) < 1e-9
