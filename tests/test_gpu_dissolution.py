"""Mineral dissolution (surfaceDissolutionKinetics3D, src/dissolutionVOP.hh:229-459) on the
GPU against the oracle's restatement, through the C ABI.

A block of solid voxels carrying phase 1 holds an inventory in an immobile solute; the
compiled hook of kinetics/abiotic_second_order dissolves it first order in the attacking
solute (1) and in the mineral, releasing solute 0 one for one. Pinned by:
  * FP64 GPU vs the oracle (same step order: water after the biotic kinetics, products
    and inventory loss joined to the step's increments before streaming), 1e-10;
  * the element balances the stoichiometry fixes, in a box closed by walls: mineral +
    released solute, and attacker + released solute, both conserved to rounding;
  * the fused speciation path (deferred increments) against the separate one, with a
    tableau whose mineral row is zero;
  * reopening: a phase voxel whose inventory falls below the threshold becomes pore."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact

NX, NY, NZ = 24, 12, 10
MIN = 4          # the immobile solute holding the inventory
S = 5


def geometry(closed):
    mask = np.full((NZ, NY, NX), 2, dtype=np.uint8)
    mask[3:7, 4:8, 9:14] = 1          # the mineral block: a bounce-back wall, as a sealed voxel is
    phase = np.zeros_like(mask)
    phase[3:7, 4:8, 9:14] = 1
    phase[4:6, 5:7, 10:13] = 0        # its core is inert rock (no phase): only the rim dissolves
    if closed:  # walls across x: bounce-back next to the pore space, no-dynamics behind it
        mask[:, :, :2] = 0
        mask[:, :, -2:] = 0
        mask[:, :, 2] = 1
        mask[:, :, -3] = 1
    return mask.ravel(), phase.ravel()


def setup(clb, closed, prec=0, mode=None, tableau=False):
    mask, phase = geometry(closed)
    N = mask.size
    kind = np.full(256, 2, dtype=np.int32)
    kind[0], kind[1] = 0, 1
    om = np.full((S, 256), 1.0 / 0.8)
    bct = [0, 1] * S
    if tableau:  # carbonate waters at equilibrium (species 0-3), no mineral in solution
        ci, c0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
        bcv = [ci[0], 0, ci[1], 0, ci[2], 0, ci[3], 0, 0.0, 0]
        c0 = list(c0) + [0.0]
    else:
        bcv = [0.0, 0, 2e-2, 0, 0.0, 0, 0.0, 0, 0.0, 0]
        c0 = [1e-4, 5e-3, 1e-4, 1e-4, 0.0]
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    if mode is not None:
        lat.set_speciation_mode(mode)
    lat.set_dynamics(kind, np.full(256, 1.25))
    lat.upload_mask(mask)
    lat.set_solutes(om, bct, bcv)
    lat.set_immobile([0, 0, 0, 0, 1])
    lat.set_velocity(np.zeros((N, 3)))
    lat.set_kinetics(clb.KINETICS_NONE, 1.0, 0, 1)
    if tableau:
        Sm = np.vstack([CARB_S, np.zeros((1, CARB_S.shape[1]))])
        lk = np.append(CARB_LOGK, 0.0)
        lat.set_tableau(Sm, lk, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(c0)
    minf = np.where(phase > 0, 10.0, 0.0)
    lat.set_solute_field(MIN, minf)
    lat.set_phase(phase)
    lat.set_dissolution(True, [1], [MIN], True)
    ref = oracle.AdeRun(NX, NY, NZ, mask, kind, om, bct, bcv, np.zeros((N, 3)), c0, immobile=[0, 0, 0, 0, 1],
                        dissolution=(phase, {1: MIN}, True),
                        tableau=None if not tableau else (np.vstack([CARB_S, np.zeros((1, CARB_S.shape[1]))]),
                                                          np.append(CARB_LOGK, 0.0), CARB_COMP))
    ref.set_field(MIN, minf)
    return lat, ref, mask, phase


@pytest.mark.gpu
def test_dissolution_fp64_vs_oracle(clb):
    print("case: dissolution FP64 vs oracle, open channel, 120 steps", flush=True)
    lat, ref, mask, phase = setup(clb, closed=False)
    lat.step(120)
    ref.run(120)
    cg, cr = lat.conc(), ref.conc()
    m0 = 10.0 * (phase > 0).sum()
    assert cg[MIN].sum() < m0 - 1.0  # it did dissolve
    for s in range(S):
        assert rel_l2(cg[s], cr[s]) < 1e-10, s
    lat.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-6)])
def test_dissolution_element_balance(clb, prec, tol):
    print(f"case: dissolution element balance, closed box, precision {prec}", flush=True)
    lat, ref, mask, phase = setup(clb, closed=True, prec=prec)
    c = lat.conc()
    m0 = c[MIN].sum()
    a_rel0 = c[1].sum() + c[0].sum()
    # first step, closed form: uniform attacker 5e-3 and inventory 10 at every phase voxel
    # with an open face: each loses k_diss * A * mineral * dt = 2e-2 * 5e-3 * 10 * 1
    wet = 0
    m3 = mask.reshape(NZ, NY, NX)
    for z, y, x in zip(*np.nonzero(phase.reshape(NZ, NY, NX))):
        wet += any(0 <= z + dz < NZ and 0 <= y + dy < NY and 1 <= x + dx <= NX - 2 and m3[z + dz, y + dy, x + dx] == 2
                   for dz, dy, dx in ((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)))
    lat.step(1)
    c = lat.conc()
    assert wet > 0 and abs((m0 - c[MIN].sum()) - wet * 1e-3) <= max(tol, 1e-13) * m0
    # the attacker is consumed one for one with the product released: their sum is conserved
    # (both travel together, into and out of the walls' reflected populations too)
    for _ in range(3):
        lat.step(60)
        c = lat.conc()
        assert abs(c[1].sum() + c[0].sum() - a_rel0) <= tol * a_rel0
    assert c[0].sum() > 1.0 and c[MIN].sum() < m0 - 1.0
    lat.close()


@pytest.mark.gpu
def test_dissolution_fused_matches_separate(clb):
    print("case: dissolution with speciation, fused (graph) vs separate vs oracle", flush=True)
    out = []
    for mode, graph in ((0, True), (1, False)):
        lat, ref, mask, phase = setup(clb, closed=False, mode=mode, tableau=True)
        if graph:
            lat.set_graph(True, 20)
        lat.step(80)
        out.append(lat.conc())
        lat.close()
    ref.run(80)
    cr = ref.conc()
    for s in range(S):
        assert rel_l2(out[0][s], out[1][s]) < 2e-6, s
        assert rel_l2(out[1][s], cr[s]) < 2e-6, s


@pytest.mark.gpu
def test_dissolution_reopens_voxels(clb):
    print("case: dissolution then reopening by node conversion", flush=True)
    lat, ref, mask, phase = setup(clb, closed=False)
    lat.step(300)
    c = lat.conc()
    rim = phase > 0
    thr = float(np.quantile(c[MIN][rim], 0.5))
    changed = lat.node_conversion(MIN, 2, 1, 1, -1.0, thr)
    m = lat.download_mask()
    ph = lat.phase()
    reopened = rim & (m == 2)
    assert changed == reopened.sum() > 0
    assert not ph[reopened].any()
    assert np.array_equal(reopened, rim & (c[MIN] < thr))
    lat.close()


@pytest.mark.gpu
def test_dissolution_autotune_before_the_first_step(clb):
    """The autotuner runs the step on its own; the water buffer the dissolution kernels read
    is made for it as for clb_step (regression: it was only made by the first clb_step)."""
    print("case: dissolution, autotune before any step, then 40 steps vs no autotune", flush=True)
    out = []
    for tune in (True, False):
        lat, ref, mask, phase = setup(clb, closed=False)
        if tune:
            assert lat.autotune(2)["trials"] > 0
            lat.set_block_shape(clb.KERNEL_ADE, 64, 4)
        lat.step(40)
        out.append(lat.conc())
        lat.close()
    assert np.array_equal(out[0], out[1])
