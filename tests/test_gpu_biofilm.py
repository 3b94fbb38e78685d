"""Impermeable biofilm: a biofilm code that is bounce-back for the flow and advection-diffusion
for the solutes, against the FP64 oracle.

soluteDomainSetup (src/complab_functions.hh:296-340) gives every biofilm code
AdvectionDiffusionBGK(substr_bMassOmega) whatever the biofilm permeability; only
NSdomainSetup (:271-283) turns an impermeable biofilm (permRatio = 0, the setting of every
reference example) into BounceBack. So the biomass in such voxels still reacts with the
solutes that diffuse in, while the flow goes around them. The GPU's solute lattices take
their own table (clb_set_solute_dynamics); the oracle's transport run is handed that same
solute table, and the velocity comes from a flow solved with the flow table.

Case: the 64 x 32 x 32 sphere pack; pore voxels with a wall face neighbour are biofilm
(code 3) with probability 1/2; biomass U[10, 100] in every biofilm voxel and in 1 % of the
pore voxels; four solutes with dual-Monod kinetics; the pore <-> biofilm threshold rule
(updateLocalMaskNtotalLattices3D, src/complab3d_processors_part2.hh:99-210) at B = 30, so
voxels flip both ways during the run."""
import numpy as np
import pytest

from conftest import rel_l2

NX, NY, NZ = 64, 32, 32
BCT = [0, 1, 0, 1, 1, 1, 1, 1]
BCV = [1.0, 0, 0.5, 0, 0, 0, 0, 0]
C0 = [0.2, 0.1, 0.05, 0.0]
BIO = 3
THR = 30.0


def biofilm_case(clb):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    m = pack.reshape(NZ, NY, NX).copy()
    wall = m <= 1
    near = np.zeros_like(wall)
    for ax in range(3):
        for sh in (1, -1):
            near |= np.roll(wall, sh, axis=ax)
    rng = np.random.default_rng(5)
    inner = np.zeros_like(wall)
    inner[:, :, 1:NX - 1] = True
    bio = (m == 2) & near & inner & (rng.random(m.shape) < 0.5)
    m[bio] = BIO
    mask = m.ravel()
    b = np.zeros(mask.size)
    sel = mask == BIO
    b[sel] = rng.uniform(10.0, 100.0, sel.sum())
    pores = np.flatnonzero((mask == 2) & inner.ravel())
    seeds = rng.choice(pores, size=pores.size // 100, replace=False)
    b[seeds] = rng.uniform(10.0, 100.0, seeds.size)
    # flow: biofilm impermeable (bounce-back); solutes: biofilm is advection-diffusion
    flow_kind = np.full(256, 2, dtype=np.int32)
    flow_kind[0], flow_kind[1], flow_kind[BIO] = 0, 1, 1
    ade_kind = np.full(256, 2, dtype=np.int32)
    ade_kind[0], ade_kind[1] = 0, 1
    flow_om = np.full(256, 1.0 / 0.8)
    # the biofilm's own solute relaxation (substr_bMassOmega: a smaller diffusivity)
    ade_om = np.full((4, 256), 1.0 / 0.8)
    ade_om[:, BIO] = 1.0 / 0.65
    return mask, b, flow_kind, ade_kind, flow_om, ade_om


@pytest.fixture(scope="module")
def case(clb, oracle):
    mask, b, flow_kind, ade_kind, flow_om, ade_om = biofilm_case(clb)
    f, _ = oracle.flow_run(NX, NY, NZ, mask, flow_kind, flow_om, 1.0, 0.999, 400)
    u, _ = oracle.flow_velocity(NX, NY, NZ, mask, flow_kind, 1.0, 0.999, f)
    return mask, b, flow_kind, ade_kind, flow_om, ade_om, u


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_impermeable_biofilm_reacts(clb, oracle, case, prec, tol):
    mask, b, flow_kind, ade_kind, flow_om, ade_om, u = case
    steps = 121
    print(f"case: impermeable biofilm, prec={prec}, {steps} steps", flush=True)
    run = oracle.AdeRun(NX, NY, NZ, mask, ade_kind, ade_om, BCT, BCV, u, C0, kinetics=True, biomass=b,
                        mask_update=(THR, 2, BIO))
    run.run(steps)
    ref = run.conc()
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(flow_kind, flow_om)
    lat.set_solute_dynamics(ade_kind)
    lat.upload_mask(mask)
    lat.set_solutes(ade_om, BCT, BCV)
    lat.set_velocity(u)
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, THR, 2, BIO)
    lat.solutes_init(C0)
    lat.step(steps)
    c = lat.conc()
    m_gpu = lat.download_mask()
    bm = lat.biomass()
    lat.close()
    bio0 = mask == BIO
    # no flow through the biofilm, yet the donor is consumed inside it
    assert np.abs(u.reshape(-1, 3)[bio0]).max() == 0.0
    assert ref[0][bio0].mean() > 0.0 and ref[2][bio0].max() > 0.06
    # both directions of the mask rule happened in the reference run
    assert ((mask == 2) & (run.mask == BIO)).any() and ((mask == BIO) & (run.mask == 2)).any()
    for s in range(4):
        assert rel_l2(c[s], ref[s]) < tol, s
    assert rel_l2(bm.ravel(), run.bmass) < tol
    if prec == 0:
        assert np.array_equal(m_gpu, run.mask)
    else:
        assert np.count_nonzero(m_gpu != run.mask) <= 2  # FP32 biomass at the threshold
