"""The compact (flow-fluid only) D3Q19 layout against the dense one.

The compact flow kernel reads and writes the same populations as the dense kernel, with the
same arithmetic, only at other addresses (the flow-fluid voxels packed in brick order, a ghost
slot standing for each wall slot a bounce-back link uses). In FP64 velocities and densities
after any number of steps are identical bit for bit, whatever the step parity, downloads in
between (which bring the dense array up to date and leave the compact state in place), and
mask changes: precipitation (a table node turned wall, kept in place), reopening (the dense
array written, the table rebuilt) and a biofilm rule whose biofilm code is a wall for the
flow. In FP32 the compiler still rounds a few x-face nodes' arithmetic differently in the two
kernels (the first difference, in the last bit of some inlet nodes, after 4 steps), so there
the layouts agree to the FP32 store's rounding: relative L2 1e-6 on the velocity, 1e-7 on the
density (the FP32 store itself is held to 1e-5 of the FP64 oracle). The energy tally sums the
same per-node terms in another order, so it agrees to rounding. Checkpoints and the autotuner
leave the compact state as an uninterrupted run's, bit for bit. (The face regularization keeps
its multiply-adds unfused and the collision is one explicit fused multiply-add per
population: in FP64 the compiler had fused them differently in the two kernels, and the first
face node to differ did so in its last bit after 76 steps on the 48 x 24 x 24 pack of seed 2.)"""
import os

import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2

NX, NY, NZ = 64, 32, 32


def same(clb, prec, a, b, rho=False):
    """Bit for bit in FP64; to the FP32 store's rounding in FP32."""
    if prec == clb.PREC_FP64:
        assert np.array_equal(a, b)
    elif rho:
        assert np.max(np.abs(np.asarray(a) - np.asarray(b))) < 1e-7
    else:
        assert rel_l2(b, a) < 1e-6


def flow_only(clb, layout, prec, segments):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    out = []
    for n in segments:
        e = lat.flow_steps(n, energy=True)
        u, rho = lat.velocity()
        out.append((e, u, rho, lat.flow_stats()))
    lat.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_flow_compact_equals_dense(clb, prec):
    print(f"case: flow compact vs dense 64x32x32 prec={prec}", flush=True)
    segs = [301, 200, 1]  # odd counts: each download sits mid-pair of the AA pattern
    dense = flow_only(clb, clb.LAYOUT_DENSE, prec, segs)
    comp = flow_only(clb, clb.LAYOUT_COMPACT, prec, segs)
    for (ed, ud, rd, sd), (ec, uc, rc, sc) in zip(dense, comp):
        same(clb, prec, ud, uc)
        same(clb, prec, rd, rc, rho=True)
        assert np.allclose(ec, ed, rtol=1e-12 if prec == clb.PREC_FP64 else 1e-5, atol=0.0)
        assert np.allclose(sc, sd, rtol=1e-12 if prec == clb.PREC_FP64 else 1e-5, atol=0.0)
    assert np.abs(comp[-1][1]).max() > 0.0  # the flow moved


def conversion_run(clb, layout, prec):
    nx, ny, nz = 48, 24, 24
    N = nx * ny * nz
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, seed=1)
    kind, om = d3q19_kind_omega(0.8)
    lat = clb.Lattice(nx, ny, nz, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(150)
    lat.set_solutes(np.full((2, 256), 1.0 / 0.8), [0, 1, 0, 1], [1.0, 0.0, 0.5, 0.0])
    lat.couple_velocity()
    lat.solutes_init([0.2, 0.1])
    lat.set_phase(np.zeros(N, dtype=np.uint8))
    x = np.arange(N) % nx
    inner = (pack >= 2) & (x >= 1) & (x <= nx - 2)
    out = []
    for k in range(4):
        lat.step(20)
        c = lat.conc()[1]
        thr = float(np.quantile(c[inner], 0.9))
        # precipitation (fluid -> wall) and reopening (wall -> fluid: every phase voxel, whatever
        # its walls hold) in turn
        ch = lat.node_conversion(1, 2, 1, 9, thr if k % 2 == 0 else -1.0, 1e3 * thr if k % 2 else -1.0)
        e = lat.flow_steps(31, energy=True)
        lat.couple_velocity()
        out.append((ch, e))
    lat.step(11)
    u, rho = lat.velocity()
    res = (out, u, rho, lat.conc(), lat.download_mask())
    lat.close()
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0])  # FP64: bit for bit (an FP32 threshold could fall between the layouts)
def test_flow_compact_through_node_conversion(clb, prec):
    print(f"case: flow compact vs dense through precipitation and reopening prec={prec}", flush=True)
    dense = conversion_run(clb, clb.LAYOUT_DENSE, prec)
    comp = conversion_run(clb, clb.LAYOUT_COMPACT, prec)
    changes = [ch for ch, _ in dense[0]]
    assert changes == [ch for ch, _ in comp[0]]
    assert changes[0] > 0 and changes[1] > 0  # both kinds of conversion happened
    for (_, ed), (_, ec) in zip(dense[0], comp[0]):
        assert np.allclose(ec, ed, rtol=1e-12 if prec == clb.PREC_FP64 else 1e-5, atol=0.0)
    u, rho, conc, mask = zip(dense[1:], comp[1:])
    same(clb, prec, *u)
    same(clb, prec, *rho, rho=True)
    for a, b in zip(*conc):
        same(clb, prec, a, b)
    assert np.array_equal(*mask)


def biofilm_run(clb, layout, prec):
    nx, ny, nz = 48, 24, 24
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.45, seed=2)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.03, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    kind, om = d3q19_kind_omega(0.8)
    kind[3] = clb.KIND_BB  # an impermeable biofilm: a wall for the flow ...
    akind = kind.copy()
    akind[3] = clb.KIND_FLUID  # ... and advection-diffusion for the solutes
    lat = clb.Lattice(nx, ny, nz, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.set_solute_dynamics(akind)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(120)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1, 0, 1, 1, 1, 1, 1], [1.0, 0, 0.5, 0, 0, 0, 0, 0])
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 30.0, 2, 3)
    lat.solutes_init([0.2, 0.1, 0.05, 0.0])
    masks = []
    for _ in range(3):
        lat.step(25)  # pore <-> biofilm flips change which voxels are walls for the flow
        masks.append(lat.download_mask())
        lat.flow_steps(17)
        lat.couple_velocity()
    u, rho = lat.velocity()
    res = (u, rho, lat.conc(), lat.biomass(), np.stack(masks))
    lat.close()
    return res


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0])  # FP64: bit for bit (an FP32 threshold could fall between the layouts)
def test_flow_compact_with_impermeable_biofilm(clb, prec):
    print(f"case: flow compact vs dense, biofilm a flow wall with the mask rule prec={prec}", flush=True)
    dense = biofilm_run(clb, clb.LAYOUT_DENSE, prec)
    comp = biofilm_run(clb, clb.LAYOUT_COMPACT, prec)
    same(clb, prec, dense[0], comp[0])
    same(clb, prec, dense[1], comp[1], rho=True)
    for a, b in zip(dense[2], comp[2]):
        same(clb, prec, a, b)
    same(clb, prec, dense[3], comp[3])
    assert np.array_equal(dense[4], comp[4])
    masks = dense[-1]
    assert not np.array_equal(masks[0], masks[-1]) or (masks[0] == 3).sum() > 0  # biofilm present


@pytest.mark.gpu
def test_flow_compact_checkpoint_and_autotune(clb, tmp_path):
    print("case: flow compact checkpoint round trip and autotune", flush=True)
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)

    def make():
        lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP32)
        lat.set_dynamics(kind, om)
        lat.upload_mask(pack)
        lat.flow_init(1.0, 0.999)
        return lat

    ref = make()
    ref.flow_steps(301)
    ref.flow_steps(200)
    u_ref, r_ref = ref.velocity()
    ref.close()

    lat = make()
    lat.flow_steps(301)
    path = os.path.join(str(tmp_path), "flow.chk")
    lat.save_state(path)
    rep = lat.autotune(10)  # compact flow blocks tried; the state comes back unchanged
    assert rep["flow_by"] == 1 and rep["flow_bx"] in (64, 128, 256)
    lat.flow_steps(200)
    u1, r1 = lat.velocity()
    lat.close()
    assert np.array_equal(u1, u_ref) and np.array_equal(r1, r_ref)

    lat = make()
    lat.flow_steps(4)  # a different state, replaced by the checkpoint
    lat.load_state(path)
    lat.flow_steps(200)
    u2, r2 = lat.velocity()
    lat.close()
    assert np.array_equal(u2, u_ref) and np.array_equal(r2, r_ref)
