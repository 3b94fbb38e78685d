"""The pore <-> biofilm rule with several material types on the GPU (clb_set_mask_codes;
updateLocalMaskNtotalLattices3D, src/complab3d_processors_part2.hh:117-198), through the C
ABI: the oracle's hand-built fixture (tests/test_oracle_mask_rule.py) in the step, on both
layouts, and in the CA's mask / total pass; then a coupled run with kinetics against the
oracle, mask bit-exact and concentrations to 1e-10 (FP64)."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2
from test_oracle_mask_rule import NX, NY, NZ, THR, fixture

FLUID = (2, 3, 4, 5, 7, 9)


def lattice(clb, m, b, prec, layout, pores, bios):
    kind, om = d3q19_kind_omega(codes_fluid=FLUID)
    lat = clb.Lattice(m.size // (NY * NZ), NY, NZ, precision=prec)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(m)
    lat.set_solutes(np.full((1, 256), 1.0 / 0.8), [1, 1], [0.0, 0.0])
    lat.set_velocity(np.zeros(3 * m.size))
    lat.set_biomass(b)
    lat.set_mask_update(True, THR, pores[0], bios[0])
    lat.set_mask_codes(pores, bios)
    lat.solutes_init([0.0])
    return lat


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_step_rule_on_the_fixture(clb, prec, layout):
    print(f"case: multi-type mask rule fixture in the step, prec={prec} layout={layout}", flush=True)
    m, b, expect = fixture()
    lat = lattice(clb, m, b, prec, layout, [2, 5], [3, 4])
    lat.step(1)
    got = lat.download_mask()
    lat.close()
    assert np.array_equal(got, expect), np.flatnonzero(got != expect)


@pytest.mark.gpu
def test_ca_pass_on_the_fixture(clb):
    print("case: multi-type mask rule fixture in the CA pass", flush=True)
    m, b, expect = fixture()
    lat = lattice(clb, m, b, 0, 1, [2, 5], [3, 4])
    lat.ca_init()
    lat.ca_redistribute(1e9)
    got = lat.download_mask()
    lat.close()
    assert np.array_equal(got, expect), np.flatnonzero(got != expect)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [0, 1])
def test_coupled_run_against_oracle(clb, oracle, layout):
    print(f"case: multi-type mask rule, coupled kinetics 32x16x16, layout={layout}", flush=True)
    nx, ny, nz = 32, 16, 16
    pack = clb.gen_sphere_pack(nx, ny, nz, 3.0, 5.0, 0.5, seed=0)
    rng = np.random.default_rng(21)
    m = pack.copy()
    pore = np.flatnonzero(pack == 2)
    m[rng.choice(pore, pore.size // 3, replace=False)] = 5            # a second pore type
    b = np.zeros((2, m.size))
    seeds = rng.choice(pore, pore.size // 10, replace=False)
    b[0, seeds] = rng.uniform(5.0, 60.0, seeds.size)
    b[1, seeds[::2]] = rng.uniform(1.0, 30.0, seeds[::2].size)
    labels = rng.choice(pore, pore.size // 10, replace=False)          # biofilm voxels, most below
    m[labels] = rng.choice([3, 4, 7], labels.size)                     # the threshold: they reopen
    kind, om = d3q19_kind_omega(0.8, codes_fluid=FLUID)
    f, _ = oracle.flow_run(nx, ny, nz, pack, kind, om, 1.0, 0.999, 200)
    u, _ = oracle.flow_velocity(nx, ny, nz, pack, kind, 1.0, 0.999, f)
    bct, bcv, c0 = [0, 1, 0, 1, 1, 1, 1, 1], [1.0, 0, 0.5, 0, 0, 0, 0, 0], [0.2, 0.1, 0.0, 0.0]
    rule = oracle.mask_rule([2, 5], [3, 4])
    run = oracle.AdeRun(nx, ny, nz, m, kind, np.full((4, 256), 1.0 / 0.8), bct, bcv, u, c0, kinetics=True,
                        biomass=b.ravel(), mask_update=(40.0, 2, 3), rule=rule)
    m0 = m.copy()
    run.run(60)
    lat = clb.Lattice(nx, ny, nz, precision=0)
    lat.set_layout(layout)
    lat.set_dynamics(kind, om)
    lat.upload_mask(m)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), bct, bcv)
    lat.set_velocity(u)
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 40.0, 2, 3)
    lat.set_mask_codes([2, 5], [3, 4])
    lat.solutes_init(c0)
    lat.step(60)
    got = lat.download_mask()
    c = lat.conc()
    lat.close()
    # the rule did its work: voxels reopened into both pore types and seeds became biofilm
    assert np.any((m0 != 2) & (m0 != 5) & (run.mask == 5)) and np.any((m0 == 5) & (run.mask >= 3) & (run.mask != 5))
    assert np.array_equal(got, run.mask), np.flatnonzero(got != run.mask)[:10]
    ref = run.conc()
    for s in range(4):
        assert rel_l2(c[s], ref[s]) < 1e-10, s
