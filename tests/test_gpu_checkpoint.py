"""Checkpoint / restart (clb_save_state / clb_load_state: the evolving state of the
reference's saveBinaryBlock / loadBinaryBlock checkpoints, src/complab.cpp:1122-1132,
1238-1247, in one file of this library's format).

A run is saved mid-way and continued; reloading the file, into the same context or into a
fresh one set up the same way, and continuing the same number of steps must give the same
fields bit for bit: populations with the fused step's deferred increments folded in,
biomass, mask, flow, CA age, and the step parity (the AA layout alternates)."""
import numpy as np
import pytest

from conftest import d3q19_kind_omega
from test_oracle_speciation import CARB_COMP, CARB_LOGK, CARB_S, carbonate_exact

NX, NY, NZ = 32, 16, 12


def build(clb, prec):
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.5, seed=0)
    b = clb.gen_biomass_seed(pack, pore=2, fraction=0.05, bmin=10, bmax=100, seed_sites=1, seed_density=2)
    kind, om = d3q19_kind_omega(0.8)
    c_in, c_0 = carbonate_exact(2e-3, 4e-4), carbonate_exact(1e-3, 1e-4)
    lat = clb.Lattice(NX, NY, NZ, precision=prec)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    lat.flow_steps(50)
    lat.set_solutes(np.full((4, 256), 1.0 / 0.8), [0, 1] * 4, [c_in[0], 0, c_in[1], 0, c_in[2], 0, c_in[3], 0])
    lat.couple_velocity()
    lat.set_biomass(b)
    lat.set_kinetics(clb.KINETICS_PLUGIN, 1.0, 0, 1)
    lat.set_mask_update(True, 10.0, 2, 3)
    lat.set_tableau(CARB_S, CARB_LOGK, CARB_COMP, tol=1e-10, max_iter=200)
    lat.solutes_init(c_0)
    lat.set_phase(np.zeros(pack.size, dtype=np.uint8))
    lat.ca_init()
    return lat


def advance(lat, steps):
    lat.step(steps)
    lat.flow_steps(7)  # odd: the flow parity matters too
    rep = lat.ca_redistribute(60.0, seed=5)
    u, rho = lat.velocity()
    return lat.conc(), lat.biomass(), lat.download_mask(), u, rho, lat.ca_age(), rep["sweeps"]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [1, 0])
def test_checkpoint_restart_bit_exact(clb, tmp_path, prec):
    print(f"case: checkpoint restart, precision {prec}", flush=True)
    path = str(tmp_path / "state.clb")
    a = build(clb, prec)
    a.step(31)  # odd step count: the next step has odd parity, with deferred increments pending
    a.save_state(path)
    x = advance(a, 20)
    a.load_state(path)
    y = advance(a, 20)
    b = build(clb, prec)
    b.load_state(path)
    z = advance(b, 20)
    for p, q, r in zip(x, y, z):
        assert np.array_equal(p, q) and np.array_equal(p, r)
    assert x[6] > 0  # the CA had work to do
    a.close()
    b.close()


@pytest.mark.gpu
def test_checkpoint_rejects_other_setup(clb, tmp_path):
    print("case: checkpoint into a differently set-up context", flush=True)
    path = str(tmp_path / "state.clb")
    a = build(clb, 1)
    a.save_state(path)
    a.close()
    b = clb.Lattice(NX, NY, NZ, precision=1)
    with pytest.raises(clb.ClbError):
        b.load_state(path)
    b.close()
    with open(path, "r+b") as f:
        f.write(b"NOTASTAT")
    c = build(clb, 1)
    with pytest.raises(clb.ClbError):
        c.load_state(path)
    c.close()
