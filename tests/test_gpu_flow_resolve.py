"""Flow solves to convergence as the reference's driver runs them, against the FP64 oracle.

The driver steps the flow until util::ValueTracer(1, 1000, epsilon) has converged
(src/complab.cpp:607-620), and after a geometry change re-solves with a tracer that keeps its
window from the previous solve (ns_convg2, declared once at src/complab.cpp:1149 and never
reset; re-solves at :1794-1800, 1836-1840). The GPU's device-reduced energies drive
clb.Lattice.flow_resolve; the oracle's energies, fed to the same tracer, must stop at the same
step, and the velocities there agree to FP64 rounding. A re-solve that uses all its steps is
followed by the percolation test (mean u_x over the plane x = nx-2 above 1e-12,
src/complab.cpp:1801-1805)."""
import copy

import numpy as np
import pytest

from conftest import d3q19_kind_omega, rel_l2

NX, NY, NZ = 64, 32, 32
EPS = 1e-4  # ns_converge_iT2 of the reference's examples (examples/*/CompLaB.xml)


def stop_step(tracer, energies):
    for k, v in enumerate(energies):
        tracer.take_value(float(v))
        if tracer.has_converged():
            return k + 1
    return -1


@pytest.mark.gpu
def test_solve_and_resolve_stop_where_the_oracle_does(clb, oracle):
    print("case: flow solve + re-solve with a persistent tracer, 64x32x32 FP64", flush=True)
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    # the oracle: energies of the first solve
    f, e = oracle.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, 1600, energy=True)
    t_ref = clb.ValueTracer(1.0, 1000.0, EPS)
    k1 = stop_step(t_ref, e)
    assert 500 <= k1 < 1600
    f1, _ = oracle.flow_run(NX, NY, NZ, pack, kind, om, 1.0, 0.999, k1)
    u1, _ = oracle.flow_velocity(NX, NY, NZ, pack, kind, 1.0, 0.999, f1)
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP64)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    tracer = clb.ValueTracer(1.0, 1000.0, EPS)
    steps, conv = lat.flow_resolve(tracer, 5000)
    assert conv and steps == k1
    u, _ = lat.velocity()
    assert rel_l2(u, u1) < 1e-10
    # geometry changes: interior pore voxels become bounce-back (precipitation). The tracer
    # keeps its window, so the re-solve may end after a single step
    rng = np.random.default_rng(3)
    x = np.arange(pack.size) % NX
    cand = np.flatnonzero((pack == 2) & (x >= 20) & (x <= 40))
    pick = rng.choice(cand, 15, replace=False)
    f_prev, t_prev = f1, t_ref
    for nv, limit in ((5, 2000), (15, 300)):
        mask2 = pack.copy()
        mask2[pick[:nv]] = 1
        lat.upload_mask(mask2)
        _, e2 = oracle.flow_run(NX, NY, NZ, mask2, kind, om, 1.0, 0.999, min(limit, 400), energy=True,
                                f=f_prev.copy())
        k2 = stop_step(copy.deepcopy(t_prev), e2)
        expect = k2 if k2 > 0 else limit
        steps2, conv2 = lat.flow_resolve(tracer, limit)
        assert conv2 == (k2 > 0) and steps2 == expect, (nv, k2, steps2)
        f2, _ = oracle.flow_run(NX, NY, NZ, mask2, kind, om, 1.0, 0.999, steps2, f=f_prev.copy())
        u2, _ = oracle.flow_velocity(NX, NY, NZ, mask2, kind, 1.0, 0.999, f2)
        ug, _ = lat.velocity()
        assert rel_l2(ug, u2) < 1e-10
        # the oracle-side tracer after exactly the steps the GPU ran
        for v in e2[:steps2]:
            t_prev.take_value(float(v))
        f_prev = f2
    lat.close()


@pytest.mark.gpu
def test_clogged_outlet_does_not_percolate(clb):
    print("case: flow re-solve against a clogged plane, percolation test", flush=True)
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    m = pack.reshape(NZ, NY, NX).copy()
    m[:, :, NX - 2] = 1                     # the outlet-reading plane sealed
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP64)
    lat.set_dynamics(kind, om)
    lat.upload_mask(m.ravel())
    lat.flow_init(1.0, 0.999)
    steps, conv = lat.flow_resolve(clb.ValueTracer(1.0, 1000.0, EPS), 300)
    assert steps == 300 and not conv        # the window (500) never filled
    assert not lat.percolates()
    lat.close()
    # the same step limit on the open pack: not converged either, but the flow percolates
    lat = clb.Lattice(NX, NY, NZ, precision=clb.PREC_FP64)
    lat.set_dynamics(kind, om)
    lat.upload_mask(pack)
    lat.flow_init(1.0, 0.999)
    steps, conv = lat.flow_resolve(clb.ValueTracer(1.0, 1000.0, EPS), 300)
    assert steps == 300 and not conv
    assert lat.percolates()
    lat.close()


def _resolve_pair(clb, prec):
    """Two identical lattices: one solved with the host tracer loop, one with the device
    tracer."""
    pack = clb.gen_sphere_pack(NX, NY, NZ, 3.0, 5.0, 0.45, seed=0)
    kind, om = d3q19_kind_omega(0.8)
    lats = []
    for _ in range(2):
        lat = clb.Lattice(NX, NY, NZ, precision=prec)
        lat.set_dynamics(kind, om)
        lat.upload_mask(pack)
        lat.flow_init(1.0, 0.999)
        lats.append(lat)
    return pack, lats


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_device_tracer_stops_where_the_host_tracer_does(clb, prec):
    """clb_flow_resolve runs the driver's loop with the tracer on the device: the same energies,
    the same stopping step and the same velocities as the host loop (flow_resolve with
    clb.ValueTracer), for the initial solve, a capped solve, re-solves whose window persists
    (ending after one step), and a fresh tracer."""
    print(f"case: device tracer vs host tracer, 64x32x32 prec={prec}", flush=True)
    pack, (host, dev) = _resolve_pair(clb, prec)
    tracer = clb.ValueTracer(1.0, 1000.0, EPS)
    dev.flow_tracer(1.0, 1000.0, EPS)
    e_host = []
    for max_steps in (700, 5000, 3, 1, 50):  # capped, to convergence, then re-solves
        t0 = len(tracer.values)
        n_h, c_h = host.flow_resolve(tracer, max_steps)
        n_d, c_d, e_d = dev.flow_resolve_device(max_steps, energy=True)
        assert (n_d, c_d) == (n_h, c_h)
        assert n_d >= 1
        if max_steps == 5000:
            assert c_d and n_d < 5000
        assert np.array_equal(host.velocity()[0], dev.velocity()[0])
        e_host.append(e_d)
    # a re-solve with a full window that still meets the threshold ends after one step
    assert len(e_host[2]) == 1 and len(e_host[3]) == 1
    # a fresh tracer: an empty window again
    tracer = clb.ValueTracer(1.0, 1000.0, 1e-6)
    dev.flow_tracer(1.0, 1000.0, 1e-6)
    n_h, c_h = host.flow_resolve(tracer, 800)
    n_d, c_d = dev.flow_resolve_device(800)
    assert (n_d, c_d) == (n_h, c_h) and n_d >= 500
    assert np.array_equal(host.velocity()[0], dev.velocity()[0])
    with pytest.raises(clb.ClbError):
        dev.flow_tracer(1.0, 2.0, 1e-4)  # a window of one value
    host.close()
    dev.close()
