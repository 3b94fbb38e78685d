"""Pins the oracle's kinetics processors (run_kinetics + update_rxnLattices,
src/complab3d_processors_part1.hh:72-127, 345-415) and its MLP restatement.

Kinetics: the dual-Monod plugin against its closed form, the dC * t_i injection, and the
reaction's mass balance (donor + product and biomass - yield * product conserved, as the
reference's examples 03 / 13-15 check by hand). MLP: the FP64 gradient against central
finite differences (surrogate_training/trainSurrogate.py:132-166 uses the same tanh /
1 - a^2 back-propagation)."""
import numpy as np

from conftest import d3q19_kind_omega


def test_dual_monod_plugin_closed_form(oracle):
    B, D, A = 7.0, 0.2, 0.05
    sr, br = oracle.kinetics_plugin([B], [D, A, 0.0, 0.0], 2)
    r = 1e-3 * B * D / (1e-2 + D) * A / (1e-2 + A)
    assert r < 0.5 * min(D, A)
    assert np.allclose(sr, [-r, -r, r, r], rtol=1e-15, atol=0)
    assert np.allclose(br, [0.4 * r], rtol=1e-15, atol=0)
    # per-step stability cap: no reactant loses more than half of what it holds
    B, D, A = 90.0, 0.2, 0.01
    sr, br = oracle.kinetics_plugin([B], [D, A, 0.0, 0.0], 2)
    assert np.allclose(sr, [-0.005, -0.005, 0.005, 0.005], rtol=1e-15, atol=0)
    assert np.allclose(br, [0.4 * 0.005], rtol=1e-15, atol=0)
    # no reactions in solid / bounce-back voxels (mask < 2)
    sr, br = oracle.kinetics_plugin([B], [D, A, 0.0, 0.0], 1)
    assert not sr.any() and not br.any()


def test_kinetics_mass_balance_closed_box(oracle):
    nx, ny, nz = 10, 6, 6
    rng = np.random.default_rng(4)
    m = np.full((nz, ny, nx), 2, dtype=np.uint8)
    m[:, :, 0] = 1
    m[:, :, nx - 1] = 1
    m = m.ravel()
    kind, _ = d3q19_kind_omega()
    om = np.full((4, 256), 1.0 / 0.8)
    b = np.where(rng.random(m.size) < 0.3, rng.uniform(10, 100, m.size), 0.0)
    b[m != 2] = 0.0
    run = oracle.AdeRun(nx, ny, nz, m, kind, om, [1] * 8, [0.0] * 8, np.zeros((m.size, 3)),
                        [0.5, 0.3, 0.0, 0.0], kinetics=True, dt=1.0, biomass=b)
    tot = lambda: run.g.reshape(4, 7, -1).sum(axis=(1, 2))
    t0 = tot()
    b0 = run.bmass.sum()
    run.run(200)
    t1 = tot()
    d = t1 - t0
    assert d[0] < -1e-3                                  # donor consumed
    assert abs(d[0] - d[1]) < 1e-11                      # acceptor consumed at the same rate
    assert abs(d[0] + d[2]) < 1e-11 and abs(d[2] - d[3]) < 1e-11
    assert abs((run.bmass.sum() - b0) - 0.4 * d[2]) < 1e-9


def test_mlp_gradient_finite_differences(oracle):
    din, H, dout, B = 5, 7, 3, 9
    rng = np.random.default_rng(8)
    P = H * din + H + H * H + H + dout * H + dout
    theta = 0.5 * rng.standard_normal(P)
    X = rng.uniform(-1, 1, (B, din))
    T = rng.uniform(-1, 1, (B, dout))
    g, loss = oracle.mlp_grad(din, H, dout, theta, X, T)
    Y = oracle.mlp_forward(din, H, dout, theta, X)
    assert abs(loss - np.mean((Y - T) ** 2)) < 1e-14
    h = 1e-6
    for k in rng.choice(P, 40, replace=False):
        tp = theta.copy(); tp[k] += h
        tm = theta.copy(); tm[k] -= h
        fd = (oracle.mlp_grad(din, H, dout, tp, X, T)[1] - oracle.mlp_grad(din, H, dout, tm, X, T)[1]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(fd))


def test_abiotic_plugin_closed_form(oracle):
    C = [0.2, 0.05, 0.03, 0.0]
    sr = oracle.abiotic_plugin(C, 2)
    R = 5e-2 * 0.03 * 0.05
    assert np.allclose(sr, [0.0, -R, -R, R], rtol=1e-15, atol=0)
    assert not oracle.abiotic_plugin(C, 1).any()           # walls do not react


def test_abiotic_mass_balance_and_immobile(oracle):
    # closed box: P1 + A -> P2 keeps P1 + P2 and A + P2 constant; P2 immobile stays where
    # it formed (its total changes only by the reaction)
    nx, ny, nz = 10, 6, 6
    m = np.full((nz, ny, nx), 2, dtype=np.uint8)
    m[:, :, 0] = 1
    m[:, :, nx - 1] = 1
    m = m.ravel()
    kind, _ = d3q19_kind_omega()
    om = np.full((4, 256), 1.0 / 0.8)
    run = oracle.AdeRun(nx, ny, nz, m, kind, om, [1] * 8, [0.0] * 8, np.zeros((m.size, 3)), [0.0, 0.3, 0.2, 0.0],
                        abiotic=1, immobile=[0, 0, 0, 1])
    tot = lambda: run.g.reshape(4, 7, -1).sum(axis=(1, 2))
    t0 = tot()
    run.run(100)
    d = tot() - t0
    assert d[3] > 1e-3
    assert abs(d[2] + d[3]) < 1e-11 and abs(d[1] + d[3]) < 1e-11 and abs(d[0]) < 1e-12
    # the immobile product has not spread into the wall nodes
    assert np.all(run.conc()[3][m != 2] == 0.0)


def test_node_conversion_rules(oracle):
    # precipNodeConversion3D (src/precipitationVOP.hh:176-204) converts every voxel that is
    # neither solid (0) nor bounce-back (1) once its mineral reaches the threshold;
    # dissolNodeConversion3D (src/dissolutionVOP.hh:469-511) reopens every voxel stamped
    # with the phase whose mineral falls below the reopen threshold, whatever its mask
    mask = np.array([0, 1, 2, 3, 2, 3, 1, 0, 2], dtype=np.uint8)
    phase = np.array([0, 0, 0, 0, 0, 0, 7, 7, 7], dtype=np.uint8)
    mineral = np.array([5.0, 5.0, 5.0, 5.0, 0.5, 0.5, 0.1, 2.0, 0.1])
    m, ph, n = oracle.node_conversion(mask, phase, mineral, 2, 0, 1, 1, 7, 1.0, -1.0)
    assert m.tolist() == [0, 1, 1, 1, 2, 3, 1, 0, 2] and ph.tolist() == [0, 0, 7, 7, 0, 0, 7, 7, 7] and n == 2
    m, ph, n = oracle.node_conversion(m, ph, mineral, 2, 0, 1, 1, 7, -1.0, 1.0)
    # the bb voxel (0.1) and the pore voxel already open (0.1) reopen; only the first
    # changes its mask code; voxels at 5.0 and 2.0 stay
    assert m.tolist() == [0, 1, 1, 1, 2, 3, 2, 0, 2] and ph.tolist() == [0, 0, 7, 7, 0, 0, 0, 7, 0] and n == 1
    # thresholds <= 0 switch a direction off
    m2, ph2, n = oracle.node_conversion(m, ph, mineral, 2, 0, 1, 1, 7, -1.0, -1.0)
    assert n == 0 and np.array_equal(m2, m) and np.array_equal(ph2, ph)
