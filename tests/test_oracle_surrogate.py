"""CPU pins of the oracle's surrogate processor (oracle/surrogate.cpp).

The reference holds no golden vectors for run_surrogate3D
(src/complab3d_processors_surrogate.hh:74-295) or for defineSurrogateModel
(surrogateModel.hh:252-500). Every rule is pinned here by a closed form the rule fixes:
  * the network: a one-neuron tanh net written with np.tanh (the oracle uses MATLAB's
    tansig form), an asymmetric two-input layer (a transposed weight matrix fails), the
    mapminmax direction (the training box maps onto [-1, 1]), the 1e-8 / non-finite guards;
  * the processor: with no Vmax the uptake is the supply-limited flux, and the units of
    exhaustionFlux and fluxToDeltaC cancel to dC = -C^2 / (Kc + C) (a slip in either unit
    fails); Vmax-limited uptake in mmol/gDW/h; the 1e4 flux cap; the substrate floor; growth
    per hour and decay per second; the shared-supply clamp; fix_concentration; component
    totals split over the species carrying them; the two closures; the deposit threshold;
  * the step: on a uniform field at rest the interior takes exactly the one-voxel increment.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def const_net(g, n_in=2):
    """A network whose output is the constant growth g (1/h): zero weights, output bias."""
    return dict(n_in=n_in, in_sub=list(range(n_in)), widths=[1], params=[0.0] * n_in + [2.0 * g - 1.0],
                x_offset=[0.0] * n_in, x_gain=[1.0] * n_in, x_ymin=-1.0, y_ymin=-1.0, y_gain=2.0, y_offset=0.0)


def one(S=2, Kc=0.1, decay=1e-6, net=None, **kw):
    return oracle.Surrogate(S, [0], [[Kc] * S], [decay], [net], **kw)


def test_net_one_neuron_closed_form():
    w1, b1, w2, b2 = 1.7, 0.3, -0.9, 0.25
    off, gain, ymin = 0.2, 0.5, -1.0
    net = dict(n_in=1, in_sub=[1], widths=[1, 1], params=[w1, b1, w2, b2], x_offset=[off], x_gain=[gain],
               x_ymin=ymin, y_ymin=-1.0, y_gain=3.0, y_offset=0.05)
    s = one(S=3, net=net)
    for x in (0.0, 0.4, 1.3, 3.7):
        xp = (x - off) * gain + ymin
        want = ((w2 * np.tanh(w1 * xp + b1) + b2) + 1.0) / 3.0 + 0.05
        got = s.growth(0, [99.0, x, -5.0])  # only Fin[1] is an input
        assert got == pytest.approx(want, rel=1e-14)


def test_net_weight_orientation():
    # W[r][c] multiplies input c into neuron r: swapping rows and columns changes the answer
    W = np.array([[0.8, -1.3], [0.1, 0.6]])
    b = np.array([0.2, -0.4])
    v = np.array([1.1, -0.7])
    net = dict(n_in=2, in_sub=[0, 1], widths=[2, 1], params=list(W.ravel()) + list(b) + list(v) + [0.05],
               x_offset=[0.0, 0.0], x_gain=[1.0, 1.0], x_ymin=0.0, y_ymin=0.0, y_gain=1.0, y_offset=0.0)
    s = one(net=net)
    x = np.array([0.9, 0.35])
    want = v @ np.tanh(W @ x + b) + 0.05
    assert s.growth(0, x) == pytest.approx(want, rel=1e-14)
    assert abs(v @ np.tanh(W.T @ x + b) + 0.05 - want) > 1e-3


def test_net_guards():
    assert one(net=const_net(0.5)).growth(0, [1, 1]) == pytest.approx(0.5, rel=1e-15)
    assert one(net=const_net(5e-9)).growth(0, [1, 1]) == 0.0   # below 1e-8: numerical dust
    assert one(net=const_net(-0.3)).growth(0, [1, 1]) == 0.0
    big = const_net(400.0)
    big["log_output"] = True                                    # 10^400 overflows: not finite
    assert one(net=big).growth(0, [1, 1]) == 0.0
    lg = const_net(0.25)
    lg["log_output"] = True
    assert one(net=lg).growth(0, [1, 1]) == pytest.approx(10 ** 0.25, rel=1e-14)
    nan_in = dict(const_net(0.5), params=[1.0, 0.0, 0.0])
    assert one(net=nan_in).growth(0, [np.nan, 1]) == 0.0


def test_mapminmax_direction_on_the_shipped_scaling():
    """The shipped network's scaling states its training box: input k spans
    x_offset_k .. x_offset_k + 2 / x_gain_k (surrogateModel.hh:379-386). A linear read-out
    of the first input, growth = xp + 1, is 0 at the bottom of the box and 2 at the top."""
    with open(os.path.join(HERE, "golden", "surrogate_geobacter.json")) as fh:
        g = json.load(fh)["networks"]["0"]
    lin = dict(n_in=2, in_sub=[0, 1], widths=[1], params=[1.0, 0.0, 0.0], x_offset=g["x_offset"],
               x_gain=g["x_gain"], x_ymin=g["x_ymin"], y_ymin=-1.0, y_gain=1.0, y_offset=0.0)
    s = one(net=lin)
    lo, hi = g["x_offset"][0], g["x_offset"][0] + 2.0 / g["x_gain"][0]
    assert lo == pytest.approx(8.9016e-4, rel=1e-4) and hi == pytest.approx(10.0, rel=1e-3)
    assert s.growth(0, [lo, 0]) == 0.0
    assert s.growth(0, [hi, 0]) == pytest.approx(2.0, rel=1e-14)
    assert s.growth(0, [(lo + hi) / 2, 0]) == pytest.approx(1.0, rel=1e-14)


def test_shipped_network_shape():
    with open(os.path.join(HERE, "golden", "surrogate_geobacter.json")) as fh:
        g = json.load(fh)["networks"]["0"]
    assert g["n_in"] == 2 and g["in_sub"] == [0, 1] and g["widths"] == [10, 10, 10, 10, 1]
    assert len(g["params"]) == 2 * 10 + 10 + 3 * (100 + 10) + 10 + 1
    assert g["y_gain"] == pytest.approx(37.9181676275123, rel=1e-15)
    # inside its box the fitted growth is a finite non-negative rate (1/h)
    s = one(net=g)
    rng = np.random.default_rng(0)
    for _ in range(20):
        x = [rng.uniform(1e-3, 10.0), rng.uniform(4e-5, 0.5)]
        v = s.growth(0, x)
        assert np.isfinite(v) and v >= 0.0


B, MW, DT = 0.8, 24.6, 50.0


def test_supply_limited_uptake_closed_form():
    # no Vmax: Fin = (C * 3.6e6 / (B MW dt)) * C / (Kc + C); the draw Fin B MW dt / 3.6e6 = C^2 / (Kc + C)
    s = one(Kc=0.3, net=const_net(0.0))
    C = np.array([0.5, 0.02])
    dC, dB = s.voxel(C, [B], DT)
    np.testing.assert_allclose(dC, -C * C / (0.3 + C), rtol=1e-14)
    assert dB[0] == pytest.approx(-1e-6 * B * DT, rel=1e-14)  # growth 0: decay, per second


def test_vmax_and_caps():
    C = np.array([0.5, 0.4])
    mm = C / (0.3 + C)
    s = one(Kc=0.3, net=const_net(0.0), vmax=[[2.0, 1e9]])
    dC, _ = s.voxel(C, [B], DT)
    assert dC[0] == pytest.approx(-2.0 * mm[0] * B * MW * DT / 3.6e6, rel=1e-14)  # enzyme-limited
    assert dC[1] == pytest.approx(-C[1], rel=1e-14)  # capped at the supply
    # the 1e4 flux cap: a tiny biomass would need more than 1e4 mmol/gDW/h to empty the voxel
    b = 1e-6
    assert C[0] * 3.6e6 / (b * MW * DT) > 1e4
    dC, _ = one(Kc=0.3, net=const_net(0.0)).voxel(C, [b], DT)
    assert dC[0] == pytest.approx(-1e4 * mm[0] * b * MW * DT / 3.6e6, rel=1e-14)
    # the substrate floor (negative lower bound) limits the uptake magnitude
    s = one(Kc=0.3, net=const_net(0.0), vmax=[[2.0, 2.0]], max_uptake=[[-0.5, -1e30]])
    dC, _ = s.voxel(C, [B], DT)
    assert dC[0] == pytest.approx(-0.5 * B * MW * DT / 3.6e6, rel=1e-14)
    assert dC[1] == pytest.approx(-2.0 * mm[1] * B * MW * DT / 3.6e6, rel=1e-14)
    # molar mass scales the draw
    s = one(Kc=0.3, net=const_net(0.0), vmax=[[2.0, 2.0]], molar_mass=[50.0])
    dC, _ = s.voxel(C, [B], DT)
    assert dC[0] == pytest.approx(-2.0 * mm[0] * B * 50.0 * DT / 3.6e6, rel=1e-14)


def test_growth_and_decay():
    C = [0.5, 0.4]
    _, dB = one(net=const_net(0.7)).voxel(C, [B], DT)
    assert dB[0] == pytest.approx(0.7 * B * DT / 3600.0, rel=1e-14)  # per hour
    _, dB = one(net=None, decay=2e-4).voxel(C, [B], DT)               # no network: decay
    assert dB[0] == pytest.approx(-2e-4 * B * DT, rel=1e-14)
    _, dB = one(net=None, decay=1.0).voxel(C, [B], DT)                # never below -B
    assert dB[0] == -B
    dC, dB = one(net=const_net(0.7)).voxel(C, [1e-12], DT)             # absent: nothing
    assert not dC.any() and not dB.any()


def test_shared_supply_clamped_once():
    # two organisms each able to take the whole supply (Kc = 0, no Vmax): summed, clamped to -C
    s = oracle.Surrogate(2, [0, 1], [[0.0, 0.0], [0.0, 0.0]], [0.0, 0.0], [const_net(0.1), const_net(0.2)])
    C = np.array([0.5, 0.25])
    dC, dB = s.voxel(C, [B, 0.3], DT)
    np.testing.assert_allclose(dC, -C, rtol=1e-15)
    np.testing.assert_allclose(dB, [0.1 * B * DT / 3600, 0.2 * 0.3 * DT / 3600], rtol=1e-14)
    # one present, one absent: the absent one neither draws nor grows
    dC, dB = s.voxel(C, [B, 0.0], DT)
    np.testing.assert_allclose(dC, -C, rtol=1e-15)
    assert dB[1] == 0.0


def test_fix_concentration_and_deposit_threshold():
    C = np.array([0.5, 0.4])
    dC, _ = one(Kc=0.3, net=const_net(0.0), fix_conc=[0, 1]).voxel(C, [B], DT)
    assert dC[0] == pytest.approx(-C[0] ** 2 / 0.8, rel=1e-14) and dC[1] == 0.0
    # a draw of 1e-13 is below the deposit threshold
    s = one(Kc=0.3, net=const_net(0.0), vmax=[[1e-13 * 3.6e6 / (B * MW * DT) / (0.5 / 0.8), 2.0]])
    dC, _ = s.voxel(C, [B], DT)
    assert dC[0] == 0.0 and dC[1] != 0.0


def test_component_totals_split():
    # species 0 and 1 are components A and B, species 2 the complex AB
    Sm = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    K = 0.2
    s = oracle.Surrogate(3, [0], [[K] * 3], [0.0], [const_net(0.0, n_in=3)], use_totals=True,
                         tableau=(Sm, [0, 1]))
    C = np.array([0.3, 0.1, 0.05])
    dC, _ = s.voxel(C, [B], DT)
    T0, T1 = C[0] + C[2], C[1] + C[2]
    m0, m1 = T0 / (K + T0), T1 / (K + T1)
    # component A: amount T0 m0 taken from A and AB in proportion to the A they hold
    want0 = -C[0] * m0
    want2 = -C[2] * m0
    # component B: amount T1 m1 from B and AB; AB only has what A's draw left
    want1 = -C[1] * m1
    want2 -= min(C[2] * m1, C[2] + want2)
    np.testing.assert_allclose(dC, [want0, want1, want2], rtol=1e-13)
    # without totals each species is its own substrate
    s = oracle.Surrogate(3, [0], [[K] * 3], [0.0], [const_net(0.0, n_in=3)])
    dC, _ = s.voxel(C, [B], DT)
    np.testing.assert_allclose(dC, -C * C / (K + C), rtol=1e-14)


def test_closures():
    C = np.array([0.5, 0.4])
    K = 0.3
    fin = (C * 3.6e6 / (B * MW * DT)) * C / (K + C)  # supply-limited estimates
    Y = np.array([0.04, 0.0])
    g = 0.7
    # closure 1: consume g / Y, at most the estimate; other substrates at the estimate
    n1 = dict(const_net(g), closure=1)
    n1["yield"] = list(Y)
    dC, dB = one(Kc=K, net=n1).voxel(C, [B], DT)
    need = min(g / Y[0], fin[0])
    assert dC[0] == pytest.approx(-need * B * MW * DT / 3.6e6, rel=1e-13)
    assert dC[1] == pytest.approx(-C[1] ** 2 / (K + C[1]), rel=1e-13)
    assert dB[0] == pytest.approx(g * B * DT / 3600, rel=1e-14)
    # closure 2: growth limited by the scarcest substrate (a large yield makes it g itself)
    n2 = dict(const_net(g), closure=2)
    n2["yield"] = [1e-9, 0.0]
    dC, dB = one(Kc=K, net=n2).voxel(C, [B], DT)
    gl = min(g, fin[0] * 1e-9)
    assert dB[0] == pytest.approx(gl * B * DT / 3600, rel=1e-13)
    assert dC[0] == pytest.approx(-(gl / 1e-9) * B * MW * DT / 3.6e6, rel=1e-13)


def test_uniform_field_step():
    """One coupled step on a uniform field at rest (no flow, Neumann faces): every interior
    voxel reacts alike, so the voxels whose neighbours all reacted (2 <= x <= nx-3) hold
    C + dC and every reactive voxel's biomass B + dB."""
    nx, ny, nz = 10, 4, 4
    N = nx * ny * nz
    mask = np.full(N, 2, dtype=np.uint8)
    kind = np.full(256, 2, dtype=np.int32)
    kind[0], kind[1] = 0, 1
    om = np.full((2, 256), 1.0 / 0.8)
    c0 = [0.5, 0.2]
    b = np.full(N, B)
    net = const_net(0.4)
    srg = oracle.Surrogate(2, [0], [[0.3, 0.3]], [1e-6], [net])
    run = oracle.AdeRun(nx, ny, nz, mask, kind, om, [1, 1, 1, 1], [0, 0, 0, 0], np.zeros((N, 3)), c0,
                        dt=DT, biomass=b, surrogate=srg)
    run.run(1)
    dC, dB = srg.voxel(c0, [B], DT)
    c = run.conc().reshape(2, nz, ny, nx)
    for s in range(2):
        np.testing.assert_allclose(c[s, :, :, 2:nx - 2], c0[s] + dC[s], rtol=1e-13)
    bm = run.bmass.reshape(nz, ny, nx)
    np.testing.assert_allclose(bm[:, :, 1:nx - 1], B + dB[0], rtol=1e-14)
    np.testing.assert_allclose(bm[:, :, [0, nx - 1]], B, rtol=0)
