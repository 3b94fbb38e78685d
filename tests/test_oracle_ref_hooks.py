"""The oracle built with the reference's own example kinetics hooks
(examples/05_biotic_cellular_automaton/defineKinetics.hh:53-86, Monod growth with decay, and
examples/03_abiotic_kinetics/defineAbioticKinetics.hh, A + B -> C at k_AB A B), compiled
where they lie by `make -C oracle ref-hooks` (oracle/_ref/liboracle_ref_hooks.so). Pinned
against the equations the two files state in their headers, with the parameter tables they
hold: growth = mu_max S / (Ks + S) B, dB/dt = growth - kd B, donor -growth / Y, product
+growth fP; R = k_AB A B. Skipped where the reference tree was not there to build from."""
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.skipif(not os.path.exists(orc.REF_HOOKS_PATH),
                                reason="built only where the reference tree is present (make -C oracle ref-hooks)")

MU, KS, Y, FP, KD = (2.0e-4, 1.0e-4), (5.0e-2, 5.0e-3), 0.4, 2.0, 1.0e-6


def test_example05_monod_hook():
    S, P = 0.03, 0.01
    for B in ([1.5], [1.5, 0.7]):
        sr, br = orc.kinetics_plugin(B, [S, P, 0.0], 2, library=orc.REF_HOOKS_PATH)
        g = [MU[m] * S / (KS[m] + S) * B[m] for m in range(len(B))]
        assert np.allclose(br, [g[m] - KD * B[m] for m in range(len(B))], rtol=1e-14, atol=0)
        assert np.isclose(sr[0], -sum(g) / Y, rtol=1e-14, atol=0)
        assert np.isclose(sr[1], sum(g) * FP, rtol=1e-14, atol=0)
        assert sr[2] == 0.0
    # walls do not react; a negative donor is floored at zero (no growth, decay only)
    sr, br = orc.kinetics_plugin([1.5], [S, P, 0.0], 1, library=orc.REF_HOOKS_PATH)
    assert not sr.any() and not br.any()
    sr, br = orc.kinetics_plugin([1.5], [-1e-3, P, 0.0], 2, library=orc.REF_HOOKS_PATH)
    assert sr[0] == 0.0 and np.isclose(br[0], -KD * 1.5, rtol=1e-14)


def test_example03_second_order_hook():
    A, B = 0.02, 0.01
    sr = orc.abiotic_plugin([A, B, 0.003], 2, library=orc.REF_HOOKS_PATH)
    R = 5.0e-2 * A * B
    assert np.allclose(sr, [-R, -R, R], rtol=1e-15, atol=0)
    assert not orc.abiotic_plugin([A, B, 0.0], 1, library=orc.REF_HOOKS_PATH).any()


def test_default_oracle_keeps_its_own_hooks():
    # the two libraries are separate: the default one still runs the package's dual-Monod
    B, D, A = 7.0, 0.2, 0.05
    sr, _ = orc.kinetics_plugin([B], [D, A, 0.0, 0.0], 2)
    r = 1e-3 * B * D / (1e-2 + D) * A / (1e-2 + A)
    assert np.allclose(sr, [-r, -r, r, r], rtol=1e-15, atol=0)
