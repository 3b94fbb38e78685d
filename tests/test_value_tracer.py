"""The flow-convergence tracer the driver's flow loops use: Palabos' util::ValueTracer
(Palabos v2.3.0, not in the reference tree), restated in clb.ValueTracer and
complab_gpu::ValueTracer (through tests/cpp/tracer_check). The reference builds its tracers as
ValueTracer(1.0, 1000.0, epsilon) (src/complab.cpp:593, 1149): a window of 500 values.

Pinned by closed forms: a constant sequence converges exactly when the window fills; an
alternating sequence a + (-1)^k d has sample standard deviation d sqrt(n / (n - 1)) over an even
window n, so it converges iff that over a is below epsilon, on either side of the threshold;
the window slides (an early outlier stops mattering 500 values later); a zero mean with
spread is infinite, not converged; an all-zero window is NaN, reported as converged (a
diverged run)."""
import math
import os
import subprocess

import numpy as np

import clb

HERE = os.path.dirname(os.path.abspath(__file__))


def py_flags(vals, eps, u=1.0, L=1000.0):
    t = clb.ValueTracer(u, L, eps)
    out = []
    for v in vals:
        t.take_value(v)
        out.append(t.has_converged())
    return out


def cpp_flags(vals, eps, u=1.0, L=1000.0):
    exe = os.path.join(HERE, "cpp", "tracer_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), "tracer_check"], check=True,
                       stdout=subprocess.DEVNULL)
    txt = f"{u} {L} {eps}\n" + "\n".join(repr(float(v)) for v in vals) + "\n"
    r = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True)
    return [line.strip() == "1" for line in r.stdout.split()]


def both(vals, eps, **kw):
    a, b = py_flags(vals, eps, **kw), cpp_flags(vals, eps, **kw)
    assert a == b
    return a


def first(flags):
    return flags.index(True) + 1 if True in flags else -1


def test_window_of_the_reference_tracers():
    assert clb.ValueTracer(1.0, 1000.0, 1e-6).delta_t == 500
    assert clb.ValueTracer(2.0, 1000.0, 1e-6).delta_t == 250


def test_constant_sequence_converges_when_the_window_fills():
    f = both([0.3] * 600, 1e-12)
    assert first(f) == 500 and all(f[499:])


def test_alternating_sequence_threshold():
    a, d, n = 1.0, 1e-4, 500
    sd = d * math.sqrt(n / (n - 1))             # sample std of a +- d over an even window
    vals = [a + (d if k % 2 == 0 else -d) for k in range(700)]
    assert first(both(vals, sd / a * 1.001)) == 500
    assert first(both(vals, sd / a * 0.999)) == -1


def test_window_slides_past_an_outlier():
    vals = [5.0] + [1.0] * 700
    f = both(vals, 1e-9)
    assert first(f) == 501                      # the outlier leaves the window with value 501


def test_zero_mean_and_all_zero():
    spread = [(1.0 if k % 2 == 0 else -1.0) for k in range(520)]
    assert first(both(spread, 1e-3)) == -1      # infinite ratio: never converged
    assert first(both([0.0] * 520, 1e-3)) == 500  # NaN ratio: reported as converged


def test_decaying_energy_stops_where_the_closed_form_says():
    # E_k = 1 + 0.5 exp(-k / 200): the first window of 500 values whose relative spread is
    # below 1e-4, found by brute force in numpy
    k = np.arange(1, 4001)
    e = 1.0 + 0.5 * np.exp(-k / 200.0)
    eps = 1e-4
    stop = -1
    for m in range(500, 4001):
        w = e[m - 500:m]
        if abs(w.std(ddof=1) / w.mean()) < eps:
            stop = m
            break
    assert stop > 500
    assert first(both(list(e), eps)) == stop
