"""Pins the oracle's D3Q7 advection-diffusion restatement (Palabos
AdvectionDiffusionBGKdynamics with Dirichlet / Neumann x faces,
src/complab_functions.hh:296-340) by closed forms: the linear steady diffusion profile
of the reference's example 02 (examples/README.md:54-56), the 1-D advection-diffusion
erfc front (which fixes c_s^2 = 1/4 and the advection sign), and exact mass
conservation in a closed box."""
import math

import numpy as np

from conftest import d3q19_kind_omega


def open_box(nx, ny, nz):
    return np.full(nx * ny * nz, 2, dtype=np.uint8)


def ade_setup(nx, ny, nz, tau, S=1):
    kind, _ = d3q19_kind_omega()
    om = np.full((S, 256), 1.0 / tau)
    return kind, om


def test_linear_steady_diffusion_profile(oracle):
    nx, ny, nz, tau = 12, 3, 3, 1.0
    kind, om = ade_setup(nx, ny, nz, tau)
    m = open_box(nx, ny, nz)
    u = np.zeros((nx * ny * nz, 3))
    run = oracle.AdeRun(nx, ny, nz, m, kind, om, [0, 0], [1.0, 0.0], u, [0.0])
    run.run(6000)
    c = run.conc()[0].reshape(nz, ny, nx)
    exact = 1.0 - np.arange(nx) / (nx - 1)
    assert np.max(np.abs(c - exact[None, None, :])) < 1e-8


def test_neumann_outlet_flat_profile(oracle):
    # Dirichlet 1 at the inlet, Neumann at the outlet, no flow: the steady state is C = 1.
    nx, ny, nz = 10, 2, 2
    kind, om = ade_setup(nx, ny, nz, 0.9)
    run = oracle.AdeRun(nx, ny, nz, open_box(nx, ny, nz), kind, om, [0, 1], [1.0, 0.0],
                        np.zeros((nx * ny * nz, 3)), [0.0])
    run.run(8000)
    assert np.max(np.abs(run.conc() - 1.0)) < 1e-6


def ogata_banks(x, t, U, D):
    """C(0, t) = 1, C(x, 0) = 0 on x > 0:
    C = 1/2 [erfc((x - U t)/(2 sqrt(D t))) + exp(U x / D) erfc((x + U t)/(2 sqrt(D t)))]."""
    from scipy.special import erfc, erfcx
    s = 2.0 * np.sqrt(D * t)
    b = (x + U * t) / s
    return 0.5 * (erfc((x - U * t) / s) + np.exp(U * x / D - b * b) * erfcx(b))


def test_advection_diffusion_erfc_front(oracle):
    # Dirichlet inlet front with D = c_s^2 (tau - 1/2), c_s^2 = 1/4 (D3Q7 weights 1/4, 1/8).
    nx, ny, nz, tau, U, t = 300, 1, 1, 0.9, 0.05, 2000
    kind, om = ade_setup(nx, ny, nz, tau)
    u = np.zeros((nx * ny * nz, 3))
    u[:, 0] = U
    run = oracle.AdeRun(nx, ny, nz, open_box(nx, ny, nz), kind, om, [0, 1], [1.0, 0.0], u, [0.0])
    run.run(t)
    c = run.conc()[0]
    D = 0.25 * (tau - 0.5)
    x = np.arange(nx, dtype=np.float64)
    exact = ogata_banks(x, t, U, D)
    sl = slice(1, 200)
    err = np.max(np.abs(c[sl] - exact[sl]))
    assert err < 5e-3
    # a wrong lattice sound speed (1/3) or a reversed advection must fit far worse
    wrong_cs = ogata_banks(x, t, U, (1.0 / 3.0) * (tau - 0.5))
    wrong_u = ogata_banks(x, t, -U, D)
    assert err < 0.25 * np.max(np.abs(c[sl] - wrong_cs[sl]))
    assert err < 0.25 * np.max(np.abs(c[sl] - wrong_u[sl]))


def test_closed_box_conserves_mass(oracle):
    nx, ny, nz = 10, 6, 5
    rng = np.random.default_rng(2)
    m = np.full((nz, ny, nx), 2, dtype=np.uint8)
    m[:, :, 0] = 1
    m[:, :, nx - 1] = 1
    m[rng.random((nz, ny, nx)) < 0.2] = 1
    m = m.ravel()
    kind, om = ade_setup(nx, ny, nz, 0.8, S=2)
    u = 0.02 * rng.standard_normal((m.size, 3))
    u[m != 2] = 0
    run = oracle.AdeRun(nx, ny, nz, m, kind, om, [1, 1, 1, 1], [0, 0, 0, 0], u, [0.3, 0.7])
    g0 = run.g.reshape(2, 7, -1).sum(axis=(1, 2))
    run.run(500)
    g1 = run.g.reshape(2, 7, -1).sum(axis=(1, 2))
    assert np.max(np.abs(g1 - g0)) < 1e-11
