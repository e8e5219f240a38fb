"""The oracle's SST k-omega 2003 cell model (turbKWSst, src/turbulence.cpp:
573-760; constants include/turbulence.hpp:489-501), pinned by what the model's
published form fixes (Menter, Kuntz and Langtry 2003): the blending limits,
the two branches of the eddy-viscosity limiter, the production of a simple
shear and its limiter, the clipped production of a pure dilatation, the
blended destruction and cross-diffusion terms, the source-Jacobian diagonal,
and invariance under a rotation of the frame."""
import math

import numpy as np
import pytest

import pyoracle

BETA_STAR, SIGMA_W2, A1 = 0.09, 0.856, 0.31
BETA1, BETA2 = 0.075, 0.0828


def cell(rho=1.2, k=2.0e-3, om=50.0, G=None, kg=(0, 0, 0), wg=(0, 0, 0), mu=1.8e-5, d=0.1,
         vol=1.0e-6, sc=1.0):
    G = np.zeros(9) if G is None else np.asarray(G, dtype=float).ravel()
    return np.concatenate([[rho, 10.0, 0.0, 0.0, 1.0e5, k, om], G, kg, wg,
                           [mu, d, vol, sc, 1.0 / sc]])


def run(c):
    return pyoracle.sst_cells(c)[0]


def shear(g):
    G = np.zeros((3, 3))
    G[0, 1] = g  # du/dy
    return G


def test_blending_is_one_at_the_wall_and_vanishes_far_from_it():
    near, far = run(cell(d=1.0e-7)), run(cell(d=1.0e4))
    assert near[0] == 1.0 and near[1] == 1.0
    assert far[0] < 1e-20 and far[1] < 1e-10
    # far from the wall the blended coefficients are the outer (k-epsilon) set
    f = run(cell(d=1.0e4, vol=2.0, sc=1.0))
    assert f[6] == pytest.approx(-2.0 * BETA2 * 50.0 * 2.0, rel=1e-12)
    n = run(cell(d=1.0e-7, vol=2.0))
    assert n[6] == pytest.approx(-2.0 * BETA1 * 50.0 * 2.0, rel=1e-15)


def test_blending_functions_follow_the_published_arguments():
    rng = np.random.default_rng(3)
    for _ in range(40):
        rho, k, om = rng.uniform(0.5, 2), 10 ** rng.uniform(-4, -1), 10 ** rng.uniform(0, 4)
        mu, d = 10 ** rng.uniform(-6, -4), 10 ** rng.uniform(-5, 0)
        kg, wg = rng.normal(size=3), rng.normal(size=3) * 100
        out = run(cell(rho, k, om, kg=kg, wg=wg, mu=mu, d=d))
        nu = mu / rho
        cd = max(2 * rho * SIGMA_W2 / om * float(np.dot(kg, wg)), 1e-10)
        arg1 = min(max(math.sqrt(k) / (BETA_STAR * om * d), 500 * nu / (d * d * om)),
                   4 * rho * SIGMA_W2 * k / (cd * d * d))
        arg2 = max(2 * math.sqrt(k) / (BETA_STAR * om * d), 500 * nu / (d * d * om))
        assert out[0] == pytest.approx(math.tanh(arg1 ** 4), rel=1e-12, abs=1e-300)
        assert out[1] == pytest.approx(math.tanh(arg2 ** 2), rel=1e-12, abs=1e-300)


def test_eddy_viscosity_is_k_over_omega_without_strain():
    out = run(cell(rho=1.3, k=3e-3, om=40.0))
    assert out[2] == pytest.approx(1.3 * 3e-3 / 40.0, rel=1e-15)


def test_eddy_viscosity_limiter_in_a_strong_shear():
    # near the wall F2 = 1; for a simple shear g the mean strain rate
    # sqrt(2 S:S) is |g|, so mut = rho a1 k / |g| once |g| > a1 omega
    g = -1.0e4
    out = run(cell(rho=1.1, k=2e-3, om=50.0, G=shear(g), d=1e-7))
    assert out[1] == 1.0
    assert out[2] == pytest.approx(1.1 * A1 * 2e-3 / abs(g), rel=1e-15)
    # the transposed shear and a rotation of the shear plane give the same
    assert run(cell(rho=1.1, k=2e-3, om=50.0, G=shear(g).T, d=1e-7))[2] == out[2]


def test_production_of_a_simple_shear_and_its_limiter():
    rho, k, om = 1.2, 2e-3, 50.0
    weak = 0.5  # a1 omega > |g| F2: mut = rho k / omega
    out = run(cell(rho, k, om, G=shear(weak), d=1e-7))
    mut = rho * k / om
    dest = BETA_STAR * rho * k * om
    assert out[3] == pytest.approx(mut * weak * weak - dest, rel=1e-13)
    # production capped at 10 x destruction
    strong = 200.0
    out = run(cell(rho, k, om, G=shear(strong), d=1e-7))
    assert out[3] == pytest.approx(10 * dest - dest, rel=1e-13)


def test_dilatation_production_is_clipped_to_zero():
    # G = a I: tau : G = -2 rho k a < 0 for an expansion, which the 2003
    # limiter clips, leaving the destruction
    rho, k, om = 1.2, 2e-3, 50.0
    out = run(cell(rho, k, om, G=np.eye(3) * 30.0, d=1e4))
    assert out[3] == -(BETA_STAR * rho * k * om)
    # compression produces 2 rho k |a| (below the cap)
    a = -0.5
    out = run(cell(rho, k, om, G=np.eye(3) * a, d=1e4))
    assert out[3] == pytest.approx(2 * rho * k * abs(a) - BETA_STAR * rho * k * om, rel=1e-12)


def test_omega_source_destruction_and_cross_diffusion():
    rho, k, om = 1.0, 1e-3, 20.0
    kg, wg = np.array([0.3, -0.2, 0.5]), np.array([40.0, 10.0, 60.0])
    far = run(cell(rho, k, om, kg=kg, wg=wg, d=1e4))
    cd = 2 * rho * SIGMA_W2 / om * float(np.dot(kg, wg))
    assert far[0] < 1e-12
    assert far[4] == pytest.approx(-BETA2 * rho * om * om + cd, rel=1e-12)
    # opposed gradients: the cross-diffusion coefficient floors at 1e-10
    far = run(cell(rho, k, om, kg=kg, wg=-wg, d=1e4))
    assert far[4] == pytest.approx(-BETA2 * rho * om * om + 1e-10, rel=1e-12)
    near = run(cell(rho, k, om, kg=kg, wg=wg, d=1e-7))
    assert near[4] == pytest.approx(-BETA1 * rho * om * om, rel=1e-15)


def test_nondimensional_scaling():
    # the destruction terms and the Jacobian carry 1/scaling; with no
    # gradients the k source is -beta* rho k omega / scaling
    sc = 4.0
    out = run(cell(k=2e-3, om=50.0, sc=sc, d=1e4, vol=3.0))
    assert out[3] == pytest.approx(-BETA_STAR * 1.2 * 2e-3 * 50.0 / sc, rel=1e-15)
    assert out[5] == pytest.approx(-2 * BETA_STAR * 50.0 * 3.0 / sc, rel=1e-15)


def test_outputs_are_invariant_under_a_rotation():
    rng = np.random.default_rng(11)
    for _ in range(10):
        G = rng.normal(size=(3, 3)) * 20
        kg, wg = rng.normal(size=3), rng.normal(size=3) * 50
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        base = dict(rho=1.1, k=3e-3, om=30.0, mu=2e-5, d=10 ** rng.uniform(-4, -1))
        a = run(cell(G=G, kg=kg, wg=wg, **base))
        b = run(cell(G=Q @ G @ Q.T, kg=Q @ kg, wg=Q @ wg, **base))
        np.testing.assert_allclose(b, a, rtol=1e-10, atol=1e-14)


def test_a_planted_transpose_in_the_production_is_seen():
    # a non-symmetric gradient whose production tau : G differs from
    # tau : G^T if tau were not symmetric: the oracle's production must
    # equal the one from the symmetric part alone
    rho, k, om = 1.2, 2e-3, 50.0
    G = np.array([[0.1, 0.4, 0.0], [-0.3, 0.05, 0.2], [0.1, 0.0, -0.02]])
    out = run(cell(rho, k, om, G=G, d=1e4))
    mut = out[2]
    S = 0.5 * (G + G.T)
    tau = 2 * mut * S - (2.0 / 3.0) * mut * np.trace(G) * np.eye(3) - (2.0 / 3.0) * rho * k * np.eye(3)
    prod = min(max(float(np.sum(tau * G)), 0.0), 10 * BETA_STAR * rho * k * om)
    assert out[3] == pytest.approx(prod - BETA_STAR * rho * k * om, rel=1e-12)
