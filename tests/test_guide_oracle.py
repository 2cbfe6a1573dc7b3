"""Pins the float64 guide oracle (oracle/guide_oracle.py) before the GPU guide
is checked against it: its backpropagated gradient against finite
differences of its own loss, its densities against quadrature over the
sphere, its sampler against the closed-form mean cosine of a vMF lobe, the
encoding's bump centres, and Adam against a hand-worked step. The reference
has no guide, so these checks, not a reference run, are what fixes it."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as go  # noqa: E402

LO, HI = np.array([-2.0, -1.0, -3.0]), np.array([2.0, 3.0, 1.0])


def random_params(seed, out_scale=0.3):
    rng = np.random.default_rng(seed)
    p = np.zeros(go.N_PARAMS)
    for l in range(5):
        no = go.n_out(l)
        b = np.sqrt(6 / go.IN) * (out_scale if l == 4 else 1.0)
        p[go.w_off(l):go.w_off(l) + no * go.IN] = rng.uniform(-b, b, no * go.IN)
        p[go.b_off(l):go.b_off(l) + no] = rng.uniform(-0.1, 0.1, no)
    return p


def random_samples(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(LO, HI, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(0.1, 2.0, n)
    s = np.zeros((n, 12))
    s[:, 0:3], s[:, 3], s[:, 4:7], s[:, 8:11] = pos, w, nrm, d
    return s


def sphere_grid(n_theta=800, n_phi=1600):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)).ravel()
    return dirs, dA


def test_gradient_matches_finite_differences():
    p = random_params(1)
    s = random_samples(24, 2)
    _, g = go.loss_and_grad(p, s, LO, HI)
    rng = np.random.default_rng(3)
    idx = []
    for l in range(5):   # weights and biases of every layer
        no = go.n_out(l)
        idx += list(go.w_off(l) + rng.integers(0, no * go.IN, 8)) + list(go.b_off(l) + rng.integers(0, no, 3))
    h = 1e-5
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (go.loss(p + e, s, LO, HI) - go.loss(p - e, s, LO, HI)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_output_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    o = rng.normal(size=(10, go.OUT))
    o[:, 2::4] = rng.uniform(-2.5, 9.5, (10, go.LOBES))     # log kappa inside the clamp
    d = rng.normal(size=(10, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ws = rng.uniform(0.5, 2, 10)
    grad, _ = go.output_grad(o, d, ws)
    h = 1e-6
    for j in range(go.OUT):
        e = np.zeros_like(o)
        e[:, j] = h
        fd = (-ws * np.log(go.pdf(o + e, d)) + ws * np.log(go.pdf(o - e, d))) / (2 * h)
        assert np.allclose(fd, grad[:, j], rtol=1e-5, atol=1e-7), j


def test_clamped_kappa_passes_no_gradient():
    o = np.zeros((1, go.OUT))
    o[0, 2::4] = 11.0          # above the clamp
    grad, _ = go.output_grad(o, np.array([[0, 0, 1.0]]), np.array([1.0]))
    assert np.all(grad[0, 2::4] == 0)


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 300.0])
def test_vmf_density_integrates_to_one(kappa):
    dirs, dA = sphere_grid(2000 if kappa > 50 else 800, 1600)
    mu = np.array([[[0.3, -0.5, np.sqrt(1 - 0.34)]]])
    p = go.vmf_pdf(np.repeat(mu, len(dirs), 0), np.full((len(dirs), 1), kappa), dirs)[:, 0]
    assert abs((p * dA).sum() - 1) < 2e-3


def test_mixture_density_integrates_to_one():
    rng = np.random.default_rng(5)
    o = rng.normal(size=(1, go.OUT))
    o[0, 2::4] = rng.uniform(-1, 3, go.LOBES)
    dirs, dA = sphere_grid()
    p = go.pdf(np.repeat(o, len(dirs), 0), dirs)
    assert abs((p * dA).sum() - 1) < 1e-3


@pytest.mark.parametrize("kappa", [0.3, 4.0, 50.0])
def test_sampler_mean_cosine_is_the_closed_form(kappa):
    o = np.full((1, go.OUT), -20.0)
    o[0, 0:4] = [0.7, -0.4, np.log(kappa), 20.0]   # one lobe carries all the weight
    n = 200000
    u = np.random.default_rng(6).uniform(size=(n, 3))
    s = go.sample(np.repeat(o, n, 0), u)
    _, mu, _ = go.mixture(o)
    cos = s @ mu[0, 0]
    want = 1 / np.tanh(kappa) - 1 / kappa
    sd = np.sqrt(np.var(cos) / n)
    assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-9)
    assert abs(cos.mean() - want) < 5 * sd + 1e-9


def test_encoding_peaks_at_bin_centres():
    for i in range(16):
        x = LO + (i + 0.5) / 16 * (HI - LO)
        f = go.encode(x[None], np.array([[0, 0, 1.0]]), LO, HI)[0]
        for a in range(3):
            assert np.isclose(f[16 * a + i], 1.0) and np.argmax(f[16 * a:16 * a + 16]) == i
    f = go.encode(np.zeros((1, 3)), np.array([[0, 0, 1.0]]), LO, HI)[0]
    assert np.argmax(f[48:56]) == 0    # polar angle 0


def test_adam_first_step_is_lr_times_sign():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    q, m, v = go.adam(p, np.zeros(3), np.zeros(3), g, 1, lr=0.01)
    # step 1: m_hat = g, v_hat = g^2, so the update is lr * g / (|g| + eps)
    assert np.allclose(q, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.01 * g * g)
