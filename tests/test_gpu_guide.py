"""The learned guide's GPU kernels (FP16 MFMA forward and backward, mixture
pdf and sampling, Adam) against the float64 guide oracle, which
tests/test_guide_oracle.py pins by finite differences and quadrature."""
import os
import sys

import numpy as np
import pytest

import photon_tracer_amd as pt
import render_check as rc

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as go  # noqa: E402

pytestmark = pytest.mark.gpu

LO, HI = np.array([-2.0, -1.0, -3.0]), np.array([2.0, 3.0, 1.0])


@pytest.fixture(scope="module")
def scene():
    with pt.Scene(rc.load_scene("room")[1]) as s:
        yield s


def make_guide(scene, seed=1, **kw):
    return pt.Guide(scene, seed=seed, box=(LO, HI), **kw)


def random_samples(n, seed, weight=(0.1, 2.0)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(LO, HI, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = np.zeros((n, 12), np.float32)
    s[:, 0:3], s[:, 3], s[:, 4:7], s[:, 8:11] = pos, rng.uniform(*weight, n), nrm, d
    return s


def test_forward_matches_oracle(scene):
    """FP16 weights, inputs and activations with FP32 accumulation: the outputs equal the oracle's
    FP16-quantised float64 forward pass to 1e-2 relative (1e-3 absolute)."""
    with make_guide(scene, seed=3) as g:
        s = random_samples(1000, 4)   # not a multiple of 64: the last block is partial
        w = g.weights.astype(np.float64)
        out = g.eval(s[:, 0:3], s[:, 4:7])
        ref, _ = go.forward(w, go.encode(s[:, 0:3], s[:, 4:7], LO, HI), quantize=True)
    err = np.abs(out - ref)
    print(f"forward: max abs err {err.max():.2e}, output scale {np.abs(ref).max():.2e}", flush=True)
    assert np.all(err <= 1e-3 + 1e-2 * np.abs(ref))


def test_pdf_and_sampling_match_oracle(scene):
    """pdf and samples from the GPU's own outputs, against the oracle's mixture math."""
    with make_guide(scene, seed=5) as g:
        s = random_samples(777, 6)
        x, n = s[:, 0:3], s[:, 4:7]
        o = g.eval(x, n).astype(np.float64)
        u = np.random.default_rng(7).uniform(size=(777, 3)).astype(np.float32)
        ds = g.sample(x, n, u)
        pdf_at = g.pdf(x, n, s[:, 8:11])
    ref_dir = go.sample(o, u)
    assert np.allclose(ds[:, 0:3], ref_dir, atol=2e-4)
    assert np.allclose(ds[:, 3], go.pdf(o, ds[:, 0:3]), rtol=2e-4, atol=1e-7)
    assert np.allclose(pdf_at, go.pdf(o, s[:, 8:11]), rtol=2e-4, atol=1e-7)


def test_gradient_matches_oracle(scene):
    """The MFMA backward pass (FP16 operands, FP32 gradients) against float64 backpropagation of the
    same quantised network: relative L2 error below 2e-2 over all 18,720 parameters and in each
    layer, cosine above 0.999."""
    with make_guide(scene, seed=8) as g:
        s = random_samples(4000, 9)
        grad, loss = g.gradient(s)
        grad2, _ = g.gradient(s)
        w = g.weights.astype(np.float64)
    rloss, rgrad = go.loss_and_grad(w, s.astype(np.float64), LO, HI, quantize=True)
    assert np.array_equal(grad, grad2)                      # deterministic: fixed reduction order
    assert abs(loss - rloss) <= 1e-3 * abs(rloss)
    rel = np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad)
    cos = grad @ rgrad / (np.linalg.norm(grad) * np.linalg.norm(rgrad))
    print(f"gradient: rel L2 err {rel:.2e}, cosine {cos:.6f}", flush=True)
    assert rel < 2e-2 and cos > 0.999
    for l in range(5):
        a, b = go.w_off(l), go.b_off(l) + go.n_out(l)
        r = np.linalg.norm(grad[a:b] - rgrad[a:b]) / np.linalg.norm(rgrad[a:b])
        assert r < 3e-2, (l, r)


def test_adam_step_matches_oracle(scene):
    """Two Adam steps on the GPU equal the oracle's Adam fed the GPU's own gradients."""
    with make_guide(scene, seed=10, learning_rate=2e-3) as g:
        s1, s2 = random_samples(512, 11), random_samples(640, 12)
        p0 = g.weights.astype(np.float64)
        g1, _ = g.gradient(s1)
        g.step(s1)
        g2, _ = g.gradient(s2)
        g.step(s2)
        p2 = g.weights.astype(np.float64)
    p, m, v = go.adam(p0, 0 * p0, 0 * p0, g1.astype(np.float64), 1, lr=2e-3)
    p, m, v = go.adam(p, m, v, g2.astype(np.float64), 2, lr=2e-3)
    assert np.allclose(p2, p, rtol=1e-5, atol=2e-6)


def test_training_learns_a_direction_field(scene):
    """Samples whose directions cluster (kappa 40) around an axis that turns with position: after
    Adam steps the guide puts far more density on the axis than the untrained mixture, and its loss
    falls below the uniform density's."""
    rng = np.random.default_rng(13)

    def batch(n):
        pos = rng.uniform(LO, HI, (n, 3))
        ang = (pos[:, 0] - LO[0]) / (HI[0] - LO[0]) * np.pi
        axis = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], 1)
        o = np.full((n, go.OUT), -30.0)
        th = np.arccos(axis[:, 2])
        ph = np.mod(np.arctan2(axis[:, 1], axis[:, 0]), 2 * np.pi)
        # logits of the wanted angles, one lobe of kappa 40
        o[:, 0] = np.log((th / np.pi) / (1 - th / np.pi))
        o[:, 1] = np.log((ph / (2 * np.pi) + 1e-9) / (1 - ph / (2 * np.pi) + 1e-9))
        o[:, 2], o[:, 3] = np.log(40.0), 30.0
        d = go.sample(o, rng.uniform(size=(n, 3)))
        s = np.zeros((n, 12), np.float32)
        s[:, 0:3], s[:, 3], s[:, 4:7], s[:, 8:11] = pos, 1.0, [0, 0, 1], d
        return s, axis

    test, axis = batch(2000)
    with make_guide(scene, seed=14, learning_rate=5e-3) as g:
        before = g.pdf(test[:, 0:3], test[:, 4:7], axis)
        losses = [g.step(batch(8192)[0]) for _ in range(300)]
        after = g.pdf(test[:, 0:3], test[:, 4:7], axis)
        _, test_loss = g.gradient(test)
    print(f"pdf on the axis: {np.median(before):.3f} -> {np.median(after):.3f}; loss {losses[0]:.3f} -> "
          f"{losses[-1]:.3f}, held-out {test_loss:.3f}, uniform {np.log(4 * np.pi):.3f}", flush=True)
    assert np.median(after) > 20 * np.median(before)
    assert test_loss < np.log(4 * np.pi) - 1.0


def test_bad_guide_arguments_are_refused(scene):
    with make_guide(scene) as g:
        with pytest.raises(pt.PTError):
            g.set_weights(np.zeros(10, np.float32))
        w = g.weights
        w[5] = np.nan
        with pytest.raises(pt.PTError):
            g.set_weights(w)
    with pytest.raises(pt.PTError):
        pt.Guide(scene, box=((0, 0, 0), (1, 0, 1)))
