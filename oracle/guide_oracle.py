"""Float64 restatement of the learned path guide (test infrastructure only).

Only the tests may import this module; it is the checker the GPU guide is
compared with, never part of the product.

The reference renderer has no path guiding, so nothing here follows a
reference file: this module restates the guide as the library defines it
(include/photon_tracer.h, "Learned path guide"):

* inputs: one-blob encoding of the position normalised to a box (3 x 16 bins)
  and of the normal's polar angle / pi and azimuth / (2 pi) (2 x 8 bins), a
  Gaussian bump of one bin's width per bin;
* network: 64 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 32, weights [out][in]
  row-major then biases, layer by layer (18,720 parameters);
* outputs per lobe k: polar logit a, azimuth logit b, log kappa c, weight
  logit d; theta = pi sigmoid(a), phi = 2 pi sigmoid(b), kappa =
  exp(clip(c, -3, 10)), weights = softmax(d);
* vMF density kappa / (2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)};
* loss of a sample: weight * -log p(direction); Adam (Kingma and Ba 2015).

What pins it (tests/test_guide_oracle.py): central finite differences of its
own loss against its backpropagated gradient, in every layer; quadrature of
the vMF and mixture densities over the sphere (they integrate to 1); the
closed-form mean cosine of a vMF lobe, coth(kappa) - 1/kappa, against the
sampling inversion; the encoding's bump centres; Adam against a hand-worked
step. With quantize=True the forward pass rounds weights, inputs and hidden
activations to FP16 where the GPU does, so only summation order differs.
"""
import numpy as np

LOBES = 8
IN, HID, OUT = 64, 64, 32
LAYER_STRIDE = HID * IN + HID
OUT_W = 4 * LAYER_STRIDE
OUT_B = OUT_W + OUT * HID
N_PARAMS = OUT_B + OUT
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -3.0, 10.0


def w_off(l):
    return l * LAYER_STRIDE if l < 4 else OUT_W


def b_off(l):
    return l * LAYER_STRIDE + HID * IN if l < 4 else OUT_B


def n_out(l):
    return HID if l < 4 else OUT


def layer(params, l):
    """(W [out][in], b [out]) of layer l."""
    no = n_out(l)
    W = params[w_off(l):w_off(l) + no * IN].reshape(no, IN)
    b = params[b_off(l):b_off(l) + no]
    return W, b


def _h(a):
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def blob(v, bins):
    """(N,) values in [0,1] -> (N, bins) one-blob features."""
    centres = (np.arange(bins) + 0.5)
    d = v[:, None] * bins - centres[None, :]
    return np.exp(-0.5 * d * d)


def encode(pos, nrm, lo, hi):
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    nrm = np.asarray(nrm, np.float64).reshape(-1, 3)
    lo = np.asarray(lo, np.float64)
    ext = np.maximum(np.asarray(hi, np.float64) - lo, 1e-6)
    c = np.clip((pos - lo) / ext, 0, 1)
    th = np.clip(np.arccos(np.clip(nrm[:, 2], -1, 1)) / np.pi, 0, 1)
    ph = np.clip((np.arctan2(nrm[:, 1], nrm[:, 0]) + np.pi) / (2 * np.pi), 0, 1)
    return np.concatenate([blob(c[:, 0], 16), blob(c[:, 1], 16), blob(c[:, 2], 16), blob(th, 8), blob(ph, 8)], 1)


def forward(params, X, quantize=False):
    """Outputs (N, 32) and the list of layer inputs [X0 .. X4]."""
    params = np.asarray(params, np.float64)
    acts = [_h(X) if quantize else np.asarray(X, np.float64)]
    for l in range(5):
        W, b = layer(params, l)
        if quantize:
            W = _h(W)
        z = acts[-1] @ W.T + b
        if l < 4:
            a = np.maximum(z, 0)
            acts.append(_h(a) if quantize else a)
        else:
            return z, acts
    raise AssertionError


def sigmoid(v):
    return 1 / (1 + np.exp(-v))


def mixture(o):
    """weights (N,8), mu (N,8,3), kappa (N,8) of outputs (N,32)."""
    o = np.asarray(o, np.float64).reshape(-1, LOBES, 4)
    d = o[:, :, 3]
    e = np.exp(d - d.max(1, keepdims=True))
    w = e / e.sum(1, keepdims=True)
    th = np.pi * sigmoid(o[:, :, 0])
    ph = 2 * np.pi * sigmoid(o[:, :, 1])
    mu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    kappa = np.exp(np.clip(o[:, :, 2], LOG_KAPPA_MIN, LOG_KAPPA_MAX))
    return w, mu, kappa


def vmf_pdf(mu, kappa, dirs):
    """Per-lobe densities: mu (N,K,3), kappa (N,K), dirs (N,3) -> (N,K)."""
    cos = np.einsum("nkc,nc->nk", mu, np.asarray(dirs, np.float64).reshape(-1, 3))
    return kappa / (2 * np.pi * -np.expm1(-2 * kappa)) * np.exp(kappa * (cos - 1))


def pdf(o, dirs):
    w, mu, kappa = mixture(o)
    return (w * vmf_pdf(mu, kappa, dirs)).sum(1)


def sample(o, u):
    """Mixture samples from uniforms u (N,3): lobe by u0, then the vMF inversion."""
    w, mu, kappa = mixture(o)
    u = np.asarray(u, np.float64).reshape(-1, 3)
    cw = np.cumsum(w, 1)
    k = np.minimum((u[:, :1] >= cw).sum(1), LOBES - 1)
    n = np.arange(len(u))
    m, kp = mu[n, k], kappa[n, k]
    cz = np.clip(1 + np.log(u[:, 1] + (1 - u[:, 1]) * np.exp(-2 * kp)) / kp, -1, 1)
    s = np.sqrt(np.maximum(0, 1 - cz * cz))
    phi = 2 * np.pi * u[:, 2]
    # any orthonormal frame around m gives the same distribution
    sign = np.copysign(1.0, m[:, 2])
    a = -1 / (sign + m[:, 2])
    b = m[:, 0] * m[:, 1] * a
    t1 = np.stack([1 + sign * m[:, 0] ** 2 * a, sign * b, -sign * m[:, 0]], 1)
    t2 = np.stack([b, sign + m[:, 1] ** 2 * a, -m[:, 1]], 1)
    return t1 * (s * np.cos(phi))[:, None] + t2 * (s * np.sin(phi))[:, None] + m * cz[:, None]


def output_grad(o, dirs, ws):
    """d(ws * -log p)/d outputs, (N,32), and the losses (N,)."""
    o = np.asarray(o, np.float64).reshape(-1, LOBES, 4)
    dirs = np.asarray(dirs, np.float64).reshape(-1, 3)
    ws = np.asarray(ws, np.float64).reshape(-1)
    w, mu, kappa = mixture(o.reshape(-1, OUT))
    v = vmf_pdf(mu, kappa, dirs)
    p = np.maximum((w * v).sum(1), 1e-300)
    r = w * v / p[:, None]
    g = -ws[:, None] * r
    sa, sb = sigmoid(o[:, :, 0]), sigmoid(o[:, :, 1])
    th, ph = np.pi * sa, 2 * np.pi * sb
    dth = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], -1)
    dph = np.stack([-np.sin(th) * np.sin(ph), np.sin(th) * np.cos(ph), np.zeros_like(th)], -1)
    cosm = np.einsum("nkc,nc->nk", mu, dirs)
    with np.errstate(over="ignore"):   # expm1 overflows to inf for large kappa: 2 / inf = 0, as wanted
        dlv_dk = 1 / kappa - 2 / np.expm1(2 * kappa) + (cosm - 1)
    free = (o[:, :, 2] > LOG_KAPPA_MIN) & (o[:, :, 2] < LOG_KAPPA_MAX)
    d = np.zeros_like(o)
    d[:, :, 0] = g * kappa * np.einsum("nkc,nc->nk", dth, dirs) * np.pi * sa * (1 - sa)
    d[:, :, 1] = g * kappa * np.einsum("nkc,nc->nk", dph, dirs) * 2 * np.pi * sb * (1 - sb)
    d[:, :, 2] = np.where(free, g * dlv_dk * kappa, 0.0)
    d[:, :, 3] = ws[:, None] * (w - r)
    return d.reshape(-1, OUT), -ws * np.log(p)


def split_samples(samples):
    s = np.asarray(samples, np.float64).reshape(-1, 12)
    return s[:, 0:3], s[:, 3], s[:, 4:7], s[:, 8:11]


def loss(params, samples, lo, hi, quantize=False):
    pos, ws, nrm, dirs = split_samples(samples)
    o, _ = forward(params, encode(pos, nrm, lo, hi), quantize)
    w, mu, kappa = mixture(o)
    p = (w * vmf_pdf(mu, kappa, dirs)).sum(1)
    return float(np.mean(-ws * np.log(p)))


def loss_and_grad(params, samples, lo, hi, quantize=False):
    """Mean loss and its gradient with respect to the parameters (N_PARAMS,)."""
    params = np.asarray(params, np.float64)
    pos, ws, nrm, dirs = split_samples(samples)
    n = len(ws)
    o, acts = forward(params, encode(pos, nrm, lo, hi), quantize)
    d, losses = output_grad(o, dirs, ws)
    grad = np.zeros(N_PARAMS)
    dy = d / n
    for l in range(4, -1, -1):
        W, _ = layer(params, l)
        if quantize:
            W = _h(W)
        X = acts[l]
        no = n_out(l)
        grad[w_off(l):w_off(l) + no * IN] = (dy.T @ X).ravel()
        grad[b_off(l):b_off(l) + no] = dy.sum(0)
        if l > 0:
            dy = (dy @ W) * (acts[l] > 0)
    return float(losses.mean()), grad


def adam(params, m, v, grad, step, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    """One Adam step (Kingma and Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
