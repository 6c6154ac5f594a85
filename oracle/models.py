"""RBM_Dense and ResConv wavefunctions in float64 (oracle).

RBM_Dense — ``quantax/model/shallow_nets.py:35-126``: h = W s + b (eqx Linear,
weight (M, N) then bias (M,)), psi = prod cosh(h) as a LogArray
(``nn/activation.py:17-23``, ``utils/big_array.py:390-396``), local update
h += 2 W[:, j] s_j for each changed site (``shallow_nets.py:101-104``).
Log-derivative O = [tanh(h) (x) s, tanh(h)] in ravel_pytree order (weight
row-major, then bias; ``state/variational.py:228-259, 424-511``).

ResConv — ``quantax/model/conv_nets.py:26-183``, ``nn/conv.py:13-68``:
spins reshaped to (1, L0, L1); block i: r = x; x /= sqrt(i+1); block 0:
x /= sqrt(2), else GELU; conv3x3 CIRCULAR + bias; GELU; conv3x3 (+ bias except
the last block); channel-broadcast r when channels grow; x + r.  Head:
x /= sqrt(nblocks+1); exp_by_scale; mean over channels and translations
(``ConvSymmetrize`` with ``TransND`` sector 0, ``symmetry/symmetry.py:343-392``)
=> log psi = logsumexp(h) - log(C * N), sign +1.
Parameters in ravel_pytree order: per block conv1.weight (C, Cin, 3, 3),
conv1.bias (C,), conv2.weight (C, C, 3, 3), conv2.bias (C,) (absent in the
last block).  Convolution is cross-correlation with circular 'SAME' padding.
"""
from __future__ import annotations

import numpy as np

SQRT_2_OVER_PI = np.sqrt(2.0 / np.pi)


# ----------------------------------------------------------------- RBM
def logcosh(x):
    ax = np.abs(x)
    return ax + np.log1p(np.exp(-2.0 * ax)) - np.log(2.0)


def rbm_h(W, b, s):
    return np.asarray(s, np.float64) @ np.asarray(W, np.float64).T + np.asarray(b, np.float64)


def rbm_logpsi(W, b, s):
    """Returns (sign, logabs) float64 (sign is +1: cosh > 0)."""
    la = logcosh(rbm_h(W, b, s)).sum(axis=1)
    return np.ones_like(la), la


def rbm_ref_update(W, h, s_new, s_old):
    """Local update of the hidden pre-activations for every changed site."""
    ds = (np.asarray(s_new, np.float64) - np.asarray(s_old, np.float64))  # = 2 s_new on flips
    return h + ds @ np.asarray(W, np.float64).T


def rbm_jacobian(W, b, s):
    s = np.asarray(s, np.float64)
    t = np.tanh(rbm_h(W, b, s))
    Ns, M = t.shape
    return np.concatenate([(t[:, :, None] * s[:, None, :]).reshape(Ns, -1), t], axis=1)


def rbm_pack(W, b):
    return np.concatenate([np.ravel(W), np.ravel(b)])


def rbm_unpack(theta, M, N):
    theta = np.asarray(theta)
    return theta[: M * N].reshape(M, N), theta[M * N: M * N + M]


# ----------------------------------------------------------------- ResConv
def gelu(x, approximate=True):
    """GELU, tanh form (approximate=True, jax.nn.gelu's default) or erf form;
    approximate="relu" selects the opt-in ReLU variant, max(x, 0), which the
    reference (model/conv_nets.py:85, 87) does not have."""
    if isinstance(approximate, str):
        assert approximate == "relu"
        return np.maximum(x, 0.0)
    if approximate:
        return 0.5 * x * (1.0 + np.tanh(SQRT_2_OVER_PI * (x + 0.044715 * x ** 3)))
    from scipy.special import erf
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x, approximate=True):
    if isinstance(approximate, str):
        assert approximate == "relu"
        return (np.asarray(x) > 0.0).astype(np.float64)  # jax.nn.relu: 0 at x = 0
    if approximate:
        u = SQRT_2_OVER_PI * (x + 0.044715 * x ** 3)
        t = np.tanh(u)
        return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * SQRT_2_OVER_PI * (1.0 + 3 * 0.044715 * x * x)
    from scipy.special import erf
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


class ResConvSpec:
    """Shape bookkeeping for ResConv(nblocks, channels, 3) on an (L0, L1) grid."""

    def __init__(self, nblocks, channels, shape, kernel_size=3):
        assert kernel_size == 3
        self.nblocks = int(nblocks)
        self.C = int(channels)
        self.shape = tuple(int(v) for v in shape)
        self.N = int(np.prod(self.shape))
        self.layers = []  # (cin, cout, has_bias)
        for i in range(self.nblocks):
            self.layers.append((1 if i == 0 else self.C, self.C, True))
            self.layers.append((self.C, self.C, i != self.nblocks - 1))
        self.offsets = []
        off = 0
        for cin, cout, hb in self.layers:
            w = off
            off += cout * cin * 9
            b = off if hb else -1
            off += cout if hb else 0
            self.offsets.append((w, b))
        self.nparams = off

    def unpack(self, theta):
        theta = np.asarray(theta, np.float64)
        out = []
        for (cin, cout, hb), (w, b) in zip(self.layers, self.offsets):
            W = theta[w: w + cout * cin * 9].reshape(cout, cin, 3, 3)
            B = theta[b: b + cout] if hb else np.zeros(cout)
            out.append((W, B))
        return out

    def he_normal(self, rng, dtype=np.float32):
        """He-normal weights (std sqrt(2 / fan_in), fan_in = cin*9), zero biases
        (``nn/initializers.py:91-119``) from a numpy Generator."""
        theta = np.zeros(self.nparams, np.float64)
        for (cin, cout, hb), (w, b) in zip(self.layers, self.offsets):
            n = cout * cin * 9
            theta[w: w + n] = rng.standard_normal(n) * np.sqrt(2.0 / (cin * 9))
        return theta.astype(dtype)


def _shift(x, da, db):
    """x[..., (a + da - 1) % L0, (b + db - 1) % L1] for tap (da, db)."""
    return np.roll(x, shift=(1 - da, 1 - db), axis=(-2, -1))


def _unshift(x, da, db):
    return np.roll(x, shift=(da - 1, db - 1), axis=(-2, -1))


def conv3x3(x, W, B):
    """Circular 'SAME' cross-correlation: x (Bt, Cin, L0, L1), W (Cout, Cin, 3, 3)."""
    out = np.zeros((x.shape[0], W.shape[0]) + x.shape[2:])
    for da in range(3):
        for db in range(3):
            out += np.einsum("oc,bcyx->boyx", W[:, :, da, db], _shift(x, da, db), optimize=True)
    return out + B[None, :, None, None]


def resconv_forward(spec, theta, s, approximate=True, cache=False):
    """(Ns, N) +-1 -> (sign, logabs) float64; optional activation cache."""
    s = np.asarray(s, np.float64)
    Ns = s.shape[0]
    params = spec.unpack(theta)
    x = s.reshape((Ns, 1) + spec.shape)
    caches = []
    for i in range(spec.nblocks):
        W1, B1 = params[2 * i]
        W2, B2 = params[2 * i + 1]
        r = x
        xin = x / np.sqrt(i + 1)
        if i == 0:
            y = xin / np.sqrt(2.0)
        else:
            y = gelu(xin, approximate)
        z1 = conv3x3(y, W1, B1)
        a1 = gelu(z1, approximate)
        z2 = conv3x3(a1, W2, B2)
        if z2.shape[1] > r.shape[1]:
            r = np.repeat(r, z2.shape[1] // r.shape[1], axis=1)
        x = z2 + r
        if cache:
            caches.append((xin, y, z1, a1))
    h = (x / np.sqrt(spec.nblocks + 1)).reshape(Ns, -1)
    hmax = h.max(axis=1, keepdims=True)
    la = hmax[:, 0] + np.log(np.exp(h - hmax).sum(axis=1)) - np.log(h.shape[1])
    if cache:
        return (np.ones(Ns), la), (caches, h)
    return np.ones(Ns), la


def resconv_jacobian(spec, theta, s, approximate=True):
    """Per-sample d log psi / d theta, (Ns, nparams) float64 (manual backward)."""
    s = np.asarray(s, np.float64)
    Ns = s.shape[0]
    params = spec.unpack(theta)
    (_, la), (caches, h) = resconv_forward(spec, theta, s, approximate, cache=True)
    p = np.exp(h - h.max(axis=1, keepdims=True))
    p /= p.sum(axis=1, keepdims=True)
    dx = (p / np.sqrt(spec.nblocks + 1)).reshape((Ns, spec.C) + spec.shape)
    O = np.zeros((Ns, spec.nparams))
    for i in reversed(range(spec.nblocks)):
        W1, B1 = params[2 * i]
        W2, B2 = params[2 * i + 1]
        xin, y, z1, a1 = caches[i]
        (w1o, b1o), (w2o, b2o) = spec.offsets[2 * i], spec.offsets[2 * i + 1]
        cin1 = W1.shape[1]
        dz2 = dx
        # conv2 grads
        gW2 = np.zeros((Ns,) + W2.shape)
        da1 = np.zeros_like(a1)
        for da in range(3):
            for db in range(3):
                gW2[:, :, :, da, db] = np.einsum("boyx,bcyx->boc", dz2, _shift(a1, da, db), optimize=True)
                da1 += np.einsum("oc,boyx->bcyx", W2[:, :, da, db], _unshift(dz2, da, db), optimize=True)
        O[:, w2o: w2o + W2.size] = gW2.reshape(Ns, -1)
        if b2o >= 0:
            O[:, b2o: b2o + W2.shape[0]] = dz2.sum(axis=(2, 3))
        dz1 = da1 * gelu_grad(z1, approximate)
        gW1 = np.zeros((Ns,) + W1.shape)
        dy = np.zeros_like(y)
        for da in range(3):
            for db in range(3):
                gW1[:, :, :, da, db] = np.einsum("boyx,bcyx->boc", dz1, _shift(y, da, db), optimize=True)
                dy += np.einsum("oc,boyx->bcyx", W1[:, :, da, db], _unshift(dz1, da, db), optimize=True)
        O[:, w1o: w1o + W1.size] = gW1.reshape(Ns, -1)
        O[:, b1o: b1o + W1.shape[0]] = dz1.sum(axis=(2, 3))
        if i == 0:
            dxin = dy / np.sqrt(2.0)
        else:
            dxin = dy * gelu_grad(xin, approximate)
        dxprev = dxin / np.sqrt(i + 1)
        # residual path: r was repeated over channels when they grew
        if cin1 < spec.C:
            dxprev = dxprev + dx.reshape((Ns, cin1, spec.C // cin1) + spec.shape).sum(axis=2)
        else:
            dxprev = dxprev + dx
        dx = dxprev
    return O
