"""FP32 ResConv forward on the host cores (oracle package; TEST AND BASELINE
INFRASTRUCTURE ONLY).

The same network as ``oracle.models.resconv_forward`` (``quantax/model/
conv_nets.py:26-183``, ``nn/conv.py:13-68``: block i: r = x; x /= sqrt(i+1);
block 0: x /= sqrt(2), else GELU; circular 3x3 conv + bias; GELU; circular 3x3
conv (+ bias except the last block); channel-broadcast r when the channel
count grows; x + r; head logsumexp(x / sqrt(nb+1)) - log(C N)), evaluated in
float32 -- the reference's parameter dtype (``model/conv_nets.py:115``) --
with torch's CPU convolution (circular padding, cross-correlation), i.e. the
kind of compiled CPU convolution a JAX-CPU run of the reference uses.  Only
``bench.py``'s ``cpu_baseline`` leg and the tests use it; it is pinned against
the float64 restatement in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import math

import numpy as np


def resconv_forward_f32(spec, theta, s, approximate=True):
    """(Ns, N) +-1 -> log psi (Ns,) float64, computed in float32 on the CPU."""
    import torch
    import torch.nn.functional as F

    L0, L1 = spec.shape
    x = torch.from_numpy(np.asarray(s, np.float32)).reshape(-1, 1, L0, L1)
    params = [(torch.from_numpy(np.asarray(W, np.float32)), torch.from_numpy(np.asarray(B, np.float32)))
              for W, B in spec.unpack(theta)]
    if approximate == "relu":
        act = F.relu
    else:
        act = (lambda v: F.gelu(v, approximate="tanh")) if approximate else F.gelu

    def conv(v, W, B):
        return F.conv2d(F.pad(v, (1, 1, 1, 1), mode="circular"), W, B)

    with torch.no_grad():
        for i in range(spec.nblocks):
            (W1, B1), (W2, B2) = params[2 * i], params[2 * i + 1]
            r = x
            xin = x / math.sqrt(i + 1)
            y = xin / math.sqrt(2.0) if i == 0 else act(xin)
            z2 = conv(act(conv(y, W1, B1)), W2, B2)
            if z2.shape[1] > r.shape[1]:
                r = r.repeat(1, z2.shape[1] // r.shape[1], 1, 1)
            x = z2 + r
        h = (x / math.sqrt(spec.nblocks + 1)).reshape(x.shape[0], -1)
        la = torch.logsumexp(h, dim=1) - math.log(h.shape[1])
    return la.double().numpy()
