"""The allocation-free ResConv Jacobian entry point against the workspace one
(needs an MI355X): bit-identical O rows on two calls in a row, the second
with fewer samples, through the C ABI.  Both run the same per-sample kernel,
so any difference is a wrong scratch address, an overlap of launches or a
stale record."""
import numpy as np
import pytest
import torch

from oracle import models as om
from tests.helpers import rand_spins

pytestmark = pytest.mark.gpu

import quantax_amd as qtx  # noqa: E402
from quantax_amd import _lib  # noqa: E402


def _jac_ws(model, sd):
    ns = sd.shape[0]
    O = torch.full((ns, model.nparams), float("nan"), dtype=torch.float32, device="cuda")
    nws = int(_lib.lib().qtx_resconv_jacobian_ws_bytes(model.desc, ns))
    ws = torch.empty(nws, dtype=torch.uint8, device="cuda")
    _lib.call("qtx_resconv_jacobian", _lib.stream(), model.desc, model.packed.data_ptr(), sd.data_ptr(), ns,
              O.data_ptr(), ws.data_ptr(), nws)
    return O


def _jac_noalloc(model, sd):
    ns = sd.shape[0]
    O = torch.full((ns, model.nparams), float("nan"), dtype=torch.float32, device="cuda")
    _lib.call("qtx_resconv_jacobian_noalloc", _lib.stream(), model.desc, model.packed.data_ptr(), sd.data_ptr(), ns,
              O.data_ptr())
    return O


@pytest.mark.parametrize("L,nb,C,sizes", [(16, 8, 16, (600, 77)), (20, 3, 32, (300, 5)), (4, 2, 16, (2100, 9))])
def test_jacobian_noalloc_equals_workspace_entry_point(L, nb, C, sizes):
    """16 x 16, 8 blocks: 256 samples per launch of the 128 MiB array, so 600
    samples take three launches; 20 x 20, 3 blocks, 32 channels: 218 per
    launch; 4 x 4: one launch of 2100 samples split into 2048-sample kernels."""
    rng = np.random.default_rng(L * 100 + nb)
    qtx.sites.Square(L)
    model = qtx.model.ResConv(nb, C, precision="fp32", seed=21)
    spec = om.ResConvSpec(nb, C, (L, L))
    theta = model.theta.cpu().numpy().astype(np.float64)
    for ns in sizes:
        s = rand_spins(rng, ns, L * L)
        sd = torch.from_numpy(s).cuda()
        a = _jac_noalloc(model, sd)
        b = _jac_ws(model, sd)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (ns, (a - b).abs().max().item())
        rows = np.unique(np.array([0, ns // 2, ns - 1]))
        ref = om.resconv_jacobian(spec, theta, s[rows])
        assert np.max(np.abs(a.cpu().numpy()[rows] - ref)) < 1e-5 * max(1.0, np.abs(ref).max())
