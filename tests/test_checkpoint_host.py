"""Reading parameters from a leaf stream in the layout of
``equinox.tree_serialise_leaves`` (one .npy record per array-like leaf, in
flatten order).  Equinox and JAX are absent here, so the streams are written
with numpy by that documented layout — the parse is pinned to the layout, not
to a file quantax wrote ("parity unpinned" for real checkpoints)."""
import io

import numpy as np
import pytest

from quantax_amd.state import params_from_file, read_leaf_stream


def _stream(path, leaves):
    with open(path, "wb") as f:
        for x in leaves:
            np.save(f, x, allow_pickle=False)


def test_rbm_like_stream(tmp_path, rng):
    # SingleDense: layers = (Linear(weight (M, N), bias (M,)), Lambda, fn), holomorphic
    W = rng.standard_normal((8, 6)).astype(np.float32)
    b = rng.standard_normal(8).astype(np.float32)
    p = tmp_path / "rbm.eqx"
    _stream(p, [W, b, np.asarray(False)])
    assert len(read_leaf_stream(str(p))) == 3
    th = params_from_file(str(p), 8 * 6 + 8)
    assert np.array_equal(th, np.concatenate([W.ravel(), b]))


def test_resconv_like_stream(tmp_path, rng):
    # ResConv: nblocks, channels, kernel_size (ints), then conv weights/biases, holomorphic
    arrs = [rng.standard_normal(sh).astype(np.float32) for sh in [(4, 1, 3, 3), (4, 1, 1), (4, 4, 3, 3)]]
    p = tmp_path / "rc.eqx"
    _stream(p, [np.asarray(1), np.asarray(4), np.asarray(3), *arrs, np.asarray(False)])
    th = params_from_file(str(p), sum(a.size for a in arrs))
    assert np.array_equal(th, np.concatenate([a.ravel() for a in arrs]))
    with pytest.raises(ValueError, match="floating-point parameters"):
        params_from_file(str(p), 10)


def test_single_array_file(tmp_path, rng):
    th = rng.standard_normal(17)
    p = tmp_path / "flat.npy"
    np.save(p, th)
    assert np.array_equal(params_from_file(str(p), 17), th)
    buf = io.BytesIO(p.read_bytes())
    assert np.array_equal(read_leaf_stream(buf)[0], th)
