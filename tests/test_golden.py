"""The committed golden vectors (tests/golden/golden_small.npz) are reproduced
by the oracle: guards the oracle against silent drift."""
import os

import numpy as np

from tests.golden.make_golden import OUT, generate


def test_golden_fixture_reproduced_by_oracle():
    ref = np.load(OUT, allow_pickle=False)
    new = generate()
    assert set(ref.files) == set(new)
    for k in ref.files:
        a, b = ref[k], new[k]
        assert a.shape == b.shape and a.dtype == b.dtype, k
        if np.issubdtype(a.dtype, np.floating):
            assert np.allclose(a, b, rtol=1e-12, atol=1e-13), k
        else:
            assert np.array_equal(a, b), k


def test_golden_known_answer():
    ref = np.load(OUT, allow_pickle=False)
    assert abs(ref["ed_j1j2_4x4"][0] - (-33.831693405579394)) < 1e-8
