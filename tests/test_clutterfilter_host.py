"""Host logic of the clutter-filter drop-ins (CPU only): the kept-eigenvector
masks follow the reference conventions exactly as the FP64 oracle states them,
and the F-order frame layout round-trips."""
import itertools

import numpy as np
import pytest

from oracle import svd_oracle as S
from simu_rca_ab_amd import clutterfilter as CF


@pytest.mark.parametrize("nt", [1, 7, 16, 200, 201])
def test_keep_python_matches_oracle(nt):
    for a, b in itertools.product([0.0, 0.01, 0.1, 0.333, 0.5], [0.5, 0.9, 0.999, 1.0]):
        np.testing.assert_array_equal(CF.keep_python([a, b], nt), S.keep_python([a, b], nt))


@pytest.mark.parametrize("nt", [8, 20, 200, 201])
@pytest.mark.parametrize("kind", ["band", "low", "high", "stop"])
def test_keep_matlab_matches_oracle(nt, kind):
    for a, b in itertools.product([0.0, 0.0125, 0.05, 0.1, 0.25], [0.5, 0.8, 1.0]):
        w = [a, b]
        try:
            ref = S.keep_matlab(w, nt, kind)
        except IndexError:
            with pytest.raises(IndexError):
                CF.keep_matlab(w, nt, kind)
            continue
        np.testing.assert_array_equal(CF.keep_matlab(w, nt, kind), ref)
    with pytest.raises(ValueError):
        CF.keep_matlab([0.1, 1.2], nt, kind)


def test_matlab_round_halves_away_from_zero():
    # round(2.5) = 3 in MATLAB (numpy's round gives 2)
    assert CF.keep_matlab([0.0125], 200, "low").sum() == 3
    assert CF.keep_matlab([0.0075], 200, "low").sum() == 2   # round(1.5) = 2
