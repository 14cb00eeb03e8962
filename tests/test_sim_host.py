"""Host side of the simulator drop-in on the CPU: the product's grid/spectra
preparation equals the oracle's, value for value, and argument validation
follows the reference wrapper (SimUS_CUDARCA.m:122-173)."""
import numpy as np
import pytest

from oracle import simus_oracle as S
from simu_rca_ab_amd import simulator as SIM
from tests.simus_cases import rca_case


@pytest.mark.parametrize("mode", ["NS200BW", "BS100BW", "BS50BW"])
def test_prepare_matches_oracle(mode):
    M, TT, TR, TX, RX = rca_case(mode=mode, n_sample=300, n_tx=3)
    a = SIM.prepare(M, TT, TR, TX, RX)
    b = S.prepare(M, TT, TR, TX, RX)
    for k in ("nfft", "nf", "sub"):
        assert a[k] == b[k], k
    for k in ("agl", "k", "pulse", "ir_tx", "ir_rx", "ipfilt"):
        assert a[k].dtype == b[k].dtype, k
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    np.testing.assert_array_equal(a["delay"], b["delay"].T)
    np.testing.assert_array_equal(a["apod"], b["apod"].T)


def test_element_layout_and_sub_element_count():
    M, TT, TR, TX, RX = rca_case(n_ele=5, n_sub=3)
    e = SIM._element_array(TT)
    assert e.shape == (3, 5, 5)
    np.testing.assert_array_equal(e[2, 1, 4], TT["ElementPos"][4, 1, 2])
    with pytest.raises(SIM.SimInputError) as x:
        SIM._element_array(TR, n_sub=4)
    assert x.value.identifier == "SimUS_CUDARCA:subElementMismatch"


@pytest.mark.parametrize("value", [1.5, -2])
def test_decrease_factor_validation(value):
    with pytest.raises(SIM.SimInputError) as x:
        SIM._validate_options({"decreaseFactorGPU": value})
    assert x.value.identifier == "SimUS_CUDARCA:decreaseFactorGPUInvalid"
    assert SIM._validate_options(None)["decreaseFactorGPU"] == 1


def test_unknown_sample_mode():
    M, TT, TR, TX, RX = rca_case()
    RX = [dict(r, sampleMode="XX") for r in RX]
    with pytest.raises(SIM.SimInputError) as x:
        SIM.prepare(M, TT, TR, TX, RX)
    assert x.value.identifier == "SimUS_CUDARCA:sampleModeInvalid"
