"""Host-side logic of the drop-in (validation, data conversion, geometry prep)
on the CPU; the beamforming itself is covered by the GPU tests."""
import warnings

import numpy as np
import pytest
import torch

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C
from simu_rca_ab_amd.beamformer import RcaInputError, RcaWarning, _to_iq, _validate
from simu_rca_ab_amd.plan import prepare_geometry
from tests import helpers as H

CFG = C.scaled(C.get("cfg0"), nx=4, ny=3, nz=5)


def _args():
    G, T, P, Op = H.case(CFG)
    return np.zeros((CFG.n_sample, 64, 1, 11), np.complex64), G, T, P, Op


@pytest.mark.parametrize("mutate,ident", [
    (lambda d, G, T, P, O: ("x", G, T, P, O), "dataNotNumeric"),
    (lambda d, G, T, P, O: (d, G, T, P, 3), "OptionsNotStruct"),
    (lambda d, G, T, P, O: (d, G, T, P, {**O, "flagCompounding": 1}), "flagCompoundingInvalid"),
    (lambda d, G, T, P, O: (d, G, T, P, {**O, "flagMaskTransmit": "no"}), "flagMaskTransmitInvalid"),
    (lambda d, G, T, P, O: (d, G, T, P, {**O, "flagInterpolationLanczos": 0}),
     "flagInterpolationLanczosInvalid"),
    (lambda d, G, T, P, O: (d, G, T, P, {**O, "decreaseFactorGPU": 1.5}), "decreaseFactorGPUInvalid"),
    (lambda d, G, T, P, O: (d, {k: v for k, v in G.items() if k != "gridBfY"}, T, P, O), "gridBfMissing"),
    (lambda d, G, T, P, O: (d, {**G, "gridBfZ": G["gridBfZ"][:2]}, T, P, O), "gridBfInvalid"),
    (lambda d, G, T, P, O: (d, {k: v for k, v in G.items() if k != "posEleZ"}, T, P, O), "posEleMissing"),
    (lambda d, G, T, P, O: (d, {**G, "apertureX": [1.0, 2.0]}, T, P, O), "apertureXInvalid"),
    (lambda d, G, T, P, O: (d, G, [1], P, O), "TransmitNotStruct"),
    (lambda d, G, T, P, O: (d, G, {**T, "virtualSourceX": [0.0]}, P, O), "ThetaORVirtualSourceRequired"),
    (lambda d, G, T, P, O: (d, G, {**T, "thetaX": T["thetaX"][:3]}, P, O), "tiltsNotSameLength"),
    (lambda d, G, T, P, O: (d, G, {"thetaX": [0.0], "flagReceiveX": True}, P, O),
     "ThetaORVirtualSourceInvalid"),
    (lambda d, G, T, P, O: (d, G, {"virtualSourceX": [0.0], "virtualSourceY": [0.0, 1.0],
                                   "virtualSourceZ": [0.0]}, P, O), "virtualSourcesNotSameLength"),
    (lambda d, G, T, P, O: (d, G, {**T, "flagReceiveX": 1}, P, O), "flagReceiveXInvalid"),
    (lambda d, G, T, P, O: (d, G, T, 5, O), "ParamNotStruct"),
    (lambda d, G, T, P, O: (d, G, T, {k: v for k, v in P.items() if k != "samplingFrequency"}, O),
     "samplingFrequencyMissing"),
    (lambda d, G, T, P, O: (d, G, T, {**P, "demodulationFrequency": [1, 2]}, O),
     "demodulationFrequencyInvalid"),
    (lambda d, G, T, P, O: (d, G, T, {**P, "fnumber": "one"}, O), "startTimeInvalid"),
])
def test_validation_errors(mutate, ident):
    with pytest.raises(RcaInputError) as e:
        _validate(*mutate(*_args()))
    assert e.value.identifier == f"bfVirtualSourceGeneric_RCA:{ident}"


def test_defaults_and_warnings():
    d, G, T, P, Op = _args()
    G = {k: v for k, v in G.items() if not k.startswith("aperture")}
    T = {k: v for k, v in T.items() if k != "flagReceiveX"}
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        T2, P2, O2 = _validate(d.real, G, T, {"samplingFrequency": 1.0, "demodulationFrequency": 1.0}, None)
    ids = {str(x.message).split(":")[1] for x in w if issubclass(x.category, RcaWarning)}
    assert {"dataReal", "apertureXDefault", "apertureYDefault", "apertureZDefault",
            "thetaToVirtualSource", "speedOfSoundDefault", "startTimeDefault",
            "fnumberDefault"} <= ids
    assert T2["flagReceiveX"] is True and P2["speedOfSound"] == 1540 and P2["fnumber"] == 1
    assert O2["flagInterpolationLanczos"] is False and O2["flagCompounding"] is True
    assert O2["flagMaskTransmit"] is True and O2["decreaseFactorGPU"] == 1


def test_real_interleaved_conversion():
    rng = np.random.default_rng(0)
    rf = rng.standard_normal((8, 3, 1, 4)).astype(np.float32)   # Ns=4 I/Q, 3 channels, 2 tilts x 2 frames
    iq, ns = _to_iq(rf, 3, 2, "cpu")
    assert ns == 4 and tuple(iq.shape) == (2, 2, 3, 4)
    # IQ(n) = RF(2n) - j RF(2n+1) along the sample axis (RCA.m:476)
    want = (rf[0::2] - 1j * rf[1::2])                      # [4, 3, 1, 4]
    want = want[:, :, 0, :].transpose(2, 1, 0).reshape(2, 2, 3, 4)
    np.testing.assert_array_equal(iq.numpy(), want.astype(np.complex64))
    # torch input with the same values gives the same tensor
    iq2, _ = _to_iq(torch.from_numpy(rf), 3, 2, "cpu")
    np.testing.assert_array_equal(iq2.numpy(), iq.numpy())


def test_complex_conversion_and_size_checks():
    x = (np.arange(2 * 3 * 6) + 1j).reshape(2, 3, 1, 6).astype(np.complex64)
    iq, ns = _to_iq(x, 3, 3, "cpu")
    assert ns == 2 and tuple(iq.shape) == (2, 3, 3, 2)
    np.testing.assert_array_equal(iq.numpy().transpose(3, 2, 1, 0).reshape(2, 3, 6, order="F"),
                                  x[:, :, 0, :])
    with pytest.raises(RcaInputError):
        _to_iq(x, 4, 3, "cpu")
    with pytest.raises(RcaInputError):
        _to_iq(x, 3, 4, "cpu")


@pytest.mark.parametrize("half_index", [0, 1])
def test_product_prep_matches_oracle_prep(half_index):
    cfg = C.get("cfg1")
    G, T, P, Op = H.case(cfg, half_index)
    a = prepare_geometry(G, T)
    b = O.prepare(G, T)
    for i, k in enumerate(("ex", "ey", "ez")):
        np.testing.assert_array_equal(a["ele"][i], b[k])
    for i, k in enumerate(("vx", "vy", "vz")):
        np.testing.assert_array_equal(a["vs"][i], b[k])
    np.testing.assert_array_equal(a["d0"], b["d0"])
    np.testing.assert_array_equal(np.float32(a["aperture"]), b["ap"])


@pytest.mark.parametrize("fn,drop,ident", [
    ("bfVirtualSourceGeneric_RCA_ab_1", "AbLawRx", "AbLawRxMissing"),
    ("bfVirtualSourceGeneric_RCA_ab_2", "AbLawRx", "AbLawRxMissing"),
    ("bfVirtualSourceGeneric_RCA_Nosum_ab", "AbLawRx", "AbLawRxMissing"),
    ("bfVirtualSourceGeneric_RCA_TxRx", "AbLawRx", "AbLawRxMissing"),
    ("bfVirtualSourceGeneric_RCA_TxRx", "AbLawTx", "AbLawTxMissing"),
])
def test_missing_aberration_law(fn, drop, ident):
    import simu_rca_ab_amd.beamformer as B
    d, G, T, P, Op = _args()
    P = {**P, "AbLawRx": np.zeros(64, np.float32), "AbLawTx": np.zeros(64, np.float32)}
    P.pop(drop)
    with pytest.raises(RcaInputError) as e:
        getattr(B, fn)(d, G, T, P, Op)
    assert e.value.identifier == f"bfVirtualSourceGeneric_RCA:{ident}"


def test_law_layout_from_size():
    from simu_rca_ab_amd.plan import _law_mode
    n_ch, n_tx, n_pix = 64, 11, 300
    z = lambda n: torch.zeros(n)   # noqa: E731
    assert _law_mode(z(n_ch), None, n_ch, n_tx, n_pix) == "channel"
    assert _law_mode(z(n_ch * n_tx), None, n_ch, n_tx, n_pix) == "tx_channel"
    assert _law_mode(z(n_ch * n_pix), None, n_ch, n_tx, n_pix) == "pixel_channel"
    assert _law_mode(None, z(n_ch * n_pix), n_ch, n_tx, n_pix) == "pixel_channel"
    with pytest.raises(ValueError, match="matches no layout"):
        _law_mode(z(n_ch + 1), None, n_ch, n_tx, n_pix)
    with pytest.raises(ValueError, match="ambiguous"):
        _law_mode(z(n_ch * n_tx), None, n_ch, n_tx, n_tx)
