"""GPU parity of the realigned per-channel dataset (save_simus_sh.m:351-403):
every frame's 17-voxel column, beamformed by the NoSum kernel in one launch
over the union of the columns, equals one drop-in call per frame bit for bit
and the FP64 oracle to 1e-5 (rel-L2), with and without the per-line law."""
import numpy as np
import pytest

from oracle import rca_oracle as O
from simu_rca_ab_amd import configs as C
from simu_rca_ab_amd.beamformer import bfVirtualSourceGeneric_RCA_Nosum
from simu_rca_ab_amd.dataset import realigned_dataset, realigned_half, realignment_grids
from tests import helpers as H

pytestmark = pytest.mark.gpu

CFG = C.scaled(C.get("cfg1"), n_elem=32, n_sample=512)


def say(*a):
    print(*a, flush=True)


def _rf(n_frame, n_tx, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((2 * CFG.n_sample, CFG.n_elem, n_tx, n_frame)).astype(np.float32)


def _oracle_frame(rf_f, G, T, P, grids, f, ablaw=None):
    X, Y, Z = grids
    prep = O.prepare(G, T)
    grid = tuple(np.ascontiguousarray(a[:, f]).astype(np.float32) for a in (X, Y, Z))
    il = rf_f.ravel(order="F").astype(np.float64).copy()
    il[1::2] = -il[1::2]                      # IQ = RF(2n) - j RF(2n+1)
    params = H.oracle_params(CFG, CFG.halves[0])
    params = dict(params, receive_x=bool(T["flagReceiveX"]))
    out, _ = O.kernel_forward(il, grid, prep, params, O.INTERP_LANCZOS, 3, nosum=True,
                              ablaw=ablaw, ablaw_mode=1 if ablaw is not None else 0)
    return np.transpose(out[:, 0], (2, 1, 0))   # [nCh, nTx, Nz] -> [Nz, nTx, nCh]


@pytest.mark.parametrize("with_law", [False, True])
def test_realigned_half_vs_per_frame_and_oracle(dev, with_law):
    say(f"case realigned NoSum columns, 5 frames, law={with_law}")
    G, T, P, Op = H.case(CFG, 0)
    Op = dict(Op, interpolation="lanczos3")
    n_tx, n_fr = len(CFG.halves[0].theta), 5
    rf = _rf(n_fr, n_tx, 3)
    lam = CFG.c / CFG.fd
    rng = np.random.default_rng(4)
    bub = np.stack([rng.uniform(-2e-3, 2e-3, n_fr), rng.uniform(-2e-3, 2e-3, n_fr),
                    rng.uniform(8e-3, 14e-3, n_fr)])
    grids = realignment_grids(bub, lam)
    law = (rng.standard_normal(CFG.n_elem) * 20e-9).astype(np.float32) if with_law else None
    got = realigned_half(rf, G, T, P, Op, grids, ab_law=law, device=dev)
    assert got.shape == (17, n_tx, CFG.n_elem, n_fr)
    X, Y, Z = grids
    for f in range(n_fr):
        if not with_law:   # one drop-in call per frame, as the script does
            Gf = dict(G, gridBfX=X[:, f:f + 1, None], gridBfY=Y[:, f:f + 1, None],
                      gridBfZ=Z[:, f:f + 1, None])
            one, _ = bfVirtualSourceGeneric_RCA_Nosum(rf[:, :, :, f:f + 1].reshape(
                rf.shape[0], CFG.n_elem, 1, n_tx, order="F"), Gf, T, P, dict(Op, flagCompounding=False),
                device=dev)
            np.testing.assert_array_equal(got[..., f], np.squeeze(one, axis=(1, 2, 4)))
        ref = _oracle_frame(rf[..., f], G, T, P, grids, f, law)
        r = H.rel_l2(got[..., f], ref)
        say(f"frame {f}: rel-L2 vs oracle {r:.3e}")
        assert r <= 1e-5


def test_realigned_dataset_concatenates_halves(dev):
    say("case realigned dataset: both halves, laws, [17, nTx, 2 nCh, nFrame]")
    n_tx, n_fr = len(CFG.halves[0].theta), 3
    GX, TX_, P, Op = H.case(CFG, 0)
    GY, TY_, _, _ = H.case(CFG, 1)
    Op = dict(Op, interpolation="lanczos3")
    rfx, rfy = _rf(n_fr, n_tx, 5), _rf(n_fr, n_tx, 6)
    bub = np.array([[0.5e-3, -1e-3, 0.0], [0.2e-3, 0.8e-3, -0.3e-3], [10e-3, 11e-3, 12e-3]])
    lam = CFG.c / CFG.fd
    lx = np.full(CFG.n_elem, 10e-9, np.float32)
    ly = np.full(CFG.n_elem, -5e-9, np.float32)
    al, al_c = realigned_dataset(rfx, rfy, GX, GY, TX_, TY_, P, Op, bub, lam, lx, ly, device=dev)
    assert al.shape == (17, n_tx, 2 * CFG.n_elem, n_fr) and al_c.shape == al.shape
    grids = realignment_grids(bub, lam)
    np.testing.assert_array_equal(al[:, :, :CFG.n_elem], realigned_half(rfx, GX, TX_, P, Op, grids, device=dev))
    np.testing.assert_array_equal(al[:, :, CFG.n_elem:], realigned_half(rfy, GY, TY_, P, Op, grids, device=dev))
    assert H.rel_l2(al_c, al) > 1e-3     # the law moves the samples


def test_realigned_set_saved_and_read_back(dev, tmp_path):
    """One simulation's set (save_simus_sh.m:349-407) beamformed on the GPU,
    written as MAT-files and read back with scipy.io.loadmat: equal to the
    arrays in memory, and dataBf_al_ab / dataBf_al_ab_c are the aberrated RF
    beamformed without / with the law."""
    from simu_rca_ab_amd.dataset import (delays_ab, load_simulation, realigned_set,
                                         save_simulation)
    say("case realigned set -> MAT-files -> loadmat")
    n_tx, n_fr = len(CFG.halves[0].theta), 2
    GX, TX_, P, Op = H.case(CFG, 0)
    GY, TY_, _, _ = H.case(CFG, 1)
    Op = dict(Op, interpolation="lanczos3")
    rf = [_rf(n_fr, n_tx, s) for s in (21, 22, 23, 24)]
    bub = np.array([[0.5e-3, -1e-3], [0.2e-3, 0.8e-3], [10e-3, 11e-3]])
    lam = CFG.c / CFG.fd
    rng = np.random.default_rng(25)
    lx = (rng.standard_normal(CFG.n_elem) * 10e-9).astype(np.float32)
    ly = (rng.standard_normal(CFG.n_elem) * 10e-9).astype(np.float32)
    s = realigned_set(rf[0], rf[1], rf[2], rf[3], GX, GY, TX_, TY_, P, Op, bub, lam, lx, ly, device=dev)
    ab, ab_c = realigned_dataset(rf[2], rf[3], GX, GY, TX_, TY_, P, Op, bub, lam, lx, ly, device=dev)
    np.testing.assert_array_equal(s["dataBf_al_ab"], ab)
    np.testing.assert_array_equal(s["dataBf_al_ab_c"], ab_c)
    s["DelaysAb"] = delays_ab(lx, ly, n_tx)
    s["Pos"] = rng.standard_normal((400, 4, 672))
    s["Pos_bubble"] = np.vstack([bub, np.ones((1, n_fr))])
    save_simulation(tmp_path / "1", s)
    back = load_simulation(tmp_path / "1")
    for k, v in s.items():
        np.testing.assert_array_equal(back[k], v)
    assert back["dataBf_al"].shape == (17, n_tx, 2 * CFG.n_elem, n_fr)
