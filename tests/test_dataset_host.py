"""Host helpers of the realigned dataset (CPU)."""
import numpy as np

from simu_rca_ab_amd.dataset import normalize_rf, realignment_grids


def test_realignment_grid_closed_form():
    lam = 0.2e-3
    bub = np.array([[1e-3, 2e-3], [-1e-3, 0.5e-3], [10e-3, 12e-3]])
    X, Y, Z = realignment_grids(bub, lam)
    assert X.shape == Y.shape == Z.shape == (17, 2)
    np.testing.assert_allclose(X[:, 1], 2e-3 + lam / 4)
    np.testing.assert_allclose(Y[:, 0], -1e-3 + lam / 4)
    # z = (1:17) * lam/4 + z_bubble - 2 lam: the column spans the bubble
    np.testing.assert_allclose(Z[:, 0], np.arange(1, 18) * lam / 4 + 10e-3 - 2 * lam)
    assert Z[7, 0] == 10e-3   # 8th depth (1-based) is the bubble


def test_normalize_rf():
    rf = np.array([[-3.0, 1.5], [0.75, 0.0]])
    np.testing.assert_allclose(normalize_rf(rf), rf / 3 * 32768)
    z = np.zeros(4)
    np.testing.assert_array_equal(normalize_rf(z), z)


def test_delays_ab_closed_form():
    """DelaysAb = cat(3, repmat(lawX, 128, 1, naAz), repmat(lawY.', 1, 128, naAz))
    (save_simus_sh.m:307-310): the X law along the second axis, the Y law
    along the first."""
    from simu_rca_ab_amd.dataset import delays_ab
    lx = np.arange(4) * 1e-9
    ly = -np.arange(4) * 2e-9
    d = delays_ab(lx, ly, 3)
    assert d.shape == (4, 4, 6) and d.dtype == np.float64
    for k in range(3):
        np.testing.assert_array_equal(d[:, :, k], np.tile(lx, (4, 1)))
        np.testing.assert_array_equal(d[:, :, 3 + k], np.tile(ly[:, None], (1, 4)))


def _sim_arrays(rng, n_scat, nz=17, ntx=21, nch=256, nfr=16, n_az=21):
    from simu_rca_ab_amd.dataset import delays_ab
    c = lambda: (rng.standard_normal((nz, ntx, nch, nfr)) +
                 1j * rng.standard_normal((nz, ntx, nch, nfr))).astype(np.complex64)
    return {"dataBf_al": c(), "dataBf_al_ab": c(), "dataBf_al_ab_c": c(),
            "DelaysAb": delays_ab(rng.standard_normal(128) * 1e-8, rng.standard_normal(128) * 1e-8, n_az),
            "Pos": rng.standard_normal((n_scat, 4, 672)),
            "Pos_bubble": rng.standard_normal((4, nfr))}


def test_mat_files_round_trip(tmp_path):
    """The six MAT-files of save_simus_sh.m:396-407, read back with
    scipy.io.loadmat by file name as Convert_job.py:13-14,31 reads them:
    same names, shapes, dtypes (complex single for the beamformed arrays,
    double otherwise) and values."""
    from scipy.io import loadmat
    from simu_rca_ab_amd.dataset import MAT_NAMES, load_simulation, save_simulation
    rng = np.random.default_rng(11)
    arrays = _sim_arrays(rng, 437)
    save_simulation(tmp_path / "1", arrays)
    assert sorted(p.name for p in (tmp_path / "1").iterdir()) == sorted(n + ".mat" for n in MAT_NAMES)
    back = load_simulation(tmp_path / "1")
    assert sorted(back) == sorted(MAT_NAMES)
    for n in MAT_NAMES:
        assert back[n].shape == np.shape(arrays[n]), n
        assert back[n].dtype == (np.complex64 if n.startswith("data") else np.float64), n
        np.testing.assert_array_equal(back[n], arrays[n])
    # the MATLAB variable name inside each file is the file name
    assert "dataBf_al_ab_c" in loadmat(tmp_path / "1" / "dataBf_al_ab_c.mat")


def test_job_arrays_layout(tmp_path):
    """Convert_job.py:27-84 layout: data* split into float64 _r/_i with a
    leading simulation axis, Pos padded to 500 scatterers, the rest stacked."""
    from simu_rca_ab_amd.dataset import job_arrays, save_simulation
    rng = np.random.default_rng(12)
    sims = []
    for i, n in enumerate((350, 500, 412)):
        a = _sim_arrays(rng, n, nfr=4)
        save_simulation(tmp_path / str(i + 1), a)
        sims.append(a)
    job = job_arrays([tmp_path / str(i + 1) for i in range(3)])
    assert sorted(job) == sorted(["dataBf_al_r", "dataBf_al_i", "dataBf_al_ab_r", "dataBf_al_ab_i",
                                  "dataBf_al_ab_c_r", "dataBf_al_ab_c_i", "DelaysAb", "Pos",
                                  "Pos_bubble"])
    assert job["Pos"].shape == (3, 500, 4, 672) and job["Pos"].dtype == np.float64
    assert job["dataBf_al_ab_r"].shape == (3, 17, 21, 256, 4)
    for i, a in enumerate(sims):
        n = a["Pos"].shape[0]
        np.testing.assert_array_equal(job["Pos"][i, :n], a["Pos"])
        assert not job["Pos"][i, n:].any()
        np.testing.assert_array_equal(job["dataBf_al_i"][i], np.imag(a["dataBf_al"]).astype(np.float64))
        np.testing.assert_array_equal(job["DelaysAb"][i], a["DelaysAb"])
