"""Snapshot and MAT-file formats on the host (mpsamd.checkpoint): round trips
of sites and metadata, the reference's write_data_file fields
(rydberg_density_matrix_gpu.jl:273-316) including a last site with a trailing
bond of 1, and the rejection of files of another format."""
import os

import numpy as np
import pytest

from mpsamd import checkpoint as CK


def _sites(rng, dims, d=9):
    return [rng.standard_normal((dims[j], d, dims[j + 1])) + 1j * rng.standard_normal((dims[j], d, dims[j + 1]))
            for j in range(len(dims) - 1)]


def test_snapshot_round_trip(rng, tmp_path):
    sites = _sites(rng, [1, 9, 16, 9, 1])
    path = str(tmp_path / "s.npz")
    CK.save_sites(path, sites, step=7, rng={"state": {"state": 2 ** 100 + 3, "inc": 5}}, p=[0.5, 0.25])
    got, meta = CK.load_sites(path)
    assert meta == {"step": 7, "rng": {"state": {"state": 2 ** 100 + 3, "inc": 5}}, "p": [0.5, 0.25]}
    assert all(np.array_equal(a, b) for a, b in zip(got, sites))
    assert [f for f in os.listdir(tmp_path) if "tmp" in f] == []


def test_snapshot_rejects_other_files(tmp_path):
    path = str(tmp_path / "x.npz")
    np.savez(path, format=np.array("something else"), dims=np.array([1, 1]), meta=np.array("{}"))
    with pytest.raises(ValueError):
        CK.load_sites(path)


def test_mat_data_file_round_trip(rng, tmp_path):
    sites = _sites(rng, [1, 9, 12, 1])
    obs = {"tr_rho": np.ones(4, complex), "e_pop": rng.standard_normal(4) + 0j,
           "e_pop_j": rng.standard_normal((3, 4)) + 0j, "I_r": rng.standard_normal(4) + 1j}
    path = str(tmp_path / "run_temp.mat")
    CK.write_data_file(path, 20, np.arange(4) * 0.025, obs, sites, times=np.ones(4),
                       params={"na": 3, "dt": 0.005, "d_max": 12})
    back = CK.read_data_file(path)
    assert back["i_last"] == 20 and float(back["dt"]) == 0.005 and int(back["na"]) == 3
    assert [a.shape for a in back["rho"]] == [a.shape for a in sites]
    assert all(np.array_equal(a, b) for a, b in zip(back["rho"], sites))
    for k, v in obs.items():
        assert np.allclose(back[k], v)
    assert str(back["TN"]) == "Complex{Float64}"
