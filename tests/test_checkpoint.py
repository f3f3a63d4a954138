"""Checkpoint formats (CPU only): the DMFT checkpoint against the reference's
own round-trip test (tk-dmft/src/loop/mod.rs:490-512) and field names
(loop/mod.rs:453-468); the DMRG checkpoint file round trip."""
import json

import numpy as np

import tk_dmft as D
import tk_hip


def test_dmft_checkpoint_reference_roundtrip(tmp_path):
    cp = D.DMFTCheckpoint(5, False, "{}", [-1.0, 0.0, 1.0], [0.5, 0.5, 0.5], [-1.0, 0.0, 1.0], [0.1, 0.3, 0.1])
    path = tmp_path / "test_dmft_checkpoint.json"
    cp.write_to_file(path)
    loaded = D.DMFTCheckpoint.read_from_file(path)
    assert loaded.iteration == 5
    assert not loaded.converged
    assert loaded.bath_epsilon == [-1.0, 0.0, 1.0]
    assert loaded.spectral_values == [0.1, 0.3, 0.1]
    # the reference's serde field names, nothing else
    assert set(json.loads(path.read_text())) == {"iteration", "converged", "config_json", "bath_epsilon", "bath_v",
                                                  "spectral_omega", "spectral_values"}
    # written atomically: no temporary file left behind
    assert not (tmp_path / "test_dmft_checkpoint.tmp").exists()


def test_dmrg_checkpoint_file_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    bonds = [[((0, 0), 1)], [((1, 0), 2), ((0, 1), 2), ((-1, 3), 1)], [((1, 1), 3), ((2, 0), 1)], [((2, 2), 1)]]
    sites = [rng.standard_normal(k) for k in (4, 17, 9)]
    cp = tk_hip.DMRGCheckpoint(7, -4.258035207282882, {"max_bond_dim": 3.0, "matvecs": 42.0}, bonds, sites)
    path = tmp_path / "state.npz"
    cp.write_to_file(path)
    back = tk_hip.DMRGCheckpoint.read_from_file(path)
    assert back.sweep_index == 7 and back.energy == cp.energy and back.stats == cp.stats
    assert back.bonds == bonds
    assert all(np.array_equal(a, b) for a, b in zip(back.sites, sites))
    with np.load(path, allow_pickle=False) as z:  # plain arrays only
        assert z["bond_1_q"].dtype == np.int32 and z["bond_1_q"].shape == (3, 2)
    assert not (tmp_path / "state.npz.tmp").exists()
