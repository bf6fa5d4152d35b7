"""Host-side parts of the scan driver: incident-angle channel selection and
the on-disk formats (no GPU)."""
import numpy as np
import pytest

import oracle.rgf_oracle as o
from rgf_amd.sweep import channels_at_angles, load_scan, save_scan


def test_angle_selection_returns_channels_at_their_own_angles():
    # oracle channels (same keys as the package's): asking for a channel's
    # own incident angle returns that channel (examples/example_script.m:45-48)
    for bc, ny in (("periodic", 40), ("Dirichlet", 33), (0.8, 25)):
        ch = o.setup_channels(ny, bc, 2 * np.pi / 9, 1.0, 1.0)
        prop = ch["L"]["ind_prop"] - 1
        ang = np.degrees(np.arctan2(ch["kydx"][prop].real, ch["L"]["kxdx"][prop].real))
        for q in range(len(prop)):
            got = channels_at_angles(ch, [ang[q]])
            assert np.allclose(ang[got - 1], ang[q])


def test_angle_outside_the_propagating_range_is_an_error():
    ch = o.setup_channels(20, "periodic", 2 * np.pi / 9, 1.0, 1.0)
    with pytest.raises(ValueError, match="outside"):
        channels_at_angles(ch, [89.9])


@pytest.mark.parametrize("ext", [".npz", ".mat"])
def test_scan_file_round_trip(tmp_path, ext):
    rng = np.random.default_rng(0)
    res = {"k0dx": np.array([0.5, 0.6, 0.7 + 0.01j]), "bc": ["periodic", "periodic", "periodic"],
           "S": [rng.standard_normal((3, 2)) + 1j * rng.standard_normal((3, 2)) for _ in range(3)],
           "in_index": [np.array([1, 2]), np.array([2, 3]), np.array([1, 3])]}
    path = str(tmp_path / ("scan" + ext))
    save_scan(res, path)
    back = load_scan(path)
    assert np.allclose(back["k0dx"], res["k0dx"])
    assert len(back["S"]) == 3
    for a, b in zip(back["S"], res["S"]):
        assert a.shape == b.shape and np.array_equal(a, b)
    for a, b in zip(back["in_index"], res["in_index"]):
        assert np.array_equal(np.asarray(a).ravel(), b)


def test_bloch_phases_are_stored_as_numbers(tmp_path):
    res = {"k0dx": np.array([0.5, 0.5]), "bc": [0.1, 0.7], "S": [np.eye(2), np.eye(2)],
           "in_index": [np.array([1, 2])] * 2}
    path = str(tmp_path / "b.npz")
    save_scan(res, path)
    assert np.allclose(load_scan(path)["bc"].ravel(), [0.1, 0.7])
