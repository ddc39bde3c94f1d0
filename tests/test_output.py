"""The run output writer/reader in the reference's HDF5 layout (writeAllSpecies,
DiffusionModel.cpp:1096-1296; read_gmp_hdf5, python/gpgmp/io.py:189-291): group and
dataset names, the (W, H) dataset view of the x + y*W state buffer, the reader's
by-name dump order and its returned tuple."""
import numpy as np

from inchman_amd import output


def test_dump_names():
    # boost::lexical_cast<std::string>(float): shortest form with 9 significant digits
    assert output.dump_name(0.0) == "Dump_0"
    assert output.dump_name(100.0) == "Dump_100"
    assert output.dump_name(0.5) == "Dump_0.5"
    assert output.dump_name(0.1) == "Dump_0.100000001"
    assert output.dump_name(1e-5) == "Dump_9.99999975e-06"


def test_round_trip_layout(tmp_path):
    rng = np.random.default_rng(1)
    W, H, S = 6, 4, 2
    times = np.array([0.0, 2.0, 10.0, 4.5], np.float32)
    runs = [(rng.integers(0, 100, (4, S, H, W)).astype(np.int32), times) for _ in range(3)]
    fn = output.write_gmp(str(tmp_path / "out"), ["A", "B"], runs, W, H, parameters={"k2": 0.5})
    with np.load(fn) as f:
        assert int(f["Model/numRuns"][0]) == 3 and int(f["Model/numSpecies"][0]) == 2
        assert list(f["Model/Compartments/World"]) == [0, 0, W, H]
        assert float(f["Model/Parameters/k2"][0]) == 0.5
        d = f["Runs/1/Dump_4.5/B"]
        assert d.shape == (W, H)
        # the reference's dataset is the flat buffer x + y*W viewed as W rows of H
        flat = runs[1][0][3, 1].ravel()
        for i in range(W):
            for j in range(H):
                assert d[i, j] == flat[i * H + j]
    n, t, sp, nr = output.read_gmp(fn)
    assert nr == 3 and list(sp) == ["A", "B"] and n.shape == (3, 4, S, W, H)
    # h5py lists members by name: Dump_0, Dump_10, Dump_2, Dump_4.5
    order = [0, 2, 1, 3]
    np.testing.assert_array_equal(t, times[order])
    for r in range(3):
        for i, k in enumerate(order):
            for s in range(S):
                np.testing.assert_array_equal(n[r, i, s], runs[r][0][k, s].reshape(W, H))
