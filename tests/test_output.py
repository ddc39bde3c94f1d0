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


# ---- the HDF5 file itself (through the system's libhdf5) ------------------------------
import os  # noqa: E402
import shutil  # noqa: E402
import subprocess  # noqa: E402

import pytest  # noqa: E402

from inchman_amd import hdf5  # noqa: E402

needs_hdf5 = pytest.mark.skipif(not hdf5.available(), reason="no HDF5 C library")


def _h5dump():
    return shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


@needs_hdf5
def test_hdf5_file_layout(tmp_path):
    """The file as h5dump (an HDF5 tool independent of this writer) reads it: the Model
    attributes and species strings, compartments, the dump groups with their time
    attribute, and each species dataset as the x + y*W buffer with dims (W, H)."""
    rng = np.random.default_rng(3)
    W, H, S = 6, 4, 2
    times = np.array([0.0, 2.0, 4.5], np.float32)
    runs = [(rng.integers(0, 100, (3, S, H, W)).astype(np.int32), times) for _ in range(2)]
    fn = hdf5.write_gmp_hdf5(str(tmp_path / "out"), ["A", "B"], runs, W, H,
                             compartments={"World": (0, 0, W - 1, H - 1), "Box": (1, 1, 2, 3)})
    assert fn.endswith(".h5")
    tool = _h5dump()
    if tool is None:
        pytest.skip("no h5dump")

    def dump(*args):
        return subprocess.run([tool, *args, fn], capture_output=True, text=True, check=True).stdout

    head = dump("-H")
    for token in ('ATTRIBUTE "numRuns"', 'ATTRIBUTE "numSpecies"', 'DATASET "Species"', "STRSIZE H5T_VARIABLE",
                  'GROUP "Compartments"', 'DATASET "Box"', 'GROUP "Dump_4.5"', 'ATTRIBUTE "time"',
                  "H5T_IEEE_F32LE", "SIMPLE { ( 6, 4 ) / ( 6, 4 ) }", 'GROUP "Individuals"'):
        assert token in head, token
    assert '"A", "B"' in dump("-d", "/Model/Species")
    assert "(0): 2" in dump("-a", "/Model/numRuns")
    assert "(0): 1, 1, 2, 3" in dump("-d", "/Model/Compartments/Box")
    # values of one dataset, row by row as h5dump prints them
    txt = dump("-d", "/Runs/1/Dump_4.5/B", "-w", "0")
    flat = runs[1][0][2, 1].ravel()
    for i in range(W):
        row = ", ".join(str(v) for v in flat[i * H:(i + 1) * H])
        assert "(%d,0): %s" % (i, row) in txt, i


@needs_hdf5
def test_hdf5_round_trip_matches_npz_reader(tmp_path):
    rng = np.random.default_rng(4)
    W, H, S = 8, 16, 3
    times = np.array([0.0, 0.1, 10.0, 2.0], np.float32)
    runs = [(rng.integers(0, 1000, (4, S, H, W)).astype(np.int32), times) for _ in range(3)]
    fn = hdf5.write_gmp_hdf5(str(tmp_path / "a.h5"), ["U", "V", "W"], runs, W, H)
    got = hdf5.read_gmp_hdf5(fn)
    want = output.read_gmp(output.write_gmp(str(tmp_path / "a"), ["U", "V", "W"], runs, W, H))
    assert got[3] == want[3] == 3 and list(got[2]) == list(want[2])
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    # deterministic runs store floats
    fruns = [(rng.random((2, S, H, W)).astype(np.float32), times[:2])]
    fn2 = hdf5.write_gmp_hdf5(str(tmp_path / "d"), ["U", "V", "W"], fruns, W, H, deterministic=True)
    n, t, _, _ = hdf5.read_gmp_hdf5(fn2, deterministic=True)
    np.testing.assert_array_equal(n[0], fruns[0][0].reshape(2, S, W, H))


@needs_hdf5
def test_hdf5_fields_individuals_and_jobs(tmp_path):
    W, H = 4, 4
    st = np.arange(2 * H * W, dtype=np.int32).reshape(2, H, W)
    pos = (np.arange(5, dtype=np.uint64) << np.uint64(32)) | np.array([0, 3, 3, 9, 15], np.uint64)
    props = np.arange(20, dtype=np.float32).reshape(4, 5)
    # the reader sizes its arrays from Dump_0 (the initial state oclSolve writes first)
    extra = [{}, {"fields": {"slit": np.full((H, W), 0.5, np.float32)}, "individuals": {"B": (pos, props)}}]
    two = np.stack([st, st + 1])
    tt = np.array([0.0, 1.5], np.float32)
    fn = hdf5.write_gmp_hdf5(str(tmp_path / "j"), ["A", "B"], [(two, tt, extra)], W, H, job_id=1)
    n, t, sp, nr = hdf5.read_gmp_hdf5(fn)   # the reference reader takes /Jobs/1/Runs
    assert nr == 1 and t[1] == 1.5
    np.testing.assert_array_equal(n[0, 1], (st + 1).reshape(2, W, H))
    tool = _h5dump()
    if tool is not None:
        head = subprocess.run([tool, "-H", fn], capture_output=True, text=True, check=True).stdout
        for token in ('GROUP "Jobs"', 'GROUP "Fields"', 'DATASET "slit"', 'DATASET "B.DriftY"', "H5T_STD_U64LE"):
            assert token in head, token
    fn2 = hdf5.write_gmp_hdf5(str(tmp_path / "k"), ["A", "B"], [(two, tt, extra)], W, H)
    p2, q2 = hdf5.read_individuals(fn2, 0, 1.5, "B")
    np.testing.assert_array_equal(p2, pos)
    np.testing.assert_array_equal(q2, props)
