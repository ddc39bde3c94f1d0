"""Run output in the layout of the reference's HDF5 file.

The reference writes one HDF5 file per run set (DiffusionModel::writeAllSpecies and
writeSingleTimeHDF5, DiffusionModel.cpp:934-1296; layout in doc/source/hdf5.rst) and reads
it back with python/gpgmp/io.py:read_gmp_hdf5 (io.py:189-291).  inchman_amd.hdf5 writes that
file through the system's HDF5 C library; this module stores the same tree as a NumPy
``.npz`` archive whose keys are the HDF5 paths, for hosts without libhdf5:

    Model/numRuns               int32[1]
    Model/numSpecies            int32[1]
    Model/Species               the species ids
    Model/Compartments/World    int32[4]  (x, y, width, height) = (0, 0, W, H)
    Model/Parameters/<name>     float32[1]
    Runs/<run>/Dump_<t>/time    float32[1]
    Runs/<run>/Dump_<t>/<id>    int32 (stochastic) or float32 [W][H]

``<t>`` is the dump time as boost::lexical_cast<std::string>(float) prints it (9 significant
digits, shortest form).  A species dataset has the reference's dimensions (W, H) over the
state buffer x + y*W, i.e. the [H][W] counts reinterpreted row-major as W rows of H.

read_gmp returns what read_gmp_hdf5 returns: (n[runs, dumps, species, W, H], times[dumps],
species ids, numRuns), with the dumps of every run in the order h5py lists the group
members (by name, so "Dump_10" comes before "Dump_2"), and the times of the last run read,
as the reference reader leaves them.
"""
import numpy as np


def dump_name(time):
    """Group name of a dump: "Dump_" + the float time with 9 significant digits."""
    return "Dump_" + format(float(np.float32(time)), ".9g")


def write_gmp(filename, species, runs, width, height, deterministic=False, parameters=None):
    """runs: one (counts [n][S][H][W], times [n]) pair per run, e.g. from
    Solver.getOutput() (with the final state appended by the caller, as oclSolve writes it
    after its loop).  Writes ``filename`` (.npz appended if missing)."""
    species = [str(s) for s in species]
    out = {
        "Model/numRuns": np.array([len(runs)], np.int32),
        "Model/numSpecies": np.array([len(species)], np.int32),
        "Model/Species": np.array(species),
        "Model/Compartments/World": np.array([0, 0, width, height], np.int32),
    }
    for k, v in (parameters or {}).items():
        out["Model/Parameters/" + k] = np.array([v], np.float32)
    dt = np.float32 if deterministic else np.int32
    for r, (counts, times) in enumerate(runs):
        counts = np.asarray(counts)
        assert counts.shape[1:] == (len(species), height, width) and len(times) == counts.shape[0]
        for d in range(counts.shape[0]):
            g = "Runs/%d/%s/" % (r, dump_name(times[d]))
            out[g + "time"] = np.array([times[d]], np.float32)
            for s, sid in enumerate(species):
                out[g + sid] = np.ascontiguousarray(counts[d, s], dt).reshape(-1).reshape(width, height)
    if not filename.endswith(".npz"):
        filename += ".npz"
    np.savez(filename, **out)
    return filename


def read_gmp(filename, deterministic=False):
    """(n[runs, dumps, species, W, H], times[dumps], species ids, numRuns), as
    read_gmp_hdf5 (python/gpgmp/io.py:189-291) returns them."""
    if not filename.endswith(".npz"):
        filename += ".npz"
    with np.load(filename, allow_pickle=False) as f:
        keys = list(f.keys())
        num_runs = int(f["Model/numRuns"][0])
        species = np.array([str(s) for s in f["Model/Species"]])

        def dumps_of(run):
            pre = "Runs/%d/" % run
            return sorted({k[len(pre):].split("/")[0] for k in keys if k.startswith(pre)})

        first = dumps_of(0)
        dx, dy = f["Runs/0/%s/%s" % (first[0], species[0])].shape
        n = np.zeros((num_runs, len(first), len(species), dx, dy), float if deterministic else int)
        times = np.zeros(len(first))
        for run in range(num_runs):
            for i, name in enumerate(dumps_of(run)):
                if i >= len(first):
                    break
                g = "Runs/%d/%s/" % (run, name)
                times[i] = float(f[g + "time"][0])
                for s, sid in enumerate(species):
                    n[run, i, s] = f[g + sid]
    return n, times, species, num_runs
