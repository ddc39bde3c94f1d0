"""Run output as an HDF5 file in the reference's layout, through the HDF5 C library.

The reference writes its results with the HDF5 C API (DiffusionModel::createOutputFile,
prepareOutputFile, writeAllSpecies, writeSingleTimeHDF5; DiffusionModel.cpp:1000-1296) and
reads them back with python/gpgmp/io.py:read_gmp_hdf5 (io.py:189-291).  This module makes
the same calls on the system's libhdf5 (1.8 or later) through ctypes, so the file is the
reference's file:

    /Model                      attributes numSpecies int[1], numRuns int[1]
    /Model/Species              variable-length strings [S]
    /Model/Compartments/<id>    int[4] = x0, y0, x1, y1 (Compartment::saveHdf5)
    [/Jobs/<job>]/Runs/<run>/Dump_<t>
                                attribute time float[1]
                                <species id>  int (stochastic) or float [W][H]: the state
                                              buffer x + y*W viewed as W rows of H
                                Fields/<name> float [W][H] (field parameters, if any)
                                Individuals/  <id> uint64[n] key pairs, <id>.DiffusivityX,
                                              .DiffusivityY, .DriftX, .DriftY float[n]

The library is looked up as $GPGMP_HDF5_LIB, then by ctypes.util.find_library("hdf5"), then
in /opt/conda/lib; without it the writer raises (the .npz form of inchman_amd.output needs
no library).
"""
import ctypes as C
import ctypes.util
import os

import numpy as np

from .output import dump_name

_CANDIDATES = ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
               "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so")

H5F_ACC_RDONLY, H5F_ACC_RDWR, H5F_ACC_TRUNC = 0, 1, 2
H5P_DEFAULT = 0
H5S_ALL = 0
H5_INDEX_NAME, H5_ITER_INC = 0, 0
H5T_VARIABLE = C.c_size_t(-1).value


class Hdf5Error(RuntimeError):
    pass


class _H5:
    """The handful of HDF5 C entry points the writer and reader use."""

    def __init__(self, path):
        L = C.CDLL(path)
        L.H5open.restype = C.c_int
        if L.H5open() < 0:
            raise Hdf5Error("H5open failed")
        maj, mnr, rel = C.c_uint(), C.c_uint(), C.c_uint()
        L.H5get_libversion(C.byref(maj), C.byref(mnr), C.byref(rel))
        self.version = (maj.value, mnr.value, rel.value)
        if self.version < (1, 8, 0):
            raise Hdf5Error("libhdf5 %d.%d.%d is older than 1.8" % self.version)
        hid = C.c_int64 if self.version >= (1, 10, 0) else C.c_int
        self.hid = hid
        hs = C.POINTER(C.c_uint64)
        sig = {
            "H5Fcreate": (hid, [C.c_char_p, C.c_uint, hid, hid]),
            "H5Fopen": (hid, [C.c_char_p, C.c_uint, hid]),
            "H5Fclose": (C.c_int, [hid]),
            "H5Gcreate2": (hid, [hid, C.c_char_p, hid, hid, hid]),
            "H5Gopen2": (hid, [hid, C.c_char_p, hid]),
            "H5Gclose": (C.c_int, [hid]),
            "H5Screate_simple": (hid, [C.c_int, hs, hs]),
            "H5Sclose": (C.c_int, [hid]),
            "H5Sget_simple_extent_ndims": (C.c_int, [hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [hid, hs, hs]),
            "H5Acreate2": (hid, [hid, C.c_char_p, hid, hid, hid, hid]),
            "H5Aopen": (hid, [hid, C.c_char_p, hid]),
            "H5Awrite": (C.c_int, [hid, hid, C.c_void_p]),
            "H5Aread": (C.c_int, [hid, hid, C.c_void_p]),
            "H5Aexists": (C.c_int, [hid, C.c_char_p]),
            "H5Aclose": (C.c_int, [hid]),
            "H5Dcreate2": (hid, [hid, C.c_char_p, hid, hid, hid, hid, hid]),
            "H5Dopen2": (hid, [hid, C.c_char_p, hid]),
            "H5Dwrite": (C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dread": (C.c_int, [hid, hid, hid, hid, hid, C.c_void_p]),
            "H5Dget_space": (hid, [hid]),
            "H5Dclose": (C.c_int, [hid]),
            "H5Dvlen_reclaim": (C.c_int, [hid, hid, hid, C.c_void_p]),
            "H5Tcopy": (hid, [hid]),
            "H5Tset_size": (C.c_int, [hid, C.c_size_t]),
            "H5Tclose": (C.c_int, [hid]),
            "H5Lexists": (C.c_int, [hid, C.c_char_p, hid]),
            "H5Eset_auto2": (C.c_int, [hid, C.c_void_p, C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.iter_cb = C.CFUNCTYPE(C.c_int, hid, C.c_char_p, C.c_void_p, C.c_void_p)
        L.H5Literate.restype = C.c_int
        L.H5Literate.argtypes = [hid, C.c_int, C.c_int, hs, self.iter_cb, C.c_void_p]
        L.H5Eset_auto2(0, None, None)  # errors come back as return codes, not stderr prints
        self.L = L
        g = lambda n: hid.in_dll(L, n).value  # noqa: E731 (the H5T_* macros read these globals)
        self.NATIVE_INT = g("H5T_NATIVE_INT_g")
        self.NATIVE_FLOAT = g("H5T_NATIVE_FLOAT_g")
        self.NATIVE_ULONG = g("H5T_NATIVE_ULONG_g")
        self.C_S1 = g("H5T_C_S1_g")

    def __getattr__(self, name):
        f = getattr(self.L, name)

        def checked(*a):
            r = f(*a)
            if (isinstance(r, int) and r < 0 and name not in ("H5Aexists", "H5Lexists")):
                raise Hdf5Error("%s failed (%d)" % (name, r))
            return r
        return checked

    @staticmethod
    def dims(*d):
        return (C.c_uint64 * len(d))(*d)


_h5 = None


def library_path():
    """Path of the libhdf5 this module uses, or None."""
    cands = [os.environ.get("GPGMP_HDF5_LIB"), ctypes.util.find_library("hdf5")] + list(_CANDIDATES)
    for p in cands:
        if not p:
            continue
        try:
            C.CDLL(p)
            return p
        except OSError:
            continue
    return None


def h5():
    global _h5
    if _h5 is None:
        p = library_path()
        if p is None:
            raise Hdf5Error("no HDF5 C library found (set GPGMP_HDF5_LIB)")
        _h5 = _H5(p)
    return _h5


def available():
    try:
        h5()
        return True
    except (Hdf5Error, OSError):
        return False


def _b(s):
    return s.encode() if isinstance(s, str) else s


class GmpHdf5Writer:
    """The reference's output calls in order: the constructor is createOutputFile (the
    /Model group, the runs' parent), begin_run(r) prepareOutputFile (group /Runs/<r>),
    write_dump(...) one writeAllSpecies call, close() the end of the run set."""

    def __init__(self, filename, species, width, height, num_runs, deterministic=False,
                 compartments=None, job_id=None):
        H = h5()
        self.H = H
        self.species = [str(s) for s in species]
        self.W, self.Hh = int(width), int(height)
        self.deterministic = bool(deterministic)
        self.f = H.H5Fcreate(_b(filename), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
        model = H.H5Gcreate2(self.f, b"/Model", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        self._attr_int(model, "numSpecies", len(self.species))
        st = H.H5Tcopy(H.C_S1)
        H.H5Tset_size(st, H5T_VARIABLE)
        sp = H.H5Screate_simple(1, H.dims(len(self.species)), None)
        ds = H.H5Dcreate2(model, b"Species", st, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        names = (C.c_char_p * len(self.species))(*[_b(s) for s in self.species])
        H.H5Dwrite(ds, st, H5S_ALL, H5S_ALL, H5P_DEFAULT, names)
        H.H5Dclose(ds)
        H.H5Sclose(sp)
        H.H5Tclose(st)
        self._attr_int(model, "numRuns", int(num_runs))
        comp = H.H5Gcreate2(model, b"Compartments", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        for name, b in (compartments or {}).items():
            self._dataset(comp, name, np.asarray(b, np.int32).reshape(4), H.NATIVE_INT)
        H.H5Gclose(comp)
        H.H5Gclose(model)
        parent = self.f
        self._close = []
        if job_id is not None:
            jobs = H.H5Gcreate2(self.f, b"Jobs", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
            parent = H.H5Gcreate2(jobs, _b(str(job_id)), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
            self._close += [parent, jobs]
        self.runs = H.H5Gcreate2(parent, b"Runs", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        self.run = None

    def _attr_int(self, obj, name, v):
        H = self.H
        sp = H.H5Screate_simple(1, H.dims(1), None)
        a = H.H5Acreate2(obj, _b(name), H.NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT)
        val = C.c_int(int(v))
        H.H5Awrite(a, H.NATIVE_INT, C.byref(val))
        H.H5Aclose(a)
        H.H5Sclose(sp)

    def _dataset(self, loc, name, arr, htype):
        H = self.H
        sp = H.H5Screate_simple(arr.ndim, H.dims(*arr.shape), None)
        ds = H.H5Dcreate2(loc, _b(name), htype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        H.H5Dwrite(ds, htype, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data_as(C.c_void_p))
        H.H5Dclose(ds)
        H.H5Sclose(sp)

    def begin_run(self, run_index):
        if self.run is not None:
            self.H.H5Gclose(self.run)
        self.run = self.H.H5Gcreate2(self.runs, _b(str(int(run_index))), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)

    def write_dump(self, time, state, fields=None, individuals=None):
        """state [S][H][W]; fields: name -> [H][W] floats; individuals: species id ->
        (key pairs uint64[n] = cell key | particle id << 32, properties float[4][n] in the
        order DiffusivityX, DiffusivityY, DriftX, DriftY)."""
        H = self.H
        if self.run is None:
            raise Hdf5Error("begin_run first")
        state = np.asarray(state)
        if state.shape != (len(self.species), self.Hh, self.W):
            raise ValueError("state must be [S][H][W] = %s" % ((len(self.species), self.Hh, self.W),))
        grp = H.H5Gcreate2(self.run, _b(dump_name(time)), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        sp = H.H5Screate_simple(1, H.dims(1), None)
        a = H.H5Acreate2(grp, b"time", H.NATIVE_FLOAT, sp, H5P_DEFAULT, H5P_DEFAULT)
        t = C.c_float(float(np.float32(time)))
        H.H5Awrite(a, H.NATIVE_FLOAT, C.byref(t))
        H.H5Aclose(a)
        H.H5Sclose(sp)
        dt, ht = (np.float32, H.NATIVE_FLOAT) if self.deterministic else (np.int32, H.NATIVE_INT)
        for s, sid in enumerate(self.species):
            # the x + y*W buffer written with dims (W, H)
            buf = np.ascontiguousarray(state[s], dt).reshape(-1).reshape(self.W, self.Hh)
            self._dataset(grp, sid, buf, ht)
        if fields:
            fg = H.H5Gcreate2(grp, b"Fields", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
            for name, arr in fields.items():
                buf = np.ascontiguousarray(arr, np.float32).reshape(-1).reshape(self.W, self.Hh)
                self._dataset(fg, name, buf, H.NATIVE_FLOAT)
            H.H5Gclose(fg)
        ig = H.H5Gcreate2(grp, b"Individuals", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
        for sid, (pos, props) in (individuals or {}).items():
            pos = np.ascontiguousarray(pos, np.uint64)
            if pos.size == 0:
                continue
            self._dataset(ig, sid, pos, H.NATIVE_ULONG)
            props = np.asarray(props, np.float32)
            for k, suffix in enumerate(("DiffusivityX", "DiffusivityY", "DriftX", "DriftY")):
                self._dataset(ig, sid + "." + suffix, np.ascontiguousarray(props[k][:pos.size]), H.NATIVE_FLOAT)
        H.H5Gclose(ig)
        H.H5Gclose(grp)

    def close(self):
        H = self.H
        if self.run is not None:
            H.H5Gclose(self.run)
            self.run = None
        if self.runs is not None:
            H.H5Gclose(self.runs)
            self.runs = None
        for g in self._close:
            H.H5Gclose(g)
        self._close = []
        if self.f is not None:
            H.H5Fclose(self.f)
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def write_gmp_hdf5(filename, species, runs, width, height, deterministic=False, compartments=None,
                   job_id=None):
    """runs: one (counts [n][S][H][W], times [n]) pair per run (optionally a third item:
    a list of per-dump dicts with "fields" / "individuals" for write_dump).  ".h5" is
    appended when missing, as the reference's file names carry it."""
    if os.path.splitext(filename)[1] != ".h5":
        filename += ".h5"
    with GmpHdf5Writer(filename, species, width, height, len(runs), deterministic, compartments, job_id) as w:
        for r, run in enumerate(runs):
            counts, times = run[0], run[1]
            extra = run[2] if len(run) > 2 else None
            w.begin_run(r)
            for d in range(len(times)):
                e = extra[d] if extra else {}
                w.write_dump(times[d], counts[d], fields=e.get("fields"), individuals=e.get("individuals"))
    return filename


def _children(H, gid):
    names = []

    @H.iter_cb
    def cb(g, name, info, data):
        names.append(name.decode())
        return 0
    idx = C.c_uint64(0)
    H.L.H5Literate(gid, H5_INDEX_NAME, H5_ITER_INC, C.byref(idx), cb, None)
    return names


def _read_dataset(H, loc, name, dtype, htype):
    ds = H.H5Dopen2(loc, _b(name), H5P_DEFAULT)
    sp = H.H5Dget_space(ds)
    nd = H.H5Sget_simple_extent_ndims(sp)
    dims = (C.c_uint64 * max(nd, 1))()
    H.H5Sget_simple_extent_dims(sp, dims, None)
    out = np.zeros(tuple(dims[:nd]), dtype)
    H.H5Dread(ds, htype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data_as(C.c_void_p))
    H.H5Sclose(sp)
    H.H5Dclose(ds)
    return out


def _read_attr_scalar(H, obj, name, ctype, htype):
    a = H.H5Aopen(obj, _b(name), H5P_DEFAULT)
    v = ctype()
    H.H5Aread(a, htype, C.byref(v))
    H.H5Aclose(a)
    return v.value


def read_gmp_hdf5(filenamebase, deterministic=False):
    """(n[runs, dumps, species, W, H], times[dumps], species ids, numRuns) as
    read_gmp_hdf5 (io.py:189-291) returns them: dumps of each run in name order, the
    times of the last run read, /Jobs/1/Runs when the file has a Jobs group."""
    H = h5()
    filename = os.path.splitext(filenamebase)[0] + ".h5"
    f = H.H5Fopen(_b(filename), H5F_ACC_RDONLY, H5P_DEFAULT)
    opened = []
    try:
        model = H.H5Gopen2(f, b"Model", H5P_DEFAULT)
        opened.append(model)
        num_runs = _read_attr_scalar(H, model, "numRuns", C.c_int, H.NATIVE_INT)
        # the species list (variable-length strings)
        ds = H.H5Dopen2(model, b"Species", H5P_DEFAULT)
        sp = H.H5Dget_space(ds)
        dims = (C.c_uint64 * 1)()
        H.H5Sget_simple_extent_dims(sp, dims, None)
        st = H.H5Tcopy(H.C_S1)
        H.H5Tset_size(st, H5T_VARIABLE)
        buf = (C.c_char_p * dims[0])()
        H.H5Dread(ds, st, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf)
        species = np.array([b.decode() for b in buf])
        H.H5Dvlen_reclaim(st, sp, H5P_DEFAULT, buf)
        H.H5Tclose(st)
        H.H5Sclose(sp)
        H.H5Dclose(ds)
        parent = f
        if H.L.H5Lexists(f, b"Jobs", H5P_DEFAULT) > 0:
            jobs = H.H5Gopen2(f, b"Jobs", H5P_DEFAULT)
            opened.append(jobs)
            parent = H.H5Gopen2(jobs, b"1", H5P_DEFAULT)
            opened.append(parent)
        runs = H.H5Gopen2(parent, b"Runs", H5P_DEFAULT)
        opened.append(runs)
        first = H.H5Gopen2(runs, b"0", H5P_DEFAULT)
        opened.append(first)
        num_dumps = len(_children(H, first))
        d0 = H.H5Gopen2(first, b"Dump_0", H5P_DEFAULT)
        opened.append(d0)
        dt, ht = (np.float32, H.NATIVE_FLOAT) if deterministic else (np.int32, H.NATIVE_INT)
        dx, dy = _read_dataset(H, d0, species[0], dt, ht).shape
        n = np.zeros((num_runs, num_dumps, len(species), dx, dy), float if deterministic else int)
        times = np.zeros(num_dumps)
        for run in range(num_runs):
            rg = H.H5Gopen2(runs, _b(str(run)), H5P_DEFAULT)
            for i, name in enumerate(_children(H, rg)):
                if i >= num_dumps:
                    break
                g = H.H5Gopen2(rg, _b(name), H5P_DEFAULT)
                times[i] = _read_attr_scalar(H, g, "time", C.c_float, H.NATIVE_FLOAT)
                for s, sid in enumerate(species):
                    n[run, i, s] = _read_dataset(H, g, sid, dt, ht)
                H.H5Gclose(g)
            H.H5Gclose(rg)
    finally:
        for g in reversed(opened):
            H.L.H5Gclose(g)
        H.L.H5Fclose(f)
    return n, times, species, num_runs


def read_individuals(filename, run, time, species_id):
    """(key pairs uint64[n], properties float[4][n]) of one dump's individual species."""
    H = h5()
    f = H.H5Fopen(_b(filename), H5F_ACC_RDONLY, H5P_DEFAULT)
    try:
        g = H.H5Gopen2(f, _b("/Runs/%d/%s/Individuals" % (run, dump_name(time))), H5P_DEFAULT)
        try:
            if H.L.H5Lexists(g, _b(species_id), H5P_DEFAULT) <= 0:
                return np.zeros(0, np.uint64), np.zeros((4, 0), np.float32)
            pos = _read_dataset(H, g, species_id, np.uint64, H.NATIVE_ULONG)
            props = np.stack([_read_dataset(H, g, species_id + "." + k, np.float32, H.NATIVE_FLOAT)
                              for k in ("DiffusivityX", "DiffusivityY", "DriftX", "DriftY")])
            return pos, props
        finally:
            H.L.H5Gclose(g)
    finally:
        H.L.H5Fclose(f)
