"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_threefry.so")

MAXS, MAXR, MAXF = 16, 32, 8


class RxTable(C.Structure):
    _fields_ = [
        ("num_species", C.c_int), ("num_reactions", C.c_int),
        ("lit_is_int", C.c_int * MAXR), ("lit_int", C.c_int32 * MAXR), ("lit_dbl", C.c_double * MAXR),
        ("n_factors", C.c_int * MAXR),
        ("factor_species", (C.c_int * MAXF) * MAXR), ("factor_offset", (C.c_int * MAXF) * MAXR),
        ("net", (C.c_int * MAXS) * MAXR),
        ("localized", C.c_uint32), ("act", C.POINTER(C.c_uint16)),
        ("hazards_fn", C.c_void_p), ("rk_fn", C.c_void_p), ("aqss_f32_fn", C.c_void_p), ("aqss_f64_fn", C.c_void_p),
        ("grid_w", C.c_int),
    ]


class Model(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("S", C.c_int),
        ("physical_length", C.c_float), ("p0", C.c_float), ("inhomogeneous", C.c_int),
        ("diffusion", C.POINTER(C.c_float)),
        ("Dx", C.POINTER(C.c_float)), ("Dy", C.POINTER(C.c_float)),
        ("rx", C.POINTER(C.c_float)), ("ry", C.POINTER(C.c_float)),
        ("boundary_mask", C.c_int * 4), ("source_mask", C.c_int * 4),
        ("max_time", C.c_float),
    ]


class RddParams(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("h", C.c_double), ("dt", C.c_double),
                ("D0u", C.c_double), ("D0v", C.c_double), ("k", C.c_double * 4)]


class ParticleParams(C.Structure):
    _fields_ = [("side", C.c_int), ("D", C.c_float), ("dt", C.c_float), ("vortex_amp", C.c_float),
                ("p_annihilate", C.c_float), ("seed", C.c_uint32)]


class IndParams(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("physical_length", C.c_float), ("boundary_mask", C.c_int32 * 4),
                ("species", C.c_int)]


class NewRule(C.Structure):
    _fields_ = [("mode", C.c_int32 * 4), ("a", C.c_double * 4), ("b", C.c_double * 4)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_u01.restype = C.c_float
        L.oracle_u01.argtypes = [C.c_uint32]
        L.oracle_logf.restype = C.c_float
        L.oracle_logf.argtypes = [C.c_float]
        L.oracle_logd.restype = C.c_double
        L.oracle_logd.argtypes = [C.c_double]
        for f in ("oracle_stay_probability", "oracle_header_float"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [C.c_float]
        L.oracle_homogeneous_dt.restype = C.c_float
        L.oracle_homogeneous_dt.argtypes = [C.c_float] * 3
        L.oracle_inhomogeneous_cell_dt.restype = C.c_float
        L.oracle_inhomogeneous_cell_dt.argtypes = [C.c_float] * 3
        L.oracle_inhomogeneous_dt.restype = C.c_float
        L.oracle_inhomogeneous_dt.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_float]
        L.oracle_ssa.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int32, C.c_float, C.c_int]
        L.oracle_hom_diffuse.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int32, C.c_float,
                                         C.c_void_p, C.c_void_p, C.c_int]
        L.oracle_inhom_sweep.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int32,
                                         C.c_float, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int]
        L.oracle_fpe_cell_dt.restype = C.c_float
        L.oracle_fpe_cell_dt.argtypes = [C.c_float] * 4
        L.oracle_fpe_sweep.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int32,
                                       C.c_float, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int]
        L.oracle_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_int,
                                 C.c_void_p]
        L.oracle_run_from.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_int,
                                      C.c_void_p]
        L.oracle_rdd_grad.restype = C.c_double
        L.oracle_rdd_loss.restype = C.c_double
        L.oracle_particle_step.restype = C.c_int
        L.oracle_expd.restype = C.c_double
        L.oracle_expd.argtypes = [C.c_double]
        L.oracle_expm1d.restype = C.c_double
        L.oracle_expm1d.argtypes = [C.c_double]
        L.oracle_ind_cell_dt.restype = C.c_float
        L.oracle_ind_cell_dt.argtypes = [C.c_float] * 3
        L.oracle_ind_sweep.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int32,
                                                          C.c_float]
        L.oracle_ind_constants.argtypes = [C.c_void_p] * 5 + [C.c_int, C.c_void_p, C.c_int] + [C.c_void_p] * 5
        L.oracle_ind_update.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_void_p, C.c_int] + [C.c_void_p] * 9
        L.oracle_ind_run.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_uint32,
                                                        C.c_void_p, C.c_void_p]
        L.oracle_det_aqss_f32.restype = C.c_int
        L.oracle_det_aqss_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_float, C.c_int]
        L.oracle_det_aqss_f64.restype = C.c_int
        L.oracle_det_aqss_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int]
        L.oracle_run_output.restype = C.c_int
        L.oracle_run_output.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_uint32, C.c_int, C.c_float, C.c_float,
                                                           C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_ind_maintain.restype = C.c_int
        L.oracle_ind_maintain.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int32,
                                                             C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own Threefry built from its vendored header, when present."""
    if os.path.exists(REF_LIB):
        L = C.CDLL(REF_LIB)
        return L
    return None


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def threefry(c0, c1, k0, k1, which="oracle"):
    L = lib() if which == "oracle" else ref_lib()
    fn = L.oracle_threefry2x32_20 if which == "oracle" else L.ref_threefry2x32_20
    a, b = C.c_uint32(), C.c_uint32()
    fn(C.c_uint32(c0), C.c_uint32(c1), C.c_uint32(k0), C.c_uint32(k1), C.byref(a), C.byref(b))
    return a.value, b.value


def u01(w):
    return lib().oracle_u01(C.c_uint32(w))


def logf(y):
    return lib().oracle_logf(float(y))


def seeds(seed0, n):
    out = np.zeros(n, np.int32)
    lib().oracle_mt19937_seeds(C.c_uint32(seed0), int(n), _p(out))
    return out


def compile_reactions(S, rates, reactant, product):
    """rates[R], reactant/product [R][S] stoichiometry."""
    R = len(rates)
    t = RxTable()
    r = np.ascontiguousarray(rates, np.float32)
    rs = np.ascontiguousarray(reactant, np.int32).reshape(R, S) if R else np.zeros((1, S), np.int32)
    ps = np.ascontiguousarray(product, np.int32).reshape(R, S) if R else np.zeros((1, S), np.int32)
    rc = lib().oracle_compile_reactions(S, R, _p(r), _p(rs), _p(ps), C.byref(t))
    assert rc == 0, rc
    return t


def table_from_model(model):
    S = model.numSpecies()
    R = len(model.reactions)
    rs = np.zeros((R, S), np.int32)
    ps = np.zeros((R, S), np.int32)
    for i, r in enumerate(model.reactions):
        for sp, n in r.reactants.items():
            rs[i, sp.index] += n
        for sp, n in r.products.items():
            ps[i, sp.index] += n
    t = compile_reactions(S, [r.rate for r in model.reactions], rs, ps)
    localized = 0
    for i, r in enumerate(model.reactions):
        if getattr(r, "compartments", None):
            localized |= 1 << i
    if localized:
        # the cells' active-reaction bits from the reactions' compartments
        # (Reaction::reactionMask, Reaction.cpp:216-235), x + y*W order
        act = np.zeros((model.height, model.width), np.uint16)
        for i, r in enumerate(model.reactions):
            if not (localized >> i) & 1:
                act |= np.uint16(1 << i)
                continue
            for cp in r.compartments:
                x0, y0, x1, y1 = cp.getBounds()
                for y in range(max(y0, 0), min(y1, model.height - 1) + 1):
                    for x in range(max(x0, 0), min(x1, model.width - 1) + 1):
                        act[y, x] |= np.uint16(1 << i)
        t._act = np.ascontiguousarray(act.ravel())
        t.localized = localized
        t.act = t._act.ctypes.data_as(C.POINTER(C.c_uint16))
    t.grid_w = model.width
    if any(getattr(r, "customLaw", None) for r in model.reactions):
        from . import lawgen

        lawgen.attach(t, model)
    return t


def hazards(table, x, cell=0):
    """Cumulative hazards [h0, h1..hR] of the state x[S] in cell `cell`."""
    xs = np.ascontiguousarray(x, np.int32)
    h = np.zeros(table.num_reactions + 1, np.float32)
    lib().oracle_hazards(C.byref(table), _p(xs), int(cell), h.ctypes.data_as(C.POINTER(C.c_float)))
    return h


def ssa(state, table, seed, deltat, nthreads=0):
    st = np.ascontiguousarray(state, np.int32).copy()
    S, H, W = st.shape
    lib().oracle_ssa(_p(st), W, H, C.byref(table), int(seed), float(deltat), nthreads)
    return st


def hom_diffuse(state, species, seed, stay, bm=(0, 0, 0, 0), sm=(0, 0, 0, 0), nthreads=0):
    st = np.ascontiguousarray(state, np.int32).copy()
    S, H, W = st.shape
    b = np.array(bm, np.int32)
    s = np.array(sm, np.int32)
    lib().oracle_hom_diffuse(_p(st), W, H, species, int(seed), float(stay), _p(b), _p(s), nthreads)
    return st


def inhom_sweep(state, species, axis, seed, dt, Dax, rax, h, bm=(0, 0, 0, 0), nthreads=0):
    st = np.ascontiguousarray(state, np.int32).copy()
    S, H, W = st.shape
    D = np.ascontiguousarray(Dax, np.float32)
    r = np.ascontiguousarray(rax, np.float32)
    b = np.array(bm, np.int32)
    lib().oracle_inhom_sweep(_p(st), W, H, species, axis, int(seed), float(dt), _p(D), _p(r), float(h),
                             _p(b), nthreads)
    return st


def fpe_cell_dt(D, r, h, stay):
    return lib().oracle_fpe_cell_dt(float(D), float(r), float(h), float(stay))


def fpe_sweep(state, species, axis, seed, dt, Dax, rax, h, bm=(0, 0, 0, 0), nthreads=0):
    st = np.ascontiguousarray(state, np.int32).copy()
    S, H, W = st.shape
    D = np.ascontiguousarray(Dax, np.float32)
    r = np.ascontiguousarray(rax, np.float32)
    b = np.array(bm, np.int32)
    lib().oracle_fpe_sweep(_p(st), W, H, species, axis, int(seed), float(dt), _p(D), _p(r), float(h),
                           _p(b), nthreads)
    return st


def stay_probability(p0):
    return lib().oracle_stay_probability(float(p0))


def header_float(v):
    return lib().oracle_header_float(float(v))


def homogeneous_dt(p0, h, D):
    return lib().oracle_homogeneous_dt(float(p0), float(h), float(D))


def inhomogeneous_dt(D, r, h):
    D = np.ascontiguousarray(D, np.float32)
    r = np.ascontiguousarray(r, np.float32)
    return lib().oracle_inhomogeneous_dt(_p(D), _p(r), D.size, float(h))


def _model_struct(model):
    S = model.numSpecies()
    m = Model()
    m.W, m.H, m.S = model.width, model.height, S
    m.physical_length, m.p0 = model.physical_length, model.p0
    # 0 homogeneous, 1 inhomogeneous, 2 inhomogeneous with the Fokker-Planck discretisation
    m.inhomogeneous = {1: 1, 6: 2}.get(model.solver, 0)
    diff = np.array([sp.diffusionConstant for sp in model.species], np.float32)
    keep = [diff]
    m.diffusion = diff.ctypes.data_as(C.POINTER(C.c_float))
    if model.fields is not None:
        for name, arr in zip(("Dx", "Dy", "rx", "ry"), model.fields):
            a = np.ascontiguousarray(arr, np.float32)
            keep.append(a)
            setattr(m, name, a.ctypes.data_as(C.POINTER(C.c_float)))
    for i in range(4):
        m.boundary_mask[i] = model.boundary_mask[i]
        m.source_mask[i] = model.source_mask[i]
    m.max_time = model.max_time
    return m, keep


def run_output(model, state, steps, seed0, interval, max_time=float("inf"), max_dumps=64, nthreads=0):
    """The host loop with output events every `interval` and the stop at max_time.
    Returns (final state, sim time, iterations, dumps [n][S][H][W], dump times, writes)."""
    st = np.ascontiguousarray(state, np.int32).copy()
    m, keep = _model_struct(model)  # noqa: F841 (keeps the field arrays alive)
    t = table_from_model(model)
    dumps = np.zeros((max_dumps,) + st.shape, np.int32)
    times = np.zeros(max_dumps, np.float32)
    sim = C.c_float()
    it = C.c_int64()
    w = lib().oracle_run_output(C.byref(m), C.byref(t), _p(st), int(steps), C.c_uint32(seed0), nthreads,
                                C.c_float(interval), C.c_float(max_time), _p(dumps), _p(times), int(max_dumps),
                                C.byref(it), C.byref(sim))
    k = min(w, max_dumps)
    return st, sim.value, it.value, dumps[:k], times[:k], w


def run(model, state, steps, seed0, nthreads=0, skip=0):
    """The host loop (DiffusionModel::oclSolve) on the oracle; model is a
    inchman_amd.gpgmp.DiffusionModel.  skip > 0: `state` is the state after `skip`
    iterations of the loop from seed0, which only advance the schedule and the seeds here,
    and `steps` more iterations run."""
    st = np.ascontiguousarray(state, np.int32).copy()
    m, keep = _model_struct(model)  # noqa: F841 (keeps the field arrays alive)
    t = table_from_model(model)
    sim = C.c_float()
    if skip:
        lib().oracle_run_from(C.byref(m), C.byref(t), _p(st), int(skip), int(steps), C.c_uint32(seed0), nthreads,
                              C.byref(sim))
    else:
        lib().oracle_run(C.byref(m), C.byref(t), _p(st), int(steps), C.c_uint32(seed0), nthreads, C.byref(sim))
    return st, sim.value


# ---- deterministic / RDD --------------------------------------------------------------
def rdd_params(W, H, h, dt, D0u, D0v, k):
    p = RddParams()
    p.W, p.H, p.h, p.dt, p.D0u, p.D0v = W, H, h, dt, D0u, D0v
    for i in range(4):
        p.k[i] = k[i]
    return p


def rdd_forward(p, D, u, v, steps, nthreads=0):
    D = np.ascontiguousarray(D, np.float64)
    u = np.ascontiguousarray(u, np.float64).copy()
    v = np.ascontiguousarray(v, np.float64).copy()
    lib().oracle_rdd_forward(C.byref(p), _p(D), _p(u), _p(v), int(steps), nthreads)
    return u, v


def rdd_loss(u, v, ut, vt):
    arr = [np.ascontiguousarray(a, np.float64) for a in (u, v, ut, vt)]
    return lib().oracle_rdd_loss(arr[0].size, *[_p(a) for a in arr])


def rdd_grad(p, D, u0, v0, ut, vt, steps, ckpt_every, nthreads=0):
    arr = [np.ascontiguousarray(a, np.float64) for a in (D, u0, v0, ut, vt)]
    gD = np.zeros_like(arr[0])
    gk = np.zeros(4, np.float64)
    loss = lib().oracle_rdd_grad(C.byref(p), *[_p(a) for a in arr], int(steps), int(ckpt_every), _p(gD),
                                 _p(gk), nthreads)
    return loss, gD, gk


def det_hom_diffuse(state, D, dt, h):
    st = np.ascontiguousarray(state, np.float64).copy()
    S, H, W = st.shape
    d = np.ascontiguousarray(D, np.float64)
    lib().oracle_det_hom_diffuse(_p(st), W, H, S, _p(d), C.c_double(dt), C.c_double(h))
    return st


def det_inhom_sweep(state, species, axis, dt, Dax, rax, h):
    st = np.ascontiguousarray(state, np.float64).copy()
    S, H, W = st.shape
    D = np.ascontiguousarray(Dax, np.float64)
    r = np.ascontiguousarray(rax, np.float64)
    lib().oracle_det_inhom_sweep(_p(st), W, H, species, axis, C.c_double(dt), _p(D), _p(r), C.c_double(h))
    return st


# ---- particles ------------------------------------------------------------------------
def stable_sort_pairs(keys, vals):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().oracle_stable_sort_pairs(_p(k), _p(v), k.size)
    return k, v


def cell_offsets(keys, ncells):
    k = np.ascontiguousarray(keys, np.uint32)
    off = np.zeros(ncells + 1, np.uint32)
    lib().oracle_cell_offsets(_p(k), k.size, ncells, _p(off))
    return off


def cell_left_right(keys, ncells):
    k = np.ascontiguousarray(keys, np.uint32)
    left = np.zeros(ncells, np.uint32)
    right = np.zeros(ncells, np.uint32)
    lib().oracle_cell_left_right(_p(k), k.size, ncells, _p(left), _p(right))
    return left, right


def particle_params(side, D, dt, vortex_amp, p_annihilate, seed):
    p = ParticleParams()
    p.side, p.D, p.dt, p.vortex_amp, p.p_annihilate, p.seed = side, D, dt, vortex_amp, p_annihilate, seed
    return p


def particle_steps(p, step0, steps, x, y, ids, nthreads=0):
    x = np.ascontiguousarray(x, np.float32).copy()
    y = np.ascontiguousarray(y, np.float32).copy()
    ids = np.ascontiguousarray(ids, np.uint32).copy()
    n = x.size
    keys = np.zeros(max(n, 1), np.uint32)
    off = np.zeros(p.side * p.side + 1, np.uint32)
    for s in range(step0, step0 + steps):
        n = lib().oracle_particle_step(C.byref(p), s, n, _p(x), _p(y), _p(ids), _p(keys), _p(off), nthreads)
    return x[:n], y[:n], ids[:n], keys[:n], off


# ---- individual-species lattice hops ---------------------------------------------------
def ind_params(W, H, physical_length, bm=(0, 0, 0, 0), species=0):
    p = IndParams()
    p.W, p.H, p.physical_length, p.species = W, H, physical_length, species
    for i in range(4):
        p.boundary_mask[i] = bm[i]
    return p


def _props(props):
    pr = np.ascontiguousarray(props, np.float32)
    assert pr.ndim == 2 and pr.shape[0] == 4
    return pr, pr.shape[1]


def ind_sweep(p, keys, vals, left, right, props, axis, seed, dt):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32)
    le = np.ascontiguousarray(left, np.uint32)
    ri = np.ascontiguousarray(right, np.uint32)
    pr, cap = _props(props)
    lib().oracle_ind_sweep(C.byref(p), _p(k), _p(v), _p(le), _p(ri), k.size, _p(pr), cap, axis, seed, dt)
    return k


def ind_constants(p, keys, vals, left, right, props):
    A = p.W * p.H
    k = np.ascontiguousarray(keys, np.uint32)
    v = np.ascontiguousarray(vals, np.uint32)
    le = np.ascontiguousarray(left, np.uint32)
    ri = np.ascontiguousarray(right, np.uint32)
    pr, cap = _props(props)
    out = [np.zeros(A, np.float32) for _ in range(4)] + [np.zeros(A, np.int32)]
    lib().oracle_ind_constants(C.byref(p), _p(k), _p(v), _p(le), _p(ri), k.size, _p(pr), cap, *[_p(o) for o in out])
    return out


def ind_update(p, keys, vals, props):
    A = p.W * p.H
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    pr, cap = _props(props)
    le, ri = np.zeros(A, np.uint32), np.zeros(A, np.uint32)
    out = [np.zeros(A, np.float32) for _ in range(4)] + [np.zeros(A, np.int32)]
    dtx, dty = C.c_float(), C.c_float()
    lib().oracle_ind_update(C.byref(p), _p(k), _p(v), k.size, _p(pr), cap, _p(le), _p(ri), *[_p(o) for o in out],
                            C.byref(dtx), C.byref(dty))
    return k, v, le, ri, out, dtx.value, dty.value


def ind_run(p, keys, vals, props, steps, seed0):
    A = p.W * p.H
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    pr, cap = _props(props)
    count = np.zeros(A, np.int32)
    t = C.c_float()
    lib().oracle_ind_run(C.byref(p), _p(k), _p(v), k.size, _p(pr), cap, steps, C.c_uint32(seed0), _p(count),
                         C.byref(t))
    return k, v, count, t.value


def new_rule(mode=(0, 0, 0, 0), a=(0, 0, 0, 0), b=(0, 0, 0, 0)):
    r = NewRule()
    for i in range(4):
        r.mode[i], r.a[i], r.b[i] = int(mode[i]), float(a[i]), float(b[i])
    return r


def ind_maintain(p, keys, vals, n, props, goal, seed, rule=None):
    """maintainPopulation on the reference's full list: keys/vals [cap] (live sorted in
    [0, n), dead ids after), props [4][cap] by id.  Returns (status, keys, vals, props, n_new)."""
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    pr = np.ascontiguousarray(props, np.float32).copy()
    assert pr.shape == (4, k.size) and v.size == k.size
    g = np.ascontiguousarray(goal, np.int32)
    assert g.size == p.W * p.H
    n_out = C.c_int(n)
    st = lib().oracle_ind_maintain(C.byref(p), _p(k), _p(v), n, k.size, _p(pr), _p(g), seed,
                                   C.byref(rule) if rule is not None else None, C.byref(n_out))
    return st, k, v, pr, n_out.value


def det_aqss(state, table, dtg, dtype=np.float32, fixed=False):
    """alpha-QSS reactions over dtg on state [S][H][W] in float (the reference's
    promotions) or double; fixed: the predictor's Pade factor in the corrector too.
    Returns (new state, largest inner-pass count)."""
    st = np.ascontiguousarray(state, dtype).copy()
    S, H, W = st.shape
    if dtype == np.float32:
        it = lib().oracle_det_aqss_f32(_p(st), W, H, C.byref(table), float(dtg), int(fixed))
    else:
        it = lib().oracle_det_aqss_f64(_p(st), W, H, C.byref(table), float(dtg), int(fixed))
    return st, it


def det_rk4(state, table, dt):
    """deterministic_performReaction_rk4 (Deterministic.cl:143-238) in double."""
    st = np.ascontiguousarray(state, np.float64).copy()
    S, H, W = st.shape
    lib().oracle_det_rk4(_p(st), W, H, C.byref(table), C.c_double(dt))
    return st


def det_run(model, state, steps, flags=0):
    """DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1479-1603) with a deterministic
    solver, in double: the multi-rate schedule of the stochastic loop with the deterministic
    time steps (homogeneous: dt = min_s (float)1e-2 h^2/(4 D_s) for every species,
    HomogeneousSolver.cpp:97-121; inhomogeneous: the per-cell minimum per axis,
    InhomogeneousSolver.cpp:189-212), doReactions = RK4 or alpha-QSS over the iteration's dt
    (Solver.cpp:250-266), diffuse = the Jacobi stencil of the due species and axes
    (HomogeneousSolver.cpp:157-160, InhomogeneousSolver.cpp:282-286, 374-378).  flags bit 1:
    the reference's own behaviour (no reactions for the inhomogeneous types; a homogeneous
    diffuse call moves every species); bit 0: the alpha-QSS corrector with the predictor's
    Pade factor.  Returns (state fp64 [S][H][W], simulation time as the float clock)."""
    f32 = np.float32
    st = np.ascontiguousarray(state, np.float64).copy()
    S, H, W = st.shape
    A = W * H
    inh = model.solver in (3, 5)
    aqss = model.solver in (4, 5)
    quirks = bool(flags & 2)
    tab = table_from_model(model)
    R = len(model.reactions)
    if not inh:
        h = f32(model.physical_length) / f32(model.width)
        dts = []
        for sp in model.species:
            D = f32(sp.diffusionConstant)
            dts.append(f32(f32(f32(f32(1e-2) * h) * h) / f32(f32(4.0) * D)) if D > 0 else f32(model.max_time))
        common = min(dts)
        dtx = [common] * S
        dty = dtx
        D_h = np.array([sp.diffusionConstant for sp in model.species], np.float32).astype(np.float64)
    else:
        Dx, Dy, rx, ry = (np.ascontiguousarray(a, np.float32) for a in model.fields)
        hx = header_float(model.physical_length / model.width)
        hy = header_float(model.physical_length / model.height)
        dtx = [f32(inhomogeneous_dt(Dx[s], rx[s], hx)) for s in range(S)]
        dty = [f32(inhomogeneous_dt(Dy[s], ry[s], hx)) for s in range(S)]
    nx, ny = list(dtx), list(dty)
    sim = f32(0.0)
    for _ in range(steps):
        nxt = [nx[s] if not inh else min(nx[s], ny[s]) for s in range(S)]
        mindt, minspec = nxt[0], 0
        for i in range(1, S):
            if nxt[i] < mindt:
                mindt, minspec = nxt[i], i
        deltat = f32(mindt - sim)
        if R > 0 and deltat > 0 and not (inh and quirks):
            if aqss:
                st, _ = det_aqss(st, tab, float(deltat), np.float64, bool(flags & 1))
            else:
                st = det_rk4(st, tab, float(deltat))
        for i in range(S):
            if not (i == minspec or nxt[i] <= mindt) or not (nxt[i] > 0):
                continue
            if not inh:
                for k in (range(S) if quirks else (i,)):
                    if D_h[k] > 0:
                        st[k:k + 1] = det_hom_diffuse(st[k:k + 1], D_h[k:k + 1], float(dtx[i]), float(h))
                nx[i] = f32(nx[i] + dtx[i])
            else:
                if nx[i] <= mindt:
                    st = det_inhom_sweep(st, i, 0, float(dtx[i]), Dx.astype(np.float64), rx.astype(np.float64), hx)
                    nx[i] = f32(nx[i] + dtx[i])
                if ny[i] <= mindt:
                    st = det_inhom_sweep(st, i, 1, float(dty[i]), Dy.astype(np.float64), ry.astype(np.float64), hy)
                    ny[i] = f32(ny[i] + dty[i])
        sim = mindt
    return st, float(sim)


def gauss2(id_, seed, step):
    """The particle step's Gaussian pair for one particle."""
    gx, gy = C.c_float(), C.c_float()
    lib().oracle_gauss2(C.c_uint32(id_), C.c_uint32(seed), C.c_uint32(step), C.byref(gx), C.byref(gy))
    return gx.value, gy.value
