"""Custom kinetic laws in the oracle: the reference's generated kernel code for a model,
written as C++ and compiled with the host compiler.

TEST INFRASTRUCTURE ONLY (used through oracle.table_from_model by tests/ and bench.py's
cpu_baseline leg).  The reference pastes every reaction's law text into its kernel source
and lets the OpenCL compiler type it:

    const Real <param> = <value>;                      (Solver.cpp:834-841)
    #define <species> _localState[...]                 (int counts / Real concentrations)
    _h[0] = 0.;
    _h[i+1] = _h[0] + [mask *] <law_i>; _h[0] = _h[i+1];          (Solver.cpp:843-857)
    _deriv[s] = 0. + n. * [mask *] <law> - m. * ...;              (Solver.cpp:918-955)
    _q[s] = 0. + n. * [mask *] <law> ...; _d[s] = 0. + ...;       (Solver.cpp:960-1010)

Here the same lines are generated from the same texts and compiled by g++ (no contraction),
so the C typing of the formulas (int, float and double literals; float parameters; int or
floating species) is the C compiler's, independent of the HIP library's own law compiler.
The laws' math functions call the oracle's exp/log/pow (oracle_law_fn_d, oracle_law_pow_d);
OpenCL's built-ins are implementation-defined (parity unpinned for them).
"""
import ctypes as C
import hashlib
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PRELUDE = r"""
#include <stdint.h>
#include <type_traits>
extern "C" double oracle_law_fn_d(int fn, double x);
extern "C" double oracle_law_pow_d(double x, double y);
namespace lawfns {
template <class A> struct dbl { static const bool v = std::is_same<A, double>::value; };
template <class A, class B> using pow_t = typename std::conditional<dbl<A>::v || dbl<B>::v, double, float>::type;
template <class A> using fn_t = typename std::conditional<dbl<A>::v, double, float>::type;
template <class A, class B> static inline pow_t<A, B> pow(A a, B b) {
    if (dbl<A>::v || dbl<B>::v) return (pow_t<A, B>)oracle_law_pow_d((double)a, (double)b);
    return (pow_t<A, B>)(float)oracle_law_pow_d((double)(float)a, (double)(float)b);
}
template <class A> static inline fn_t<A> fnx(int f, A a) {
    if (dbl<A>::v) return (fn_t<A>)oracle_law_fn_d(f, (double)a);
    return (fn_t<A>)(float)oracle_law_fn_d(f, (double)(float)a);
}
template <class A> static inline fn_t<A> exp(A a) { return fnx(0, a); }
template <class A> static inline fn_t<A> log(A a) { return fnx(1, a); }
template <class A> static inline fn_t<A> log10(A a) { return fnx(2, a); }
static inline float sqrt(float a) { return __builtin_sqrtf(a); }
static inline double sqrt(double a) { return __builtin_sqrt(a); }
static inline float sqrt(int a) { return __builtin_sqrtf((float)a); }
static inline float fabs(float a) { return __builtin_fabsf(a); }
static inline double fabs(double a) { return __builtin_fabs(a); }
static inline float fabs(int a) { return __builtin_fabsf((float)a); }
static inline float floor(float a) { return __builtin_floorf(a); }
static inline double floor(double a) { return __builtin_floor(a); }
static inline float floor(int a) { return __builtin_floorf((float)a); }
static inline float ceil(float a) { return __builtin_ceilf(a); }
static inline double ceil(double a) { return __builtin_ceil(a); }
static inline float ceil(int a) { return __builtin_ceilf((float)a); }
static inline int abs(int a) { return (a < 0) ? (int)(0u - (unsigned)a) : a; }
static inline float abs(float a) { return __builtin_fabsf(a); }
static inline double abs(double a) { return __builtin_fabs(a); }
}
using namespace lawfns;
"""


def _flit(v):
    """A float's exact value as a C++ hex-float literal of type float."""
    return float(np.float32(v)).hex() + "f"


def _counts(model, stoich):
    """{species index: count} of a reaction's reactant or product map (keys Species or index)."""
    out = {}
    for sp, n in (stoich or {}).items():
        i = sp.index if hasattr(sp, "index") else int(sp)
        out[i] = out.get(i, 0) + int(n)
    return out


def law_texts(model, deterministic):
    """Per reaction the law text the reference pastes: the custom law, or the generated one
    (Reaction.cpp:119-162, species in index order)."""
    out = []
    for r in model.reactions:
        if getattr(r, "customLaw", None):
            out.append(r.customLaw)
            continue
        t = "%g" % np.float32(r.rate)
        rc = _counts(model, r.reactants)
        for sp in model.species:
            n = rc.get(sp.index, 0)
            for k in range(n):
                t += (" * " + sp.id) if (k == 0 or deterministic) else " * (%s-%d)" % (sp.id, k)
        out.append(t)
    return out


def source(model, header):
    """The generated C++ of a model's laws.  header: the header constants as floats."""
    S = len(model.species)
    ids = [sp.id for sp in model.species]
    net = [[_counts(model, r.products).get(s, 0) - _counts(model, r.reactants).get(s, 0) for s in range(S)]
           for r in model.reactions]
    loc = [bool(getattr(r, "compartments", None)) for r in model.reactions]
    params = dict(getattr(model, "parameters", {}) or {})
    src = [PRELUDE]

    def consts(real):
        lines = ["    const %s %s = %s;" % (real, k, _flit(v) if real == "float" else repr(float(np.float32(v))))
                 for k, v in header.items()]
        # `const Real k = <value>;` with the parameter printed with 6 significant digits
        lines += ["    const %s %s = %s;" % (real, k, "%g" % np.float32(v)) for k, v in params.items()]
        lines += ["    (void)%s;" % k for k in list(header) + list(params)]
        return lines

    def masks():
        return ["    const float _mask%d = ((_act >> %d) & 1u) ? 1.f : 0.f; (void)_mask%d;" % (i, i, i)
                for i in range(len(model.reactions))]

    # hazards: int counts
    sl = law_texts(model, False)
    src.append('extern "C" void law_hazards(const int32_t *_x, int GridX, int GridY, uint32_t _act, float *_h) {')
    src.append("    (void)GridX; (void)GridY;")
    src += consts("float") + masks()
    src += ["#define %s (_x[%d])" % (ids[s], s) for s in range(S)]
    src.append("    _h[0] = 0.;")
    for i, law in enumerate(sl):
        m = "_mask%d * " % i if loc[i] else ""
        src.append("    _h[%d] = _h[0] + %s%s; _h[0] = _h[%d];" % (i + 1, m, law, i + 1))
    src += ["#undef %s" % ids[s] for s in range(S)]
    src.append("}")
    dl = law_texts(model, True)

    def rhs_terms(s, sign_filter):
        terms = []
        for i, law in enumerate(dl):
            n = net[i][s]
            if n == 0 or (sign_filter == 1 and n < 0) or (sign_filter == -1 and n > 0):
                continue
            m = "_mask%d * " % i if loc[i] else ""
            op = " + " if (n > 0 or sign_filter != 0) else " - "
            terms.append("%s%d. * %s%s" % (op, abs(n), m, law))
        return "0." + "".join(terms)

    for name, real in (("law_rk_f64", "double"),):
        src.append('extern "C" void %s(const %s *_x, int GridX, int GridY, uint32_t _act, %s *_deriv) {'
                   % (name, real, real))
        src.append("    (void)GridX; (void)GridY;")
        src += consts(real) + masks()
        src += ["#define %s (_x[%d])" % (ids[s], s) for s in range(S)]
        src += ["    _deriv[%d] = %s;" % (s, rhs_terms(s, 0)) for s in range(S)]
        src += ["#undef %s" % ids[s] for s in range(S)]
        src.append("}")
    for name, real in (("law_aqss_f32", "float"), ("law_aqss_f64", "double")):
        src.append('extern "C" void %s(const %s *_y, int GridX, int GridY, uint32_t _act, %s *_q, %s *_d) {'
                   % (name, real, real, real))
        src.append("    (void)GridX; (void)GridY;")
        src += consts(real) + masks()
        src += ["#define %s (_y[%d])" % (ids[s], s) for s in range(S)]
        for s in range(S):
            src.append("    _q[%d] = %s;" % (s, rhs_terms(s, 1)))
            src.append("    _d[%d] = %s;" % (s, rhs_terms(s, -1)))
        src += ["#undef %s" % ids[s] for s in range(S)]
        src.append("}")
    return "\n".join(src) + "\n"


_loaded = {}


def build(src):
    key = hashlib.sha1(src.encode()).hexdigest()[:16]
    if key in _loaded:
        return _loaded[key]
    d = os.path.join(tempfile.gettempdir(), "gpgmp_lawgen")
    os.makedirs(d, exist_ok=True)
    cpp = os.path.join(d, key + ".cpp")
    so = os.path.join(d, key + ".so")
    if not os.path.exists(so):
        with open(cpp, "w") as f:
            f.write(src)
        tmp = so + ".%d" % os.getpid()
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                        cpp, "-o", tmp, "-L", HERE, "-loracle", "-Wl,-rpath," + HERE], check=True)
        os.replace(tmp, so)
    lib = C.CDLL(so)
    _loaded[key] = lib
    return lib


def attach(table, model):
    """Points the oracle table's custom-law slots at the compiled code of `model`."""
    from . import oracle as o

    L = o.lib()
    W = model.width
    header = {
        "PhysicalCellWidth": L.oracle_header_float(C.c_float(np.float32(model.physical_length) / np.float32(W))),
        "PhysicalCellHeight": L.oracle_header_float(C.c_float(np.float32(model.physical_length) /
                                                              np.float32(model.height))),
        "PhysicalModelWidth": L.oracle_header_float(C.c_float(model.physical_length)),
        "PhysicalModelHeight": L.oracle_header_float(C.c_float(model.physical_length)),
        "StayProbability": o.stay_probability(model.p0),
    }
    lib = build(source(model, header))
    table._lawlib = lib
    for field, fn in (("hazards_fn", "law_hazards"), ("rk_fn", "law_rk_f64"), ("aqss_f32_fn", "law_aqss_f32"),
                      ("aqss_f64_fn", "law_aqss_f64")):
        setattr(table, field, C.cast(getattr(lib, fn), C.c_void_p).value)
    table.grid_w = W
    return table


# ---- drift/diffusivity methods -----------------------------------------------------------
METHOD_PRELUDE = r"""
struct _species_t {
    const int State;
    const int *SpeciesState;
    const float MeanX;
    float DiffusivityX, DiffusivityY, DriftX, DriftY;
};
static int _gx, _gy, _W;
// getCentralDifferenceX/Y (cl_include/helpers.cl:39-87): the mirrored neighbour at an edge,
// the Y form bounded by the grid width as written there
static inline float getCentralDifferenceX(_species_t *species) {
    const int *d = species->SpeciesState;
    const int sl = (_gx > 0) ? d[(_gx - 1) + _gy * _W] : d[(_gx + 1) + _gy * _W];
    const int sr = (_gx < _W - 1) ? d[(_gx + 1) + _gy * _W] : d[(_gx - 1) + _gy * _W];
    return sr - sl;
}
static inline float getCentralDifferenceY(_species_t *species) {
    const int *d = species->SpeciesState;
    const int st = (_gy > 0) ? d[_gx + (_gy - 1) * _W] : d[_gx + (_gy + 1) * _W];
    const int sb = (_gy < _W - 1) ? d[_gx + (_gy + 1) * _W] : d[_gx + (_gy - 1) * _W];
    return sb - st;
}
"""


def method_source(model, method, field_names, header):
    """The reference's computeDiffusionConstants_forSpecies_<id> (Solver.cpp:659-777) for every
    species as one C++ function over all cells."""
    S = len(model.species)
    ids = [sp.id for sp in model.species]
    params = dict(getattr(model, "parameters", {}) or {})
    src = [PRELUDE, METHOD_PRELUDE, "typedef float Real;"]
    src.append('extern "C" void method_fields(const int *_state, const float *_field, const float *_cur, int W, '
               'int H, float PhysicalSimTime, float *_out) {')
    src.append("    const int __area = W * H; _W = W; (void)PhysicalSimTime;")
    for k, v in header.items():
        src.append("    const float %s = %s; (void)%s;" % (k, _flit(v), k))
    for k, v in params.items():
        src.append("    const Real %s = %s; (void)%s;" % (k, "%g" % np.float32(v), k))
    src.append("    for (int __sp = 0; __sp < %d; ++__sp)" % S)
    src.append("    for (int __cell = 0; __cell < __area; ++__cell) {")
    src.append("        const int GridX = __cell % W, GridY = __cell / W; _gx = GridX; _gy = GridY; (void)GridY;")
    # getPhysicalX/Y (cl_include/helpers.cl:29-36)
    src.append("        const Real PhysicalX = (Real)((double)(PhysicalModelWidth / (Real)W) * "
               "((double)(Real)GridX - (double)(Real)W / 2.)); (void)PhysicalX;")
    src.append("        const Real PhysicalY = (Real)((double)(PhysicalModelHeight / (Real)H) * "
               "((double)(Real)GridY - (double)(Real)H / 2.)); (void)PhysicalY;")
    for s, sid in enumerate(ids):
        src.append("        _species_t _%s = {_state[%d * __area + __cell], _state + %d * __area, 0.f, _cur[(0 * %d + %d) * __area + __cell], "
                   "_cur[(1 * %d + %d) * __area + __cell], _cur[(2 * %d + %d) * __area + __cell], _cur[(3 * %d + %d) * __area + __cell]};"
                   % (sid, s, s, S, s, S, s, S, s, S, s))
        src.append("        _species_t *%s = &_%s; (void)%s;" % (sid, sid, sid))
    src.append("        _species_t *__all[%d] = {%s};" % (S, ", ".join("&_" + sid for sid in ids)))
    src.append("        _species_t *All = __all[__sp]; (void)All;")
    for k, name in enumerate(field_names):
        src.append("#define %s (_field[%d * __area + __cell])" % (name, k))
    src.append("        {")
    src.append(method)
    src.append("        }")
    for name in field_names:
        src.append("#undef %s" % name)
    src.append("        _out[(0 * %d + __sp) * __area + __cell] = All->DiffusivityX;" % S)
    src.append("        _out[(1 * %d + __sp) * __area + __cell] = All->DiffusivityY;" % S)
    src.append("        _out[(2 * %d + __sp) * __area + __cell] = All->DriftX;" % S)
    src.append("        _out[(3 * %d + __sp) * __area + __cell] = All->DriftY;" % S)
    src.append("    }")
    src.append("}")
    return "\n".join(src) + "\n"


def method_fields(model, method, field_names=(), fields=None, state=None, current=None, sim_time=0.0):
    """The drift/diffusivity fields [4][S][H][W] the method gives (g++-compiled)."""
    from . import oracle as o

    L = o.lib()
    W, H, S = model.width, model.height, len(model.species)
    header = {
        "PhysicalCellWidth": L.oracle_header_float(C.c_float(np.float32(model.physical_length) / np.float32(W))),
        "PhysicalCellHeight": L.oracle_header_float(C.c_float(np.float32(model.physical_length) /
                                                              np.float32(H))),
        "PhysicalModelWidth": L.oracle_header_float(C.c_float(model.physical_length)),
        "PhysicalModelHeight": L.oracle_header_float(C.c_float(model.physical_length)),
        "StayProbability": o.stay_probability(model.p0),
    }
    lib = build(method_source(model, method, list(field_names), header))
    st = np.ascontiguousarray(state if state is not None else np.zeros((S, H, W)), np.int32)
    fl = np.ascontiguousarray(fields if fields is not None and len(field_names) else np.zeros(1), np.float32)
    cur = np.ascontiguousarray(current if current is not None else np.zeros((4, S, H, W)), np.float32)
    out = np.zeros((4, S, H, W), np.float32)
    fn = lib.method_fields
    fn.argtypes = [C.c_void_p] * 3 + [C.c_int, C.c_int, C.c_float, C.c_void_p]
    fn(st.ctypes.data, fl.ctypes.data, cur.ctypes.data, W, H, float(sim_time), out.ctypes.data)
    return out
