"""Model front end: the Inchman editor's annotated SBML read into a DiffusionModel.

Restates InchmanDiffusionModel::load (exec/src/inchman/InchmanDiffusionModel.cpp:75-439) with
the standard XML parser instead of libSBML:

* parameters first (a `domain="field"` annotation makes a per-cell field parameter), then the
  overrides (:99-119, :285-304);
* the <im:sbml> annotation: solver, runs, steps, time, outputInterval, the init and events
  scripts with their time stamps, the drift/diffusivity method with its nonLinear and
  computeMoments flags, the setField and newIndividuals methods (:122-210; the latter as
  the library's per-property rule, new_individual_rule);
* the <im:model> annotation: grid width/height and the physical width (:213-220);
* compartments with x, y, width, height (Compartment::setX0/setY0/setWidth/setHeight,
  Compartment.h:82-100; :307-318), species with their diffusion constant, initial amounts per
  compartment and the individual flag (:321-356), reactions with their kinetic law as formula
  text, reactant and product stoichiometry and their compartment localization (:359-439).

MathML contents become the formula text libSBML's SBML_formulaToString writes (numbers
"%.15g", n-ary operators, parentheses by precedence, power as pow(), ln as log()), and
numeric contents are that text evaluated with the parameters in scope, as the reference does
with its embedded Python (:475-500).  The initial state follows
DiffusionModel::generateInitialStates (DiffusionModel.cpp:567-700) and the init script runs
with the namespace of DiffusionModel::callInitScript (DiffusionModel.cpp:728-920).
"""
import copy
import ctypes as C
import math
import re
import xml.etree.ElementTree as ET

import numpy as np

from .gpgmp import (Compartment, DETERMINISTIC_HOMOGENEOUS_AQSS, DETERMINISTIC_HOMOGENEOUS_RK4,
                    DETERMINISTIC_INHOMOGENEOUS_AQSS, DETERMINISTIC_INHOMOGENEOUS_RK4, DiffusionModel,
                    STOCHASTIC_HOMOGENEOUS, STOCHASTIC_INHOMOGENEOUS, STOCHASTIC_INHOMOGENEOUS_FPE)

NS_SBML = "http://www.sbml.org/sbml/level2"
NS_IM = "http://www.csse.monash.edu.au/~berndm/inchman"
NS_M = "http://www.w3.org/1998/Math/MathML"

# DiffusionModel::setSolver(const std::string &) (DiffusionModel.cpp:250-265); an unknown
# name keeps stochastic_inhomogeneous
SOLVERS = {
    "stochastic_homogeneous": STOCHASTIC_HOMOGENEOUS,
    "stochastic_inhomogeneous": STOCHASTIC_INHOMOGENEOUS,
    "stochastic_inhomogeneous_fpe": STOCHASTIC_INHOMOGENEOUS_FPE,
    "deterministic_homogeneous": DETERMINISTIC_HOMOGENEOUS_RK4,
    "deterministic_homogeneous_RK4": DETERMINISTIC_HOMOGENEOUS_RK4,
    "deterministic_homogeneous_aqss": DETERMINISTIC_HOMOGENEOUS_AQSS,
    "deterministic_inhomogeneous_RK4": DETERMINISTIC_INHOMOGENEOUS_RK4,
    "deterministic_inhomogeneous_aqss": DETERMINISTIC_INHOMOGENEOUS_AQSS,
}


def _local(tag):
    return tag.rsplit("}", 1)[-1]


# ---- MathML -> libSBML formula text ----------------------------------------------------
# ASTNode precedences (unary minus 5, + - 2, * / 3, functions and leaves 6) and the grouping
# rule of FormulaFormatter_isGrouped: a child is parenthesized when its precedence is lower
# than its parent's, or equal while it is the last (right) child of a different operator, or
# of a - or /.
_BIN = {"plus": "+", "minus": "-", "times": "*", "divide": "/"}
_FUN = {"exp": "exp", "ln": "log", "abs": "abs", "floor": "floor", "ceiling": "ceil"}


def _real(v):
    """libSBML's real formatting (%.15g)."""
    s = "%.15g" % v
    if s in ("inf", "-inf", "nan"):
        raise ValueError("non-finite number in MathML")
    return s


class _Node:
    def __init__(self, kind, text="", kids=()):
        self.kind = kind  # "num", "name", "uminus", "+", "-", "*", "/", "fn"
        self.text = text
        self.kids = list(kids)

    def precedence(self):
        if self.kind == "uminus":
            return 5
        if self.kind in ("+", "-"):
            return 2
        if self.kind in ("*", "/"):
            return 3
        return 6


def _num(el):
    t = el.get("type", "real")
    txt = "".join(el.itertext()).strip()
    if t == "integer":
        return _Node("num", "%d" % int(txt))
    if t == "e-notation":
        mant = (el.text or "").strip()
        exp_ = "".join(c.tail or "" for c in el).strip()
        return _Node("num", _real(float(mant) * 10.0 ** float(exp_)))
    if t == "rational":
        num = (el.text or "").strip()
        den = "".join(c.tail or "" for c in el).strip()
        return _Node("num", _real(float(num) / float(den)))
    return _Node("num", _real(float(txt)))


def _ast(el):
    tag = _local(el.tag)
    if tag == "math":
        kids = [c for c in el]
        if len(kids) != 1:
            raise ValueError("math element with %d children" % len(kids))
        return _ast(kids[0])
    if tag == "cn":
        return _num(el)
    if tag == "ci":
        return _Node("name", "".join(el.itertext()).strip())
    if tag == "apply":
        kids = list(el)
        op = _local(kids[0].tag)
        args = [_ast(c) for c in kids[1:] if _local(c.tag) not in ("degree", "logbase")]
        if op in _BIN:
            if op == "minus" and len(args) == 1:
                return _Node("uminus", "", args)
            if op == "plus" and len(args) == 1:
                return args[0]
            return _Node(_BIN[op], "", args)
        if op == "power":
            return _Node("fn", "pow", args)
        if op == "root":
            deg = [c for c in kids[1:] if _local(c.tag) == "degree"]
            if deg and _formula(_ast(list(deg[0])[0])) != "2":
                return _Node("fn", "pow", [args[0], _Node("/", "", [_Node("num", "1"), _ast(list(deg[0])[0])])])
            return _Node("fn", "sqrt", args)
        if op == "log":
            base = [c for c in kids[1:] if _local(c.tag) == "logbase"]
            if base and _formula(_ast(list(base[0])[0])) != "10":
                raise ValueError("log with a base other than 10")
            return _Node("fn", "log10", args)
        if op in _FUN:
            return _Node("fn", _FUN[op], args)
        raise ValueError("unsupported MathML operator <%s>" % op)
    if tag in ("true", "false"):
        return _Node("name", tag)
    raise ValueError("unsupported MathML element <%s>" % tag)


def _grouped(parent, child, last):
    if parent.kind == "fn":
        return False
    pp, cp = parent.precedence(), child.precedence()
    if pp > cp:
        return True
    if pp == cp and last and parent.kind != "uminus":
        return parent.kind != child.kind or parent.kind in ("-", "/")
    return False


def _formula(n):
    if n.kind in ("num", "name"):
        return n.text
    if n.kind == "fn":
        return "%s(%s)" % (n.text, ", ".join(_formula(k) for k in n.kids))
    if n.kind == "uminus":
        k = n.kids[0]
        s = _formula(k)
        return "-(" + s + ")" if _grouped(n, k, True) else "-" + s
    out = []
    for i, k in enumerate(n.kids):
        s = _formula(k)
        out.append("(" + s + ")" if _grouped(n, k, i == len(n.kids) - 1) else s)
    return (" %s " % n.kind).join(out)


def mathml_formula(math_el):
    """The formula text of a MathML <math> element (SBML_formulaToString's form)."""
    return _formula(_ast(math_el))


# ---- the model ------------------------------------------------------------------------
class InchmanModel:
    """What InchmanDiffusionModel::load sets on its DiffusionModel, plus the run settings."""

    def __init__(self):
        self.model = None            # DiffusionModel
        self.solver_name = None
        self.num_runs = 1            # DiffusionModel.cpp:151-158 defaults
        self.max_steps = None
        self.max_time = 100.0
        self.output_interval = 100.0
        self.init_script = ""
        self.events_script = ""
        self.event_times = []
        self.drift_diffusivity_method = ""
        self.nonlinear_diffusivity = False
        self.compute_moments = False
        self.set_field_method = ""
        self.new_individuals_method = ""
        self.field_parameters = []
        self.compartments = {}       # id -> Compartment (in document order)
        self.initial_amounts = {}    # compartment id -> {species id: amount}
        self.individual_species = []
        self.namespace = {}          # the evaluation namespace (parameters, true/false)

    def eval(self, formula):
        """A MathML formula's value with the parameters in scope (the reference evals it in its
        embedded Python, InchmanDiffusionModel.cpp:484-488)."""
        env = {"__builtins__": {}, "pow": pow, "exp": math.exp, "log": math.log, "log10": math.log10,
               "sqrt": math.sqrt, "abs": abs, "floor": math.floor, "ceil": math.ceil}
        env.update(self.namespace)
        return float(eval(formula, env))  # noqa: S307 (the model file's own MathML, as the reference)


def _find(el, ns, name):
    return el.find("{%s}%s" % (ns, name)) if el is not None else None


def _annotation(el, name):
    ann = _find(el, NS_SBML, "annotation")
    return _find(ann, NS_IM, name)


def _math_child(el):
    for c in el:
        if _local(c.tag) == "math":
            return c
    return None


def _eval_child(im, node, name):
    c = _find(node, NS_IM, name)
    if c is None:
        return None
    m = _math_child(c)
    if m is None:
        return None
    return im.eval(mathml_formula(m))


def _text(node, name):
    c = _find(node, NS_IM, name)
    return (c.text or "") if c is not None else ""


def load(filename_or_text, param_overrides=None):
    """InchmanDiffusionModel::fromFile: the annotated SBML (a path or the XML text) as an
    InchmanModel whose .model is the DiffusionModel for the solvers."""
    if filename_or_text.lstrip().startswith("<"):
        root = ET.fromstring(filename_or_text)
    else:
        root = ET.parse(filename_or_text).getroot()
    im = InchmanModel()
    im.namespace = {"false": 0.0, "true": 1.0}
    sm = _find(root, NS_SBML, "model")
    if sm is None:
        raise ValueError("no <model> element")
    params = {}
    # parameters first: every later MathML may refer to them
    for p in sm.iter("{%s}parameter" % NS_SBML):
        a = _annotation(p, "parameter")
        domain = a.get("domain", "single") if a is not None else "single"
        if domain == "field":
            im.field_parameters.append(p.get("id"))
        elif domain == "single":
            v = float(np.float32(float(p.get("value", "0"))))  # a Real
            params[p.get("id")] = v
            im.namespace[p.get("id")] = v
        else:
            raise ValueError("parameter %s has domain %s (not supported by the reference either)"
                             % (p.get("id"), domain))
    for k, v in (param_overrides or {}).items():
        params[k] = float(np.float32(v))
        im.namespace[k] = params[k]

    top = _annotation(root, "sbml")
    if top is not None:
        s = _find(top, NS_IM, "solver")
        if s is not None and s.text:
            im.solver_name = s.text.strip()
        for name, attr, conv in (("runs", "num_runs", int), ("steps", "max_steps", int),
                                 ("time", "max_time", float), ("outputInterval", "output_interval", float)):
            v = _eval_child(im, top, name)
            if v is not None:
                setattr(im, attr, conv(v))
        init = _find(top, NS_IM, "init")
        if init is not None:
            im.init_script = _text(init, "script")
        ev = _find(top, NS_IM, "events")
        if ev is not None:
            im.events_script = _text(ev, "script")
            ts = _find(ev, NS_IM, "listOfTimeStamps")
            for t in (ts if ts is not None else []):
                m = _math_child(t)
                if m is not None:
                    im.event_times.append(im.eval(mathml_formula(m)))
        dd = _find(top, NS_IM, "driftDiffusivity")
        if dd is not None:
            im.drift_diffusivity_method = _text(dd, "method")
            im.nonlinear_diffusivity = dd.get("nonLinear", "").lower() in ("true", "1")
            im.compute_moments = dd.get("computeMoments", "").lower() in ("true", "1")
        sf = _find(top, NS_IM, "setField")
        if sf is not None:
            im.set_field_method = _text(sf, "method")
        ni = _find(top, NS_IM, "newIndividualsMethod")
        if ni is not None:
            im.new_individuals_method = _text(ni, "method")

    W = H = 0
    L = 0.0
    ma = _annotation(sm, "model")
    if ma is not None:
        W = int(_eval_child(im, ma, "gridWidth") or 0)
        H = int(_eval_child(im, ma, "gridHeight") or 0)
        L = float(np.float32(_eval_child(im, ma, "physicalWidth") or 0.0))
    solver = SOLVERS.get(im.solver_name or "", STOCHASTIC_INHOMOGENEOUS)
    m = DiffusionModel(W, H, physical_length=L, solver=solver,
                       max_time=im.max_time if im.max_time is not None else 1e30)
    for k, v in params.items():
        m.setParameter(k, v)

    for c in sm.iter("{%s}compartment" % NS_SBML):
        cp = Compartment(c.get("id"))
        a = _annotation(c, "compartment")
        x0 = y0 = 0
        x1 = y1 = 0
        if a is not None:
            v = _eval_child(im, a, "x")
            x0 = int(v) if v is not None else 0
            v = _eval_child(im, a, "y")
            y0 = int(v) if v is not None else 0
            v = _eval_child(im, a, "width")
            x1 = x0 + int(v) - 1 if v is not None else 0     # Compartment::setWidth
            v = _eval_child(im, a, "height")
            y1 = y0 + int(v) - 1 if v is not None else 0
        cp.setBounds(x0, y0, x1, y1)
        im.compartments[cp.id] = cp
        im.initial_amounts[cp.id] = {}

    for sp in sm.iter("{%s}species" % NS_SBML):
        a = _annotation(sp, "species")
        D = 0.0
        if a is not None:
            v = _eval_child(im, a, "diffusionConstant")
            D = float(np.float32(v)) if v is not None else 0.0
        s = m.addSpecies(sp.get("id"), D)
        if a is not None:
            lst = _find(a, NS_IM, "listOfCompartmentParameters")
            for cpar in (lst if lst is not None else []):
                if _local(cpar.tag) != "compartmentParameters":
                    continue
                cid = cpar.get("compartment")
                if cid in im.compartments:
                    v = _eval_child(im, cpar, "initialAmount")
                    im.initial_amounts[cid][s.id] = int(v) if v is not None else 0
            if _find(a, NS_IM, "individual") is not None:
                im.individual_species.append(s.id)

    by_id = {s.id: s for s in m.species}
    for r in sm.iter("{%s}reaction" % NS_SBML):
        reac, prod = {}, {}
        for lst, out in (("listOfReactants", reac), ("listOfProducts", prod)):
            el = _find(r, NS_SBML, lst)
            for ref in (el if el is not None else []):
                s = by_id.get(ref.get("species"))
                if s is None:
                    continue
                out[s] = out.get(s, 0) + int(float(ref.get("stoichiometry", "1")))
        rx = m.addReaction(1.0, reactants=reac, products=prod, id=r.get("id"))
        kl = _find(r, NS_SBML, "kineticLaw")
        mm = _math_child(kl) if kl is not None else None
        if mm is not None:
            rx.setCustomLaw(mathml_formula(mm))
        a = _annotation(r, "reaction")
        lst = _find(a, NS_IM, "listOfCompartmentParameters") if a is not None else None
        for cpar in (lst if lst is not None else []):
            if _local(cpar.tag) != "compartmentParameters":
                continue
            allowed = _eval_child(im, cpar, "isAllowed")
            cid = cpar.get("compartment")
            if allowed and cid in im.compartments:
                rx.addCompartment(im.compartments[cid])
    im.model = m
    return im


def _print_statements(script):
    """The reference runs the scripts in an embedded Python 2 (boost-python): its `print x`
    statements become print(x) calls; the rest of the script is run as written."""
    import re

    return "\n".join(re.sub(r"^(\s*)print(?!\s*\()\s+(.*)$", r"\1print(\2)", line)
                     for line in script.split("\n"))


def initial_state(im, run_init_script=True):
    """DiffusionModel::generateInitialStates (DiffusionModel.cpp:567-700) for the homogeneous
    distribution the editor's initial amounts use, then the init script.  Returns (int32 state
    [S][H][W], float32 deterministic state [S][H][W], field parameters {name: float32 [S*H*W]})."""
    m = im.model
    W, H, S = m.width, m.height, m.numSpecies()
    A = W * H
    states = np.zeros(S * A, np.int32)
    for cid, cp in im.compartments.items():
        x0, y0, x1, y1 = cp.getBounds()
        area = (x1 - x0 + 1) * (y1 - y0 + 1)
        for si, sp in enumerate(m.species):
            amount = im.initial_amounts.get(cid, {}).get(sp.id, 0)
            if amount == 0 or area <= 0:
                continue
            per, rem = amount // area, amount % area
            for y in range(y0, y1 + 1):
                for x in range(x0, x1 + 1):
                    if 0 <= x < W and 0 <= y < H:
                        off = y * W + x
                        # the reference compares the grid offset, not the compartment-local
                        # index, with the remainder (DiffusionModel.cpp:614-617)
                        states[si * A + off] += per + 1 if off < rem else per
    real = states.astype(np.float32)
    fields = {name: np.zeros(S * A, np.float32) for name in im.field_parameters}
    if run_init_script and im.init_script.strip():
        solver_names = {STOCHASTIC_HOMOGENEOUS: "stochastic_homogeneous",
                        DETERMINISTIC_HOMOGENEOUS_RK4: "deterministic_homogeneous",
                        DETERMINISTIC_HOMOGENEOUS_AQSS: "deterministic_homogeneous",
                        STOCHASTIC_INHOMOGENEOUS: "stochastic_inhomogeneous"}
        params = dict(m.parameters)
        params.update(fields)
        ns = {"__name__": "__main__",
              "runtimeInformation": {"nx": W, "ny": H, "length": m.physical_length, "nspecies": S,
                                     "solver": solver_names.get(m.solver, "unknown")},
              "parameters": params,
              "species": {sp.id: i for i, sp in enumerate(m.species)},
              "state": states}
        ns.update(m.parameters)
        if m.solver in (DETERMINISTIC_HOMOGENEOUS_RK4, DETERMINISTIC_HOMOGENEOUS_AQSS,
                        DETERMINISTIC_INHOMOGENEOUS_RK4, DETERMINISTIC_INHOMOGENEOUS_AQSS):
            ns["deterministicState"] = real
        exec(compile(_print_statements(im.init_script), "<init script>", "exec"), ns)  # noqa: S102
    return states.reshape(S, H, W), real.reshape(S, H, W), fields


def drift_diffusivity_fields(im, state, fields=None, current=None, sim_time=0.0, stream=None, on_host=False):
    """The model's drift/diffusivity method run on the device for every cell and species
    (gpgmp_model_eval_drift_diffusivity): (Dx, Dy, rx, ry), each [S][H][W] float32 -- the
    fields an inhomogeneous solver takes.  fields: {name: array with at least H*W values} of
    the field parameters (the init script's, first H*W values per field)."""
    from ._lib import call

    m = im.model
    S, H, W = m.numSpecies(), m.height, m.width
    tmp = copy.copy(m)  # a handle for the names and constants (no fields needed)
    tmp.solver = STOCHASTIC_HOMOGENEOUS
    tmp.fields = None
    cm = tmp.compile()
    names = list(im.field_parameters)
    fv = np.zeros((max(len(names), 1), H, W), np.float32)
    for k, name in enumerate(names):
        if fields and name in fields:
            fv[k] = np.asarray(fields[name], np.float32).reshape(-1)[: H * W].reshape(H, W)
    st = np.ascontiguousarray(state, np.int32).reshape(S, H, W)
    cur = None if current is None else np.ascontiguousarray(current, np.float32).reshape(4, S, H, W)
    out = np.zeros((4, S, H, W), np.float32)
    cnames = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    if on_host:  # the compiler checked on the host (gpgmp_model_method_host)
        cur = np.zeros((4, S, H, W), np.float32) if cur is None else cur
        call("gpgmp_model_method_host", cm.handle, im.drift_diffusivity_method.encode(), cnames, len(names),
             fv.ctypes.data_as(C.POINTER(C.c_float)), st.ctypes.data_as(C.POINTER(C.c_int32)),
             cur.ctypes.data_as(C.POINTER(C.c_float)), C.c_float(sim_time), out.ctypes.data_as(C.POINTER(C.c_float)))
        return out[0], out[1], out[2], out[3]
    stream_ptr = None if stream is None else int(getattr(stream, "cuda_stream", stream))
    call("gpgmp_model_eval_drift_diffusivity", cm.handle, im.drift_diffusivity_method.encode(), cnames, len(names),
         fv.ctypes.data_as(C.POINTER(C.c_float)), st.ctypes.data_as(C.POINTER(C.c_int32)),
         None if cur is None else cur.ctypes.data_as(C.POINTER(C.c_float)), C.c_float(sim_time),
         out.ctypes.data_as(C.POINTER(C.c_float)), stream_ptr)
    return out[0], out[1], out[2], out[3]


def method_dependencies(im):
    """What the drift/diffusivity method reads besides the cell's position, the constants and
    the field parameters (gpgmp_model_method_dependencies): bit 0 the counts, bit 1
    PhysicalSimTime, bit 2 the current fields."""
    from ._lib import call

    tmp = copy.copy(im.model)
    tmp.solver = STOCHASTIC_HOMOGENEOUS
    tmp.fields = None
    cm = tmp.compile()
    names = list(im.field_parameters)
    cnames = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    deps = C.c_uint32()
    call("gpgmp_model_method_dependencies", cm.handle, im.drift_diffusivity_method.encode(), cnames, len(names),
         C.byref(deps))
    return deps.value


def check_nonlinear(im):
    """A non-linear model (fields recomputed after every sweep, InhomogeneousSolver.cpp:304-336,
    384-417): True when its method reads what a sweep changes -- the counts, the clock, the
    current fields -- so that the solver must re-evaluate it (the stochastic inhomogeneous
    solvers do, gpgmp_solver_set_nonlinear_method); False when re-evaluating it reproduces
    the first fields and time steps.  Other solver types with such a method raise
    NotImplementedError with what it reads."""
    if not im.nonlinear_diffusivity:
        return False
    deps = method_dependencies(im)
    if not deps:
        return False
    if im.model.solver in (STOCHASTIC_INHOMOGENEOUS, STOCHASTIC_INHOMOGENEOUS_FPE):
        return True
    what = [n for b, n in ((1, "the counts"), (2, "PhysicalSimTime"), (4, "the current fields")) if deps & b]
    raise NotImplementedError("non-linear diffusivity whose method reads %s (fields recomputed after every "
                              "sweep) is supported by the stochastic inhomogeneous solvers only" % " and ".join(what))


class _Affine:
    """a + b * RANDOM for the newIndividuals method's expressions (RANDOM the uniform draw)."""

    def __init__(self, a, b=0.0):
        self.a, self.b = float(a), float(b)

    @staticmethod
    def of(v):
        return v if isinstance(v, _Affine) else _Affine(v)

    def __add__(self, o):
        o = _Affine.of(o)
        return _Affine(self.a + o.a, self.b + o.b)

    __radd__ = __add__

    def __sub__(self, o):
        o = _Affine.of(o)
        return _Affine(self.a - o.a, self.b - o.b)

    def __rsub__(self, o):
        return _Affine.of(o) - self

    def __neg__(self):
        return _Affine(-self.a, -self.b)

    def __pos__(self):
        return self

    def __mul__(self, o):
        o = _Affine.of(o)
        if self.b and o.b:
            raise NotImplementedError("a newIndividuals expression with RANDOM * RANDOM")
        return _Affine(self.a * o.a, self.a * o.b + self.b * o.a)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = _Affine.of(o)
        if o.b:
            raise NotImplementedError("a newIndividuals expression dividing by RANDOM")
        return _Affine(self.a / o.a, self.b / o.a)

    def __rtruediv__(self, o):
        if self.b:
            raise NotImplementedError("a newIndividuals expression dividing by RANDOM")
        return _Affine(float(o) / self.a)


class _CInt(int):
    """An integer literal: division of two of them truncates toward zero, as in C."""

    def __truediv__(self, o):
        if isinstance(o, _CInt):
            q = abs(int(self)) // abs(int(o))
            return _CInt(q if (self < 0) == (o < 0) else -q)
        return _Affine.of(float(self)) / o

    def __rtruediv__(self, o):
        return _Affine.of(o) / float(self)

    def __neg__(self):
        return _CInt(-int(self))

    def _int_op(self, o, f):
        if isinstance(o, int) and not isinstance(o, bool):
            return _CInt(f(int(self), int(o)))
        return f(_Affine.of(float(self)), o)

    def __add__(self, o):
        return self._int_op(o, lambda p, q: p + q)

    def __radd__(self, o):
        return self._int_op(o, lambda p, q: q + p)

    def __sub__(self, o):
        return self._int_op(o, lambda p, q: p - q)

    def __rsub__(self, o):
        return self._int_op(o, lambda p, q: q - p)

    def __mul__(self, o):
        return self._int_op(o, lambda p, q: p * q)

    def __rmul__(self, o):
        return self._int_op(o, lambda p, q: q * p)


_NEW_IND_PROPS = ("DiffusivityX", "DiffusivityY", "DriftX", "DriftY")


def new_individual_rule(im):
    """The newIndividuals method (DiffusionModel::getNewIndividualsMethod, pasted into
    updateNewParticles, IndividualSolver.cl:1284 and IndividualSolver.cpp:264-305) as the
    library's rule: ((mode per property), (a), (b)) with mode 0 keep, 1 the constant a,
    2 a + b * RANDOM, for the properties DiffusivityX, DiffusivityY, DriftX, DriftY; None for
    an empty method.  Statements `All-><property> = <expr>;` with parameters, numbers (C
    literals, integer division truncating) and RANDOM entering affinely; anything else
    raises NotImplementedError (the reference pastes arbitrary C there)."""
    text = re.sub(r"//[^\n]*", "", im.new_individuals_method or "")
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    modes, a, b = [0] * 4, [0.0] * 4, [0.0] * 4
    env = {k: float(v) for k, v in im.model.parameters.items()}
    env["RANDOM"] = _Affine(0.0, 1.0)
    any_stmt = False
    for stmt in text.split(";"):
        stmt = stmt.strip()
        if not stmt:
            continue
        m = re.fullmatch(r"All\s*->\s*(\w+)\s*=\s*(.+)", stmt, re.S)
        if not m or m.group(1) not in _NEW_IND_PROPS:
            raise NotImplementedError("newIndividuals statement not supported: %r" % stmt)
        expr = m.group(2)
        if re.search(r"[^\w\s.+\-*/()]", expr):
            raise NotImplementedError("newIndividuals expression not supported: %r" % expr)
        # C literals: float suffixes dropped; plain integers keep C's integer division
        lit = r"(?<![\w.])(\d+(?:\.\d*)?[eE][+-]?\d+|\d+\.\d*|\.\d+|\d+)[fF]?(?![\w.])"
        names = set(re.findall(r"[A-Za-z_]\w*", re.sub(lit, " ", expr)))
        expr = re.sub(lit, lambda t: t.group(1) if re.search(r"[.eE]", t.group(1)) else "_CInt(%s)" % t.group(1), expr)
        unknown = names - set(env)
        if unknown:
            raise NotImplementedError("newIndividuals expression names %s" % sorted(unknown))
        v = _Affine.of(eval(expr, {"__builtins__": {}, "_CInt": _CInt}, env))  # noqa: S307 (model text)
        k = _NEW_IND_PROPS.index(m.group(1))
        modes[k], a[k], b[k] = (2 if v.b else 1), v.a, v.b
        any_stmt = True
    return (tuple(modes), tuple(a), tuple(b)) if any_stmt else None


def build(im, run_init_script=True):
    """The model ready for its solver: the initial state (initial amounts and the init
    script), and for the inhomogeneous types the drift/diffusivity fields of its method at
    t = 0 (computed on the device).  Returns (int32 state, float32 deterministic state,
    field parameters)."""
    state, real, fields = initial_state(im, run_init_script)
    m = im.model
    if m.solver in (STOCHASTIC_INHOMOGENEOUS, STOCHASTIC_INHOMOGENEOUS_FPE, DETERMINISTIC_INHOMOGENEOUS_RK4,
                    DETERMINISTIC_INHOMOGENEOUS_AQSS):
        recompute = check_nonlinear(im)
        S, H, W = m.numSpecies(), m.height, m.width
        cur = np.zeros((4, S, H, W), np.float32)
        for s, sp in enumerate(m.species):  # the species' constant as the starting value
            cur[0, s] = cur[1, s] = sp.diffusionConstant
        Dx, Dy, rx, ry = drift_diffusivity_fields(im, state, fields, cur)
        m.setDriftDiffusivityFields(Dx, Dy, rx, ry)
        if recompute:  # the solvers re-evaluate the method after every sweep (runPerCall)
            names = list(im.field_parameters)
            fv = np.stack([np.asarray(fields[n], np.float32).reshape(-1)[: H * W].reshape(H, W) for n in names]) \
                if names else None
            m.setNonlinearMethod(im.drift_diffusivity_method, names, fv)
    return state, real, fields
