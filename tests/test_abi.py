"""The C-ABI library loads and exports every entry point include/gpgmp_hip.h declares;
model validation returns status codes (host-only calls, no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "gpgmp_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(gpgmp_\w+)\s*\(", src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from inchman_amd import _lib

    L = C.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers the whole header
    assert set(names) - {"gpgmp_last_error"} == set(_lib.SIGNATURES), \
        set(names) ^ set(_lib.SIGNATURES)


def test_binding_loads():
    import inchman_amd

    L = inchman_amd.lib()
    maj, mi = C.c_int32(), C.c_int32()
    assert L.gpgmp_version(C.byref(maj), C.byref(mi)) == 0
    assert maj.value == 1


def _model(**kw):
    from inchman_amd import configs

    m = configs.plain_gmp_model(32)
    for k, v in kw.items():
        setattr(m, k, v)
    return m


@pytest.mark.parametrize("field,value", [("width", 30), ("height", 0), ("physical_length", -1.0),
                                         ("solver", 9), ("boundary_mask", (2, 0, 0, 0))])
def test_invalid_models_return_status(field, value):
    from inchman_amd import GpgmpError

    m = _model(**{field: value})
    with pytest.raises(GpgmpError) as e:
        m.compile()
    assert e.value.code == -1


def test_too_many_species():
    from inchman_amd import DiffusionModel

    m = DiffusionModel(32, 32, 32.0)
    for i in range(9):
        m.addSpecies("S%d" % i, 1.0)
    with pytest.raises(ValueError):
        m.compile()


def test_negative_rate_rejected():
    from inchman_amd import GpgmpError

    m = _model()
    m.reactions[0].rate = -1.0
    with pytest.raises(GpgmpError):
        m.compile()


def test_rate_literals_follow_reference_printing():
    """Reaction.cpp:122 prints the rate with `ss << rate` (%g, 6 digits): 1 and 3 are
    int literals, 0.01, 0.1 and 0.0001 double literals with their decimal value."""
    from inchman_amd import configs

    is_int, val = configs.plain_gmp_model(32).compile().rateLiterals()
    assert list(is_int) == [True, True, False, False]
    assert list(val) == [1.0, 1.0, 0.01, 0.1]
    is_int, val = configs.schnakenberg_model(32).compile().rateLiterals()
    assert list(is_int) == [True, True, False, True]
    assert val[2] == 0.0001  # the decimal, not (double)1e-4f
    assert val[2] != float(np.float32(1e-4))


def test_rate_literals_match_oracle(oracle):
    from inchman_amd import configs

    for m in (configs.plain_gmp_model(32), configs.schnakenberg_model(32)):
        is_int, val = m.compile().rateLiterals()
        t = oracle.table_from_model(m)
        for i in range(len(m.reactions)):
            assert bool(t.lit_is_int[i]) == bool(is_int[i])
            assert t.lit_dbl[i] == val[i]
