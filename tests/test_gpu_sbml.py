"""Models read by the front end (InchmanDiffusionModel::load restated in inchman_amd.sbml) run
on the GPU: the drift/diffusivity method evaluated by the device kernel equals the method text
compiled by g++ (oracle/lawgen.py) bit for bit, and whole runs of the models equal the oracle."""
import os

import numpy as np
import pytest

from inchman_amd import sbml

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_homogeneous_model_file_run(gpu, oracle):
    im = sbml.load(os.path.join(GOLDEN, "inchman_homogeneous.xml"))
    st, real, fields = sbml.build(im)
    m = im.model
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(200, 17)
    want, tw = oracle.run(m, st, 200, 17)
    assert t == tw
    got = s.synchronizeStateBufferFromDevice()
    assert np.array_equal(got, want)
    assert got[2].sum() > st[2].sum()  # the catalysis made product


def test_drift_method_fields_and_run(gpu, oracle):
    from oracle import lawgen

    im = sbml.load(os.path.join(GOLDEN, "inchman_drift.xml"))
    st, real, fields = sbml.initial_state(im)
    m = im.model
    cur = np.zeros((4, 1, 64, 64), np.float32)
    cur[0] = cur[1] = 1.0
    got = np.stack(sbml.drift_diffusivity_fields(im, st, fields, cur))
    want = lawgen.method_fields(m, im.drift_diffusivity_method, im.field_parameters,
                                fields["gradient"][: 64 * 64].reshape(1, 64, 64), st, cur)
    assert got.tobytes() == want.tobytes()
    st2, _, _ = sbml.build(im)
    assert np.array_equal(st2, st)
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(60, 3)
    want_st, tw = oracle.run(m, st, 60, 3)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want_st)


def test_nonlinear_model_with_fixed_fields_runs(gpu, oracle):
    """A non-linear model whose method reads only positions, parameters and field parameters
    (as the reference's OrnsteinUhlenbeck and MultiplicativeNoise examples): its fields after
    any sweep are the first ones, so it builds and runs as the oracle's loop over them; one that
    reads the counts is refused."""
    im = sbml.load(os.path.join(GOLDEN, "inchman_drift.xml"))
    im.nonlinear_diffusivity = True
    im.drift_diffusivity_method = ("All->DiffusivityX = 0.8 + 0.2 * gradient;\nAll->DiffusivityY = 0.6;\n"
                                   "All->DriftX = -(0.01 * PhysicalX);\nAll->DriftY = -(0.005 * PhysicalY);")
    st, real, fields = sbml.build(im)
    m = im.model
    again = np.stack(sbml.drift_diffusivity_fields(im, st * 3 + 1, fields, np.ones((4, 1, 64, 64), np.float32),
                                                   sim_time=5.0))
    first = np.stack(sbml.drift_diffusivity_fields(im, st, fields))
    assert again.tobytes() == first.tobytes()  # other counts, clock and current fields: same fields
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(40, 9)
    want, tw = oracle.run(m, st, 40, 9)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)
    im.drift_diffusivity_method = "All->DriftX = 0.001 * All->State;"
    with pytest.raises(NotImplementedError):
        sbml.build(im)
