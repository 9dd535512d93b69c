"""C-ABI boundary on the CPU: the library loads, exports every entry point include/ns_fsi.h declares,
parses the reference ConfigFile format with libconfig's strict typing, and reports errors as status
codes (no GPU is touched by these tests)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import nsfsi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(ROOT, "tests", "golden", "ConfigFile_reference_values")


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ns_fsi.h")).read()
    return sorted(set(re.findall(r"\b(ns_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = nsfsi.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert L.ns_abi_version() == 2


def test_struct_layout_matches_header():
    # ns_config ends with two int32 (precision, kahan); size is a multiple of 8
    assert C.sizeof(nsfsi.NSConfig) % 8 == 0
    c = nsfsi.default_config()
    assert (c.NI, c.NJ, c.NK) == (161, 83, 3)
    assert c.body_type == nsfsi.BODY["cylinder"] and c.body_axis == 2
    assert c.body_radius == pytest.approx(0.1) and list(c.body_centre) == [0.5, 0.5, 0.0]


def test_parse_reference_configfile_equals_defaults():
    c = nsfsi.parse_config(open(FIXTURE).read())
    d = nsfsi.default_config()
    for name, _ in nsfsi.NSConfig._fields_:
        a, b = getattr(c, name), getattr(d, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    assert c.machine_precision_buffer == np.finfo(float).eps * (2.0 + 1.0 + 0.05)


@pytest.mark.parametrize("edit,fragment", [
    (("NI = 161", "NI = 161."), "NI"),                     # integer written as float
    (("Gamma = 1.4", "Gamma = 1"), "Gamma"),              # float written as integer
    (("stopCriterion = \"end_time\"", "stopCriterion = \"sometime\""), "stopCriterion"),
    (("saveConvergenceHistory = \"all\"", "saveConvergenceHistory = \"most\""), "saveConvergenceHistory"),
    (("filterIntervalChangeTimes = [30., 60., 100.]", "filterIntervalChangeTimes = [30.]"), "filterIntervalChangeTimes"),
    (("L_z = 0.05", ""), "L_z"),                           # missing key
    (("Pr = 0.72", "Pr = 0.72 }"), "line"),                # syntax error
])
def test_parse_errors(edit, fragment):
    text = open(FIXTURE).read().replace(*edit)
    with pytest.raises(nsfsi.NSError) as e:
        nsfsi.parse_config(text)
    assert e.value.status == 1
    assert fragment in str(e.value)


def test_parse_extension_keys():
    text = open(FIXTURE).read() + """
boundaryConditions = ["inlet", "outlet", "symmetry", "symmetry", "periodic", "periodic"];
bodyType = "sphere"; bodyCentre = [2.0, 0.5, 0.025]; bodyRadius = 0.2;
initialCondition = "freestream_gaussian"; perturbationSeed = 7; perturbationAmplitude = 1e-3;
inletRampTime = 5.0; precision = "fp32"; kahan = false;
"""
    c = nsfsi.parse_config(text)
    assert list(c.bc) == [0, 1, 3, 3, 2, 2]
    assert c.body_type == 2 and list(c.body_centre) == [2.0, 0.5, 0.025] and c.body_radius == 0.2
    assert c.ic_type == 2 and c.ic_seed == 7 and c.ic_amplitude == 1e-3
    assert c.inlet_ramp_time == 5.0 and c.precision == 1 and c.kahan == 0


def test_invalid_config_rejected_before_device_use():
    """A configuration error is reported as NS_ERR_CONFIG (periodic faces must pair); on a host without a
    gfx950 GPU a valid configuration reports NS_ERR_HIP. Either way no exception crosses the ABI."""
    bad = nsfsi.make_config(bc=["periodic", "outlet", "symmetry", "symmetry", "periodic", "periodic"])
    h = C.c_void_p()
    s = nsfsi.lib().ns_create(C.byref(bad), C.byref(h))
    assert s in (1, 4)
    msg = nsfsi.lib().ns_last_error(h).decode()
    nsfsi.lib().ns_destroy(h)
    if s == 1:
        assert "periodic" in msg
    assert nsfsi.lib().ns_init(None) == 5  # null handle
    assert nsfsi.lib().ns_step(None, 1, None, None, None) == 5
