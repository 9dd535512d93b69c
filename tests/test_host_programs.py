"""The C++ drop-in surface (include/ns_fsi.hpp) and the command-line program on the CPU: ConfigFile text written by
nsfsi.config_text parses back to the same configuration; the program returns -1 when ./ConfigFile is missing or
malformed (src/main.cpp:14-16), before any device is touched."""
import os
import subprocess

import pytest

import cases
import nsfsi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "3d-ns-fsi_amd", "nsfsi_run")


@pytest.mark.parametrize("cfg", [cases.sphere48(N=24, steps=5), cases.cylinder256(N=32), nsfsi.default_config(),
                                 cases.sweep64(n=2)[1], cases.two_bodies(), cases.three_bodies()],
                         ids=["sphere", "cylinder", "reference_default", "sweep_case", "two_bodies", "three_bodies"])
def test_config_text_round_trip(cfg):
    back = nsfsi.parse_config(nsfsi.config_text(cfg))
    for name, _ in nsfsi.NSConfig._fields_:
        a, b = getattr(cfg, name), getattr(back, name)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, name


@pytest.mark.parametrize("entry,message", [
    ('{ type = "cube"; centre = [1.0, 1.0, 1.0]; radius = 0.5; }', "type must be cylinder or sphere"),
    ('{ type = "sphere"; centre = [1.0, 1.0]; radius = 0.5; }', "centre needs 3 floats"),
    ('{ type = "sphere"; centre = [1.0, 1.0, 1.0]; }', "needs type, centre and radius"),
    ('{ type = "sphere"; centre = [1.0, 1.0, 1.0]; radius = 1; }', "radius must be a floating-point value"),
    ('{ type = "cylinder"; axis = "w"; centre = [1.0, 1.0, 1.0]; radius = 0.5; }', "axis must be x, y or z"),
    ('{ type = "sphere"; centre = [1.0, 1.0, 1.0]; radius = -0.5; }', "extra body radius must be positive"),
    ('{ type = "sphere"; centre = [1.0, 1.0, 1.0]; radius = 0.5; colour = "red"; }', "unknown member colour"),
    ('"sphere"', "must be a group"),
])
def test_extra_bodies_parse_errors(entry, message):
    text = nsfsi.config_text(cases.sphere48(N=24, steps=5)) + f"extraBodies = ({entry})\n"
    with pytest.raises(nsfsi.NSError, match=message):
        nsfsi.parse_config(text)


def test_extra_bodies_need_a_first_body():
    cfg = cases.two_bodies()
    cfg.body_type = 0
    with pytest.raises(nsfsi.NSError, match="extra bodies need a first body"):
        nsfsi.parse_config(nsfsi.config_text(cfg))


@pytest.mark.skipif(not os.path.exists(CLI), reason="host program not built")
def test_cli_without_configfile_returns_minus_one(tmp_path):
    r = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 255  # -1
    assert "ConfigFile" in r.stderr


@pytest.mark.skipif(not os.path.exists(CLI), reason="host program not built")
def test_cli_with_malformed_configfile_returns_minus_one(tmp_path):
    text = nsfsi.config_text(cases.sphere48(N=24, steps=5)).replace("NI = 24", "NI = 24.")
    (tmp_path / "ConfigFile").write_text(text)
    r = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert "NI" in r.stderr
    assert not (tmp_path / "output").exists()  # nothing ran
