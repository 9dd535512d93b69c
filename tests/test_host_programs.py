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
                                 cases.sweep64(n=2)[1]], ids=["sphere", "cylinder", "reference_default", "sweep_case"])
def test_config_text_round_trip(cfg):
    back = nsfsi.parse_config(nsfsi.config_text(cfg))
    for name, _ in nsfsi.NSConfig._fields_:
        a, b = getattr(cfg, name), getattr(back, name)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, name


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
