"""The rtm3d driver's input handling (no GPU needed: every case fails
before the device is touched), mirroring the reference's checks."""
import json
import os
import subprocess

import pytest

from rtm3d_demo_amd import DRIVER


def _run(tmp_path, params):
    path = tmp_path / "input.json"
    path.write_text(json.dumps(params))
    return subprocess.run([DRIVER, str(path)], capture_output=True, text=True, timeout=60)


def test_no_argument():
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "forgot the input.json" in r.stdout


def test_missing_file(tmp_path):
    r = subprocess.run([DRIVER, str(tmp_path / "nope.json")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "doesn't exists" in r.stdout


def test_missing_key(tmp_path, cases):
    p = dict(cases["small_mod"])
    del p["taper_factor"]
    r = _run(tmp_path, p)
    assert r.returncode == 1 and "taper_factor" in r.stdout


def test_wrong_type(tmp_path, cases):
    p = dict(cases["small_mod"])
    p["modeling"] = 1
    r = _run(tmp_path, p)
    assert r.returncode == 1 and "boolean" in r.stdout


def test_nt_smaller_than_ntstep(tmp_path, cases):
    p = dict(cases["small_mod"])
    p["ntstep"] = p["nt"] + 1
    r = _run(tmp_path, p)
    assert r.returncode == 1 and "NT smaller than NTSTEP" in r.stdout


def test_malformed_json(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text('{"nx": 12, ')
    r = subprocess.run([DRIVER, str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "JSON parse error" in r.stdout
