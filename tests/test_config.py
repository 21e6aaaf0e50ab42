"""Device options from the run's JSON configuration (the file the reference's
ConfigParser reads, ConfigParser.cpp:18-226): the reference's own configs
leave the defaults, the optional "Device ..." keys set precision, em_push
compat, seed, density mode, sort interval, write-back, movement capacity and
per-type charges and diameters, and malformed input is reported with its
reason.  Through shim_check --print-config (host only)."""
from __future__ import annotations

import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "start_amd", "shim_check")
CONFIGS = os.path.join(ROOT, "tests", "golden", "configs")
E = 1.602176565e-19
DEFAULTS = {"precision": "fp64", "em_push": "1", "seed": "1", "density": "barycentric", "sort_interval": "20",
            "write_back": "0", "movement_max_particles": "200000"}


def _options(path):
    r = subprocess.run([EXE, "--print-config", str(path)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def _parse(out):
    o, charges, diam = {}, {}, {}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "charge":
            charges[int(f[1])] = float(f[2])
        elif f[0] == "diameter":
            diam[int(f[1])] = float(f[2])
        else:
            o[f[0]] = f[1]
    return o, charges, diam


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(CONFIGS) if f.endswith(".json")))
def test_reference_configs_leave_the_defaults(name):
    rc, out, err = _options(os.path.join(CONFIGS, name))
    assert rc == 0, err
    o, c, d = _parse(out)
    assert o == DEFAULTS and not c and not d


def test_device_keys(tmp_path):
    cfg = json.load(open(os.path.join(CONFIGS, "config2.json")))
    cfg.update({"Device Precision": "Mixed", "Device EM Push": False, "Device Seed": 18446744073709551615,
                "Device Density Mode": "TetAverage", "Device Sort Interval": 0, "Device Write Back Each Step": True,
                "Device Movement Max Particles": 12345, "Device Charges": {"Ti": 2, "Ar": 1.5},
                "Device Diameters": {"Ti": 2.1e-10}})
    p = tmp_path / "c.json"
    p.write_text(json.dumps(cfg, indent=4))
    rc, out, err = _options(p)
    assert rc == 0, err
    o, c, d = _parse(out)
    assert o == {"precision": "mixed", "em_push": "0", "seed": "18446744073709551615", "density": "tet_average",
                 "sort_interval": "0", "write_back": "1", "movement_max_particles": "12345"}
    # ParticleType order of Constants.hpp:78-93: O2 Ar Ne He Ti ...
    assert c == {1: 1.5 * E, 4: 2 * E} and d == {4: 2.1e-10}


def test_json_reading(tmp_path):
    """Escapes, surrogate pairs, nested arrays, exponents and literals in keys
    the reader skips."""
    p = tmp_path / "c.json"
    p.write_text('{"a\\u00e9\\ud83d\\ude00\\"\\\\\\/\\b\\f\\n\\r\\t": [[1, -2.5e-3, 1E+2], {"x": null}, true, false],'
                 ' "Device Seed": 7}')
    rc, out, err = _options(p)
    assert rc == 0, err
    assert _parse(out)[0]["seed"] == "7"


@pytest.mark.parametrize("text,msg", [
    ('{"Device Precision": "FP16"}', '"Device Precision" must be'),
    ('{"Device Seed": -1}', '"Device Seed" must be a non-negative integer'),
    ('{"Device Seed": 1.5}', '"Device Seed" must be a non-negative integer'),
    ('{"Device EM Push": 1}', '"Device EM Push" must be true or false'),
    ('{"Device Charges": {"Xe": 1}}', 'unknown particle type "Xe"'),
    ('{"Device Diameters": {"Ti": -1e-10}}', 'must be a positive number'),
    ('{"Device Precison": "Mixed"}', 'unknown device key "Device Precison"'),
    ('{"Device Density Mode": "nodal"}', '"Device Density Mode" must be'),
    ('{"Time Step": 1e-9,\n "Gas": "Ar"\n "P": 1}', "JSON line 3: expected '}'"),
    ('{"Gas": "Ar"} x', "text after the top-level value"),
    ('{"Gas": "A\x01r"}', "control character"),
    ('[1, 2]', "not a JSON object"),
])
def test_errors(tmp_path, text, msg):
    p = tmp_path / "bad.json"
    p.write_text(text)
    rc, out, err = _options(p)
    assert rc == 1 and msg in err, err
    rc, out, err = _options(tmp_path / "missing.json")
    assert rc == 1 and "cannot open" in err
