"""The Node.js host surface (spectral-2d-turbulence_amd/js) over the N-API addon."""

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

JS = os.path.join(ROOT, "spectral-2d-turbulence_amd", "js")
NODE = shutil.which("node")
ADDON = os.path.join(JS, "s2t_napi.node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON),
                                reason="node or the built addon is missing")


def node(code):
    return subprocess.run([NODE, "-e", code], cwd=JS, capture_output=True, text=True, timeout=120)


def test_js_surface_loads_without_gpu():
    r = node("const m=require('./index.js');"
             "console.log(JSON.stringify(m.radixSchedule(128)), m.SCENARIO_NAMES.length);"
             "try { m.radixSchedule(96); } catch (e) { console.log(e.message); }")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0] == '{"stages":3,"hasRadix2":true} 5'
    assert lines[1] == "grid size 96 must be 4^a or 2 * 4^a"


def test_js_scenarios_match_oracle():
    import json

    import numpy as np

    from oracle.gpusem import scenario_field
    r = node("const m=require('./index.js'); const out={};"
             "for (const k of m.SCENARIO_NAMES) out[k]=Array.from(m.SCENARIOS[k].build(64,{mergerSep:0.9,polygonN:7}));"
             "console.log(JSON.stringify(out));")
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    for name, vals in got.items():
        want = scenario_field(name, 64, {"mergerSep": 0.9, "polygonN": 7}).astype(np.float32)
        assert np.allclose(np.asarray(vals, np.float32).reshape(64, 64), want, rtol=1e-6,
                           atol=1e-5 * np.max(np.abs(want))), name


@pytest.mark.gpu
def test_js_harness_fixtures():
    r = subprocess.run([NODE, os.path.join(JS, "harness.js")], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL CHECKS PASS" in r.stdout
