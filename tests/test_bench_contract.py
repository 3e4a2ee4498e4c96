"""bench.py's command-line contract, checked without a GPU: more than one
GPU is refused with a JSON message, and the form table names the fields the
measurement reads."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_forms_and_gate():
    sys.path.insert(0, ROOT)
    import bench
    for name, (precision, tol, schedule, save_dtype, controller) in bench.FORMS.items():
        assert precision in ("fp64", "mixed")
        assert 0 < tol <= 1e-6
        assert schedule in ("lane", "persistent")
        assert save_dtype in ("float64", "float32")
        assert controller in (0, 1)
    # the plain form is the configuration's FP64 Tsit5 at 1e-8 with the
    # reference controller; every measured traffic entry names a form
    assert bench.FORMS["plain"] == ("fp64", 1e-8, "lane", "float64", 0)
    assert set(bench.MEASURED_TRAFFIC) <= set(bench.FORMS)
    assert set(bench.SAVE_LAYOUT) <= set(bench.FORMS)
    assert set(bench.SAVE_LAYOUT.values()) <= {0, 1}
    assert bench.FLOPS_PER_STEP == 262 and bench.FLOPS_PER_SAVE == 98


def test_headline_only_at_the_stated_tolerance():
    """The headline is chosen among forms at their precision's stated
    tolerance (FP64 1e-8, FP32 stages 1e-6); a looser form is reported
    beside it, never as it."""
    sys.path.insert(0, ROOT)
    import bench
    assert bench.CONFIG_TOL == {"fp64": 1e-8, "mixed": 1e-6}
    on = [f for f, v in bench.FORMS.items() if v[1] == bench.CONFIG_TOL[v[0]]]
    assert {"plain", "tuned", "mixed"} <= set(on)
    assert "tuned_tol" not in on


def test_headline_only_with_soa_output():
    """The configuration's output is structure-of-arrays [100][3][N]; a form
    with another layout is reported beside the headline, never as it."""
    sys.path.insert(0, ROOT)
    import bench
    src = open(os.path.join(ROOT, "bench.py")).read()
    assert 'SAVE_LAYOUT.get(f, 0) == 0]' in src
    assert bench.SAVE_LAYOUT.get("tuned", 0) == 0 and bench.SAVE_LAYOUT.get("plain", 0) == 0
