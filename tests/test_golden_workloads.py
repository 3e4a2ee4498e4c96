"""The seeded benchmark inputs against their committed fixture
(tests/golden/workloads.json, written by tests/golden/make_workload_fixtures.py):
first 8 rows of every table exactly and a SHA-256 of each full table; the
generator's mixing function against published SplitMix64 outputs; and the
configured distributions (ranges of every parameter)."""
import json
import os

import numpy as np
import pytest

from mideq import workloads

HERE = os.path.dirname(os.path.abspath(__file__))


def fixture():
    with open(os.path.join(HERE, "golden", "workloads.json")) as f:
        return json.load(f)


def test_splitmix64_known_outputs():
    """uniform(0, i, k) is the (16 i + k + 1)-th output of SplitMix64 started
    from state 0, top 53 bits.  Published first outputs of that generator:
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    known = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = workloads.uniform(0, np.zeros(3, dtype=np.uint64), 0)
    assert got[0] == (known[0] >> 11) / 2.0**53
    for k, v in enumerate(known):
        assert workloads.uniform(0, [0], k)[0] == (v >> 11) / 2.0**53


@pytest.mark.slow
def test_tables_match_fixture():
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "make_workload_fixtures", os.path.join(HERE, "golden", "make_workload_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    want = fixture()
    got = json.loads(json.dumps(mk.tables()))
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name] == want[name], name


def test_distributions():
    w = workloads.lorenz(20000, seed=1234)
    base = np.array([10.0, 28.0, 8.0 / 3.0])[:, None]
    r = w["p"] / base
    assert (r >= 0.9).all() and (r < 1.1).all()
    assert r.min() < 0.901 and r.max() > 1.099
    assert (w["u0"][:, 0] == [1.0, 0.0, 0.0]).all()
    w = workloads.robertson(20000, seed=1234)
    d = np.log10(w["p"] / np.array([0.04, 3e7, 1e4])[:, None])
    assert (d >= -0.5 - 1e-12).all() and (d < 0.5 + 1e-12).all()
    w = workloads.vdp(20000, seed=1234)
    mu = w["p"][0]
    assert (mu >= 1.0).all() and (mu < 1e4).all()
    assert np.array_equal(w["tf"], 2.0 * mu)
    w = workloads.brusselator(20000, seed=1234)
    g = w["nvec"] // 2
    assert g.min() == 8 and g.max() == 64 and len(np.unique(g)) == 57
    assert (w["p"][1] >= 2.0).all() and (w["p"][1] < 4.0).all()
    assert (w["tf"] >= 1.0).all() and (w["tf"] < 20.0).all()
    # initial state: u_j = 1 + sin(2 pi j / Ngrid), v_j = 3, zero past nvec
    for k in range(5):
        n = w["nvec"][k]
        x = np.arange(n // 2) / (n // 2)
        assert np.array_equal(w["u0"][0:n:2, k], 1.0 + np.sin(2.0 * np.pi * x))
        assert (w["u0"][1:n:2, k] == 3.0).all() and (w["u0"][n:, k] == 0.0).all()


def test_fixture_heads_equal_fresh_draws():
    """The fixture's first rows are what the generator gives for an ensemble
    started at any index offset (draws depend on the index only)."""
    want = fixture()["lorenz_1048576"]["p"]["head"]
    w = workloads.lorenz(8, seed=1234)
    assert w["p"].tolist() == want
    w2 = workloads.lorenz(4, seed=1234, start=4)
    assert w2["p"].tolist() == [row[4:] for row in want]
