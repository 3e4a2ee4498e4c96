"""The C-ABI library loads and exports every symbol include/mideq.h declares;
argument validation works without a GPU (no compute calls here)."""
import ctypes

import numpy as np
import pytest

import mideq
from mideq import _abi


def test_library_exports_every_declared_symbol():
    L = _abi.load()
    names = _abi.header_functions()
    assert "mideq_ensemble_solve" in names and "mideq_ensemble_solve_device" in names
    for nm in names:
        assert hasattr(L, nm), nm
    assert L.mideq_version() == 10000


def test_struct_layouts_match_header(tmp_path):
    """Compile a probe against include/mideq.h and compare every field offset
    and struct size with the ctypes mirror (the layout the Julia shim binds)."""
    import os
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"mideq_problem": _abi.Problem, "mideq_options": _abi.Options,
               "mideq_outputs": _abi.Outputs, "mideq_stats": _abi.Stats,
               "mideq_user_model": _abi.UserModelSpec}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "mideq.h"', "int main(){"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    inc = os.path.dirname(_abi.HEADER)
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if ln:
            a, b, v = ln.split()
            got[(a, b)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_no_device_is_an_error_not_a_crash():
    L = _abi.load()
    h = ctypes.c_void_p()
    rc = L.mideq_ctx_create(ctypes.byref(h), 0)
    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert rc == -6
    else:
        assert rc == 0
        L.mideq_ctx_destroy(h)


def test_saveat_scalar_expansion_matches_core_jl():
    # test/core.jl:95: saveat=0.1 on (0,1) gives 11 points, both endpoints
    from mideq.scimlbase import _saveat_vector
    v = _saveat_vector(0.1, (0.0, 1.0))
    assert len(v) == 11 and v[0] == 0.0 and v[-1] == 1.0
    # backward: t0, t0 - 0.3, ... and tf, handed over ascending
    b = _saveat_vector(0.3, (1.0, 0.0))
    np.testing.assert_allclose(b, [0.0, 0.1, 0.4, 0.7, 1.0], atol=1e-15)
    assert b[0] == 0.0 and b[-1] == 1.0
    f = _saveat_vector(0.3, (0.0, 1.0))
    np.testing.assert_allclose(f, [0.0, 0.3, 0.6, 0.9, 1.0], atol=1e-15)


def test_saveat_scalar_range_rounds_as_julia():
    """A scalar saveat expands as Julia's Float64 range t0:d:tf: with
    rational t0, d, tf the k-th point is k d rounded once (0:0.1:0.3 ends at
    0.3, not 0.30000000000000004), so tf itself is always among the points.
    Expected values are the correctly rounded quotients k / 10 (one IEEE
    division of exact integers)."""
    from mideq.scimlbase import _saveat_vector, julia_range
    for stop, npts in ((0.3, 4), (0.7, 8), (1.0, 11)):
        v = _saveat_vector(0.1, (0.0, stop))
        assert len(v) == npts
        assert list(v) == [k / 10 for k in range(npts)]
        assert v[-1] == stop
    b = _saveat_vector(0.1, (0.3, 0.0))  # backward span, handed over ascending
    assert list(b) == [0.0, 0.1, 0.2, 0.3]
    # a step that does not divide the span: the range stops short, tf appended
    v = _saveat_vector(0.1, (0.0, 0.25))
    assert list(v) == [0.0, 0.1, 0.2, 0.25]
    # non-rational inputs: t0 + k d rounded once, length round((tf - t0)/d) + 1
    r = julia_range(0.0, 1.0 / 3.0, 1.0)
    assert len(r) == 4 and r[-1] == 1.0
    assert list(julia_range(1.0, -0.25, 0.0)) == [1.0, 0.75, 0.5, 0.25, 0.0]
    assert len(julia_range(0.0, 0.1, -1.0)) == 0


def test_user_model_limits_match_header():
    import re
    txt = open(_abi.HEADER).read()
    get = lambda k: int(re.search(r"#define %s (\d+)" % k, txt).group(1))
    assert (get("MIDEQ_USER_MODEL_BASE"), get("MIDEQ_USER_NMAX"), get("MIDEQ_USER_NPMAX")) == (
        _abi.USER_MODEL_BASE, _abi.USER_NMAX, _abi.USER_NPMAX)


def test_c_abi_example_builds_and_loads():
    """The C example (examples/abi_example.c) is built with the library and
    links libmideq.so; without a GPU it stops at mideq_ctx_create."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "differentialequations.jl_amd", "_build", "abi_example")
    assert os.path.exists(exe)
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present: tests/test_gpu_api.py runs it")
    except ImportError:
        pass
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "no GPU" in r.stdout


def test_stream_only_for_device_resident_solves():
    """A stream names where a device-resident solve (u0 a GPU tensor) is
    enqueued; with host arrays it is an error, raised before any GPU call."""
    from mideq import EnsembleMI355X, device_solve, lorenz
    from mideq.scimlbase import Tsit5
    with pytest.raises(ValueError, match="device-resident"):
        device_solve(lorenz, np.zeros((3, 2)), np.zeros((3, 2)), np.zeros(2), np.ones(2), None,
                     Tsit5(), EnsembleMI355X(), stream=object())


def test_device_buffers_dtypes_and_shapes():
    """The buffers of a device-resident solve, made here in host memory by
    the same code: shapes given as an int or a tuple, every dtype the C ABI
    takes, inputs converted (or used in place when they already fit), and
    each buffer kept for the solve's lifetime."""
    torch = pytest.importorskip("torch")
    from mideq.scimlbase import _DeviceMem
    m = _DeviceMem(torch, torch.device("cpu"))
    a = m.full(5, 0, np.int32)
    assert a.shape == (5,) and a.dtype == torch.int32 and (a == 0).all()
    b = m.full((2, 3, 4), np.nan, np.float64)
    assert b.shape == (2, 3, 4) and b.dtype == torch.float64 and b.isnan().all()
    for dt, tdt in ((np.float64, torch.float64), (np.float32, torch.float32),
                    (np.int64, torch.int64), (np.int32, torch.int32), (np.int8, torch.int8)):
        assert m.empty(3, dt).dtype == tdt and m.empty((2, 3), dt).shape == (2, 3)
    x = np.arange(6.0).reshape(2, 3)
    t = m.arr(x.T, np.float64)  # a strided view: copied contiguous
    assert t.is_contiguous() and np.array_equal(t.numpy(), x.T)
    y = torch.arange(4, dtype=torch.float64)
    assert m.arr(y, np.float64).data_ptr() == y.data_ptr()  # already fits: in place
    assert m.arr(np.array([1, 2]), np.int8).dtype == torch.int8
    assert len(m.keep) == 15


def test_saveat_save_start_save_end_flags():
    """save_start / save_end as OrdinaryDiffEq applies them: only the points
    strictly inside the span come from saveat; t0 and tf are saved by the
    flags, whose defaults are true for a scalar saveat and membership for a
    vector."""
    from mideq.scimlbase import _saveat_vector
    np.testing.assert_array_equal(_saveat_vector(0.25, (0.0, 1.0), save_start=False),
                                  [0.25, 0.5, 0.75, 1.0])
    np.testing.assert_array_equal(_saveat_vector(0.25, (0.0, 1.0), save_end=False),
                                  [0.0, 0.25, 0.5, 0.75])
    np.testing.assert_array_equal(_saveat_vector([0.5, 0.2], (0.0, 1.0)), [0.2, 0.5])
    np.testing.assert_array_equal(_saveat_vector([0.5, 0.2], (0.0, 1.0), save_start=True,
                                                 save_end=True), [0.0, 0.2, 0.5, 1.0])
    np.testing.assert_array_equal(_saveat_vector([0.0, 0.5, 1.0], (0.0, 1.0), save_start=False),
                                  [0.5, 1.0])
    # backward span: the ends are t0 = 1 and tf = 0
    np.testing.assert_array_equal(_saveat_vector(0.5, (1.0, 0.0), save_end=False), [0.5, 1.0])
