"""The C ABI: liblepic_hip.so loads without a GPU, exports every entry point
include/lepic.h declares, and its host-side routines (ran2, seed, the host
loader) agree with the oracle.  No device calls."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import _native, config as cfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "lepic.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lepic_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(names) == sorted(_native.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", _native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (lepic_\w+)", out))
    assert set(names) <= exported


def test_host_ran2_and_seed_match_known_answers():
    lib = _native.load()
    s = C.c_int32(1)
    for _ in range(10000):
        lib.lepic_ran2(C.byref(s))
    assert s.value == 1043618065
    assert lib.lepic_seed(1, 0) == 123456
    assert lib.lepic_version().startswith(b"lepic-hip")


def test_host_loader_matches_oracle(oracle):
    """The product's host loader (same ran2 stream) against the oracle's."""
    c = cfg.cfg1_plain_box(n=6, np_cell=4)
    lib = _native.load()
    cap = 2000
    host = []
    psets = (_native.PSet * 2)()
    for t in range(2):
        a = {f: np.zeros(cap) for f in ("x", "y", "z", "vx", "vy", "vz")}
        a["cell"] = np.zeros(cap, np.int32)
        a["dead"] = np.zeros(cap, np.int8)
        host.append(a)
        for f in a:
            setattr(psets[t], f, a[f].ctypes.data)
        psets[t].n, psets[t].cap = 0, cap
    bc = np.ascontiguousarray(c.bcnd)
    kq = np.ascontiguousarray(c.kq())
    npt = np.zeros((2,) + bc.shape)
    seed = (C.c_int32 * 1)(123456)
    d = lambda v: (C.c_double * 2)(*v)
    st = lib.lepic_load_particles_host(C.byref(_native.grid(c.n, c.h)), bc.ctypes.data, kq.ctypes.data,
                                       0.0, c.np_cell, 2, d(c.vt0[:2]), d(c.Nm[:2]), d([1.0, 1.0]),
                                       seed, psets, npt.ctypes.data)
    assert st == 0
    po, npo, so = oracle.load_particles(c.n, c.h, bc, kq, c.np_cell, c.vt0[:2], c.Nm[:2],
                                        [1.0, 1.0], 123456, cap=cap)
    assert seed[0] == so
    for t in range(2):
        m = psets[t].n
        o = po[t].arrays()
        assert m == po[t].n
        for f in ("x", "y", "z"):
            np.testing.assert_array_equal(host[t][f][:m], o[f])
        for f in ("vx", "vy", "vz"):   # log/cos/sin of two libm builds: 1 ulp
            np.testing.assert_allclose(host[t][f][:m], o[f], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(npt, npo)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stdout


def test_load_fails_loudly_without_library(tmp_path, monkeypatch):
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(ImportError):
        _native.load()
