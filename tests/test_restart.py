"""Backup files (restart.py) against a Fortran compiler's own unformatted
I/O: a backup written here is read by Fortran statements laid out as
Src/restart.f90:63-70 reads it, and one written by statements laid out as
Src/main.f90:1321-1327 writes it is byte-identical to ours (oracle/_ref,
amdflang).  The subrecord split of records over 2 GiB is checked on a
scaled-down limit; np_dup duplication and ran2 thinning against their
definitions."""
import ctypes as C
import os

import numpy as np
import pytest

from essaitemporaire_lepic_module_amd import restart as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_lepic.so")
need_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="oracle/_ref not built")


def _species(rng, counts):
    return [rng.normal(0, 1, (k, 6)) * np.array([1e-2, 1e-2, 1e-2, 1e6, 1e6, 1e6]) for k in counts]


def _ref_io(kind, path, time, np_tot, vxp, tag_neg, cex):
    lib = C.CDLL(LIB)
    lib.ref_backup_io.argtypes = [C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_double),
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
    t = C.c_double(time)
    name = path.encode()
    lib.ref_backup_io(kind, name, len(name), C.byref(t), len(np_tot), np_tot.ctypes.data,
                      vxp.ctypes.data, tag_neg, cex.ctypes.data)
    return t.value


@need_lib
@pytest.mark.parametrize("tag_neg", [0, 2])
def test_backup_bytes_match_fortran(tmp_path, tag_neg):
    rng = np.random.default_rng(tag_neg)
    counts = [137, 0, 55]
    sp = _species(rng, counts)
    cex = rng.integers(-3, 3, counts[tag_neg - 1]).astype(np.int32) if tag_neg else np.zeros(1, np.int32)
    mine = str(tmp_path / "mine.bak")
    R.write_backup(mine, 1.25e-6, [sp], [cex] if tag_neg else None)
    ref = str(tmp_path / "ref.bak")
    vxp = np.ascontiguousarray(np.concatenate(sp))
    _ref_io(1, ref, 1.25e-6, np.array(counts, np.int32), vxp, tag_neg, cex)
    assert open(mine, "rb").read() == open(ref, "rb").read()
    # and Fortran reads ours back
    got = np.zeros_like(vxp)
    nt = np.zeros(3, np.int32)
    cx = np.zeros_like(cex)
    t = _ref_io(2, mine, 0.0, nt, got, tag_neg, cx)
    assert t == 1.25e-6 and list(nt) == counts
    np.testing.assert_array_equal(got, vxp)
    if tag_neg:
        np.testing.assert_array_equal(cx, cex)
    time, np_tot, back, bc = R.read_backup(mine, 3, 1, tag_neg)
    assert time == 1.25e-6 and list(np_tot[:, 0]) == counts
    for a, b in zip(back[0], sp):
        np.testing.assert_array_equal(a, b)


def test_subrecords(tmp_path, monkeypatch):
    """Records longer than the subrecord limit: leading marker negated when
    another subrecord follows, trailing marker negated when one came before."""
    monkeypatch.setattr(R, "MAX_SUBRECORD", 40)
    payload = np.arange(13, dtype="<f8")          # 104 bytes: 40 + 40 + 24
    path = tmp_path / "r.bin"
    with open(path, "wb") as f:
        R.write_record(f, payload)
        R.write_record(f, np.array([7.0]))
    raw = open(path, "rb").read()
    marks = np.frombuffer(raw[:4], "<i4")[0], np.frombuffer(raw[44:48], "<i4")[0]
    assert marks == (-40, 40)
    assert np.frombuffer(raw[48:52], "<i4")[0] == -40
    assert np.frombuffer(raw[92:96], "<i4")[0] == -40
    assert np.frombuffer(raw[96:100], "<i4")[0] == 24
    assert np.frombuffer(raw[124:128], "<i4")[0] == -24
    with open(path, "rb") as f:
        assert np.array_equal(np.frombuffer(R.read_record(f), "<f8"), payload)
        assert np.frombuffer(R.read_record(f), "<f8")[0] == 7.0


def test_truncated_backup_is_refused(tmp_path):
    rng = np.random.default_rng(1)
    path = str(tmp_path / "p.bak")
    R.write_backup(path, 0.0, [_species(rng, [10, 10])])
    data = open(path, "rb").read()
    open(path, "wb").write(data[:-30])
    with pytest.raises(R.BackupError):
        R.read_backup(path, 2)


def test_external_backups_dealt_round_robin(tmp_path):
    rng = np.random.default_rng(2)
    paths = []
    for rank in range(2):
        sp = [_species(rng, [3 + rank, 2]) for _ in range(3)]     # 3 threads per file
        p = R.backup_path(str(tmp_path), rank)
        R.write_backup(p, 5e-6, sp)
        paths.append((p, sp))
    time, out = R.read_external_backups([p for p, _ in paths], 2, omp_rank_max=3, nproc=2)
    assert time == 5e-6
    # file 0 threads 0,1,2 -> 0,1,0; file 1 the same, appended
    for t in range(2):
        want0 = np.concatenate([paths[0][1][0][t], paths[0][1][2][t], paths[1][1][0][t],
                                paths[1][1][2][t]])
        np.testing.assert_array_equal(out[0][t], want0)
        want1 = np.concatenate([paths[0][1][1][t], paths[1][1][1][t]])
        np.testing.assert_array_equal(out[1][t], want1)
    assert os.path.basename(R.backup_path("d", 12)) == "particles12.bak"


def test_duplicate_and_thin():
    rng = np.random.default_rng(3)
    v = _species(rng, [1000])[0]
    d = R.duplicate_or_thin(v, 3.0, [1])
    assert len(d) == 3000 and np.array_equal(d[2000:], v)
    with pytest.raises(R.BackupError):
        R.duplicate_or_thin(v, 2.5, [1])
    seed = [123456]
    keep = R.duplicate_or_thin(v, 0.3, seed)
    from essaitemporaire_lepic_module_amd import _native
    lib = _native.load()
    s = C.c_int32(123456)
    draws = np.array([lib.lepic_ran2(C.byref(s)) for _ in range(1000)])
    np.testing.assert_array_equal(keep, v[draws <= 0.3])
    assert seed[0] == s.value and 200 < len(keep) < 400
    assert R.duplicate_or_thin(v, 1.0, [1]) is v
