"""The library's .npy / .npz reader and writer (include/evoc_mi355x.h) against
NumPy's own format implementation: every committed golden read back equal,
round trips of every supported element type and shape in both directions, the
1.0 / 2.0 / 3.0 header versions, zip64 members and archive comments, and the
error paths.  Host-only; no GPU."""
import glob
import io
import os
import zipfile

import numpy as np
import pytest

import evoc_mi355x as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

TYPES = [np.bool_, np.int8, np.uint8, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]
SHAPES = [(), (0,), (5,), (3, 4), (2, 3, 4), (0, 7), (1, 1, 1, 1, 1, 1, 1, 3)]


def _sample(t, shape, seed=0):
    rng = np.random.default_rng(seed)
    if t is np.bool_:
        return rng.random(shape) < 0.5
    if np.issubdtype(t, np.integer):
        info = np.iinfo(t)
        return rng.integers(info.min, info.max, size=shape, dtype=t, endpoint=True)
    x = rng.standard_normal(shape).astype(t)
    if x.size:
        x.flat[0] = np.inf
    return x


def _same(a, b):
    assert a.dtype == b.dtype and a.shape == b.shape
    if a.dtype.kind == "f":
        np.testing.assert_array_equal(a, b)  # NaN-free inputs; inf compares equal
    else:
        assert np.array_equal(a, b)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*", "*.npy"))),
                         ids=lambda p: os.path.relpath(p, GOLDEN))
def test_goldens_npy_read_equal_numpy(path):
    ref = np.load(path, allow_pickle=False)
    _same(E.npy_read(path), ref)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*", "*.npz"))),
                         ids=lambda p: os.path.relpath(p, GOLDEN))
def test_goldens_npz_members_equal_numpy(path):
    with np.load(path, allow_pickle=False) as z:
        names = list(z.files)
        assert names
        for name in names:
            _same(E.npz_read(path, name), z[name])
            _same(E.npz_read(path, name + ".npy"), z[name])  # by_name takes either


@pytest.mark.parametrize("t", TYPES, ids=lambda t: np.dtype(t).name)
def test_roundtrip_every_type_and_shape(tmp_path, t):
    for i, shape in enumerate(SHAPES):
        x = _sample(t, shape, seed=i)
        p = str(tmp_path / f"ours_{i}.npy")
        E.npy_write(p, x)
        _same(np.load(p, allow_pickle=False), x)  # NumPy reads what the library writes
        with open(p, "rb") as f:
            head = f.read(10)
        hlen = int.from_bytes(head[8:10], "little")
        assert head[:8] == b"\x93NUMPY\x01\x00" and (10 + hlen) % 64 == 0
        q = str(tmp_path / f"numpy_{i}.npy")
        np.save(q, x)
        _same(E.npy_read(q), x)  # and the library reads what NumPy writes


@pytest.mark.parametrize("version", [(1, 0), (2, 0), (3, 0)])
def test_header_versions(tmp_path, version):
    x = _sample(np.float32, (17, 5), seed=3)
    p = str(tmp_path / "v.npy")
    with open(p, "wb") as f:
        np.lib.format.write_array(f, x, version=version)
    _same(E.npy_read(p), x)


def test_npz_from_numpy_savez_and_zip64_members_with_comment(tmp_path):
    a, b = _sample(np.int32, (100,), 1), _sample(np.float64, (7, 3), 2)
    p = str(tmp_path / "s.npz")
    np.savez(p, rows=a, vals=b)
    _same(E.npz_read(p, "rows"), a)
    _same(E.npz_read(p, "vals"), b)
    # zip64 extra fields on every member and an archive comment after the
    # central directory (the end record is found by scanning back over it)
    q = str(tmp_path / "z.npz")
    with zipfile.ZipFile(q, "w", compression=zipfile.ZIP_STORED, allowZip64=True) as zf:
        for name, arr in (("rows", a), ("vals", b)):
            buf = io.BytesIO()
            np.lib.format.write_array(buf, arr)
            with zf.open(name + ".npy", "w", force_zip64=True) as w:
                w.write(buf.getvalue())
        zf.comment = b"x" * 300
    _same(E.npz_read(q, "vals"), b)
    _same(E.npz_read(q, "rows.npy"), a)


def _code(excinfo):
    return excinfo.value.code


def test_error_paths(tmp_path):
    good = str(tmp_path / "g.npy")
    np.save(good, np.arange(12, dtype=np.float32).reshape(3, 4))
    with pytest.raises(E.BackendError) as ei:
        E.npy_read(str(tmp_path / "missing.npy"))
    assert _code(ei) == E.ERR_INVALID_ARGUMENT
    bad = str(tmp_path / "bad.npy")
    with open(bad, "wb") as f:
        f.write(b"NOTNUMPY" + bytes(100))
    with pytest.raises(E.BackendError) as ei:
        E.npy_read(bad)
    assert _code(ei) == E.ERR_INVALID_ARGUMENT and "magic" in str(ei.value)
    trunc = str(tmp_path / "t.npy")
    with open(good, "rb") as f:
        blob = f.read()
    with open(trunc, "wb") as f:
        f.write(blob[:-4])
    with pytest.raises(E.BackendError) as ei:
        E.npy_read(trunc)
    assert _code(ei) == E.ERR_INVALID_ARGUMENT and "shorter" in str(ei.value)
    fort = str(tmp_path / "f.npy")
    np.save(fort, np.asfortranarray(np.arange(12, dtype=np.float32).reshape(3, 4)))
    with pytest.raises(E.BackendError) as ei:
        E.npy_read(fort)
    assert _code(ei) == E.ERR_UNSUPPORTED and "Fortran" in str(ei.value)
    for arr in (np.arange(4, dtype=">f4"), np.zeros(3, np.complex64), np.zeros(3, np.float16)):
        p = str(tmp_path / "u.npy")
        np.save(p, arr)
        with pytest.raises(E.BackendError) as ei:
            E.npy_read(p)
        assert _code(ei) == E.ERR_UNSUPPORTED
    z = str(tmp_path / "c.npz")
    np.savez_compressed(z, a=np.arange(10))
    with pytest.raises(E.BackendError) as ei:
        E.npz_read(z, "a")
    assert _code(ei) == E.ERR_UNSUPPORTED and "compressed" in str(ei.value)
    s = str(tmp_path / "s.npz")
    np.savez(s, a=np.arange(10))
    with pytest.raises(E.BackendError) as ei:
        E.npz_read(s, "b")
    assert _code(ei) == E.ERR_INVALID_ARGUMENT and "no member" in str(ei.value)
    with pytest.raises(E.BackendError) as ei:
        E.npz_read(good, "a")  # a .npy is not a zip archive
    assert _code(ei) == E.ERR_INVALID_ARGUMENT
    with pytest.raises(E.BackendError):
        E.npy_write(str(tmp_path / "x.npy"), np.zeros(2, np.complex128))
