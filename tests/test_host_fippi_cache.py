"""The FIPPI cache file format of the reference (FIPPICache.cc:189-414):
"FIPPICACHE\\0", then 352-byte records of 15 float32 key values, 4 bytes
of padding and the 36 doubles of QIFIPPIData.  Host-side: preload, the
first-record-wins rule, byte-exact store, the skip when storing back to
the file just preloaded, and the warnings on malformed files."""
import warnings

import numpy as np
import pytest

REC = np.dtype([("key", "<f4", 15), ("pad", "<u4"), ("rec", "<f8", 36)])
SIG = b"FIPPICACHE\x00"


def write_cache(path, keys, recs):
    a = np.zeros(len(keys), dtype=REC)
    a["key"] = keys
    a["rec"] = recs
    path.write_bytes(SIG + a.tobytes())


def test_record_layout():
    assert REC.itemsize == 352


def test_preload_store_roundtrip(sx, tmp_path):
    rng = np.random.default_rng(7)
    n = 5
    # keys far outside any mesh (|x| ~ 1e30), so the records never meet a real geometry
    keys = (1e30 * (1 + rng.random((n, 15)))).astype(np.float32)
    recs = rng.standard_normal((n, 36))
    f1 = tmp_path / "a.cache"
    write_cache(f1, keys, recs)
    before = sx.CacheSize()
    assert sx.PreloadCache(str(f1)) == n
    assert sx.CacheSize() == before + n
    # storing back to the file just preloaded is skipped (FIPPICache.cc:240-261)
    assert sx.StoreCache(str(f1)) == 0
    f2 = tmp_path / "b.cache"
    m = sx.StoreCache(str(f2))
    assert m == before + n
    raw = f2.read_bytes()
    assert raw[:11] == SIG and (len(raw) - 11) % 352 == 0
    got = np.frombuffer(raw[11:], dtype=REC)
    tail = got[-n:]
    assert np.array_equal(tail["key"], keys) and np.array_equal(tail["rec"], recs)
    assert np.all(tail["pad"] == 0)
    # a duplicated key keeps its first record
    f3 = tmp_path / "c.cache"
    write_cache(f3, keys[:1], recs[1:2])
    assert sx.PreloadCache(str(f3)) == 1
    assert sx.CacheSize() == before + n
    f4 = tmp_path / "d.cache"
    sx.StoreCache(str(f4))
    got = np.frombuffer(f4.read_bytes()[11:], dtype=REC)
    row = np.nonzero(np.all(got["key"] == keys[0], axis=1))[0]
    assert len(row) == 1 and np.array_equal(got["rec"][row[0]], recs[0])


@pytest.mark.parametrize("content,why", [
    (b"FIPPICACHX\x00" + bytes(352), "signature"),
    (SIG + bytes(351), "size"),
    (b"FIPPI", "short"),
])
def test_malformed_files_warn_and_skip(sx, tmp_path, content, why):
    f = tmp_path / f"bad_{why}.cache"
    f.write_bytes(content)
    before = sx.CacheSize()
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert sx.PreloadCache(str(f)) == 0
    assert any("PreloadCache" in str(x.message) for x in w)
    assert sx.CacheSize() == before


def test_missing_file_warns(sx, tmp_path):
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        assert sx.PreloadCache(str(tmp_path / "none.cache")) == 0
    assert w
