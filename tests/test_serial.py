"""Binary caches of parsed edge lists and update records (host only): a round trip
equals the parse, the layout equals an independent writer of the documented format,
and corrupted, truncated or inconsistent files are refused."""
import struct

import numpy as np
import pytest

from suminc_amd import _lib as L
from suminc_amd import textio


def fnv1a(data: bytes) -> int:
    h = 0xcbf29ce484222325
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def _edge_file(tmp_path, weights):
    lines = ["# comment", "0 1 5", "1 2 7", "2 0 3", "3 1 9", "0 2 4", "bad line", "1 2 8", ""]
    p = tmp_path / "g.e"
    p.write_text("\n".join(lines) + "\n")
    return textio.EdgeFile(p, weights=weights)


@pytest.mark.parametrize("weights", [None, "u32", "f32"])
def test_edge_cache_round_trip(tmp_path, weights):
    ef = _edge_file(tmp_path, weights)
    cache = tmp_path / "cache.s"
    ef.save(cache)
    back = textio.EdgeFile(cache, cache=True)
    assert (back.n, back.m, back.weight_kind, back.skipped_lines) == (ef.n, ef.m, ef.weight_kind, ef.skipped_lines)
    for a, b in zip(ef.csr(), back.csr()):
        assert (a is None and b is None) or np.array_equal(a, b)
    # the documented layout, written independently
    off, col, w = ef.csr()
    body = b"SUMINCE1" + struct.pack("<IIQQ", ef.weight_kind, ef.n, ef.m, ef.skipped_lines)
    body += off.astype("<u8").tobytes() + col.astype("<u4").tobytes()
    if w is not None:
        body += w.view("<u4").tobytes()
    body += struct.pack("<Q", fnv1a(body))
    assert cache.read_bytes() == body


def _raw(tmp_path, n, off, col, fix_sum=True):
    body = b"SUMINCE1" + struct.pack("<IIQQ", 0, n, len(col), 0)
    body += np.asarray(off, "<u8").tobytes() + np.asarray(col, "<u4").tobytes()
    body += struct.pack("<Q", fnv1a(body) if fix_sum else 0)
    p = tmp_path / "raw.s"
    p.write_bytes(body)
    return p


def test_edge_cache_refuses_bad_files(tmp_path):
    ef = _edge_file(tmp_path, "u32")
    good = tmp_path / "good.s"
    ef.save(good)
    data = bytearray(good.read_bytes())
    cases = {"flip": bytes(data[:40]) + bytes([data[40] ^ 1]) + bytes(data[41:]),
             "truncated": bytes(data[:-3]),
             "magic": b"SUMINCX1" + bytes(data[8:])}
    for name, blob in cases.items():
        p = tmp_path / f"{name}.s"
        p.write_bytes(blob)
        with pytest.raises(L.SumincError) as e:
            textio.EdgeFile(p, cache=True)
        assert e.value.code == L.E_INVALID, name
    # consistent checksums over inconsistent content
    for off, col in (([0, 2, 1, 3], [1, 2, 0]), ([0, 1, 2, 3], [1, 2, 7]), ([1, 1, 2, 3], [1, 2, 0])):
        with pytest.raises(L.SumincError):
            textio.EdgeFile(_raw(tmp_path, 3, off, col), cache=True)
    textio.EdgeFile(_raw(tmp_path, 3, [0, 1, 2, 3], [1, 2, 0]), cache=True)  # the valid one loads
    with pytest.raises(L.SumincError):
        textio.EdgeFile(tmp_path / "missing.s", cache=True)


def test_update_cache_round_trip_and_refusals(tmp_path):
    p = tmp_path / "u.txt"
    p.write_text("a 0 1 5\nd 1 2\na 3 4 9\n")
    ops = textio.read_updates(p, weights="u32")
    cache = tmp_path / "u.s"
    textio.save_updates(cache, ops)
    back = textio.load_updates(cache)
    assert back.dtype == ops.dtype and np.array_equal(back, ops)
    raw = cache.read_bytes()
    assert raw[:8] == b"SUMINCU1" and struct.unpack("<Q", raw[8:16])[0] == 3 and len(raw) == 24 + 3 * 24
    assert struct.unpack("<Q", raw[-8:])[0] == fnv1a(raw[:-8])
    bad = tmp_path / "bad.s"
    bad.write_bytes(raw[:30] + bytes([raw[30] ^ 4]) + raw[31:])
    with pytest.raises(L.SumincError):
        textio.load_updates(bad)
    bad.write_bytes(raw[:-5])
    with pytest.raises(L.SumincError):
        textio.load_updates(bad)
    textio.save_updates(tmp_path / "empty.s", np.zeros(0, L.UPDATE_DTYPE))
    assert len(textio.load_updates(tmp_path / "empty.s")) == 0


def test_serialization_path_follows_the_reference_rule():
    # SetSerialize: prefix + "/" + (efile + vfile + worker count, '/' -> '_'), then
    # kSerializationFilenameFormat "%s/frag_%d.s" with the fragment id
    assert textio.serialization_path("/tmp/ser", "/data/g.e", "", 1, 0) == "/tmp/ser/_data_g.e1/frag_0.s"
    assert textio.serialization_path("p", "a/b.e", "a/b.v", 4, 3) == "p/a_b.ea_b.v4/frag_3.s"
