"""Arena files (hmv_plan_save / hmv_plan_load): round trip, integrity checks
and index validation, on the host."""
import struct

import numpy as np
import pytest

import plan_emulator as PE

HEADER = 512
SEC_AT = 232  # offset of the section table in the header: 8 x {offset, count, bytes, checksum}
SECTIONS = ["arena", "work1", "work2", "zseg", "outidx", "codes", "tfin", "seg"]


def _fnv(data):
    """The file's section checksum: FNV-1a over little-endian 8-byte words,
    then the tail bytes."""
    h, p, m = 0xcbf29ce484222325, 0x100000001b3, (1 << 64) - 1
    n8 = len(data) // 8 * 8
    for w in np.frombuffer(data[:n8], "<u8").tolist():
        h = ((h ^ w) * p) & m
    for b in data[n8:]:
        h = ((h ^ b) * p) & m
    return h


def _sections(raw):
    return {name: struct.unpack_from("<4Q", raw, SEC_AT + 32 * i) for i, name in enumerate(SECTIONS)}


def _same_tables(a, b):
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


@pytest.mark.parametrize("prec,tuning", [("f64", None), ("bf16", None),
                                         ("f32", {"p1_small_threads": 64, "p2_max_rows": 32})])
def test_round_trip(H, O, sphere4, tmp_path, prec, tuning):
    path = tmp_path / "a.hmv"
    H.save_plan(sphere4, path, prec, tuning)
    T = H.plan_tables(sphere4, prec, tuning)
    L = H.load_plan_tables(path)
    _same_tables(T, L)
    raw = path.read_bytes()
    assert raw[:8] == b"HMVARENA" and struct.unpack_from("<I", raw, 8)[0] == 1
    for name, (off, cnt, nbytes, chk) in _sections(raw).items():
        assert off % 64 == 0 and off + nbytes <= len(raw)
        if nbytes < 4 << 20:
            assert _fnv(raw[off:off + nbytes]) == chk, name
    x = H.uniform(3, sphere4.nd)
    _, y = PE.run(L, x, sphere4.nd)
    y_ref = O.adot_body_lfmtx(sphere4, x)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < {"f64": 1e-14, "f32": 1e-6, "bf16": 1e-2}[prec]


def test_code_table_round_trip(H, O, tmp_path):
    import synthetic_leaves as SL
    hm = SL.many_runs(H)
    path = tmp_path / "runs.hmv"
    H.save_plan(hm, path, "f64")
    L = H.load_plan_tables(path)
    assert L["codes"].size > 0
    _same_tables(H.plan_tables(hm, "f64"), L)


def _write(path, raw):
    path.write_bytes(bytes(raw))
    return path


def test_corrupt_files_are_rejected(H, sphere3, tmp_path):
    good = tmp_path / "good.hmv"
    H.save_plan(sphere3, good, "f64")
    raw = bytearray(good.read_bytes())
    sec = _sections(raw)
    # a flipped bit in each of the table sections and in the image
    for name in ("arena", "work1", "zseg", "outidx"):
        off, cnt, nbytes, chk = sec[name]
        bad = bytearray(raw)
        bad[off + nbytes // 2] ^= 0x10
        with pytest.raises(H.HmvError, match="checksum"):
            H.load_plan_tables(_write(tmp_path / ("bit_" + name), bad))
    with pytest.raises(H.HmvError, match="magic"):
        H.load_plan_tables(_write(tmp_path / "magic", b"NOTARENA" + raw[8:]))
    bad = bytearray(raw)
    struct.pack_into("<I", bad, 8, 99)
    with pytest.raises(H.HmvError, match="version 99"):
        H.load_plan_tables(_write(tmp_path / "version", bad))
    with pytest.raises(H.HmvError, match="out of the file|shorter"):
        H.load_plan_tables(_write(tmp_path / "short", raw[: len(raw) // 2]))
    with pytest.raises(H.HmvError, match="shorter"):
        H.load_plan_tables(_write(tmp_path / "tiny", raw[:100]))
    with pytest.raises(H.HmvError, match="cannot open"):
        H.load_plan_tables(tmp_path / "missing.hmv")


def _patch_section(raw, name, data):
    """Replaces a section's bytes (same length) and its checksum."""
    sec = _sections(raw)
    off, cnt, nbytes, _ = sec[name]
    assert len(data) == nbytes
    raw[off:off + nbytes] = data
    i = SECTIONS.index(name)
    struct.pack_into("<Q", raw, SEC_AT + 32 * i + 24, _fnv(bytes(data)))
    return raw


def test_out_of_range_tables_are_rejected(H, sphere3, tmp_path):
    """Files with valid checksums whose tables address memory outside their
    buffers never reach a kernel."""
    good = tmp_path / "good.hmv"
    H.save_plan(sphere3, good, "f64")
    raw = bytearray(good.read_bytes())
    sec = _sections(raw)
    cases = []
    w1 = np.frombuffer(raw[sec["work1"][0]:sec["work1"][0] + sec["work1"][2]], H.WORK_DTYPE).copy()
    w = w1.copy()
    w["off0"][0] = sec["arena"][2]  # block starts at the end of the image
    cases.append(("work1", w.tobytes(), "outside the arena"))
    w = w1.copy()
    w["rows"][1] = w["rows_p"][1] + 1
    cases.append(("work1", w.tobytes(), "rows"))
    w = w1.copy()
    w["out_base"][2] = 1 << 30
    cases.append(("work1", w.tobytes(), "outputs"))
    zs = np.frombuffer(raw[sec["zseg"][0]:sec["zseg"][0] + sec["zseg"][2]], H.ZSEG_DTYPE).copy()
    z = zs.copy()
    z["src"][0] = sphere3.nd
    cases.append(("zseg", z.tobytes(), "x run out of range"))
    z = zs.copy()
    k = int(np.nonzero(z["kind"] == 2)[0][0])
    z["src"][k] = 1 << 29
    cases.append(("zseg", z.tobytes(), "t run out of range"))
    oi = np.frombuffer(raw[sec["outidx"][0]:sec["outidx"][0] + sec["outidx"][2]], "<i4").copy()
    oi[5] = -1
    cases.append(("outidx", oi.tobytes(), "output index"))
    for i, (name, data, msg) in enumerate(cases):
        bad = _patch_section(bytearray(raw), name, data)
        with pytest.raises(H.HmvError, match=msg):
            H.load_plan_tables(_write(tmp_path / ("case%d" % i), bad))
