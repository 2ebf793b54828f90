"""CPU emulation of the engine's two streaming phases over a host plan.

Test helper: walks the same work/zseg tables the HIP kernel walks
(hmv_kernels.hip), asserts that every access the kernel would make stays
inside its buffer, and computes t partials and y in float64 from the stored
(FP64, FP32 or BF16) entries.  Used to validate the repack on the CPU before
any GPU run; it is not an engine path.
"""
import numpy as np

ZCHUNK = 2048
MAX_ROWS = 256
THREADS = 256
DT_BYTES = {0: 8, 1: 4, 2: 2}
Z_X, Z_CODES, Z_T = 0, 1, 2
RUN_MAX = 64
MAX_TPARTS = 32


def _decode(raw, dt):
    if dt == 0:
        return raw.view("<f8").astype(np.float64)
    if dt == 1:
        return raw.view("<f4").astype(np.float64)
    return (raw.view("<u2").astype(np.uint32) << 16).view("<f4").astype(np.float64)


def _part(arena, off, rows_p, ncols, dt):
    """Column-major rows_p x ncols block at byte offset off (bounds-checked)."""
    es = DT_BYTES[dt]
    assert off >= 0 and off % 16 == 0, "part offsets are 16-byte aligned"
    end = off + rows_p * ncols * es
    assert end <= arena.size, "part reads past the arena"
    return _decode(arena[off:end], dt).reshape(ncols, rows_p).T


def _lane_layout(w, nt=THREADS, phase=1):
    """Lane layout of a workgroup of nt threads (short blocks may run with 64
    or 128); asserts the block fits it."""
    # rows per lane follow the first part's storage type (two-row lanes for
    # the narrow types in phase 2); the second part uses the same rows
    vr = {0: 1, 1: 2, 2: 4 if phase == 1 else 2}[int(w["dt0"])]
    rows, rows_p = int(w["rows"]), int(w["rows_p"])
    assert 0 < rows <= rows_p <= MAX_ROWS * vr and rows_p % vr == 0 and rows_p - rows < vr
    q = rows_p // vr
    assert 1 <= q <= nt, "more lanes per column than threads"
    m = nt // q
    assert m * rows_p <= vr * nt <= ZCHUNK * nt // THREADS, "reduction buffer overflow"
    if int(w["ncol0"]) > 0:
        assert int(w["off0"]) % (vr * DT_BYTES[int(w["dt0"])]) == 0, "lane loads are aligned"
    return vr, q, m


def _stage(tables, w, x, tbuf, tsize, nd):
    zsegs, codes = tables["zseg"], tables["codes"]
    zb, ze = int(w["zseg_begin"]), int(w["zseg_end"])
    assert 0 <= zb <= ze <= len(zsegs)
    kinds = [int(k) for k in zsegs["kind"][zb:ze]]
    # runs (x / t ranges) fit the workgroup's run table; a code table comes alone
    assert (Z_CODES not in kinds and len(kinds) <= RUN_MAX) or kinds == [Z_CODES]
    parts = []
    col = 0
    for zs in zsegs[zb:ze]:
        kind, n, src = int(zs["kind"]), int(zs["ncols"]), int(zs["src"])
        assert n > 0
        if kind != Z_CODES:
            assert int(zs["col"]) == col, "runs are in column order and cover every column"
        col += n
        if kind == Z_X:
            assert 0 <= src and src + n <= nd
            parts.append(x[src:src + n])
        elif kind == Z_T:
            assert 0 <= src and src + n <= tsize, "runs address only final t slots"
            parts.append(tbuf[src:src + n])
        else:
            assert kind == Z_CODES and 0 <= src and src + n <= codes.size
            out = np.empty(n)
            for j, code in enumerate(codes[src:src + n]):
                code = int(code)
                if code >= 0:
                    assert code < nd
                    out[j] = x[code]
                else:
                    slot = -code - 1
                    assert 0 <= slot < tsize, "codes address only final t slots"
                    out[j] = tbuf[slot]
            parts.append(out)
    z = np.concatenate(parts) if parts else np.zeros(0)
    assert z.size == int(w["ncol0"]) + int(w["ncol1"]), "zsegs must cover every column"
    return z


def _block_product(arena, w, z):
    rows_p = int(w["rows_p"])
    y = np.zeros(rows_p)
    n0, n1 = int(w["ncol0"]), int(w["ncol1"])
    if n0:
        y += _part(arena, int(w["off0"]), rows_p, n0, int(w["dt0"])) @ z[:n0]
    if n1:
        y += _part(arena, int(w["off1"]), rows_p, n1, int(w["dt1"])) @ z[n0:]
    return y[:int(w["rows"])]


def _threads(tables):
    """Per phase: work index -> workgroup threads (short blocks first)."""
    out = []
    for k in (1, 2):
        ns, nt = int(tables.get("nwork%d_small" % k, 0)), int(tables.get("nt%d_small" % k, 0))
        assert ns == 0 or nt in (64, 128)
        out.append(lambda i, ns=ns, nt=nt: nt if i < ns else THREADS)
    return out


def run(tables, x, nd):
    """Returns (t, y) computed as the kernel launches would (phase 1, the
    chunk-partial sums, phase 2), with bounds checks.  t is the tsize final
    slots."""
    arena = tables["arena"]
    tsize = int(tables["tsize"])
    tpart_size = int(tables["tpart_size"])
    assert tpart_size >= tsize
    outidx = tables["outidx"]
    tbuf = np.full(max(tpart_size, 1), np.nan)
    written = np.zeros(max(tpart_size, 1), bool)
    nts = _threads(tables)
    for i, w in enumerate(tables["work1"]):
        _lane_layout(w, nts[0](i))
        assert int(w["ncol1"]) == 0
        z = _stage(tables, w, x, tbuf, tsize, nd)
        out = _block_product(arena, w, z)
        ob, rows = int(w["out_base"]), int(w["rows"])
        assert 0 <= ob and ob + rows <= outidx.size
        idx = outidx[ob:ob + rows]
        assert idx.min() >= 0 and idx.max() < tpart_size
        assert not written[idx].any(), "a t entry is written twice"
        written[idx] = True
        tbuf[idx] = out
    fin = tables["tfin"]
    for f in fin:
        slot, base, k = int(f["slot"]), int(f["base"]), int(f["np"])
        assert 0 <= slot < tsize and tsize <= base and base + k <= tpart_size and 2 <= k <= MAX_TPARTS
        assert written[base:base + k].all() and not written[slot]
        s = tbuf[base]
        for p in range(1, k):
            s += tbuf[base + p]
        tbuf[slot] = s
        written[slot] = True
    if tsize:
        assert written[:tsize].all(), "some t slots are never produced"
    y = np.full(nd, np.nan)
    wy = np.zeros(nd, bool)
    for i, w in enumerate(tables["work2"]):
        _lane_layout(w, nts[1](i), phase=2)
        z = _stage(tables, w, x, tbuf, tsize, nd)
        out = _block_product(arena, w, z) if z.size else np.zeros(int(w["rows"]))
        ob, rows = int(w["out_base"]), int(w["rows"])
        assert 0 <= ob and ob + rows <= nd
        assert not wy[ob:ob + rows].any(), "a y row is written twice"
        wy[ob:ob + rows] = True
        y[ob:ob + rows] = out
    assert wy.all(), "some y rows are never written"
    return tbuf[:tsize], y
