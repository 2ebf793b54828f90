"""CPU emulation of the engine's two streaming phases over a host plan.

Test helper: walks the same segment/work tables the HIP kernels walk
(hmv_kernels.hip), asserts that every access the kernel would make stays
inside its buffer, and computes t and y in float64 from the stored (FP64,
FP32 or BF16) entries.  Used to validate the repack on the CPU before any
GPU run; it is not an engine path.
"""
import numpy as np

ZCHUNK = 2048
MAX_ROWS = 256
DT_BYTES = {0: 8, 1: 4, 2: 2}


def _decode(raw_bytes, dtype_code):
    if dtype_code == 0:
        return raw_bytes.view("<f8").astype(np.float64)
    if dtype_code == 1:
        return raw_bytes.view("<f4").astype(np.float64)
    u16 = raw_bytes.view("<u2").astype(np.uint32) << 16
    return u16.view("<f4").astype(np.float64)


def _rows(arena, seg, c0, cwp):
    es = DT_BYTES[int(seg["dtype"])]
    ld, nrows = int(seg["ld"]), int(seg["nrows"])
    start = int(seg["data_off"]) + c0 * es
    assert start % 16 == 0, "segment rows must be 16-byte aligned"
    assert (ld * es) % 16 == 0, "row stride must be a multiple of 16 bytes"
    last = start + ((nrows - 1) * ld + cwp) * es
    assert 0 <= start and last <= arena.size, "segment reads outside the arena"
    out = np.empty((nrows, cwp))
    for r in range(nrows):
        b0 = start + r * ld * es
        out[r] = _decode(arena[b0:b0 + cwp * es], int(seg["dtype"]))
    return out


def check_lpr(seg):
    lpr = int(seg["lpr"])
    assert lpr in (4, 8, 16, 32, 64, 256), lpr


def run(tables, x, nd):
    """Returns (t, y) computed as the two kernels would, with bounds checks."""
    arena = tables["arena"]
    tsize = int(tables["tsize"])
    slotmap, zidx = tables["slotmap"], tables["zidx"]
    t = np.full(max(tsize, 1), np.nan)
    written_t = np.zeros(max(tsize, 1), bool)
    for w in tables["work1"]:
        nout = int(w["nout"])
        assert 0 < nout <= MAX_ROWS
        acc = np.zeros(nout)
        for si in range(int(w["seg_begin"]), int(w["seg_end"])):
            s = tables["seg1"][si]
            check_lpr(s)
            nx, ncols, nrows = int(s["nx"]), int(s["ncols"]), int(s["nrows"])
            assert nx == ncols, "phase-1 segments read only x"
            assert 0 <= int(s["acc_row"]) and int(s["acc_row"]) + nrows <= nout
            for c0 in range(0, ncols, ZCHUNK):
                cw = min(ZCHUNK, ncols - c0)
                vec = 16 // DT_BYTES[int(s["dtype"])]
                cwp = (cw + vec - 1) // vec * vec
                assert 0 <= int(s["xsrc"]) + c0 and int(s["xsrc"]) + c0 + cw <= nd
                z = np.zeros(cwp)
                z[:cw] = x[int(s["xsrc"]) + c0:int(s["xsrc"]) + c0 + cw]
                A = _rows(arena, s, c0, cwp)
                acc[int(s["acc_row"]):int(s["acc_row"]) + nrows] += A @ z
        ob = int(w["out_base"])
        assert 0 <= ob and ob + nout <= slotmap.size
        slots = slotmap[ob:ob + nout]
        assert slots.min() >= 0 and slots.max() < tsize
        assert not written_t[slots].any(), "a t entry is written twice"
        written_t[slots] = True
        t[slots] = acc
    if tsize:
        assert written_t[:tsize].all(), "some t entries are never written"
    y = np.full(nd, np.nan)
    written_y = np.zeros(nd, bool)
    for w in tables["work2"]:
        nout = int(w["nout"])
        assert 0 < nout <= MAX_ROWS
        acc = np.zeros(nout)
        for si in range(int(w["seg_begin"]), int(w["seg_end"])):
            s = tables["seg2"][si]
            check_lpr(s)
            nx, ncols, nrows = int(s["nx"]), int(s["ncols"]), int(s["nrows"])
            ar = int(s["acc_row"])
            assert 0 <= ar and ar + nrows <= nout
            for c0 in range(0, ncols, ZCHUNK):
                cw = min(ZCHUNK, ncols - c0)
                vec = 16 // DT_BYTES[int(s["dtype"])]
                cwp = (cw + vec - 1) // vec * vec
                z = np.zeros(cwp)
                gc = np.arange(c0, c0 + cw)
                xm = gc < nx
                if xm.any():
                    zi = int(s["xsrc"]) + gc[xm]
                    assert zi.min() >= 0 and zi.max() < zidx.size
                    xi = zidx[zi]
                    assert xi.min() >= 0 and xi.max() < nd
                    z[:cw][xm] = x[xi]
                if (~xm).any():
                    ti = int(s["tsrc"]) + gc[~xm] - nx
                    assert ti.min() >= 0 and ti.max() < tsize
                    z[:cw][~xm] = t[ti]
                A = _rows(arena, s, c0, cwp)
                acc[ar:ar + nrows] += A @ z
        ob = int(w["out_base"])
        assert 0 <= ob and ob + nout <= nd
        assert not written_y[ob:ob + nout].any(), "a y row is written twice"
        written_y[ob:ob + nout] = True
        y[ob:ob + nout] = acc
    assert written_y.all(), "some y rows are never written"
    return t[:tsize], y
