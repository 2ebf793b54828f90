"""Hand-made leaf sets (stc_HACApK_leafmtx records, HACApK_MAGMA.h:12-20) for
tests that need block structures the mesh generator does not produce.

Layout of the factor storage follows the reference's leaf records: a dense
leaf holds a1 as ndt x ndl (column-major, i.e. A row-major); a low-rank leaf
holds a1 = V (ndt x kt, column-major) followed in the same allocation by
a2 = U (ndl x kt, column-major), a1size bytes after a1.
"""
import ctypes as C

import numpy as np


class LeafSet:
    """nd unknowns and a list of leaves (kind, row0, nrows, col0, ncols, rank),
    0-based, filled from a seeded generator."""

    def __init__(self, H, nd, leaves, seed=0):
        rng = np.random.default_rng(seed)
        self.nd = nd
        self.nlf = len(leaves)
        self._rec = (H.LeafRec * self.nlf)()
        self._bufs = []
        self._table = []
        for i, (kind, r0, nr, c0, nc, k) in enumerate(leaves):
            assert 0 <= r0 and r0 + nr <= nd and 0 <= c0 and c0 + nc <= nd
            rec = self._rec[i]
            rec.nstrtl, rec.ndl, rec.nstrtt, rec.ndt = r0 + 1, nr, c0 + 1, nc
            if kind == "dense":
                buf = rng.uniform(-1, 1, nc * nr)
                rec.ltmtx, rec.kt = 2, 0
                rec.a1size = buf.nbytes
                rec.a1 = buf.ctypes.data_as(C.POINTER(C.c_double))
                self._table.append([2, 0, r0 + 1, nr, c0 + 1, nc])
            else:
                v = rng.uniform(-1, 1, nc * k)
                u = rng.uniform(-1, 1, nr * k)
                buf = np.concatenate([v, u])
                rec.ltmtx, rec.kt = 1, k
                rec.a1size = v.nbytes
                rec.a1 = buf.ctypes.data_as(C.POINTER(C.c_double))
                rec.a2 = C.cast(buf.ctypes.data + v.nbytes, C.POINTER(C.c_double))
                self._table.append([1, k, r0 + 1, nr, c0 + 1, nc])
            self._bufs.append(buf)
        self.leaves_ptr = C.addressof(self._rec)
        self.stride = C.sizeof(H.LeafRec)

    def leaf_table(self):
        return np.array(self._table, dtype=np.int64).reshape(-1, 6)


def many_runs(H, seed=0):
    """Row slices with more input runs than a workgroup's run table holds
    (100 one-column dense leaves at every other column), with some runs
    (30 three-column dense leaves), and ordinary dense and low-rank leaves."""
    nd = 512
    lv = []
    lv += [("dense", 0, 16, 2 * j, 1, 0) for j in range(100)]
    lv += [("dense", 16, 32, 4 * j, 3, 0) for j in range(30)]
    lv += [("lowrank", 0, 48, 300, 100, 5)]
    lv += [("lowrank", 48, 464, 0, 256, 7)]
    lv += [("dense", 48, 64, 256, 64, 0)]
    lv += [("lowrank", 112, 400, 256, 256, 3)]
    return LeafSet(H, nd, lv, seed)


class DenseBlocks:
    """nd unknowns covered by dense leaves of given values: blocks is a list
    of (row0, col0, M), 0-based, M an (ndl, ndt) array; a1 holds M row-major
    (ndt x ndl column-major, as the reference stores a dense leaf)."""

    def __init__(self, H, nd, blocks):
        self.nd = nd
        self.nlf = len(blocks)
        self._rec = (H.LeafRec * self.nlf)()
        self._bufs = []
        self._table = []
        for i, (r0, c0, M) in enumerate(blocks):
            M = np.asarray(M, np.float64)
            nr, nc = M.shape
            assert 0 <= r0 and r0 + nr <= nd and 0 <= c0 and c0 + nc <= nd
            buf = np.ascontiguousarray(M.reshape(-1))
            rec = self._rec[i]
            rec.ltmtx, rec.kt = 2, 0
            rec.nstrtl, rec.ndl, rec.nstrtt, rec.ndt = r0 + 1, nr, c0 + 1, nc
            rec.a1size = buf.nbytes
            rec.a1 = buf.ctypes.data_as(C.POINTER(C.c_double))
            self._bufs.append(buf)
            self._table.append([2, 0, r0 + 1, nr, c0 + 1, nc])
        self.leaves_ptr = C.addressof(self._rec)
        self.stride = C.sizeof(H.LeafRec)
        self.dense = np.zeros((nd, nd))
        for r0, c0, M in blocks:
            M = np.asarray(M, np.float64)
            self.dense[r0:r0 + M.shape[0], c0:c0 + M.shape[1]] += M

    def leaf_table(self):
        return np.array(self._table, dtype=np.int64).reshape(-1, 6)


SKEW = [[0.0, 1.0], [-1.0, 0.0]]  # (b, A b) = 0 for every b: r^.v = 0 at step 1


def skew_system(H):
    """2 x 2 skew-symmetric: BiCGSTAB's alpha = (r^,r)/(r^,A p) divides by 0."""
    return DenseBlocks(H, 2, [(0, 0, SKEW)])


def skew_and_scaled_identity(H):
    """4 x 4 block diagonal [SKEW, 2 I]: right-hand sides on the first two
    unknowns break down, on the last two the first half step solves
    exactly (t = r - alpha A p = 0)."""
    return DenseBlocks(H, 4, [(0, 0, SKEW), (2, 2, [[2.0, 0.0], [0.0, 2.0]])])


def scaled_identity(H):
    """2 I on two unknowns: the first half step of BiCGSTAB solves exactly."""
    return DenseBlocks(H, 2, [(0, 0, [[2.0, 0.0], [0.0, 2.0]])])


def with_nan(H, value=float("nan")):
    """[[2, value], [0, 3]]: a leaf holding a non-finite entry."""
    return DenseBlocks(H, 2, [(0, 0, [[2.0, value], [0.0, 3.0]])])
