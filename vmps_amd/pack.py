"""Flat packing of block-sparse tensors for the bulk builder calls
(vmps_pivot2_add_env_sectors, vmps_pivot2_add_W_blocks, vmps_pivot2_add_V_blocks,
vmps_env_add_old_sectors, vmps_env_add_A_blocks, vmps_env_add_W_blocks): one
C call per Tripod, W slot, MPS tensor or vector layout instead of one per
sector or block.

Layouts follow include/vmps_hxv.h: quantum numbers nq ints per item;
environment and MPS blocks are passed as one pointer per block (column-major
float64, no packing copy), W nonzeros back to back; an absent environment
block is rows = cols = 0."""
import ctypes

import numpy as np

P_int = ctypes.POINTER(ctypes.c_int)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_dbl = ctypes.POINTER(ctypes.c_double)
P_ptr = ctypes.POINTER(ctypes.c_void_p)


def _qns(seq, nq):
    a = np.array(list(seq), dtype=np.int32).reshape(-1)
    return np.ascontiguousarray(a if a.size else np.zeros(max(1, nq), dtype=np.int32))


def _cat(parts, dtype):
    if not parts:
        return np.zeros(1, dtype=dtype)
    return np.ascontiguousarray(np.concatenate(parts), dtype=dtype)


class Packed:
    """Arrays of one bulk call; `args()` gives the ctypes pointers in
    declaration order (the arrays stay alive with this object)."""

    def __init__(self, count, arrays, kinds):
        self.count = count
        self.arrays = arrays
        self.kinds = kinds

    def args(self):
        return [a.ctypes.data_as(k) for a, k in zip(self.arrays, self.kinds)]


def _block_ptrs(blocks):
    """Pointers to column-major float64 copies-if-needed of `blocks` (None:
    absent, NULL).  Returns (pointer array, arrays to keep alive)."""
    keep, ptrs = [], np.zeros(max(1, len(blocks)), dtype=np.uint64)
    for i, blk in enumerate(blocks):
        if blk is None or blk.size == 0:
            continue
        a = np.asfortranarray(blk, dtype=np.float64)
        keep.append(a)
        ptrs[i] = a.ctypes.data
    return ptrs, keep


def tripod(sectors, nq):
    """[(qin, qout, qmid, [block or None, ...]), ...] -> nsec, qin, qout,
    qmid, adim, rows, cols, block pointers."""
    qin = _qns((x for s in sectors for x in s[0]), nq)
    qout = _qns((x for s in sectors for x in s[1]), nq)
    qmid = _qns((x for s in sectors for x in s[2]), nq)
    adim = np.array([len(s[3]) for s in sectors] or [0], dtype=np.int32)
    blocks = [blk for s in sectors for blk in s[3]]
    rows = np.array([0 if b is None else b.shape[0] for b in blocks] or [0], dtype=np.int64)
    cols = np.array([0 if b is None else b.shape[1] for b in blocks] or [0], dtype=np.int64)
    ptrs, keep = _block_ptrs(blocks)
    pk = Packed(len(sectors), [qin, qout, qmid, adim, rows, cols, ptrs],
                [P_int, P_int, P_int, P_int, P_i64, P_i64, P_ptr])
    pk.keep = keep
    return pk


def w_blocks(blocks, nq):
    """[(s1, s2, k, qin, qout, rows, cols, row[], col[], val[]), ...] -> nblk,
    idx, qin, qout, dims, nnz, row, col, val."""
    idx = np.array([(b[0], b[1], b[2]) for b in blocks] or [(0, 0, 0)], dtype=np.int32).reshape(-1)
    qin = _qns((x for b in blocks for x in b[3]), nq)
    qout = _qns((x for b in blocks for x in b[4]), nq)
    dims = np.array([(b[5], b[6]) for b in blocks] or [(0, 0)], dtype=np.int32).reshape(-1)
    nnz = np.array([len(b[9]) for b in blocks] or [0], dtype=np.int64)
    row = _cat([np.asarray(b[7]).ravel() for b in blocks], np.int32)
    col = _cat([np.asarray(b[8]).ravel() for b in blocks], np.int32)
    val = _cat([np.asarray(b[9]).ravel() for b in blocks], np.float64)
    return Packed(len(blocks), [np.ascontiguousarray(idx), qin, qout, np.ascontiguousarray(dims), nnz, row, col, val],
                  [P_int, P_int, P_int, P_int, P_i64, P_int, P_int, P_dbl])


def a_blocks(A, nq):
    """A[s] = [(qin, qout, block), ...] -> nblk, s, qin, qout, rows, cols,
    block pointers."""
    items = [(s, qi, qo, blk) for s, blocks in enumerate(A) for (qi, qo, blk) in blocks]
    s = np.array([it[0] for it in items] or [0], dtype=np.int32)
    qin = _qns((x for it in items for x in it[1]), nq)
    qout = _qns((x for it in items for x in it[2]), nq)
    rows = np.array([it[3].shape[0] for it in items] or [0], dtype=np.int64)
    cols = np.array([it[3].shape[1] for it in items] or [0], dtype=np.int64)
    ptrs, keep = _block_ptrs([it[3] for it in items])
    pk = Packed(len(items), [s, qin, qout, rows, cols, ptrs], [P_int, P_int, P_int, P_i64, P_i64, P_ptr])
    pk.keep = keep
    return pk


def v_blocks(V, nq):
    """V[s] = [(qin, qout, rows, cols), ...] -> nblk, s, qin, qout, rows, cols
    (the offsets array is the caller's)."""
    items = [(s, qi, qo, r, c) for s, blocks in enumerate(V) for (qi, qo, r, c) in blocks]
    s = np.array([it[0] for it in items] or [0], dtype=np.int32)
    qin = _qns((x for it in items for x in it[1]), nq)
    qout = _qns((x for it in items for x in it[2]), nq)
    rows = np.array([it[3] for it in items] or [0], dtype=np.int64)
    cols = np.array([it[4] for it in items] or [0], dtype=np.int64)
    return Packed(len(items), [s, qin, qout, rows, cols], [P_int, P_int, P_int, P_i64, P_i64])
