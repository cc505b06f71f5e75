"""Coupling graphs of the workloads: the C60 fullerene (truncated icosahedron)
and the Cuthill-McKee site ordering that turns a graph into a chain with a
small bandwidth, so its long-range MPO stays narrow.

The reference takes the fullerene from hopping_fullerene(60) with the default
VARIANT 0 (models/ParamCollection.h:1540-1735, 2002-2010): the numbering of
the IUPAC/Kivelson listing (variant 2) followed by compress_CuthillMcKee.  The
compressor itself lives in a header outside the reference (TOOLS), so the
ordering here is the textbook Cuthill-McKee algorithm (George & Liu, "Computer
Solution of Large Sparse Positive Definite Systems", 1981, section 4.3): a
breadth-first search from a vertex of minimal degree that visits the
unvisited neighbours of each vertex in order of increasing degree, ties broken
by the lower index.  Any relabelling gives the same physics; the ordering only
decides the MPO's bond dimensions.
"""
import numpy as np

# Edges of the truncated icosahedron in the Kivelson numbering: vertex i and
# its neighbours j > i (each edge listed once; 90 edges, every vertex of degree 3).
_C60_UPPER = {
    0: (1, 4, 8), 1: (2, 11), 2: (3, 14), 3: (4, 17), 4: (5,), 5: (6, 19), 6: (7, 21), 7: (8, 22), 8: (9,),
    9: (10, 23), 10: (11, 24), 11: (12,), 12: (13, 25), 13: (14, 26), 14: (15,), 15: (16, 27), 16: (17, 28),
    17: (18,), 18: (19, 29), 19: (20,), 20: (21, 31), 21: (32,), 22: (23, 33), 23: (34,), 24: (25, 35),
    25: (36,), 26: (27, 37), 27: (38,), 28: (29, 39), 29: (30,), 30: (31, 41), 31: (43,), 32: (33, 44),
    33: (46,), 34: (35, 47), 35: (49,), 36: (37, 50), 37: (52,), 38: (39, 53), 39: (40,), 40: (41, 54),
    41: (42,), 42: (43, 56), 43: (44,), 44: (45,), 45: (46, 57), 46: (47,), 47: (48,), 48: (49, 58), 49: (50,),
    50: (51,), 51: (52, 59), 52: (53,), 53: (54,), 54: (55,), 55: (56, 59), 56: (57,), 57: (58,), 58: (59,),
}


def fullerene_c60_edges():
    """The 90 nearest-neighbour bonds (i, j), i < j, in the Kivelson numbering."""
    return [(i, j) for i in sorted(_C60_UPPER) for j in _C60_UPPER[i]]


def adjacency(edges, n, value=1.0):
    """Symmetric coupling matrix with `value` on every edge."""
    A = np.zeros((n, n))
    for i, j in edges:
        A[i, j] = A[j, i] = value
    return A


def cuthill_mckee(A):
    """Cuthill-McKee permutation of the graph with adjacency A (nonzero =
    edge): perm[new] = old.  Each connected component starts at its
    lowest-index vertex of minimal degree."""
    n = A.shape[0]
    nbrs = [np.flatnonzero(A[i]) for i in range(n)]
    deg = np.array([len(x) for x in nbrs])
    seen = np.zeros(n, dtype=bool)
    perm = []
    while len(perm) < n:
        rest = np.flatnonzero(~seen)
        start = int(rest[np.argmin(deg[rest])])  # argmin takes the first (lowest index) minimum
        seen[start] = True
        queue = [start]
        head = 0
        while head < len(queue):
            v = queue[head]
            head += 1
            nxt = sorted((int(u) for u in nbrs[v] if not seen[u]), key=lambda u: (deg[u], u))
            for u in nxt:
                seen[u] = True
                queue.append(u)
        perm.extend(queue)
    return np.array(perm, dtype=np.int64)


def reorder(A, perm):
    """Coupling matrix in the new site order: B[a, b] = A[perm[a], perm[b]]."""
    return A[np.ix_(perm, perm)]


def bandwidth(A):
    i, j = np.nonzero(A)
    return int(np.max(np.abs(i - j))) if len(i) else 0


def c60_couplings(J=1.0, order="cuthill_mckee"):
    """J_ij of the nearest-neighbour Heisenberg model on C60 in chain order
    (`order` "cuthill_mckee" or "kivelson")."""
    A = adjacency(fullerene_c60_edges(), 60, J)
    if order == "kivelson":
        return A
    if order == "cuthill_mckee":
        return reorder(A, cuthill_mckee(A))
    raise ValueError(order)


def cut_ranks(J, tol=1e-12):
    """Rank of the couplings across every bond of the chain: r[b] =
    rank(J[0:b, b:L]) for b = 0..L.  A two-body Hamiltonian sum J_ij O_i O_j
    needs exactly r[b] operator channels at bond b besides the two identity
    channels (done / not yet started)."""
    L = J.shape[0]
    out = [0]
    for b in range(1, L):
        C = J[:b, b:]
        s = np.linalg.svd(C, compute_uv=False) if C.size else np.zeros(0)
        out.append(int(np.sum(s > tol * max(1.0, s[0] if len(s) else 0.0))))
    out.append(0)
    return out


def polygon_counts(A, sizes=(5, 6)):
    """Number of chordless cycles of each length in `sizes` (faces of a
    fullerene are the chordless 5- and 6-cycles)."""
    n = A.shape[0]
    nbrs = [set(np.flatnonzero(A[i]).tolist()) for i in range(n)]
    found = {k: set() for k in sizes}
    kmax = max(sizes)

    def extend(path):
        v = path[-1]
        for u in nbrs[v]:
            if u == path[0] and len(path) in found:
                cyc = path
                # chordless: no edges between non-consecutive vertices
                ok = all(cyc[b] not in nbrs[cyc[a]] for a in range(len(cyc)) for b in range(a + 2, len(cyc))
                         if not (a == 0 and b == len(cyc) - 1))
                if ok:
                    found[len(cyc)].add(frozenset(cyc))
            elif u > path[0] and u not in path and len(path) < kmax:
                extend(path + [u])

    for s in range(n):
        extend([s])
    return {k: len(v) for k, v in found.items()}
