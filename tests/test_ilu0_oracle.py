"""Pins of the CPU block ILU(0) (oracle/ilu0.py), the checker of the
library's preconditioner 3.

ILU(0) is fixed by two facts independent of how it is computed: the product
of its factors equals the matrix on the matrix's own pattern (diagonal and
off-diagonal blocks), and on a tree-shaped pattern eliminated leaves first no
fill arises, so it is the exact LU and M^-1 A = I.  A transposed block product,
a wrong elimination order or a dropped update fails one of them.  The
greedy colouring is checked to be proper.  CPU only.
"""
import numpy as np

from adhv8_kraken_amd import mesh
from oracle.ilu0 import BlockILU0, colour_order, greedy_colors


def _adjacency(p):
    e = np.asarray(p.e2d).reshape(-1, 3)
    nb = [set() for _ in range(p.nnodes)]
    for a in range(3):
        for b in range(3):
            if a != b:
                for i, j in zip(e[:, a], e[:, b]):
                    nb[i].add(int(j))
    row_ptr = np.zeros(p.nnodes + 1, dtype=np.int64)
    col = []
    for i in range(p.nnodes):
        s = sorted(nb[i])
        col.extend(s)
        row_ptr[i + 1] = row_ptr[i] + len(s)
    return row_ptr, np.array(col, dtype=np.int64)


def _random_system(row_ptr, col, n, seed):
    rng = np.random.default_rng(seed)
    diag = rng.uniform(-1, 1, (n, 3, 3)) + 8.0 * np.eye(3)
    offd = {}
    for i in range(n):
        for j in col[row_ptr[i]:row_ptr[i + 1]]:
            offd[(i, int(j))] = rng.uniform(-1, 1, (3, 3))
    return diag, offd


def _dense(diag, offd, n):
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = diag[i]
    for (i, j), b in offd.items():
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = b
    return A


def test_greedy_colouring_is_proper():
    p = mesh.bumpy_channel(20, 10)
    rp, col = _adjacency(p)
    color = greedy_colors(rp, col, p.nnodes)
    for i in range(p.nnodes):
        assert all(color[j] != color[i] for j in col[rp[i]:rp[i + 1]])
    assert color.max() + 1 <= 7


def test_factors_reproduce_the_matrix_on_its_pattern():
    p = mesh.bumpy_channel(12, 6)
    rp, col = _adjacency(p)
    n = p.nnodes
    diag, offd = _random_system(rp, col, n, 3)
    _, pos = colour_order(greedy_colors(rp, col, n))
    M = BlockILU0(diag, offd, pos)
    pd, po = M.product_on_pattern()
    scale = max(np.abs(diag).max(), 1.0)
    assert np.abs(pd - diag).max() <= 1e-12 * scale
    assert max(np.abs(po[k] - offd[k]).max() for k in offd) <= 1e-12 * scale
    # and the apply inverts the product of the factors
    v = np.random.default_rng(4).uniform(-1, 1, 3 * n)
    x = M.apply(v)
    LU = np.zeros((3 * n, 3 * n))
    for i in range(n):
        LU[3 * i:3 * i + 3, 3 * i:3 * i + 3] = pd[i]
    # the full product (with fill outside the pattern) applied to x gives v
    full = np.zeros((3 * n, 3 * n))
    Lf = np.eye(3 * n)
    Uf = np.zeros((3 * n, 3 * n))
    for i in range(n):
        Uf[3 * i:3 * i + 3, 3 * i:3 * i + 3] = M.Dblk[i]
        for k, b in M.L[i].items():
            Lf[3 * i:3 * i + 3, 3 * k:3 * k + 3] = b
        for j, b in M.U[i].items():
            Uf[3 * i:3 * i + 3, 3 * j:3 * j + 3] = b
    assert np.abs(Lf @ (Uf @ x) - v).max() <= 1e-12 * np.abs(v).max()


def test_ilu0_on_a_tree_pattern_is_the_exact_lu():
    # a binary tree of 40 nodes eliminated leaves first (every node's only
    # later neighbour is its parent): elimination creates no fill, so the
    # incomplete factorisation is the complete one
    n = 40
    edges = [(i, (i - 1) // 2) for i in range(1, n)]
    nb = [set() for _ in range(n)]
    for a, b in edges:
        nb[a].add(b)
        nb[b].add(a)
    rp = np.zeros(n + 1, dtype=np.int64)
    col = []
    for i in range(n):
        col.extend(sorted(nb[i]))
        rp[i + 1] = rp[i] + len(nb[i])
    col = np.array(col)
    diag, offd = _random_system(rp, col, n, 7)
    pos = np.arange(n)[::-1].copy()   # children (higher index) before parents
    M = BlockILU0(diag, offd, pos)
    A = _dense(diag, offd, n)
    b = np.random.default_rng(8).uniform(-1, 1, 3 * n)
    x = M.apply(b)
    assert np.abs(A @ x - b).max() <= 1e-11 * np.abs(b).max()
