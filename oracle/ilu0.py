"""Block ILU(0) in multicolour order -- TEST INFRASTRUCTURE ONLY.

A CPU statement of the library's preconditioner 3 (no counterpart in the
reference, whose serial path preconditions with the skyline LU of
src/solver/solv_blk_*.c): 3x3 blocks, nodes coloured greedily in index order,
rows in colour-major order, ILU(0) by the IKJ variant without fill.  Used by
the tests as the checker of the device factors; pinned itself by
`tests/test_ilu0_oracle.py` (the product of the factors equals the matrix on
its pattern; on a tree-shaped pattern ILU(0) is the exact LU).  Plain Python
loops over blocks: small meshes only.
"""
import numpy as np


def greedy_colors(row_ptr, col, n):
    """Each node takes the smallest colour none of its lower-index neighbours has."""
    color = np.full(n, -1, dtype=np.int64)
    for i in range(n):
        used = {int(color[j]) for j in col[row_ptr[i]:row_ptr[i + 1]] if j < i}
        k = 0
        while k in used:
            k += 1
        color[i] = k
    return color


def colour_order(color):
    """Rows in colour-major, then node, order: pos[node]."""
    order = np.lexsort((np.arange(color.size), color))
    pos = np.empty(color.size, dtype=np.int64)
    pos[order] = np.arange(color.size)
    return order, pos


class BlockILU0:
    """ILU(0) of a block matrix given as diag[n,3,3] and {(i, j): block} for
    the off-diagonal blocks (i != j), in the order `pos`."""

    def __init__(self, diag, offd, pos):
        n = diag.shape[0]
        self.n = n
        self.pos = pos
        rows = [dict() for _ in range(n)]
        for (i, j), b in offd.items():
            rows[i][j] = np.array(b, dtype=np.float64)
        A = [dict(r) for r in rows]
        D = [np.array(d, dtype=np.float64) for d in diag]
        order = np.argsort(pos)
        self.L = [dict() for _ in range(n)]
        self.U = [dict() for _ in range(n)]
        self.Dinv = [None] * n
        self.Dblk = [None] * n   # U_ii
        for i in order:
            lower = sorted((j for j in A[i] if pos[j] < pos[i]), key=lambda j: pos[j])
            for k in lower:
                lik = A[i][k] @ self.Dinv[k]
                A[i][k] = lik
                for j, ukj in self.U[k].items():   # pos[j] > pos[k]
                    if j == i:
                        D[i] = D[i] - lik @ ukj
                    elif j in A[i]:
                        A[i][j] = A[i][j] - lik @ ukj
            for j, a in A[i].items():
                (self.L if pos[j] < pos[i] else self.U)[i][j] = a
            self.Dblk[i] = D[i]
            self.Dinv[i] = np.linalg.inv(D[i])

    def apply(self, v):
        """M^-1 v for v[3n] (per node u, v, h)."""
        order = np.argsort(self.pos)
        v = np.asarray(v, dtype=np.float64).reshape(self.n, 3)
        z = np.zeros_like(v)
        for i in order:
            s = v[i].copy()
            for k, l in self.L[i].items():
                s -= l @ z[k]
            z[i] = s
        x = np.zeros_like(v)
        for i in order[::-1]:
            s = z[i].copy()
            for j, u in self.U[i].items():
                s -= u @ x[j]
            x[i] = self.Dinv[i] @ s
        return x.reshape(-1)

    def product_on_pattern(self):
        """(L U) restricted to the diagonal and the off-diagonal pattern."""
        n = self.n
        out_d = np.zeros((n, 3, 3))
        out_o = {}
        eye = np.eye(3)
        for i in range(n):
            Li = dict(self.L[i])
            Li[i] = eye
            Ui_rows = {}
            for k, lik in Li.items():
                Uk = dict(self.U[k])
                Uk[k] = self.Dblk[k]
                for j, ukj in Uk.items():
                    Ui_rows[j] = Ui_rows.get(j, 0.0) + lik @ ukj
            out_d[i] = Ui_rows[i]
            for j in list(self.L[i]) + list(self.U[i]):
                out_o[(i, j)] = Ui_rows[j]
        return out_d, out_o
