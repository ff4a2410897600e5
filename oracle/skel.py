"""Recursive skeletonization: ID, elimination, root LU and the solve sweeps.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Restates ``kern_interp/ki_mat.cpp`` (``id`` ``:505-541``, ``LU_factorize``
``:427-433``, ``*_multiply_inverse`` ``:435-493``) and
``kern_interp/skel_factorization/skel_factorization.cpp`` (``id_compress``
``:32-52``, ``decouple`` ``:55-91``, ``skeletonize`` ``:94-243``, Schur
update gathering ``:246-294``, sweeps ``:301-338``, ``solve`` ``:341-403``,
``multiply_connected_solve`` ``:406-551``).  LAPACK is called exactly where
the reference calls it: ``dgeqp3`` and ``dtrtrs`` in the ID, ``dgetrf`` /
``dgetrs`` for every LU, BLAS GEMM through numpy.

Pinned by (``tests/test_oracle_skel.py``): the ID is checked against
scipy.linalg.qr(pivoting=True) and the ID error bound on matrices with a
designed singular-value gap; the factor/solve is checked against a dense
LAPACK solve (backward error <= 10 id_tol, the reference's own criterion,
kern_interp_test.cpp:26-95), against the circle's Sherman-Morrison closed
form, and against analytic interior solutions (kern_interp_test.cpp:297-316).
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack

MIN_DOFS_TO_COMPRESS = 16   # skel_factorization.h:11


class RankError(RuntimeError):
    pass


def interpolative_decomposition(A: np.ndarray, tol: float):
    """``ki_Mat::id`` (ki_mat.cpp:505-541).

    Returns ``(k, perm, T)``: k = first i >= 1 with |R_ii| < |tol R_00|
    (0 means the box is not compressed), ``perm`` the 0-based dgeqp3 pivots,
    ``T = R11^{-1} R12`` (k x (n-k)).
    """
    m, n = A.shape
    assert m != 0 and n != 0
    qr, jpvt, tau, work, info = lapack.dgeqp3(np.asfortranarray(A))
    assert info == 0
    thresh = abs(tol * qr[0, 0])
    k = 0
    for i in range(1, n):
        if i >= m:
            # the reference reads R(i, i) past the last row here (an assert)
            raise RankError(f"ID matrix has {m} rows < {n} columns")
        if abs(qr[i, i]) < thresh:
            k = i
            break
    if k == 0:
        return 0, None, None
    perm = (jpvt - 1).astype(np.int64)
    T, info2 = lapack.dtrtrs(qr[:k, :k], qr[:k, k:], lower=0, trans=0, unitdiag=0)
    assert info2 == 0
    return k, perm, np.asarray(T)


def lu_factor(X):
    """``ki_Mat::LU_factorize`` (ki_mat.cpp:427-433); dgetrf info is not checked there."""
    lu, piv, info = lapack.dgetrf(np.asfortranarray(X))
    return lu, piv


def lu_solve(lu, piv, B, trans=0):
    x, info = lapack.dgetrs(lu, piv, np.asfortranarray(B), trans=trans)
    assert info == 0
    return x


class SkelFactorization:
    """``class SkelFactorization`` (skel_factorization.h:19-62)."""

    def __init__(self, id_tol: float, fact_threads: int = 1):
        assert id_tol > 0
        self.id_tol = id_tol
        self.fact_threads = fact_threads
        self.stats = []   # (level, node id, n, m, k) per compressed box

    # -- factor -------------------------------------------------------------
    def id_compress(self, kernel, tree, node):
        """``id_compress`` (skel_factorization.cpp:32-52).  Returns #redundant or 0."""
        A = kernel.id_matrix(tree, node)
        if A.shape[0] == 0:
            return 0
        k, perm, T = interpolative_decomposition(A, self.id_tol)
        node.m_rows = A.shape[0]
        if k == 0:
            return 0
        ab = node.active_box
        node.T = T
        node.permutation = perm.tolist()
        node.skel = [ab[p] for p in perm[:k]]
        node.redundant = [ab[p] for p in perm[k:]]
        return T.shape[1]

    def schur_updates(self, BN, node):
        """``get_all_schur_updates`` (skel_factorization.cpp:246-294): the sum
        of every compressed descendant's Schur update restricted to ``BN``,
        accumulated in the reference's depth-first order."""
        upd = np.zeros((len(BN), len(BN)))
        if node.is_leaf:
            return upd
        pos = {d: i for i, d in enumerate(BN)}

        def visit(nd):
            for ch in nd.children:
                if ch.compressed:
                    si, bi = [], []
                    for a, d in enumerate(ch.skel):
                        j = pos.get(d)
                        if j is not None:
                            si.append(a)
                            bi.append(j)
                    if si:
                        upd[np.ix_(bi, bi)] += ch.schur[np.ix_(si, si)]
                if not ch.is_leaf:
                    visit(ch)
        visit(node)
        return upd

    def decouple(self, kernel, node):
        """``decouple`` (skel_factorization.cpp:55-91)."""
        k = node.T.shape[0]
        BN = node.active_box
        K = kernel(BN, BN) - self.schur_updates(BN, node)
        s = np.asarray(node.permutation[:k])
        r = np.asarray(node.permutation[k:])
        T = node.T
        Kss, Ksr = K[np.ix_(s, s)], K[np.ix_(s, r)]
        Krs, Krr = K[np.ix_(r, s)], K[np.ix_(r, r)]
        K_r_sn = Krs - T.T @ Kss
        X = Krr - T.T @ Ksr - K_r_sn @ T
        K_sn_r = Ksr - Kss @ T
        lu, piv = lu_factor(X)
        node.X_lu, node.X_piv = lu, piv
        node.L = lu_solve(lu, piv, K_sn_r.T, trans=1).T
        node.U = lu_solve(lu, piv, K_r_sn, trans=0)
        node.schur = node.L @ K_r_sn
        node.compressed = True

    def skeletonize(self, kernel, tree):
        """``skeletonize`` (skel_factorization.cpp:94-243)."""
        nl = len(tree.levels)

        def box(level, node):
            if node.compressed or len(node.active_box) < MIN_DOFS_TO_COMPRESS:
                return None
            if self.id_compress(kernel, tree, node) == 0:
                return None
            self.decouple(kernel, node)
            return (level, node.id, len(node.active_box), node.m_rows, node.T.shape[0])

        for level in range(nl - 1, 0, -1):
            tree.refresh_level(level)
            nodes = tree.levels[level]
            if self.fact_threads > 1 and len(nodes) >= 2 * self.fact_threads:
                self.stats.extend(_level_in_processes(box, level, nodes, self.fact_threads))
            else:
                self.stats.extend(r for r in (box(level, nd) for nd in nodes) if r is not None)
        tree.refresh_all()
        allskel = list(tree.root.active_box)
        tree.allskel = allskel
        if allskel:
            A = kernel(allskel, allskel) - self.schur_updates(allskel, tree.root)
            tree.allskel_mat = A
        if tree.U is None or tree.U.shape[1] == 0:
            if allskel:
                tree.allskel_lu, tree.allskel_piv = lu_factor(tree.allskel_mat)
            return
        self._bordered(kernel, tree, allskel)

    def _compressed_by_level(self, tree):
        for level in range(len(tree.levels) - 1, -1, -1):
            yield [nd for nd in tree.levels[level] if nd.compressed]

    def _bordered(self, kernel, tree, allskel):
        """Bordered system for holes (skel_factorization.cpp:130-242)."""
        ndof = kernel.W.shape[0] * kernel.sd
        red = _complement(allskel, ndof)
        if not red:
            raise RuntimeError("No compression possible")   # the reference exit(0)s
        mU = np.array(tree.U, dtype=np.float64, copy=True)
        mPsiT = np.array(tree.Psi.T, dtype=np.float64, copy=True)
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                sweep(-nd.T, mU, nd.skel, nd.redundant, True)
                sweep(-nd.L, mU, nd.redundant, nd.skel, False)
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                sweep(-nd.T, mPsiT, nd.skel, nd.redundant, True)
                sweep(-nd.U, mPsiT, nd.redundant, nd.skel, True)
        mPsi = mPsiT.T
        rpos = {d: i for i, d in enumerate(red)}
        DinvC = mU[red, :].copy()
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                if nd.redundant:
                    idx = [rpos[d] for d in nd.redundant]
                    DinvC[idx, :] = lu_solve(nd.X_lu, nd.X_piv, DinvC[idx, :])
        h = tree.Psi.shape[0]
        ns = len(allskel)
        S = np.zeros((ns + h, ns + h))
        S[:ns, :ns] = tree.allskel_mat
        S[ns:, :ns] = mPsi[:, allskel]
        S[:ns, ns:] = mU[allskel, :]
        S[ns:, ns:] = -np.eye(h) - mPsi[:, red] @ DinvC
        tree.S_mat = S
        tree.S_lu, tree.S_piv = lu_factor(S)

    # -- solve ----------------------------------------------------------------
    def solve(self, tree, b):
        """``solve`` (skel_factorization.cpp:341-403); any number of columns."""
        x = np.array(b, dtype=np.float64, copy=True)
        squeeze = x.ndim == 1
        if squeeze:
            x = x[:, None]
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                sweep(-nd.T, x, nd.skel, nd.redundant, True)
                sweep(-nd.L, x, nd.redundant, nd.skel, False)
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                if nd.redundant:
                    x[nd.redundant, :] = lu_solve(nd.X_lu, nd.X_piv, x[nd.redundant, :])
        if tree.allskel:
            x[tree.allskel, :] = lu_solve(tree.allskel_lu, tree.allskel_piv,
                                          x[tree.allskel, :])
        for level in range(len(tree.levels)):
            for nd in reversed(tree.levels[level]):
                if not nd.compressed:
                    continue
                sweep(-nd.U, x, nd.skel, nd.redundant, False)
                sweep(-nd.T, x, nd.redundant, nd.skel, False)
        return x[:, 0] if squeeze else x

    def solve_transpose(self, tree, b):
        """x = F^{-T} b for the factorization F that :meth:`solve` inverts.

        No reference counterpart (the reference only solves with F).  The
        sweeps of ``solve`` (skel_factorization.cpp:341-403) transposed and
        taken in reverse order: the down pass's steps, finest level first,
        as x_r -= T^T x_s, x_s -= U^T x_r; then X_rr^{-T} and
        allskel_mat^{-T} (``dgetrs`` with 'T'); then the up pass's steps,
        coarsest level first, as x_r -= L^T x_s, x_s -= T x_r.  Pinned by the
        adjoint identity <F^{-T} g, b> = <g, F^{-1} b> and, at tight id_tol,
        by the dense K^{-T} g.
        """
        x = np.array(b, dtype=np.float64, copy=True)
        squeeze = x.ndim == 1
        if squeeze:
            x = x[:, None]
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                sweep(-nd.T, x, nd.skel, nd.redundant, True)
                sweep(-nd.U, x, nd.redundant, nd.skel, True)
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                if nd.redundant:
                    x[nd.redundant, :] = lu_solve(nd.X_lu, nd.X_piv, x[nd.redundant, :],
                                                  trans=1)
        if tree.allskel:
            x[tree.allskel, :] = lu_solve(tree.allskel_lu, tree.allskel_piv,
                                          x[tree.allskel, :], trans=1)
        for level in range(len(tree.levels)):
            for nd in tree.levels[level]:
                if not nd.compressed:
                    continue
                sweep(-nd.L, x, nd.skel, nd.redundant, True)
                sweep(-nd.T, x, nd.redundant, nd.skel, False)
        return x[:, 0] if squeeze else x

    def multiply_connected_solve(self, tree, b):
        """``multiply_connected_solve`` (skel_factorization.cpp:406-551). Returns (mu, alpha)."""
        allskel = tree.allskel
        red = _complement(allskel, b.shape[0])
        mu = np.array(b, dtype=np.float64, copy=True)[:, None]
        mU = np.array(tree.U, dtype=np.float64, copy=True)
        mPsiT = np.array(tree.Psi.T, dtype=np.float64, copy=True)
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                sweep(-nd.T, mu, nd.skel, nd.redundant, True)
                sweep(-nd.L, mu, nd.redundant, nd.skel, False)
                sweep(-nd.T, mU, nd.skel, nd.redundant, True)
                sweep(-nd.L, mU, nd.redundant, nd.skel, False)
        w = mu[red, :].copy()
        z = mu[allskel, :].copy()
        for nodes in self._compressed_by_level(tree):
            for nd in nodes:
                sweep(-nd.T, mPsiT, nd.skel, nd.redundant, True)
                sweep(-nd.U, mPsiT, nd.redundant, nd.skel, True)
        mPsi = mPsiT.T
        rpos = {d: i for i, d in enumerate(red)}

        def dinv(v):
            v = v.copy()
            for nodes in self._compressed_by_level(tree):
                for nd in nodes:
                    if nd.redundant:
                        idx = [rpos[d] for d in nd.redundant]
                        v[idx, :] = lu_solve(nd.X_lu, nd.X_piv, v[idx, :])
            return v

        Dw = dinv(w)
        M = np.vstack([z, -(mPsi[:, red] @ Dw)])
        y = lu_solve(tree.S_lu, tree.S_piv, M)
        ns = len(allskel)
        alpha = y[ns:, :]
        N = w - mU[red, :] @ alpha
        mu[red, :] = dinv(N)
        mu[allskel, :] = y[:ns, :]
        for level in range(len(tree.levels)):
            for nd in reversed(tree.levels[level]):
                if not nd.compressed:
                    continue
                sweep(-nd.U, mu, nd.skel, nd.redundant, False)
                sweep(-nd.T, mu, nd.redundant, nd.skel, False)
        return mu[:, 0], alpha[:, 0]


_FORK_STATE = {}
_BOX_FIELDS = ("T", "permutation", "skel", "redundant", "m_rows", "X_lu", "X_piv", "L", "U",
               "schur", "compressed")


def _fork_worker(args):
    """Runs the boxes [a, b) of the level in a forked copy of the state."""
    a, b = args
    box, level, nodes = _FORK_STATE["job"]
    out = []
    for i in range(a, b):
        st = box(level, nodes[i])
        if st is not None:
            nd = nodes[i]
            out.append((i, st, {f: getattr(nd, f) for f in _BOX_FIELDS}))
    return out


def _level_in_processes(box, level, nodes, workers):
    """The boxes of one level in `workers` forked processes (the
    reference runs them on `fact_threads` OpenMP threads,
    skel_factorization.cpp:97-113); each child works on a copy-on-write image
    of the tree, and the compressed boxes' results are copied back."""
    import multiprocessing as mp
    _FORK_STATE["job"] = (box, level, nodes)
    n = len(nodes)
    cuts = [n * w // workers for w in range(workers + 1)]
    with mp.get_context("fork").Pool(workers) as pool:
        parts = pool.map(_fork_worker, [(cuts[w], cuts[w + 1]) for w in range(workers)])
    stats = []
    for part in parts:
        for i, st, fields in part:
            for f, v in fields.items():
                setattr(nodes[i], f, v)
            stats.append(st)
    _FORK_STATE.clear()
    return stats


def sweep(M, vec, a, b, transpose):
    """``apply_sweep_matrix`` (skel_factorization.cpp:301-318): vec[b] += op(M) vec[a]."""
    if len(a) * len(b) == 0:
        return
    prod = (M.T if transpose else M) @ vec[a, :]
    vec[b, :] = prod + vec[b, :]


def _complement(allskel, ndof):
    s = set(allskel)
    return [i for i in range(ndof) if i not in s]
