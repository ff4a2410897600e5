"""Nystrom matrix entries, proxy blocks and the per-box ID matrix.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Restates ``kern_interp/kernel/kernel.cpp``: Laplace double layer and its
Neumann variant (``:11-36``), the Stokes double layer in 2x2 DOF blocks
(``:47-96``), proxy rings (``:238-334``), the ID matrix (``:153-235``) and
the boundary-to-domain ``forward`` matrix (``:337-357``).  Each entry is
evaluated with the same sequence of IEEE double operations as the reference
(no fused multiply-add), so entries are reproducible bit for bit.

Pinned by (``tests/test_oracle_kernel.py``): on a circle of radius R every
off-diagonal Laplace double-layer entry equals ``w / (4 pi R)`` and the
diagonal ``1/2 + w / (4 pi R)``; Gauss' law for the double layer
(``sum_s K(t, s) = 1`` on the boundary for the Dirichlet form); the Stokes
double layer reproduces a rigid rotation's traction-free identity; proxy
rows equal direct kernel evaluations with the proxy ring as sources/targets.
"""
from __future__ import annotations

import math

import numpy as np

LAPLACE, LAPLACE_NEUMANN, STOKES = 0, 1, 2
NUM_PROXY_POINTS = 128          # kernel.h:10
RADIUS_RATIO = 1.5              # kernel.h:11
INV_2PI = 1.0 / (2 * math.pi)   # the constant (1.0 / (2 * M_PI)) of kernel.cpp:17-29
INV_PI = 1.0 / math.pi          # kernel.cpp:48


class Kernel:
    """``struct Kernel`` (kernel.h:15-69): owns copies of the boundary arrays."""

    def __init__(self, sd: int, pde: int, boundary, domain_points=None):
        self.sd = sd
        self.pde = pde
        self.P = np.asarray(boundary.points, np.float64).reshape(-1, 2)
        self.Nm = np.asarray(boundary.normals, np.float64).reshape(-1, 2)
        self.W = np.asarray(boundary.weights, np.float64)
        self.C = np.asarray(boundary.curvatures, np.float64)
        dp = np.zeros(0) if domain_points is None else np.asarray(domain_points, np.float64)
        self.D = dp.reshape(-1, 2)

    # -- dense blocks -----------------------------------------------------
    def __call__(self, tgt, src, forward=False):
        tgt = np.asarray(tgt, np.int64)
        src = np.asarray(src, np.int64)
        if self.pde == STOKES:
            return self._stokes(tgt, src, forward)
        return self._laplace(tgt, src, forward)

    def _laplace(self, tgt, src, forward):
        """``laplace_get`` (kernel.cpp:114-150) over ``laplace`` (:11-36)."""
        sp, sn = self.P[src], self.Nm[src]
        sw, sc = self.W[src][None, :], self.C[src][None, :]
        if forward:
            tp = self.D[tgt]
            tn = np.zeros_like(tp)
        else:
            tp, tn = self.P[tgt], self.Nm[tgt]
        r1 = tp[:, 0:1] - sp[None, :, 0]
        r2 = tp[:, 1:2] - sp[None, :, 1]
        return laplace_entries(self.pde, r1, r2, sw, sc, sn[None, :, 0], sn[None, :, 1],
                               tn[:, 0:1], tn[:, 1:2], forward)

    def _stokes(self, tgt, src, forward):
        """``stokes_get`` (kernel.cpp:47-96): DOF = 2*point + component."""
        sp_i, tp_i = src // 2, tgt // 2
        scomp, tcomp = src % 2, tgt % 2
        sp, sn = self.P[sp_i], self.Nm[sp_i]
        sw, sc = self.W[sp_i][None, :], self.C[sp_i][None, :]
        if forward:
            tp = self.D[tp_i]
            tn = np.zeros_like(tp)
        else:
            tp, tn = self.P[tp_i], self.Nm[tp_i]
        r0 = tp[:, 0:1] - sp[None, :, 0]
        r1 = tp[:, 1:2] - sp[None, :, 1]
        rt = np.where(tcomp[:, None] == 0, r0, r1)
        rs = np.where(scomp[None, :] == 0, r0, r1)
        tn_t = np.where(tcomp == 0, tn[:, 0], tn[:, 1])[:, None]
        sn_s = np.where(scomp == 0, sn[:, 0], sn[:, 1])[None, :]
        with np.errstate(divide="ignore", invalid="ignore"):
            rr = r0 * r0 + r1 * r1
            off = sw * INV_PI * (r0 * sn[None, :, 0] + r1 * sn[None, :, 1]) / (rr * rr) \
                * rt * rs + sw * tn_t * sn_s
        same = (tp[:, 0:1] == sp[None, :, 0]) & (tp[:, 1:2] == sp[None, :, 1])
        if same.any():
            parity = (tgt[:, None] + src[None, :] + 1) % 2
            sn_a = np.where((tcomp[:, None] + 1) % 2 == 0, sn[None, :, 0], sn[None, :, 1])
            sn_b = np.where((scomp[None, :] + 1) % 2 == 0, sn[None, :, 0], sn[None, :, 1])
            diag = -0.5 * parity + (1 - 2 * parity) * 0.5 * sc * sw * INV_PI * sn_a * sn_b \
                + sw * tn_t * sn_s
            off = np.where(same, diag, off)
        return off

    def forward(self):
        """``Kernel::forward`` (kernel.cpp:337-357): domain targets x boundary sources."""
        nt = self.D.shape[0] * (1 if self.sd == 1 else 2)
        ns = self.P.shape[0] * self.sd
        return self(np.arange(nt), np.arange(ns), forward=True)

    # -- proxy and ID matrix ------------------------------------------------
    def proxy_block(self, cx, cy, r, box):
        """``get_proxy_mat`` + ``laplace_proxy_get`` / ``stokes_proxy_get`` (kernel.cpp:238-334)."""
        px, py, pnx, pny = proxy_ring(cx, cy, r)
        pw = 2.0 * math.pi * r / NUM_PROXY_POINTS
        box = np.asarray(box, np.int64)
        if self.pde == STOKES:
            return self._stokes_proxy(px, py, pnx, pny, pw, box)
        bp, bn, bw = self.P[box], self.Nm[box], self.W[box][None, :]
        r1 = px[:, None] - bp[None, :, 0]
        r2 = py[:, None] - bp[None, :, 1]
        zero = np.zeros_like(bw)
        top = laplace_entries(self.pde, r1, r2, bw, zero, bn[None, :, 0], bn[None, :, 1],
                              pnx[:, None], pny[:, None], False)
        bot = laplace_entries(self.pde, -r1, -r2, np.full_like(bw, pw), zero,
                              pnx[:, None], pny[:, None], bn[None, :, 0], bn[None, :, 1], False)
        return np.vstack([top, bot])

    def _stokes_proxy(self, px, py, pnx, pny, pw, box):
        pi_, comp = box // 2, box % 2
        sp, sn, sw = self.P[pi_], self.Nm[pi_], self.W[pi_][None, :]
        r0 = px[:, None] - sp[None, :, 0]
        r1 = py[:, None] - sp[None, :, 1]
        rs = np.where(comp[None, :] == 0, r0, r1)
        sns = np.where(comp == 0, sn[:, 0], sn[:, 1])[None, :]
        rr = r0 * r0 + r1 * r1
        pot = sw * INV_PI * (r0 * sn[None, :, 0] + r1 * sn[None, :, 1]) / (rr * rr)
        a0 = pot * r0 * rs + sw * pnx[:, None] * sns
        a1 = pot * r1 * rs + sw * pny[:, None] * sns
        pot2 = -pw * INV_PI * (r0 * pnx[:, None] + r1 * pny[:, None]) / (rr * rr)
        b0 = pot2 * r0 * rs + pw * sns * pnx[:, None]
        b1 = pot2 * r1 * rs + pw * sns * pny[:, None]
        n = box.size
        top = np.empty((2 * NUM_PROXY_POINTS, n))
        top[0::2], top[1::2] = a0, a1
        bot = np.empty((2 * NUM_PROXY_POINTS, n))
        bot[0::2], bot[1::2] = b0, b1
        return np.vstack([top, bot])

    def id_rows(self, tree, node):
        """Row DOF lists of the ID matrix (kernel.cpp:153-235).

        Returns ``(rows, with_proxy)``: the matrix is
        ``[K(rows, B); K(B, rows)^T]`` plus the proxy block when ``with_proxy``.
        """
        if node.level in (1, 2):
            rows = []
            for other in tree.levels[node.level]:
                if other is not node:
                    rows.extend(other.active_box)
            if node.level == 2:
                for other in tree.levels[1]:
                    if other.is_leaf:
                        rows.extend(other.original_box)
            return rows, False
        cx = node.corners[0] + node.side / 2.0
        cy = node.corners[1] + node.side / 2.0
        lim = RADIUS_RATIO * node.side
        near = np.asarray(node.near, np.int64)
        if near.size == 0:
            return [], True
        p = self.P[near // self.sd]
        dx = cx - p[:, 0]
        dy = cy - p[:, 1]
        keep = np.sqrt(dx * dx + dy * dy) < lim
        return near[keep].tolist(), True

    def id_matrix(self, tree, node):
        """``Kernel::get_id_mat`` (kernel.cpp:153-235)."""
        box = node.active_box
        rows, with_proxy = self.id_rows(tree, node)
        blocks = []
        if rows:
            blocks += [self(rows, box), self(box, rows).T]
        if with_proxy:
            cx = node.corners[0] + node.side / 2.0
            cy = node.corners[1] + node.side / 2.0
            blocks.append(self.proxy_block(cx, cy, node.side * RADIUS_RATIO, box))
        if not blocks:
            return np.zeros((0, len(box)))
        return np.vstack(blocks)


def proxy_ring(cx, cy, r):
    """Proxy points c + r (cos a_i, sin a_i), a_i = 2 pi i / 128, with libm cos/sin (kernel.cpp:242-252)."""
    ang = [2 * math.pi * i * (1.0 / NUM_PROXY_POINTS) for i in range(NUM_PROXY_POINTS)]
    c = np.array([math.cos(a) for a in ang])
    s = np.array([math.sin(a) for a in ang])
    return cx + r * c, cy + r * s, c, s


def laplace_entries(pde, r1, r2, sw, sc, sn1, sn2, tn1, tn2, forward):
    """Entry formula of ``Kernel::laplace`` (kernel.cpp:11-36), broadcast over arrays."""
    with np.errstate(divide="ignore", invalid="ignore"):
        rr = r1 * r1 + r2 * r2
        diag = (r1 == 0.0) & (r2 == 0.0)
        if pde == LAPLACE:
            off = -sw * INV_2PI * (r1 * sn1 + r2 * sn2) / rr
            dval = 0.5 + 0.5 * sc * sw * INV_2PI
        else:
            if forward:
                return sw * INV_2PI * np.log(np.sqrt(rr))
            off = sw + sw * INV_2PI * (r1 * tn1 + r2 * tn2) / rr
            dval = sw - 0.5 + 0.5 * sc * sw * INV_2PI
    if diag.any():
        off = np.where(diag, np.broadcast_to(dval, off.shape), off)
    return off
