"""Dense restatement of the operator build (TEST INFRASTRUCTURE ONLY).

An independent, loop-for-loop restatement in numpy/Python of the reference's
setup arithmetic, used to check the product's (row-parallel C++) operator
arrays on small meshes:

  grad_coeff        StaticSolver::calcCvGradCoeff, StaticSolver.cpp:5598-5765
                    (the same formula as StripedMesh.hpp:5624-5800, step 4)
  D(i,j), n_ef, c_ef  StripedMesh::buildExtendedFaces, EF_OPS DEFAULT,
                    StripedMesh.hpp:6013-6131 (global loop over cells, faces
                    and stencil, accumulated into a (row, column) map exactly
                    as the reference's coeffVec), transposition 6589-6603,
                    n = D - D^T, c = D + D^T  6740-6861
  area_over_delta   StaticSolver.cpp:4447-4459
  dcmag_hat_cv, sa_to_vol_ratio   FlowSolver.hpp:702-758, 859-878

The SVD step uses numpy's SVD; the reference's own calcSvd33 pins that
substitution (tests/test_oracle_pins.py::test_reference_svd_regularised_inverse).
Periodic images are identified by (owner, shift) the way the reference's
rank-bits-index keys identify ghost images.
"""
from __future__ import annotations

import numpy as np


def regularised_update(A, C, sigma_thresh):
    """g = c + A^+ (c - A c) for every row c of C, with singular values of A
    below sigma_thresh dropped from the pseudo-inverse (StaticSolver.cpp:5704-5756)."""
    U, s, Vt = np.linalg.svd(A)
    sinv = np.array([0.0 if x < sigma_thresh else 1.0 / x for x in s])
    Aplus = Vt.T @ np.diag(sinv) @ U.T
    out = np.empty_like(C)
    for k in range(C.shape[0]):
        c = C[k]
        out[k] = c + Aplus @ (c - A @ c)
    return out


class DenseOperators:
    def __init__(self, x_cv, vol, cvofa, shift, n_fa, x_fa, L, interp_factor=1.0 / 3.0, sigma_thresh=0.45):
        self.x_cv = np.asarray(x_cv, float)
        self.vol = np.asarray(vol, float)
        self.cvofa = np.asarray(cvofa)
        self.shift = np.asarray(shift)
        self.n_fa = np.asarray(n_fa, float)
        self.x_fa = np.asarray(x_fa, float)
        self.L = np.asarray(L, float)
        self.ncv = len(vol)
        self.nfa = len(cvofa)
        self.f = interp_factor
        self.sigma = sigma_thresh
        self._faces_of_cell()
        self._grad()
        self._extended()

    def _faces_of_cell(self):
        faces = [[] for _ in range(self.ncv)]
        for f in range(self.nfa):
            faces[self.cvofa[f, 0]].append(f)
            faces[self.cvofa[f, 1]].append(f)
        self.faces = faces

    def view(self, c, f):
        """(sign, neighbour owner, neighbour shift, x_fa in c's frame, x_nbr image in c's frame)"""
        c0, c1 = self.cvofa[f]
        s = self.shift[f]
        if c0 == c:
            sign, nb, ns, xf = 1.0, c1, s.copy(), self.x_fa[f].copy()
        else:
            sign, nb, ns, xf = -1.0, c0, -s, self.x_fa[f] - s * self.L
        return sign, nb, tuple(int(v) for v in ns), xf, self.x_cv[nb] + np.array(ns) * self.L

    def _grad(self):
        self.stencil = []   # per cell: [(owner, shift)], diagonal first
        self.gcoef = []     # per cell: array [len][3]
        for c in range(self.ncv):
            st = [(c, (0, 0, 0))]
            A = np.eye(3) * self.vol[c]
            coef = [np.zeros(3)]
            for f in self.faces[c]:
                sign, nb, ns, xf, xn = self.view(c, f)
                st.append((nb, ns))
                r = xf - 0.5 * (self.x_cv[c] + xn)
                A -= np.outer(self.n_fa[f] * sign, r)  # A[i][j] = delta_ij V - sum n_i r_j
                cc = 0.5 * self.n_fa[f] * sign
                coef.append(cc.copy())
                coef[0] = coef[0] - cc
            A /= self.vol[c]
            C = np.array(coef) / self.vol[c]
            self.stencil.append(st)
            self.gcoef.append(regularised_update(A, C, self.sigma))

    def _extended(self):
        f = self.f
        D = {}

        def add(row, col, v):
            key = (row, col)
            D[key] = D.get(key, np.zeros(3)) + v

        for i in range(self.ncv):
            for fa in self.faces[i]:
                sign, nb, ns, xf, xn = self.view(i, fa)
                nout = sign * self.n_fa[fa]
                # central part (i, nbr)
                add((i, (0, 0, 0)), (nb, ns), 0.5 * nout)
                dx = 0.5 * xf - (0.5 - 0.25 * f) * self.x_cv[i] - 0.25 * f * xn
                for (k, ks), g in zip(self.stencil[i], self.gcoef[i]):
                    dp = g @ dx
                    if (k, ks) != (i, (0, 0, 0)):
                        add((i, (0, 0, 0)), (k, ks), dp * nout)
                    if (k, ks) != (nb, ns):
                        # row nbr expressed in its own frame: shift columns by -ns
                        kr = tuple(int(a - b) for a, b in zip(ks, ns))
                        add((nb, (0, 0, 0)), (k, kr), -dp * nout)
        # pairs (lo < hi): n = D(lo,hi) - D(hi,lo), c = D(lo,hi) + D(hi,lo)
        pairs = {}
        for ((r, _), (cidx, cs)), v in D.items():
            if r == cidx:
                raise ValueError("cell couples to its own image")
            if r < cidx:
                key = (r, cidx, cs)
                n, c = pairs.get(key, (np.zeros(3), np.zeros(3)))
                pairs[key] = (n + v, c + v)
            else:
                key = (cidx, r, tuple(-a for a in cs))
                n, c = pairs.get(key, (np.zeros(3), np.zeros(3)))
                pairs[key] = (n - v, c + v)
        self.pairs = pairs

    def area_over_delta(self):
        out = np.zeros(self.nfa)
        for fa in range(self.nfa):
            c0, c1 = self.cvofa[fa]
            n = self.n_fa[fa]
            a2 = n @ n
            d = (self.x_cv[c1] + self.shift[fa] * self.L) - self.x_cv[c0]
            dp = d @ n
            out[fa] = a2 / dp if (a2 > 0 and dp > 0) else 0.0
        return out

    def dcmag_and_sa(self):
        dc = np.zeros(self.ncv)
        for (lo, hi, s), (n, c) in self.pairs.items():
            delta = np.linalg.norm(self.x_cv[hi] + np.array(s) * self.L - self.x_cv[lo])
            cm = np.linalg.norm(c)
            dc[lo] += delta * cm
            dc[hi] += delta * cm
        area = np.zeros(self.ncv)
        for fa in range(self.nfa):
            a = np.linalg.norm(self.n_fa[fa])
            area[self.cvofa[fa, 0]] += a
            area[self.cvofa[fa, 1]] += a
        return dc / area, area / self.vol
