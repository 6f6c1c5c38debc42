"""Reference-format mesh writer used to build test fixtures (test side only).

Writes ``Mesh/D2node.dat`` and ``Mesh/BCconnectors.dat`` in the layout the
reference reads (serial/FileIO.c:14-174) and the shipped cases use
(cases/coarse/Mesh/*): node lines ``i    j    x    y    fluid`` in i-fastest
order, connector lines ``i    j    k  type   x_int   y_int   boundary_id`` for
every fluid->solid link, numbers in C ``%g``.  Lattice directions follow the
reference (serial/CellFunctions.c:281-289): 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW,
7 SW, 8 SE.

The cavity convention is the shipped one: Delta = 1/N, x = (i - 1/2) Delta,
fluid iff 1 <= i, j <= N, n = N + 2, m = N + 3; the lid row is type 2 /
boundary id 2, west 1, east 3, bottom 4, and the top-right corner's NE link
belongs to the east wall (cases/coarse/Mesh/BCconnectors.dat:1484).  Links of
one node are written in increasing k, which gives every cell the same
BoundaryID as the shipped files' ordering does.
"""
from __future__ import annotations

import math
import os

import numpy as np

CX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
CY = [0, 0, 1, 0, -1, 1, 1, -1, -1]


def _g(v: float) -> str:
    return "%g" % v


class Mesh:
    def __init__(self, n: int, m: int, delta: float):
        self.n, self.m, self.delta = n, m, delta
        self.fluid = np.zeros((m, n), dtype=np.int32)
        self.x = np.array([[(i - 0.5) * delta for i in range(n)] for _ in range(m)])
        self.y = np.array([[(j - 0.5) * delta for _ in range(n)] for j in range(m)])
        self.links: list[tuple] = []  # (i, j, k, type, xi, yi, bid)

    def node_lines(self):
        for j in range(self.m):
            for i in range(self.n):
                yield "%d    %d    %s    %s    %d" % (
                    i, j, _g(self.x[j, i]), _g(self.y[j, i]), self.fluid[j, i])

    def conn_lines(self):
        for (i, j, k, t, xi, yi, b) in self.links:
            yield "%d    %d    %d  %d   %s   %s   %d" % (i, j, k, t, _g(xi), _g(yi), b)

    def write(self, case_dir: str) -> None:
        os.makedirs(os.path.join(case_dir, "Mesh"), exist_ok=True)
        with open(os.path.join(case_dir, "Mesh", "D2node.dat"), "w") as fh:
            fh.write("\n".join(self.node_lines()) + "\n")
        with open(os.path.join(case_dir, "Mesh", "BCconnectors.dat"), "w") as fh:
            fh.write("\n".join(self.conn_lines()) + "\n")

    def arrays(self):
        """(nodes float32 [N,5], conn float32 [Nc,7]) as the reference parses them."""
        nodes = np.array([[float(t) for t in ln.split()] for ln in self.node_lines()],
                         dtype=np.float32)
        conn = np.array([[float(t) for t in ln.split()] for ln in self.conn_lines()],
                        dtype=np.float32).reshape(-1, 7)
        return nodes, conn


def _box_links(mesh: Mesh, classify, extra_solid_q=None):
    """Links from every fluid node to solid neighbours, increasing k."""
    n, m, d = mesh.n, mesh.m, mesh.delta
    for j in range(m):
        for i in range(n):
            if not mesh.fluid[j, i]:
                continue
            for k in range(1, 9):
                ii, jj = i + CX[k], j + CY[k]
                if 0 <= ii < n and 0 <= jj < m and mesh.fluid[jj, ii]:
                    continue
                t, b, q = classify(i, j, k, ii, jj)
                if q is None:
                    q = 0.5
                xi = mesh.x[j, i] + q * CX[k] * d
                yi = mesh.y[j, i] + q * CY[k] * d
                mesh.links.append((i, j, k, t, xi, yi, b))


def cavity(N: int, lid_type: int = 2) -> Mesh:
    """Lid-driven cavity in the shipped-case convention (n=N+2, m=N+3)."""
    mesh = Mesh(N + 2, N + 3, 1.0 / N)
    mesh.fluid[1:N + 1, 1:N + 1] = 1

    def classify(i, j, k, ii, jj):
        if jj == 0:
            return 1, 4, None
        if ii == N + 1 and jj == N + 1:
            return 1, 3, None
        if jj == N + 1:
            return lid_type, 2, None
        if ii == 0:
            return 1, 1, None
        return 1, 3, None

    _box_links(mesh, classify)
    return mesh


def cavity_obstacle(N: int, cx: float, cy: float, r: float) -> Mesh:
    """Cavity with a circular solid obstacle (node units); links into the
    circle are walls (type 1, boundary id 5) with the exact cut distance q."""
    mesh = Mesh(N + 2, N + 3, 1.0 / N)
    mesh.fluid[1:N + 1, 1:N + 1] = 1
    for j in range(N + 3):
        for i in range(N + 2):
            if (i - cx) ** 2 + (j - cy) ** 2 <= r * r:
                mesh.fluid[j, i] = 0

    def classify(i, j, k, ii, jj):
        if 1 <= ii <= N and 1 <= jj <= N:  # the obstacle
            # smallest t in (0,1] with |(i,j) + t c_k - centre| = r
            dx, dy = i - cx, j - cy
            a = CX[k] ** 2 + CY[k] ** 2
            b = 2 * (dx * CX[k] + dy * CY[k])
            c = dx * dx + dy * dy - r * r
            t = (-b - math.sqrt(max(b * b - 4 * a * c, 0.0))) / (2 * a)
            return 1, 5, min(max(t, 1e-3), 1.0)
        if jj == 0:
            return 1, 4, None
        if ii == N + 1 and jj == N + 1:
            return 1, 3, None
        if jj == N + 1:
            return 2, 2, None
        if ii == 0:
            return 1, 1, None
        return 1, 3, None

    _box_links(mesh, classify)
    return mesh


def channel(NX: int, NY: int) -> Mesh:
    """Channel: west inlet (type 2, id 1), east outlet (type 3, id 3),
    bottom/top walls (type 1, ids 4/2); n=NX+2, m=NY+2, Delta = 1/NY."""
    mesh = Mesh(NX + 2, NY + 2, 1.0 / NY)
    mesh.fluid[1:NY + 1, 1:NX + 1] = 1

    def classify(i, j, k, ii, jj):
        if jj == 0:
            return 1, 4, None
        if jj == NY + 1:
            return 1, 2, None
        if ii == 0:
            return 2, 1, None
        return 3, 3, None

    _box_links(mesh, classify)
    return mesh


def tall_cavity(NXf: int, NYf: int) -> Mesh:
    """Cavity narrower than it is tall (m > 2n), so the reference residual's
    first-m-cells sweep reaches fluid cells."""
    mesh = Mesh(NXf + 2, NYf + 3, 1.0 / NYf)
    mesh.fluid[1:NYf + 1, 1:NXf + 1] = 1

    def classify(i, j, k, ii, jj):
        if jj == 0:
            return 1, 4, None
        if ii == NXf + 1 and jj == NYf + 1:
            return 1, 3, None
        if jj == NYf + 1:
            return 2, 2, None
        if ii == 0:
            return 1, 1, None
        return 1, 3, None

    _box_links(mesh, classify)
    return mesh


def cavity_arrays_fast(N: int, lid_type: int = 2):
    """Vectorised cavity (same links, classes and order as ``cavity``), with
    coordinates at full float precision as the native generator
    lbm_case_cavity stores them in memory (no %g rounding).  For the large
    lattices the oracle is timed on."""
    n, m, d = N + 2, N + 3, 1.0 / N
    jj, ii = np.mgrid[0:m, 0:n]
    fl = ((ii >= 1) & (ii <= N) & (jj >= 1) & (jj <= N)).astype(np.float32)
    nodes = np.stack([ii.ravel().astype(np.float32), jj.ravel().astype(np.float32),
                      ((ii.ravel() - 0.5) * d).astype(np.float32),
                      ((jj.ravel() - 0.5) * d).astype(np.float32), fl.ravel()], axis=1)
    # boundary nodes in row-major order, each with its links in increasing k
    rows = []
    fluid = fl.astype(bool)
    bj, bi = np.nonzero(fluid & ((ii == 1) | (ii == N) | (jj == 1) | (jj == N)))
    for k in range(1, 9):
        ti, tj = bi + CX[k], bj + CY[k]
        solid = ~fluid[tj, ti]
        i, j, ti, tj = bi[solid], bj[solid], ti[solid], tj[solid]
        t = np.where(tj == 0, 1, np.where((ti == N + 1) & (tj == N + 1), 1,
                                          np.where(tj == N + 1, lid_type, 1)))
        b = np.where(tj == 0, 4, np.where((ti == N + 1) & (tj == N + 1), 3,
                                          np.where(tj == N + 1, 2, np.where(ti == 0, 1, 3))))
        xi = (i - 0.5) * d + 0.5 * CX[k] * d
        yi = (j - 0.5) * d + 0.5 * CY[k] * d
        rows.append(np.stack([i, j, np.full_like(i, k), t, xi, yi, b, j * n + i], axis=1))
    allr = np.concatenate(rows)
    order = np.lexsort((allr[:, 2], allr[:, 7]))
    conn = allr[order][:, :7].astype(np.float32)
    return nodes.astype(np.float32), conn
