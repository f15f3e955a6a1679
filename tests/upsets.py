"""Independent reference for the fixpoint: Kleene iteration over upward-closed
subsets of {0..3}^6 as 4096-entry boolean arrays (numpy).

It shares nothing with the oracle's list algorithms (pairwise suprema, O(n^2)
Pareto filter, the round loop) nor with the GPU's bitmap kernels: membership is
computed point by point over all 4096 points.  Point p has coordinates
e_k = (p >> 2k) & 3, which is also the packed STANDARD energy.
"""
from __future__ import annotations

import numpy as np

N = 4096
_P = np.arange(N, dtype=np.int64)
_COORD = np.stack([(_P >> (2 * k)) & 3 for k in range(6)])      # (6, 4096)
# leq[a, b] = a <= b componentwise (4096 x 4096 bool, 16 MiB)
_LEQ = None


def leq_matrix() -> np.ndarray:
    global _LEQ
    if _LEQ is None:
        m = np.ones((N, N), dtype=bool)
        for k in range(6):
            m &= _COORD[k][:, None] <= _COORD[k][None, :]
        _LEQ = m
    return _LEQ


def inv_table(w: int, maxv: int = 3, variant: str = "ok") -> np.ndarray:
    """inv_update of every point for update word w (shadergen.rs:103-128).
    Other variants plant plausible slips: "min_first" (Min fields before the
    decrements), "transposed" (e[i] = max(e[i], e[j])), "no_min" (Min ignored)."""
    e = _COORD.copy()
    u = [(w >> (4 * k)) & 15 for k in range(6)]

    def dec():
        for k in range(6):
            e[k] = np.minimum(e[k] + (1 if u[k] == 1 else 0), maxv)

    def mins():
        for i in range(6):
            if u[i] > 1 and variant != "no_min":
                j = u[i] - 2
                if variant == "transposed":
                    e[i] = np.maximum(e[i], e[j])
                else:
                    e[j] = np.maximum(e[j], e[i])

    if variant == "min_first":
        mins(); dec()
    else:
        dec(); mins()
    return sum(e[k] << (2 * k) for k in range(6)).astype(np.int64)


def up(mask: np.ndarray) -> np.ndarray:
    """Upward closure by brute force: x in up(S) iff some s in S has s <= x."""
    if not mask.any():
        return np.zeros(N, dtype=bool)
    return leq_matrix()[mask].any(axis=0)


def minimal(mask: np.ndarray) -> np.ndarray:
    """Minimal elements of a set by brute force: no other member strictly below."""
    idx = np.nonzero(mask)[0]
    if idx.size == 0:
        return mask.copy()
    sub = leq_matrix()[np.ix_(idx, idx)]
    strictly_below = sub & ~np.eye(idx.size, dtype=bool)
    keep = ~strictly_below.any(axis=0)
    out = np.zeros(N, dtype=bool)
    out[idx[keep]] = True
    return out


def image(mask: np.ndarray, table: np.ndarray) -> np.ndarray:
    out = np.zeros(N, dtype=bool)
    out[table[mask]] = True
    return out


def kleene(row, col, wt, atk, reach, maxv: int = 3, variant: str = "ok", max_iter: int = 10000):
    """Least fixpoint of the energy game operator from the with_reach start
    (reach nodes = up({0}) = everything).  Returns the (V, 4096) up-sets."""
    V = len(row) - 1
    W = np.zeros((V, N), dtype=bool)
    for r in reach:
        W[int(r)] = True
    tables = {}

    def tab(w):
        if w not in tables:
            tables[w] = inv_table(int(w), maxv, variant)
        return tables[w]

    for _ in range(max_iter):
        changed = False
        for u in range(V):
            a, b = int(row[u]), int(row[u + 1])
            if a == b:
                continue
            if atk[u]:
                new = W[u].copy()
                for k in range(a, b):
                    new |= up(image(W[int(col[k])], tab(int(wt[k]))))
            else:
                if any(not W[int(col[k])].any() for k in range(a, b)):
                    continue
                new = np.ones(N, dtype=bool)
                for k in range(a, b):
                    new &= up(image(W[int(col[k])], tab(int(wt[k]))))
            if not np.array_equal(new, W[u]):
                W[u] = new
                changed = True
        if not changed:
            return W
    raise RuntimeError("no fixpoint")


def as_points(mask: np.ndarray) -> tuple:
    return tuple(int(x) for x in np.nonzero(mask)[0])


# ---------------------------------------------------------------------------
# Any small EnergyConf: the points of {0..max}^elements enumerated explicitly.
# ---------------------------------------------------------------------------
class Grid:
    """All points of a conf with at most a few thousand points; point i has
    coordinates coord[:, i] (mixed radix max + 1, component 0 fastest)."""

    def __init__(self, elements: int, maxv: int):
        self.elements, self.maxv = elements, maxv
        r = maxv + 1
        self.n = r ** elements
        idx = np.arange(self.n, dtype=np.int64)
        self.coord = np.stack([(idx // r ** k) % r for k in range(elements)])       # (elements, n)
        self.radix = r
        m = np.ones((self.n, self.n), dtype=bool)
        for k in range(elements):
            m &= self.coord[k][:, None] <= self.coord[k][None, :]
        self.leq = m

    def index(self, coords) -> int:
        return int(sum(int(c) * self.radix ** k for k, c in enumerate(coords)))

    def image_table(self, fields) -> np.ndarray:
        """inv_update of every point for an update given as field values
        (0 Zero, 1 Decrement, v >= 2 Min(v - 1)): decrements first, then the
        Min fields in index order."""
        e = self.coord.copy()
        for k, u in enumerate(fields):
            if u == 1:
                e[k] = np.minimum(e[k] + 1, self.maxv)
        for i, u in enumerate(fields):
            if u >= 2:
                j = u - 2
                e[j] = np.maximum(e[j], e[i])
        return sum(e[k] * self.radix ** k for k in range(self.elements)).astype(np.int64)

    def up(self, mask):
        return self.leq[mask].any(axis=0) if mask.any() else np.zeros(self.n, dtype=bool)

    def minimal(self, mask):
        idx = np.nonzero(mask)[0]
        out = np.zeros(self.n, dtype=bool)
        if idx.size:
            sub = self.leq[np.ix_(idx, idx)] & ~np.eye(idx.size, dtype=bool)
            out[idx[~sub.any(axis=0)]] = True
        return out


def kleene_grid(grid: Grid, row, col, upd_fields, atk, reach, max_iter: int = 10000):
    """Least fixpoint on up-sets of the grid's points (reach nodes start with
    everything).  upd_fields: per edge, the list of its field values."""
    V = len(row) - 1
    W = np.zeros((V, grid.n), dtype=bool)
    for r in reach:
        W[int(r)] = True
    tabs = [grid.image_table(f) for f in upd_fields]

    def img(mask, t):
        out = np.zeros(grid.n, dtype=bool)
        out[t[mask]] = True
        return out

    for _ in range(max_iter):
        changed = False
        for u in range(V):
            a, b = int(row[u]), int(row[u + 1])
            if a == b:
                continue
            if atk[u]:
                new = W[u].copy()
                for k in range(a, b):
                    new |= grid.up(img(W[int(col[k])], tabs[k]))
            else:
                if any(not W[int(col[k])].any() for k in range(a, b)):
                    continue
                new = np.ones(grid.n, dtype=bool)
                for k in range(a, b):
                    new &= grid.up(img(W[int(col[k])], tabs[k]))
            if not np.array_equal(new, W[u]):
                W[u] = new
                changed = True
        if not changed:
            return W
    raise RuntimeError("no fixpoint")


def one_dimension_values(row, col, upd_fields, atk, reach, max_iter: int = 100000):
    """The one-component game in closed form: a node's minimal budgets are one
    value or none (a chain has one minimum).  Attacker: the least over its
    successors' values after the update (+1 for a Decrement, unchanged
    otherwise; Min(1) on the only component is the identity); defender with
    every successor valued: the greatest of them.  Least fixpoint by value
    iteration from the nodes to reach at 0.  Returns values (None: no budget)."""
    V = len(row) - 1
    val = [None] * V
    for r in reach:
        val[int(r)] = 0
    for _ in range(max_iter):
        changed = False
        for u in range(V):
            a, b = int(row[u]), int(row[u + 1])
            if a == b:
                continue
            imgs = [None if val[int(col[k])] is None else val[int(col[k])] + (1 if upd_fields[k][0] == 1 else 0)
                    for k in range(a, b)]
            if atk[u]:
                c = [x for x in imgs if x is not None] + ([] if val[u] is None else [val[u]])
                new = min(c) if c else None
            else:
                new = None if any(x is None for x in imgs) else max(imgs)
                if new is not None and val[u] is not None:
                    new = min(new, val[u])
            if new != val[u]:
                val[u] = new
                changed = True
        if not changed:
            return val
    raise RuntimeError("no fixpoint")
