"""Results of comparing many process pairs (src/equivalence.rs)."""
from __future__ import annotations

from typing import Sequence

from .energy import Energy, EnergyArray


class StartInfo:
    """Which pairs the game was seeded with (src/equivalence.rs:17-38).  Given
    the pairs as an array ([k, 2] of (p, q)), the AttackPosition list is made
    when first asked for."""

    def __init__(self, starting_points, n_processes: int, pairs=None):
        import numpy as np
        self._sp = None if starting_points is None else list(starting_points)
        self._pairs = None if pairs is None else np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        if self._sp is None and self._pairs is None:
            raise ValueError("starting points or pairs")
        self.n_processes = int(n_processes)

    @property
    def starting_points(self) -> list:
        if self._sp is None:
            from .gamebuild import AttackPosition
            self._sp = [AttackPosition(p, (q,)) for p, q in self._pairs.tolist()]
        return self._sp

    def pairs(self):
        """The starting points' (p, q) as an int64 array [k, 2]."""
        import numpy as np
        if self._pairs is None:
            self._pairs = np.array([(pt.p, pt.q[0]) for pt in self._sp], dtype=np.int64).reshape(-1, 2)
        return self._pairs

    def __len__(self) -> int:
        return len(self._sp) if self._sp is not None else len(self._pairs)

    def equivalence(self, energies: Sequence[EnergyArray]) -> "Equivalence":
        return Equivalence(self, energies)


class Equivalence:
    """Preorders / equivalences between processes (src/equivalence.rs:41-175)."""

    def __init__(self, start_info: StartInfo, energies: Sequence[EnergyArray]):
        if len(energies) < len(start_info.starting_points):
            raise ValueError("Not enough energies")
        self.start_info = start_info
        self.energies = list(energies)[:len(start_info.starting_points)]
        self.pos_to_idx = {(sp.p, sp.q[0]): i for i, sp in enumerate(start_info.starting_points)}
        self.minimization = None

    def set_minimization(self, minimization: Sequence[int]) -> None:
        self.minimization = list(minimization)

    def _mapped(self, p: int) -> int:
        return self.minimization[p] if self.minimization is not None else p

    def _energies_minimized(self, p: int, q: int):
        if not (0 <= p < self.start_info.n_processes and 0 <= q < self.start_info.n_processes):
            raise IndexError("process out of range")
        idx = self.pos_to_idx.get((p, q))
        return None if idx is None else self.energies[idx]

    def energies_of(self, p: int, q: int):
        """Budgets of the pair (p, q), or None if it was not a starting point (:96-98)."""
        return self._energies_minimized(self._mapped(p), self._mapped(q))

    def preorder(self, p: int, q: int, equivalence: Energy) -> bool:
        """p <= q for the preorder of `equivalence` (:125-132)."""
        p, q = self._mapped(p), self._mapped(q)
        if p == q:
            return True
        e = self._energies_minimized(p, q)
        return e is not None and e.test_equivalence(equivalence)

    def equiv(self, p: int, q: int, equivalence: Energy) -> bool:
        return self.preorder(p, q, equivalence) and self.preorder(q, p, equivalence)

    def preorder_matrix(self, equivalences: Sequence[Energy]):
        """preorder(p, q, e) for every ordered pair of processes and every e
        in `equivalences`, as a bool array [n, n, len(equivalences)]:
        vectorised over the packed budgets of all seeded pairs (a pair's
        preorder holds unless one of its budgets lies below e, :125-132)."""
        import numpy as np
        k = len(equivalences)
        conf = equivalences[0].conf if k else None
        eqs = np.array([e.to_vec() for e in equivalences], dtype=np.int64).reshape(k, -1)
        sp = self.start_info.starting_points
        below_any = np.zeros((len(sp), k), dtype=bool)
        if sp and k:
            for e in self.energies:
                if e.conf != conf:
                    raise ValueError("Incompatible Energy configurations")
            lens = np.array([e.n_energies() for e in self.energies], dtype=np.int64)
            if lens.sum():
                data = np.concatenate([e.array for e in self.energies]).astype(np.int64)   # [r, words]
                # the fields' word and shift, as Energy's packing lays them out
                bits, mask = conf.energy_bits(), conf.energy_mask()
                vals = np.empty((data.shape[0], conf.elements), dtype=np.int64)
                shift = word = 0
                for f in range(conf.elements):
                    vals[:, f] = (data[:, word] >> shift) & mask
                    shift += bits
                    if shift >= 32:
                        shift, word = 0, word + 1
                below = np.all(vals[:, None, :] <= eqs[None, :, :], axis=2).astype(np.uint8)   # [r, k]
                starts = np.concatenate(([0], np.cumsum(lens)[:-1]))
                has = lens > 0
                below_any[has] = np.maximum.reduceat(below, starts[has], axis=0).astype(bool)
        return preorder_matrix_of(self.start_info, self.minimization, ~below_any)

    def relation(self, equivalence: Energy) -> "EquivalenceRelation":
        """Classes of processes equivalent under `equivalence` (:154-170)."""
        uf = EquivalenceRelation(self.start_info.n_processes)
        for sp, e in zip(self.start_info.starting_points, self.energies):
            p, q = sp.p, sp.q[0]
            e2 = self._energies_minimized(q, p)
            if e2 is None:
                raise ValueError("Symmetric position not included in starting points for energy game")
            if e.test_equivalence(equivalence) and e2.test_equivalence(equivalence):
                uf.union(p, q)
        if self.minimization is not None:
            return uf.with_mapping(self.minimization)
        return uf


class EquivalenceRelation:
    """Union-find over processes (src/equivalence.rs:178-247)."""

    def __init__(self, n: int):
        self.parent = list(range(n))

    def find(self, x: int) -> int:
        while self.parent[x] != x:
            self.parent[x] = self.parent[self.parent[x]]
            x = self.parent[x]
        return x

    def union(self, a: int, b: int) -> None:
        ra, rb = self.find(a), self.find(b)
        if ra != rb:
            self.parent[max(ra, rb)] = min(ra, rb)

    def __len__(self) -> int:
        return len(self.parent)

    def get_classes(self) -> list:
        idx, classes = {}, []
        for i in range(len(self)):
            c = self.find(i)
            if c not in idx:
                idx[c] = len(classes)
                classes.append([])
            classes[idx[c]].append(i)
        return classes

    def count_classes(self) -> int:
        return len({self.find(i) for i in range(len(self))})

    def class_of(self, proc: int):
        c = self.find(proc)
        return (i for i in range(len(self)) if self.find(i) == c)

    def with_mapping(self, mapping: Sequence[int]) -> "EquivalenceRelation":
        out = EquivalenceRelation(len(mapping))
        rep = {}
        for proc, m in enumerate(mapping):
            part = self.find(m)
            if part in rep:
                out.union(proc, rep[part])
            else:
                rep[part] = proc
        return out


def preorder_matrix_of(start_info: StartInfo, minimization, holds):
    """preorder(p, q, e) for every ordered pair of processes and every
    equivalence e, as Equivalence.preorder_matrix returns it, from the
    starting points' verdicts: holds[i, k] is True when no budget of starting
    point i lies below equivalence k (src/equivalence.rs:125-132;
    GameBuild.preorder_bits computes them on the device).  A pair that is not
    a starting point is not in the preorder unless p and q map to the same
    process."""
    import numpy as np
    holds = np.asarray(holds, dtype=bool)
    pr = start_info.pairs()
    n_sp = pr.shape[0]
    n = start_info.n_processes
    k = holds.shape[1] if holds.ndim == 2 else 0
    if holds.shape[0] != n_sp:
        raise ValueError("one row of verdicts per starting point")
    part = np.asarray(minimization if minimization is not None else np.arange(n), dtype=np.int64)
    m = int(part.max()) + 1 if minimization is not None and part.size else n
    if k == 0:
        return np.zeros((part.size, part.size, 0), dtype=bool)
    # the verdicts of a pair as k bits in nb bytes: the two gathers over the
    # minimisation move nb bytes per pair instead of k
    nb = (k + 7) // 8
    red = np.zeros((max(m, 1), max(m, 1), nb), dtype=np.uint8)
    if n_sp:
        P, Q = pr[:, 0], pr[:, 1]
        # a repeated pair takes its last entry, as the per-pair loop would
        _, first_rev = np.unique((P * max(m, 1) + Q)[::-1], return_index=True)
        last = n_sp - 1 - first_rev
        red[P[last], Q[last]] = np.packbits(holds[last], axis=1, bitorder="little")
    idx = np.arange(m)
    red[idx, idx] = np.packbits(np.ones(k, dtype=bool), bitorder="little")
    full = red.take(part, axis=0).take(part, axis=1)
    bits = np.unpackbits(full.reshape(-1), bitorder="little").view(bool)
    return bits.reshape(part.size, part.size, nb * 8)[:, :, :k]
