"""Energy data model: EnergyConf, Energy, EnergyArray, Upd, Update, UpdateArray.

Host-side mirror of the reference crate's ``src/energy.rs`` (Gobbel2000/gpuequiv
v1.0.0).  Packing, partial order and the equivalence test follow the cited
lines exactly; the kernels consume the packed words produced here.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable, Sequence

import numpy as np


def _next_pow2(x: int) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


@dataclass(frozen=True)
class EnergyConf:
    """Number of energy components and their saturation value (src/energy.rs:36-78)."""

    elements: int
    max: int

    def energy_bits(self) -> int:
        return _next_pow2(int(self.max).bit_length())

    def energy_size(self) -> int:
        return -(-(self.energy_bits() * self.elements) // 32)

    def energy_mask(self) -> int:
        return 0xFFFFFFFF >> (32 - self.energy_bits())

    def update_bits(self) -> int:
        return _next_pow2(int(self.elements + 2).bit_length())

    def update_size(self) -> int:
        return -(-(self.update_bits() * self.elements) // 32)

    def update_mask(self) -> int:
        return 0xFFFFFFFF >> (32 - self.update_bits())


EnergyConf.STANDARD = EnergyConf(6, 3)            # src/energy.rs:44
EnergyConf.SILENT_STEP = EnergyConf(8, 2)         # :45
EnergyConf.ONE_DIMENSION = EnergyConf(1, 0xFFFFFFFF)  # :46


def _pack(values: Sequence[int], bits: int, nwords: int) -> list:
    """Pack fields, moving to the next word when a field would not fit
    (src/energy.rs:143-162, :543-565)."""
    out = [0] * nwords
    shift = word = 0
    for v in values:
        out[word] |= (int(v) << shift) & 0xFFFFFFFF
        shift += bits
        if shift + bits > 32:
            shift = 0
            word += 1
    return out


def _unpack(words: Sequence[int], n: int, bits: int, mask: int) -> list:
    out = []
    shift = word = 0
    for _ in range(n):
        out.append((int(words[word]) >> shift) & mask)
        shift += bits
        if shift >= 32:
            shift = 0
            word += 1
    return out


class Energy:
    """One packed energy tuple (src/energy.rs:81-141)."""

    __slots__ = ("data", "conf")

    def __init__(self, data: Sequence[int], conf: EnergyConf):
        self.data = tuple(int(x) for x in data)
        self.conf = conf

    @classmethod
    def new(cls, values: Sequence[int], conf: EnergyConf) -> "Energy | None":
        if len(values) > conf.elements or any(int(v) > conf.max for v in values):
            return None
        return cls(_pack(values, conf.energy_bits(), conf.energy_size()), conf)

    @classmethod
    def zero(cls, conf: EnergyConf) -> "Energy":
        return cls([0] * conf.energy_size(), conf)

    def to_vec(self) -> list:
        c = self.conf
        return _unpack(self.data, c.elements, c.energy_bits(), c.energy_mask())

    def raw_data(self) -> tuple:
        return self.data

    def partial_cmp(self, other: "Energy"):
        """-1 / 0 / 1 / None as Ordering::{Less, Equal, Greater} / incomparable (:121-141)."""
        if self.conf.elements != other.conf.elements:
            raise ValueError("Incompatible Energy configurations")
        less = greater = True
        for a, b in zip(self.to_vec(), other.to_vec()):
            if a < b:
                greater = False
            elif a > b:
                less = False
        if less and greater:
            return 0
        if less:
            return -1
        if greater:
            return 1
        return None

    def __le__(self, other: "Energy") -> bool:
        return self.partial_cmp(other) in (-1, 0)

    def __lt__(self, other: "Energy") -> bool:
        return self.partial_cmp(other) == -1

    def __eq__(self, other) -> bool:
        return isinstance(other, Energy) and self.conf == other.conf and self.data == other.data

    def __hash__(self) -> int:
        return hash((self.data, self.conf))

    def __repr__(self) -> str:
        return "Energy(" + " ".join("∞" if v == self.conf.max else str(v) for v in self.to_vec()) + ")"


class Upd:
    """One element of an update tuple (src/energy.rs:444-500): Zero, Decrement, Min(k)."""

    __slots__ = ("kind", "idx")

    def __init__(self, kind: str, idx: int = 0):
        self.kind = kind
        self.idx = idx

    @staticmethod
    def from_int(n: int) -> "Upd":
        if n == 0:
            return Upd.Zero
        if n == -1:
            return Upd.Decrement
        if n <= -2:
            raise ValueError("Invalid update entry")
        return Upd.Min(n)

    def to_int(self) -> int:
        return {"zero": 0, "dec": -1}.get(self.kind, self.idx)

    def packed_repr(self) -> int:
        return {"zero": 0, "dec": 1}.get(self.kind, self.idx + 1)

    def __eq__(self, other) -> bool:
        return isinstance(other, Upd) and self.kind == other.kind and self.idx == other.idx

    def __repr__(self) -> str:
        return {"zero": "0", "dec": "-1"}.get(self.kind, f"Min({self.idx})")


Upd.Zero = Upd("zero")
Upd.Decrement = Upd("dec")
Upd.Min = staticmethod(lambda k: Upd("min", int(k)))  # type: ignore[assignment]


def update(*xs) -> list:
    """The crate's ``update!`` macro (src/energy.rs:615-620): ints or Upd values."""
    return [x if isinstance(x, Upd) else Upd.from_int(int(x)) for x in xs]


class Update:
    """Packed update tuple (src/energy.rs:503-613)."""

    __slots__ = ("data", "conf")

    def __init__(self, data: Sequence[int], conf: EnergyConf):
        self.data = tuple(int(x) for x in data)
        self.conf = conf

    @classmethod
    def new(cls, values: Sequence, conf: EnergyConf) -> "Update | None":
        vals = [v if isinstance(v, Upd) else Upd.from_int(int(v)) for v in values]
        if len(vals) > conf.elements:
            return None
        for v in vals:
            if v.kind == "min" and v.idx > conf.elements:
                return None
        return cls(_pack([v.packed_repr() for v in vals], conf.update_bits(), conf.update_size()), conf)

    @classmethod
    def zero(cls, conf: EnergyConf) -> "Update":
        return cls([0] * conf.update_size(), conf)

    def to_vec(self) -> list:
        c = self.conf
        out = []
        for v in _unpack(self.data, c.elements, c.update_bits(), c.update_mask()):
            out.append(Upd.Zero if v == 0 else Upd.Decrement if v == 1 else Upd.Min(v - 1))
        return out

    def to_ints(self) -> list:
        return [u.to_int() for u in self.to_vec()]

    def raw_data(self) -> tuple:
        return self.data

    def __eq__(self, other) -> bool:
        return isinstance(other, Update) and self.conf == other.conf and self.data == other.data

    def __repr__(self) -> str:
        return "<" + " ".join(repr(u) for u in self.to_vec()) + ">"


class EnergyArray:
    """A list of energies, one row of packed words each (src/energy.rs:227-332).

    Equality is order-insensitive like the crate's PartialEq (:334-369): same
    number of rows and every row of ``other`` is a row of ``self``.
    """

    __slots__ = ("array", "conf")

    def __init__(self, array: np.ndarray, conf: EnergyConf):
        array = np.asarray(array, dtype=np.uint32)
        if array.ndim != 2 or array.shape[1] != conf.energy_size():
            raise ValueError("array width does not match the energy configuration")
        self.array = array
        self.conf = conf

    @classmethod
    def empty(cls, conf: EnergyConf) -> "EnergyArray":
        return cls.zero(0, conf)

    @classmethod
    def zero(cls, n: int, conf: EnergyConf) -> "EnergyArray":
        return cls(np.zeros((n, conf.energy_size()), dtype=np.uint32), conf)

    @classmethod
    def from_conf(cls, rows: Iterable, conf: EnergyConf) -> "EnergyArray":
        """From component lists (``FromEnergyConf<&[Vec<u32>]>``, :431-441) or Energy values."""
        packed = []
        for r in rows:
            if isinstance(r, Energy):
                if r.conf != conf:
                    raise ValueError("EnergyConf in list doesn't match provided EnergyConf")
                packed.append(list(r.data))
            else:
                e = Energy.new(list(r), conf)
                if e is None:
                    raise ValueError("Energies don't match configuration")
                packed.append(list(e.data))
        arr = np.array(packed, dtype=np.uint32).reshape(len(packed), conf.energy_size())
        return cls(arr, conf)

    def n_energies(self) -> int:
        return self.array.shape[0]

    def is_empty(self) -> bool:
        return self.array.size == 0

    def get(self, idx: int) -> "Energy | None":
        if idx >= self.n_energies():
            return None
        return Energy(self.array[idx], self.conf)

    def __iter__(self):
        for i in range(self.n_energies()):
            yield Energy(self.array[i], self.conf)

    def __len__(self) -> int:
        return self.n_energies()

    def to_lists(self) -> list:
        return [e.to_vec() for e in self]

    def test_equivalence(self, equivalence: Energy) -> bool:
        """True iff no minimal budget lies below the equivalence's coordinates (:307-318)."""
        if self.conf != equivalence.conf:
            raise ValueError("Incompatible Energy configurations")
        return not any(e <= equivalence for e in self)

    def __eq__(self, other) -> bool:
        if isinstance(other, EnergyArray):
            if self.conf != other.conf:
                return False
            other_arr = other.array
        else:
            other_arr = np.asarray(other, dtype=np.uint32)
        if self.array.shape != other_arr.shape:
            return False
        rows = {tuple(r) for r in self.array.tolist()}
        return all(tuple(r) in rows for r in other_arr.tolist())

    def __repr__(self) -> str:
        return "[" + ", ".join(repr(e) for e in self) + "]"


class UpdateArray:
    """Rows of packed updates, one per edge (src/energy.rs:623-711)."""

    __slots__ = ("array", "conf")

    def __init__(self, array: np.ndarray, conf: EnergyConf):
        array = np.asarray(array, dtype=np.uint32)
        if array.ndim != 2 or array.shape[1] != conf.update_size():
            raise ValueError("array width does not match the update configuration")
        self.array = array
        self.conf = conf

    @classmethod
    def from_conf(cls, rows: Iterable, conf: EnergyConf) -> "UpdateArray":
        packed = []
        for r in rows:
            if isinstance(r, Update):
                packed.append(list(r.data))
            else:
                u = Update.new(list(r), conf)
                if u is None:
                    raise ValueError("Updates don't match configuration")
                packed.append(list(u.data))
        arr = np.array(packed, dtype=np.uint32).reshape(len(packed), conf.update_size())
        return cls(arr, conf)

    def n_updates(self) -> int:
        return self.array.shape[0]

    def get(self, idx: int) -> "Update | None":
        if idx >= self.n_updates():
            return None
        return Update(self.array[idx], self.conf)

    def set(self, idx: int, upd: Update) -> None:
        if upd.conf != self.conf or idx >= self.n_updates():
            raise IndexError("update index or configuration mismatch")
        self.array[idx] = upd.data

    def __eq__(self, other) -> bool:
        return isinstance(other, UpdateArray) and self.conf == other.conf and np.array_equal(self.array, other.array)


class std_equivalences:
    """Energy coordinates of the 13 equivalences (src/energy.rs:198-225)."""

    _table = {
        "enabledness": (1, 1),
        "traces": (3, 1),
        "failures": (3, 2, 0, 0, 1, 1),
        "failure_traces": (3, 3, 3, 0, 1, 1),
        "readiness": (3, 2, 1, 1, 1, 1),
        "readiness_traces": (3, 3, 3, 1, 1, 1),
        "revivals": (3, 2, 1, 0, 1, 1),
        "impossible_futures": (3, 2, 0, 0, 3, 1),
        "possible_futures": (3, 2, 3, 3, 3, 1),
        "simulation": (3, 3, 3, 3, 0, 0),
        "ready_simulation": (3, 3, 3, 3, 1, 1),
        "nested_2simulation": (3, 3, 3, 3, 3, 1),
        "bisimulation": (3, 3, 3, 3, 3, 3),
    }
    NAMES = tuple(_table)

    @classmethod
    def get(cls, name: str) -> Energy:
        return Energy.new(cls._table[name], EnergyConf.STANDARD)


for _name in std_equivalences.NAMES:
    setattr(std_equivalences, _name, staticmethod((lambda n: (lambda: std_equivalences.get(n)))(_name)))
