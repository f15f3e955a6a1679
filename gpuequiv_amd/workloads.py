"""Seeded synthetic workloads (numpy PCG64; same seed -> same bytes everywhere).

* ``random_game``  - standalone energy game: Bernoulli attacker positions,
  geometric out-degree on {0,1,...} with the given mean, uniform targets, and
  updates drawn as identity / one Decrement / one Min(j) (the update shapes of
  graphgen/graphgen.py:14-27, restated with j != i).
* ``random_lts`` / ``powerlaw_lts`` - labelled transition systems.
* ``game_batch``  - many random games with log-uniform sizes.
"""
from __future__ import annotations

import numpy as np

from .energy import EnergyConf
from .energygame import GameGraph


def random_update_words(rng: np.random.Generator, n: int, p_ident: float = 0.5, p_dec: float = 0.35,
                        elements: int = 6) -> np.ndarray:
    """Packed STANDARD updates (4-bit fields): identity, a Decrement at one
    uniform component, or Min(j) at component i with j uniform over the others
    (1-based, as Upd::Min)."""
    r = rng.random(n)
    comp = rng.integers(0, elements, n)
    other = rng.integers(0, elements - 1, n)
    other = other + (other >= comp)           # uniform over the other components, 0-based
    w = np.zeros(n, dtype=np.uint32)
    dec = (r >= p_ident) & (r < p_ident + p_dec)
    mn = r >= p_ident + p_dec
    w[dec] = (np.uint32(1) << (4 * comp[dec]).astype(np.uint32))
    # Min(k) with k = other + 1 is packed as k + 1 = other + 2
    w[mn] = ((other[mn] + 2).astype(np.uint32) << (4 * comp[mn]).astype(np.uint32))
    return w


def any_update_words(rng: np.random.Generator, n: int) -> np.ndarray:
    """Packed STANDARD updates with every field drawn independently from
    Zero, Decrement and Min(1..6): several Decrements and Mins per update,
    chained Mins and self-Mins included."""
    nib = rng.integers(0, 8, (n, 6)).astype(np.uint32)
    return (nib << (4 * np.arange(6, dtype=np.uint32))).sum(axis=1).astype(np.uint32)


def random_game(n_vertices: int, seed: int, p_attacker: float = 0.5, mean_degree: float = 4.0,
                p_ident: float = 0.5, p_dec: float = 0.35, any_updates: bool = False) -> GameGraph:
    """Standalone game: out-degree Geometric on {0,1,...} with the given mean,
    targets uniform, self-loops and multi-edges allowed (GameGraph::new allows
    both, src/energygame.rs:141-194).  any_updates: every update field random
    (any_update_words) instead of the identity / one Decrement / one Min mix."""
    rng = np.random.default_rng(seed)
    V = int(n_vertices)
    atk = (rng.random(V) < p_attacker).astype(np.uint8)
    p = 1.0 / (mean_degree + 1.0)
    deg = rng.geometric(p, V) - 1
    row = np.zeros(V + 1, dtype=np.uint64)
    np.cumsum(deg, out=row[1:])
    E = int(row[-1])
    col = rng.integers(0, V, E, dtype=np.uint32)
    w = any_update_words(rng, E) if any_updates else random_update_words(rng, E, p_ident, p_dec)
    return GameGraph(EnergyConf.STANDARD, row.astype(np.uint32), col, w.reshape(E, 1), atk)


def pack_fields(values: np.ndarray, bits: int, words: int) -> np.ndarray:
    """Rows of field values (n, fields) packed like the crate (src/energy.rs:143-162):
    32 // bits fields per word, field k of a word at bits [k * bits, (k + 1) * bits)."""
    n, f = values.shape
    per = 32 // bits
    out = np.zeros((n, words), dtype=np.uint64)
    for k in range(f):
        out[:, k // per] |= values[:, k].astype(np.uint64) << np.uint64((k % per) * bits)
    return out.astype(np.uint32)


def random_game_conf(n_vertices: int, seed: int, conf: EnergyConf, p_attacker: float = 0.5, mean_degree: float = 4.0,
                     p_ident: float = 0.5, p_dec: float = 0.35, any_updates: bool = False) -> GameGraph:
    """``random_game`` for any EnergyConf: the same degree and target
    distributions; each update is the identity, a Decrement at one uniform
    component, or Min(j) at component i with j uniform over the others (over
    all components when there is only one), packed in the conf's update
    fields.  any_updates: every field drawn from Zero, Decrement and
    Min(1..elements)."""
    rng = np.random.default_rng(seed)
    V = int(n_vertices)
    atk = (rng.random(V) < p_attacker).astype(np.uint8)
    deg = rng.geometric(1.0 / (mean_degree + 1.0), V) - 1
    row = np.zeros(V + 1, dtype=np.uint64)
    np.cumsum(deg, out=row[1:])
    E = int(row[-1])
    col = rng.integers(0, V, E, dtype=np.uint32)
    el = conf.elements
    f = np.zeros((E, el), dtype=np.int64)               # packed field values: 0 Zero, 1 Decrement, k + 1 Min(k)
    if any_updates:
        f[:] = rng.integers(0, el + 2, (E, el))
    else:
        r = rng.random(E)
        comp = rng.integers(0, el, E)
        if el > 1:
            other = rng.integers(0, el - 1, E)
            other = other + (other >= comp)
        else:
            other = np.zeros(E, dtype=np.int64)
        dec = (r >= p_ident) & (r < p_ident + p_dec)
        mn = r >= p_ident + p_dec
        f[np.nonzero(dec)[0], comp[dec]] = 1
        f[np.nonzero(mn)[0], comp[mn]] = other[mn] + 2
    w = pack_fields(f, conf.update_bits(), conf.update_size())
    return GameGraph(conf, row.astype(np.uint32), col, w, atk)


def game_batch(n_games: int, seed: int, v_min: int = 1000, v_max: int = 200000):
    """``n_games`` random games with log-uniform position counts in [v_min, v_max]."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(v_min), np.log(v_max), n_games)).astype(np.int64)
    seeds = rng.integers(0, 2**63 - 1, n_games)
    return [random_game(int(s), int(sd)) for s, sd in zip(sizes, seeds)]


def random_lts(n_states: int, n_transitions: int, n_labels: int, seed: int):
    """Uniform sources, targets and labels.  Returns (n, src, dst, lbl) edge lists."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n_states, n_transitions, dtype=np.uint32)
    dst = rng.integers(0, n_states, n_transitions, dtype=np.uint32)
    lbl = rng.integers(0, n_labels, n_transitions).astype(np.int32)
    return n_states, src, dst, lbl


def powerlaw_lts(n_states: int, n_transitions: int, n_labels: int, seed: int, alpha: float = 2.1, max_deg: int = 200,
                 zipf_s: float = 1.2, p_tau: float = 0.1):
    """Power-law out-degrees (alpha, capped, rescaled to the exact transition
    count), uniform targets, Zipf(s) label frequencies over labels 1..n-1 and
    label 0 ("i", src/lib.rs:376) with probability p_tau."""
    rng = np.random.default_rng(seed)
    u = rng.random(n_states)
    raw = np.minimum(np.floor((1 - u) ** (-1.0 / (alpha - 1.0))), max_deg)
    deg = np.maximum(0, np.round(raw * (n_transitions / raw.sum()))).astype(np.int64)
    deg = np.minimum(deg, max_deg)
    diff = n_transitions - int(deg.sum())
    while diff != 0:
        idx = rng.integers(0, n_states, abs(diff))
        if diff > 0:
            ok = deg[idx] < max_deg
            np.add.at(deg, idx[ok], 1)
        else:
            ok = deg[idx] > 0
            np.add.at(deg, idx[ok], -1)
        deg = np.clip(deg, 0, max_deg)
        diff = n_transitions - int(deg.sum())
    src = np.repeat(np.arange(n_states, dtype=np.uint32), deg)
    dst = rng.integers(0, n_states, n_transitions, dtype=np.uint32)
    ranks = np.arange(1, n_labels, dtype=np.float64)
    pz = ranks ** (-zipf_s)
    pz /= pz.sum()
    lbl = (rng.choice(n_labels - 1, n_transitions, p=pz) + 1).astype(np.int32)
    lbl[rng.random(n_transitions) < p_tau] = 0
    return n_states, src, dst, lbl


def sample_pairs(n_states: int, k: int, seed: int) -> list:
    """k distinct ordered pairs (p, q), p != q, drawn uniformly (seeded)."""
    rng = np.random.default_rng(seed)
    pairs, seen = [], set()
    while len(pairs) < k:
        p, q = (int(x) for x in rng.integers(0, n_states, 2))
        if p != q and (p, q) not in seen:
            seen.add((p, q))
            pairs.append((p, q))
    return pairs


def serve_job(rng: np.random.Generator, n_min: int = 100, n_max: int = 20000, per_state: int = 4,
              labels_min: int = 4, labels_max: int = 64, p_tau: float = 0.05, pairs_max: int = 256):
    """One spectroscopy job of the serving workload: states log-uniform in
    [n_min, n_max], per_state transitions per state (uniform sources and
    targets), the label count uniform in [labels_min, labels_max] with label 0
    ("i") on a p_tau share of the transitions and the visible labels uniform,
    and 1..pairs_max distinct pairs (p, q), p != q.
    Returns (n, src, dst, lbl, pairs)."""
    n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
    m = per_state * n
    n_labels = int(rng.integers(labels_min, labels_max + 1))
    src = rng.integers(0, n, m, dtype=np.uint32)
    dst = rng.integers(0, n, m, dtype=np.uint32)
    lbl = rng.integers(1, n_labels, m).astype(np.int32)
    lbl[rng.random(m) < p_tau] = 0
    k = int(rng.integers(1, pairs_max + 1))
    p = rng.integers(0, n, 4 * k + 8)
    q = rng.integers(0, n, 4 * k + 8)
    seen, pairs = set(), []
    for a, b in zip(p.tolist(), q.tolist()):
        if a != b and (a, b) not in seen:
            seen.add((a, b))
            pairs.append((a, b))
            if len(pairs) == k:
                break
    return n, src, dst, lbl, pairs
