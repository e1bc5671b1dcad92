"""Seeded synthetic workloads (random binary CSPs, ternary tables, uneven batches).

Base seed 0x5EED2026; the seed of instance `index` of workload `config_id` is
splitmix64(base ^ config_id ^ index).  Pairs are drawn uniformly without
replacement over the n(n-1)/2 variable pairs; every tuple of a binary
relation is forbidden independently with probability p.  All draws are
integer arithmetic in libcpim_workload.so, so instances are bit-identical
everywhere.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _native as N
from .model import ModelDesc, words_for

BASE_SEED = 0x5EED2026


def instance_seed(config_id: int, index: int, base: int = BASE_SEED) -> int:
    return int(N.workload_lib().cpimw_instance_seed(base, config_id, index))


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def random_binary(n: int, d: int, e: int, p_forbid: float, seed: int,
                  preremove: float = 0.0, name: str = "") -> ModelDesc:
    """Random binary CSP with dense relations (every tuple forbidden w.p. p)."""
    lib = N.workload_lib()
    W = words_for(d)
    scope = np.zeros(2 * e, dtype=np.int32)
    if lib.cpimw_random_pairs(seed, n, e, _p(scope, C.c_int32)) != 0:
        raise ValueError(f"cannot draw {e} distinct pairs over {n} variables")
    rel = np.zeros((e, d, W), dtype=np.uint64)
    if lib.cpimw_random_relations(seed, e, d, d, W, p_forbid, _p(rel, C.c_uint64)) != 0:
        raise ValueError("bad relation parameters")
    init = None
    if preremove > 0.0:
        init = np.zeros((n, W), dtype=np.uint64)
        lib.cpimw_domains_with_removals(seed, n, d, W, preremove, _p(init, C.c_uint64))
    return ModelDesc(np.full(n, d, np.int32), np.full(e, 2, np.int32), scope, rel,
                     None, None, init, name)


def phase_transition_tightness(n: int, d: int, density: float) -> float:
    """Model-B estimate p* = 1 - d^(-2 / (density (n - 1)))."""
    return 1.0 - d ** (-2.0 / (density * (n - 1)))


def random_table(seed: int, arity: int, d: int, ntuples: int) -> np.ndarray:
    t = np.zeros(ntuples * arity, dtype=np.int32)
    if N.workload_lib().cpimw_random_table(seed, arity, d, ntuples, _p(t, C.c_int32)) != 0:
        raise ValueError("bad table parameters")
    return t.reshape(ntuples, arity)


def uniform_int(seed: int, k: int, lo: int, hi: int) -> int:
    return int(N.workload_lib().cpimw_uniform_int(seed, k, lo, hi))


def uniform_real(seed: int, k: int, lo: float, hi: float) -> float:
    return float(N.workload_lib().cpimw_uniform_real(seed, k, lo, hi))


# ---------------------------------------------------------------------------
# The five benchmark workloads
# ---------------------------------------------------------------------------

def small_plain(index: int = 0) -> ModelDesc:
    """n=30, d=16, e=120, p=0.40: root AC3bit fixpoint then MAC (CPU case)."""
    return random_binary(30, 16, 120, 0.40, instance_seed(1, index), name=f"c1-{index}")


def single_large(index: int = 0) -> ModelDesc:
    """n=2000, d=64, e=40000, t=0.25, 10% of values pre-removed: one GAC fixpoint."""
    return random_binary(2000, 64, 40000, 0.25, instance_seed(2, index), preremove=0.10,
                         name=f"c2-{index}")


def batched_sac(index: int = 0, n: int = 500, d: int = 256, e: int = 20000,
                p: float = 0.5) -> ModelDesc:
    """n=500, d=256, e=20000, t=0.5: every singleton test is one GAC fixpoint."""
    return random_binary(n, d, e, p, instance_seed(3, index), name=f"c3-{index}")


# Tightness of the cascade variants of the two large shapes.  At t=0.25 the
# single fixpoint deletes nothing and at t=0.5 every singleton test survives;
# these tightnesses were found with the oracle so that the same shapes exercise
# long propagation chains and failures.
CASCADE_FIXPOINT_TIGHTNESS = 0.892
CASCADE_SAC_TIGHTNESS = 0.86


def single_large_cascade(index: int = 1) -> ModelDesc:
    """n=2000, d=64, e=40000, t=0.892, 10% of values pre-removed: the root
    fixpoint runs 25 frontier rounds and deletes 13% of the values (15,369 of
    115,143) without a wipe-out (index 1)."""
    return random_binary(2000, 64, 40000, CASCADE_FIXPOINT_TIGHTNESS, instance_seed(2, index),
                         preremove=0.10, name=f"c2-cascade-{index}")


def batched_sac_cascade(index: int = 1) -> ModelDesc:
    """n=500, d=256, e=20000, t=0.86: about 42% of the 128,000 singleton tests
    wipe out after long propagation chains (index 1)."""
    return random_binary(500, 256, 20000, CASCADE_SAC_TIGHTNESS, instance_seed(3, index),
                         name=f"c3-cascade-{index}")


UNEVEN_DOMAINS = (8, 16, 32, 64, 128)


@dataclass
class UnevenParams:
    n: int
    d: int
    density: float
    e: int
    p: float


def uneven_params(index: int) -> UnevenParams:
    s = instance_seed(4, index)
    n = uniform_int(s, 1, 20, 200)
    d = UNEVEN_DOMAINS[uniform_int(s, 2, 0, len(UNEVEN_DOMAINS) - 1)]
    rho = uniform_real(s, 3, 0.1, 0.5)
    e = int(round(rho * n * (n - 1) / 2))
    p = phase_transition_tightness(n, d, rho)
    return UnevenParams(n, d, rho, e, p)


def uneven_instance(index: int) -> ModelDesc:
    """One instance of the 4096-instance uneven batch (n~U[20,200], d in {8..128},
    density~U[0.1,0.5], tightness at the model-B phase transition)."""
    q = uneven_params(index)
    return random_binary(q.n, q.d, q.e, q.p, instance_seed(4, index) ^ 0xB,
                         name=f"c4-{index}")


def serving_request_c3(index: int, n: int = 500, d: int = 256, e: int = 20000, p: float = 0.5,
                       n_tables: int = 4, table_tuples: int = 500, table_dom: int = 16) -> ModelDesc:
    """A serving request of the benchmark SAC network's shape: a random binary
    CSP n=500, d=256, e=20000, t=0.5 (its own seed per request) plus n_tables
    arity-3 positive tables over distinct random triples of its variables,
    each with table_tuples random distinct tuples over the values
    0..table_dom-1 (the tables therefore restrict their variables to those 16
    values)."""
    s = instance_seed(6, index)
    base = random_binary(n, d, e, p, s ^ 0xC, name=f"c5c3-{index}")
    triples, tables = [], []
    for t in range(n_tables):
        a = uniform_int(s, 100 + 3 * t, 0, n - 1)
        b = uniform_int(s, 101 + 3 * t, 0, n - 2)
        b = b + 1 if b >= a else b
        c = uniform_int(s, 102 + 3 * t, 0, n - 3)
        lo, hi = sorted((a, b))
        c = c + 1 if c >= lo else c
        c = c + 1 if c >= hi else c
        triples.append(np.array([a, b, c], np.int32))
        tables.append(random_table(s ^ (0x7000 + t), 3, table_dom, table_tuples))
    arity = np.concatenate([base.con_arity, np.full(n_tables, 3, np.int32)])
    scope = np.concatenate([base.con_scope] + triples).astype(np.int32)
    counts = [0] * e + [t.size for t in tables]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    flat = np.concatenate([t.reshape(-1) for t in tables]).astype(np.int32)
    return ModelDesc(base.dom_size, arity, scope, base.bin_rel, off, flat, None, f"c5c3-{index}")


def serving_request(index: int, n_lo: int = 20, n_hi: int = 40, d: int = 16,
                    density: float = 0.25, n_tables: int = 4, table_tuples: int = 500,
                    tight_scale: float = 0.85) -> ModelDesc:
    """A serving request: random binary CSP (tightness slightly below the
    phase transition so that MAC has real search) plus arity-3 positive tables
    with table_tuples random tuples over d=16 (binary part dense, tables as
    tuple lists)."""
    s = instance_seed(5, index)
    n = uniform_int(s, 1, n_lo, n_hi)
    e = int(round(density * n * (n - 1) / 2))
    p = tight_scale * phase_transition_tightness(n, d, density)
    base = random_binary(n, d, e, p, s ^ 0xC, name=f"c5-{index}")
    # Ternary tables over distinct random triples.
    tables = []
    triples = []
    for t in range(n_tables):
        a = uniform_int(s, 100 + 3 * t, 0, n - 1)
        b = uniform_int(s, 101 + 3 * t, 0, n - 2)
        b = b + 1 if b >= a else b
        c = uniform_int(s, 102 + 3 * t, 0, n - 3)
        lo, hi = sorted((a, b))
        c = c + 1 if c >= lo else c
        c = c + 1 if c >= hi else c
        triples.append(np.array([a, b, c], np.int32))
        tables.append(random_table(s ^ (0x7000 + t), 3, d, table_tuples))
    arity = np.concatenate([base.con_arity, np.full(n_tables, 3, np.int32)])
    scope = np.concatenate([base.con_scope] + triples).astype(np.int32)
    counts = [0] * e + [t.size for t in tables]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    flat = np.concatenate([t.reshape(-1) for t in tables]).astype(np.int32)
    return ModelDesc(base.dom_size, arity, scope, base.bin_rel, off, flat, None,
                     f"c5-{index}")
