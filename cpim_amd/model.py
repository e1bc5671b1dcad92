"""Normalized constraint networks as plain arrays (the cpim_model_desc of the C-ABI).

A ModelDesc holds numpy arrays and hands a ctypes view of them to the C
libraries; the arrays stay owned by the ModelDesc, so keep it alive while a
call runs.  Values are canonical indices 0..d-1, relations are supports
(the reference normalizes every model to that form,
src/model/model_normalizer.cpp:174-316).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _native as N


def words_for(d: int) -> int:
    return max(1, (int(d) + 63) // 64)


@dataclass
class ModelDesc:
    dom_size: np.ndarray                  # int32 [n]
    con_arity: np.ndarray                 # int32 [e]
    con_scope: np.ndarray                 # int32 [sum(arity)]
    bin_rel: Optional[np.ndarray] = None  # uint64 [n_binary, max_dom, rel_words]
    tup_off: Optional[np.ndarray] = None  # int64 [e + 1], int32 units into tuples
    tuples: Optional[np.ndarray] = None   # int32 [total tuples * arity]
    init_dom: Optional[np.ndarray] = None  # uint64 [n, words]
    name: str = ""
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.dom_size = np.ascontiguousarray(self.dom_size, dtype=np.int32)
        self.con_arity = np.ascontiguousarray(self.con_arity, dtype=np.int32)
        self.con_scope = np.ascontiguousarray(self.con_scope, dtype=np.int32)
        if self.bin_rel is not None:
            self.bin_rel = np.ascontiguousarray(self.bin_rel, dtype=np.uint64)
        if self.tup_off is not None:
            self.tup_off = np.ascontiguousarray(self.tup_off, dtype=np.int64)
        if self.tuples is not None:
            self.tuples = np.ascontiguousarray(self.tuples, dtype=np.int32)
        if self.init_dom is not None:
            self.init_dom = np.ascontiguousarray(self.init_dom, dtype=np.uint64)

    # ---- shape helpers -------------------------------------------------
    @property
    def num_vars(self) -> int:
        return int(self.dom_size.shape[0])

    @property
    def num_cons(self) -> int:
        return int(self.con_arity.shape[0])

    @property
    def max_dom(self) -> int:
        return int(self.dom_size.max()) if self.num_vars else 0

    @property
    def words(self) -> int:
        return words_for(self.max_dom)

    def scopes(self) -> List[np.ndarray]:
        off = np.concatenate([[0], np.cumsum(self.con_arity)])
        return [self.con_scope[off[c]:off[c + 1]] for c in range(self.num_cons)]

    def full_domains(self) -> np.ndarray:
        dom = np.zeros((self.num_vars, self.words), dtype=np.uint64)
        for v, d in enumerate(self.dom_size):
            for w in range(self.words):
                k = min(64, max(0, int(d) - 64 * w))
                if k == 64:
                    dom[v, w] = np.uint64(0xFFFFFFFFFFFFFFFF)
                elif k > 0:
                    dom[v, w] = np.uint64((1 << k) - 1)
        return dom

    def initial_domains(self) -> np.ndarray:
        return self.init_dom.copy() if self.init_dom is not None else self.full_domains()

    def with_init_dom(self, init_dom: Optional[np.ndarray]) -> "ModelDesc":
        return ModelDesc(self.dom_size, self.con_arity, self.con_scope, self.bin_rel,
                         self.tup_off, self.tuples, init_dom, self.name)

    # ---- C view -----------------------------------------------------------
    def c_desc(self) -> N.CModelDesc:
        def ptr(arr, ctype):
            if arr is None:
                return C.cast(None, C.POINTER(ctype))
            return arr.ctypes.data_as(C.POINTER(ctype))
        d = N.CModelDesc()
        d.num_vars = self.num_vars
        d.dom_size = ptr(self.dom_size, C.c_int32)
        d.num_cons = self.num_cons
        d.con_arity = ptr(self.con_arity, C.c_int32)
        d.con_scope = ptr(self.con_scope, C.c_int32)
        d.bin_rel = ptr(self.bin_rel, C.c_uint64)
        d.rel_words = int(self.bin_rel.shape[2]) if self.bin_rel is not None else 0
        d.tup_off = ptr(self.tup_off, C.c_int64)
        d.tuples = ptr(self.tuples, C.c_int32)
        d.init_dom = ptr(self.init_dom, C.c_uint64)
        return d

    # ---- conversions -----------------------------------------------------
    def tuple_lists(self) -> List[np.ndarray]:
        """Support tuples of every constraint as [k, arity] int arrays."""
        out = []
        scopes = self.scopes()
        bk = 0
        for c in range(self.num_cons):
            ar = int(self.con_arity[c])
            if ar == 2 and self.bin_rel is not None:
                x, y = scopes[c]
                rows = self.bin_rel[bk]
                bk += 1
                bits = np.unpackbits(rows.view(np.uint8), bitorder="little").reshape(rows.shape[0], -1)
                a, b = np.nonzero(bits[: self.dom_size[x], : self.dom_size[y]])
                out.append(np.stack([a, b], axis=1).astype(np.int32))
            else:
                t0, t1 = int(self.tup_off[c]), int(self.tup_off[c + 1])
                out.append(self.tuples[t0:t1].reshape(-1, ar))
        return out


def from_tuples(dom_size: Sequence[int], scopes: Sequence[Sequence[int]],
                tuples: Sequence[np.ndarray], init_dom: Optional[np.ndarray] = None,
                name: str = "") -> ModelDesc:
    """Build a ModelDesc from per-constraint support tuple lists."""
    arity = np.array([len(s) for s in scopes], dtype=np.int32)
    scope = np.concatenate([np.asarray(s, dtype=np.int32) for s in scopes]) if scopes else \
        np.zeros(0, np.int32)
    counts = [int(np.asarray(t).size) for t in tuples]  # int32 units
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    flat = np.concatenate([np.asarray(t, dtype=np.int32).reshape(-1) for t in tuples]) if tuples else \
        np.zeros(0, np.int32)
    return ModelDesc(np.asarray(dom_size, np.int32), arity, scope, None, off, flat, init_dom, name)


def dense_from_tuples(desc: ModelDesc) -> ModelDesc:
    """Same network with every binary relation as dense rows (bin_rel)."""
    md, W = desc.max_dom, desc.words
    scopes = desc.scopes()
    lists = desc.tuple_lists()
    bins = [c for c in range(desc.num_cons) if desc.con_arity[c] == 2]
    rel = np.zeros((len(bins), md, W), dtype=np.uint64)
    for k, c in enumerate(bins):
        t = lists[c]
        if len(t):
            a = t[:, 0].astype(np.int64)
            b = t[:, 1].astype(np.int64)
            np.bitwise_or.at(rel[k], (a, b // 64), (np.uint64(1) << (b % 64).astype(np.uint64)))
    nary = [c for c in range(desc.num_cons) if desc.con_arity[c] != 2]
    if nary:
        counts = [lists[c].size if desc.con_arity[c] != 2 else 0 for c in range(desc.num_cons)]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        flat = np.concatenate([lists[c].reshape(-1) for c in nary]).astype(np.int32)
    else:
        off, flat = None, None
    return ModelDesc(desc.dom_size, desc.con_arity, desc.con_scope, rel, off, flat,
                     desc.init_dom, desc.name)


def domains_to_sets(dom: np.ndarray, dom_size: np.ndarray) -> List[List[int]]:
    out = []
    for v in range(dom.shape[0]):
        vals = []
        for w in range(dom.shape[1]):
            x = int(dom[v, w])
            while x:
                low = x & -x
                vals.append(w * 64 + low.bit_length() - 1)
                x ^= low
        out.append([a for a in vals if a < dom_size[v]])
    return out
