"""Minimal XCSP 2.x reader + normalizer for the reference's fixture instances.

Follows the reference's reading and normalization rules:
  * domain text containing ".." is a range min..max, otherwise a list of
    integers, sorted and deduplicated (src/model/libxml2_parser.cpp:743-782);
  * tuples are '|'-separated, values separated by blanks or commas (:804-840);
  * values become canonical indices (position in the domain's value list) and
    conflicts become the complement over the product of the scope's canonical
    domains, enumerated lexicographically (src/model/model_normalizer.cpp:99-171);
  * constraints keep file order; only extensional constraints are read.
"""
from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from itertools import product
from typing import Dict, List

import numpy as np


def _ints(text: str) -> List[int]:
    return [int(t) for t in re.split(r"[ ,\t\n\r]+", text.strip()) if t]


def parse_domain(text: str) -> List[int]:
    text = text.strip()
    if ".." in text:
        lo, hi = text.split("..", 1)
        return list(range(int(lo), int(hi) + 1))
    return sorted(set(_ints(text)))


def parse_tuples(text: str, arity: int) -> List[tuple]:
    out = []
    for part in (text or "").split("|"):
        vals = _ints(part)
        if not vals:
            continue
        if len(vals) != arity:
            raise ValueError(f"tuple size mismatch: expected {arity}, got {len(vals)}")
        out.append(tuple(vals))
    return out


def read(path: str):
    """Return (dom_size list, scopes list, support tuple lists) in canonical form."""
    root = ET.parse(path).getroot()
    domains: Dict[str, List[int]] = {}
    for d in root.iter("domain"):
        domains[d.get("name")] = parse_domain(d.text or "")
    var_index: Dict[str, int] = {}
    var_values: List[List[int]] = []
    for v in root.iter("variable"):
        var_index[v.get("name")] = len(var_values)
        var_values.append(domains[v.get("domain")])
    relations = {}
    for r in root.iter("relation"):
        arity = int(r.get("arity"))
        relations[r.get("name")] = (arity, r.get("semantics"), parse_tuples(r.text or "", arity))
    scopes, tuple_lists = [], []
    for c in root.iter("constraint"):
        ref = c.get("reference")
        if ref not in relations:
            continue  # intension / global constraints are not read
        arity, semantics, tuples = relations[ref]
        scope = [var_index[s] for s in re.split(r"[ ,\t\n]+", c.get("scope").strip()) if s]
        maps = [{val: i for i, val in enumerate(var_values[x])} for x in scope]
        mapped = [tuple(maps[i][t[i]] for i in range(arity)) for t in tuples]
        if semantics == "conflicts":
            forbidden = set(mapped)
            sizes = [len(var_values[x]) for x in scope]
            mapped = [t for t in product(*[range(s) for s in sizes]) if t not in forbidden]
        elif semantics != "supports":
            raise ValueError(f"unknown semantics {semantics}")
        scopes.append(scope)
        tuple_lists.append(np.array(mapped, dtype=np.int32).reshape(-1, arity))
    return [len(v) for v in var_values], scopes, tuple_lists
