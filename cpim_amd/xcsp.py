"""XCSP 2.x reader and normalizer (extension constraints) for the device engine.

Reading follows the reference's XML front end:
  * a domain is a range "min..max" when its text contains "..", else a list
    of integers separated by blanks or commas, sorted and deduplicated
    (src/model/libxml2_parser.cpp:743-782);
  * tuples are separated by '|', values by blanks or commas, and every tuple
    must have the relation's arity (:804-840);
  * constraints are read in file order from their arity, scope and
    reference attributes (:595-658).
Normalization follows src/model/model_normalizer.cpp:
  * values become canonical indices 0..d-1 (position in the sorted domain)
    and a tuple value outside its domain is an error (:46-56, :60-95);
  * conflicts become the complement over the product of the scope's
    canonical domains, enumerated lexicographically with the last position
    fastest, and an expanded space above max_expanded_tuples (default
    1,000,000) is an error (:99-171; include/model/model_normalizer.h:23).
Constraints that reference no relation (predicates, globals) are an error,
or skipped and listed when skip_unsupported is set.
"""
from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Sequence, Union

import numpy as np

from .model import ModelDesc, from_tuples

_SEP = re.compile(r"[ ,\t\n\r]+")


class XcspError(ValueError):
    """Malformed or unsupported XCSP input."""


def _ints(text: str, what: str) -> List[int]:
    out = []
    for part in _SEP.split(text.strip()):
        if not part:
            continue
        try:
            out.append(int(part))
        except ValueError:
            raise XcspError(f"Failed to parse {what}: {part}") from None
    return out


def parse_domain(text: str) -> List[int]:
    """Sorted distinct values of a domain element's text."""
    text = (text or "").strip()
    if ".." in text:
        lo, hi = text.split("..", 1)
        try:
            lo_v, hi_v = int(lo.strip()), int(hi.strip())
        except ValueError:
            raise XcspError(f"Failed to parse range: {text}") from None
        return list(range(lo_v, hi_v + 1))
    values = sorted(set(_ints(text, "domain value")))
    if not values:
        raise XcspError("Domain values are empty")
    return values


def parse_tuples(text: str, arity: int) -> np.ndarray:
    """[T, arity] int64 array of the tuples of a relation element's text."""
    rows = []
    for part in (text or "").split("|"):
        vals = _ints(part, "tuple value")
        if not vals:
            continue
        if len(vals) != arity:
            raise XcspError(f"Tuple size mismatch: expected {arity}, got {len(vals)}")
        rows.append(vals)
    return np.array(rows, dtype=np.int64).reshape(-1, arity)


def _canonical(values: np.ndarray, domain: np.ndarray) -> np.ndarray:
    idx = np.searchsorted(domain, values)
    bad = (idx >= len(domain)) | (domain[np.minimum(idx, len(domain) - 1)] != values)
    if bad.any():
        v = int(values[np.argmax(bad)])
        raise XcspError(f"Value {v} not present in original domain during normalization")
    return idx


@dataclass
class XcspInstance:
    name: str
    var_names: List[str]
    var_values: List[List[int]]  # canonical index -> original value, per variable
    desc: ModelDesc
    skipped: List[str] = field(default_factory=list)

    @property
    def num_vars(self) -> int:
        return len(self.var_names)

    @property
    def num_cons(self) -> int:
        return self.desc.num_cons

    def original_solution(self, canonical: Sequence[int]) -> List[int]:
        return [self.var_values[v][int(a)] for v, a in enumerate(canonical)]


def read(source: Union[str, bytes], max_expanded_tuples: int = 1_000_000,
         skip_unsupported: bool = False) -> XcspInstance:
    """Read an XCSP 2.x instance from a path or from XML text."""
    if isinstance(source, bytes) or (isinstance(source, str) and source.lstrip().startswith("<")):
        root = ET.fromstring(source)
    else:
        root = ET.parse(source).getroot()
    pres = root.find("presentation")
    name = pres.get("name", "") if pres is not None else ""
    domains: Dict[str, np.ndarray] = {}
    for d in root.iter("domain"):
        domains[d.get("name")] = np.array(parse_domain(d.text or ""), dtype=np.int64)
    var_index: Dict[str, int] = {}
    var_names: List[str] = []
    var_dom: List[np.ndarray] = []
    for v in root.iter("variable"):
        dn = v.get("domain")
        if dn not in domains:
            raise XcspError(f"Variable {v.get('name')} names an unknown domain {dn}")
        var_index[v.get("name")] = len(var_names)
        var_names.append(v.get("name"))
        var_dom.append(domains[dn])
    relations = {}
    for r in root.iter("relation"):
        try:
            arity = int(r.get("arity"))
        except (TypeError, ValueError):
            raise XcspError(f"Relation {r.get('name')} has no arity attribute") from None
        sem = r.get("semantics")
        if sem not in ("supports", "conflicts"):
            raise XcspError(f"Relation {r.get('name')} has unknown semantics {sem}")
        relations[r.get("name")] = (arity, sem, parse_tuples(r.text or "", arity))
    scopes: List[List[int]] = []
    lists: List[np.ndarray] = []
    skipped: List[str] = []
    for i, c in enumerate(root.iter("constraint")):
        label = c.get("name") or f"constraint {i}"
        if c.get("arity") is None:
            raise XcspError(f"Constraint {i} has no arity attribute")
        if c.get("scope") is None:
            raise XcspError(f"Constraint {i} has no scope attribute")
        ref = c.get("reference")
        if ref is None:
            raise XcspError(f"Constraint {i} has no reference attribute")
        try:
            scope = [var_index[s] for s in _SEP.split(c.get("scope").strip()) if s]
        except KeyError as e:
            raise XcspError(f"Variable not found in scope: {e.args[0]}") from None
        if ref not in relations:
            if skip_unsupported:
                skipped.append(label)
                continue
            raise XcspError(f"{label} references {ref}, which is not an extension relation")
        arity, sem, tuples = relations[ref]
        if arity != len(scope) or int(c.get("arity")) != arity:
            raise XcspError(f"{label}: scope of {len(scope)} variables for a relation of arity {arity}")
        doms = [var_dom[x] for x in scope]
        mapped = np.stack([_canonical(tuples[:, k], doms[k]) for k in range(arity)], 1) \
            if len(tuples) else np.zeros((0, arity), np.int64)
        if sem == "conflicts":
            sizes = [len(d) for d in doms]
            total = 1
            for s in sizes:
                total *= s
                if total > max_expanded_tuples:
                    raise XcspError(f"Expanded tuple space exceeds limit ({total}) while normalizing conflicts.")
            keep = np.ones(total, dtype=bool)
            if len(mapped):
                keep[np.ravel_multi_index(tuple(mapped.T), sizes)] = False
            mapped = np.stack(np.unravel_index(np.nonzero(keep)[0], sizes), 1)
        scopes.append(scope)
        lists.append(mapped.astype(np.int32).reshape(-1, arity))
    desc = from_tuples([len(d) for d in var_dom], scopes, lists, name=name or "xcsp")
    return XcspInstance(name, var_names, [d.tolist() for d in var_dom], desc, skipped)
