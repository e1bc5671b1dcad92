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

import os
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


# ---------------------------------------------------------------------------
# Bench manifests (BMPath.xml): <BenchMark><BMFiles><BMFile [type=...]>path
# </BMFile>...</BMFiles></BenchMark>, src/model/libxml2_parser.cpp:165-383.
# ---------------------------------------------------------------------------
@dataclass
class BenchFile:
    path: str
    format: str  # the presentation's format attribute ("" when absent)


@dataclass
class BenchEntry:
    original_path: str
    resolved_path: str
    kind: str  # "file" or "dir"
    files: List[BenchFile]


def detect_format(path: str) -> str:
    """The <presentation format="..."> attribute of an XCSP file, "" if none
    (DetectFormatVersion, :385-410)."""
    try:
        root = ET.parse(path).getroot()
    except ET.ParseError:
        raise XcspError(f"Failed to parse XML file: {path}") from None
    pres = root.find("presentation")
    return pres.get("format", "") if pres is not None else ""


def describe_bench_path(original: str, resolved: str) -> BenchEntry:
    """A manifest entry resolved to a file or to every *.xml below a directory
    (BuildBenchEntry / CollectDirectoryFiles, :284-383).  Directory files are
    listed in sorted order (the reference walks them in directory order)."""
    resolved = os.path.realpath(os.path.abspath(resolved))
    if not os.path.exists(resolved):
        raise FileNotFoundError(f"Bench path not found: {resolved}")
    if os.path.isdir(resolved):
        found = []
        for root, _, files in os.walk(resolved):
            found += [os.path.join(root, f) for f in files if f.endswith(".xml")]
        if not found:
            raise FileNotFoundError(f"No XML benchmarks found under {resolved}")
        return BenchEntry(original, resolved, "dir",
                          [BenchFile(os.path.abspath(f), detect_format(f)) for f in sorted(found)])
    if os.path.isfile(resolved):
        return BenchEntry(original, resolved, "file", [BenchFile(resolved, detect_format(resolved))])
    raise XcspError(f"Bench path is neither file nor directory: {resolved}")


def is_manifest(root: ET.Element) -> bool:
    """A bench manifest: a <BMFiles> section or <BMFile> entries under the root.
    (The reference's Parse looks for a <BMFile> child of the root, :88-90,
    while its own BMPath.xml and LoadBenchManifest keep the entries in a
    <BMFiles> section; both are accepted here.)"""
    return root.find("BMFiles") is not None or root.find("BMFile") is not None


def load_bench_manifest(path: str) -> List[BenchEntry]:
    """Entries of a manifest (LoadBenchManifest, :165-282).  A relative manifest
    path is tried against the working directory, then its parent; an entry
    against the manifest's directory, then (relative entries) the working
    directory and the manifest's grandparent directory; the first that exists
    is used.  type="dir" / "directory" / "folder" must resolve to a directory,
    any other type to a file."""
    cands = [os.path.abspath(path)]
    if not os.path.isabs(path):
        cands.append(os.path.abspath(os.path.join(os.path.dirname(os.getcwd()), path)))
    manifest = next((c for c in cands if os.path.exists(c)), None)
    if manifest is None:
        raise FileNotFoundError(f"Bench manifest not found: {cands[0]}")
    try:
        root = ET.parse(manifest).getroot()
    except ET.ParseError:
        raise XcspError(f"Failed to parse XML file: {manifest}") from None
    section = root.find("BMFiles")
    if section is None:
        raise XcspError("Bench manifest missing <BMFiles> section")
    mdir = os.path.dirname(manifest)
    entries: List[BenchEntry] = []
    for node in section.findall("BMFile"):
        raw = (node.text or "").strip()
        if not raw:
            continue
        tries = [os.path.join(mdir, raw)]
        if not os.path.isabs(raw):
            tries += [os.path.abspath(raw), os.path.abspath(os.path.join(os.path.dirname(mdir), raw))]
        entry = None
        for t in tries:
            try:
                entry = describe_bench_path(raw, t)
                break
            except FileNotFoundError:
                continue
        if entry is None:
            raise FileNotFoundError(f"Bench path not found: {os.path.realpath(tries[0])}")
        kind = node.get("type")
        if kind is not None:
            want = "dir" if kind.lower() in ("dir", "directory", "folder") else "file"
            if entry.kind != want:
                raise XcspError(f"Manifest entry {raw} expected {'directory' if want == 'dir' else 'file'} "
                                f"but resolved to {'directory' if entry.kind == 'dir' else 'file'}")
        entries.append(entry)
    if not entries:
        raise XcspError("Bench manifest did not contain usable BMFile entries")
    return entries


def read(source: Union[str, bytes], max_expanded_tuples: int = 1_000_000,
         skip_unsupported: bool = False) -> XcspInstance:
    """Read an XCSP 2.x instance from a path or from XML text.  A bench
    manifest path reads the first file its entries resolve to, as the
    reference's parser does (:88-116)."""
    if isinstance(source, bytes) or (isinstance(source, str) and source.lstrip().startswith("<")):
        root = ET.fromstring(source)
    else:
        root = ET.parse(source).getroot()
        if is_manifest(root):
            entries = load_bench_manifest(source)
            first = next((e for e in entries if e.files), None)
            if first is None:
                raise XcspError("Manifest did not resolve to any XML files")
            return read(first.files[0].path, max_expanded_tuples, skip_unsupported)
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
