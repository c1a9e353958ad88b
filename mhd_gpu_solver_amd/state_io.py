"""State snapshots in the reference's golden text format.

The format (tests/golden_io.hpp:12-81 of the reference) is::

    MHD_GOLDEN 1
    <key> <value>                 one line per metadata entry, keys in byte order
    field <name> <n>              then n values, one per line, "%.17e"
    ...

``write_golden`` produces the same bytes as the reference writer for the same snapshot (the
reference keeps meta and fields in std::map, so both come out sorted by name); ``read_golden``
accepts what the reference reader accepts and raises ``ValueError`` where it throws.
``snapshot`` builds a snapshot of a solver's interior fields with the metadata keys the
reference's regression tests record (tests/test_regression_golden.cpp:64-78).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

MAGIC = "MHD_GOLDEN"
VERSION = 1


@dataclasses.dataclass
class GoldenSnapshot:
    meta: dict
    fields: dict


def cpp_to_string(x) -> str:
    """std::to_string: integers as is, doubles with six decimals ("%f")."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    return f"{float(x):f}"


def write_golden(path: os.PathLike, snap: GoldenSnapshot) -> None:
    lines = [f"{MAGIC} {VERSION}"]
    for key in sorted(snap.meta):
        lines.append(f"{key} {snap.meta[key]}")
    for name in sorted(snap.fields):
        vals = np.asarray(snap.fields[name], dtype=np.float64).ravel()
        lines.append(f"field {name} {vals.size}")
        lines.extend(f"{v:.17e}" for v in vals.tolist())
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_golden(path: os.PathLike) -> GoldenSnapshot:
    with open(path) as f:
        tok = f.read().split()
    if len(tok) < 2 or tok[0] != MAGIC or tok[1] != str(VERSION):
        raise ValueError(f"read_golden: bad header in {path}")
    meta, fields = {}, {}
    k = 2
    while k < len(tok):
        key = tok[k]
        if key == "field":
            if k + 2 >= len(tok):
                raise ValueError(f"read_golden: truncated field header in {path}")
            name, n = tok[k + 1], int(tok[k + 2])
            vals = tok[k + 3:k + 3 + n]
            if len(vals) != n:
                raise ValueError(f"read_golden: truncated field {name}")
            fields.setdefault(name, np.array([float(v) for v in vals]))
            k += 3 + n
        else:
            if k + 1 >= len(tok):
                raise ValueError(f"read_golden: truncated meta {key}")
            meta.setdefault(key, tok[k + 1])
            k += 2
    return GoldenSnapshot(meta, fields)


def interior_field(solver, name: str) -> np.ndarray:
    """Interior values of one field, row by row (tests/test_regression_golden.cpp:33-48)."""
    ng = solver.ng
    full = solver.field(name)
    return np.ascontiguousarray(full[ng:full.shape[0] - ng, ng:full.shape[1] - ng]).ravel()


def snapshot(solver, name: str, fields, meta=None) -> GoldenSnapshot:
    """Snapshot of `fields` (interior) with the reference's recorded keys: name, nx, ny, ng,
    steps and max_div_b (as std::to_string writes them), plus `meta` (strings, as given)."""
    m = {"name": name, "nx": cpp_to_string(solver.nx), "ny": cpp_to_string(solver.ny),
         "ng": cpp_to_string(solver.ng), "steps": cpp_to_string(solver.steps),
         "max_div_b": cpp_to_string(solver.max_div_b())}
    m.update({k: str(v) for k, v in (meta or {}).items()})
    return GoldenSnapshot(m, {f: interior_field(solver, f) for f in fields})
