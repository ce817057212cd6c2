"""MLIP ``.cfg`` training data: reader and writer (no ASE).

The format is the one the reference reads through ASE
(motep_original_files/old_io/mlip/cfg.py:14-134): blocks
``BEGIN_CFG .. END_CFG`` holding ``Size`` (atom count), ``Supercell``
(three lattice rows; absent for a non-periodic configuration),
``AtomData:`` with named columns (``id type`` and either ``cartes_x/y/z``
or ``direct_x/y/z``, optionally ``fx fy fz``), ``Energy``, ``PlusStress:``
(named components xx yy zz yz xz xy, the virial in eV) and ``Feature``
lines.  Types are 0-based species indices, as the MTP uses them.

Labels the trainer uses (motep_jax_train_import.py:188-193):
  energy; forces [n, 3]; stress = -PlusStress / volume in the reference's
  Voigt order (xx, yy, zz, yz, xz, xy), i.e. ASE's ``get_stress(voigt=True)``
  (cfg.py:131-134).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

VOIGT = ("xx", "yy", "zz", "yz", "xz", "xy")
_CART = ("cartes_x", "cartes_y", "cartes_z")
_DIRECT = ("direct_x", "direct_y", "direct_z")
_FORCE = ("fx", "fy", "fz")


@dataclass
class Configuration:
    types: np.ndarray                     # [n] species index (0-based)
    positions: np.ndarray                 # [n, 3] Cartesian, Angstrom
    cell: np.ndarray | None = None        # [3, 3] rows are lattice vectors; None: not periodic
    energy: float | None = None           # eV
    forces: np.ndarray | None = None      # [n, 3] eV/A
    plus_stress: np.ndarray | None = None  # [6] virial (eV), order xx yy zz yz xz xy
    features: dict = field(default_factory=dict)
    ids: np.ndarray | None = None

    @property
    def natoms(self) -> int:
        return len(self.types)

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.cell))) if self.cell is not None else float("nan")

    @property
    def stress(self) -> np.ndarray | None:
        """-PlusStress / V, reference Voigt order (cfg.py:131-134); None without a cell or stress."""
        if self.plus_stress is None or self.cell is None:
            return None
        return -np.asarray(self.plus_stress, dtype=float) / self.volume

    def orthogonal_box(self) -> np.ndarray:
        """Edge lengths of an orthogonal cell (the packed training batches take orthogonal cells)."""
        c = np.asarray(self.cell, dtype=float)
        if np.any(np.abs(c - np.diag(np.diag(c))) > 1e-12):
            raise ValueError("only orthogonal cells are supported by the training batches")
        return np.diag(c).copy()


def _block_lines(lines, start):
    end = start
    while end < len(lines) and not lines[end].strip().startswith("END_CFG"):
        end += 1
    if end == len(lines):
        raise ValueError(f"BEGIN_CFG at line {start} has no END_CFG")
    return lines[start + 1:end], end


def _parse_block(body) -> Configuration:
    size, cell, energy, ps = None, None, None, None
    cols, rows = None, []
    features = {}
    i = 0
    while i < len(body):
        words = body[i].split()
        i += 1
        if not words:
            continue
        head = words[0]
        if head == "Size":
            size = int(body[i].split()[0]); i += 1
        elif head in ("Supercell", "SuperCell"):
            cell = np.array([[float(v) for v in body[i + r].split()[:3]] for r in range(3)]); i += 3
        elif head in ("AtomData:", "Atomic_data:"):
            cols = words[1:]
            if size is None:
                raise ValueError("AtomData before Size")
            rows = [body[i + r].split() for r in range(size)]; i += size
            if any(len(r) != len(cols) for r in rows):
                raise ValueError("an AtomData row does not match its columns")
        elif head == "Energy":
            energy = float(body[i].split()[0]); i += 1
        elif head == "PlusStress:":
            vals = dict(zip(words[1:], (float(v) for v in body[i].split())))
            i += 1
            ps = np.array([vals[k] for k in VOIGT])
        elif head == "Feature":
            key = words[1] if len(words) > 1 else ""
            val = " ".join(words[2:])
            try:
                features[key] = float(val)
            except ValueError:
                features[key] = val
    if cols is None:
        raise ValueError("configuration without AtomData")
    col = {name: k for k, name in enumerate(cols)}
    num = np.array([[float(v) for v in r] for r in rows]).reshape(len(rows), len(cols))
    types = num[:, col["type"]].astype(np.int32)
    if all(c in col for c in _CART):
        pos = num[:, [col[c] for c in _CART]]
    elif all(c in col for c in _DIRECT):
        if cell is None:
            raise ValueError("direct coordinates need a Supercell")
        pos = num[:, [col[c] for c in _DIRECT]] @ cell
    else:
        raise ValueError("AtomData holds neither cartes_* nor direct_* columns")
    forces = num[:, [col[c] for c in _FORCE]] if all(c in col for c in _FORCE) else None
    ids = num[:, col["id"]].astype(np.int64) if "id" in col else None
    return Configuration(types, pos, cell, energy, forces, ps, features, ids)


def parse_cfg_text(text: str) -> list[Configuration]:
    lines = text.splitlines()
    out, i = [], 0
    while i < len(lines):
        if lines[i].strip().startswith("BEGIN_CFG"):
            body, end = _block_lines(lines, i)
            out.append(_parse_block(body))
            i = end + 1
        else:
            i += 1
    return out


def read_cfg(path) -> list[Configuration]:
    with open(path, encoding="utf-8") as f:
        return parse_cfg_text(f.read())


def format_cfg(configs) -> str:
    """The .cfg text of configurations (Cartesian coordinates; forces, energy, PlusStress and features
    when present)."""
    out = []
    for c in configs:
        n = c.natoms
        out.append("BEGIN_CFG")
        out.append(" Size")
        out.append(f"    {n}")
        if c.cell is not None:
            out.append(" Supercell")
            for r in np.asarray(c.cell, dtype=float):
                out.append("    " + " ".join(f"{v:14.6f}" for v in r))
        cols = ["id", "type", *_CART] + (list(_FORCE) if c.forces is not None else [])
        out.append(" AtomData:  " + " ".join(f"{k:>13s}" for k in cols))
        ids = c.ids if c.ids is not None else np.arange(1, n + 1)
        for a in range(n):
            vals = [f"{int(ids[a]):13d}", f"{int(c.types[a]):13d}"]
            vals += [f"{v:13.6f}" for v in c.positions[a]]
            if c.forces is not None:
                vals += [f"{v:13.6f}" for v in c.forces[a]]
            out.append("            " + " ".join(vals))
        if c.energy is not None:
            out.append(" Energy")
            out.append(f"    {c.energy:20.12f}")
        if c.plus_stress is not None:
            out.append(" PlusStress:  " + " ".join(f"{k:>11s}" for k in VOIGT))
            out.append("    " + " ".join(f"{v:11.5f}" for v in c.plus_stress))
        for k, v in c.features.items():
            out.append(f" Feature   {k}\t{v}")
        out.append("END_CFG")
        out.append("")
    return "\n".join(out)


def write_cfg(path, configs) -> None:
    with open(path, "w", encoding="utf-8") as f:
        f.write(format_cfg(configs))
