"""The ``.mtp`` potential format and its parameter container (host side).

The format is the contract with the reference and with LAMMPS users: its
keys, their order, the ``21.15e`` number format, the ``{a, b, ...}`` lists
and the radial block (one ``k0-k1`` header per (center, neighbor) species
pair, then one ``{...}`` row of Chebyshev coefficients per radial function)
are those of motep_original_files/mtp.py:14-139.  The code here is driven by
one table, ``SCHEMA``: every key's value kind, indentation and position.
The parser and the writer both walk it, and the C++ reader
(csrc/mtp_format.cpp) reads the same text.

``MTPData`` holds the values under the reference's field names
(motep_original_files/data.py:13-39); its flat parameter vector follows the
group order of data.py:60-102 (scaling, moment, species, radial), described
by the table ``PARAM_GROUPS``.
"""
from __future__ import annotations

import pathlib
import re
from dataclasses import dataclass, field, fields, replace

import numpy as np

# alpha_moments_count -> level (conversion.py:14-29)
MOMENTS_COUNT_TO_LEVEL = {
    1: 2, 2: 4, 8: 6, 18: 8, 41: 10, 84: 12, 174: 14, 350: 16, 718: 18,
    1352: 20, 2621: 22, 4991: 24, 9396: 26, 17366: 28,
}

POTENTIALS = pathlib.Path(__file__).resolve().parent / "potentials"

# (key, kind, tabs) in file order.  Kinds: text, int, real, reals (flat list),
# ints (flat list), quads (list of 4-int rows), radial (the coefficient block).
SCHEMA = (
    ("version", "text", 0),
    ("potential_name", "text", 0),
    ("scaling", "real", 0),
    ("species_count", "int", 0),
    ("potential_tag", "text", 0),
    ("radial_basis_type", "text", 0),
    ("min_dist", "real", 1),
    ("max_dist", "real", 1),
    ("radial_basis_size", "int", 1),
    ("radial_funcs_count", "int", 1),
    ("radial_coeffs", "radial", 1),
    ("alpha_moments_count", "int", 0),
    ("alpha_index_basic_count", "int", 0),
    ("alpha_index_basic", "quads", 0),
    ("alpha_index_times_count", "int", 0),
    ("alpha_index_times", "quads", 0),
    ("alpha_scalar_moments", "int", 0),
    ("alpha_moment_mapping", "ints", 0),
    ("species_coeffs", "reals", 0),
    ("moment_coeffs", "reals", 0),
)
KIND = {key: kind for key, kind, _ in SCHEMA}


@dataclass
class MTPData:
    version: str = ""
    potential_name: str = ""
    scaling: float = 1.0
    species_count: int = 0
    potential_tag: str = ""
    radial_basis_type: str = ""
    min_dist: float = float("nan")
    max_dist: float = float("nan")
    radial_funcs_count: int = 0
    radial_basis_size: int = 0
    radial_coeffs: np.ndarray | None = None
    alpha_moments_count: int = 0
    alpha_index_basic_count: int = 0
    alpha_index_basic: np.ndarray | None = None
    alpha_index_times_count: int = 0
    alpha_index_times: np.ndarray | None = None
    alpha_scalar_moments: int = 0
    alpha_moment_mapping: np.ndarray | None = None
    species_coeffs: np.ndarray | None = None
    moment_coeffs: np.ndarray | None = None
    species: np.ndarray | None = None
    optimized: list = field(default_factory=lambda: ["species_coeffs", "moment_coeffs", "radial_coeffs"])

    @property
    def level(self) -> int:
        return MOMENTS_COUNT_TO_LEVEL[self.alpha_moments_count]

    def group_shape(self, name: str) -> tuple:
        """Shape of one trainable parameter group."""
        sp = self.species_count
        return {"scaling": (1,), "moment_coeffs": (self.alpha_scalar_moments,), "species_coeffs": (sp,),
                "radial_coeffs": (sp, sp, self.radial_funcs_count, self.radial_basis_size)}[name]

    def initialize(self, rng: np.random.Generator) -> None:
        """Draw every missing coefficient group uniformly from its PARAM_GROUPS range (data.py:41-58)."""
        for name, half_width in PARAM_GROUPS:
            if half_width is not None and getattr(self, name) is None:
                shape = self.group_shape(name)
                v = rng.uniform(-half_width, half_width, shape)
                setattr(self, name, v if len(shape) > 1 else v.reshape(shape[0]))

    @property
    def parameters(self) -> np.ndarray:
        """The optimized groups concatenated in PARAM_GROUPS order (data.py:60-72)."""
        parts = [np.atleast_1d(np.asarray(getattr(self, name), dtype=float)).ravel()
                 for name, _ in PARAM_GROUPS if name in self.optimized]
        return np.concatenate(parts) if parts else np.zeros(0)

    @parameters.setter
    def parameters(self, flat) -> None:
        """Split a flat vector back into the optimized groups (data.py:74-102)."""
        flat = np.asarray(flat, dtype=float).ravel()
        pos = 0
        for name, _ in PARAM_GROUPS:
            if name not in self.optimized:
                continue
            shape = self.group_shape(name)
            n = int(np.prod(shape))
            chunk = flat[pos:pos + n]
            pos += n
            setattr(self, name, float(chunk[0]) if name == "scaling" else chunk.reshape(shape).copy())

    def copy(self) -> "MTPData":
        out = replace(self)
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, (np.ndarray, list)):
                setattr(out, f.name, v.copy())
        return out


# Trainable groups in the order of the flat parameter vector, with the
# half-width of the uniform initial draw (None: not drawn).
PARAM_GROUPS = (
    ("scaling", None),
    ("moment_coeffs", 5.0),
    ("species_coeffs", 5.0),
    ("radial_coeffs", 0.1),
)

# ---------------------------------------------------------------------------
# values <-> text, per kind
_BRACED = re.compile(r"\{([^{}]*)\}")


def _numbers(text: str, conv):
    body = text.strip().strip("{}").strip()
    return [conv(tok) for tok in body.split(",")] if body else []


def _decode(kind: str, text: str):
    text = text.strip()
    if kind == "int":
        return int(text)
    if kind == "real":
        return float(text)
    if kind == "reals":
        return np.array(_numbers(text, float), dtype=float)
    if kind == "ints":
        return np.array(_numbers(text, int), dtype=int)
    if kind == "quads":
        rows = [_numbers(m, int) for m in _BRACED.findall(text[1:-1] if text.startswith("{{") else text)]
        rows = [r for r in rows if r]
        return np.array(rows, dtype=int).reshape(-1, 4)
    return text


def _real(v) -> str:
    return f"{float(v):21.15e}"


def _braced(items) -> str:
    return "{" + ", ".join(items) + "}"


def _encode(kind: str, value) -> str:
    if kind == "int":
        return f"{int(value):d}"
    if kind == "real":
        return _real(value)
    if kind == "reals":
        return _braced(_real(v) for v in np.ravel(value))
    if kind == "ints":
        return _braced(f"{int(v):d}" for v in np.ravel(value))
    if kind == "quads":
        return _braced(_braced(f"{int(v):d}" for v in row) for row in np.asarray(value).reshape(-1, 4))
    return str(value).strip()


# ---------------------------------------------------------------------------
def parse_mtp_text(text: str) -> MTPData:
    """Read the text of a .mtp file (mtp.py:31-70 semantics)."""
    values: dict = {}
    lines = [ln.strip() for ln in text.splitlines()]
    i = 0
    while i < len(lines):
        line = lines[i]
        i += 1
        if not line or line == "MTP":
            continue
        if line == "radial_coeffs":
            i = _read_radial(lines, i, values)
            continue
        if "=" not in line:
            continue
        key, rest = (s.strip() for s in line.split("=", 1))
        kind = KIND.get(key)
        if kind is None:  # a key outside the format: kept only if MTPData has a field of that name
            values[key] = int(rest) if rest.isdigit() else rest
        else:
            values[key] = _decode(kind, rest)
    names = {f.name for f in fields(MTPData)}
    data = MTPData(**{k: v for k, v in values.items() if k in names})
    if data.radial_coeffs is not None and data.radial_coeffs.shape != data.group_shape("radial_coeffs"):
        raise ValueError("radial_coeffs does not match the final species_count, radial_funcs_count and "
                         "radial_basis_size")
    return data


def _read_radial(lines: list, i: int, values: dict) -> int:
    """The radial block: species_count^2 headers "k0-k1", each followed by radial_funcs_count rows."""
    sp, nf, nb = values["species_count"], values["radial_funcs_count"], values["radial_basis_size"]
    coeffs = np.zeros((sp, sp, nf, nb))
    for _ in range(sp * sp):
        while not lines[i]:
            i += 1
        k0, k1 = (int(t) for t in lines[i].split("-"))
        i += 1
        for mu in range(nf):
            row = _numbers(lines[i], float)
            if len(row) != nb:
                raise ValueError(f"radial_coeffs {k0}-{k1}: row {mu} holds {len(row)} values, expected {nb}")
            coeffs[k0, k1, mu] = row
            i += 1
    values["radial_coeffs"] = coeffs
    return i


def read_mtp(path) -> MTPData:
    return parse_mtp_text(pathlib.Path(path).read_text(encoding="utf-8"))


def format_mtp(data: MTPData) -> str:
    """The .mtp text of data (mtp.py:93-139 layout); groups that are None are left out."""
    out = ["MTP"]
    for key, kind, tabs in SCHEMA:
        value = getattr(data, key)
        if value is None:
            continue
        pad = "\t" * tabs
        if kind != "radial":
            out.append(f"{pad}{key} = {_encode(kind, value)}")
            continue
        out.append(f"{pad}radial_coeffs")
        rc = np.asarray(value)
        for k0 in range(data.species_count):
            for k1 in range(data.species_count):
                out.append(f"{pad}\t{k0}-{k1}")
                out.extend(f"{pad}\t\t{_encode('reals', row)}" for row in rc[k0, k1])
    return "\n".join(out) + "\n"


def write_mtp(path, data: MTPData) -> None:
    pathlib.Path(path).write_text(format_mtp(data), encoding="utf-8")


def untrained(level: int) -> MTPData:
    """Alpha tables of a level (the reference's training/untrained_mtps/NN.mtp)."""
    return read_mtp(POTENTIALS / f"untrained_{level:02d}.mtp")
