"""The ``.mtp`` potential format and its parameter container (host side).

Mirrors the reference's ``MTPData`` (motep_original_files/data.py:13-102)
and ``read_mtp`` / ``write_mtp`` (motep_original_files/mtp.py:14-139): same
keys, same radial block layout ``[species, species, radial_funcs,
radial_basis]`` indexed (center, neighbor), same number formatting on write.
"""
from __future__ import annotations

import itertools
import pathlib
from dataclasses import asdict, dataclass, field, fields, replace

import numpy as np

# conversion.py:14-29
MOMENTS_COUNT_TO_LEVEL = {
    1: 2, 2: 4, 8: 6, 18: 8, 41: 10, 84: 12, 174: 14, 350: 16, 718: 18,
    1352: 20, 2621: 22, 4991: 24, 9396: 26, 17366: 28,
}

POTENTIALS = pathlib.Path(__file__).resolve().parent / "potentials"


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

    def initialize(self, rng: np.random.Generator) -> None:
        """data.py:41-58: uniform initial parameters for missing groups."""
        if self.species_coeffs is None:
            self.species_coeffs = rng.uniform(-5.0, +5.0, self.species_count)
        if self.moment_coeffs is None:
            self.moment_coeffs = rng.uniform(-5.0, +5.0, self.alpha_scalar_moments)
        if self.radial_coeffs is None:
            sp, rf, rb = self.species_count, self.radial_funcs_count, self.radial_basis_size
            self.radial_coeffs = rng.uniform(-0.1, +0.1, (sp, sp, rf, rb))

    @property
    def parameters(self) -> np.ndarray:
        """data.py:60-72: flat parameter vector in (scaling, moment, species, radial) order."""
        tmp = []
        if "scaling" in self.optimized:
            tmp.append(np.atleast_1d(self.scaling))
        if "moment_coeffs" in self.optimized:
            tmp.append(self.moment_coeffs)
        if "species_coeffs" in self.optimized:
            tmp.append(self.species_coeffs)
        if "radial_coeffs" in self.optimized:
            tmp.append(np.asarray(self.radial_coeffs).ravel())
        return np.hstack(tmp)

    @parameters.setter
    def parameters(self, parameters) -> None:
        """data.py:74-102."""
        sp, rf, rb, asm = self.species_count, self.radial_funcs_count, self.radial_basis_size, self.alpha_scalar_moments
        n = 0
        if "scaling" in self.optimized:
            self.scaling = float(parameters[n]); n += 1
        if "moment_coeffs" in self.optimized:
            self.moment_coeffs = np.asarray(parameters[n:n + asm], dtype=float); n += asm
        if "species_coeffs" in self.optimized:
            self.species_coeffs = np.asarray(parameters[n:n + sp], dtype=float); n += sp
        if "radial_coeffs" in self.optimized:
            self.radial_coeffs = np.asarray(parameters[n:], dtype=float).reshape(sp, sp, rf, rb)

    def copy(self) -> "MTPData":
        out = replace(self)
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, np.ndarray):
                setattr(out, f.name, v.copy())
            elif isinstance(v, list):
                setattr(out, f.name, list(v))
        return out


def _parse_radial_coeffs(lines, data: dict) -> np.ndarray:
    """mtp.py:14-28."""
    coeffs = []
    for _ in range(data["species_count"] ** 2):
        next(lines)
        for _ in range(data["radial_funcs_count"]):
            coeffs.append([float(v) for v in next(lines).strip().strip("{}").split(",")])
    shape = (data["species_count"], data["species_count"], data["radial_funcs_count"], data["radial_basis_size"])
    return np.array(coeffs).reshape(shape)


def parse_mtp_text(text: str) -> MTPData:
    """mtp.py:31-70 on a string."""
    data: dict = {}
    lines = iter(text.splitlines())
    for line in lines:
        if line.strip() == "MTP":
            continue
        if "=" in line:
            key, value = (s.strip() for s in line.strip().split("=", 1))
            if key in {"scaling", "min_dist", "max_dist"}:
                data[key] = float(value)
            elif value.isdigit():
                data[key] = int(value)
            elif key == "alpha_moment_mapping":
                data[key] = np.fromiter((v for v in value.strip("{}").split(",")), dtype=int)
            elif key in {"species_coeffs", "moment_coeffs"}:
                data[key] = np.fromiter((v for v in value.strip().strip("{}").split(",")), dtype=float)
            elif key in {"alpha_index_basic", "alpha_index_times"}:
                rows = [[int(v) for v in s.split(",")] for s in value.strip("{}").split("}, {") if s != ""]
                arr = np.array(rows, dtype=int)
                data[key] = arr if arr.size else np.zeros((0, 4), dtype=int)
            else:
                data[key] = value.strip()
        elif line.strip() == "radial_coeffs":
            data["radial_coeffs"] = _parse_radial_coeffs(lines, data)
    known = {f.name for f in fields(MTPData)}
    return MTPData(**{k: v for k, v in data.items() if k in known})


def read_mtp(path) -> MTPData:
    return parse_mtp_text(pathlib.Path(path).read_text(encoding="utf-8"))


def _format_value(value) -> str:
    """mtp.py:73-82."""
    if isinstance(value, (float, np.floating)):
        return f"{value:21.15e}"
    if isinstance(value, (int, np.integer)):
        return f"{value:d}"
    if isinstance(value, list):
        return _format_list(value)
    if isinstance(value, np.ndarray):
        return _format_list(value.tolist())
    return str(value).strip()


def _format_list(value: list) -> str:
    """mtp.py:85-90."""
    if len(value) == 0:
        return "{}"
    if isinstance(value[0], list):
        return "{" + ", ".join(_format_list(v) for v in value) + "}"
    return "{" + ", ".join(f"{_format_value(v)}" for v in value) + "}"


def format_mtp(data: MTPData) -> str:
    """mtp.py:93-139 key order and layout."""
    keys0 = ["version", "potential_name", "scaling", "species_count", "potential_tag", "radial_basis_type"]
    keys1 = ["min_dist", "max_dist", "radial_basis_size", "radial_funcs_count"]
    keys2 = ["alpha_moments_count", "alpha_index_basic_count", "alpha_index_basic", "alpha_index_times_count",
             "alpha_index_times", "alpha_scalar_moments", "alpha_moment_mapping", "species_coeffs", "moment_coeffs"]
    d = asdict(data)
    sp = d["species_count"]
    out = ["MTP\n"]
    for key in keys0:
        if d[key] is not None:
            out.append(f"{key} = {_format_value(d[key])}\n")
    for key in keys1:
        if d[key] is not None:
            out.append(f"\t{key} = {_format_value(d[key])}\n")
    if d["radial_coeffs"] is not None:
        out.append("\tradial_coeffs\n")
        rc = np.asarray(d["radial_coeffs"])
        for k0, k1 in itertools.product(range(sp), repeat=2):
            out.append(f"\t\t{k0}-{k1}\n")
            for m in range(d["radial_funcs_count"]):
                out.append(f"\t\t\t{_format_list(rc[k0, k1][m].tolist())}\n")
    for key in keys2:
        if d[key] is not None:
            out.append(f"{key} = {_format_value(d[key])}\n")
    return "".join(out)


def write_mtp(path, data: MTPData) -> None:
    pathlib.Path(path).write_text(format_mtp(data), encoding="utf-8")


def untrained(level: int) -> MTPData:
    """Alpha tables of a level (the reference's training/untrained_mtps/NN.mtp)."""
    return read_mtp(POTENTIALS / f"untrained_{level:02d}.mtp")
