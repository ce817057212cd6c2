"""FP64 CPU oracle for the MTP evaluation path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It is the checker, never the measured or shipped path.

Contents, each restating the reference (sebastian57/JAX-MTP):

* ``read_mtp``            mtp.py:14-70 (the ``.mtp`` text format), numpy.
* ``load_tree``           moment_jax.py:116-211, 342-373 and the
                          ``_flatten_computation_graph`` order of
                          compile_cuda.py:144-176, from the reference's
                          precomputed JSON trees (copied as data into
                          oracle/data/).
* ``remap_coeffs``        conversion.py:112-168 (BasisConverter), with the
                          test vectors of jax_engine/utils.py:3-13.
* ``Oracle.compact``      conversion.py:39-102 compact form + hand adjoint.
* ``Oracle.tree``         jax_backend.py:162-247 full-tensor form + adjoint.
* ``Oracle.forces`` etc.  jax_backend.py:263-303, 355-370, 431-456.
* ``Oracle.pack/stage``   pair_jax_mtp.cpp:96-173, zero_overhead.cpp:61-117.

The numerical kernels live in mtp_oracle.c (built by oracle/Makefile).

Parity pins (see tests/test_oracle.py): compact vs tree agreement, the
reference's own remap check at its test vectors, closed forms (no
neighbors, level-2 pair potential, single neighbor), central finite
differences, invariances.  Parity unpinned: nothing in this module beyond
those checks; the reference's JAX outputs cannot be produced here (no JAX).
"""
from __future__ import annotations

import ctypes
import json
import os
import pathlib
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
DATA = HERE / "data"
LIB_PATH = HERE / "libmtp_oracle.so"

# jax_engine/utils.py:3-13
TEST_RS = np.array(
    [
        [0.234, 0.456, 0.789],
        [0.987, 0.423, 0.765],
        [-1.120, 0.123, 0.563],
        [-0.120, 0.012, -0.001],
    ]
)
TEST_R_ABS = np.linalg.norm(TEST_RS, axis=1)
TEST_RB_VALUES = np.array([TEST_R_ABS ** (mu + 1) for mu in range(10)])
TEST_R_UNITS = TEST_RS / TEST_R_ABS[:, None]

# conversion.py:14-29
MOMENTS_COUNT_TO_LEVEL = {
    1: 2, 2: 4, 8: 6, 18: 8, 41: 10, 84: 12, 174: 14, 350: 16, 718: 18,
    1352: 20, 2621: 22, 4991: 24, 9396: 26, 17366: 28,
}


# ---------------------------------------------------------------------------
# .mtp reader (mtp.py:14-70)
# ---------------------------------------------------------------------------
@dataclass
class OracleMTP:
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
    extra: dict = field(default_factory=dict)

    @property
    def level(self) -> int:
        return MOMENTS_COUNT_TO_LEVEL[self.alpha_moments_count]


def read_mtp(path) -> OracleMTP:
    data: dict = {}
    with open(path, "r", encoding="utf-8") as fd:
        lines = iter(fd.readlines())
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
                data[key] = np.array([int(v) for v in value.strip("{}").split(",") if v.strip()], dtype=np.int64)
            elif key in {"species_coeffs", "moment_coeffs"}:
                data[key] = np.array([float(v) for v in value.strip().strip("{}").split(",") if v.strip()])
            elif key in {"alpha_index_basic", "alpha_index_times"}:
                rows = [[int(v) for v in s.split(",")] for s in value.strip("{}").split("}, {") if s != ""]
                arr = np.array(rows, dtype=np.int64)
                data[key] = arr if arr.size else np.zeros((0, 4), dtype=np.int64)
            else:
                data[key] = value.strip()
        elif line.strip() == "radial_coeffs":
            sp, rf, rb = data["species_count"], data["radial_funcs_count"], data["radial_basis_size"]
            coeffs = []
            for _ in range(sp * sp):
                next(lines)
                for _ in range(rf):
                    coeffs.append([float(v) for v in next(lines).strip().strip("{}").split(",")])
            data["radial_coeffs"] = np.array(coeffs).reshape(sp, sp, rf, rb)
    known = set(OracleMTP.__dataclass_fields__)
    extra = {k: v for k, v in data.items() if k not in known}
    for k in extra:
        data.pop(k)
    return OracleMTP(**data, extra=extra)


# ---------------------------------------------------------------------------
# JSON contraction tree (moment_jax.py), execution order per compile_cuda.py
# ---------------------------------------------------------------------------
def _to_tuple(x):
    return tuple(_to_tuple(i) if isinstance(i, list) else i for i in x)


def _is_basic(node) -> bool:
    return not isinstance(node[0], tuple)


def _flatten_leaves(node, out):
    if _is_basic(node):
        out.append(node)
        return
    _flatten_leaves(node[0], out)
    _flatten_leaves(node[1], out)


def _pairs_rec(node):
    """Post-order list of pair contractions (moment_jax.py:362-373)."""
    if _is_basic(node):
        return []
    res = []
    for c in _pairs_rec(node[0]) + _pairs_rec(node[1]):
        if c not in res:
            res.append(c)
    res.append(node)
    return res


@dataclass
class OracleTree:
    level: int
    scalar_contractions: tuple
    basic: list
    pairs: list
    order: list  # node keys in execution order
    index: dict


def load_tree(level: int) -> OracleTree:
    path = DATA / f"moments_level{level}.json"
    with open(path) as f:
        scalars = _to_tuple(json.load(f))
    basic = []
    for s in scalars:
        leaves = []
        _flatten_leaves(s, leaves)
        for m in leaves:
            if m not in basic:
                basic.append(m)
    pairs = []
    for s in scalars:
        for c in _pairs_rec(s):
            if c not in pairs:
                pairs.append(c)
    # compile_cuda.py:144-176: basics, then repeatedly the first ready contraction
    order = list(basic)
    ready = set(basic)
    remaining = list(pairs)
    while remaining:
        for i, c in enumerate(remaining):
            if c[0] in ready and c[1] in ready:
                order.append(c)
                ready.add(c)
                remaining.pop(i)
                break
        else:
            raise ValueError("Circular dependency in contraction graph")
    index = {k: i for i, k in enumerate(order)}
    return OracleTree(level, scalars, basic, pairs, order, index)


# ---------------------------------------------------------------------------
# ctypes binding
# ---------------------------------------------------------------------------
class _CMTP(ctypes.Structure):
    _fields_ = [
        ("Sp", ctypes.c_int), ("R", ctypes.c_int), ("N", ctypes.c_int), ("B", ctypes.c_int),
        ("T", ctypes.c_int), ("amc", ctypes.c_int), ("S", ctypes.c_int),
        ("scaling", ctypes.c_double), ("min_dist", ctypes.c_double), ("max_dist", ctypes.c_double),
        ("radial_coeffs", ctypes.c_void_p), ("species_coeffs", ctypes.c_void_p),
        ("moment_coeffs", ctypes.c_void_p), ("basic", ctypes.c_void_p), ("times", ctypes.c_void_p),
        ("mapping", ctypes.c_void_p),
    ]


_lib = None


def build(force: bool = False) -> None:
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < (HERE / "mtp_oracle.c").stat().st_mtime:
        subprocess.check_call(["make", "-s", "-C", str(HERE), "-B" if force else "all"])


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        I = ctypes.c_int
        L.orc_compact_atoms.argtypes = [P, I, I, P, P, P, P, P, I]
        L.orc_compact_basis_and_dc.argtypes = [P, I, I, P, P, P, P, P, P]
        L.orc_compact_atoms_xi.argtypes = [P, I, I, P, P, P, P, I, P, P, P, I]
        L.orc_compact_basis_rb.argtypes = [P, I, P, P, I, P]
        L.orc_tree_build.argtypes = [I, P, P, P, P, P, P, P, P, I, P]
        L.orc_tree_build.restype = P
        L.orc_tree_free.argtypes = [P]
        L.orc_tree_basis_rb.argtypes = [P, I, P, P, I, P]
        L.orc_tree_atoms.argtypes = [P, P, P, I, I, P, P, P, P, P, I]
        L.orc_accumulate_forces.argtypes = [I, I, P, P, I, P]
        L.orc_stress.argtypes = [I, I, P, P, I, ctypes.c_double, P, P]
        L.orc_energy.argtypes = [I, P, I]
        L.orc_energy.restype = ctypes.c_double
        L.orc_pack.argtypes = [I, I, P, P, P, I, P, P, P, ctypes.c_double, P, P, I, I, P, P, P, P, P]
        L.orc_pack.restype = ctypes.c_longlong
        L.orc_stage.argtypes = [I, I, I, P, P, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """FP64 evaluator bound to one set of MTP parameters."""

    def __init__(self, mtp: OracleMTP, with_tree: bool = False):
        self.mtp = mtp
        self._keep = dict(
            radial=_f64(mtp.radial_coeffs), species=_f64(mtp.species_coeffs), moment=_f64(mtp.moment_coeffs),
            basic=_i32(mtp.alpha_index_basic).reshape(-1, 4), times=_i32(mtp.alpha_index_times).reshape(-1, 4),
            mapping=_i32(mtp.alpha_moment_mapping),
        )
        k = self._keep
        self.c = _CMTP(
            mtp.species_count, mtp.radial_funcs_count, mtp.radial_basis_size, len(k["basic"]), len(k["times"]),
            mtp.alpha_moments_count, len(k["mapping"]), mtp.scaling, mtp.min_dist, mtp.max_dist,
            k["radial"].ctypes.data, k["species"].ctypes.data, k["moment"].ctypes.data, k["basic"].ctypes.data,
            k["times"].ctypes.data, k["mapping"].ctypes.data,
        )
        self.tree = None
        self.xi_tree = None
        if with_tree:
            self._build_tree()

    # -- tree ---------------------------------------------------------------
    def _build_tree(self):
        t = load_tree(self.mtp.level)
        n = len(t.order)
        kind = np.zeros(n, np.int32); mu = np.zeros(n, np.int32); nu = np.zeros(n, np.int32)
        left = np.full(n, -1, np.int32); right = np.full(n, -1, np.int32); naxes = np.zeros(n, np.int32)
        laxes = np.zeros((n, 16), np.int32); raxes = np.zeros((n, 16), np.int32)
        for i, key in enumerate(t.order):
            if _is_basic(key):
                mu[i], nu[i] = key[0], key[1]
            else:
                kind[i] = 1
                left[i] = t.index[key[0]]; right[i] = t.index[key[1]]
                la, ra = key[3]
                naxes[i] = len(la)
                laxes[i, : len(la)] = la; raxes[i, : len(ra)] = ra
        scal = np.array([t.index.get(s, -1) for s in t.scalar_contractions], np.int32)
        self._tree_arrays = (kind, mu, nu, left, right, naxes, laxes, raxes, scal)
        self.tree = lib().orc_tree_build(n, _p(kind), _p(mu), _p(nu), _p(left), _p(right), _p(naxes), _p(laxes),
                                         _p(raxes), len(scal), _p(scal))
        self.tree_info = t
        self.xi_tree = remap_coeffs(self)

    def __del__(self):
        if getattr(self, "tree", None):
            try:
                lib().orc_tree_free(self.tree)
            except Exception:
                pass
            self.tree = None

    # -- basis at given unit vectors / radial values (remap support) --------
    def compact_basis_rb(self, u, rb):
        u = _f64(u); rb = _f64(rb)
        out = np.zeros(len(self._keep["mapping"]))
        lib().orc_compact_basis_rb(ctypes.byref(self.c), u.shape[0], _p(u), _p(rb), rb.shape[1], _p(out))
        return out

    def tree_basis_rb(self, u, rb):
        u = _f64(u); rb = _f64(rb)
        out = np.zeros(len(self.tree_info.scalar_contractions))
        lib().orc_tree_basis_rb(self.tree, u.shape[0], _p(u), _p(rb), rb.shape[1], _p(out))
        return out

    # -- per-atom energies and slot gradients --------------------------------
    def compact(self, itypes, rij, jtypes, grad=True, nthreads=0):
        itypes = _i32(itypes); rij = _f64(rij); jtypes = _i32(jtypes)
        A, K = jtypes.shape
        Ei = np.zeros(A)
        g = np.zeros((A, K, 3)) if grad else None
        rc = lib().orc_compact_atoms(ctypes.byref(self.c), A, K, _p(itypes), _p(rij), _p(jtypes), _p(Ei),
                                     _p(g) if grad else None, nthreads)
        if rc:
            raise RuntimeError(f"orc_compact_atoms failed ({rc})")
        return Ei, g

    def compact_xi(self, itypes, rij, jtypes, xi, grad=True, nthreads=0):
        """Per-row coefficient vectors xi[A][S] (or one vector [S]) in place of
        moment_coeffs: (Ei, g, basis)."""
        itypes = _i32(itypes); rij = _f64(rij); jtypes = _i32(jtypes); xi = _f64(xi)
        A, K = jtypes.shape
        S = len(self._keep["mapping"])
        stride = 0 if xi.ndim == 1 else S
        assert xi.shape[-1] == S and (xi.ndim == 1 or xi.shape[0] == A)
        Ei = np.zeros(A)
        g = np.zeros((A, K, 3)) if grad else None
        basis = np.zeros((A, S))
        rc = lib().orc_compact_atoms_xi(ctypes.byref(self.c), A, K, _p(itypes), _p(rij), _p(jtypes), _p(xi), stride,
                                        _p(Ei), _p(g) if grad else None, _p(basis), nthreads)
        if rc:
            raise RuntimeError(f"orc_compact_atoms_xi failed ({rc})")
        return Ei, g, basis

    def basis_jacobian(self, itypes, rij, jtypes, nthreads=0):
        """J[A][S][K][3] = d basis / d r (one per-row-coefficient pass per s)."""
        A, K = np.asarray(jtypes).shape
        S = len(self._keep["mapping"])
        J = np.zeros((A, S, K, 3))
        for s in range(S):
            e = np.zeros(S)
            e[s] = 1.0
            J[:, s] = self.compact_xi(itypes, rij, jtypes, e, nthreads=nthreads)[1]
        return J

    def basis_and_dc(self, itypes, rij, jtypes):
        """Per-atom basis [A,S] and d(sum_i E_i)/d radial_coeffs."""
        itypes = _i32(itypes); rij = _f64(rij); jtypes = _i32(jtypes)
        A, K = jtypes.shape
        Ei = np.zeros(A)
        basis = np.zeros((A, len(self._keep["mapping"])))
        dc = np.zeros_like(self._keep["radial"])
        lib().orc_compact_basis_and_dc(ctypes.byref(self.c), A, K, _p(itypes), _p(rij), _p(jtypes), _p(Ei),
                                       _p(basis), _p(dc))
        return Ei, basis, dc

    def tree_eval(self, itypes, rij, jtypes, grad=True, nthreads=0):
        if self.tree is None:
            self._build_tree()
        itypes = _i32(itypes); rij = _f64(rij); jtypes = _i32(jtypes)
        A, K = jtypes.shape
        Ei = np.zeros(A)
        g = np.zeros((A, K, 3)) if grad else None
        xi = _f64(self.xi_tree)
        rc = lib().orc_tree_atoms(ctypes.byref(self.c), self.tree, _p(xi), A, K, _p(itypes), _p(rij), _p(jtypes),
                                  _p(Ei), _p(g) if grad else None, nthreads)
        if rc:
            raise RuntimeError(f"orc_tree_atoms failed ({rc})")
        return Ei, g

    # -- system-level outputs ------------------------------------------------
    @staticmethod
    def forces(g, js, natoms_force):
        g = _f64(g); js = _i32(js)
        A, K = js.shape
        F = np.zeros((A, 3))
        lib().orc_accumulate_forces(A, K, _p(g), _p(js), natoms_force, _p(F))
        return F

    @staticmethod
    def stress(rij, g, volume, cell_rank=3):
        rij = _f64(rij); g = _f64(g)
        A, K = g.shape[:2]
        s6 = np.zeros(6); W = np.zeros(9)
        lib().orc_stress(A, K, _p(rij), _p(g), cell_rank, float(volume), _p(s6), _p(W))
        return s6, W.reshape(3, 3)

    @staticmethod
    def energy(Ei, natoms_energy):
        Ei = _f64(Ei)
        return lib().orc_energy(len(Ei), _p(Ei), natoms_energy)

    def calc_energy_forces_stress_padded(self, itypes, all_js, all_rijs, all_jtypes, cell_rank, volume,
                                         natoms_energy, natoms_force, form="compact", nthreads=0):
        """FP64 restatement of jax_backend.py:374-460 (returns E, F[A,3], stress6, Ei, g).

        The reference passes natoms_energy positionally into the natoms_force
        parameter of calc_energy_forces_stress (jax_backend.py:414-418), which
        is the mask of accumulate_forces (:279-301): forces are masked by
        natoms_energy and natoms_force has no effect."""
        if form == "compact":
            Ei, g = self.compact(itypes, all_rijs, all_jtypes, True, nthreads)
        else:
            Ei, g = self.tree_eval(itypes, all_rijs, all_jtypes, True, nthreads)
        F = self.forces(g, all_js, natoms_energy)
        s6, _ = self.stress(all_rijs, g, volume, cell_rank)
        E = self.energy(Ei, natoms_energy)
        return E, F, s6, Ei, g


def remap_coeffs(orc: Oracle) -> np.ndarray:
    """conversion.py:112-168: map .mtp moment_coeffs onto the tree's scalar order."""
    mtp = orc.mtp
    mlip_vals = orc.compact_basis_rb(TEST_R_UNITS, TEST_RB_VALUES)
    bc_map = {}
    for coef, v in zip(mtp.moment_coeffs, mlip_vals):
        bc_map[float(v)] = coef
    tree_vals = orc.tree_basis_rb(TEST_R_UNITS, TEST_RB_VALUES)
    remapped = []
    remaining = list(bc_map.keys())
    missing = []
    for bv in tree_vals:
        for mv, coef in bc_map.items():
            if np.isclose(mv, bv, rtol=1e-8):
                remapped.append(coef)
                remaining.remove(mv)
                break
        else:
            missing.append(bv)
            remapped.append(0.0)
    if remaining:
        raise RuntimeError(f"Not all MLIP contractions found: {remaining}")
    orc.remap_missing = missing
    return np.array(remapped)


# ---------------------------------------------------------------------------
# Training: packed-batch outputs and the exact parameter gradient
# ---------------------------------------------------------------------------
def batch_outputs(orc: Oracle, itypes, js, rij, jtypes, conf_off, volumes, nthreads=0):
    """Per-configuration E[C], forces[N,3], stress[C,6] (reference Voigt), per-slot g."""
    Ei, g = orc.compact(itypes, rij, jtypes, True, nthreads)
    N = len(itypes)
    F = Oracle.forces(g, js, N)
    C = len(volumes)
    E = np.array([Ei[conf_off[c]:conf_off[c + 1]].sum() for c in range(C)])
    S = np.zeros((C, 6))
    for c in range(C):
        sl = slice(conf_off[c], conf_off[c + 1])
        S[c], _ = Oracle.stress(rij[sl], g[sl], volumes[c])
    return E, F, S, g


def _alpha_delta(itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS):
    """alpha_i and Delta_ik of J = sum alpha E + sum g . Delta (the loss's linearisation)."""
    N = len(itypes)
    conf_of = np.repeat(np.arange(len(volumes)), np.diff(conf_off))
    alpha = np.asarray(gE)[conf_of]
    gF = np.asarray(gF).reshape(N, 3)
    valid_j = js < N
    Delta = gF[:, None, :] - np.where(valid_j[..., None], gF[np.minimum(js, N - 1)], 0.0)
    for c in range(len(volumes)):
        s = np.asarray(gS)[c]
        M = np.array([[s[0], s[5], s[4]], [0.0, s[1], s[3]], [0.0, 0.0, s[2]]])
        Lam = (M + M.T) / (2.0 * volumes[c])
        sl = slice(conf_off[c], conf_off[c + 1])
        Delta[sl] += rij[sl] @ Lam.T
    return alpha, Delta


def param_grad_exact(orc: Oracle, itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS, h=0.5, nthreads=0):
    """d/dtheta of sum gE.E + sum gF.F + sum gS.S, theta = (moment, species, radial) coefficients.

    J(theta) = sum_i alpha_i E_i + sum_ik g_ik . Delta_ik has the same parameter gradient.  Along any single
    coordinate J is a polynomial of degree <= 4 (every basis function is a product of at most four basic
    moments, each linear in the coefficients), so the 5-point central stencil is exact up to rounding for any
    step h; tests check the degree bound itself.
    """
    alpha, Delta = _alpha_delta(itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS)
    k = orc._keep

    def J():
        Ei, g = orc.compact(itypes, rij, jtypes, True, nthreads)
        return float(alpha @ Ei + np.sum(g * Delta))

    grads = []
    for arr in (k["moment"], k["species"], k["radial"].reshape(-1)):
        for q in range(arr.size):
            x0 = arr[q]
            vals = []
            for t in (-2, -1, 1, 2):
                arr[q] = x0 + t * h
                vals.append(J())
            arr[q] = x0
            grads.append((vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h))
    return np.array(grads)


def J_along(orc: Oracle, itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS, which, q, ts, nthreads=0):
    """J at theta + t e_q for each t (degree checks)."""
    alpha, Delta = _alpha_delta(itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS)
    arr = {"moment": orc._keep["moment"], "species": orc._keep["species"],
           "radial": orc._keep["radial"].reshape(-1)}[which]
    x0 = arr[q]
    out = []
    for t in ts:
        arr[q] = x0 + t
        Ei, g = orc.compact(itypes, rij, jtypes, True, nthreads)
        out.append(float(alpha @ Ei + np.sum(g * Delta)))
    arr[q] = x0
    return np.array(out)


def J_direction(orc: Oracle, itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS, v, ts, nthreads=0):
    """J at theta + t v for each t, v = (moment, species, radial) direction in the order of param_grad_exact."""
    alpha, Delta = _alpha_delta(itypes, js, rij, jtypes, conf_off, volumes, gE, gF, gS)
    k = orc._keep
    arrs = (k["moment"], k["species"], k["radial"].reshape(-1))
    v = np.asarray(v, dtype=np.float64)
    parts, o = [], 0
    for a in arrs:
        parts.append(v[o:o + a.size])
        o += a.size
    if o != v.size:
        raise ValueError("direction has the wrong length")
    x0 = [a.copy() for a in arrs]
    out = []
    for t in ts:
        for a, a0, d in zip(arrs, x0, parts):
            a[:] = a0 + t * d
        Ei, g = orc.compact(itypes, rij, jtypes, True, nthreads)
        out.append(float(alpha @ Ei + np.sum(g * Delta)))
    for a, a0 in zip(arrs, x0):
        a[:] = a0
    return np.array(out)


# ---------------------------------------------------------------------------
# Neighbor packing restatement (pair_jax_mtp.cpp, zero_overhead.cpp)
# ---------------------------------------------------------------------------
def pack_full_list(x, type_, tag, ilist, neigh, nlocal, cutoff, prd, periodic, max_atoms, max_nbr):
    """Returns (pos[max_atoms,3K], nbr[max_atoms,K], ntype[max_atoms,K], counts, atom_types, nneigh_total)."""
    x = _f64(x); type_ = _i32(type_); tag = _i32(tag); ilist = _i32(ilist)
    numneigh = np.array([len(n) for n in neigh], np.int32)
    rows = [_i32(n) for n in neigh]
    ptrs = (ctypes.c_void_p * len(rows))(*[r.ctypes.data if len(r) else 0 for r in rows])
    pos = np.zeros((max_atoms, 3 * max_nbr)); nbr = np.zeros((max_atoms, max_nbr), np.int32)
    ntype = np.zeros((max_atoms, max_nbr), np.int32); counts = np.zeros(max_atoms, np.int32)
    atypes = np.zeros(max_atoms, np.int32)
    prd = _f64(prd); per = _i32(periodic)
    tot = lib().orc_pack(nlocal, len(type_), _p(x), _p(type_), _p(tag), len(ilist), _p(ilist), _p(numneigh),
                         ctypes.cast(ptrs, ctypes.c_void_p), cutoff, _p(prd), _p(per), max_atoms, max_nbr, _p(pos),
                         _p(nbr), _p(ntype), _p(counts), _p(atypes))
    return pos, nbr, ntype, counts, atypes, int(tot)


def stage(max_atoms, max_nbr, natoms, pos, atom_types, nbr, counts, ntype):
    pos = _f64(pos); atom_types = _i32(atom_types); nbr = _i32(nbr); counts = _i32(counts); ntype = _i32(ntype)
    it = np.zeros(max_atoms, np.int32); js = np.zeros((max_atoms, max_nbr), np.int32)
    rij = np.zeros((max_atoms, max_nbr, 3)); jt = np.zeros((max_atoms, max_nbr), np.int32)
    lib().orc_stage(max_atoms, max_nbr, natoms, _p(pos), _p(atom_types), _p(nbr), _p(counts), _p(ntype), _p(it),
                    _p(js), _p(rij), _p(jt))
    return it, js, rij, jt
