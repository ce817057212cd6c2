"""Training path: packed batches of uneven configurations, the weighted
energy/force/stress loss and its parameter gradient on the GPU.

Reference semantics (motep_jax_train_function.py:67-100, 154-191):
    L = w_E sum_c (E_c - E_c^ref)^2 + w_F sum_{c,a} |F_a - F_a^ref|^2
        + w_S sum_c |sigma_c - sigma_c^ref|^2,   (w_E, w_F, w_S) = (1, 0.01, 0.001)
summed over configurations without per-atom normalisation, differentiated
with respect to {moment_coeffs, species_coeffs, radial_coeffs} (scaling is
not trained).  The reference pads every configuration to a global
(A_max, K_max) (motep_jax_train_import.py:219-295); here configurations are
concatenated and described by a segment-offset table, so one launch covers
the whole batch without padded atoms.

Gradients come from a hand-written second backward (the C ABI's
``mtp_batch_param_grad``): with the upstream gradients (gE per
configuration, gF per atom, gS per configuration) it evaluates
    J = sum_i [ gE_c(i) E_i + sum_k g_ik . Delta_ik ],
    Delta_ik = gF_i - gF_{j_ik} + Lambda_c r_ik,  Lambda = (M + M^T) / (2 V_c)
(M holds gS in the reference's Voigt order), whose parameter gradient equals
that of the loss: a forward tangent of the moments along Delta followed by a
reverse pass over moments and tangents.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi
from .neighbors import padded_neighbors

LOSS_WEIGHTS = (1.0, 0.01, 0.001)


@dataclass
class PackedBatch:
    conf_off: np.ndarray   # [C+1] atom offsets
    itypes: np.ndarray     # [N]
    js: np.ndarray         # [N, K] global row index; N marks an empty slot
    rij: np.ndarray        # [N, K, 3]
    jtypes: np.ndarray     # [N, K]
    volumes: np.ndarray    # [C]
    conf_of: np.ndarray    # [N]

    @property
    def nconf(self) -> int:
        return len(self.volumes)

    @property
    def natoms(self) -> int:
        return len(self.itypes)

    @property
    def nslots(self) -> int:
        return self.js.shape[1]

    def to_device(self):
        import torch

        dev = torch.device("cuda")
        return {
            "conf_off": torch.as_tensor(self.conf_off, dtype=torch.int32, device=dev),
            "conf_of": torch.as_tensor(self.conf_of, dtype=torch.int32, device=dev),
            "itypes": torch.as_tensor(self.itypes, dtype=torch.int32, device=dev),
            "js": torch.as_tensor(self.js, dtype=torch.int32, device=dev),
            "rij": torch.as_tensor(self.rij, dtype=torch.float64, device=dev),
            "jtypes": torch.as_tensor(self.jtypes, dtype=torch.int32, device=dev),
            "volumes": torch.as_tensor(self.volumes, dtype=torch.float64, device=dev),
        }


def pack_configs(configs, cutoff: float, max_nbr: int) -> PackedBatch:
    """configs: iterable of (positions[n,3], types[n], cell[3]) (orthogonal periodic cells)."""
    rows_it, rows_js, rows_r, rows_jt, vols, offs = [], [], [], [], [], [0]
    for pos, types, cell in configs:
        nb = padded_neighbors(pos, types, cell, cutoff, max_nbr)
        n = len(pos)
        rows_it.append(nb["itypes"]); rows_r.append(nb["rij"]); rows_jt.append(nb["jtypes"])
        rows_js.append(nb["js"].astype(np.int64) + offs[-1])  # local sentinel n -> fixed below
        vols.append(float(np.prod(cell)))
        offs.append(offs[-1] + n)
    N = offs[-1]
    js = np.concatenate(rows_js)
    jt = np.concatenate(rows_jt)
    js[jt < 0] = N
    conf_off = np.array(offs, dtype=np.int32)
    conf_of = np.repeat(np.arange(len(vols), dtype=np.int32), np.diff(conf_off))
    return PackedBatch(conf_off, np.concatenate(rows_it).astype(np.int32), js.astype(np.int32),
                       np.concatenate(rows_r), jt.astype(np.int32), np.array(vols), conf_of)


def random_cells(nconf: int, rng, natoms_range=(8, 256), density_range=(0.06, 0.09), species: int = 2,
                 kinds=("fcc", "bcc", "random"), min_sep: float = 1.9):
    """Seeded FCC / BCC / random-dense periodic cells of uneven atom counts."""
    out = []
    for c in range(nconf):
        kind = kinds[c % len(kinds)]
        n_target = int(rng.integers(natoms_range[0], natoms_range[1] + 1))
        rho = float(rng.uniform(*density_range))
        if kind in ("fcc", "bcc"):
            per = 4 if kind == "fcc" else 2
            m = max(1, int(round((n_target / per) ** (1.0 / 3.0))))
            n = per * m ** 3
            a = (per / rho) ** (1.0 / 3.0)
            base = (np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) if kind == "fcc"
                    else np.array([[0, 0, 0], [0.5, 0.5, 0.5]]))
            g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)
            pos = (g[:, None] + base[None]).reshape(-1, 3) * a
            pos = pos + rng.normal(0, 0.05 * a, pos.shape)
            cell = np.full(3, m * a)
        else:
            n = n_target
            L = (n / rho) ** (1.0 / 3.0)
            cell = np.full(3, L)
            pts = []
            tries = 0
            while len(pts) < n and tries < 200000:
                tries += 1
                p = rng.uniform(0, L, 3)
                if pts:
                    d = np.asarray(pts) - p
                    d -= L * np.round(d / L)
                    if np.min(np.einsum("ij,ij->i", d, d)) < min_sep * min_sep:
                        continue
                pts.append(p)
            pos = np.array(pts)
            n = len(pos)
        types = rng.integers(0, species, n).astype(np.int32)
        out.append((pos, types, cell))
    return out


# ---------------------------------------------------------------------------
# GPU calls
# ---------------------------------------------------------------------------
def batch_forward(pot, dbatch, stream=None):
    """Per-configuration energies [C], forces [N,3] and stress [C,6] (reference Voigt order)."""
    import torch

    L = capi.lib()
    N, K = dbatch["js"].shape
    C = dbatch["volumes"].shape[0]
    E = torch.empty(C, dtype=torch.float64, device="cuda")
    F = torch.empty((N, 3), dtype=torch.float64, device="cuda")
    S = torch.empty((C, 6), dtype=torch.float64, device="cuda")
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    capi.check(L.mtp_batch_forward(pot.ctx, N, K, C, dbatch["conf_off"].data_ptr(), dbatch["conf_of"].data_ptr(),
                                   dbatch["itypes"].data_ptr(), dbatch["js"].data_ptr(), dbatch["rij"].data_ptr(),
                                   dbatch["jtypes"].data_ptr(), dbatch["volumes"].data_ptr(), E.data_ptr(),
                                   F.data_ptr(), S.data_ptr(), s), pot.ctx)
    return E, F, S


def batch_param_grad(pot, dbatch, gE, gF, gS, stream=None):
    """d/dtheta of sum gE.E + sum gF.F + sum gS.S, theta = (moment, species, radial) coefficients."""
    import torch

    L = capi.lib()
    N, K = dbatch["js"].shape
    C = dbatch["volumes"].shape[0]
    p = pot.params
    n_theta = len(p.moment_coeffs) + p.species_count + int(np.prod(p.radial_coeffs.shape))
    grad = torch.empty(n_theta, dtype=torch.float64, device="cuda")
    gE = gE.contiguous().to(torch.float64)
    gF = gF.contiguous().to(torch.float64)
    gS = gS.contiguous().to(torch.float64)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    capi.check(L.mtp_batch_param_grad(pot.ctx, N, K, C, dbatch["conf_off"].data_ptr(), dbatch["conf_of"].data_ptr(),
                                      dbatch["itypes"].data_ptr(), dbatch["js"].data_ptr(), dbatch["rij"].data_ptr(),
                                      dbatch["jtypes"].data_ptr(), dbatch["volumes"].data_ptr(), gE.data_ptr(),
                                      gF.data_ptr(), gS.data_ptr(), grad.data_ptr(), s), pot.ctx)
    return grad


def theta_of(params) -> np.ndarray:
    return np.concatenate([np.asarray(params.moment_coeffs, float), np.asarray(params.species_coeffs, float),
                           np.asarray(params.radial_coeffs, float).ravel()])


def _split_theta(params, theta):
    S = len(params.moment_coeffs)
    sp = params.species_count
    return theta[:S], theta[S:S + sp], theta[S + sp:]


def make_function():
    import torch

    class MTPBatchFunction(torch.autograd.Function):
        """(theta) -> (E[C], F[N,3], S[C,6]) with the hand-written parameter backward."""

        @staticmethod
        def forward(ctx, theta, pot, dbatch):
            th = theta.detach().cpu().numpy()
            m, s, r = _split_theta(pot.params, th)
            pot.set_params(m, s, r)
            E, F, S = batch_forward(pot, dbatch)
            ctx.pot = pot
            ctx.dbatch = dbatch
            return E, F, S

        @staticmethod
        def backward(ctx, gE, gF, gS):
            import torch as _t

            C = ctx.dbatch["volumes"].shape[0]
            N = ctx.dbatch["js"].shape[0]
            if gE is None:
                gE = _t.zeros(C, dtype=_t.float64, device="cuda")
            if gF is None:
                gF = _t.zeros((N, 3), dtype=_t.float64, device="cuda")
            if gS is None:
                gS = _t.zeros((C, 6), dtype=_t.float64, device="cuda")
            return batch_param_grad(ctx.pot, ctx.dbatch, gE, gF, gS), None, None

    return MTPBatchFunction


def loss_fn(E, F, S, E_ref, F_ref, S_ref, weights=LOSS_WEIGHTS):
    """motep_jax_train_function.py:67-74."""
    wE, wF, wS = weights
    return wE * ((E - E_ref) ** 2).sum() + wF * ((F - F_ref) ** 2).sum() + wS * ((S - S_ref) ** 2).sum()


def lls_init(pot, dbatch, E_ref, F_ref, S_ref, weights=LOSS_WEIGHTS):
    """Weighted linear least squares for moment_coeffs at fixed radial and species coefficients.

    Mirrors solve_lls_for_basis (motep_jax_train_import.py:605-683): every output (E, F, stress) is linear in
    the moment coefficients, so the design matrix column s is the batch output with coefficient e_s minus the
    output with all coefficients zero (computed here by the GPU forward kernels), and the weights are sqrt(w).
    Returns the new coefficients (also loaded into ``pot``).
    """
    import torch

    p = pot.params
    S = len(p.moment_coeffs)
    wE, wF, wS = (float(np.sqrt(w)) for w in weights)

    def flat(E, F, St):
        return torch.cat([wE * E.reshape(-1), wF * F.reshape(-1), wS * St.reshape(-1)])

    pot.set_params(moment_coeffs=np.zeros(S))
    off = flat(*batch_forward(pot, dbatch))
    cols = []
    for s in range(S):
        e = np.zeros(S)
        e[s] = 1.0
        pot.set_params(moment_coeffs=e)
        cols.append(flat(*batch_forward(pot, dbatch)) - off)
    X = torch.stack(cols, dim=1)
    y = flat(torch.as_tensor(E_ref, device="cuda"), torch.as_tensor(F_ref, device="cuda"),
             torch.as_tensor(S_ref, device="cuda")) - off
    sol = torch.linalg.lstsq(X.cpu(), y.cpu().unsqueeze(1)).solution.squeeze(1).numpy()
    pot.set_params(moment_coeffs=sol)
    return sol


class GraphedLossGrad:
    """One training evaluation theta -> (loss, gradient) replayed as a HIP graph.

    Captured once: device parameter update, batch forward, the loss's
    upstream gradients (elementwise), the hand-written parameter backward.
    The loss value itself (three reductions) runs eagerly after the replay:
    multi-block reductions of the framework are kept out of the graph.
    Nothing in the captured region allocates, synchronises or memsets.
    """

    def __init__(self, pot, dbatch, E_ref, F_ref, S_ref, weights=LOSS_WEIGHTS):
        import torch

        self.pot, self.dbatch, self.weights = pot, dbatch, weights
        self.E_ref, self.F_ref, self.S_ref = E_ref, F_ref, S_ref
        self.theta = torch.tensor(theta_of(pot.params), dtype=torch.float64, device="cuda")
        self._body()  # warm-up: allocates the library's training buffers outside the capture
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self._body()

    def _body(self):
        wE, wF, wS = self.weights
        self.pot.set_params_device(self.theta)
        E, F, S = batch_forward(self.pot, self.dbatch)
        gE = (2.0 * wE) * (E - self.E_ref)
        gF = (2.0 * wF) * (F - self.F_ref)
        gS = (2.0 * wS) * (S - self.S_ref)
        grad = batch_param_grad(self.pot, self.dbatch, gE, gF, gS)
        return E, F, S, grad

    def __call__(self, theta):
        self.theta.copy_(theta)
        self.graph.replay()
        E, F, S, grad = self.out
        return loss_fn(E, F, S, self.E_ref, self.F_ref, self.S_ref, self.weights), grad


class LBFGS:
    """Limited-memory BFGS with Armijo backtracking on a flat parameter tensor.

    The MTP loss is a high-degree polynomial in the radial coefficients with gradients of 1e6-1e8 at random
    initial parameters; a backtracking line search that only asks for sufficient decrease is robust there
    (torch.optim.LBFGS's strong-Wolfe search can return a zero step on such curvature).
    """

    def __init__(self, fun, history: int = 20, c1: float = 1e-4, max_backtrack: int = 60):
        self.fun = fun          # theta -> (loss float, grad tensor)
        self.m = history
        self.c1 = c1
        self.max_backtrack = max_backtrack
        self.S, self.Y = [], []

    def _direction(self, g):
        q = g.clone()
        alphas = []
        for s, y in zip(reversed(self.S), reversed(self.Y)):
            rho = 1.0 / float(y @ s)
            a = rho * float(s @ q)
            q -= a * y
            alphas.append((rho, a))
        if self.S:
            s, y = self.S[-1], self.Y[-1]
            q *= float(s @ y) / float(y @ y)
        else:
            q /= max(1.0, float(g.abs().max()))
        for (s, y), (rho, a) in zip(zip(self.S, self.Y), reversed(alphas)):
            b = rho * float(y @ q)
            q += (a - b) * s
        return -q

    def step(self, theta):
        f, g = self.fun(theta)
        d = self._direction(g)
        gd = float(g @ d)
        if gd >= 0:  # not a descent direction: restart from steepest descent
            self.S, self.Y = [], []
            d = -g / max(1.0, float(g.abs().max()))
            gd = float(g @ d)
        t = 1.0
        for _ in range(self.max_backtrack):
            f_new, g_new = self.fun(theta + t * d)
            if f_new <= f + self.c1 * t * gd:
                s = t * d
                y = g_new - g
                if float(s @ y) > 1e-30:
                    self.S.append(s)
                    self.Y.append(y)
                    if len(self.S) > self.m:
                        self.S.pop(0)
                        self.Y.pop(0)
                return theta + s, f_new
            t *= 0.5
        return theta, f


class Trainer:
    """Full-batch training: LLS start, then L-BFGS (Armijo backtracking) over the hand-written gradient.

    Replaces the reference's LLS + optax L-BFGS loop (motep_jax_train_function.py:160-213).  Each loss/gradient
    evaluation is the batch forward plus the hand-written parameter backward, through torch.autograd.
    """

    def __init__(self, pot, batch: PackedBatch, E_ref, F_ref, S_ref, weights=LOSS_WEIGHTS, lr=1.0,
                 lls: bool = True, graph: bool = False, autotune: bool = True):
        import torch

        self.pot = pot
        self.dbatch = batch.to_device()
        if autotune:  # kernel variant for this potential on this batch
            d = self.dbatch
            pot.autotune(d["itypes"], d["js"], d["rij"], d["jtypes"], reps=2)
        self.E_ref = torch.as_tensor(E_ref, dtype=torch.float64, device="cuda")
        self.F_ref = torch.as_tensor(F_ref, dtype=torch.float64, device="cuda")
        self.S_ref = torch.as_tensor(S_ref, dtype=torch.float64, device="cuda")
        self.weights = weights
        self.fn = make_function()
        if lls:  # motep_jax_train_function.py:160-172: LLS for the moment coefficients first
            lls_init(pot, self.dbatch, self.E_ref, self.F_ref, self.S_ref, weights)
        self.theta = torch.tensor(theta_of(pot.params), dtype=torch.float64, device="cuda", requires_grad=True)
        self.graphed = GraphedLossGrad(pot, self.dbatch, self.E_ref, self.F_ref, self.S_ref, weights) if graph else None
        self.opt = LBFGS(self.value_and_grad)

    def loss(self, theta=None):
        th = self.theta if theta is None else theta
        E, F, S = self.fn.apply(th, self.pot, self.dbatch)
        return loss_fn(E, F, S, self.E_ref, self.F_ref, self.S_ref, self.weights)

    def value_and_grad(self, theta):
        import torch

        if self.graphed is not None:
            L, g = self.graphed(theta.detach())
            return float(L), g.clone()
        th = theta.detach().clone().requires_grad_(True)
        L = self.loss(th)
        (g,) = torch.autograd.grad(L, th)
        return float(L.detach()), g.detach()

    def step(self) -> float:
        """One L-BFGS iteration; returns the loss after it."""
        import torch

        new, f = self.opt.step(self.theta.detach())
        with torch.no_grad():
            self.theta.copy_(new)
        return f


# ---------------------------------------------------------------------------
# Training from MLIP .cfg data (motep_jax_train_function.py:17-229)
def batch_from_cfg(configs, cutoff: float, max_nbr: int):
    """A packed batch and its labels from .cfg configurations (orthogonal periodic cells): energies [C],
    forces [N, 3] and stresses [C, 6] = -PlusStress / V in the reference's Voigt order."""
    cells = [(c.positions, c.types, c.orthogonal_box()) for c in configs]
    batch = pack_configs(cells, cutoff, max_nbr)
    E = np.array([c.energy for c in configs], dtype=np.float64)
    F = np.concatenate([c.forces for c in configs]).astype(np.float64)
    S = np.stack([c.stress for c in configs]).astype(np.float64)
    return batch, E, F, S


def train_cfg(cfg_path, level: int = 12, steps: int = 50, species_count: int = 1, min_dist: float = 0.5,
              max_dist: float = 5.0, scaling: float = 1.0, radial_basis_size: int | None = None, seed: int = 10,
              max_nbr: int = 64, split: float = 0.75, log=None):
    """The reference's training recipe on a .cfg file (motep_jax_train_function.py:17-50, 154-213): the untrained
    level's alpha tables, species_count / min_dist / max_dist / scaling set, parameters drawn uniformly from a
    seeded generator (data.py:41-58), the first `split` of the configurations as the training set, linear least
    squares for the moment coefficients, then L-BFGS on the weighted E/F/S loss with the hand-written gradient.

    Returns (trained MTPData, dict of losses and per-set RMSEs).  The reference takes its validation set as
    images[1 - n_train:] (motep_jax_train_function.py:25), which overlaps the training set; here the held-out
    configurations are images[n_train:].
    """
    import torch

    from .cfg_io import read_cfg
    from .mtp_io import untrained
    from .potential import MTPPotential

    configs = read_cfg(cfg_path)
    n_train = int(split * len(configs))
    p = untrained(level)
    p.species_count = species_count
    p.min_dist, p.max_dist, p.scaling = min_dist, max_dist, scaling
    if radial_basis_size:
        p.radial_basis_size = radial_basis_size
    p.initialize(np.random.default_rng(seed))
    batch, E, F, S = batch_from_cfg(configs[:n_train], max_dist, max_nbr)
    pot = MTPPotential(p, max_atoms=max(batch.natoms, 1), max_nbr=max_nbr)
    tr = Trainer(pot, batch, E, F, S, lls=True, graph=False, autotune=False)
    history = [float(tr.loss().detach())]
    for _ in range(steps):
        history.append(tr.step())
    theta = tr.theta.detach().cpu().numpy()
    S_n, sp = len(p.moment_coeffs), p.species_count
    pot.set_params(moment_coeffs=theta[:S_n], species_coeffs=theta[S_n:S_n + sp],
                   radial_coeffs=theta[S_n + sp:].reshape(np.asarray(pot.params.radial_coeffs).shape))
    trained = pot.params.copy()
    pot.close()

    def errors(cfgs):
        if not cfgs:
            return {}
        b, Eg, Fg, Sg = batch_from_cfg(cfgs, max_dist, max_nbr)
        q = MTPPotential(trained, max_atoms=b.natoms, max_nbr=max_nbr)
        Ep, Fp, Sp = (t.cpu().numpy() for t in batch_forward(q, b.to_device()))
        torch.cuda.synchronize()
        q.close()
        n = np.diff(b.conf_off)
        return {"configs": len(cfgs), "energy_rmse_meV_per_atom": float(np.sqrt(np.mean(((Ep - Eg) / n) ** 2))) * 1e3,
                "force_rmse_meV_per_A": float(np.sqrt(np.mean((Fp - Fg) ** 2))) * 1e3,
                "stress_rmse": float(np.sqrt(np.mean((Sp - Sg) ** 2)))}

    res = {"level": level, "loss_initial_after_lls": history[0], "loss_final": history[-1], "lbfgs_steps": steps,
           "train": errors(configs[:n_train]), "held_out": errors(configs[n_train:])}
    if log:
        log(f"[train_cfg] {res}")
    return trained, res
