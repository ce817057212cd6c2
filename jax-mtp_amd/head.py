"""Distilled head: an MLP over a low-level MTP's basis values that stands in
for a high-level MTP's contraction tree, with the high-level MTP as the
exact fallback and judge.

The reference has no learned head; this module is the host side of the
C ABI's ``mtp_head_*`` functions (network layout in include/mtp_hip.h):

* ``HeadModel``       parameters, input normalisation, per-species training
                      box, energy shift/scale; saved as a plain ``.npz``
                      (no pickled objects) with the descriptor ``.mtp`` text.
* ``DistilledHead``   inference: ``calc_energy_forces_stress_padded`` with the
                      same arguments and outputs as ``MTPPotential`` (which
                      mirrors jax_backend.py:374-460), plus the count of rows
                      re-evaluated exactly.
* ``HeadTrainer``     force-and-energy matching of the network to a
                      teacher's site energies and forces on the GPU: the
                      network's forward/backward and its parameter gradient
                      are hand-written kernels behind a torch.autograd
                      Function; basis Jacobians are precomputed once per
                      training set (the positions do not change).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi
from .mtp_io import MTPData, format_mtp, parse_mtp_text
from .potential import MTPPotential, _f64, _i32

HIDDEN = 64
N_IN = 64
N_THETA = N_IN * HIDDEN + HIDDEN + HIDDEN * HIDDEN + 2 * HIDDEN + 1


def init_theta(S: int, Sp: int, rng) -> np.ndarray:
    """Scaled-normal initial weights; rows of W1 past the S + Sp inputs are zero."""
    W1 = rng.normal(0.0, 1.0 / np.sqrt(S + Sp), (N_IN, HIDDEN))
    W1[S + Sp:] = 0.0
    W2 = rng.normal(0.0, 1.0 / np.sqrt(HIDDEN), (HIDDEN, HIDDEN))
    w3 = rng.normal(0.0, 1.0 / np.sqrt(HIDDEN), HIDDEN)
    return np.concatenate([W1.ravel(), np.zeros(HIDDEN), W2.ravel(), np.zeros(HIDDEN), w3, [0.0]])


@dataclass
class HeadModel:
    desc: MTPData               # low-level potential whose basis values are the inputs
    theta: np.ndarray           # [N_THETA]
    mean: np.ndarray            # [S]
    inv_std: np.ndarray         # [S]
    lo: np.ndarray | None       # [Sp, S] training box of the normalised inputs
    hi: np.ndarray | None
    e_shift: np.ndarray         # [Sp]
    e_scale: float
    residual: bool = False      # add the descriptor potential's own (linear) energy

    def save(self, path) -> None:
        np.savez(path, theta=self.theta, mean=self.mean, inv_std=self.inv_std, residual=np.array(int(self.residual)),
                 lo=self.lo if self.lo is not None else np.zeros(0), hi=self.hi if self.hi is not None else np.zeros(0),
                 e_shift=self.e_shift, e_scale=np.array(self.e_scale), desc=np.array(format_mtp(self.desc)),
                 desc_radial=np.asarray(self.desc.radial_coeffs, dtype=np.float64),
                 desc_species=np.asarray(self.desc.species_coeffs, dtype=np.float64),
                 desc_moment=np.asarray(self.desc.moment_coeffs, dtype=np.float64))

    @classmethod
    def load(cls, path) -> "HeadModel":
        d = np.load(path)  # allow_pickle=False
        lo, hi = d["lo"], d["hi"]
        desc = parse_mtp_text(str(d["desc"]))
        # the .mtp text keeps 16 significant digits; the arrays are exact
        desc.radial_coeffs = d["desc_radial"].copy()
        desc.species_coeffs = d["desc_species"].copy()
        desc.moment_coeffs = d["desc_moment"].copy()
        return cls(desc, d["theta"].copy(), d["mean"].copy(), d["inv_std"].copy(),
                   lo.copy() if lo.size else None, hi.copy() if hi.size else None, d["e_shift"].copy(),
                   float(d["e_scale"]), bool(int(d["residual"])))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class DistilledHead:
    """The head on one GPU: descriptor potential + network + exact fallback."""

    def __init__(self, model: HeadModel, exact: MTPPotential | None, max_atoms: int, max_nbr: int,
                 desc_pot: MTPPotential | None = None):
        self.model = model
        self.desc = desc_pot if desc_pot is not None else MTPPotential(model.desc, max_atoms, max_nbr)
        self.exact = exact
        self._own_desc = desc_pot is None
        arrs = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None
                for x in (model.theta, model.mean, model.inv_std, model.lo, model.hi, model.e_shift)]
        self._keep = arrs
        L = capi.lib()
        if L.mtp_head_theta_size() != N_THETA:
            raise RuntimeError("head parameter layout mismatch")
        h = ctypes.c_void_p()
        capi.check(L.mtp_head_open(self.desc.ctx, exact.ctx if exact is not None else None, *[_ptr(a) for a in arrs],
                                   float(model.e_scale), int(model.residual), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            capi.lib().mtp_head_close(self.h)
            self.h = None
        if getattr(self, "_own_desc", False) and self.desc is not None:
            self.desc.close()
            self.desc = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_theta(self, theta) -> None:
        """From numpy (synchronous) or a device tensor (stream-ordered)."""
        import torch

        L = capi.lib()
        if isinstance(theta, torch.Tensor) and theta.is_cuda:
            t = theta.detach().contiguous()
            capi.check(L.mtp_head_set_theta(self.h, t.data_ptr(), 1, torch.cuda.current_stream().cuda_stream),
                       head=self.h)
            self._theta_dev = t  # keep alive until the copy has run
        else:
            a = np.ascontiguousarray(theta, dtype=np.float64)
            capi.check(L.mtp_head_set_theta(self.h, _ptr(a), 0, None), head=self.h)
            self.model.theta = a.copy()

    def calc_energy_forces_stress_padded(self, itypes, all_js, all_rijs, all_jtypes, cell_rank, volume,
                                         natoms_energy, natoms_force, return_all: bool = False, sync: bool = True):
        """Same contract as MTPPotential.calc_energy_forces_stress_padded.  With
        ``return_all`` the dict also holds ``flags`` [A] and ``n_fallback``."""
        import torch

        dev = torch.device("cuda")
        it, js, rij, jt = _i32(itypes, dev), _i32(all_js, dev), _f64(all_rijs, dev), _i32(all_jtypes, dev)
        A, K = js.shape
        if rij.shape != (A, K, 3) or jt.shape != (A, K) or it.shape != (A,):
            raise ValueError("inconsistent padded shapes")
        out = torch.empty(14, dtype=torch.float64, device=dev)
        F = torch.empty((A, 3), dtype=torch.float64, device=dev)
        eatom = torch.empty(A, dtype=torch.float64, device=dev) if return_all else None
        flags = torch.empty(A, dtype=torch.int32, device=dev) if return_all else None
        nfb = torch.zeros(1, dtype=torch.int32, device=dev)
        st = torch.cuda.current_stream().cuda_stream
        capi.check(capi.lib().mtp_head_compute_padded_device(
            self.h, A, K, it.data_ptr(), js.data_ptr(), rij.data_ptr(), jt.data_ptr(), int(cell_rank), float(volume),
            int(natoms_energy), int(natoms_force), out.data_ptr(), F.data_ptr(),
            eatom.data_ptr() if eatom is not None else None, flags.data_ptr() if flags is not None else None,
            nfb.data_ptr(), st), head=self.h)
        if sync:
            torch.cuda.current_stream().synchronize()
        E, stress = out[0], out[1:7]
        if not return_all:
            return E, F, stress
        return E, F, stress, dict(virial6=out[7:13], eatom=eatom, status=out[13], flags=flags,
                                  n_fallback=int(nfb.item()))

    # network alone (device tensors)
    def eval_network(self, itypes, basis):
        import torch

        it, b = itypes, basis
        A, S = b.shape
        E = torch.empty(A, dtype=torch.float64, device=b.device)
        xi = torch.empty((A, S), dtype=torch.float64, device=b.device)
        capi.check(capi.lib().mtp_head_eval_device(self.h, A, it.data_ptr(), b.data_ptr(), E.data_ptr(), xi.data_ptr(),
                                                   None, torch.cuda.current_stream().cuda_stream), head=self.h)
        return E, xi

    def param_grad(self, itypes, basis, alpha, v):
        import torch

        A = basis.shape[0]
        g = torch.empty(N_THETA, dtype=torch.float64, device=basis.device)
        a, vv = alpha.contiguous(), v.contiguous()
        capi.check(capi.lib().mtp_head_param_grad_device(self.h, A, itypes.data_ptr(), basis.data_ptr(), a.data_ptr(),
                                                         vv.data_ptr(), g.data_ptr(),
                                                         torch.cuda.current_stream().cuda_stream), head=self.h)
        return g


def _function():
    import torch

    class HeadForward(torch.autograd.Function):
        """theta -> (site energies E [N], forces F [N, 3]) of a fixed training set."""

        @staticmethod
        def forward(ctx, theta, tr):
            tr.head.set_theta(theta)
            E, xi = tr.head.eval_network(tr.itypes, tr.basis)
            g = torch.empty((tr.N, tr.K, 3), dtype=torch.float64, device=theta.device)
            st = torch.cuda.current_stream().cuda_stream
            L = capi.lib()
            capi.check(L.mtp_jacobian_apply_device(tr.N, tr.K, tr.S, tr.J.data_ptr(), xi.data_ptr(), g.data_ptr(), st))
            F = tr.desc.gather_forces(tr.js, g, tr.N)
            ctx.tr = tr
            return E, F

        @staticmethod
        def backward(ctx, gE, gF):
            tr = ctx.tr
            gE = torch.zeros(tr.N, dtype=torch.float64, device=tr.basis.device) if gE is None else gE.contiguous()
            gF = torch.zeros((tr.N, 3), dtype=torch.float64, device=tr.basis.device) if gF is None else gF.contiguous()
            v = torch.empty((tr.N, tr.S), dtype=torch.float64, device=tr.basis.device)
            capi.check(capi.lib().mtp_jacobian_tapply_device(tr.N, tr.K, tr.S, tr.N, tr.J.data_ptr(), tr.js.data_ptr(),
                                                             gF.data_ptr(), v.data_ptr(),
                                                             torch.cuda.current_stream().cuda_stream))
            return tr.head.param_grad(tr.itypes, tr.basis, gE, v), None

    return HeadForward


class HeadTrainer:
    """Fits the head's network to a teacher's site energies E_ref [N] and
    forces F_ref [N, 3] on a packed training set (training.PackedBatch)."""

    def __init__(self, desc_pot: MTPPotential, batch, E_ref, F_ref, weights=(1.0, 1.0), margin: float = 0.0,
                 seed: int = 0, residual: bool = True):
        import torch

        dev = torch.device("cuda")
        self.desc = desc_pot
        self.N, self.K = batch.natoms, batch.nslots
        self.S = len(desc_pot.params.moment_coeffs)
        Sp = desc_pot.params.species_count
        self.itypes = torch.as_tensor(batch.itypes, dtype=torch.int32, device=dev)
        self.js = torch.as_tensor(batch.js, dtype=torch.int32, device=dev)
        rij = torch.as_tensor(batch.rij, dtype=torch.float64, device=dev)
        jt = torch.as_tensor(batch.jtypes, dtype=torch.int32, device=dev)
        self.basis = desc_pot.descriptor(self.itypes, rij, jt)
        self.J = desc_pot.basis_jacobian(self.itypes, rij, jt)
        self.E_ref = _f64(E_ref, dev)
        self.F_ref = _f64(F_ref, dev).reshape(self.N, 3)
        self.weights = weights
        b = self.basis.cpu().numpy()
        t = batch.itypes
        mean = b.mean(0)
        std = b.std(0)
        inv_std = np.where(std > 1e-300, 1.0 / np.maximum(std, 1e-300), 0.0)
        z = (b - mean) * inv_std
        lo = np.zeros((Sp, self.S))
        hi = np.zeros((Sp, self.S))
        Er = self.E_ref.cpu().numpy()
        e_shift = np.zeros(Sp)
        for s in range(Sp):
            m = t == s
            if m.any():
                lo[s], hi[s] = z[m].min(0), z[m].max(0)
        span = hi - lo
        lo -= margin * span
        hi += margin * span
        ts = np.where(t < 0, t + Sp, t).clip(0, Sp - 1)
        if residual:  # the network fits what the descriptor potential's linear energy leaves
            lin = np.asarray(desc_pot.params.moment_coeffs, dtype=np.float64)
            Er = Er - (np.asarray(desc_pot.params.species_coeffs, dtype=np.float64)[ts] + b @ lin)
        for s in range(Sp):
            m = ts == s
            e_shift[s] = Er[m].mean() if m.any() else 0.0
        resid = Er - e_shift[ts]
        e_scale = float(resid.std()) if resid.std() > 0 else 1.0
        rng = np.random.default_rng(seed)
        theta = init_theta(self.S, Sp, rng)
        if residual:  # start at the linear potential: small output layer
            theta[-HIDDEN - 1:-1] *= 0.1
        self.model = HeadModel(desc_pot.params, theta, mean, inv_std, lo, hi, e_shift, e_scale, residual)
        self.head = DistilledHead(self.model, None, self.N, self.K, desc_pot=desc_pot)
        self.theta = torch.tensor(self.model.theta, dtype=torch.float64, device=dev, requires_grad=True)
        self._fn = _function()

    def predict(self, theta=None):
        return self._fn.apply(self.theta if theta is None else theta, self)

    def loss(self, theta=None):
        E, F = self.predict(theta)
        wE, wF = self.weights
        return wE * ((E - self.E_ref) ** 2).mean() + wF * ((F - self.F_ref) ** 2).mean()

    def fit(self, steps: int = 1000, lr: float = 3e-3, log_every: int = 0):
        import torch

        opt = torch.optim.Adam([self.theta], lr=lr)
        sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=max(1, steps), eta_min=lr * 0.02)
        hist = []
        for it in range(steps):
            opt.zero_grad()
            L = self.loss()
            L.backward()
            opt.step()
            sched.step()
            if log_every and (it % log_every == 0 or it == steps - 1):
                hist.append((it, float(L.detach())))
        self.model.theta = self.theta.detach().cpu().numpy().copy()
        self.head.set_theta(self.model.theta)
        return hist

    def rmse(self):
        """(energy RMSE per atom, force RMSE per component) on the training set."""
        import torch

        with torch.no_grad():
            E, F = self.predict(self.theta.detach())
        return (float(torch.sqrt(((E - self.E_ref) ** 2).mean())), float(torch.sqrt(((F - self.F_ref) ** 2).mean())))
