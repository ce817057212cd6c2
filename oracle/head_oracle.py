"""FP64 numpy oracle of the distilled head — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.

The reference has no learned head, so nothing here restates a reference
file; the head's definition is the network documented in
include/mtp_hip.h (mtp_head_*) and jax-mtp_amd/csrc/mtp_head.hpp:

  z   = [(b - mean) * inv_std, onehot(type)]  padded to 64
  a1  = z W1 + b1,  h1 = silu(a1)             W1 [64, 64]
  a2  = h1 W2 + b2, h2 = silu(a2)             W2 [64, 64]
  E   = e_shift[type] + e_scale * (h2 . w3 + b3)
        [+ lin_sc[type] + lin . b   in residual form]

and its composition with the MTP oracle (mtp_oracle.c): basis values b of
the descriptor potential, slot gradients of the head energy through the
descriptor (per-row coefficient vectors dE/db), the exact potential for
rows outside the training box, then the reference's force gather and stress
(jax_backend.py:263-303, 355-370).

Pins (tests/test_head.py): the network and its two gradients against torch
autograd in float64 on the CPU (an independent implementation, including
double backward for the force-matching gradient) and central differences;
the per-row-coefficient slot gradients against central differences of the
descriptor energy.
"""
from __future__ import annotations

import numpy as np

H = 64
IN = 64
N_THETA = IN * H + H + H * H + H + H + 1


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    assert theta.shape == (N_THETA,)
    o = 0
    W1 = theta[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = theta[o:o + H]; o += H
    W2 = theta[o:o + H * H].reshape(H, H); o += H * H
    b2 = theta[o:o + H]; o += H
    w3 = theta[o:o + H]; o += H
    b3 = theta[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]]).astype(np.float64)


def silu(a):
    # logistic function without overflow: exp of a non-positive argument only
    e = np.exp(-np.abs(a))
    g = np.where(a >= 0, 1.0 / (1.0 + e), e / (1.0 + e))
    return a * g, g * (1.0 + a * (1.0 - g)), g * (1.0 - g) * (2.0 + a * (1.0 - 2.0 * g))


def wrap_species(t, Sp):
    t = np.asarray(t).copy()
    t[t < 0] += Sp
    return np.clip(t, 0, Sp - 1)


def inputs(basis, itypes, mean, inv_std, Sp):
    A, S = basis.shape
    z = np.zeros((A, IN))
    z[:, :S] = (basis - mean) * inv_std
    ts = wrap_species(itypes, Sp)
    z[np.arange(A), S + ts] = 1.0
    return z, ts


def net_and_grad(theta, z):
    """net [A] and d net / d z [A, IN]."""
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    h1, s1, _ = silu(z @ W1 + b1)
    h2, s2, _ = silu(h1 @ W2 + b2)
    net = h2 @ w3 + b3
    dz = (((w3 * s2) @ W2.T) * s1) @ W1.T
    return net, dz


def eval_head(model, basis, itypes):
    """(E [A], xi [A, S] = dE/db, flags [A]) of HeadModel-like `model`."""
    S = basis.shape[1]
    Sp = len(model.e_shift)
    z, ts = inputs(basis, itypes, model.mean, model.inv_std, Sp)
    net, dz = net_and_grad(model.theta, z)
    E = model.e_shift[ts] + model.e_scale * net
    xi = model.e_scale * dz[:, :S] * model.inv_std
    lin = getattr(model, "lin", None)
    if lin is not None:
        E = E + model.lin_sc[ts] + basis @ lin
        xi = xi + lin
    if model.lo is None:
        flags = np.zeros(len(E), dtype=np.int32)
    else:
        zz = z[:, :S]
        out = (zz < model.lo[ts]) | (zz > model.hi[ts])
        flags = out.any(axis=1).astype(np.int32)
    return E, xi, flags


def param_grad(model, basis, itypes, alpha, v):
    """d/dtheta sum_a [alpha_a E_a + v_a . dE_a/db_a]  (v in basis units)."""
    S = basis.shape[1]
    Sp = len(model.e_shift)
    z, _ = inputs(basis, itypes, model.mean, model.inv_std, Sp)
    vz = np.zeros_like(z)
    vz[:, :S] = v * model.inv_std
    W1, b1, W2, b2, w3, b3 = unpack(model.theta)
    al = np.asarray(alpha, dtype=np.float64)
    a1 = z @ W1 + b1
    ad1 = vz @ W1
    h1, s1, q1 = silu(a1)
    hd1 = s1 * ad1
    a2 = h1 @ W2 + b2
    ad2 = hd1 @ W2
    h2, s2, q2 = silu(a2)
    hd2 = s2 * ad2
    gw3 = (al[:, None] * h2 + hd2).sum(0)
    gb3 = al.sum()
    ac2 = s2 * w3
    ab2 = w3 * (al[:, None] * s2 + q2 * ad2)
    gW2 = h1.T @ ab2 + hd1.T @ ac2
    gb2 = ab2.sum(0)
    hb1 = ab2 @ W2.T
    hcb1 = ac2 @ W2.T
    ac1 = s1 * hcb1
    ab1 = s1 * hb1 + q1 * ad1 * hcb1
    gW1 = z.T @ ab1 + vz.T @ ac1
    gb1 = ab1.sum(0)
    return model.e_scale * pack(gW1, gb1, gW2, gb2, gw3, gb3)


def head_compute(model, orc_desc, orc_exact, itypes, js, rij, jtypes, cell_rank, volume, n_e, n_f):
    """Head energies/forces/stress with exact re-evaluation of out-of-box rows.

    Returns (E, F [A,3], stress6, eatom [A], flags [A], g [A,K,3])."""
    from oracle import Oracle

    S = len(model.mean)
    _, _, basis = orc_desc.compact_xi(itypes, rij, jtypes, np.zeros(S), grad=False)
    eatom, xi, flags = eval_head(model, basis, itypes)
    _, g, _ = orc_desc.compact_xi(itypes, rij, jtypes, xi)
    rows = np.nonzero((flags != 0) & (np.asarray(itypes) >= 0))[0] if orc_exact is not None else np.zeros(0, int)
    if len(rows):
        Ee, ge = orc_exact.compact(np.asarray(itypes)[rows], np.asarray(rij)[rows], np.asarray(jtypes)[rows])
        eatom[rows] = Ee
        g[rows] = ge
    # forces masked by natoms_energy; n_f is unused (jax_backend.py:414-418 wiring)
    F = Oracle.forces(g, js, n_e)
    stress, _ = Oracle.stress(rij, g, volume, cell_rank)
    E = Oracle.energy(eatom, n_e)
    return E, F, stress, eatom, flags, g
