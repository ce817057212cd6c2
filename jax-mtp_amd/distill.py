"""Distillation of a level-24 MTP into the head, and its evaluation.

Workload (the distilled-head configuration): a level-24, 3-species MTP is
the teacher; the head reads level-12 basis values (29 per atom; 52 at the
optional level 14) through a 2x64 SiLU network; it is trained on 1e5 site
environments of 3-species FCC cells taken from Langevin MD at 300-1500 K
(one 1 ps trajectory per 256-atom cell, see md_snapshots) and labelled by
the teacher, and evaluated on a 100,000-atom 3-species FCC cell, with exact
level-24 re-evaluation of every atom outside the training box.  Reported:
atom-steps/s of the head (with fallback) and of the plain level-24
evaluation, the head's force RMSE against the level-24 forces, and the same
RMSE for the descriptor MTP alone (the method without the network).

Synthetic teacher.  All parameters are seeded; there is no trained
level-24 potential to start from.  Coefficients drawn N(0, 0.1) for every
basis function (as in the parity workloads) give energies of 1e3-1e4 eV per
atom and forces of 1e5 eV/A, which is no physical target for a force error
in meV/A.  The teacher therefore normalises each scalar basis function by
its RMS over 300 K snapshots and damps it with its MTP level (level of
M_{mu,nu} = 2 + 4 mu + nu, levels add under contraction), a converging
body-order expansion, then scales everything so that the RMS force at
300 K is 0.3 eV/A.  The random-parameter teacher has no stable lattice, so
it labels MD frames but does not drive the MD (see md_snapshots).  The
100,000-atom evaluation cell has Gaussian thermal displacements,
sigma = 0.05 A sqrt(900 K / 300 K) per component.
"""
from __future__ import annotations

import time

import numpy as np

from . import configs, training
from .head import DistilledHead, HeadModel, HeadTrainer
from .mtp_io import MTPData, read_mtp, write_mtp
from .potential import MTPPotential

A_FCC = 3.6
CUTOFF = 5.5
MAX_NBR = 80
FRACTIONS = (0.4, 0.3, 0.3)


def basis_levels(p: MTPData) -> np.ndarray:
    """MTP level of every alpha moment: 2 + 4 mu + nu for basic moments, sums for products."""
    lev = np.zeros(p.alpha_moments_count, dtype=np.int64)
    for b, (mu, a, bb, c) in enumerate(p.alpha_index_basic):
        lev[b] = 2 + 4 * mu + a + bb + c
    for i1, i2, _, i3 in p.alpha_index_times:
        lev[i3] = lev[i1] + lev[i2]
    return lev


# Langevin molecular dynamics of independent periodic cells, for training
# snapshots.  The synthetic teacher cannot drive it: it has no repulsive core
# and no stable lattice (under it a 300 K FCC cell collapses within about
# 60 fs, the closest pair going from 2.55 A to 1.8 A, where its radial basis
# is extrapolated).  The frames therefore come from a Morse pair potential
# with copper's parameters (Girifalco and Weizer, Phys. Rev. 114, 687
# (1959): D = 0.3429 eV, alpha = 1.3588 / A, r0 = 2.866 A), whose FCC lattice
# constant is close to the 3.6 A of the cells, and the teacher labels them.
AMU = 103.6427     # 1 amu A^2/fs^2 in eV
KB = 8.617333e-5   # eV/K
MASSES = (58.6934, 26.9815, 51.9961)  # amu of species 0, 1, 2 (Ni, Al, Cr)
MORSE = (0.3429, 1.3588, 2.866)       # D (eV), alpha (1/A), r0 (A)


def _min_image(x, box):
    """Pair vectors d[c, i, j] = x_j - x_i of cells x [C, n, 3] under the minimum image."""
    import torch

    d = x[:, None, :, :] - x[:, :, None, :]
    return d - box * torch.round(d / box)


def _cell_rows(x, types, box, cutoff, K):
    """Padded rows of C independent periodic cells (box >= 2 cutoff): x [C, n, 3],
    types [C, n] (device).  Slots hold the K nearest neighbours (minimum
    image), those beyond the cutoff marked empty (j = C n, type -1, r = 2000)."""
    import torch

    C, n, _ = x.shape
    d = _min_image(x, box)
    r2 = (d * d).sum(-1)
    r2.diagonal(dim1=1, dim2=2).fill_(float("inf"))
    k = min(K, n - 1)
    vals, idx = torch.topk(r2, k, dim=2, largest=False)
    if k == K and bool((vals[:, :, -1] < cutoff * cutoff).any()):
        raise RuntimeError("more than K neighbours within the cutoff")
    valid = vals < cutoff * cutoff
    rij = torch.gather(d, 2, idx[..., None].expand(C, n, k, 3))
    rij = torch.where(valid[..., None], rij, torch.full_like(rij, 2000.0))
    off = (torch.arange(C, device=x.device) * n)[:, None, None]
    js = torch.where(valid, idx + off, torch.full_like(idx, C * n)).to(torch.int32)
    jt = torch.gather(types[:, None, :].expand(C, n, n), 2, idx)
    jt = torch.where(valid, jt, torch.full_like(jt, -1)).to(torch.int32)
    pad = K - k
    if pad:
        rij = torch.cat([rij, torch.full((C, n, pad, 3), 2000.0, dtype=rij.dtype, device=x.device)], 2)
        js = torch.cat([js, torch.full((C, n, pad), C * n, dtype=js.dtype, device=x.device)], 2)
        jt = torch.cat([jt, torch.full((C, n, pad), -1, dtype=jt.dtype, device=x.device)], 2)
    return (types.reshape(-1).to(torch.int32), js.reshape(C * n, K), rij.reshape(C * n, K, 3),
            jt.reshape(C * n, K))


def _morse_forces(x, box, cutoff):
    """Forces [C, n, 3] of the Morse pair potential MORSE, force shifted to zero at the cutoff."""
    import torch

    D, al, r0 = MORSE
    d = _min_image(x, box)
    r = torch.sqrt((d * d).sum(-1))
    r.diagonal(dim1=1, dim2=2).fill_(float("inf"))

    def dv(rr):  # dV/dr
        e = torch.exp(-al * (rr - r0))
        return 2.0 * al * D * (e - e * e)

    dvc = float(dv(torch.tensor(cutoff, dtype=torch.float64)))
    w = torch.where(r < cutoff, (dv(r) - dvc) / r, torch.zeros_like(r))
    return (w[..., None] * d).sum(2)   # dV/dr > 0 pulls i toward j


def md_snapshots(n_cells: int, ncell=(4, 4, 4), T_range=(300.0, 1500.0), steps=1000, dt=1.0, gamma=0.01,
                 seed=0, driver=None, stats_every=0, log=None):
    """Langevin MD (BAOAB, time step dt fs, friction gamma 1/fs) of n_cells
    independent 3-species FCC cells (a = 3.6 A, species 40/30/30, masses
    MASSES), each at its own temperature drawn uniformly from T_range,
    started from the lattice with Maxwell velocities.  The forces are those
    of the Morse potential MORSE, or of an MTP when driver is one.  Returns
    the last frame of every cell as (pos, types, cell, T) and, with
    stats_every, a list of per-interval statistics (step, mean kinetic
    temperature, smallest pair distance, largest |F|)."""
    import torch

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda")
    base, cell = configs.fcc(A_FCC, ncell)
    n = len(base)
    types_np = np.stack([configs.species_by_fraction(n, FRACTIONS, rng) for _ in range(n_cells)])
    T = rng.uniform(*T_range, n_cells)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    x = torch.as_tensor(np.broadcast_to(base, (n_cells, n, 3)).copy(), dtype=torch.float64, device=dev)
    types = torch.as_tensor(types_np, dtype=torch.int64, device=dev)
    box = torch.as_tensor(np.asarray(cell, dtype=float), dtype=torch.float64, device=dev)
    m = torch.as_tensor(np.asarray(MASSES)[types_np], dtype=torch.float64, device=dev)[..., None] * AMU
    kT = torch.as_tensor(KB * T, dtype=torch.float64, device=dev)[:, None, None]
    sv = torch.sqrt(kT / m)
    v = sv * torch.randn(x.shape, dtype=torch.float64, device=dev, generator=gen)
    N = n_cells * n
    pot = MTPPotential(driver, max_atoms=N, max_nbr=MAX_NBR) if driver is not None else None

    def forces(xc):
        if pot is None:
            return _morse_forces(xc, box, CUTOFF)
        it, js, rij, jt = _cell_rows(xc, types, box, CUTOFF, MAX_NBR)
        _, F, _ = pot.calc_energy_forces_stress_padded(it, js, rij, jt, 3, 1.0, N, N)
        return F.reshape(n_cells, n, 3)

    c1 = float(np.exp(-gamma * dt))
    c2 = float(np.sqrt(1.0 - c1 * c1))
    stats = []
    F = forces(x)
    for step in range(1, steps + 1):
        v = v + 0.5 * dt * F / m
        x = x + 0.5 * dt * v
        v = c1 * v + c2 * sv * torch.randn(x.shape, dtype=torch.float64, device=dev, generator=gen)
        x = x + 0.5 * dt * v
        F = forces(x)
        v = v + 0.5 * dt * F / m
        if stats_every and step % stats_every == 0:
            tk = float((m * v * v).sum() / (3 * N * KB))
            d = _min_image(x, box)
            r2 = (d * d).sum(-1)
            r2.diagonal(dim1=1, dim2=2).fill_(float("inf"))
            row = (step, tk, float(torch.sqrt(r2.min())), float(F.norm(dim=-1).max()))
            stats.append(row)
            if log:
                log(f"[md] step {step}: T_kin {row[1]:.0f} K, closest pair {row[2]:.3f} A, max |F| {row[3]:.2f} eV/A")
    if pot is not None:
        pot.close()
    xs = x.cpu().numpy()
    snaps = [(xs[c], types_np[c].astype(np.int32), cell, float(T[c])) for c in range(n_cells)]
    return (snaps, stats) if stats_every else snaps


def inference_cell(ncell=(25, 25, 40), T=900.0, sigma300=0.05, seed=1):
    """The 100,000-atom 3-species FCC cell."""
    rng = np.random.default_rng(seed)
    pos, cell = configs.fcc(A_FCC, ncell)
    pos = pos + rng.normal(0.0, sigma300 * np.sqrt(T / 300.0), pos.shape)
    types = configs.species_by_fraction(len(pos), FRACTIONS, rng)
    return configs.Workload(f"fcc{len(pos)}_head", None, pos, types, cell, CUTOFF, MAX_NBR)


def teacher_params(seed=0, decay=0.25, f_rms=0.3, calib_cells=8) -> MTPData:
    """Level-24, 3-species teacher with physical force scale (see module doc)."""
    import torch

    rng = np.random.default_rng(seed)
    p = configs.random_params(24, 3, 10, 2.0, CUTOFF, rng)
    lev = basis_levels(p)[np.asarray(p.alpha_moment_mapping)]
    raw = rng.normal(0.0, 1.0, len(lev)) * decay ** ((lev - 2) / 2.0)
    snaps = md_snapshots(calib_cells, T_range=(300.0, 300.0), seed=seed + 100)
    batch = training.pack_configs([s[:3] for s in snaps], CUTOFF, MAX_NBR)
    pot = MTPPotential(p, max_atoms=batch.natoms, max_nbr=MAX_NBR)
    b = pot.descriptor(batch.itypes, batch.rij, batch.jtypes)
    rms = torch.sqrt((b ** 2).mean(0)).cpu().numpy()
    xi = np.where(rms > 0, raw / np.maximum(rms, 1e-300), 0.0)
    pot.set_params(moment_coeffs=xi)
    _, F, _ = pot.calc_energy_forces_stress_padded(batch.itypes, batch.js, batch.rij, batch.jtypes, 3, 1.0,
                                                   batch.natoms, batch.natoms)
    scale = f_rms / float(torch.sqrt((F ** 2).mean()))
    pot.close()
    p.moment_coeffs = xi * scale
    p.potential_name = "synthetic_level24_teacher"
    return p


def label(pot: MTPPotential, batch):
    """Site energies [N], forces [N, 3], per-configuration energies and stresses."""
    _, F, _, extra = pot.calc_energy_forces_stress_padded(batch.itypes, batch.js, batch.rij, batch.jtypes, 3, 1.0,
                                                          batch.natoms, batch.natoms, return_all=True)
    E_c, F_b, S_c = training.batch_forward(pot, batch.to_device())
    return extra["eatom"].clone(), F.clone(), E_c.clone(), S_c.clone()


def with_pair_invariants(p: MTPData, mus) -> MTPData:
    """p with the pair moments M_{mu,0} = sum_k f_mu(r_k) of the given mu added
    as basic moments and scalar basis functions (inserted after the basic
    moments; later moment indices shift)."""
    mus = [int(m) for m in mus]
    k = len(mus)
    if k == 0:
        return p
    B = int(p.alpha_index_basic_count)
    basic = np.asarray(p.alpha_index_basic, dtype=np.int64).reshape(-1, 4)
    times = np.asarray(p.alpha_index_times, dtype=np.int64).reshape(-1, 4).copy()
    for col in (0, 1, 3):
        times[:, col] = np.where(times[:, col] >= B, times[:, col] + k, times[:, col])
    mapping = np.asarray(p.alpha_moment_mapping, dtype=np.int64)
    mapping = np.concatenate([np.where(mapping >= B, mapping + k, mapping), B + np.arange(k)])
    p.alpha_index_basic = np.concatenate([basic, np.array([[m, 0, 0, 0] for m in mus], dtype=np.int64)])
    p.alpha_index_basic_count = B + k
    p.alpha_moments_count = int(p.alpha_moments_count) + k
    p.alpha_index_times = times
    p.alpha_moment_mapping = mapping
    p.alpha_scalar_moments = int(p.alpha_scalar_moments) + k
    p.moment_coeffs = np.concatenate([np.asarray(p.moment_coeffs, dtype=np.float64), np.zeros(k)])
    R0, R = int(p.radial_funcs_count), max(int(p.radial_funcs_count), max(mus) + 1)
    if R > R0:
        rc = np.asarray(p.radial_coeffs, dtype=np.float64)
        pad = np.zeros(rc.shape[:2] + (R - R0,) + rc.shape[3:])
        p.radial_coeffs = np.concatenate([rc, pad], axis=2)
    p.radial_funcs_count = R
    return p


def fit_descriptor(teacher: MTPData, batch, E_c, F, S_c, seed=0, steps=0, log=None, level=12, pair_mus=()):
    """Low-level MTP (level 12 by default) over the teacher's own invariants.

    A level-L basis uses the radial functions mu with 2 + 4 mu <= L (mu < 3
    at level 12, mu < 4 at level 14); they start as the teacher's, so the head's inputs start as
    the teacher's own low-level basis values.  Moment coefficients come from linear least
    squares; with steps > 0, L-BFGS then adjusts all coefficients (radial
    ones included) to the teacher's energies and forces.  Alone, this
    potential is the method without the network."""
    rng = np.random.default_rng(seed)
    p = configs.random_params(level, teacher.species_count, teacher.radial_basis_size, teacher.min_dist,
                              teacher.max_dist, rng, scaling=teacher.scaling)
    p = with_pair_invariants(p, pair_mus)
    R = p.radial_funcs_count
    p.radial_coeffs = np.array(teacher.radial_coeffs)[:, :, :R, :].copy()
    p.species_coeffs = np.array(teacher.species_coeffs).copy()
    p.potential_name = f"level{level}_descriptor"
    pot = MTPPotential(p, max_atoms=batch.natoms, max_nbr=MAX_NBR)
    tr = training.Trainer(pot, batch, E_c, F, S_c, weights=(1.0, 1.0, 0.0))
    hist = [float(tr.loss().detach())]
    for _ in range(steps):
        hist.append(tr.step())
    th = tr.theta.detach().cpu().numpy()
    S = len(p.moment_coeffs)
    sp = p.species_count
    pot.set_params(th[:S], th[S:S + sp], th[S + sp:])
    if log:
        log(f"[distill] level-{level} fit loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    return pot, hist


def _time(fn, reps=5, warmup=2):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        fn()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1) / reps


def run(n_env=100_000, infer_ncell=(25, 25, 40), head_steps=1500, desc_steps=0, seed=0, decay=0.25, log=print,
        save_dir=None, head_lr=3e-3, head_weights=(1.0, 1.0), desc_level=12, pair_mus=(), md_steps=1000):
    """Teacher, training set, descriptor fit, head training, then evaluate();
    with save_dir the trained head (head.npz) and the teacher (teacher.mtp)
    are written there for infer()."""
    import torch

    t_start = time.time()
    teacher = teacher_params(seed, decay=decay)
    snaps = md_snapshots(-(-n_env // 256), steps=md_steps, seed=seed + 1)
    batch = training.pack_configs([s[:3] for s in snaps], CUTOFF, MAX_NBR)
    log(f"[distill] {len(snaps)} MD snapshots ({md_steps} fs each, 300-1500 K), {batch.natoms} environments, "
        f"K={batch.nslots}")
    tpot = MTPPotential(teacher, max_atoms=batch.natoms, max_nbr=MAX_NBR)
    E_site, F, E_c, S_c = label(tpot, batch)
    tpot.close()
    desc, hist_d = fit_descriptor(teacher, batch, E_c, F, S_c, seed=seed, steps=desc_steps, log=log,
                                  level=desc_level, pair_mus=pair_mus)
    tr = HeadTrainer(desc, batch, E_site, F, weights=head_weights, margin=0.05, seed=seed)
    t0 = time.time()
    hist = tr.fit(steps=head_steps, lr=head_lr, log_every=max(1, head_steps // 10))
    torch.cuda.synchronize()
    t_fit = time.time() - t0
    e_rmse, f_rmse = tr.rmse()
    log(f"[distill] head fit {t_fit:.1f} s: loss {hist[0][1]:.4e} -> {hist[-1][1]:.4e}; train RMSE E "
        f"{e_rmse * 1e3:.2f} meV/atom, F {f_rmse * 1e3:.2f} meV/A")
    model = tr.model
    tr.head.close()
    desc.close()
    del tr
    torch.cuda.empty_cache()
    if save_dir is not None:
        import pathlib

        d = pathlib.Path(save_dir)
        d.mkdir(parents=True, exist_ok=True)
        model.save(d / "head.npz")
        write_mtp(d / "teacher.mtp", teacher)
    res = evaluate(model, teacher, infer_ncell, seed=seed + 2)
    res.update({"train_environments": batch.natoms, "train_force_rmse_meV_per_A": f_rmse * 1e3,
                "head_fit_s": t_fit, "total_s": time.time() - t_start})
    return res


def evaluate(model, teacher: MTPData, infer_ncell=(25, 25, 40), seed=2, reps=5):
    """Head (with exact fallback) vs plain teacher on the inference cell: timing and errors."""
    import torch

    w = inference_cell(infer_ncell, seed=seed)
    nb = w.neighbors()
    dev = torch.device("cuda")
    it = torch.as_tensor(nb["itypes"], dtype=torch.int32, device=dev)
    js = torch.as_tensor(nb["js"], dtype=torch.int32, device=dev)
    rij = torch.as_tensor(nb["rij"], dtype=torch.float64, device=dev)
    jt = torch.as_tensor(nb["jtypes"], dtype=torch.int32, device=dev)
    A = w.natoms
    exact = MTPPotential(teacher, max_atoms=A, max_nbr=MAX_NBR)
    exact.autotune(it, js, rij, jt)
    head = DistilledHead(model, exact, max_atoms=A, max_nbr=MAX_NBR)
    head.desc.autotune(it, js, rij, jt)
    lin = head.desc  # the fitted descriptor MTP itself
    Ex, Fx, _, ex_x = exact.calc_energy_forces_stress_padded(it, js, rij, jt, 3, w.volume, A, A, return_all=True)
    Eh, Fh, _, extra = head.calc_energy_forces_stress_padded(it, js, rij, jt, 3, w.volume, A, A, return_all=True)
    El, Fl, _, ex_l = lin.calc_energy_forces_stress_padded(it, js, rij, jt, 3, w.volume, A, A, return_all=True)
    rm = lambda a, b: float(torch.sqrt(((a - b) ** 2).mean()))
    ms_exact = _time(lambda: exact.calc_energy_forces_stress_padded(it, js, rij, jt, 3, w.volume, A, A, sync=False),
                     reps=reps)
    ms_head = _time(lambda: head.calc_energy_forces_stress_padded(it, js, rij, jt, 3, w.volume, A, A, sync=False),
                    reps=reps)
    res = {
        "workload": w.name, "natoms": A,
        "atom_steps_per_s_head": A / (ms_head * 1e-3), "atom_steps_per_s_exact": A / (ms_exact * 1e-3),
        "speedup": ms_exact / ms_head, "ms_head": ms_head, "ms_exact": ms_exact,
        "n_fallback": extra["n_fallback"],
        "force_rmse_head_meV_per_A": rm(Fh, Fx) * 1e3,
        "force_rmse_desc_meV_per_A": rm(Fl, Fx) * 1e3,
        "force_rms_exact_eV_per_A": float(torch.sqrt((Fx ** 2).mean())),
        # site energies E_i against the exact ones, atom by atom (RMSE), and the error of the cell's total
        "energy_rmse_head_meV_per_atom": rm(extra["eatom"], ex_x["eatom"]) * 1e3,
        "energy_rmse_desc_meV_per_atom": rm(ex_l["eatom"], ex_x["eatom"]) * 1e3,
        "total_energy_err_head_meV_per_atom": abs(float(Eh - Ex)) / A * 1e3,
        "total_energy_err_desc_meV_per_atom": abs(float(El - Ex)) / A * 1e3,
        "variant_exact": exact.variant, "variant_desc": head.desc.variant,
    }
    head.close()
    exact.close()
    return res


def infer(save_dir, infer_ncell=(25, 25, 40), seed=2, reps=20):
    """evaluate() of a head saved by run(save_dir=...)."""
    import pathlib

    d = pathlib.Path(save_dir)
    return evaluate(HeadModel.load(d / "head.npz"), read_mtp(d / "teacher.mtp"), infer_ncell, seed=seed, reps=reps)


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="distil a level-24 MTP into the head and evaluate it")
    ap.add_argument("--n-env", type=int, default=100_000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--decay", type=float, default=0.35)
    ap.add_argument("--desc-level", type=int, default=12, help="MTP level of the descriptor")
    ap.add_argument("--pair-mus", default="", help="comma-separated mu of pair moments added to the descriptor")
    ap.add_argument("--desc-steps", type=int, default=40, help="L-BFGS steps of the descriptor fit")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save", default=None, help="directory for head.npz and teacher.mtp")
    ap.add_argument("--infer", default=None, help="only evaluate the head saved in this directory")
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args(argv)
    if a.infer:
        res = infer(a.infer, reps=a.reps)
    else:
        res = run(n_env=a.n_env, head_steps=a.steps, desc_steps=a.desc_steps, decay=a.decay, seed=a.seed,
                  save_dir=a.save, desc_level=a.desc_level,
                  pair_mus=tuple(int(m) for m in a.pair_mus.split(",") if m))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
