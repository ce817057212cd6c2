#!/usr/bin/env python3
"""Benchmark: atom-steps/s (energy + forces + virial) of the level-24 MTP on
the 32,000-atom 3-species FCC workload, one GPU, FP64.

A step is one full force evaluation of the padded neighbor arrays already
resident in HBM: radial basis -> moments -> contraction -> energy ->
adjoint -> per-slot gradients -> Newton-3 force gather -> virial.

Prints ONE JSON line.  ``training_step`` adds the packed training
configuration (512 uneven configurations, level 16, eager vs HIP-graph
evaluation).  ``distilled_head`` adds the head configuration:
distillation of the level-24 MTP into an MLP over level-12 basis values
(1e5 training environments) and its atom-steps/s and force RMSE on a
100,000-atom cell against plain level-24 evaluation.  ``roofline`` is the atom kernel's achieved
algorithmic FP64 FLOP rate (the fixed per-atom FLOP formula of ``algorithmic_flops``)
over its average launch time, measured with HIP events on its own stream.
``cpu_baseline`` times the FP64 CPU restatement of the reference's
full-tensor algorithm on a bounded sample of the same workload on the host
cores; the same leg runs the compact CPU oracle over the whole workload
and reports the GPU force error against it.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "atom-steps/s (E+F) level-24 MTP 32k atoms, 1 GPU; force err vs FP64 ref"
PEAK_FP64_TFLOPS = 78.6


def algorithmic_flops(nbr_counts, N, R, B, T, S):
    """FP64 FLOPs of one force evaluation in the compact formulation (FMA = 2)."""
    fwd = 12 + 3 * (N - 2) + (2 * R * N + 2 * R) + 4 * B
    adj = 14 * B + (4 * R * N + 5 * N) + 15 + 24
    n = np.asarray(nbr_counts, dtype=np.float64)
    return float(n.sum() * (fwd + adj) + len(n) * (3 * T + 5 * T + 2 * S))


def algorithmic_bytes(nbr_counts):
    n = np.asarray(nbr_counts, dtype=np.float64)
    return float((32.0 * n + 40.0).sum())


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_values(name, kernel):
    """key=value lines of a committed rocprofv3 summary under profiles/ (None when absent or for another kernel)."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    kv = dict(line.split("=", 1) for line in f.read_text().splitlines() if "=" in line and not line.startswith("#"))
    if kv.get("kernel", "").strip() != kernel:
        return None
    return {k: v.strip() for k, v in kv.items()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


PEAK_LDS_TBS = 78.6    # 128 B/clk/CU (64-bit reads) x 256 CUs x 2.4 GHz
PEAK_HBM_TBS = 8.0


def traffic_bytes(workload, kernel):
    """HBM bytes per atom-kernel launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by the
    same counter's reading for a plain copy of known size), or None."""
    f = ROOT / "profiles" / f"atom_kernel_{workload}_traffic.txt"
    if not f.exists():
        return None
    kv = dict(line.split("=", 1) for line in f.read_text().splitlines() if "=" in line and not line.startswith("#"))
    if kv.get("kernel", "").strip() != kernel:
        return None
    return float(kv["traffic_bytes"])


def head_leg(seed, steps, save_dir=None, desc_level=12, pair_mus=()):
    """Distilled head vs plain level-24 on the 100,000-atom cell (see jax-mtp_amd/distill.py); the
    head's inputs are the basis values of a level-``desc_level`` MTP, plus the pair moments
    sum_k f_mu(r_k) of ``pair_mus``."""
    jm = importlib.import_module("jax-mtp_amd")
    res = jm.distill.run(n_env=100_000, infer_ncell=(25, 25, 40), head_steps=steps, desc_steps=40, seed=seed,
                         decay=0.35, log=log, save_dir=save_dir, desc_level=desc_level, pair_mus=pair_mus)
    res["desc_level"] = desc_level
    res["desc_pair_mus"] = list(pair_mus)
    res["desc_lbfgs_steps"] = 40
    res["teacher"] = "synthetic level-24, 3 species, basis functions RMS-normalised and damped 0.35^((level-2)/2)"
    return res


def training_leg(seed, lbfgs_steps=5, reps=10):
    """Config 4: one training evaluation (forward + loss gradients + hand-written parameter backward) over
    512 packed configurations (8-256 atoms, FCC/BCC/random-dense, 2 species), level-16 student, labels from
    a seeded level-20 teacher.  Eager (autograd Function) vs replayed HIP graph."""
    import torch

    jm = importlib.import_module("jax-mtp_amd")
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    cells = jm.training.random_cells(512, rng, natoms_range=(8, 256), species=2)
    b = jm.training.pack_configs(cells, cutoff=5.0, max_nbr=64)
    t_pack = time.perf_counter() - t0
    teacher = jm.configs.random_params(20, 2, 8, 1.6, 5.0, rng, sigma=0.05)
    tp = jm.MTPPotential(teacher, max_atoms=b.natoms, max_nbr=b.nslots)
    E_ref, F_ref, S_ref = [t.clone() for t in jm.training.batch_forward(tp, b.to_device())]
    tp.close()
    student = jm.configs.random_params(16, 2, 8, 1.6, 5.0, rng, sigma=0.05)
    pot = jm.MTPPotential(student, max_atoms=b.natoms, max_nbr=b.nslots)
    tr = jm.training.Trainer(pot, b, E_ref, F_ref, S_ref, lls=True, graph=True)
    th = tr.theta.detach().clone()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t) / reps

    tr_e = jm.training.Trainer.__new__(jm.training.Trainer)
    tr_e.__dict__.update(tr.__dict__)
    tr_e.graphed = None
    ms_eager = timed(lambda: tr_e.value_and_grad(th))
    ms_graph = timed(lambda: tr.value_and_grad(th))
    db = tr.dbatch
    ms_fwd = timed(lambda: jm.training.batch_forward(pot, db))
    l0 = tr.value_and_grad(th)[0]
    losses = [tr.step() for _ in range(lbfgs_steps)]
    pot.close()
    return {"configs": b.nconf, "atoms": b.natoms, "max_nbr": b.nslots, "level": 16, "teacher_level": 20,
            "ms_eval_eager": ms_eager, "ms_eval_graph": ms_graph, "ms_forward": ms_fwd,
            "atom_evals_per_s_graph": b.natoms / (ms_graph * 1e-3), "loss_after_lls": l0,
            "loss_after_lbfgs": losses[-1], "lbfgs_steps": lbfgs_steps, "host_batch_build_s": t_pack}


def lammps_full_list(pos, types, cell, cutoff, skin):
    """Locals + periodic ghost copies (owner tags, 1-based types) and a full neighbor list with a skin, as
    LAMMPS hands them to the pair style (rows unsorted, self excluded)."""
    from scipy.spatial import cKDTree

    L = np.asarray(cell, dtype=float)
    n = len(pos)
    rg = cutoff + skin
    xs, owners = [pos], [np.arange(n)]
    for sh in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1], indexing="ij")).reshape(3, -1).T:
        if not sh.any():
            continue
        p = pos + sh * L
        keep = np.all((p >= -rg) & (p < L + rg), axis=1)
        xs.append(p[keep]); owners.append(np.nonzero(keep)[0])
    x = np.concatenate(xs)
    own = np.concatenate(owners)
    tree = cKDTree(x)
    rows = tree.query_ball_point(pos, rg)
    neigh = [np.array([j for j in r if j != i], dtype=np.int32) for i, r in enumerate(rows)]
    return (np.ascontiguousarray(x), (np.asarray(types)[own] + 1).astype(np.int32), (own + 1).astype(np.int32),
            neigh)


def pair_style_leg(jm, pot, w, reps=10):
    """The pair style's per-step cost at this workload: host packing + rows through mtp_compute (the
    reference's flow), and mtp_compute_full_list (packing and staging on the GPU), both from the same
    LAMMPS-like arrays; wall-clock per call, host work and PCIe transfers included."""
    import ctypes

    t0 = time.perf_counter()
    x, typ, tag, neigh = lammps_full_list(w.pos, w.types, w.cell, w.cutoff, 1.5)
    t_list = time.perf_counter() - t0
    n, K = w.natoms, pot.max_nbr
    L = jm.capi.lib()
    vp = ctypes.c_void_p
    numneigh = np.array([len(r) for r in neigh], np.int32)
    first = (vp * n)(*[r.ctypes.data for r in neigh])
    ilist = np.arange(n, dtype=np.int32)
    prd = np.asarray(w.cell, dtype=np.float64)
    per = np.ones(3, np.int32)
    pos = np.zeros((n, 3 * K)); nbr = np.zeros((n, K), np.int32); nt = np.zeros((n, K), np.int32)
    cnt = np.zeros(n, np.int32); at = np.zeros(n, np.int32)
    prow = (vp * n)(*[pos[i].ctypes.data for i in range(n)])
    nrow = (vp * n)(*[nbr[i].ctypes.data for i in range(n)])
    trow = (vp * n)(*[nt[i].ctypes.data for i in range(n)])
    E, F = ctypes.c_double(), np.zeros((n, 3))
    s6, v6 = np.zeros(6), np.zeros(6)
    tr = ctypes.c_longlong()
    vol = float(np.prod(prd))

    def rows_step():
        tot = L.mtp_pack_full_list(n, x.ctypes.data_as(vp), typ.ctypes.data_as(vp), tag.ctypes.data_as(vp), n,
                                   ilist.ctypes.data_as(vp), numneigh.ctypes.data_as(vp), ctypes.cast(first, vp),
                                   w.cutoff, prd.ctypes.data_as(vp), per.ctypes.data_as(vp), n, K,
                                   pos.ctypes.data_as(vp), nbr.ctypes.data_as(vp), nt.ctypes.data_as(vp),
                                   cnt.ctypes.data_as(vp), at.ctypes.data_as(vp), ctypes.byref(tr))
        jm.capi.check(L.mtp_compute(pot.ctx, n, int(tot), ctypes.cast(prow, vp), at.ctypes.data_as(vp),
                                    ctypes.cast(nrow, vp), cnt.ctypes.data_as(vp), ctypes.cast(trow, vp), vol,
                                    ctypes.byref(E), F.ctypes.data_as(vp), s6.ctypes.data_as(vp),
                                    v6.ctypes.data_as(vp)), pot.ctx)
        return E.value, F.copy()

    def full_list_step():
        jm.capi.check(L.mtp_compute_full_list(pot.ctx, n, len(x), x.ctypes.data_as(vp), typ.ctypes.data_as(vp),
                                              tag.ctypes.data_as(vp), n, ilist.ctypes.data_as(vp),
                                              numneigh.ctypes.data_as(vp), ctypes.cast(first, vp), new_list[0],
                                              prd.ctypes.data_as(vp), per.ctypes.data_as(vp), vol, ctypes.byref(E),
                                              F.ctypes.data_as(vp), s6.ctypes.data_as(vp), v6.ctypes.data_as(vp),
                                              ctypes.byref(tr)), pot.ctx)
        return E.value, F.copy()

    def timed(fn):
        fn()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        return 1e3 * (time.perf_counter() - t) / reps

    new_list = [1]
    ms_rows = timed(rows_step)
    ms_full_rebuild = timed(full_list_step)   # every call a neighbor-list rebuild step
    new_list[0] = 0
    ms_full = timed(full_list_step)           # steps between rebuilds: positions only
    Er, Fr = rows_step()
    Ef, Ff = full_list_step()
    return {"ms_per_step_host_rows": ms_rows, "ms_per_step_full_list": ms_full,
            "ms_per_step_full_list_rebuild": ms_full_rebuild,
            "neighbor_entries": int(numneigh.sum()), "ghosts": int(len(x) - n), "skin": 1.5,
            "bitwise_equal": bool(Er == Ef and np.array_equal(Fr, Ff)), "list_build_s": t_list}


def cpu_baseline(w, nb, gpu_F, threads):
    """Reference-algorithm (full-tensor tree) CPU timing on a bounded sample,
    plus the compact CPU oracle on the whole workload (and the GPU force
    error against it)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    O.build()
    import tempfile

    jm = importlib.import_module("jax-mtp_amd")
    with tempfile.NamedTemporaryFile("w", suffix=".mtp", delete=False) as f:
        f.write(jm.format_mtp(w.params))
        path = f.name
    P = O.read_mtp(path)
    os.unlink(path)
    orc = O.Oracle(P, with_tree=True)
    it, js, rij, jt = nb["itypes"], nb["js"], nb["rij"], nb["jtypes"]
    # calibrate the tree path on 64 atoms, then size the sample for ~15 s
    t0 = time.perf_counter()
    orc.tree_eval(it[:64], rij[:64], jt[:64], True, threads)
    t_cal = time.perf_counter() - t0
    n_s = int(min(len(it), max(64, 64 * 5.0 / max(t_cal, 1e-3))))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        orc.tree_eval(it[:n_s], rij[:n_s], jt[:n_s], True, threads)
        times.append(time.perf_counter() - t0)
    tree_rate = n_s / float(np.mean(times))
    # compact oracle, whole workload, one pass (also the force reference)
    t0 = time.perf_counter()
    E, F, s6, Ei, g = orc.calc_energy_forces_stress_padded(it, js, rij, jt, 3, w.volume, w.natoms, w.natoms,
                                                           form="compact", nthreads=threads)
    t_comp = time.perf_counter() - t0
    ferr = float(np.max(np.abs(gpu_F - F)) / np.max(np.abs(F)))
    return dict(
        value=tree_rate, unit="atom-steps/s", cores=threads, kind="port",
        sample=f"first {n_s} of {w.natoms} atoms of the same workload, full-tensor JSON-tree FP64 restatement of "
               f"jax_backend.py (E + per-slot gradients), OpenMP, mean of 3 passes after a 64-atom calibration",
        compact_oracle=dict(value=w.natoms / t_comp, unit="atom-steps/s", cores=threads,
                            sample=f"all {w.natoms} atoms, one pass, E+F+stress"),
    ), ferr, F


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ncell", type=int, default=20, help="FCC cells per edge (20 -> 32,000 atoms)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-head", action="store_true", help="skip the distilled-head leg")
    ap.add_argument("--no-train", action="store_true", help="skip the packed training-step leg")
    ap.add_argument("--no-pair-style", action="store_true", help="skip the pair-style (host arrays) leg")
    ap.add_argument("--head-steps", type=int, default=3000)
    ap.add_argument("--head-save", default=None, help="directory to keep the trained head and its teacher")
    ap.add_argument("--head-infer", default=None,
                    help="only evaluate a head saved with --head-save (prints its JSON; for kernel profiles)")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the MTP evaluator's multi-GPU path (spatial domain decomposition "
              "with a halo exchange) is a design, not part of this build. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    jm = importlib.import_module("jax-mtp_amd")
    if args.head_infer:
        print(json.dumps(jm.distill.infer(args.head_infer, reps=args.steps)), flush=True)
        return
    w = jm.configs.fcc32k_level24(seed=args.seed, ncell=args.ncell)
    t0 = time.perf_counter()
    nb = w.neighbors()
    log(f"[bench] workload {w.name}: {w.natoms} atoms, neighbors built in {time.perf_counter() - t0:.2f}s, "
        f"mean n={nb['counts'].mean():.2f} max n={nb['counts'].max()}")
    dev = torch.device("cuda")
    it = torch.as_tensor(nb["itypes"], dtype=torch.int32, device=dev)
    js = torch.as_tensor(nb["js"], dtype=torch.int32, device=dev)
    rij = torch.as_tensor(nb["rij"], dtype=torch.float64, device=dev)
    jt = torch.as_tensor(nb["jtypes"], dtype=torch.int32, device=dev)
    A, K = js.shape
    pot = jm.MTPPotential(w.params, max_atoms=A, max_nbr=K)
    info = pot.info()
    best, tune_ms = pot.autotune(it, js, rij, jt, reps=3)
    log(f"[bench] autotune: variant {best} ({pot.VARIANTS[best]}); ms per variant {tune_ms}")
    L = jm.capi.lib()
    out = torch.empty(14, dtype=torch.float64, device=dev)
    F = torch.empty((A, 3), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def step():
        jm.capi.check(L.mtp_compute_padded_device(pot.ctx, A, K, it.data_ptr(), js.data_ptr(), rij.data_ptr(),
                                                  jt.data_ptr(), 3, w.volume, A, A, out.data_ptr(), F.data_ptr(),
                                                  None, sptr), pot.ctx)

    results = {}
    for variant in ("fp64", "radial_fp32"):
        pot.set_precision(variant == "radial_fp32")
        for _ in range(args.warmup):
            step()
        # timed steps: repeated calls with the same arrays replay the library's recorded HIP graph
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms_total = ev0.elapsed_time(ev1)
        # the same steps again with the atom kernel bracketed by HIP events on its stream (launched directly)
        import ctypes

        jm.capi.check(L.mtp_profile(pot.ctx, 1), pot.ctx)
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        ms_profiled = ev0.elapsed_time(ev1)
        kms = ctypes.c_double()
        nl = ctypes.c_int()
        jm.capi.check(L.mtp_profile_read(pot.ctx, ctypes.byref(kms), ctypes.byref(nl)), pot.ctx)
        jm.capi.check(L.mtp_profile(pot.ctx, 0), pot.ctx)
        status = float(out[13].item())
        if status != 0.0:
            raise RuntimeError("non-finite energy or virial")
        results[variant] = dict(ms_per_step=ms_total / args.steps, wall_ms_per_step=1e3 * wall / args.steps,
                                ms_per_step_direct=ms_profiled / args.steps,
                                atom_kernel_ms=kms.value / max(nl.value, 1), launches=nl.value,
                                F=F.cpu().numpy().copy(), E=float(out[0].item()))
        log(f"[bench] {variant}: {ms_total / args.steps:.3f} ms/step (events), atom kernel "
            f"{kms.value / max(nl.value, 1):.3f} ms, E={results[variant]['E']:.12e}")

    r = results["fp64"]
    flops = algorithmic_flops(nb["counts"][: w.natoms], info["radial_basis_size"], info["radial_funcs_count"],
                              info["basic_count"], info["times_count"], info["scalar_count"])
    achieved = flops / (r["atom_kernel_ms"] * 1e-3) / 1e12
    value = w.natoms / (r["ms_per_step"] * 1e-3)
    kname = {0: "atom_kernel<double, false>", 1: "atom_kernel_v2<double, 256>", 2: "atom_kernel_v2<double, 512>",
             3: "atom_kernel_v2<double, 64>", 4: "atom_kernel_pair<double, 1024>",
             5: "atom_kernel_pair<double, 512>", 6: "atom_kernel_v2<double, 128>"}[best]
    traffic = traffic_bytes(w.name, pot.VARIANTS[best])
    cnt = profile_values(f"atom_kernel_{w.name}_counters.txt", kname)
    kt = r["atom_kernel_ms"] * 1e-3
    roofs = {"fp64_valu": {"achieved_algorithmic": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                           "frac": achieved / PEAK_FP64_TFLOPS}}
    if traffic is not None:
        roofs["hbm"] = {"achieved": traffic / kt / 1e12, "peak": PEAK_HBM_TBS, "unit": "TB/s",
                        "frac": traffic / kt / 1e12 / PEAK_HBM_TBS}
    if cnt is not None:
        # counter-derived rates: the counts of the committed rocprofv3 passes over this time
        fc = float(cnt["fp64_flop_counter"])
        roofs["fp64_valu"].update(achieved_counter=fc / kt / 1e12, frac_counter=fc / kt / 1e12 / PEAK_FP64_TFLOPS,
                                  flop_counter_per_launch=fc)
        lb = float(cnt["lds_bytes"])
        roofs["lds"] = {"achieved": lb / kt / 1e12, "peak": PEAK_LDS_TBS, "unit": "TB/s",
                        "frac": lb / kt / 1e12 / PEAK_LDS_TBS, "busy_frac": float(cnt["lds_busy_frac"]),
                        "bytes_per_launch": lb}
        roofs["l2"] = {"achieved": float(cnt["l2_read_bytes"]) / kt / 1e12, "unit": "TB/s",
                       "bytes_per_launch": float(cnt["l2_read_bytes"])}
    nearest = max((k for k in roofs if "frac" in roofs[k]), key=lambda k: roofs[k]["frac"])
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "ms_per_step_direct_launches": r["ms_per_step_direct"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded FCC lattice, N(0,0.1) parameters)",
        "config": {"workload": w.name, "natoms": w.natoms, "level": info["level"], "species": 3,
                   "radial_basis_size": info["radial_basis_size"], "radial_funcs_count": info["radial_funcs_count"],
                   "cutoff": w.cutoff, "max_nbr": K, "mean_neighbors": float(nb["counts"].mean()),
                   "parallelism": "single GPU"},
        "roofline": {"bound": {"fp64_valu": "valu"}.get(nearest, nearest),
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "achieved_counter": roofs["fp64_valu"].get("achieved_counter"),
                     "lds": roofs.get("lds"), "roofs": roofs,
                     "note": "achieved/peak/frac: algorithmic FP64 FLOP/s of the atom kernel (HIP events, this run) "
                             "against the FP64 vector peak; bound: the roof with the largest fraction",
                     "kernel": f"atom_kernel ({pot.VARIANTS[best]})", "kernel_ms": r["atom_kernel_ms"],
                     "algorithmic_flop_per_launch": flops,
                     "algorithmic_bytes_per_launch": algorithmic_bytes(nb["counts"][: w.natoms])},
        "autotune": {"chosen": best, "kernel": pot.VARIANTS[best],
                     "ms_per_step_by_variant": {pot.VARIANTS[v]: tune_ms[v] for v in range(len(tune_ms))
                                                if tune_ms[v] >= 0}},
        "radial_fp32": {"value": w.natoms / (results["radial_fp32"]["ms_per_step"] * 1e-3),
                        "ms_per_step": results["radial_fp32"]["ms_per_step"],
                        "atom_kernel_ms": results["radial_fp32"]["atom_kernel_ms"]},
        "cpu_baseline": None,
    }
    if not args.no_pair_style:
        pot.set_precision(False)
        line["pair_style_step"] = pair_style_leg(jm, pot, w)
        line["pair_style_step"]["ms_per_step_device"] = r["ms_per_step"]
        log(f"[bench] pair style step: {line['pair_style_step']}")
    if not args.no_cpu_baseline:
        # every host CPU this process may run on (the OpenMP thread count when it is set)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        log(f"[bench] cpu baseline on {threads} threads")
        cb, ferr, Fref = cpu_baseline(w, nb, r["F"], threads)
        cb["cpu_model"] = cpu_model()
        cb["host_cpus_visible"] = os.cpu_count()
        cb["cpus_allowed"] = len(os.sched_getaffinity(0))
        line["cpu_baseline"] = cb
        f32 = results["radial_fp32"]["F"]
        line["force_err_vs_fp64_oracle"] = {
            "fp64": ferr,
            "radial_fp32": float(np.max(np.abs(f32 - Fref)) / np.max(np.abs(Fref))),
            "norm": "max|F-F_ref| / max|F_ref| over all atoms and components",
        }
    pot.close()
    if not args.no_train:
        torch.cuda.empty_cache()
        line["training_step"] = training_leg(args.seed)
        log(f"[bench] training: {line['training_step']}")
    if not args.no_head:
        torch.cuda.empty_cache()
        hs = args.head_save
        line["distilled_head"] = head_leg(args.seed, args.head_steps, hs and f"{hs}/level12", 12)
        line["distilled_head_level12_pairs"] = head_leg(args.seed, args.head_steps, hs and f"{hs}/level12_pairs", 12,
                                                        (3, 4, 5))
        line["distilled_head_level14"] = head_leg(args.seed, args.head_steps, hs and f"{hs}/level14", 14)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
