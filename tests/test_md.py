"""NVE dynamics through the pair style's host path (configuration: 2889-atom
BCC, level 16, 2 species, 1000 steps of 1 fs).

The standalone driver jax-mtp_amd/md/md_nve does per step what LAMMPS with
pair_style jax/mtp does: periodic ghosts carrying their owners' tags, a full
neighbor list, the pair style's packing (pair_jax_mtp.cpp:96-173, tag fold
of ghosts onto owners), mtp_compute() on the GPU, and the velocity-Verlet
half-kicks and drift.  Checked: total-energy conservation over the run, and
the forces dumped along the trajectory against the FP64 oracle evaluated on
the dumped positions (1e-10 relative).
"""
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import oracle_for

ROOT = pathlib.Path(__file__).resolve().parent.parent
MD = ROOT / "jax-mtp_amd" / "md" / "md_nve"


def _read_dump(path):
    raw = path.read_bytes()
    n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    o = 4
    L = np.frombuffer(raw[o:o + 24], dtype=np.float64); o += 24
    t = np.frombuffer(raw[o:o + 4 * n], dtype=np.int32); o += 4 * n
    x = np.frombuffer(raw[o:o + 24 * n], dtype=np.float64).reshape(n, 3); o += 24 * n
    F = np.frombuffer(raw[o:o + 24 * n], dtype=np.float64).reshape(n, 3)
    return L, t, x, F


def test_driver_is_built():
    assert MD.exists(), "build with python jax-mtp_amd/build.py"


@pytest.mark.gpu
def test_gpu_nve_energy_drift_and_force_parity(jm, O, tmp_path):
    p = jm.configs.bcc2889_nve_params(seed=0)
    mtp = tmp_path / "nve16.mtp"
    jm.write_mtp(mtp, p)
    steps, dump_every = 1000, 250
    print(f"[case] md_nve 2889 atoms level 16, {steps} steps", flush=True)
    r = subprocess.run([str(MD), str(mtp), "bcc", "3.2", "11", "11", "12", "2889", str(steps), "0.001", "300",
                        "0.08", "0", str(tmp_path / "run"), str(dump_every), "58.6934", "26.9815"],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-400:], r.stderr[-400:], flush=True)
    assert r.returncode == 0
    th = np.loadtxt(tmp_path / "run.thermo")
    assert th.shape[0] == steps + 1
    n = 2889
    etot = th[:, 3]
    ke = th[:, 2]
    drift = float(np.max(np.abs(etot - etot[0])) / n)
    slope = float(np.polyfit(th[:, 0] * 1e-3, etot / n, 1)[0])  # eV/atom/ps
    print(f"  E_tot drift max {drift:.3e} eV/atom, slope {slope:.3e} eV/atom/ps, <KE>/atom {ke.mean() / n:.4f} eV",
          flush=True)
    assert drift < 2e-4 * max(1.0, ke.mean() / n / 0.0388)
    # forces along the trajectory vs the oracle on the same potential file
    orc = O.Oracle(O.read_mtp(str(mtp)))
    for step in range(0, steps + 1, dump_every):
        L, t, x, F = _read_dump(tmp_path / f"run.dump{step:06d}")
        nb = jm.neighbors.padded_neighbors(x, t - 1, L, p.max_dist, 64)
        _, F_ref, _, _, _ = orc.calc_energy_forces_stress_padded(nb["itypes"], nb["js"], nb["rij"], nb["jtypes"], 3,
                                                                 float(np.prod(L)), n, n)
        err = float(np.max(np.abs(F - F_ref)) / np.max(np.abs(F_ref)))
        print(f"  step {step}: force rel err {err:.3e} (max |F| {np.abs(F_ref).max():.3f} eV/A)", flush=True)
        assert err <= 1e-10
