"""MLIP .cfg data (the reference's training/training_data/training.cfg:
21 DFT-labelled configurations) and training from it.

CPU: the reader against the file's own numbers, the write/read round trip,
direct coordinates and configurations without a cell, and physical pins of
the label conventions taken from the data itself (the force is minus the
derivative of the energy along the bond; the stress -PlusStress/V follows
sum r dE/dr / V, cfg.py:131-134).
GPU: the reference's training recipe (motep_jax_train_function.py:17-229:
untrained level-12 alpha tables, 1 species, min_dist 0.5, max_dist 5.0,
uniform initial parameters, the first 3/4 of the configurations, linear least
squares then L-BFGS on the weighted E/F/S loss) with the hand-written
gradient; the trained potential written as .mtp, reopened by the C++ reader,
reproduces the trainer's predictions.
"""
import pathlib

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
CFG = ROOT / "tests" / "golden" / "training.cfg"


def test_read_training_cfg(jm):
    c = jm.cfg_io.read_cfg(CFG)
    assert len(c) == 21
    assert all(x.natoms == 2 and set(x.types.tolist()) == {0} for x in c)
    assert c[0].energy == -6.54227372
    np.testing.assert_array_equal(c[0].positions, [[0, 0, 0], [0, 0, 0.65]])
    np.testing.assert_array_equal(c[0].forces, [[-0.0, -0.0, -4.867791], [0.0, 0.0, 4.867791]])
    np.testing.assert_array_equal(c[0].plus_stress, [-0.13529, -0.13529, 3.03673, 0.0, 0.0, 0.0])
    np.testing.assert_allclose(c[0].volume, 1000.0)
    np.testing.assert_allclose(c[0].stress, -np.array([-0.13529, -0.13529, 3.03673, 0, 0, 0]) / 1000.0)
    assert c[0].features == {"EFS_by": "VASP"}


def test_cfg_round_trip(jm, tmp_path):
    c = jm.cfg_io.read_cfg(CFG)
    jm.cfg_io.write_cfg(tmp_path / "x.cfg", c)
    d = jm.cfg_io.read_cfg(tmp_path / "x.cfg")
    for a, b in zip(c, d):
        np.testing.assert_array_equal(a.positions, b.positions)
        np.testing.assert_array_equal(a.forces, b.forces)
        np.testing.assert_array_equal(a.plus_stress, b.plus_stress)
        np.testing.assert_array_equal(a.cell, b.cell)
        assert a.energy == b.energy and a.features == b.features


def test_cfg_direct_coordinates_and_no_cell(jm):
    text = """BEGIN_CFG
 Size
    2
 Supercell
    4.0 0.0 0.0
    0.0 5.0 0.0
    0.0 0.0 6.0
 AtomData:  id type direct_x direct_y direct_z
    1 1 0.25 0.5 0.5
    2 0 0.0 0.0 0.0
 Energy
    -1.5
END_CFG
BEGIN_CFG
 Size
    1
 AtomData:  id type cartes_x cartes_y cartes_z
    1 0 1.0 2.0 3.0
END_CFG
"""
    a, b = jm.cfg_io.parse_cfg_text(text)
    np.testing.assert_allclose(a.positions, [[1.0, 2.5, 3.0], [0, 0, 0]])
    assert a.types.tolist() == [1, 0] and a.energy == -1.5 and a.forces is None and a.stress is None
    assert b.cell is None and np.isnan(b.volume) and b.positions.tolist() == [[1.0, 2.0, 3.0]]
    with pytest.raises(ValueError):
        jm.cfg_io.parse_cfg_text("BEGIN_CFG\n Size\n 1\n")


def test_label_conventions_from_the_data(jm):
    """Dimers along z: F_z of the second atom is -dE/dr (central differences of the energies), and the
    anisotropic part of the stress, (-PlusStress_zz + PlusStress_xx) / V, follows the bond's r dE/dr / V
    wherever the bond force is not small (the isotropic rest is the DFT cell's background)."""
    c = jm.cfg_io.read_cfg(CFG)
    r = np.array([x.positions[1, 2] - x.positions[0, 2] for x in c])
    E = np.array([x.energy for x in c])
    Fz = np.array([x.forces[1, 2] for x in c])
    dEdr = (E[2:] - E[:-2]) / (r[2:] - r[:-2])
    np.testing.assert_allclose(-dEdr, Fz[1:-1], atol=0.03 * np.abs(Fz).max())
    for x, rr, f in zip(c, r, Fz):
        if abs(f) > 1.0:
            assert x.stress[2] - x.stress[0] == pytest.approx(rr * (-f) / x.volume, rel=0.05)


@pytest.mark.gpu
def test_gpu_train_on_reference_cfg(jm, tmp_path):
    print("[case] training on training.cfg, level 12, 1 species", flush=True)
    trained, res = jm.training.train_cfg(CFG, level=12, steps=60, log=print)
    assert res["loss_final"] < res["loss_initial_after_lls"]
    assert res["train"]["force_rmse_meV_per_A"] < 100.0
    # the written potential, reopened through the C++ reader, predicts what the trainer's parameters do
    path = tmp_path / "trained.mtp"
    jm.write_mtp(path, trained)
    c = jm.cfg_io.read_cfg(CFG)
    b, E, F, S = jm.training.batch_from_cfg(c, 5.0, 64)
    pa = jm.MTPPotential(trained, max_atoms=b.natoms, max_nbr=64)
    pb = jm.MTPPotential(str(path), max_atoms=b.natoms, max_nbr=64)
    Ea, Fa, Sa = (t.cpu().numpy() for t in jm.training.batch_forward(pa, b.to_device()))
    Eb, Fb, Sb = (t.cpu().numpy() for t in jm.training.batch_forward(pb, b.to_device()))
    np.testing.assert_allclose(Eb, Ea, rtol=1e-12)
    np.testing.assert_allclose(Fb, Fa, rtol=1e-10, atol=1e-12 * np.abs(Fa).max())
    assert jm.read_mtp(path).level == 12
    pa.close()
    pb.close()
