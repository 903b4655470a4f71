"""The uneven-batch workload description (host logic, no GPU)."""
import numpy as np
import pytest

from gato_amd.batch import (BOX, GRID_SIZES, KINDS, R_EQUILIBRIUM, BatchMolecule,
                            draw_uneven_batch)
from gato_amd.solvers.lanczos import default_krylov_dim


def test_draw_is_seeded_and_covers_every_kind_and_grid():
    a = draw_uneven_batch(64, 0)
    assert a == draw_uneven_batch(64, 0)
    assert a != draw_uneven_batch(64, 1)
    assert len(a) == 64
    assert {m.kind for m in a} == set(KINDS)
    assert {m.N for m in a} == set(GRID_SIZES)


def test_bond_lengths_within_15_percent_of_equilibrium():
    for m in draw_uneven_batch(256, 3):
        f = m.R / R_EQUILIBRIUM[m.kind]
        assert 0.85 <= f <= 1.15
        assert m.L == BOX[m.kind]


def test_krylov_dimension_follows_driver_rules():
    # HCl on N >= 80 needs 200 vectors for its 3p manifold (hcl.py:40-51)
    assert BatchMolecule("HCl", 80, 2.41).lanczos_iters == 200
    assert BatchMolecule("HCl", 128, 2.41).lanczos_iters == 200
    assert BatchMolecule("HCl", 64, 2.41).lanczos_iters == default_krylov_dim(64, 4)
    assert BatchMolecule("H2", 96, 1.4).lanczos_iters == default_krylov_dim(96, 1)


def test_geometry_along_z_about_the_origin():
    nuc = BatchMolecule("HCl", 64, 2.5).nuclei()
    P = nuc.positions.numpy()
    assert np.allclose(P[:, :2], 0.0)
    assert P[1, 2] - P[0, 2] == pytest.approx(2.5)
    assert nuc.charges.tolist() == [1.0, 17.0]
    assert BatchMolecule("LiH", 64, 3.0).nuclei().charges.tolist() == [3.0, 1.0]


def test_empty_batch_rejected():
    with pytest.raises(ValueError):
        draw_uneven_batch(0)


def test_auto_eigensolver_picks_block_lanczos_for_degenerate_hcl():
    from gato_amd.batch import eigensolver_for
    # mixed precision: scf_rhf's default (seeded block Lanczos, residual control)
    for kind in ("HCl", "LiH", "H2"):
        assert eigensolver_for(BatchMolecule(kind, 96, 2.0), "auto") == "auto"
        assert eigensolver_for(BatchMolecule(kind, 96, 2.0), "auto", "mixed") == "auto"
    # other precisions: block Lanczos for HCl's degenerate pair, scalar otherwise
    assert eigensolver_for(BatchMolecule("HCl", 80, 2.41), "auto", "fp64") == "block_lanczos"
    assert eigensolver_for(BatchMolecule("HCl", 128, 2.41), "auto", "fp32") == "block_lanczos"
    assert eigensolver_for(BatchMolecule("HCl", 64, 2.41), "auto", "fp64") == "block_lanczos"
    assert eigensolver_for(BatchMolecule("LiH", 128, 3.0), "auto", "fp64") == "lanczos"
    assert eigensolver_for(BatchMolecule("H2", 96, 1.4), "block_lanczos") == "block_lanczos"
