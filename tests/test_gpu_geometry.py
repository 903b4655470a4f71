"""The alternating SCF + Adam geometry loop on the GPU against the oracle's
(geometry_opt.py:28-78): same start, same SCF iteration counts, same Adam,
positions and energies compared step by step."""
import numpy as np
import pytest
import torch

from oracle import gato_oracle as O

pytestmark = pytest.mark.gpu


def test_water_geometry_trajectory_matches_oracle(gpu):
    """H2O at 24^3 (L = 10, HGH O/H) from the equilibrium geometry plus a
    seeded displacement of up to 0.3 bohr per coordinate (no symmetry, so no
    force component sits at rounding level where Adam's g/sqrt(g^2) would
    amplify the noise).  Three Adam steps (lr 0.05), each followed by a
    warm-started FP64 SCF of a fixed number of iterations on both sides.
    Positions after every step agree to 1e-8 bohr and energies to 1e-8 Ha."""
    import gato_amd as G
    from gato_amd.geometry import Nuclei
    from gato_amd.geometry_opt import optimize_geometry_alternating
    from gato_amd.physics.water import (hellmann_feynman_force, solve_rhf_at_geometry,
                                        water_hgh_elements, water_initial_geometry)
    N, L, lanczos_iters = 24, 10.0, 72
    iters_first, iters_warm = 4, 2
    disp = np.random.default_rng(3).uniform(-0.3, 0.3, (3, 3))

    g_o = O.Grid3D(N, L)
    el_o = O.water_hgh_elements()
    base_o = O.water_initial_geometry(1.81, 104.5)

    class OPoint:
        def __init__(self, nuclei, res, E):
            self.nuclei, self.orbitals, self.energy = nuclei, res.orbitals, E

    def solve_o(nuclei, prev):
        res, E, _ = O.solve_rhf_at_geometry(
            nuclei, el_o, g_o, max_iters=iters_first if prev is None else iters_warm, tol=0.0,
            mixing=0.5, order=4, lanczos_iters=lanczos_iters,
            initial_orbitals=None if prev is None else prev.orbitals)
        return OPoint(nuclei, res, E)

    def grad_o(p):
        return -O.hellmann_feynman_force(p.nuclei.positions, p.nuclei.charges, p.orbitals,
                                         el_o, g_o)

    _, traj_o = O.optimize_geometry_alternating(O.Nuclei(base_o.positions + disp, base_o.charges),
                                                solve_at_geometry=solve_o, bo_grad=grad_o,
                                                geom_steps=3, learning_rate=0.05)

    gg = G.Grid3D(N, L)
    el = water_hgh_elements()
    base = water_initial_geometry(1.81, 104.5)
    nuc = Nuclei(base.positions + torch.tensor(disp, dtype=base.positions.dtype), base.charges)

    def solve_g(nuclei, prev):
        return solve_rhf_at_geometry(nuclei, el, gg,
                                     max_iters=iters_first if prev is None else iters_warm,
                                     tol=0.0, mixing=0.5, order=4, lanczos_iters=lanczos_iters,
                                     initial_orbitals=None if prev is None else prev.orbitals)

    def grad_g(p):
        return -hellmann_feynman_force(p, el, gg)

    _, traj_g = optimize_geometry_alternating(nuc, solve_at_geometry=solve_g, bo_grad=grad_g,
                                              geom_steps=3, learning_rate=0.05)
    assert len(traj_g) == len(traj_o) == 4
    for s, (pg, po) in enumerate(zip(traj_g, traj_o)):
        dpos = np.max(np.abs(pg.nuclei.positions.detach().cpu().numpy() - po.nuclei.positions))
        dE = abs(pg.energy - po.energy)
        print(f"step {s}: |dR|max = {dpos:.2e} bohr  |dE| = {dE:.2e} Ha  E = {pg.energy:+.10f}",
              flush=True)
        assert dpos < 1e-8, s
        assert dE < 1e-8, s
    # the loop moved the nuclei (so the comparison is not of a fixed point)
    moved = np.max(np.abs(traj_o[-1].nuclei.positions - traj_o[0].nuclei.positions))
    assert moved > 0.05
