"""Island specifications of the uneven-archipelago workload, shared by the
benchmark and the full-size tests.

64 islands; island k has a population size drawn log-uniform in [32, 4096],
a dimension drawn log-uniform in [10, 500], the problem rosenbrock,
rastrigin or lennard_jones in turn (for lennard_jones the atom count is
max(3, round((dim + 6) / 3)), so nx = 3 atoms - 6), and pso_gen with the
default parameters seeded 42 + k.  Its initial population is
population(problem, bfe, pop, 42 + k) (population.cpp:82-103), drawn and
evaluated on the device by pgh_archi_create_seeded.
"""
from __future__ import annotations

import numpy as np

from ._capi import LENNARD_JONES, RASTRIGIN, ROSENBROCK, pso_params


def archipelago_islands(n_islands: int = 64, gens: int = 200, seed: int = 42):
    """List of {kind, dim, pop, seed, algo} for Archipelago.seeded."""
    rng = np.random.default_rng(seed)
    kinds = [ROSENBROCK, RASTRIGIN, LENNARD_JONES]
    isl = []
    for k in range(n_islands):
        kind = kinds[k % 3]
        pop = int(round(np.exp(rng.uniform(np.log(32), np.log(4096)))))
        dim = int(round(np.exp(rng.uniform(np.log(10), np.log(500)))))
        if kind == LENNARD_JONES:
            dim = max(3, round((dim + 6) / 3))  # atoms
        isl.append(dict(kind=kind, dim=dim, pop=pop, seed=42 + k, algo=pso_params(gen=gens, seed=42 + k)))
    return isl
