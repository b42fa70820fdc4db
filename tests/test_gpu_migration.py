"""Migration on the device (pgh_archi_set_migration) against the oracle's
archipelago rounds (island.cpp:452-642 with select_best and fair_replace):
final populations and the migrants database, in batched and per-island
modes, for p2p / broadcast, preserve / evict, absolute and fractional rates,
weights below one (the migration engines' draws decide), and islands of
different problems sharing one dimension.

Bars: decision vectors bit-exact (they depend on fitness only through
comparisons); fitness values at the kernels' bars (tests/parity_bars.py):
bit-exact for rosenbrock, 1e-12 relative for the separable problems, 1e-12 of
the sum of the absolute pair terms for Lennard-Jones."""
import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P
from parity_bars import assert_fitness

pytestmark = pytest.mark.gpu


def islands(spec, gen, sel, rep):
    out = []
    for k, (kind, dim, pop) in enumerate(spec):
        X, F = O.population_init(kind, dim, pop, 1000 + k)
        out.append(dict(kind=kind, dim=dim, X=X, F=F, seed=50 + k, gen=gen, sel=sel, rep=rep, migr_seed=900 + k))
    return out


def same(kind, got, ref):
    gX, gF = got
    rX, rF = ref
    assert np.array_equal(gX, rX)
    if gX.shape[0]:
        assert_fitness(kind, gX, gF, rF)


def device(isl, conns, mtype, handling, rounds, mode):
    a = P.Archipelago([dict(kind=d["kind"], dim=d["dim"], X0=d["X"], F0=d["F"],
                            algo=P.pso_params(gen=d["gen"], seed=d["seed"])) for d in isl])
    try:
        a.set_migration(conns, [d["sel"] for d in isl], [d["rep"] for d in isl], [d["migr_seed"] for d in isl],
                        mtype, handling)
        for _ in range(rounds):
            a.evolve(mode)
        return [a.get(k) for k in range(len(isl))], [a.migrants(k) for k in range(len(isl))]
    finally:
        a.close()


SPEC10 = [(P.RASTRIGIN, 10, 20), (P.ROSENBROCK, 10, 33), (P.RASTRIGIN, 10, 64), (P.ACKLEY, 10, 17),
          (P.ROSENBROCK, 10, 40), (P.GRIEWANK, 10, 25)]

CASES = [
    ("ring", 1.0, P.MIGRATION_P2P, P.MIGRANTS_PRESERVE, 2, 2),
    ("ring", 0.6, P.MIGRATION_P2P, P.MIGRANTS_EVICT, 0.1, 0.5),
    ("fully_connected", 0.5, P.MIGRATION_BROADCAST, P.MIGRANTS_PRESERVE, 3, 3),
    ("fully_connected", 1.0, P.MIGRATION_BROADCAST, P.MIGRANTS_EVICT, 1, 1.0),
    ("ring", 0.3, P.MIGRATION_BROADCAST, P.MIGRANTS_PRESERVE, 0.25, 4),
]


@pytest.mark.parametrize("topo,w,mtype,handling,sel,rep", CASES)
def test_migration_matches_oracle(topo, w, mtype, handling, sel, rep):
    isl = islands(SPEC10, 6, sel, rep)
    conns = P.ring_connections(len(isl), w) if topo == "ring" else P.fully_connected_connections(len(isl), w)
    print(f"migration {topo} w={w} type={mtype} handling={handling} sel={sel} rep={rep}", flush=True)
    ref, ref_db = O.archi_migrate(isl, conns, mtype, handling, rounds=4)
    got, got_db = device(isl, conns, mtype, handling, 4, P.ARCHI_BATCHED)
    # migrants keep the fitness their source island computed, so an island's
    # rows mix the problems of SPEC10: the separable problems' 1e-12 relative
    # bar covers them all (rosenbrock rows are bit-exact)
    for k in range(len(isl)):
        same(P.RASTRIGIN, got[k], ref[k])
        same(P.RASTRIGIN, got_db[k], ref_db[k])


def test_per_island_mode_and_lennard_jones():
    # LJ islands (5 atoms, nx 9) in a ring, per-island launches
    spec = [(P.LENNARD_JONES, 5, 24), (P.LENNARD_JONES, 5, 40), (P.LENNARD_JONES, 5, 31)]
    isl = islands(spec, 5, 2, 3)
    conns = P.ring_connections(3, 0.7)
    ref, ref_db = O.archi_migrate(isl, conns, P.MIGRATION_P2P, P.MIGRANTS_PRESERVE, rounds=3)
    for mode in (P.ARCHI_PER_ISLAND, P.ARCHI_BATCHED):
        got, got_db = device(isl, conns, P.MIGRATION_P2P, P.MIGRANTS_PRESERVE, 3, mode)
        for k in range(3):
            same(P.LENNARD_JONES, got[k], ref[k])
            same(P.LENNARD_JONES, got_db[k], ref_db[k])


def test_unconnected_islands_with_migration_set():
    # an island with no inbound connection evolves as without migration;
    # its database entry is still filled (island.cpp:599-620)
    isl = islands(SPEC10[:3], 4, 2, 2)
    conns = [[], [(0, 1.0)], []]
    ref, ref_db = O.archi_migrate(isl, conns, P.MIGRATION_P2P, P.MIGRANTS_PRESERVE, rounds=2)
    got, got_db = device(isl, conns, P.MIGRATION_P2P, P.MIGRANTS_PRESERVE, 2, P.ARCHI_BATCHED)
    for k in range(3):
        same(P.RASTRIGIN, got[k], ref[k])
        same(P.RASTRIGIN, got_db[k], ref_db[k])


def test_migration_off_again():
    isl = islands(SPEC10[:2], 3, 1, 1)
    a = P.Archipelago([dict(kind=d["kind"], dim=d["dim"], X0=d["X"], F0=d["F"],
                            algo=P.pso_params(gen=d["gen"], seed=d["seed"])) for d in isl])
    a.set_migration(P.ring_connections(2))
    a.set_migration(None)
    a.evolve()
    for k, d in enumerate(isl):
        same(d["kind"], a.get(k), O.pso_gen_evolve(d["kind"], d["dim"], d["X"], d["F"], d["gen"], d["seed"]))
    assert a.migrants(0)[1].size == 0
    a.close()


def test_messages(kats):
    isl = islands(SPEC10[:3], 1, 1, 1)
    a = P.Archipelago([dict(kind=d["kind"], dim=d["dim"], X0=d["X"], F0=d["F"],
                            algo=P.pso_params(gen=1, seed=1)) for d in isl])
    ring = P.ring_connections(3)
    with pytest.raises(P.PghError) as e:
        a.set_migration(ring, select_rate=[100, 1, 1])
    assert str(e.value) == ("The absolute migration rate (100) in a 'Select best' selection policy is larger than "
                            "the number of input individuals (20)")
    with pytest.raises(P.PghError) as e:
        a.set_migration(ring, replace_rate=[1, 34, 1])
    assert str(e.value) == ("The absolute migration rate (34) in a 'fair_replace' replacement policy is larger than "
                            "the number of input individuals (33)")
    with pytest.raises(P.PghError) as e:
        a.set_migration(ring, replace_rate=2.0)
    assert str(e.value).startswith(kats["migration_throws"][3]["message_prefix"])
    with pytest.raises(P.PghError) as e:
        a.set_migration(P.ring_connections(3, 1.5))
    assert str(e.value) == kats["topology"]["weight_messages"][1]["message"]
    a.close()
    # connected islands of different dimension
    b = P.Archipelago([dict(kind=P.RASTRIGIN, dim=n, X0=np.zeros((4, n)), F0=np.zeros(4),
                            algo=P.pso_params(gen=1, seed=1)) for n in (3, 4)])
    with pytest.raises(P.PghError):
        b.set_migration(P.ring_connections(2))
    b.close()
    c = P.Archipelago([dict(kind=P.RASTRIGIN, dim=3, X0=np.zeros((4, 3)), F0=np.zeros(4),
                            algo=P.pso_params(gen=1, seed=1, memory=True))])
    with pytest.raises(P.PghError) as e:
        c.set_migration([[]])
    assert e.value.status == P.UNSUPPORTED
    c.close()
