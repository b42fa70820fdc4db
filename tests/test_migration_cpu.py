"""Migration on the CPU: the oracle's select_best, fair_replace, topologies and
archipelago rounds against the reference's known answers
(tests/select_best.cpp:143-175, tests/fair_replace.cpp:143-178,
tests/ring.cpp:50-73, tests/fully_connected.cpp:74-84) and against a
plain-Python restatement of the round on cases whose draws are forced; and the
product's connection lists against the oracle's."""
import math

import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P


def group(g):
    return np.array(g["ids"], np.uint64), np.array(g["x"], np.float64), np.array(g["f"], np.float64)


@pytest.mark.parametrize("stable", [False, True])
def test_select_best_kats(kats, stable):
    ids, X, F = group(kats["select_best"]["inds"])
    for c in kats["select_best"]["cases"]:
        oi, _, oF = O.select_best(ids, X, F, c["rate"], stable)
        assert oi.tolist() == c["ids"] and oF.tolist() == c["f"], c


@pytest.mark.parametrize("stable", [False, True])
def test_fair_replace_kats(kats, stable):
    ids, X, F = group(kats["fair_replace"]["inds"])
    mids, mX, mF = group(kats["fair_replace"]["mig"])
    for c in kats["fair_replace"]["cases"]:
        oi, _, oF = O.fair_replace(ids, X, F, mids, mX, mF, c["rate"], stable)
        assert oi.tolist() == c["ids"] and oF.tolist() == c["f"], c


def test_policy_messages(kats):
    for c in kats["migration_throws"]:
        n = c["n"]
        ids, X, F = np.arange(n, dtype=np.uint64), np.zeros((n, 1)), np.arange(n, dtype=np.float64)
        with pytest.raises(O.OracleError) as e:
            if c["policy"] == "select_best":
                O.select_best(ids, X, F, c["rate"])
            else:
                O.fair_replace(ids, X, F, ids, X, F, c["rate"])
        if "message" in c:
            assert str(e.value) == c["message"]
        else:
            assert str(e.value).startswith(c["message_prefix"])


def test_topology_kats(kats):
    # ring.cpp:56-73: two inbound connections (one for two islands), the
    # previous and the next island, every weight w
    for s in range(0, 9):
        for w in (1.0, 0.25):
            r = O.topology("ring", s, w)
            assert r == P.ring_connections(s, w)
            for i, c in enumerate(r):
                if s < 2:
                    assert c == []
                    continue
                assert len(c) == (2 if s > 2 else 1)
                assert all(x == w for _, x in c)
                srcs = [a for a, _ in c]
                assert (i + 1) % s in srcs and (i - 1) % s in srcs
            f = O.topology("fully_connected", s, w)
            assert f == P.fully_connected_connections(s, w)
            for i, c in enumerate(f):
                assert [a for a, _ in c] == [j for j in range(s) if j != i]
                assert all(x == w for _, x in c)
    for m in kats["topology"]["weight_messages"]:
        with pytest.raises(O.OracleError) as e:
            O.topology("ring", 3, m["w"])
        msg = str(e.value)
        assert msg == m["message"] if "message" in m else msg.startswith(m["message_prefix"])


def test_ring_order_small():
    # the BGL insertion order for 3 and 5 islands, worked by hand from
    # ring::push_back: island 0 hears from 1 first, the last island from its
    # left neighbour first
    assert [[a for a, _ in c] for c in O.topology("ring", 3)] == [[1, 2], [0, 2], [1, 0]]
    assert [[a for a, _ in c] for c in O.topology("ring", 5)] == [[1, 4], [0, 2], [1, 3], [2, 4], [3, 0]]


# ---- archipelago rounds against a plain-Python restatement -------------------

def key(f):
    return (1, 0.0) if math.isnan(f) else (0, f)


def py_count(rate, n):
    if isinstance(rate, float):
        return min(int(rate * n), n)
    return rate


def py_select_best(pop, rate):
    order = sorted(range(len(pop)), key=lambda i: key(pop[i][1]))
    return [pop[i] for i in order[:py_count(rate, len(pop))]]


def py_fair_replace(pop, mig, rate):
    n_migr = min(py_count(rate, len(pop)), len(mig))
    morder = sorted(range(len(mig)), key=lambda i: key(mig[i][1]))
    merged = pop + [mig[i] for i in morder[:n_migr]]
    order = sorted(range(len(merged)), key=lambda i: key(merged[i][1]))
    return [merged[i] for i in order[:len(pop)]]


def py_rounds(pops, conns, mtype, handling, rounds, sel, rep):
    """Migration-only rounds (no evolution) with every weight 1 and, for p2p,
    one inbound connection per island, so that no draw decides anything."""
    db = [[] for _ in pops]
    for _ in range(rounds):
        for k, c in enumerate(conns):
            if not c:
                continue
            srcs = [a for a, _ in c]
            if mtype == 0:
                assert len(srcs) == 1
            mig = []
            for s in srcs:
                mig += list(db[s])
                if handling == 1:
                    db[s] = []
            pops[k] = py_fair_replace(pops[k], mig, rep)
        for k in range(len(pops)):
            db[k] = py_select_best(pops[k], sel)
    return pops, db


def make_islands(pops, nx, seed, gen):
    rng = np.random.default_rng(seed)
    isl = []
    for k, p in enumerate(pops):
        X = rng.uniform(-5, 5, (p, nx))
        F = rng.permutation(p * 7)[:p].astype(np.float64) + rng.uniform(0, 0.5, p)  # distinct values
        isl.append(dict(kind=O.RASTRIGIN, dim=nx, X=X, F=F, seed=100 + k, gen=gen, sel=3, rep=2,
                        migr_seed=7 + k))
    return isl


@pytest.mark.parametrize("mtype,handling,topo", [(0, 0, "directed"), (0, 1, "directed"), (1, 0, "ring"),
                                                 (1, 1, "fully_connected"), (1, 0, "fully_connected")])
@pytest.mark.parametrize("sel,rep", [(3, 2), (0.25, 0.5), (1, 1.0)])
def test_migration_rounds_forced(mtype, handling, topo, sel, rep):
    pops = [9, 14, 6, 11, 8]
    isl = make_islands(pops, 4, 5, gen=0)
    for d in isl:
        d["sel"], d["rep"] = sel, rep
    n = len(isl)
    if topo == "directed":
        conns = [[((k - 1) % n, 1.0)] for k in range(n)]
    else:
        conns = O.topology(topo, n, 1.0)
    out, db = O.archi_migrate(isl, conns, mtype, handling, rounds=3)
    py_pops = [[(isl[k]["X"][i].copy(), isl[k]["F"][i]) for i in range(pops[k])] for k in range(n)]
    exp, exp_db = py_rounds(py_pops, conns, mtype, handling, 3, sel, rep)
    for k in range(n):
        assert np.array_equal(out[k][0], np.array([x for x, _ in exp[k]]).reshape(pops[k], -1))
        assert np.array_equal(out[k][1], np.array([f for _, f in exp[k]]))
        assert db[k][1].tolist() == [f for _, f in exp_db[k]]
        assert np.array_equal(db[k][0], np.array([x for x, _ in exp_db[k]]).reshape(len(exp_db[k]), -1))


def test_broadcast_with_zero_weights_sorts():
    # broadcast runs fair_replace even when every draw fails
    # (island.cpp:523-597): each population comes back sorted, nothing moves
    isl = make_islands([7, 5, 9], 3, 11, gen=0)
    conns = O.topology("fully_connected", 3, 0.0)
    out, db = O.archi_migrate(isl, conns, 1, 0, rounds=1)
    for k, d in enumerate(isl):
        order = np.argsort(d["F"], kind="stable")
        assert np.array_equal(out[k][1], d["F"][order])
        assert np.array_equal(out[k][0], d["X"][order])
        assert np.array_equal(db[k][1], d["F"][order][:3])
    # p2p with failing draws leaves the populations as they were
    out, _ = O.archi_migrate(isl, conns, 0, 0, rounds=1)
    for k, d in enumerate(isl):
        assert np.array_equal(out[k][1], d["F"])


def test_migration_preserves_population_sizes_and_best():
    # with evolution on: fair_replace never loses the best individual seen,
    # and the database holds select_best of the final populations
    isl = []
    for k, p in enumerate([12, 20, 16]):
        X, F = O.population_init(O.ROSENBROCK, 6, p, 40 + k)
        isl.append(dict(kind=O.ROSENBROCK, dim=6, X=X, F=F, seed=k, gen=4, sel=2, rep=2, migr_seed=k))
    out, db = O.archi_migrate(isl, O.topology("ring", 3, 1.0), 0, 0, rounds=4)
    for k, (X, F) in enumerate(out):
        assert X.shape == isl[k]["X"].shape
        assert np.all(np.diff(db[k][1]) >= 0)
        assert np.array_equal(db[k][1], np.sort(F)[:2])
