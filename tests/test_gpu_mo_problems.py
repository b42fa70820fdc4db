"""ZDT and DTLZ on the device against the oracle and the reference's known
answers (tests/zdt.cpp:67-176, tests/dtlz.cpp:69-182).

Bars: ZDT1, ZDT2 and ZDT5 bit-exact (only + - * / sqrt round); ZDT3, ZDT4,
ZDT6 and every DTLZ within 1e-12 relative per objective (the device sin, cos,
exp and pow against glibc's), with 1e-300 absolute for subnormal results.
"""
import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"
EXACT_ZDT = (1, 2, 5)


def gpu_mo(kind, prob_id, dim, x, nobj=2, alpha=100):
    x = np.ascontiguousarray(np.atleast_2d(x), np.float64)
    nx, nf = P.dims(kind, dim, nobj=nobj, prob_id=prob_id, alpha=alpha)
    assert x.shape[1] == nx
    f = torch.empty((x.shape[0], nf), dtype=torch.float64, device=DEV)
    P.bfe_eval(kind, dim, torch.from_numpy(x).to(DEV), f, nobj=nobj, prob_id=prob_id, alpha=alpha)
    torch.cuda.synchronize()
    return f.cpu().numpy()


def close(a, b, rel=1e-12):
    # values below 1e-300 (DTLZ4's x^alpha underflows into subnormals, where
    # one unit of the last place is no longer relative) compare absolutely
    return np.all(np.abs(a - b) <= rel * np.abs(b) + 1e-300)


def kat_x(spec):
    if "fill" in spec:
        n, v = spec["fill"]
        return np.full(n, v)
    n, v, at, w = spec["fill2"]
    x = np.full(n, v)
    x[at:] = w
    return x


def test_zdt_kats(kats):
    for c in kats["zdt_fitness"]["cases"]:
        f = gpu_mo(P.ZDT, c["prob_id"], c["param"], kat_x(c["x"]))[0]
        o = O.mo_fitness(O.ZDT, c["prob_id"], c["param"], kat_x(c["x"]))[0]
        if c["prob_id"] in EXACT_ZDT:
            assert np.array_equal(f, o), c
        assert close(f, np.array(c["f"])), c


def test_dtlz_kats(kats):
    for c in kats["dtlz_fitness"]["cases"]:
        f = gpu_mo(P.DTLZ, c["prob_id"], c["dim"], np.array(c["x"]), c["fdim"], c["alpha"])[0]
        assert close(f, np.array(c["f"])), c


@pytest.mark.parametrize("prob_id", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("param,n", [(2, 5), (30, 4097), (100, 300)])
def test_zdt_random(prob_id, param, n):
    _, _, lb, ub = O.mo_info(O.ZDT, prob_id, param)
    x = O.batch_uniform(lb, ub, n, 7 * prob_id + param)
    f = gpu_mo(P.ZDT, prob_id, param, x)
    o = O.mo_fitness(O.ZDT, prob_id, param, x)
    if prob_id in EXACT_ZDT:
        assert np.array_equal(f, o)
    else:
        assert close(f, o)


@pytest.mark.parametrize("param,n", [(2, 100), (11, 1000), (30, 513)])
def test_zdt5_random_bits(param, n):
    nx, _, _, _ = O.mo_info(O.ZDT, 5, param)
    rng = np.random.default_rng(param)
    bits = rng.integers(0, 2, (n, nx)).astype(np.float64)
    relaxed = bits + rng.uniform(-0.49, 0.49, (n, nx))  # rounds back to the bits
    halves = np.where(rng.uniform(0, 1, (n, nx)) < 0.5, 0.5, 1.5)  # ties: round half away from zero
    for x in (bits, relaxed, halves):
        assert np.array_equal(gpu_mo(P.ZDT, 5, param, x), O.mo_fitness(O.ZDT, 5, param, x))


@pytest.mark.parametrize("prob_id", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("dim,fdim,alpha", [(5, 3, 100), (12, 2, 100), (20, 5, 7), (30, 10, 100), (4, 3, 1)])
def test_dtlz_random(prob_id, dim, fdim, alpha):
    x = O.batch_uniform(np.zeros(dim), np.ones(dim), 777, prob_id * 100 + dim)
    f = gpu_mo(P.DTLZ, prob_id, dim, x, fdim, alpha)
    o = O.mo_fitness(O.DTLZ, prob_id, dim, x, fdim, alpha)
    assert close(f, o)


def test_host_entry_point_and_checks(kats):
    x = np.full((3, 30), 0.25)
    f = P.bfe_eval_host(P.ZDT, 30, x, prob_id=1)
    assert np.array_equal(f, O.mo_fitness(O.ZDT, 1, 30, x))
    for c in kats["dtlz_throws"]:
        with pytest.raises(P.PghError) as e:
            P.bfe_eval_host(P.DTLZ, c["dim"], np.zeros((1, max(c["dim"], 1))), nobj=c["fdim"],
                            prob_id=c["prob_id"], alpha=100)
        assert str(e.value) == c["message"]
