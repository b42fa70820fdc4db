"""pgh_local_descent_noalloc (the work buffer in the library's own device
array) against pgh_local_descent (the work buffer given by the caller), both
through the C ABI: bit for bit on the same inputs, two calls in a row with the
second batch smaller than the first, so that the second call runs over what
the first left in the array.  Rosenbrock is also checked against the oracle
(its fitness and gradient kernels are bit-exact with it).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def run_both(kind, dim, x0, iters, step0, lb, ub):
    """(x, f) of the work-buffer entry point and of the no-alloc one."""
    out = []
    for noalloc in (False, True):
        xd = dev(x0)
        f = torch.empty(x0.shape[0], dtype=torch.float64, device=DEV)
        dlb = None if lb is None else dev(lb)
        dub = None if ub is None else dev(ub)
        if noalloc:
            P.local_descent_noalloc(kind, dim, xd, f, iters, step0, dlb, dub)
        else:
            P.local_descent(kind, dim, xd, f, iters, step0, dlb, dub)
        torch.cuda.synchronize()
        out.append((xd.cpu().numpy(), f.cpu().numpy()))
    return out


@pytest.mark.parametrize("kind,dim,sizes,iters,step0", [
    (P.ROSENBROCK, 30, (4096, 1000), 12, 0.5),
    (P.RASTRIGIN, 1000, (2000, 333), 6, 0.05),
    (P.LENNARD_JONES, 48, (16384, 3001), 6, 0.05),
])
def test_noalloc_matches_work_buffer_entry_point(kind, dim, sizes, iters, step0):
    nx, _ = P.dims(kind, dim)
    _, _, lb, ub = O.problem_info(kind, dim)
    for call, n in enumerate(sizes):
        assert P.local_descent_work_len(n, nx) <= P.LOCAL_DESCENT_NOALLOC_MAX_WORK
        if kind == P.LENNARD_JONES:
            x0 = O.batch_uniform(np.full(nx, -3.0), np.full(nx, 3.0), n, 100 + call)
            lo, hi = None, None
        else:
            x0 = O.batch_uniform(lb, ub, n, 100 + call)
            lo, hi = lb, ub
        (xa, fa), (xb, fb) = run_both(kind, dim, x0, iters, step0, lo, hi)
        assert np.array_equal(xa, xb, equal_nan=True), f"call {call}: positions differ"
        assert np.array_equal(fa, fb, equal_nan=True), f"call {call}: fitness differs"
        assert not np.array_equal(xa, x0)  # the step moved something
        if kind == P.ROSENBROCK:
            xo, fo = O.local_descent(kind, dim, x0, iters, step0, lo, hi)
            assert np.array_equal(xb, xo) and np.array_equal(fb, fo)


def test_noalloc_limit_and_arguments():
    nx = 138
    most = P.LOCAL_DESCENT_NOALLOC_MAX_WORK // (2 * (nx + 1))
    u = P.make_udp(P.LENNARD_JONES, 48)
    x = torch.zeros((1, nx), dtype=torch.float64, device=DEV)
    f = torch.zeros(1, dtype=torch.float64, device=DEV)
    # one individual too many: refused before anything is launched
    rc = P.lib().pgh_local_descent_noalloc(C.byref(u), x.data_ptr(), most + 1, nx, f.data_ptr(), None, None, 1, 0.1,
                                           None)
    assert rc == P.INVALID_ARGUMENT
    assert "work buffer beyond" in P.lib().pgh_last_error().decode()
    # problems without a device gradient and bad steps as pgh_local_descent
    ua = P.make_udp(P.ACKLEY, 3)
    rc = P.lib().pgh_local_descent_noalloc(C.byref(ua), x.data_ptr(), 1, 3, f.data_ptr(), None, None, 1, 0.1, None)
    assert rc == P.UNSUPPORTED
    rc = P.lib().pgh_local_descent_noalloc(C.byref(u), x.data_ptr(), 1, nx, f.data_ptr(), None, None, 1, -1.0, None)
    assert rc == P.INVALID_ARGUMENT
    # zero individuals: nothing to do
    assert P.lib().pgh_local_descent_noalloc(C.byref(u), None, 0, nx, None, None, None, 1, 0.1, None) == P.OK
