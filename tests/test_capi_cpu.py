"""C-ABI host logic, no GPU: the library loads, exports every declared symbol,
and validates its inputs with the reference's messages before touching a
device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pagmo2_amd as P
from pagmo2_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exports_match_header():
    with open(os.path.join(ROOT, "include", "pagmo_hip", "pgh.h")) as f:
        hdr = f.read()
    declared = set(re.findall(r"\b(pgh_[A-Za-z0-9_]+)\s*\(", hdr))
    assert declared == set(_capi.EXPORTS)
    assert P.missing_exports() == []


def test_abi_version():
    assert P.lib().pgh_abi_version() == 3
    assert C.sizeof(P.pgh_udp) == 48


def test_problem_dims():
    assert P.dims(P.ROSENBROCK, 10) == (10, 1)
    assert P.dims(P.LENNARD_JONES, 48) == (138, 1)
    assert P.dims(P.RASTRIGIN, 1000) == (1000, 1)
    assert P.dims(P.MO_SPHERE, 30, nobj=2) == (30, 2)


@pytest.mark.parametrize("kind,dim,msg", [
    (1, 1, "Rosenbrock Function must have minimum 2 dimensions, 1 requested"),        # rosenbrock.cpp:43-49
    (2, 0, "Rastrigin Function must have minimum 1 dimension, 0 requested"),         # rastrigin.cpp:46-51
    (3, 0, "Ackley Function must have minimum 1 dimension, 0 requested"),
    (4, 0, "Schwefel Function must have minimum 1 dimension, 0 requested"),
    (5, 0, "Griewank Function must have minimum 1 dimension, 0 requested"),
    (6, 2, "The number of atoms in a Lennard Jones Clusters problem must be positive and greater than 2, "
           "while a number of 2 was detected."),                                        # lennard_jones.cpp:51-57
])
def test_problem_ctor_errors(kind, dim, msg):
    with pytest.raises(P.PghError) as e:
        P.dims(kind, dim)
    assert e.value.status == P.INVALID_ARGUMENT
    assert str(e.value) == msg


def test_unknown_kind_is_unsupported():
    u = P.make_udp(42, 3)
    assert P.lib().pgh_supported(C.byref(u)) == 0
    assert P.lib().pgh_supported(C.byref(P.make_udp(P.ROSENBROCK, 3))) == 1


def test_bfe_host_input_length_message(kats):
    # src/detail/bfe_impl.cpp:63-88, tested verbatim by tests/bfe.cpp:366-402
    u = P.make_udp(P.ROSENBROCK, 3)
    dvs = np.zeros(7)
    out = np.zeros(2)
    rc = P.lib().pgh_bfe_eval_host(C.byref(u), dvs.ctypes.data, 7, out.ctypes.data, 2)
    assert rc == P.INVALID_ARGUMENT
    assert P.lib().pgh_last_error().decode() == kats["bfe_messages"]["input_not_multiple"].format(n=7, nx=3)


def test_bfe_device_shape_mismatch():
    u = P.make_udp(P.RASTRIGIN, 5)
    rc = P.lib().pgh_bfe_eval(C.byref(u), None, 0, 4, None, 1, None)
    assert rc == P.INVALID_ARGUMENT
    assert "the number of dimensions of the problem is 5" in P.lib().pgh_last_error().decode()


def test_gradient_unsupported_kind():
    # ackley has no gradient in the reference (problem::gradient throws
    # not_implemented_error, problem.hpp:633-645)
    u = P.make_udp(P.ACKLEY, 4)
    rc = P.lib().pgh_batch_gradient(C.byref(u), None, 0, 4, None, None, None)
    assert rc == P.UNSUPPORTED


def test_empty_batch_is_noop():
    # zero decision vectors: nothing to launch, no device needed
    u = P.make_udp(P.ROSENBROCK, 3)
    dvs = np.zeros(0)
    out = np.zeros(0)
    assert P.lib().pgh_bfe_eval_host(C.byref(u), dvs.ctypes.data, 0, out.ctypes.data, 0) == P.OK


def test_no_device_reported():
    if P.lib().pgh_device_available():
        pytest.skip("a GPU is present")
    u = P.make_udp(P.ROSENBROCK, 3)
    dvs = np.zeros(6)
    out = np.zeros(2)
    rc = P.lib().pgh_bfe_eval_host(C.byref(u), dvs.ctypes.data, 6, out.ctypes.data, 2)
    assert rc == P.NO_DEVICE


def test_crowding_and_selection_host_checks():
    # validations and corner cases answered before any device work
    # (multi_objective.cpp:283-298, 346-360, 427-434)
    with pytest.raises(P.PghError, match="A non dominated front must contain at least two points: 1 detected."):
        P.crowding_distance_host(np.zeros((1, 2)))
    with pytest.raises(P.PghError, match="Points in the non dominated front must contain at least two objectives: "
                                         "1 detected."):
        P.crowding_distance_host(np.zeros((3, 1)))
    assert list(P.select_best_N_mo_host(np.zeros((0, 2)), 10)) == []
    assert list(P.select_best_N_mo_host(np.zeros((5, 2)), 0)) == []
    assert list(P.select_best_N_mo_host(np.zeros((1, 2)), 10)) == [0]
    assert list(P.select_best_N_mo_host(np.random.default_rng(0).random((4, 2)), 4)) == [0, 1, 2, 3]
    assert list(P.sort_population_mo_host(np.zeros((0, 2)))) == []
    assert list(P.sort_population_mo_host(np.zeros((1, 4)))) == [0]


def test_local_descent_host_checks():
    # argument checks answered before any device work
    x = np.zeros((2, 3))
    with pytest.raises(P.PghError, match="The gradient has not been implemented"):
        P.local_descent_host(P.ACKLEY, 3, x, 1, 0.1)
    for bad in (0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(P.PghError, match="initial step of a local descent"):
            P.local_descent_host(P.ROSENBROCK, 3, x, 1, bad)
    u = P.make_udp(P.ROSENBROCK, 3)
    lb = np.zeros(3)
    f = np.zeros(2)
    rc = P.lib().pgh_local_descent_host(C.byref(u), x.ctypes.data, 6, f.ctypes.data, 2, lb.ctypes.data, None, 1, 0.1)
    assert rc == P.INVALID_ARGUMENT and "both given or both absent" in P.lib().pgh_last_error().decode()
    with pytest.raises(P.PghError, match="not an exact multiple"):
        P.local_descent_host(P.ROSENBROCK, 3, np.zeros((1, 4)).reshape(2, 2), 1, 0.1)
    assert P.lib().pgh_local_descent_work_len(1000, 30) == 2 * 1000 * 30 + 2 * 1000
    x0, f0 = P.local_descent_host(P.ROSENBROCK, 3, np.zeros((0, 3)), 5, 0.1)
    assert x0.shape == (0, 3) and f0.shape == (0,)


def test_local_descent_noalloc_limit():
    # a batch whose work exceeds the library's device array is refused before
    # any HIP call; the largest batch that fits passes that check
    nx = 138
    most = P.LOCAL_DESCENT_NOALLOC_MAX_WORK // (2 * (nx + 1))
    assert P.local_descent_work_len(most, nx) <= P.LOCAL_DESCENT_NOALLOC_MAX_WORK
    assert P.local_descent_work_len(most + 1, nx) > P.LOCAL_DESCENT_NOALLOC_MAX_WORK
    u = P.make_udp(P.LENNARD_JONES, 48)
    rc = P.lib().pgh_local_descent_noalloc(C.byref(u), 8, most + 1, nx, 8, None, None, 1, 0.1, None)
    assert rc == P.INVALID_ARGUMENT and "work buffer beyond" in P.lib().pgh_last_error().decode()
    # the same argument checks as pgh_local_descent
    ua = P.make_udp(P.ACKLEY, 3)
    assert P.lib().pgh_local_descent_noalloc(C.byref(ua), 8, 1, 3, 8, None, None, 1, 0.1, None) == P.UNSUPPORTED
    assert P.lib().pgh_local_descent_noalloc(C.byref(u), 8, 1, 137, 8, None, None, 1, 0.1, None) == P.INVALID_ARGUMENT
    assert P.lib().pgh_local_descent_noalloc(C.byref(u), None, 0, nx, None, None, None, 1, 0.1, None) == P.OK



def test_zdt_dtlz_descriptions(kats):
    # shapes (zdt.cpp:108-157, dtlz.cpp:116-119) and the constructors' messages
    assert P.dims(P.ZDT, 30, prob_id=1) == (30, 2)
    assert P.dims(P.ZDT, 30, prob_id=5) == (175, 2)
    assert P.dims(P.ZDT, 11, prob_id=5) == ((11 - 1) * 5 + 30, 2)
    assert P.dims(P.DTLZ, 7, nobj=3, prob_id=5, alpha=100) == (7, 3)
    for c in kats["zdt_throws"]:
        with pytest.raises(P.PghError) as e:
            P.dims(P.ZDT, c["param"], prob_id=c["prob_id"])
        assert e.value.status == P.INVALID_ARGUMENT and str(e.value) == c["message"]
    for c in kats["dtlz_throws"]:
        with pytest.raises(P.PghError) as e:
            P.dims(P.DTLZ, c["dim"], nobj=c["fdim"], prob_id=c["prob_id"], alpha=100)
        assert e.value.status == P.INVALID_ARGUMENT and str(e.value) == c["message"]
    # no device seeding for these two families
    u = P.make_udp(P.ZDT, 30, prob_id=1)
    assert P.lib().pgh_batch_random_dv(C.byref(u), 4, 1, None, None) == P.UNSUPPORTED



def test_translation_only_on_host_buffer_paths():
    # the device entry points may not modify their input: a translated problem
    # is refused there, before any device work
    t = np.full(3, 0.5)
    u = P.make_udp(P.ROSENBROCK, 3, shift=t)
    out = np.zeros(2)
    assert P.lib().pgh_bfe_eval(C.byref(u), 1, 2, 3, out.ctypes.data, 1, None) == P.UNSUPPORTED
    assert "does not take a translated problem" in P.lib().pgh_last_error().decode()
    assert P.lib().pgh_batch_gradient(C.byref(u), 1, 2, 3, 1, None, None) == P.UNSUPPORTED
    assert P.lib().pgh_batch_random_dv(C.byref(u), 2, 1, None, None) == P.UNSUPPORTED
    with pytest.raises(P.PghError, match="does not take a translated problem"):
        P.check(P.lib().pgh_local_descent_host(C.byref(u), out.ctypes.data, 0, out.ctypes.data, 0, None, None, 1,
                                               0.1))
