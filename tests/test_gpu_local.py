"""The batched gradient in use: the memetic local step (pgh_local_descent)
against the oracle's statement of its rule, and a torch.autograd.Function
around pgh_batch_gradient (check harness only, never in the library).

Bars: rosenbrock bit-exact (its fitness and gradient kernels are, and the
step's max-norm is exact in any order); rastrigin and Lennard-Jones by the
rule's invariants at full size (fitness never increases, the returned fitness
is the bfe fitness of the returned point bit for bit, the bounds hold) and one
round against the oracle within 1e-12.
"""
import numpy as np
import pytest

import oracle as O
import pagmo2_amd as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda:0"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def bounds(kind, dim):
    _, _, lb, ub = O.problem_info(kind, dim)
    return lb, ub


def device_descent(kind, dim, x, iters, step0, lb=None, ub=None):
    xd = dev(x)
    f = torch.empty(x.shape[0], dtype=torch.float64, device=DEV)
    P.local_descent(kind, dim, xd, f, iters, step0, None if lb is None else dev(lb), None if ub is None else dev(ub))
    torch.cuda.synchronize()
    return xd.cpu().numpy(), f.cpu().numpy()


def gpu_fitness(kind, dim, x):
    out = torch.empty((x.shape[0], 1), dtype=torch.float64, device=DEV)
    P.bfe_eval(kind, dim, dev(x), out)
    torch.cuda.synchronize()
    return out.cpu().numpy()[:, 0]


def test_hand_derived_traces():
    # rosenbrock at the origin: g = (-2, 0); alpha 0.5 -> (0.5, 0), f 6.5,
    # rejected; alpha 0.25 -> (0.25, 0), f 0.953125, accepted
    x, f = device_descent(P.ROSENBROCK, 2, np.zeros((1, 2)), 2, 0.5)
    assert x.tolist() == [[0.25, 0.0]] and f.tolist() == [0.953125]
    # an upper bound of 0.125 clamps the first trial: (0.125, 0), f 0.7900390625
    x, f = device_descent(P.ROSENBROCK, 2, np.zeros((1, 2)), 1, 0.5, np.array([-5.0, -5.0]), np.array([0.125, 10.0]))
    assert x.tolist() == [[0.125, 0.0]] and f.tolist() == [0.7900390625]
    # zero rounds: only the fitness
    x, f = device_descent(P.ROSENBROCK, 2, np.array([[1.0, 1.0]]), 0, 0.5)
    assert x.tolist() == [[1.0, 1.0]] and f.tolist() == [0.0]


@pytest.mark.parametrize("n,dim,iters", [(1, 2, 5), (64, 10, 30), (257, 37, 12), (1000, 3, 40)])
def test_rosenbrock_bit_exact_with_oracle(n, dim, iters):
    lb, ub = bounds(P.ROSENBROCK, dim)
    x0 = O.batch_uniform(lb, ub, n, 11 + dim)
    xo, fo = O.local_descent(P.ROSENBROCK, dim, x0, iters, 0.5, lb, ub)
    xd, fd = device_descent(P.ROSENBROCK, dim, x0, iters, 0.5, lb, ub)
    assert np.array_equal(xd, xo) and np.array_equal(fd, fo)
    xh, fh = P.local_descent_host(P.ROSENBROCK, dim, x0, iters, 0.5, lb, ub)
    assert np.array_equal(xh, xo) and np.array_equal(fh, fo)
    # unbounded form
    xo, fo = O.local_descent(P.ROSENBROCK, dim, x0, iters, 0.5)
    xd, fd = device_descent(P.ROSENBROCK, dim, x0, iters, 0.5)
    assert np.array_equal(xd, xo) and np.array_equal(fd, fo)


@pytest.mark.parametrize("kind,dim,n,iters,step0", [
    (P.RASTRIGIN, 1000, 4096, 8, 0.05),
    (P.LENNARD_JONES, 48, 16384, 8, 0.05),
    (P.LENNARD_JONES, 13, 3000, 25, 0.1),
])
def test_invariants_at_size(kind, dim, n, iters, step0):
    lb, ub = bounds(kind, dim)
    nx = lb.size
    lo, hi = (lb, ub) if kind == P.RASTRIGIN else (np.full(nx, -3.0), np.full(nx, 3.0))
    x0 = O.batch_uniform(lo, hi, n, 5)
    x0 = np.clip(x0, lb, ub)
    f0 = gpu_fitness(kind, dim, x0)
    x, f = device_descent(kind, dim, x0, iters, step0, lb, ub)
    assert np.all(f <= f0)
    assert np.mean(f < f0) > 0.5  # most individuals improve
    assert np.array_equal(f, gpu_fitness(kind, dim, x))
    assert np.all((x >= lb) & (x <= ub))
    moved = np.any(x != x0, axis=1)
    assert np.array_equal(moved, f < f0)


@pytest.mark.parametrize("kind,dim", [(P.RASTRIGIN, 50), (P.LENNARD_JONES, 10)])
def test_one_round_against_oracle(kind, dim):
    lb, ub = bounds(kind, dim)
    nx = lb.size
    lo, hi = (lb, ub) if kind == P.RASTRIGIN else (np.full(nx, 0.5), np.full(nx, 2.5))
    x0 = np.clip(O.batch_uniform(lo, hi, 200, 9), lb, ub)
    xo, fo = O.local_descent(kind, dim, x0, 1, 0.05, lb, ub)
    xd, fd = device_descent(kind, dim, x0, 1, 0.05, lb, ub)
    same_decision = np.any(xd != x0, axis=1) == np.any(xo != x0, axis=1)
    assert np.all(same_decision)
    assert np.allclose(xd, xo, rtol=1e-12, atol=1e-12)
    assert np.allclose(fd, fo, rtol=1e-12, atol=1e-12 * np.abs(fo).max())


def test_workspace_checks():
    x = dev(np.zeros((4, 2)))
    f = torch.empty(4, dtype=torch.float64, device=DEV)
    need = P.local_descent_work_len(4, 2)
    assert need == 2 * 4 * 2 + 2 * 4
    small = torch.empty(need - 1, dtype=torch.float64, device=DEV)
    with pytest.raises(P.PghError, match="work buffer"):
        P.local_descent(P.ROSENBROCK, 2, x, f, 1, 0.5, work=small)
    with pytest.raises(P.PghError, match="gradient has not been implemented"):
        P.local_descent(P.ACKLEY, 2, x, f, 1, 0.5)


# ---- torch.autograd.Function around the batched gradient (harness only) ----

class BatchFitness(torch.autograd.Function):
    """f(x) per row with d f / d x from pgh_batch_gradient."""

    @staticmethod
    def forward(ctx, x, kind, dim):
        x = x.contiguous()
        g = torch.empty_like(x)
        f = torch.empty((x.shape[0], 1), dtype=torch.float64, device=x.device)
        P.batch_gradient(kind, dim, x, g, f)
        ctx.save_for_backward(g)
        return f[:, 0]

    @staticmethod
    def backward(ctx, go):
        (g,) = ctx.saved_tensors
        return go[:, None] * g, None, None


def lj_spread(atoms, n, seed):
    """Clusters whose atoms sit near the points of a grid of spacing 1.1 (no
    pair closer than about 0.9), so that finite differences are accurate."""
    rng = np.random.default_rng(seed)
    grid = np.array([(i, j, k) for i in range(3) for j in range(3) for k in range(3)], np.float64) * 1.1
    out = []
    for _ in range(n):
        p = grid[rng.permutation(len(grid))[:atoms]] + rng.uniform(-0.05, 0.05, (atoms, 3))
        p -= p[0]
        # pagmo's encoding fixes atom 0 at the origin, atom 1 on the z axis and
        # atom 2 in the y-z plane: rotate the cluster accordingly
        e3 = p[1] / np.linalg.norm(p[1])
        v = p[2] - np.dot(p[2], e3) * e3
        e2 = v / np.linalg.norm(v)
        e1 = np.cross(e2, e3)
        q = p @ np.stack([e1, e2, e3], axis=1)
        out.append(np.concatenate([[q[1, 2]], q[2, 1:], q[3:].reshape(-1)]))
    return np.array(out)


@pytest.mark.parametrize("kind,dim", [(P.LENNARD_JONES, 5), (P.ROSENBROCK, 6), (P.RASTRIGIN, 7)])
def test_autograd_gradcheck(kind, dim):
    lb, ub = bounds(kind, dim)
    if kind == P.LENNARD_JONES:
        x0 = lj_spread(dim, 3, 21)
    else:
        x0 = O.batch_uniform(lb / 4, ub / 4, 3, 21)
    x = dev(x0).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: BatchFitness.apply(t, kind, dim), (x,), eps=1e-6, atol=1e-5,
                                    rtol=1e-6)


def test_autograd_backward_matches_batched_gradient_and_drives_a_local_step():
    atoms = 13
    x0 = lj_spread(atoms, 64, 3)
    x = dev(x0).requires_grad_(True)
    f = BatchFitness.apply(x, P.LENNARD_JONES, atoms)
    f.sum().backward()
    g = torch.empty_like(x)
    P.batch_gradient(P.LENNARD_JONES, atoms, x.detach(), g)
    assert torch.equal(x.grad, g)
    ref = O.gradient(P.LENNARD_JONES, atoms, x0)
    gd = x.grad.cpu().numpy()
    for i in range(64):
        assert np.linalg.norm(gd[i] - ref[i]) <= 1e-12 * np.linalg.norm(ref[i])
    # a few normalised-gradient steps with backtracking through autograd lower
    # every individual's energy, as the library's local step does
    e0 = f.detach().clone()
    xs = x.detach().clone()
    e = e0.clone()
    alpha = torch.full((64, 1), 0.01, dtype=torch.float64, device=DEV)
    for _ in range(10):
        xt = xs.clone().requires_grad_(True)
        ft = BatchFitness.apply(xt, P.LENNARD_JONES, atoms)
        ft.sum().backward()
        gt = xt.grad
        m = gt.abs().amax(dim=1, keepdim=True)
        y = xs - alpha / m * gt
        fy = BatchFitness.apply(y, P.LENNARD_JONES, atoms)
        ok = (fy < e)[:, None]
        xs = torch.where(ok, y, xs)
        e = torch.where(ok[:, 0], fy, e)
        alpha = torch.where(ok, 2 * alpha, 0.5 * alpha)
    assert torch.all(e <= e0) and torch.mean((e < e0).double()) > 0.9
