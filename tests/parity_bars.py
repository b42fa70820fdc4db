"""The parity bars of the GPU tests, in one place.

* rosenbrock (and every fitness made of + - * in the reference's order):
  bit-exact;
* rastrigin, ackley, schwefel, griewank: within 1e-12 relative of the oracle;
* lennard_jones: within 1e-12 of the sum of the absolute pair terms of the
  individual, 4 sum(|r^-12| + |r^-6|), since the energy is a cancelling sum
  (lennard_jones.cpp:72-91 accumulates pow(sixth, 2) - sixth);
* gradients: 1e-12 norm-wise against the oracle's analytic gradient.
"""
import numpy as np

ROSENBROCK, RASTRIGIN, ACKLEY, SCHWEFEL, GRIEWANK, LENNARD_JONES = range(1, 7)


def lj_positions(x):
    """Atom positions of a decision vector (lennard_jones.cpp:132-151)."""
    atoms = (x.size + 6) // 3
    p = np.zeros((atoms, 3))
    p[1, 2] = x[0]
    p[2, 1:] = x[1:3]
    p[3:] = x[3:].reshape(-1, 3)
    return p


def lj_abs_scale(x):
    """4 sum over pairs of |r^-12| + |r^-6| for one decision vector."""
    p = lj_positions(np.asarray(x, np.float64))
    atoms = p.shape[0]
    d = p[:, None, :] - p[None, :, :]
    r2 = np.sum(d * d, axis=-1)[np.triu_indices(atoms, 1)]
    with np.errstate(divide="ignore", over="ignore"):
        s = r2 ** -3
        return 4 * np.sum(np.abs(s * s) + np.abs(s))


def lj_abs_scales(X):
    """lj_abs_scale of every row of X, vectorised over rows."""
    X = np.asarray(X, np.float64)
    n, nx = X.shape
    atoms = (nx + 6) // 3
    p = np.zeros((n, atoms, 3))
    p[:, 1, 2] = X[:, 0]
    p[:, 2, 1:] = X[:, 1:3]
    p[:, 3:] = X[:, 3:].reshape(n, -1, 3)
    iu, ju = np.triu_indices(atoms, 1)
    out = np.empty(n)
    for a in range(0, n, 1024):
        q = p[a:a + 1024]
        d = q[:, iu, :] - q[:, ju, :]
        r2 = np.sum(d * d, axis=-1)
        with np.errstate(divide="ignore", over="ignore"):
            s = r2 ** -3
            out[a:a + 1024] = 4 * np.sum(np.abs(s * s) + np.abs(s), axis=1)
    return out


def assert_fitness(kind, X, got, ref, what=""):
    """got (device) against ref (oracle) for the single-objective kinds, at
    the bar of the kind; X holds the decision vectors the values belong to."""
    got = np.asarray(got, np.float64).reshape(-1)
    ref = np.asarray(ref, np.float64).reshape(-1)
    assert got.shape == ref.shape, what
    if kind == ROSENBROCK:
        assert np.array_equal(got, ref), what
    elif kind == LENNARD_JONES:
        scale = lj_abs_scales(X)
        same = got == ref  # covers +inf
        bad = ~same & ~(np.abs(got - ref) <= 1e-12 * scale)
        assert not np.any(bad), (what, np.flatnonzero(bad)[:5], got[bad][:5], ref[bad][:5])
    else:
        same = got == ref
        bad = ~same & ~(np.abs(got - ref) <= 1e-12 * np.abs(ref))
        assert not np.any(bad), (what, np.flatnonzero(bad)[:5], got[bad][:5], ref[bad][:5])
