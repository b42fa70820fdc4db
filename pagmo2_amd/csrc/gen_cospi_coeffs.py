"""Coefficients of the cos(pi u) and sin(pi u) polynomials in trig_dev.hpp
(cos_2pi_small, sin_2pi_small).

cos(pi u) for |u| <= 1/2 is approximated by p(z), z = u^2 in [0, 1/4], a
degree-8 polynomial with the smallest maximum absolute error (Remez exchange,
50 significant digits).  The same polynomial times 10 is the rastrigin set.
sin(pi u) is u * q(z) with q the degree-8 minimax polynomial of
sin(pi sqrt(z)) / sqrt(z) on the same interval.  Prints the binary64
coefficients (hex) and the minimax errors.

    python pagmo2_amd/csrc/gen_cospi_coeffs.py
"""
import mpmath as mp

mp.mp.dps = 50
DEGREE = 8
A, B = mp.mpf(0), mp.mpf(1) / 4


def target_cos(z):
    return mp.cos(mp.pi * mp.sqrt(z))


def target_sin(z):
    return mp.pi if z == 0 else mp.sin(mp.pi * mp.sqrt(z)) / mp.sqrt(z)


def remez(n, target, iters=30, grid_points=4000):
    xs = [(A + B) / 2 - (B - A) / 2 * mp.cos(mp.pi * i / (n + 1)) for i in range(n + 2)]
    for _ in range(iters):
        m = mp.matrix(n + 2, n + 2)
        rhs = mp.matrix(n + 2, 1)
        for i, x in enumerate(xs):
            for j in range(n + 1):
                m[i, j] = x ** j
            m[i, n + 1] = (-1) ** i
            rhs[i] = target(x)
        sol = mp.lu_solve(m, rhs)
        coef = [sol[j] for j in range(n + 1)]
        level = sol[n + 1]

        def err(x):
            return mp.polyval(coef[::-1], x) - target(x)

        grid = [A + (B - A) * k / grid_points for k in range(grid_points + 1)]
        ev = [err(x) for x in grid]
        ext = [grid[0]] + [grid[k] for k in range(1, grid_points) if (ev[k] - ev[k - 1]) * (ev[k + 1] - ev[k]) <= 0]
        ext.append(grid[-1])
        if len(ext) != n + 2:
            ext = sorted(sorted(ext, key=lambda x: -abs(err(x)))[:n + 2])
        xs = ext
    return coef, level


if __name__ == "__main__":
    coef, level = remez(DEGREE, target_cos)
    print(f"degree {DEGREE} in u^2, minimax absolute error {mp.nstr(abs(level), 5)}")
    print("cos(pi u):", ", ".join(float(c).hex() for c in coef))
    print("10 cos(pi u):", ", ".join(float(10 * c).hex() for c in coef))
    coef, level = remez(DEGREE, target_sin)
    print(f"sin(pi u) / u: degree {DEGREE} in u^2, minimax absolute error {mp.nstr(abs(level), 5)}")
    print("sin(pi u) / u:", ", ".join(float(c).hex() for c in coef))
