"""The five synthetic workloads, built as concrete grids, bodies and
parameters (grid files in the reference's format can be written with
`write_grid_file`).

C1  one circular cylinder (D=1) at (5, 8) in a 16x16 box, 256^2 uniform,
    Re=100, dt=0.002, FP64, Poisson stop 1e-8.
C2  ellipse (chord 1, thickness ratio 0.1) at 10 deg incidence, Re=300,
    2048x1024: uniform h=0.005 in a 4x2 box around the body, geometric
    stretching outward to a 30x20 domain (per-side ratio solved so the
    cells fill the domain; <= 1.02), dt = 0.5 h (CFL 0.5 at U=1).
C3  eight seeded star bodies in a 20x20 box, 4096^2 uniform, Re=1000.
C4  256 seeded cases, n log-uniform in [128, 1024], circle or ellipse,
    dt = min(0.5 h, 1/Re) (stable_dt).
C5  Poisson-arrival request trace: star cases (64^2..2048^2) plus 20 %
    heaving cylinders, 100..1000 steps, dt = min(0.5 h, 1/Re).
"""
import math

import numpy as np

from . import circle, ellipse, heaving_circle, star


def stable_dt(h, re, cfl=0.5):
    """Time step of the uniform-grid workloads: CFL 0.5 at U = 1, capped at
    1/Re.  The predictor advances the convective term explicitly (forward
    Euler, central differences, the reference's factor f = 0.5,
    ADSolver.cu:66-74) with the viscous term implicit; its von Neumann
    amplification factor (1 - i c sin t) / (1 + 4 d sin^2(t/2)), c = f|u|dt/h,
    d = dt/(Re h^2), stays within the unit circle only for
    dt <= 2 / (Re f^2 |u|^2).  The flow around the bodies reaches |u| ~ 2.4,
    so the limit is ~1.4/Re; coarse grids at high Re (CFL 0.5 gives dt up to
    0.16 on 64^2 in 20 x 20) diverge in a few hundred steps."""
    return min(cfl * h, 1.0 / re)


def write_grid_file(path, faces):
    """The reference's grid-file format: '<1-based id> <x>' per face
    (inputs/uniformGrid.py writes it with '{i+1:>10} {value:.7E}')."""
    with open(path, "w") as f:
        for i, v in enumerate(faces):
            f.write(f"{i + 1:>10} {v:.16E}\n")


def uniform_faces(n, L):
    return np.linspace(0.0, L, n + 1)


def _solve_ratio(h, n, length):
    """r with sum_{k=1..n} h r^k = length (bisection)."""
    lo, hi = 1.0 + 1e-12, 1.5
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        s = h * mid * (mid ** n - 1.0) / (mid - 1.0)
        if s > length:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def stretched_box_faces(n_total, h, box_lo, box_hi, dom_lo, dom_hi):
    """Uniform spacing h on [box_lo, box_hi], geometric stretching outward to
    the domain ends with one ratio per side solved to fill it."""
    nb = int(round((box_hi - box_lo) / h))
    side = (n_total - nb) // 2
    rl = _solve_ratio(h, side, box_lo - dom_lo)
    rr = _solve_ratio(h, n_total - nb - side, dom_hi - box_hi)
    left = [box_lo - h * rl * (rl ** k - 1) / (rl - 1) for k in range(side, 0, -1)]
    mid = list(box_lo + h * np.arange(nb + 1))
    nr = n_total - nb - side
    right = [box_hi + h * rr * (rr ** k - 1) / (rr - 1) for k in range(1, nr + 1)]
    f = np.array(left + mid + right)
    f[0], f[-1] = dom_lo, dom_hi
    f -= dom_lo  # faces start at 0 (the west ghost mirrors about x = 0)
    return f, rl, rr


def c1():
    xf = uniform_faces(256, 16.0)
    return dict(xf=xf, yf=xf.copy(), bodies=[circle(5.0, 8.0, 0.5)], dt=0.002, re=100.0,
                steps=500, ppe_tol=1e-8, name="C1 cylinder 256^2 FP64")


def c2(n=(2048, 1024)):
    h = 0.005
    xf, rx0, rx1 = stretched_box_faces(n[0], h, 13.0, 17.0, 0.0, 30.0)
    yf, ry0, ry1 = stretched_box_faces(n[1], h, 9.0, 11.0, 0.0, 20.0)
    body = ellipse(15.0, 10.0, 0.5, 0.05, math.radians(10.0))
    return dict(xf=xf, yf=yf, bodies=[body], dt=0.5 * h, re=300.0, steps=2000,
                ppe_tol=1e-6, ratios=(rx0, rx1, ry0, ry1),
                name="C2 ellipse 2048x1024 stretched")


class SplitMix64:
    """SplitMix64 PRNG; doubles = (x >> 11) * 2^-53."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) * 2.0 ** -53)


def star_set(rng, count, lo, hi, max_tries=10000):
    """Non-overlapping star bodies by rejection (bounding radius R(1+sum a))."""
    out, placed = [], []
    for _ in range(max_tries):
        if len(out) == count:
            break
        R = rng.uniform(0.3, 0.6)
        a = [rng.uniform(0.0, 0.08) for _ in range(5)]
        ph = [rng.uniform(0.0, 2 * math.pi) for _ in range(5)]
        rb = R * (1.0 + sum(a))
        x, y = rng.uniform(lo + rb, hi - rb), rng.uniform(lo + rb, hi - rb)
        if all(math.hypot(x - px, y - py) > rb + pr for px, py, pr in placed):
            placed.append((x, y, rb))
            out.append(star(x, y, R, a, ph, 400))
    return out


def c3(n=4096, seed=3):
    rng = SplitMix64(seed)
    xf = uniform_faces(n, 20.0)
    h = 20.0 / n
    return dict(xf=xf, yf=xf.copy(), bodies=star_set(rng, 8, 2.0, 18.0), dt=0.5 * h,
                re=1000.0, steps=1000, ppe_tol=1e-7, name="C3 8 stars 4096^2")


def c4_cases(count=256, seed=4):
    rng = SplitMix64(seed)
    cases = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(128), math.log(1024)))))
        xf = uniform_faces(n, 16.0)
        h = 16.0 / n
        if rng.uniform() < 0.5:
            body = circle(5.0, 8.0, 0.5)
        else:
            asp = rng.uniform(1.0, 5.0)
            ang = rng.uniform(0.0, math.pi)
            body = ellipse(5.0, 8.0, 0.5, 0.5 / asp, ang)
        re = rng.uniform(50.0, 500.0)
        cases.append(dict(xf=xf, yf=xf.copy(), bodies=[body], dt=stable_dt(h, re), re=re, steps=200,
                          ppe_tol=1e-7))
    return cases


def c5_trace(count=500, seed=5, rate=50.0):
    """(arrival time, case) pairs of the serving workload."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        n = int(round(math.exp(rng.uniform(math.log(64), math.log(2048)))))
        steps = int(100 + math.floor(rng.uniform() * 901))
        xf = uniform_faces(n, 20.0)
        h = 20.0 / n
        if rng.uniform() < 0.2:
            bodies = [heaving_circle(6.0, 10.0, 0.5, 0.5, 0.2)]
        else:
            bodies = star_set(rng, 1 + int(rng.uniform() * 3), 2.0, 18.0)
        re = rng.uniform(100.0, 1000.0)
        out.append((t, dict(xf=xf, yf=xf.copy(), bodies=bodies, dt=stable_dt(h, re),
                            re=re, steps=steps, ppe_tol=1e-7)))
    return out
