"""Pins of the oracle's moving-body and ellipse routines (CPU).

The reference has no immersed boundary (iBlankComputeKernel is a stub,
src/include/preSim.cu:110-136; its circle test is commented out at :122-130)
and no moving bodies; the predecessor code only names the face blanking and
the body boundary condition (test/UTIL_PRE_SIM.f90:152-164).  The oracle's
routines for them are therefore held here by definitions restated
independently of the oracle's own code:

* of_phys_move_bodies: the fresh cells are the set {solid before, fluid
  after} of a brute-force classification at the two positions; each takes
  the velocity of the body that left it; nothing else changes; the new
  flags are those of a brute-force classification at the new position.
* the moving-body Neumann term |IP - GC| a_B.n: on a linear pressure
  p = -a_B.x (the field whose normal derivative is -a_B.n everywhere) the
  ghost value must come back as -a_B.x_GC exactly, for every ghost whose
  bilinear image-point stencil is four fluid cells.  A sign error, a
  normal pointing the wrong way or a length of |BI - GC| instead of
  |IP - GC| each miss by O(h |a_B|).
* the ellipse inside test: against the focal-distance definition
  (|x - F1| + |x - F2| <= 2a), a formula with no rotation in it.
* the heave kinematics the tests and the engine use: velocity and
  acceleration are the time derivatives of the position.
"""
import math

import numpy as np

import oracle_lib as O


def uniform_faces(n, L):
    return np.linspace(0.0, L, n + 1)


def _circle_solid(xc, yc, cx, cy, r):
    """Brute-force circle classification on the cell centres (FP64, the
    squared-distance form, no square root)."""
    X, Y = np.meshgrid(xc, yc)
    ex, ey = X - cx, Y - cy
    return (ex * ex + ey * ey) <= r * r


def _heave_state(y0, A, f, t):
    w = 2.0 * math.pi * f
    return y0 + A * math.sin(w * t), A * w * math.cos(w * t), -A * w * w * math.sin(w * t)


def test_move_bodies_fresh_cells_brute_force():
    """Two heaving circles moving in opposite directions at different
    speeds, plus one static circle: the fresh cells and their velocities
    against a brute-force classification at both positions."""
    xf = uniform_faces(96, 8.0)
    xc, dx = O.metric(xf)
    nx = ny = len(xc)
    specs = {0: (3.0, 0.5, 0.2), 1: (5.0, -0.4, 0.35)}  # body: (y0, A, f)
    circ = [dict(kind=O.CIRCLE, xc=2.0, yc=3.0, r=0.7), dict(kind=O.CIRCLE, xc=5.5, yc=5.0, r=0.9),
            dict(kind=O.CIRCLE, xc=6.5, yc=1.5, r=0.6)]
    c = O.PhysCase(xf, xf, circ, dict(dt=0.01, re=100.0, conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=25,
                                      ad_tol=1e-6, ppe_solver=1, ppe_omega=1.8, ppe_itermax=1000,
                                      ppe_tol=1e-8, check_every=5))
    rng = np.random.default_rng(17)
    checked = 0
    for t in (0.0, 0.4, 1.1, 2.3):
        # state before: bodies where the last call left them, random velocities
        before = [(b.xc, b.yc) for b in c.bodies.arr[:3]]
        solid_before = [_circle_solid(xc, xc, bx, by, circ[k]["r"]) for k, (bx, by) in enumerate(before)]
        c.u[:] = rng.standard_normal(nx * ny)
        c.v[:] = rng.standard_normal(nx * ny)
        u0, v0, f0 = c.u.copy(), c.v.copy(), c.flags.copy()
        fresh_n = c.heave(specs, t)
        # brute force at the new position
        vel = {}
        after = []
        for k in range(3):
            if k in specs:
                y, vb, _ = _heave_state(*specs[k], t)
                after.append((circ[k]["xc"], y))
                vel[k] = (0.0, vb)
            else:
                after.append((circ[k]["xc"], circ[k]["yc"]))
                vel[k] = (0.0, 0.0)
        solid_after = [_circle_solid(xc, xc, bx, by, circ[k]["r"]) for k, (bx, by) in enumerate(after)]
        any_before = np.logical_or.reduce(solid_before)
        any_after = np.logical_or.reduce(solid_after)
        F = c.flags.reshape(ny, nx)
        assert np.array_equal((F & O.OF_FLUID) == 0, any_after)
        assert np.array_equal((f0.reshape(ny, nx) & O.OF_FLUID) == 0, any_before)
        fresh = any_before & ~any_after
        assert fresh_n == int(fresh.sum())
        U, V, U0, V0 = (a.reshape(ny, nx) for a in (c.u, c.v, u0, v0))
        # every fresh cell carries the velocity of the body that left it
        # (bodies do not overlap, so exactly one was there)
        for k in range(3):
            left = fresh & solid_before[k]
            assert np.all(U[left] == vel[k][0]) and np.all(V[left] == vel[k][1])
            checked += int(left.sum())
        # and nothing else changed
        assert np.array_equal(U[~fresh], U0[~fresh]) and np.array_equal(V[~fresh], V0[~fresh])
        # the body ids follow the new classification
        B = c.bid.reshape(ny, nx)
        for k in range(3):
            assert np.array_equal(B == k, solid_after[k])
        # the rebuilt ghost list is the brute-force one at the new position
        fl = ~any_after
        ghost = np.zeros_like(any_after)
        ghost[1:-1, 1:-1] = any_after[1:-1, 1:-1] & (fl[1:-1, 2:] | fl[1:-1, :-2] | fl[2:, 1:-1] | fl[:-2, 1:-1])
        assert np.array_equal(np.sort(np.asarray(c.glist)), np.flatnonzero(ghost.ravel()))
    assert checked > 20  # the bodies did leave cells behind


def test_moving_body_neumann_term_reproduces_linear_pressure():
    """p = -a.x has dp/dn = -a.n; the ghost value from the image point plus
    |IP - GC| a.n must be -a.x_GC wherever the bilinear stencil at the
    image point is four fluid cells (bilinear weights reproduce a linear
    field exactly)."""
    xf = uniform_faces(80, 4.0)
    xc, dx = O.metric(xf)
    nx = ny = len(xc)
    for ax, ay in ((0.0, -3.7), (2.5, 1.25), (-1.0, 0.0)):
        bs = O.BodySet([dict(kind=O.CIRCLE, xc=2.03, yc=1.96, r=0.57, axb=ax, ayb=ay)])
        flags, bid = O.classify(xc, xc, bs)
        lst, gh, ng = O.ghost_records(xc, xc, flags, bid, bs)
        X, Y = np.meshgrid(xc, xc)
        p = (-(ax * X + ay * Y)).ravel().copy()
        q = p.copy()
        O.lib().of_phys_ghost_pressure(nx, ny, ng, gh, 1, bs.arr, O.dptr(q))
        n4 = 0
        scale = math.hypot(ax, ay) * 4.0
        for g in range(ng):
            o = gh[g]
            if all(o.node[k] != o.gidx for k in range(4)) and all(o.wn[k] != 0.0 for k in range(4)):
                n4 += 1
                assert abs(q[o.gidx] - p[o.gidx]) <= 1e-13 * max(1.0, scale)
                # the lever arm is the full IP - GC vector, along the outward normal
                gx, gy = xc[o.gidx % nx], xc[o.gidx // nx]
                nxv, nyv = gx - 2.03, gy - 1.96
                L = math.hypot(o.lnx, o.lny)
                assert abs(L - 2.0 * (0.57 - math.hypot(nxv, nyv))) < 1e-12
                assert o.lnx * nxv + o.lny * nyv > 0.0
        assert n4 >= 12
        # the same ghosts with the sign of the term flipped miss by O(h a)
        bs_neg = O.BodySet([dict(kind=O.CIRCLE, xc=2.03, yc=1.96, r=0.57, axb=-ax, ayb=-ay)])
        r = p.copy()
        O.lib().of_phys_ghost_pressure(nx, ny, ng, gh, 1, bs_neg.arr, O.dptr(r))
        assert np.max(np.abs(r[lst] - p[lst])) > 1e-3 * math.hypot(ax, ay)


def test_ellipse_inside_test_against_focal_definition():
    """Ellipse classification (rotated, both orientations of the major
    axis) against |x - F1| + |x - F2| <= 2a on a fine grid of centres; the
    points within 1e-12 of the boundary are left out (rounding)."""
    xs = np.linspace(-1.3, 1.3, 521)
    for (a, b, deg) in ((0.5, 0.05, 10.0), (0.3, 0.8, -35.0), (0.6, 0.6 / 4.2, 123.0)):
        th = math.radians(deg)
        body = dict(kind=O.ELLIPSE, xc=0.11, yc=-0.07, a=a, b=b, c=math.cos(th), s=math.sin(th))
        flags, _ = O.classify(xs, xs, O.BodySet([body]))
        # focal definition in the world frame: the major axis lies along
        # the rotated x axis when a >= b, along the rotated y axis otherwise
        big, small = max(a, b), min(a, b)
        fc = math.sqrt(big * big - small * small)
        dirx, diry = (math.cos(th), math.sin(th)) if a >= b else (-math.sin(th), math.cos(th))
        X, Y = np.meshgrid(xs, xs)
        d1 = np.hypot(X - (0.11 + fc * dirx), Y - (-0.07 + fc * diry))
        d2 = np.hypot(X - (0.11 - fc * dirx), Y - (-0.07 - fc * diry))
        s = d1 + d2 - 2.0 * big
        sure = np.abs(s) > 1e-12
        solid = (flags.reshape(len(xs), len(xs)) & O.OF_FLUID) == 0
        assert sure.sum() > 0.999 * s.size
        assert np.array_equal(solid[sure], (s <= 0)[sure])
        assert solid.sum() > 100


def test_heave_kinematics_are_derivatives():
    """The heave state used by the oracle cases and the engine: velocity and
    acceleration are the first and second time derivatives of
    y = y0 + A sin(2 pi f t) (central differences, O(h^2))."""
    y0, A, f = 10.0, 0.5, 0.2
    h = 1e-4
    for t in (0.0, 0.7, 1.9, 3.3):
        ym, _, _ = _heave_state(y0, A, f, t - h)
        yc, vc, ac = _heave_state(y0, A, f, t)
        yp, _, _ = _heave_state(y0, A, f, t + h)
        assert abs((yp - ym) / (2 * h) - vc) < 1e-6
        assert abs((yp - 2 * yc + ym) / (h * h) - ac) < 1e-3
    # the oracle case places the body with exactly this state
    xf = uniform_faces(32, 20.0)
    c = O.PhysCase(xf, xf, [dict(kind=O.CIRCLE, xc=6.0, yc=10.0, r=0.5)],
                   dict(dt=0.3, re=100.0, conv_factor=0.5, ubc=1.0, vbc=0.0, ad_itermax=5, ad_tol=1e-6,
                        ppe_solver=1, ppe_omega=1.5, ppe_itermax=10, ppe_tol=1e-6, check_every=5))
    c.heave({0: (y0, A, f)}, 0.9)
    y, v, a = _heave_state(y0, A, f, 0.9)
    b = c.bodies.arr[0]
    assert (b.yc, b.vb, b.ayb) == (y, v, a)
