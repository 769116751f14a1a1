"""Workload builders shared by the tests, the smoke check and the benchmark.

Each builder returns an ``immersedbodies.Problem``; ``oracle_for`` builds the FP64 CPU
oracle of the same problem so both run on identical inputs.
"""
from __future__ import annotations

import math

import numpy as np

import immersedbodies as ib
from immersedbodies import curves


def cylinder(nx=60, ny=60, nlevel=3, h=0.05, x0=(-1.0, -1.5), Re=100.0, dt=0.004, r=0.5,
             npts=None, center=(0.0, 0.0), U=(1.0, 0.0)):
    """Stationary circular cylinder in a uniform stream (the reference's solver tests
    and the 200^2 / 1024^2 cylinder workloads)."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: U)
    c = curves.Circle(r=r, center=center)
    pan = curves.partition(c, int(npts)) if npts else curves.panels(grid, c)
    bodies = ib.Bodies(cyl=ib.StaticBody(pan))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=dt))


def cylinder_200():
    """200 x 200 cells on [-1,3] x [-2,2], h = 0.02, 3 levels, Re 100, dt 0.004, 157 points."""
    return cylinder(nx=200, ny=200, nlevel=3, h=0.02, x0=(-1.0, -2.0), Re=100.0, dt=0.004,
                    npts=157)


def cylinder_1024(nlevel=5, dt=0.001):
    """1024 x 1024 cells on [-2,6] x [-4,4], h = 1/128, 5 levels, Re 200, 402 points."""
    return cylinder(nx=1024, ny=1024, nlevel=nlevel, h=1.0 / 128, x0=(-2.0, -4.0), Re=200.0,
                    dt=dt, npts=402)


def two_lines(nx=60, ny=60, nlevel=3, h=0.05, x0=(-1.0, -1.5), Re=100.0, dt=0.004):
    """Two static line segments (test/solver.jl:21-29)."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.0))
    l1 = curves.panels(grid, curves.LineSegment((-0.2, 0.4), (0.2, 0.1)))
    l2 = curves.panels(grid, curves.LineSegment((-0.2, -0.1), (0.2, -0.4)))
    bodies = ib.Bodies(line1=ib.StaticBody(l1), line2=ib.StaticBody(l2))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=dt))


def moving_cylinder(nx=60, ny=60, nlevel=3, h=0.05, x0=(-1.0, -1.5), Re=100.0, dt=0.004):
    """Moving cylinder of test/solver.jl:30-41."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.0))
    pan = curves.panels(grid, curves.Circle(r=0.5, center=(0.0, 0.0)))
    motion = lambda t: ib.RigidBodyTransform(pos=(0.3 * math.cos(t), 0.1 * math.sin(t)),
                                             vel=(-0.3 * math.sin(t), 0.1 * math.cos(t)),
                                             angle=0.1 * t, angular_vel=0.1)
    bodies = ib.Bodies(cyl=ib.MovingRigidBody(motion, pan))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=dt))


def pitching_plate(nx=64, ny=64, nlevel=3, h=1.0 / 32, x0=(-1.0, -1.0), Re=200.0, dt=0.002,
                   npts=40, k=0.5):
    """Flat plate of length 1 pitching about midchord, alpha = 20 + 10 sin(2 pi f t) deg,
    k = pi f (the learned-warm-start workload, reduced size)."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.0))
    pan = curves.partition(curves.LineSegment((-0.5, 0.0), (0.5, 0.0)), npts)
    f = k / math.pi
    deg = math.pi / 180

    def motion(t):
        a = 20 * deg + 10 * deg * math.sin(2 * math.pi * f * t)
        adot = 10 * deg * 2 * math.pi * f * math.cos(2 * math.pi * f * t)
        return ib.RigidBodyTransform(pos=(0.0, 0.0), vel=(0.0, 0.0), angle=-a, angular_vel=-adot)

    bodies = ib.Bodies(plate=ib.MovingRigidBody(motion, pan))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=dt))


def pitching_plate_512(k=0.5, npts=33):
    """512 x 512 cells on [-2,6] x [-4,4] (h = 1/64), 4 levels, Re 200, dt 0.002: plate of
    length 1 pitching about midchord, alpha = 20 + 10 sin(2 pi f t) deg, k = pi f. The
    points sit at the reference's panel spacing 2h (Panels(grid, curve), curves.jl:53-55):
    33 points. At 200 points (spacing 0.32 h) B is numerically singular and CG cannot reach
    a 1e-8 residual."""
    return pitching_plate(nx=512, ny=512, nlevel=4, h=1.0 / 64, x0=(-2.0, -4.0), Re=200.0,
                          dt=0.002, npts=npts, k=k)


def ellipse_body(nx=96, ny=64, nlevel=2, h=0.05, x0=(-1.2, -1.6), Re=250.0, dt=None, a=0.5,
                 b=0.2, aoa_deg=10.0, spacing=None, U=(1.0, 0.0)):
    """Stationary ellipse (semi-axes a, b) at angle of attack ``aoa_deg`` in a uniform
    stream, points spaced ``spacing`` (default h). dt defaults to CFL 0.2 (h/5)."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: U)
    curve = curves.ellipse(a, b, angle=-math.radians(aoa_deg))
    pan = curves.partition(curve, float(spacing or h))
    bodies = ib.Bodies(ellipse=ib.StaticBody(pan))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=dt or 0.2 * h))


def ensemble(nprob=64, seed=2024, sizes=(256, 512, 768, 1024), levels=(3, 4, 5)):
    """Batched-ensemble workload: finest grids N x N drawn from ``sizes`` on [-2,6] x [-4,4]
    (h = 8/N), levels from ``levels``, Re uniform in [50, 300], one ellipse per problem
    (a = 0.5, b uniform in [0.1, 0.5], angle of attack uniform in [0, 30] deg, points
    spaced h), dt = h/5 (CFL 0.2), numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        N = int(rng.choice(sizes))
        L = int(rng.choice(levels))
        Re = float(rng.uniform(50.0, 300.0))
        b = float(rng.uniform(0.1, 0.5))
        aoa = float(rng.uniform(0.0, 30.0))
        h = 8.0 / N
        out.append(ellipse_body(nx=N, ny=N, nlevel=L, h=h, x0=(-2.0, -4.0), Re=Re, a=0.5, b=b,
                                aoa_deg=aoa))
    return out


def plunging_ellipses(nx=2048, ny=1024, nlevel=6, h=1.0 / 128, x0=(-4.0, -4.0), Re=200.0,
                      dt=None, nbody=3, seed=7, amp=0.2):
    """Several ellipses plunging in a uniform stream: body k (a = 0.5, b uniform in
    [0.1, 0.3], 3 chords apart along x) moves as y = amp sin(2 pi f t + phase), vel =
    (0, 2 pi f amp cos(2 pi f t + phase)), f uniform in [0.1, 0.5], phase uniform in
    [0, 2 pi), numpy default_rng(seed). Panels spaced 2h (Panels(grid, curve)), dt from
    CFL 0.2 at the freestream speed. The full size is 2048 x 1024 on [-4,12] x [-4,4]
    with 6 levels."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.0))
    rng = np.random.default_rng(seed)
    bodies = {}
    for k in range(nbody):
        b = float(rng.uniform(0.1, 0.3))
        f = float(rng.uniform(0.1, 0.5))
        ph = float(rng.uniform(0.0, 2 * math.pi))
        xc = 3.0 * k
        pan = curves.panels(grid, curves.ellipse(0.5, b))

        def motion(t, xc=xc, f=f, ph=ph):
            w = 2 * math.pi * f
            return ib.RigidBodyTransform(pos=(xc, amp * math.sin(w * t + ph)),
                                         vel=(0.0, w * amp * math.cos(w * t + ph)),
                                         angle=0.0, angular_vel=0.0)

        bodies[f"ellipse{k + 1}"] = ib.MovingRigidBody(motion, pan)
    return ib.Problem(fluid=fluid, bodies=ib.Bodies(**bodies),
                      scheme=ib.CNAB(dt=dt or ib.default_scheme(grid, Umax=1.0, cfl=0.2).dt))


def multibody_ensemble(nprob=32, seed=2025, nx=2048, ny=1024, nlevel=6, h=1.0 / 128,
                       x0=(-4.0, -4.0), amp=0.2):
    """Combined workload: ``nprob`` independent problems of 2-4 plunging ellipses each.
    Per problem (numpy default_rng(seed)): Re uniform in [100, 1000]; 2-4 ellipses with
    semi-axes uniform in [0.1, 0.5], centres uniform in [0, 4] x [-1.5, 1.5], redrawn until
    no two bounding circles (plus the plunge travel 2 amp and 0.1) overlap; each plunges
    y = cy + amp sin(2 pi f t) with f uniform in [0.1, 0.5]. Full size: 2048 x 1024 cells
    on [-4,12] x [-4,4] (h = 1/128), 6 levels; panels spaced 2h; dt from CFL 0.2 at the
    freestream speed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
        Re = float(rng.uniform(100.0, 1000.0))
        fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.0))
        nb = int(rng.integers(2, 5))
        bodies = {}
        placed = []
        while len(placed) < nb:
            a, b = float(rng.uniform(0.1, 0.5)), float(rng.uniform(0.1, 0.5))
            cx, cy = float(rng.uniform(0.0, 4.0)), float(rng.uniform(-1.5, 1.5))
            r = max(a, b)
            if all(math.hypot(cx - px, cy - py) >= r + pr + 2 * amp + 0.1
                   for (px, py, pr) in placed):
                placed.append((cx, cy, r))
                f = float(rng.uniform(0.1, 0.5))
                pan = curves.panels(grid, curves.ellipse(a, b))

                def motion(t, cx=cx, cy=cy, f=f):
                    w = 2 * math.pi * f
                    return ib.RigidBodyTransform(pos=(cx, cy + amp * math.sin(w * t)),
                                                 vel=(0.0, w * amp * math.cos(w * t)),
                                                 angle=0.0, angular_vel=0.0)

                bodies[f"ellipse{len(placed)}"] = ib.MovingRigidBody(motion, pan)
        out.append(ib.Problem(fluid=fluid, bodies=ib.Bodies(**bodies),
                              scheme=ib.default_scheme(grid, Umax=1.0, cfl=0.2)))
    return out


def reference_solver_problem(grid_motion="static grid", bodies="single cyl"):
    """The problems of the reference's solver tests (test/solver.jl:1-63): a 3-level grid
    on [-1,2] x [-1.5,1.5] with the Re 100 default step (h = 0.02: 150 x 150 cells, not
    divisible by 4), a static or moving grid, and one cylinder, two segments, or a moving
    cylinder; dt from CFL 0.2 at Umax 2."""
    if grid_motion == "static grid":
        gm = ib.StaticGrid()
    else:
        gm = ib.MovingGrid(lambda t: ib.GridVelocity(center=(-1.0, 0.0), vel=(0.0, 0.2),
                                                     angle=0.1 * math.sin(t),
                                                     angular_vel=0.1 * math.cos(t)))
    fluid = ib.Fluid(grid=ib.fluid_grid(xlims=(-1.0, 2.0), ylims=(-1.5, 1.5), nlevel=3),
                     Re=100.0, freestream_vel=lambda t: (1.0, 0.0), grid_motion=gm)
    circle = curves.panels(fluid.grid, curves.Circle(r=0.5, center=(0.0, 0.0)))
    if bodies == "single cyl":
        b = ib.Bodies(cyl=ib.StaticBody(circle))
    elif bodies == "double line":
        b = ib.Bodies(
            line1=ib.StaticBody(curves.panels(fluid.grid, curves.LineSegment((-0.2, 0.4), (0.2, 0.1)))),
            line2=ib.StaticBody(curves.panels(fluid.grid, curves.LineSegment((-0.2, -0.1), (0.2, -0.4)))))
    else:
        motion = lambda t: ib.RigidBodyTransform(pos=(0.3 * math.cos(t), 0.1 * math.sin(t)),
                                                 vel=(-0.3 * math.sin(t), 0.1 * math.cos(t)),
                                                 angle=0.1 * t, angular_vel=0.1)
        b = ib.Bodies(cyl=ib.MovingRigidBody(motion, circle))
    return ib.Problem(fluid=fluid, bodies=b,
                      scheme=ib.default_scheme(fluid.grid, Umax=2.0, cfl=0.2))


def reference_outside_region_problem():
    """test/solver.jl:66-79: a cylinder too close to the edge of a 15 x 20 grid."""
    fluid = ib.Fluid(grid=ib.fluid_grid(h=0.1, xlims=(-1.0, 0.5), ylims=(-1.0, 1.0), nlevel=3),
                     Re=20.0, freestream_vel=lambda t: (1.0, 0.0))
    b = ib.Bodies(cyl=ib.StaticBody(curves.panels(fluid.grid, curves.Circle(r=0.5, center=(0.0, 0.0)))))
    return ib.Problem(fluid=fluid, bodies=b, scheme=ib.default_scheme(fluid.grid, Umax=2.0, cfl=0.2))


def naca_airfoil(spec="0012", aoa_deg=20.0, Re=400.0, nlevel=3):
    """The reference's NACA example (docs/src/examples/naca.md): the section, its
    mid-chord at the origin, turned to ``aoa_deg`` incidence, on [-1,2] x [-0.5,0.5] with
    the Reynolds-number default step, panels spaced 2h, dt from CFL 0.2 at Umax 2."""
    grid = ib.fluid_grid(xlims=(-1.0, 2.0), ylims=(-0.5, 0.5), nlevel=nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.0))
    airfoil = (curves.naca(spec) | curves.translate((-0.5, 0.0))
               | curves.rotate(-math.radians(aoa_deg)))
    bodies = ib.Bodies(airfoil=ib.StaticBody(curves.panels(fluid.grid, airfoil)))
    scheme = ib.default_scheme(fluid.grid, Umax=2.0, cfl=0.2)
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=scheme)


def no_body(nx=32, ny=32, nlevel=2, h=0.1, x0=(-1.6, -1.6), Re=50.0, dt=0.01):
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: (1.0, 0.3))
    return ib.Problem(fluid=fluid, bodies=ib.Bodies(), scheme=ib.CNAB(dt=dt))


def rotating_frame(nx=64, ny=64, nlevel=3, h=0.05, x0=(-1.6, -1.6), Re=100.0, dt=0.004,
                   omega=0.3, vel=(0.2, -0.1), U=(1.0, 0.0), body=True):
    """MovingGrid: the grid translates with ``vel`` and turns at ``omega`` about the origin
    (angle = omega t), a static ellipse in the grid frame (the body-fixed frame of a
    rotating body)."""
    grid = ib.MultiDomainGrid(ib.CartesianGrid(tuple(x0), (nx, ny), h), nlevel)
    motion = ib.MovingGrid(lambda t: ib.GridVelocity(center=(0.0, 0.0), vel=vel,
                                                     angle=omega * t, angular_vel=omega))
    fluid = ib.Fluid(grid=grid, Re=Re, freestream_vel=lambda t: U, grid_motion=motion)
    bodies = ib.Bodies()
    if body:
        pan = curves.partition(curves.ellipse(0.4, 0.15, angle=0.3), 2.0 * h)
        bodies = ib.Bodies(ellipse=ib.StaticBody(pan))
    return ib.Problem(fluid=fluid, bodies=bodies, scheme=ib.CNAB(dt=dt))


def oracle_for(prob, **kw):
    """FP64 CPU oracle of ``prob`` (test infrastructure)."""
    import oracle as O

    g = prob.fluid.grid
    grid = O.Grid(g.base.n[0], g.base.n[1], g.nlevel, g.base.x0[0], g.base.x0[1], g.base.h)
    b = prob.bodies
    bodies = [(off, n, body.moving) for (_, body, off, n) in b.sections]
    kw.setdefault("grid_moving", isinstance(prob.fluid.grid_motion, ib.MovingGrid))
    return O.OracleSolver(grid, prob.fluid.Re, prob.scheme.dt, b.xb(), b.ds(), bodies, **kw)


def oracle_advance(orc, prob, nsteps=1):
    """Advance the oracle with the freestream and motion the host evaluates."""
    forces = []
    gm = prob.fluid.grid_motion
    for _ in range(nsteps):
        t = orc.next_time()
        orc.advance(prob.fluid.freestream_vel(t), prob.bodies.motion_array(t),
                    gm(t).as_array() if isinstance(gm, ib.MovingGrid) else None)
        forces.append(orc.field("total_force"))
    return np.array(forces)


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))
