"""Initial-condition generators for the benchmark configurations, in the
reference's SpaceVariable3D layout (ghosted AoS, see include/m2c_amd.h).

These restate what SpaceInitializer does for planes and spheres
(SpaceInitializer.cpp:237-482: default state everywhere, then geometries in
order; level set by signed distance, SpaceInitializer.cpp:1390-1514, negative
inside the material it tracks), followed by the boundary fill.
"""
import math

import numpy as np

from . import abi


def _axis(N, a0, a1):
    """Cell centres with one ghost each side (MeshGenerator.cpp:161-173,
    SpaceOperator.cpp:259-327)."""
    h = (a1 - a0) / N
    c = np.empty(N + 2)
    c[1:-1] = a0 + 0.5 * h + np.arange(N) * h
    c[0] = c[1] - h
    c[-1] = c[-2] + h
    return c


def fill_ghosts(V, ID, Phi, grid):
    """ApplyBoundaryConditions (SpaceOperator.cpp:738-821) for OUTLET /
    SYMMETRY-type faces, UpdateMaterialIDAtGhostNodes and the level-set
    ZeroNeumann fill, on the host, for initial states."""
    bcs = list(grid.bc)
    for ax in range(3):  # array axes: V[k, j, i]
        arr_ax = 2 - ax
        for side in (0, 1):
            bc = bcs[2 * ax + side]
            g = 0 if side == 0 else -1
            im = 1 if side == 0 else -2
            sl_g = [slice(1, -1)] * 3
            sl_i = [slice(1, -1)] * 3
            sl_g[arr_ax] = g
            sl_i[arr_ax] = im
            sl_g, sl_i = tuple(sl_g), tuple(sl_i)
            ID[sl_g] = ID[sl_i]
            if Phi is not None:
                Phi[sl_g] = Phi[sl_i]
            if bc == abi.BC_INLET:
                V[sl_g] = list(grid.inlet)
            elif bc == abi.BC_INLET2:
                V[sl_g] = list(grid.inlet2)
            else:
                V[sl_g] = V[sl_i]
                if bc in (abi.BC_SYMMETRY, abi.BC_SLIPWALL):
                    V[sl_g + (ax + 1,)] *= -1.0
                elif bc == abi.BC_STICKWALL:
                    V[sl_g + (slice(1, 4),)] *= -1.0


def c1_sod(n=(128, 32, 32)):
    """3D Sod shock tube on the unit box, single stiffened-gas material
    (gamma 1.4, p_c 0), MUSCL primitive + van Albada, HLLC, RK2, CFL 0.5,
    transmissive (Outlet) boundaries."""
    g = abi.grid(n, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (abi.BC_OUTLET,) * 6)
    mats = abi.materials(abi.material("SG", gamma=1.4, pc=0.0))
    s = abi.default_scheme(flux=abi.FLUX_HLLC, limiter=abi.LIM_VANALBADA, integrator=abi.TI_RK2, cfl=0.5)
    x = _axis(n[0], 0.0, 1.0)
    V = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2, 5))
    V[..., 0] = np.where(x < 0.5, 1.0, 0.125)[None, None, :]
    V[..., 4] = np.where(x < 0.5, 1.0, 0.1)[None, None, :]
    ID = np.zeros(V.shape[:3], dtype=np.int32)
    fill_ghosts(V, ID, None, g)
    return g, mats, 1, s, V, ID, None


WATER_SG = dict(gamma=6.12, pc=3.43e8)
AIR_SG = dict(gamma=1.4, pc=0.0)
WATER_NASG = dict(gamma=1.19, pc=6.2178e8, b=6.7212e-4, q=-1.177788e6)


def sg_post_shock(rho1, p1, gam, pc, mach):
    """Rankine-Hugoniot state behind a right-moving shock of Mach `mach`
    entering quiescent stiffened gas (ideal-gas relations in p + p_c)."""
    c1 = math.sqrt(gam * (p1 + pc) / rho1)
    m2 = mach * mach
    rho2 = rho1 * (gam + 1) * m2 / ((gam - 1) * m2 + 2)
    p2 = (p1 + pc) * (1 + 2 * gam / (gam + 1) * (m2 - 1)) - pc
    u2 = mach * c1 * (1 - rho1 / rho2)
    return rho2, u2, p2


def c2_shock_bubble(n=256, L=1.0, radius=0.15, mach=1.5, shock_x=0.2, tol=1e-10):
    """Two-material shock-bubble: stiffened-gas water with one centred ideal-gas
    (SG, gamma 1.4) air bubble of radius 0.15 L, hit by a planar Mach-1.5 shock
    placed at x = 0.2 L travelling +x.  Exact interface Riemann solver with
    tolerance 1e-10, HLLC elsewhere, MUSCL primitive van Albada, RK3, CFL 0.4,
    one level set (material 1 = air) with ZeroNeumann boundaries, transmissive
    boundaries."""
    g = abi.grid((n, n, n), (0.0, 0.0, 0.0), (L, L, L), (abi.BC_OUTLET,) * 6)
    mats = abi.materials(abi.material("SG", **WATER_SG), abi.material("SG", **AIR_SG))
    s = abi.default_scheme(flux=abi.FLUX_HLLC, limiter=abi.LIM_VANALBADA, integrator=abi.TI_RK3, cfl=0.4,
                           n_levelsets=1, ls_materialid=1, riem_tol_main=tol)
    x = _axis(n, 0.0, L)
    rho_w, p_w = 1000.0, 1.0e5
    rho2, u2, p2 = sg_post_shock(rho_w, p_w, WATER_SG["gamma"], WATER_SG["pc"], mach)
    V = np.empty((n + 2, n + 2, n + 2, 5))
    V[...] = [rho_w, 0.0, 0.0, 0.0, p_w]
    post = (x < shock_x * L)
    V[:, :, post, :] = [rho2, u2, 0.0, 0.0, p2]
    ID = np.zeros(V.shape[:3], dtype=np.int32)
    c = 0.5 * L
    R = radius * L
    X = x[None, None, :]
    Y = x[None, :, None]
    Z = x[:, None, None]
    dist = np.sqrt((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2) - R
    Phi = np.array(np.broadcast_to(dist, V.shape[:3]), dtype=np.float64, order="C")
    inside = dist < 0
    V[inside] = [1.2, 0.0, 0.0, 0.0, p_w]
    ID[inside] = 1
    fill_ghosts(V, ID, Phi, g)
    return g, mats, 2, s, V, ID, Phi


def c4_instances(count=64, seed=4, edges=(64, 96, 128, 160, 192)):
    """Uneven ensemble of shock-bubble runs: per instance the grid edge is
    drawn from `edges`, the bubble radius uniformly from [0.05, 0.25] L and the
    shock Mach number from [1.2, 3.0] (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(edges))
        r = float(rng.uniform(0.05, 0.25))
        m = float(rng.uniform(1.2, 3.0))
        out.append(((n, r, m), c2_shock_bubble(n, radius=r, mach=m)))
    return out


def nasg_post_shock(rho1, p1, gam, pinf, b, q, mach):
    """Rankine-Hugoniot state behind a right-moving shock of Mach `mach`
    entering quiescent Noble-Abel stiffened gas (e = (p + gam pinf)(v - b) /
    (gam - 1) + q, VarFcnNASG.h:55-86): mass and momentum give p2 and v2 as
    functions of the particle velocity u2, and the Hugoniot energy relation
    e2 - e1 = (p1 + p2)(v1 - v2)/2 fixes u2 (one root in (0, Us(1 - b/v1)))."""
    from scipy.optimize import brentq
    v1 = 1.0 / rho1
    e1 = (p1 + gam * pinf) * (v1 - b) / (gam - 1) + q
    c1 = math.sqrt(gam * (p1 + pinf) * v1 * v1 / (v1 - b))
    us = mach * c1

    def hug(u2):
        p2 = p1 + rho1 * us * u2
        v2 = v1 * (us - u2) / us
        e2 = (p2 + gam * pinf) * (v2 - b) / (gam - 1) + q
        return e2 - e1 - 0.5 * (p1 + p2) * (v1 - v2)

    hi = us * (1.0 - b / v1) * (1.0 - 1e-12)
    lo = 1e-6 * us
    u2 = brentq(hug, lo, hi, xtol=1e-15 * us, rtol=1e-15, maxiter=500)
    p2 = p1 + rho1 * us * u2
    rho2 = 1.0 / (v1 * (us - u2) / us)
    return rho2, u2, p2


def shock_bubble_geometries(L, radius, mach, shock_x, water, rho_w=1000.0, p_w=1.0e5, centre=None):
    """The shock-bubble initial condition as geometric entities for
    Solver.initialize: default quiescent water (material 0), a plane behind
    which the post-shock state holds (x < shock_x), an air sphere (material
    1, the level-set material)."""
    if water.get("b", 0.0) or water.get("q", 0.0):
        rho2, u2, p2 = nasg_post_shock(rho_w, p_w, water["gamma"], water["pc"], water.get("b", 0.0),
                                       water.get("q", 0.0), mach)
    else:
        rho2, u2, p2 = sg_post_shock(rho_w, p_w, water["gamma"], water["pc"], mach)
    c = centre if centre is not None else (0.5 * L, 0.5 * L, 0.5 * L)
    plane = abi.geometry("plane", (shock_x, 0.0, 0.0), normal=(-1.0, 0.0, 0.0), materialid=0,
                         state=(rho2, u2, 0.0, 0.0, p2))
    sphere = abi.geometry("sphere", c, radius=radius, materialid=1, state=(1.2, 0.0, 0.0, 0.0, p_w))
    return (rho_w, 0.0, 0.0, 0.0, p_w), 0, [plane, sphere]


def c5_bubble_cloud(n=(512, 256, 256), nbubbles=200, seed=5, mach=2.0, shock_x=0.1, r_range=(0.005, 0.03),
                    gap=0.002, reinit_every=10):
    """Bubble cloud: NASG water (gamma 1.19, p_inf 6.2178e8, b 6.7212e-4,
    q -1.177788e6) on [0, 2L] x [0, L] x [0, L] with `nbubbles` non-overlapping
    ideal-gas bubbles (radius log-uniform in r_range x L, seeded centres by
    rejection, all downstream of the shock), hit by a planar Mach-2 shock.
    HLLC bulk, exact interface Riemann, RK3, the level set reinitialized every
    `reinit_every` steps (full domain, Fixed first layer, the reference's
    defaults otherwise).  Returns (grid, mats, nmat, scheme, default_state,
    default_id, geometries) for Solver.initialize."""
    L = 1.0
    lx = L * n[0] / n[2]
    g = abi.grid(n, (0.0, 0.0, 0.0), (lx, L, L), (abi.BC_OUTLET,) * 6)
    mats = abi.materials(abi.material("NASG", **WATER_NASG), abi.material("SG", **AIR_SG))
    s = abi.default_scheme(flux=abi.FLUX_HLLC, limiter=abi.LIM_VANALBADA, integrator=abi.TI_RK3, cfl=0.4,
                           n_levelsets=1, ls_materialid=1, riem_tol_main=1e-10, reinit_frequency=reinit_every)
    rng = np.random.default_rng(seed)
    placed = []
    tries = 0
    while len(placed) < nbubbles and tries < 200000:
        tries += 1
        r = L * 10 ** rng.uniform(np.log10(r_range[0]), np.log10(r_range[1]))
        cx = rng.uniform(shock_x * lx + r + gap, lx - r - gap)
        cy = rng.uniform(r + gap, L - r - gap)
        cz = rng.uniform(r + gap, L - r - gap)
        ok = all((cx - x) ** 2 + (cy - y) ** 2 + (cz - z) ** 2 > (r + rr + gap) ** 2 for x, y, z, rr in placed)
        if ok:
            placed.append((cx, cy, cz, r))
    rho_w, p_w = 1000.0, 1.0e5
    rho2, u2, p2 = nasg_post_shock(rho_w, p_w, WATER_NASG["gamma"], WATER_NASG["pc"], WATER_NASG["b"],
                                   WATER_NASG["q"], mach)
    geoms = [abi.geometry("plane", (shock_x * lx, 0.0, 0.0), normal=(-1.0, 0.0, 0.0), materialid=0,
                          state=(rho2, u2, 0.0, 0.0, p2))]
    for cx, cy, cz, r in placed:
        geoms.append(abi.geometry("sphere", (cx, cy, cz), radius=r, materialid=1, state=(1.2, 0.0, 0.0, 0.0, p_w)))
    return g, mats, 2, s, (rho_w, 0.0, 0.0, 0.0, p_w), 0, geoms


def c2_geometries(n=256, L=1.0, radius=0.15, mach=1.5, shock_x=0.2, tol=1e-10):
    """c2_shock_bubble as geometric entities for Solver.initialize (no host
    arrays): grid, materials, scheme, default state and ID, geometries."""
    g, mats, nmat, s = _c2_setup(n, L, tol)
    st, sid, geoms = shock_bubble_geometries(L, radius * L, mach, shock_x * L, WATER_SG)
    return g, mats, nmat, s, st, sid, geoms


def _c2_setup(n, L, tol):
    g = abi.grid((n, n, n), (0.0, 0.0, 0.0), (L, L, L), (abi.BC_OUTLET,) * 6)
    mats = abi.materials(abi.material("SG", **WATER_SG), abi.material("SG", **AIR_SG))
    s = abi.default_scheme(flux=abi.FLUX_HLLC, limiter=abi.LIM_VANALBADA, integrator=abi.TI_RK3, cfl=0.4,
                           n_levelsets=1, ls_materialid=1, riem_tol_main=tol)
    return g, mats, 2, s


def c4_geometries(count=64, seed=4, edges=(64, 96, 128, 160, 192)):
    """c4_instances as geometric entities: the same draws (edge, radius,
    Mach) per instance, each built on the device by Solver.initialize."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(edges))
        r = float(rng.uniform(0.05, 0.25))
        m = float(rng.uniform(1.2, 3.0))
        out.append(((n, r, m), c2_geometries(n, radius=r, mach=m)))
    return out


def brick_split(n, parts):
    """Factor `parts` into (px, py, pz) dividing the grid n into bricks as
    close to cubes as possible (cells per brick axis n_a / p_a)."""
    best = None
    for px in range(1, parts + 1):
        if parts % px:
            continue
        for py in range(1, parts // px + 1):
            if (parts // px) % py:
                continue
            pz = parts // px // py
            if n[0] % px or n[1] % py or n[2] % pz:
                continue
            b = (n[0] // px, n[1] // py, n[2] // pz)
            score = max(b) / min(b)
            if best is None or score < best[0]:
                best = (score, (px, py, pz))
    if best is None:
        return (parts, 1, 1)
    return best[1]


def initialize_host(grid, scheme, default_state, default_id, geoms, lo=(0, 0, 0), n=None):
    """What Solver.initialize builds on the device (planes and spheres in
    order, the level set as the signed distance of the nearest surface of its
    material; SpaceInitializer.cpp:237-482, 1390-1514), on the host, for the
    brick of `n` cells starting at cell `lo` of `grid`.  Returns the brick's
    grid (its own coordinates, Outlet boundaries) and its ghosted V, ID, Phi."""
    n = n or (grid.nx, grid.ny, grid.nz)
    full = (grid.nx, grid.ny, grid.nz)
    x0 = (grid.x0, grid.y0, grid.z0)
    x1 = (grid.xmax, grid.ymax, grid.zmax)
    h = [(x1[a] - x0[a]) / full[a] for a in range(3)]
    blo = tuple(x0[a] + lo[a] * h[a] for a in range(3))
    bhi = tuple(x0[a] + (lo[a] + n[a]) * h[a] for a in range(3))
    bg = abi.grid(n, blo, bhi, (abi.BC_OUTLET,) * 6)
    xs = [_axis(n[a], blo[a], bhi[a]) for a in range(3)]
    X = xs[0][None, None, :]
    Y = xs[1][None, :, None]
    Z = xs[2][:, None, None]
    shape = (n[2] + 2, n[1] + 2, n[0] + 2)
    V = np.empty(shape + (5,))
    V[...] = list(default_state)
    ID = np.full(shape, default_id, dtype=np.int32)
    diag = math.sqrt(sum((x1[a] - x0[a]) ** 2 for a in range(3)))
    Phi = np.full(shape, diag)
    for gm in geoms:
        if gm.kind == abi.GEOM_PLANE:
            nv = np.array(gm.normal[:])
            nv = nv * (1.0 / math.sqrt(nv @ nv))
            dist = (X - gm.point[0]) * nv[0] + (Y - gm.point[1]) * nv[1] + (Z - gm.point[2]) * nv[2]
            inside = dist > 0
        else:
            a, b, e = X - gm.point[0], Y - gm.point[1], Z - gm.point[2]
            dist = np.sqrt(a * a + b * b + e * e) - gm.radius
            inside = dist < 0 if gm.side == 0 else dist > 0
        dist = np.broadcast_to(dist, shape)
        inside = np.broadcast_to(inside, shape)
        V[inside] = list(gm.state)
        ID[inside] = gm.materialid
        if scheme.n_levelsets and gm.materialid == scheme.ls_materialid:
            upd = np.abs(dist) < np.abs(Phi)
            if gm.kind == abi.GEOM_PLANE:
                Phi[upd] = -dist[upd]
            else:
                Phi[upd] = (dist if gm.side == 0 else -dist)[upd]
    fill_ghosts(V, ID, Phi if scheme.n_levelsets else None, bg)
    return bg, V, ID, (Phi if scheme.n_levelsets else None)
