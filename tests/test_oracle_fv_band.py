"""The finite-volume level set in a narrow band (LevelSetOperator::
ReconstructInBand, ComputeAdvectionFluxInBand, AddSourceTermInBand,
LevelSetOperator.cpp:966-1062, 1160-1232, 1288-1306).  No reference deck
runs it, so it is pinned by hand-derived closed forms on a line where every
value is a binary fraction: the reconstruction writes the ACTIVE nodes only,
into buffers kept between calls, and the flux through a face of a useful node
that is not active reads what those buffers held -- zero before any write,
else the node's values from the last call in which it was active.  Away from
the band's edge it agrees with the full-domain scheme (pinned by Test4) up to
the summation order, which follows the useful-node list."""
import numpy as np

import oracle_lib as ol
from m2c_amd import abi, configs


def _line(bw=2, nx=10):
    g = abi.grid((nx, 1, 1), (0.0, 0.0, 0.0), (float(nx), 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, ls_materialid=1, ls_solver=abi.LSSOLVER_FINITE_VOLUME,
                           ls_flux=abi.LSFLUX_ROE, reinit_frequency=1, ls_bandwidth=bw)
    o = ol.Oracle(g, mats, 2, s)
    x = np.broadcast_to(configs._axis(nx, 0.0, float(nx))[None, None, :], o.shape)
    return o, x, s


def _state(o, x, u, x0):
    V = np.zeros(o.shape + (5,))
    V[..., 0] = 1.0
    V[..., 1] = u
    V[..., 4] = 1.0
    phi = x - x0
    o.set_state(V, np.where(phi < 0, 1, 0).astype(np.int32), phi)
    o.construct_narrow_band()


def _roe(pm, pp, um, up, delta):
    """the Roe flux of LevelSetOperator::ComputeLocalAdvectionFlux (:1248-1253)"""
    lam = abs(0.5 * (um + up))
    tol = max(abs(um), abs(up)) * delta
    if lam < tol:
        lam = (lam * lam + tol * tol) / (2 * tol)
    return 0.5 * (pm * um + pp * up) - 0.5 * lam * (pp - pm)


def test_band_edge_reads_zero_before_a_first_write():
    """phi = x - 4.75, u = 1, bandwidth 2: cells 4 and 5 (level 1) are active,
    cells 3 and 6 (level 2) useful only.  The faces 3|4 and 5|6 take the
    never-written (zero) face values of cells 3 and 6."""
    o, x, s = _line()
    _state(o, x, 1.0, 4.75)
    lev, use, act, _ = o.narrow_band()
    row = lev[1, 1, 1:-1]
    assert row[3] == 2 and row[4] == 1 and row[5] == 1 and row[6] == 2
    R = o.ls_residual()[1, 1, 1:-1]
    # linear phi: slope 1, faces phi -+ 1/2; u uniform: faces 1
    f34 = _roe(0.0, -0.75, 0.0, 1.0, s.ls_delta)      # stale zeros of cell 3
    f45 = _roe(0.25, 0.25, 1.0, 1.0, s.ls_delta)
    f56 = _roe(1.25, 0.0, 1.0, 0.0, s.ls_delta)       # stale zeros of cell 6
    assert R[4] == f34 - f45 == -0.4375
    assert R[5] == f45 - f56 == -0.6875
    assert R[3] == 0.0 and R[6] == 0.0                 # useful, not active
    # the full domain gives -u dphi/dx = -1 at both
    full = ol.Oracle(o.grid, o.mats, 2, abi.default_scheme(n_levelsets=1, ls_materialid=1,
                                                           ls_solver=abi.LSSOLVER_FINITE_VOLUME))
    V = np.zeros(o.shape + (5,))
    V[..., 0] = 1.0
    V[..., 1] = 1.0
    V[..., 4] = 1.0
    full.set_state(V, np.zeros(o.shape, np.int32), x - 4.75)
    Rf = full.ls_residual()[1, 1, 1:-1]
    assert Rf[4] == -1.0 and Rf[5] == -1.0


def test_band_edge_reads_the_values_of_its_last_active_call():
    """After the first call, the band moves one cell left (phi = x - 3.75) and
    u = -1: cell 5 is now at the band's edge and still holds phi_l = 0.25 and
    u_l = 1 from the first call; cell 2 was never active (zeros)."""
    o, x, s = _line()
    _state(o, x, 1.0, 4.75)
    o.ls_residual()
    _state(o, x, -1.0, 3.75)
    R = o.ls_residual()[1, 1, 1:-1]
    f23 = _roe(0.0, -0.75, 0.0, -1.0, s.ls_delta)     # cell 2: zeros
    f34 = _roe(0.25, 0.25, -1.0, -1.0, s.ls_delta)
    f45 = _roe(1.25, 0.25, -1.0, 1.0, s.ls_delta)     # cell 5: last call's values
    assert R[3] == f23 - f34 == 0.8125
    assert R[4] == f34 - f45
    assert abs(R[4] - 0.2) < 1e-15  # fresh values would give -u dphi/dx = 1



def test_band_interior_matches_the_full_domain():
    """A sphere in a swirling velocity field: on the active nodes two or more
    levels inside the band's edge every stencil is fresh, so the in-band
    residual equals the full-domain one up to the summation order of each
    node's face fluxes (list order against (k, j, i) order): |diff| <=
    4e-15 max|R|."""
    n = 16
    g = abi.grid((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    kw = dict(n_levelsets=1, ls_materialid=1, ls_solver=abi.LSSOLVER_FINITE_VOLUME, ls_flux=abi.LSFLUX_ROE)
    band = ol.Oracle(g, mats, 2, abi.default_scheme(reinit_frequency=2, ls_bandwidth=5, **kw))
    full = ol.Oracle(g, mats, 2, abi.default_scheme(**kw))
    a = np.array(configs._axis(n, 0.0, 1.0))
    Z, Y, X = np.meshgrid(a, a, a, indexing="ij")
    V = np.zeros(band.shape + (5,))
    V[..., 0] = 1.0
    V[..., 1] = np.sin(np.pi * X) * np.cos(np.pi * Y)
    V[..., 2] = -np.cos(np.pi * X) * np.sin(np.pi * Y) + 0.3 * Z
    V[..., 3] = 0.2 * np.sin(2 * np.pi * Y)
    V[..., 4] = 1.0
    phi = np.sqrt((X - 0.45) ** 2 + (Y - 0.5) ** 2 + (Z - 0.55) ** 2) - 0.27
    ID = np.where(phi < 0, 1, 0).astype(np.int32)
    for o in (band, full):
        o.set_state(V, ID, phi)
    band.construct_narrow_band()
    Rb, Rf = band.ls_residual(), full.ls_residual()
    lev, use, act, _ = band.narrow_band()
    inner = np.zeros(lev.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    sel = inner & (lev <= 5 - 2)
    assert sel.sum() > 500
    scale = np.abs(Rf[sel]).max()
    assert np.abs(Rb[sel] - Rf[sel]).max() <= 4e-15 * scale
    # on the level just inside the edge the stale (zero) face values enter
    edge = inner & (lev == 5 - 1)
    assert np.abs(Rb[edge] - Rf[edge]).max() > 1e-3 * scale
