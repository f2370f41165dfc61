"""The direction of the interface Riemann problems (GetNormalForBimaterialRiemann,
SpaceOperator.cpp:2186-2415) in the oracle, pinned by what the mathematics
fixes: for a linear level set the face gradient is exact, so LevelSet gives
the plane's normal (oriented from the level set's material into material 0,
its axis component made non-negative), Average the normalized sum with the
axis, Mesh the axis; for a sphere the normal is radial to O(h^2)."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi, configs


def _solver(phi_fn, rnormal, n=(12, 10, 8)):
    g = abi.grid(n, (0, 0, 0), (1.2, 1.0, 0.8))
    mats = abi.materials(abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, riemann_normal=rnormal)
    x = np.array(configs._axis(n[0], 0, 1.2))[None, None, :]
    y = np.array(configs._axis(n[1], 0, 1.0))[None, :, None]
    z = np.array(configs._axis(n[2], 0, 0.8))[:, None, None]
    phi = np.broadcast_to(phi_fn(x, y, z), (n[2] + 2, n[1] + 2, n[0] + 2)).copy()
    V = np.zeros(phi.shape + (5,))
    V[..., 0] = V[..., 4] = 1.0
    ID = (phi < 0).astype(np.int32)
    o = ol.Oracle(g, mats, 2, s)
    o.set_state(V, ID, phi)
    return o


NRM = np.array([0.6, 0.48, 0.64])


def _plane(x, y, z):
    return NRM[0] * (x - 0.6) + NRM[1] * (y - 0.5) + NRM[2] * (z - 0.4)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_plane_normal(d):
    cell = (6, 5, 4)
    ijk = list(cell)
    for rn in (abi.RNORMAL_LEVEL_SET, abi.RNORMAL_AVERAGE, abi.RNORMAL_MESH):
        o = _solver(_plane, rn)
        e = np.eye(3)[d]
        # the minus-side cell in the level set's material (1), this cell in 0
        a = o.bimaterial_normal(d, *ijk, 0, 1)
        # the other way round: the gradient is flipped, then its axis part
        b = o.bimaterial_normal(d, *ijk, 1, 0)
        if rn == abi.RNORMAL_MESH:
            assert np.array_equal(a, e) and np.array_equal(b, e)
            continue
        flipped = -NRM.copy()
        flipped[d] *= -1.0
        if rn == abi.RNORMAL_LEVEL_SET:
            assert np.allclose(a, NRM, rtol=0, atol=1e-14), a
            assert np.allclose(b, flipped, rtol=0, atol=1e-14), b
        else:
            assert np.allclose(a, (NRM + e) / np.linalg.norm(NRM + e), rtol=0, atol=1e-14), a
            assert np.allclose(b, (flipped + e) / np.linalg.norm(flipped + e), rtol=0, atol=1e-14), b


def test_sphere_normal_is_radial_to_second_order():
    c0 = np.array([0.6, 0.5, 0.4])
    errs = []
    for n in ((12, 10, 8), (24, 20, 16), (48, 40, 32)):
        o = _solver(lambda x, y, z: np.sqrt((x - c0[0]) ** 2 + (y - c0[1]) ** 2 + (z - c0[2]) ** 2) - 0.3,
                    abi.RNORMAL_LEVEL_SET, n)
        h = 1.2 / n[0]
        # x face at the same physical point, right of the centre (r = 0.3,
        # angle with the axes away from 0)
        p = c0 + 0.3 * np.array([0.8, 0.36, 0.48])
        i = int(round(p[0] / h))
        j, k = int(p[1] / h), int(p[2] / h)
        xf = np.array([i * h, (j + 0.5) * h, (k + 0.5) * h])
        r = (xf - c0) / np.linalg.norm(xf - c0)
        dvec = o.bimaterial_normal(0, i, j, k, 0, 1)
        errs.append(np.abs(dvec - r).max())
    assert errs[0] < 0.05
    assert errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3, errs
