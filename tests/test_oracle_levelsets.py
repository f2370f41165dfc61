"""The oracle with several level sets (Main.cpp:285-330;
MultiPhaseOperator::UpdateMaterialIDByLevelSet, MultiPhaseOperator.cpp:
153-357; ResolveConflictsInLevelSets, :1705-1762; SpaceOperator::
CalculateGradPhiAtCellInterface, SpaceOperator.cpp:2211-2268;
TimeIntegrator.cpp:720-800).  No reference deck with two level sets stays
inside the built EOS set (the one that exists, HVImpactShafquatIslam/
1D_TaArSLG_5kms, needs the extended Mie-Gruneisen EOS), so:

- two separated bubbles of two materials with the same EOS, each tracked by
  its own level set, evolve bit for bit as one level set tracking both (the
  single-level-set path is pinned by the reference's decks): the velocities,
  pressures and densities everywhere, the material IDs (2 read as 1), and
  each level set near its own bubble;
- ResolveConflictsInLevelSets on hand-made overlaps: the owner is the level
  set with the largest |phi| (the first on ties), the others get +mean|phi|;
  nodes are taken in (k, j, i) order and see the updates of earlier nodes
  (a node that becomes a boundary node only through its updated neighbour
  is resolved too);
- the Riemann normal between two tracked materials: for two planar level
  sets the difference of their unit normals, normalized (closed form);
- the optional corrections (LevelSetCorrectionFrequency): part I's branches
  for one owner and for a node on a boundary (phi = 0), uncounted, and part
  II's isolated background nodes in (k, j, i) order, on hand-made cases
  whose results follow from the rules by hand;
- the material-ID update: a node leaving one subdomain for another takes
  the other's material; entering none, its untracked neighbours' material;
  entering two is an error.
"""
import numpy as np
import pytest

import levelset_cases as lc
import oracle_lib as ol
from m2c_amd import abi, configs


def _start(o, V, ID, phis=None, phi=None):
    o.set_state(V, ID, phi if phi is not None else phis[0])
    if phis is not None:
        for q in range(1, len(phis)):
            o.set_levelset(q, phis[q])
    o.apply_bc()
    o.update_ghost_ids()
    o.ls_apply_bc()


@pytest.mark.parametrize("rnormal", [abi.RNORMAL_MESH, abi.RNORMAL_LEVEL_SET])
def test_two_levelsets_match_one_levelset_over_both_bubbles(rnormal):
    g, mats, nmat, s, V, ID, phis, ls, union = lc.two_bubbles(rnormal=rnormal)
    two = ol.Oracle(g, mats, nmat, s, levelsets=ls)
    _start(two, V, ID, phis=phis)
    one = ol.Oracle(g, mats, nmat, s)
    _start(one, V, np.where(ID == 2, 1, ID), phi=union)
    for step in range(12):
        two.run_steps(1)
        one.run_steps(1)
        V2, ID2, _ = two.get_state()
        V1, ID1, P1 = one.get_state()
        assert np.array_equal(V2, V1), step
        assert np.array_equal(np.where(ID2 == 2, 1, ID2), ID1), step
    # each level set equals the single one where its own bubble is the
    # nearer (a few cells off the midplane: the FD3 stencil reaches 3 cells)
    x = configs._axis(g.nx, 0.0, 2.0)
    for q, side in ((0, x < 0.8), (1, x > 1.2)):
        Pq = two.levelset(q)
        assert np.array_equal(Pq[..., side], P1[..., side]), q
    c2, c1 = two.counters(), one.counters()
    assert c2["id_changed"] == c1["id_changed"] > 0
    assert c2["riemann_calls"] == c1["riemann_calls"] > 0


def _line(nx=8, n2=None, correction_frequency=-1):
    """nx x 1 x 1 cells on [0, nx]: centres 0.5, 1.5, ..."""
    g = abi.grid((nx, 1, 1), (0.0, 0.0, 0.0), (float(nx), 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"), abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(mp_ls_correction_frequency=correction_frequency)
    ls = [abi.levelset_from_scheme(s, materialid=m) for m in (1, 2, 3)[:n2 or 2]]
    o = ol.Oracle(g, mats, 4, s, levelsets=ls)
    x = configs._axis(nx, 0.0, float(nx))
    return o, np.broadcast_to(x[None, None, :], o.shape)


def test_resolve_conflicts_largest_magnitude_wins():
    o, x = _line()
    phi0 = x - 3.0           # subdomain 0: x < 3 (cells 0, 1, 2)
    phi1 = 1.8 - x           # subdomain 1: x > 1.8 (cells 2 .. 7)
    o.set_levelset(0, phi0)
    o.set_levelset(1, phi1)
    # cell 2 (x = 2.5) lies in both, next to both boundaries, in all nine
    # (j, k) rows of the ghosted cross-section
    assert o.resolve_conflicts() == 9
    p0, p1 = o.levelset(0), o.levelset(1)
    i2 = 3  # ghosted index of cell 2
    m = (abs(2.5 - 3.0) + abs(1.8 - 2.5)) / 2  # mean |phi| of the owners
    assert np.all(p0[..., i2] == m) and np.all(p1[..., i2] == -m)  # |-0.7| > |-0.5|: level set 1 owns it
    rest = np.ones(x.shape, bool)
    rest[..., i2] = False
    assert np.array_equal(p0[rest], phi0[rest]) and np.array_equal(p1[rest], phi1[rest])


def test_resolve_conflicts_tie_goes_to_the_first_level_set():
    o, x = _line()
    o.set_levelset(0, x - 3.0)
    o.set_levelset(1, 2.0 - x)  # -0.5 at cell 2, as level set 0
    assert o.resolve_conflicts() == 9
    assert np.all(o.levelset(0)[..., 3] == -0.5) and np.all(o.levelset(1)[..., 3] == 0.5)


def test_resolve_conflicts_in_order():
    """Cells 2 and 3 lie in both subdomains.  Cell 3 has level-set-0
    neighbours inside subdomain 0 on both sides, so it is a boundary node of
    level set 0 only once cell 2 has been given to level set 1 -- which the
    reference does first, in (k, j, i) order."""
    o, x = _line(10)
    phi0 = np.where((x > 2.0) & (x < 6.0), -1.0, 1.0)    # cells 2 .. 5
    phi1 = np.where((x > 2.0) & (x < 4.0), -3.0, 1.0)    # cells 2, 3
    o.set_levelset(0, phi0)
    o.set_levelset(1, phi1)
    assert o.resolve_conflicts() == 2 * 9
    p0, p1 = o.levelset(0), o.levelset(1)
    for i in (3, 4):  # ghosted indices of cells 2 and 3
        assert np.all(p0[..., i] == 2.0) and np.all(p1[..., i] == -2.0), i
    # taken all at once from the old values, cell 3 would stay in both
    assert np.all(p0[..., 5] == -1.0) and np.all(p1[..., 5] == 1.0)


@pytest.mark.parametrize("rnormal", [abi.RNORMAL_LEVEL_SET, abi.RNORMAL_AVERAGE])
def test_normal_between_two_tracked_materials(rnormal):
    n = (8, 8, 8)
    g = abi.grid(n, (0, 0, 0), (1.0, 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(riemann_normal=rnormal)
    ls = [abi.levelset_from_scheme(s, materialid=1), abi.levelset_from_scheme(s, materialid=2)]
    o = ol.Oracle(g, mats, 3, s, levelsets=ls)
    a = np.array(configs._axis(8, 0, 1.0))
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    X, Y, Z = X.transpose(2, 1, 0), Y.transpose(2, 1, 0), Z.transpose(2, 1, 0)
    n1, n2 = np.array([1.0, 0.4, -0.2]), np.array([-0.6, 0.3, 0.9])
    o.set_levelset(0, n1[0] * X + n1[1] * Y + n1[2] * Z - 0.3)
    o.set_levelset(1, n2[0] * X + n2[1] * Y + n2[2] * Z + 0.1)
    u1, u2 = n1 / np.linalg.norm(n1), n2 / np.linalg.norm(n2)
    for d in range(3):
        # face between a cell of material 2 (this cell) and material 1 (minus side)
        got = o.bimaterial_normal(d, 4, 4, 4, 2, 1)
        want = u1 - u2
        want = want / np.linalg.norm(want)
        if want[d] < 0:
            want[d] *= -1.0
        if rnormal == abi.RNORMAL_AVERAGE:
            want[d] += 1.0
            want = want / np.linalg.norm(want)
        assert np.allclose(got, want, rtol=0, atol=1e-14), (d, got, want)


def _ids_case():
    o, x = _line(8, n2=2)
    V = np.zeros(o.shape + (5,))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    ID = np.zeros(o.shape, dtype=np.int32)
    ID[..., 1:4] = 1  # cells 0 .. 2 in subdomain 0 (material 1)
    ID[..., 6:9] = 2  # cells 5 .. 7 in subdomain 1 (material 2)
    o.set_state(V, ID, None)
    return o, x


def test_material_ids_follow_every_level_set():
    o, x = _ids_case()
    old0, old1 = x - 3.0, 5.0 - x  # subdomains: x < 3 (material 1), x > 5 (material 2)
    new0, new1 = x - 2.0, 4.0 - x  # both move left by one cell
    for q, (a, b) in enumerate(((old0, new0), (old1, new1))):
        o.set_levelset(q, a, tmp=True)
        o.set_levelset(q, b)
    o.update_material_ids()
    ID = o.get_state()[1][1, 1, 1:-1]
    # cell 2 left subdomain 0 for none: its untracked neighbour's material
    # (0); cell 4 entered subdomain 1
    assert ID.tolist() == [1, 1, 0, 0, 2, 2, 2, 2]


def test_material_ids_two_subdomains_is_an_error():
    o, x = _ids_case()
    o.set_levelset(0, x - 3.0, tmp=True)
    o.set_levelset(0, x - 4.0)       # cell 3 enters subdomain 0 ...
    o.set_levelset(1, 3.6 - x, tmp=True)
    o.set_levelset(1, 3.0 - x)       # ... and subdomain 1
    with pytest.raises(RuntimeError, match="two material subdomains"):
        o.update_material_ids()


def test_optional_correction_one_owner():
    """Cell 2 lies inside subdomain 0 only, next to both boundaries: with the
    corrections both level sets take +-mean|phi| there (uncounted); cell 3,
    taken after it, then sees the updated cell 2 and is corrected too."""
    o, x = _line(correction_frequency=1)
    phi0, phi1 = x - 2.8, 2.9 - x
    o.set_levelset(0, phi0)
    o.set_levelset(1, phi1)
    assert o.resolve_conflicts() == 0  # the optional branches are not counted
    p0, p1 = o.levelset(0), o.levelset(1)
    m2 = (abs(2.5 - 2.8) + abs(2.9 - 2.5)) / 2
    m3 = (abs(3.5 - 2.8) + abs(2.9 - 3.5)) / 2
    assert np.all(p0[..., 3] == -m2) and np.all(p1[..., 3] == m2)
    assert np.all(p0[..., 4] == m3) and np.all(p1[..., 4] == -m3)
    rest = np.ones(x.shape, bool)
    rest[..., 3:5] = False
    assert np.array_equal(p0[rest], phi0[rest]) and np.array_equal(p1[rest], phi1[rest])
    # without the corrections nothing changes
    o, x = _line(correction_frequency=-1)
    o.set_levelset(0, phi0)
    o.set_levelset(1, phi1)
    assert o.resolve_conflicts() == 0
    assert np.array_equal(o.levelset(0), phi0) and np.array_equal(o.levelset(1), phi1)


def test_optional_correction_on_a_boundary():
    """phi0 = 0 at cell 2, which is outside subdomain 1 next to its boundary:
    level set 0 (the first on the boundary) takes it with -mean|phi|."""
    o, x = _line(correction_frequency=1)
    phi0, phi1 = x - 2.5, 2.9 - x
    o.set_levelset(0, phi0)
    o.set_levelset(1, phi1)
    assert o.resolve_conflicts() == 0
    p0, p1 = o.levelset(0), o.levelset(1)
    m2 = (0.0 + abs(2.9 - 2.5)) / 2
    m3 = (abs(3.5 - 2.5) + abs(2.9 - 3.5)) / 2  # cell 3 then has one owner (level set 1)
    assert np.all(p0[..., 3] == -m2) and np.all(p1[..., 3] == m2)
    assert np.all(p0[..., 4] == m3) and np.all(p1[..., 4] == -m3)


def _cube_case(nls):
    n = 8
    g = abi.grid((n, n, n), (0.0, 0.0, 0.0), (8.0, 8.0, 8.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(mp_ls_correction_frequency=3)
    ls = [abi.levelset_from_scheme(s, materialid=m) for m in (1, 2)[:nls]]
    o = ol.Oracle(g, mats, 3, s, levelsets=ls)
    return o


def test_optional_correction_isolated_background_nodes_in_order():
    """One level set covering everything but three pockets of background
    nodes (phi > 0): a single node, a line of three along x, and a 2x2x2
    block.  Part II gives a background node with at most one background
    neighbour to the level set (phi -> -phi), in (k, j, i) order: the line's
    middle node has two background neighbours at the start, but its first
    node has gone by its turn, so the whole line goes; the block stays."""
    o = _cube_case(1)
    phi = -np.ones(o.shape)
    planted = {(2, 2, 2): 0.3, (4, 5, 5): 0.5, (5, 5, 5): 0.25, (6, 5, 5): 0.75}
    block = [(i, j, k) for i in (5, 6) for j in (1, 2) for k in (1, 2)]
    for (i, j, k), v in planted.items():
        phi[k + 1, j + 1, i + 1] = v
    for (i, j, k) in block:
        phi[k + 1, j + 1, i + 1] = 1.0
    o.set_levelset(0, phi)
    assert o.resolve_conflicts() == 4
    p = o.levelset(0)
    want = phi.copy()
    for (i, j, k), v in planted.items():
        want[k + 1, j + 1, i + 1] = -v
    assert np.array_equal(p, want)


def test_optional_correction_background_goes_to_the_smallest_phi():
    """Two level sets: an isolated background node goes to the level set with
    the smaller phi (the first on ties); the other keeps its value."""
    o = _cube_case(2)
    phi0 = -np.ones(o.shape)
    phi1 = np.ones(o.shape)
    phi0[3, 3, 3] = 0.7   # node (2, 2, 2): outside both, phi1 = 0.4 smaller
    phi1[3, 3, 3] = 0.4
    phi0[6, 6, 6] = 0.5   # node (5, 5, 5): a tie, level set 0 takes it
    phi1[6, 6, 6] = 0.5
    o.set_levelset(0, phi0)
    o.set_levelset(1, phi1)
    # part I: no node lies next to both boundaries (phi1 > 0 everywhere);
    # part II: the two isolated nodes (step 0: the frequency 3 divides it)
    assert o.resolve_conflicts() == 2
    p0, p1 = o.levelset(0), o.levelset(1)
    assert p1[3, 3, 3] == -0.4 and p0[3, 3, 3] == 0.7
    assert p0[6, 6, 6] == -0.5 and p1[6, 6, 6] == 0.5


def _vote_line():
    """8 x 1 x 1 cells; materials 0 and 1 have no level set, 2 is tracked.
    Cells 0, 1 hold material 1, cells 2-4 material 2, cells 5-7 material 0;
    cells 2-4 leave subdomain 2 for none (phi_tmp < 0, phi > 0)."""
    g = abi.grid((8, 1, 1), (0.0, 0.0, 0.0), (8.0, 1.0, 1.0))
    mats = abi.materials(abi.material("SG"), abi.material("SG"), abi.material("SG"))
    s = abi.default_scheme(n_levelsets=1, ls_materialid=2)
    V = np.zeros((3, 3, 10, 5))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    ID = np.zeros((3, 3, 10), dtype=np.int32)
    ID[..., 1:3] = 1
    ID[..., 3:6] = 2
    phi = np.ones((3, 3, 10))
    phi_tmp = np.ones((3, 3, 10))
    phi_tmp[..., 3:6] = -1.0
    return g, mats, s, V, ID, phi, phi_tmp


def test_vote_in_the_reference_order_with_two_untracked_materials():
    """The neighbour vote in (i, j, k) order, a node decided in a sweep seen
    by the later ones: cell 2 takes 1 (cell 1), cell 3 then 1 (cell 2), cell
    4 the smaller of 1 (cell 3) and 0 (cell 5).  Sweeps that see only the
    previous sweep's decisions would leave cell 3 for the second sweep, where
    it takes the smaller of 1 and 0."""
    g, mats, s, V, ID, phi, phi_tmp = _vote_line()
    o = ol.Oracle(g, mats, 3, s)
    o.set_state(V, ID, phi)
    o.set_levelset(0, phi_tmp, tmp=True)
    o.update_material_ids()
    assert o.get_state()[1][1, 1, 1:-1].tolist() == [1, 1, 1, 1, 0, 0, 0, 0]
