"""Pins of the oracle's Godunov flux (Flux = Godunov, FluxFcnGodunov.h:37-60:
the physical flux of the exact Riemann solution at x/t = 0).  No reference
test deck runs it; the exact solver it samples is pinned by the reference's
Riemann-problem fixtures, and the sampling by what the mathematics fixes:

* equal states: the flux of a uniform state is its physical flux, which the
  LLF flux of (V, V) returns exactly;
* the Sod problem (ideal gas 1.4, (1, 0, 1) | (0.125, 0, 0.1)): x/t = 0 lies
  in the star region left of the contact, so the flux is that of the left
  star state -- closed form of the two-shock/rarefaction solution solved
  here to machine precision by bisection on the ideal-gas wave curves;
* a transonic rarefaction ((1, 0.75, 1) | (0.125, 0, 0.1)): x/t = 0 lies in
  the left fan, whose sonic state u = c = 2/(g+1) (c_L + (g-1)/2 u_L) has a
  closed form.  The solver samples the fan by linear interpolation between
  the integration points on either side of x/t = 0
  (ExactRiemannSolverBase.cpp:1351-1364), points spaced by the adaptive
  steps, so the sampled flux meets the closed form only to a few percent; the
  test checks that the transonic branch is taken (the flux is not the star
  state's) and lands within 3% of the sonic flux (loose: parity unpinned
  beyond that);
* tangential velocity is carried from the upwind side and the flux along
  y and z is the x flux with the components permuted.

The two closed-form cases run the rarefaction integrator with the published
Cash-Karp abscissae (the oracle's test switch): with the reference's own
third-stage abscissa (ExactRiemannSolverBase.cpp:1814) the Sod star state is
off by ~6e-4, in the reference as here, which is not what is pinned here."""
import numpy as np
import pytest

import oracle_lib as ol
from m2c_amd import abi

G = 1.4
AIR = abi.materials(abi.material("SG", gamma=G))


@pytest.fixture
def textbook():
    ol.set_ck_textbook(True)
    yield
    ol.set_ck_textbook(False)


def _godunov(d, Vm, Vp, tol=1e-12, tol_rare=1e-11):
    s = abi.default_scheme(flux=abi.FLUX_GODUNOV, riem_tol_main=tol, riem_tol_rarefaction=tol_rare)
    Vm, Vp = np.atleast_2d(Vm), np.atleast_2d(Vp)
    return ol.flux_batch(AIR, 1, s, d, Vm, Vp, np.zeros(len(Vm)))


def _phys(V):
    rho, u, v, w, p = V
    E = p / (G - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, u * (E + p)])


def _star(L, R):
    """p*, u* of the ideal-gas Riemann problem (Toro's pressure function,
    bisection to machine precision)."""
    def f(p, rho, pk):
        if p > pk:
            A, B = 2.0 / ((G + 1) * rho), (G - 1) / (G + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        c = np.sqrt(G * pk / rho)
        return 2 * c / (G - 1) * ((p / pk) ** ((G - 1) / (2 * G)) - 1)
    lo, hi = 1e-8, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid, L[0], L[2]) + f(mid, R[0], R[2]) + (R[1] - L[1]) > 0:
            hi = mid
        else:
            lo = mid
    p = 0.5 * (lo + hi)
    u = 0.5 * (L[1] + R[1]) + 0.5 * (f(p, R[0], R[2]) - f(p, L[0], L[2]))
    return p, u


def test_uniform_state_gives_the_physical_flux():
    rng = np.random.default_rng(3)
    V = np.column_stack([rng.uniform(0.2, 2, 50), rng.uniform(-1, 1, (50, 3)), rng.uniform(0.1, 3, 50)])
    s = abi.default_scheme(flux=abi.FLUX_LLF)
    for d in range(3):
        F = _godunov(d, V, V)
        Fl = ol.flux_batch(AIR, 1, s, d, V, V, np.zeros(len(V)))
        assert np.array_equal(F, Fl)


def test_sod_left_star_state(textbook):
    L, R = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    p, u = _star(L, R)
    rho = L[0] * (p / L[2]) ** (1.0 / G)  # left rarefaction: isentrope
    c_tail = np.sqrt(G * p / rho)
    assert u - c_tail < 0 < u  # x/t = 0 between the fan's tail and the contact
    F = _godunov(0, [L[0], L[1], 0, 0, L[2]], [R[0], R[1], 0, 0, R[2]])[0]
    want = _phys((rho, u, 0.0, 0.0, p))
    assert np.allclose(F, want, rtol=1e-7, atol=0)


def test_transonic_rarefaction_sonic_state(textbook):
    L, R = (1.0, 0.75, 1.0), (0.125, 0.0, 0.1)
    cL = np.sqrt(G * L[2] / L[0])
    assert L[1] - cL < 0  # head of the left fan moves left
    p_star, u_star = _star(L, R)
    c_tail = cL * (p_star / L[2]) ** ((G - 1) / (2 * G))
    assert u_star - c_tail > 0  # tail moves right: x/t = 0 inside the fan
    c = 2.0 / (G + 1) * (cL + 0.5 * (G - 1) * L[1])
    rho = L[0] * (c / cL) ** (2.0 / (G - 1))
    p = L[2] * (c / cL) ** (2.0 * G / (G - 1))
    F = _godunov(0, [L[0], L[1], 0, 0, L[2]], [R[0], R[1], 0, 0, R[2]])[0]
    want = _phys((rho, c, 0.0, 0.0, p))
    rho_s = L[0] * (p_star / L[2]) ** (1.0 / G)
    star = _phys((rho_s, u_star, 0.0, 0.0, p_star))
    assert np.all(np.abs(F - want) <= 0.03 * np.abs(want))
    assert np.all(np.abs(F - star)[[0, 1, 4]] > 1e-4 * np.abs(star)[[0, 1, 4]])


@pytest.mark.parametrize("upwind", ["left", "right"])
def test_tangential_velocity_and_axis_permutation(upwind):
    # flow to the right (left state upwind) or to the left
    s = 1.0 if upwind == "left" else -1.0
    L = np.array([1.0, 0.3 * s, 0.2, -0.4, 1.0])
    R = np.array([0.5, 0.3 * s, -0.7, 0.1, 0.6])
    Fx = _godunov(0, L, R)[0]
    up = L if upwind == "left" else R
    # tangential momentum flux = mass flux x upwind tangential velocity
    assert np.isclose(Fx[2], Fx[0] * up[2], rtol=1e-12)
    assert np.isclose(Fx[3], Fx[0] * up[3], rtol=1e-12)
    for d in (1, 2):
        perm = [0, 1, 2, 3, 4]
        perm[1], perm[1 + d] = perm[1 + d], perm[1]
        Fd = _godunov(d, L[perm], R[perm])[0]
        assert np.array_equal(Fd[perm], Fx)
