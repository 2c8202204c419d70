"""SGS_MODEL DSM_LOCAL in the oracle (IdealGasSolverSgs.cpp:84-205, Sgs.hpp:16-244),
pinned by what the model's mathematics fixes rather than by its own formulas:

* the test filter (filterCvR1/R2, StaticSolver.cpp:9849-9975) on a uniform lattice
  has every face weight h^3/6, so it is phi/2 + (1/12) sum of the six neighbours:
  a Fourier mode cos(k x) is multiplied by (5 + cos kh)/6; on any mesh it keeps
  constants;
* L and M are Galilean invariant at uniform density (the filter keeps constants),
  so adding a uniform velocity leaves L:M, M:M and mu_sgs unchanged;
* the time relaxation x <- min(0, x + eps LM)/(1 + eps), y <- (y + eps MM)/(1 + eps)
  with eps = dt (|S| + 1e-15), repeated on a fixed state, follows its closed form
  x_n = LM (1 - r^n), r = 1/(1 + eps), toward the local Germano-Lilly value
  c_delta^2 = -L:M / (2 M:M); mu_sgs = c_delta^2 rho |S|;
* a one-dimensional shear u = (A sin ky, 0, 0) has L:M = 0 (L is diagonal, M
  off-diagonal), hence no eddy viscosity;
* next to a wall the filter adds the cell's own value over the pyramid of each
  boundary face (area x delta / 3): on the lattice channel a unit spike at a wall
  cell filters to (5 x 1/2 + 1) / 6 = 7/12 (1/2 away from the walls) and 1/12 at
  each face neighbour, and the weights still sum to the cell volume.
"""
import numpy as np
import pytest

from nextgen_amd.ic import tgv_params
from nextgen_amd.ideal_gas import Mesh, parse_params
from oracle.oracle import OracleSolver

GAMMA = 1.4


def dsm_params(**kw):
    text = tgv_params(Re=1600.0, Ma=0.1, sgs="DSM_LOCAL")
    for k, v in kw.items():
        text += f"{k} {v}\n"
    return parse_params(text)


def uniform_state(m, u):
    rho = np.ones(m.ncv)
    p = np.full(m.ncv, 1.0 / (GAMMA * 0.1 ** 2))
    rhoE = p / (GAMMA - 1.0) + 0.5 * rho * np.sum(u * u, axis=1)
    return rho, u, rhoE


def smooth_velocity(x, a=1.0):
    return np.stack([a * np.sin(x[:, 1]) * np.cos(2 * x[:, 2]),
                     a * 0.7 * np.sin(x[:, 2] + 0.3) * np.cos(x[:, 0]),
                     a * 0.4 * np.cos(x[:, 0] + 2 * x[:, 1])], axis=1)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_filter_transfer_function_on_the_lattice(k):
    m = Mesh.lattice(10)
    o = OracleSolver(m, dsm_params())
    x = m.get("x_cv")
    h = 2 * np.pi / 10
    for axis in range(3):
        phi = np.cos(k * x[:, axis] + 0.2)
        np.testing.assert_allclose(o.filter_cv(phi), phi * (5 + np.cos(k * h)) / 6, rtol=0, atol=1e-14)
    # three components at once give the same numbers as one at a time
    v = np.stack([np.cos(k * x[:, 0]), np.sin(x[:, 1]), x[:, 2] * 0 + 3.0], axis=1)
    f = o.filter_cv(v)
    for c in range(3):
        np.testing.assert_array_equal(f[:, c], o.filter_cv(np.ascontiguousarray(v[:, c])))


def test_filter_weights_and_constants_on_a_jittered_mesh():
    m = Mesh.jittered(10, 0.3, seed=3)
    o = OracleSolver(m, dsm_params())
    np.testing.assert_allclose(o.filter_cv(np.full(m.ncv, 2.5)), 2.5, rtol=1e-14)
    # the weights sum, per cell, to about the cell volume (face area x generator
    # distance / 6 is the volume of the two pyramids over the face, for a face
    # midway between the generators)
    from oracle.oracle import filter_weights
    w = filter_weights(m)
    cvofa = m.get("cvofa")
    tot = np.zeros(m.ncv)
    np.add.at(tot, cvofa[:, 0], w)
    np.add.at(tot, cvofa[:, 1], w)
    np.testing.assert_allclose(tot, m.get("vol"), rtol=1e-12)
    # the Voronoi radius reaches past the centroid distance of every face
    r = m.get("r_vv")
    x_vv, x_fa = m.get("x_vv"), m.get("x_fa")
    assert np.all(np.linalg.norm(x_fa - x_vv[cvofa[:, 0]], axis=1) < r[cvofa[:, 0]])


def _dsm_after_relax(m, u, dt, n_props=1, **kw):
    o = OracleSolver(m, dsm_params(**kw))
    o.set_state(*uniform_state(m, u))
    o.set_time(0, 0.0, dt)
    o.init_complete()
    for _ in range(n_props - 1):
        o.sgs_props()
    return o


def test_dsm_is_galilean_invariant_at_uniform_density():
    m = Mesh.jittered(10, 0.3, seed=8)
    u = smooth_velocity(m.get("x_cv"))
    a = _dsm_after_relax(m, u, 0.01)
    b = _dsm_after_relax(m, u + np.array([0.3, -0.2, 0.25]), 0.01)
    lm_a, lm_b = a.get_field("lijmij"), b.get_field("lijmij")
    assert np.min(lm_a) < 0.0  # the model is active somewhere
    scale = np.max(np.abs(lm_a))
    assert np.max(np.abs(lm_a - lm_b)) <= 1e-10 * scale
    np.testing.assert_allclose(b.get_field("mijmij"), a.get_field("mijmij"), rtol=1e-9)
    mu_a, mu_b = a.get_field("mu_sgs"), b.get_field("mu_sgs")
    assert np.max(mu_a) > 0.0
    assert np.max(np.abs(mu_a - mu_b)) <= 1e-8 * np.max(mu_a)


def test_relaxation_follows_its_closed_form():
    # the state is held fixed, so L:M, M:M and eps are constants and, from
    # x0 = 0, y0 = 1e-12: x_n = LM (1 - r^n) where LM < 0 (else 0), and
    # y_n = MM' + (y0 - MM') r^n with r = 1/(1 + eps), MM' = 1e-12 + M:M
    m = Mesh.jittered(10, 0.3, seed=9)
    u = smooth_velocity(m.get("x_cv"), a=2.0)
    dt = 0.05
    o = _dsm_after_relax(m, u, dt)
    xs, ys = [o.get_field("lijmij")], [o.get_field("mijmij")]
    for _ in range(5):
        o.sgs_props()
        xs.append(o.get_field("lijmij"))
        ys.append(o.get_field("mijmij"))
    d = o.get_field("dudx")
    S = 0.5 * (d + np.transpose(d, (0, 2, 1)))
    smag = np.sqrt(2.0 * np.sum(S * S, axis=(1, 2)))
    eps = dt * (smag + 1e-15)
    r = 1.0 / (1.0 + eps)
    LM = xs[0] / (1.0 - r)
    MMp = (ys[0] - 1e-12 * r) / (1.0 - r)
    assert np.mean(LM < 0) > 0.2 and np.all(LM <= 0)
    for n in range(1, 6):
        np.testing.assert_allclose(xs[n], LM * (1.0 - r ** (n + 1)), rtol=1e-9, atol=1e-300)
        np.testing.assert_allclose(ys[n], MMp + (1e-12 - MMp) * r ** (n + 1), rtol=1e-9)
    # mu_sgs from the relaxed values: c_delta^2 = -x / (2 (y + 1e-12)), clipped to
    # [0, 0.5 dx^2], times rho |S| (rho = 1)
    x, y = xs[-1], ys[-1]
    cdel2 = np.maximum(-0.5 * x / (y + 1e-12), 0.0)
    dx2 = m.get("vol") ** (2.0 / 3.0)
    cdel2 = np.where(np.sqrt(cdel2 / dx2) > 0.5, 0.5 * dx2, cdel2)
    np.testing.assert_allclose(o.get_field("mu_sgs"), cdel2 * smag, rtol=1e-12, atol=1e-300)
    assert np.max(o.get_field("mu_sgs")) > 0.0


def test_one_dimensional_shear_has_no_eddy_viscosity():
    m = Mesh.lattice(12)
    x = m.get("x_cv")
    u = np.zeros((m.ncv, 3))
    u[:, 0] = 0.5 * np.sin(x[:, 1])
    o = _dsm_after_relax(m, u, 0.02, n_props=5)
    assert np.max(np.abs(o.get_field("lijmij"))) < 1e-14
    assert np.max(o.get_field("mu_sgs")) < 1e-14


def test_filter_boundary_faces_on_the_lattice_channel():
    m = Mesh.channel(10, jitter=0.0)
    text = tgv_params(Re=1600.0, Ma=0.1, sgs="DSM_LOCAL") + "y0 SLIP\ny1 WALL_ADIABATIC\n"
    o = OracleSolver(m, parse_params(text))
    cvobf, zone = m.get("cvobf"), m.get("zone_bf")
    wall = int(cvobf[zone == 0][3])
    cvofa = m.get("cvofa")
    inner = int(cvofa[0, 0]) if cvofa[0, 0] not in cvobf else int(np.setdiff1d(np.arange(m.ncv), cvobf)[0])
    for c, own in ((wall, 7.0 / 12.0), (inner, 0.5)):
        phi = np.zeros(m.ncv)
        phi[c] = 1.0
        f = o.filter_cv(phi)
        assert abs(f[c] - own) < 1e-14, (c, f[c], own)
        nbrs = np.concatenate([cvofa[cvofa[:, 0] == c, 1], cvofa[cvofa[:, 1] == c, 0]])
        np.testing.assert_allclose(f[nbrs], 1.0 / 12.0, rtol=1e-13)
        rest = np.setdiff1d(np.arange(m.ncv), np.concatenate([[c], nbrs]))
        assert np.abs(f[rest]).max() == 0.0
    # constants survive on a jittered channel (weights of faces and boundary faces alike)
    mj = Mesh.channel(10, jitter=0.3, seed=5)
    oj = OracleSolver(mj, parse_params(text))
    np.testing.assert_allclose(oj.filter_cv(np.full(mj.ncv, 2.5)), 2.5, rtol=1e-14)
