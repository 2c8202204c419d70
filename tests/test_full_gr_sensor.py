"""USE_FULL_GR: the full Gibbs-remainder sensor (setFgrCvsFull,
IdealGasSolver.cpp:747-838) of the oracle, pinned by thermodynamics rather
than by its own formula. At rest and at uniform pressure the remainder of a
pair reduces, through s = (ln p - gamma ln rho) / (gamma - 1) and h = cp T, to

    gr = gamma/(gamma-1) [ (T1/T0 - T0/T1) / 2 - ln(T1/T0) ]

(the trapezoidal rule against the exact integral of dT/T), so each cell's
vv2 is the largest |gr| over its extended pairs; a uniform state gives 0; the
NO_DISS override of the other sensor does not apply here."""
import numpy as np
import pytest

from nextgen_amd.ideal_gas import Mesh, parse_params
from oracle.oracle import OracleSolver

GAMMA = 1.4


def deck(extra=""):
    return (f"GAMMA {GAMMA}\nP_REF 1.0\nRHO_REF 1.0\nT_REF 1.0\nMU_REF 1e-3\nMU_POWER_LAW 0.75\n"
            f"PR_LAM 0.7\nSGS_MODEL NONE\nTIMESTEP CFL 0.8\nUSE_FULL_GR\n{extra}")


def oracle_at_rest(m, T, extra=""):
    p0 = 1.0
    R = 1.0  # p_ref / (rho_ref T_ref)
    rho = p0 / (R * T)
    u = np.zeros((m.ncv, 3))
    rhoE = np.full(m.ncv, p0 / (GAMMA - 1.0))
    o = OracleSolver(m, parse_params(deck(extra)))
    o.set_state(rho, u, rhoE)
    o.init_complete()
    return o


@pytest.fixture(scope="module")
def mesh():
    return Mesh.jittered(8, 0.3, seed=4)


def test_uniform_state_has_no_remainder(mesh):
    o = oracle_at_rest(mesh, np.ones(mesh.ncv))
    assert np.all(o.get_field("vv2") == 0.0)


def test_isobaric_remainder_matches_thermodynamics(mesh):
    x = mesh.get("x_cv")
    T = 1.0 + 0.2 * np.sin(x[:, 0]) * np.cos(x[:, 1]) + 0.1 * np.sin(2 * x[:, 2])
    o = oracle_at_rest(mesh, T)
    Tc = o.get_field("T")
    lo, hi = mesh.get("lo"), mesh.get("hi")
    r = Tc[hi] / Tc[lo]
    gr = GAMMA / (GAMMA - 1.0) * (0.5 * (r - 1.0 / r) - np.log(r))
    want = np.zeros(mesh.ncv)
    np.maximum.at(want, lo, np.abs(gr))
    np.maximum.at(want, hi, np.abs(gr))
    got = o.get_field("vv2")
    assert np.max(want) > 1e-6
    assert np.allclose(got, want, rtol=1e-8, atol=1e-15), np.max(np.abs(got - want))


def test_no_diss_does_not_override_the_full_sensor(mesh):
    x = mesh.get("x_cv")
    T = 1.0 + 0.3 * np.sin(x[:, 0])
    a = oracle_at_rest(mesh, T)
    b = oracle_at_rest(mesh, T, "NO_DISS\n")
    assert np.array_equal(a.get_field("fgr"), b.get_field("fgr"))
    assert np.all(a.get_field("fgr") > 0.0)
