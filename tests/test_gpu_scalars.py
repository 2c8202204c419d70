"""Passive scalars on the device: against the oracle (FP64 to 1e-12 of the field's
magnitude, mixed precision to 1e-5), without effect on the flow, the same in a
batch as alone, and carried by data files; on channel meshes through CBC inflows
(the zone's value coming in), CBC, NSCBC and sponge outflows (the cell's value
going out, 0 on backflow) and walls (no scalar flux)."""
import numpy as np
import pytest

from nextgen_amd.batch import Batch
from nextgen_amd.ic import tgv_params, tgv_state
from nextgen_amd.ideal_gas import CG_PREC_FP64, CG_PREC_MIXED, IdealGasSolver, Mesh, parse_params
from oracle.oracle import OracleSolver

pytestmark = pytest.mark.gpu

SC = "REGISTER_SCALAR Zmix C2\nRES_TIME\n"


def text(prec="FP64", sc=SC):
    return tgv_params(Re=400.0, Ma=0.2, sgs="VREMAN") + sc + f"PRECISION {prec}\n"


def fields(x):
    z = 0.5 + 0.5 * np.sin(x[:, 0] + 2 * x[:, 2]) * np.cos(x[:, 1])
    c2 = np.exp(-((x[:, 0] - np.pi) ** 2 + (x[:, 1] - np.pi) ** 2) / 2.0)
    return {"Zmix": z, "C2": c2}


def gpu_run(m, prec, nsteps, sc=SC):
    g = IdealGasSolver(m, parse_params(text(prec, sc)))
    g.set_state(*tgv_state(m.get("x_cv"), Ma=0.2))
    if sc:
        for name, f in fields(m.get("x_cv")).items():
            g.set_scalar(name, f)
    g.init_complete()
    g.advance_solution(nsteps)
    return g


@pytest.mark.parametrize("prec,tol", [("FP64", 1e-12), ("MIXED", 1e-5)])
def test_scalars_match_oracle(prec, tol):
    m = Mesh.jittered(12, 0.3, seed=11)
    g = gpu_run(m, prec, 10)
    assert g.scalar_names() == ["C2", "RES_TIME", "Zmix"]
    o = OracleSolver(m, parse_params(text()))
    o.set_state(*tgv_state(m.get("x_cv"), Ma=0.2))
    f = fields(m.get("x_cv"))
    o.set_scalar(0, f["C2"])
    o.set_scalar(2, f["Zmix"])
    o.init_complete()
    o.step(10)
    for i, name in enumerate(["C2", "RES_TIME", "Zmix"]):
        a, b = g.get_scalar(name), o.get_scalar(i)
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"{prec} {name}: {err:.3e}")
        assert err <= tol, name
    assert np.max(np.abs(g.get_scalar("Zmix") - f["Zmix"])) > 1e-3  # transported
    g.close()


def test_scalars_leave_the_flow_unchanged():
    m = Mesh.jittered(12, 0.3, seed=11)
    a = gpu_run(m, "FP64", 6)
    b = gpu_run(m, "FP64", 6, sc="")
    for x, y in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [CG_PREC_FP64, CG_PREC_MIXED])
def test_batch_with_and_without_scalars_matches_solo_runs(precision):
    prec = "MIXED" if precision == CG_PREC_MIXED else "FP64"
    specs = [(10, 3, SC), (12, 5, ""), (9, 7, "REGISTER_SCALAR Q\n")]
    b = Batch(precision)
    ids = []
    for n, seed, sc in specs:
        m = Mesh.jittered(n, 0.3, seed=seed)
        ids.append(b.add(m, parse_params(text(prec, sc)), *tgv_state(m.get("x_cv"), Ma=0.2)))
    b.step(5)
    for (n, seed, sc), iid in zip(specs, ids):
        m = Mesh.jittered(n, 0.3, seed=seed)
        g = IdealGasSolver(m, parse_params(text(prec, sc)))
        g.set_state(*tgv_state(m.get("x_cv"), Ma=0.2))
        g.init_complete()
        g.advance_solution(5)
        for x, y in zip(b.state(iid), g.get_state()):
            np.testing.assert_array_equal(x, y)
        g.close()
    b.close()


def test_scalars_travel_in_data_files(tmp_path):
    m = Mesh.jittered(12, 0.3, seed=6)
    path = str(tmp_path / "sc.sles")
    a = gpu_run(m, "FP64", 4)
    a.write_sles(path)
    a.advance_solution(5)
    ref = {n: a.get_scalar(n) for n in a.scalar_names()}
    a.close()
    b = IdealGasSolver(m, parse_params(text()))
    b.read_sles(path)
    b.init_complete()
    b.advance_solution(5)
    for n, v in ref.items():
        np.testing.assert_array_equal(b.get_scalar(n), v)
    b.close()


P0 = 1.0 / (1.4 * 0.1 ** 2)
CHANNEL_ZONES = {
    "cbc": {"y0": f"CBC_UPT 0.1 0.4 0 {P0!r} 1.0 Zmix 1.0 C2 0.25 RES_TIME 0.0",
            "y1": f"CBC_RUNP 1.0 -0.2 {P0 * 0.999!r} Zmix 0.0 C2 0.0 RES_TIME 0.0"},
    "nscbc": {"y0": f"CBC_UPT 0.1 0.4 0 {P0!r} 1.0 Zmix 1.0 C2 0.25 RES_TIME 0.0",
              "y1": f"NSCBC_OUTLET_P {P0 * 0.995!r} 0.3 6.28"},
    "sponge+wall": {"y0": f"SPONGE TYPE L_Y LENGTH 2.0 STRENGTH 5.0 RELAX_TIME 0.05 SPONGE_PRESSURE {P0 * 1.002!r}",
                    "y1": "WALL_ADIABATIC"},
}


def channel_state(m, Ma=0.1, gamma=1.4):
    """A shear layer between the channel's walls with cellular perturbations."""
    x = m.get("x_cv")
    p0 = 1.0 / (gamma * Ma * Ma)
    y = (x[:, 1] - np.pi) / np.pi
    u = np.empty((m.ncv, 3))
    u[:, 0] = 1.0 - y * y + 0.2 * np.sin(x[:, 0]) * np.cos(x[:, 2])
    u[:, 1] = 0.1 * np.sin(x[:, 2]) * np.sin(x[:, 1])
    u[:, 2] = -0.2 * np.cos(x[:, 0]) * np.sin(x[:, 2])
    p = p0 + 0.05 * np.cos(x[:, 0]) * np.cos(2 * x[:, 1])
    rho = p / p0
    return rho, u, p / (gamma - 1.0) + 0.5 * rho * (u * u).sum(1)


def channel_scalar_run(zones, prec, nsteps):
    m = Mesh.channel(10, jitter=0.3, seed=21)
    deck = (tgv_params(sgs="VREMAN") + "".join(f"{z} {w}\n" for z, w in zones.items()) + SC)
    st = channel_state(m)
    f = fields(m.get("x_cv"))
    g = IdealGasSolver(m, parse_params(deck + f"PRECISION {prec}\n"))
    g.set_state(*st)
    for name, v in f.items():
        g.set_scalar(name, v)
    g.init_complete()
    o = OracleSolver(m, parse_params(deck))
    o.set_state(*st)
    o.set_scalar(0, f["C2"])
    o.set_scalar(2, f["Zmix"])
    o.init_complete()
    g.advance_solution(nsteps)
    o.step(nsteps)
    return m, g, o


@pytest.mark.parametrize("zones", sorted(CHANNEL_ZONES))
@pytest.mark.parametrize("prec,tol", [("FP64", 1e-12), ("MIXED", 1e-5)])
def test_channel_scalars_through_boundaries_match_oracle(zones, prec, tol):
    m, g, o = channel_scalar_run(CHANNEL_ZONES[zones], prec, 12)
    for i, name in enumerate(["C2", "RES_TIME", "Zmix"]):
        a, b = g.get_scalar(name), o.get_scalar(i)
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"{zones} {prec} {name}: {err:.3e}", flush=True)
        assert err <= tol, name
    g.close()
