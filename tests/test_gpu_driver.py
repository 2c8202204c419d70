"""The C++ host side (charles_ig_gpu over include/charles_gpu.h) on a small deck,
checked against the oracle's history of the same run."""
import os
import re
import subprocess

import numpy as np
import pytest

from nextgen_amd.ic import tgv_params, tgv_state
from nextgen_amd.ideal_gas import Mesh, parse_params
from oracle.oracle import OracleSolver

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nextgen_amd", "lib", "charles_ig_gpu")


def test_driver_runs_deck_and_matches_oracle(tmp_path):
    deck = tmp_path / "charles.in"
    deck.write_text(tgv_params(sgs="VREMAN") + "MESH JITTERED 12 0.3 11\nINIT TGV 0.1\nNSTEPS 10\nCHECK_INTERVAL 5\n")
    out = subprocess.run([EXE, "-i", str(deck)], capture_output=True, text=True, timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    steps = [int(s) for s in re.findall(r"> step: (\d+)", out.stdout)]
    kes = [float(s) for s in re.findall(r"ke: ([0-9.eE+-]+)", out.stdout)]
    assert steps[-1] == 10
    m = Mesh.jittered(12, 0.3, seed=11)
    p = parse_params(tgv_params(sgs="VREMAN"))
    o = OracleSolver(m, p)
    o.set_state(*tgv_state(m.get("x_cv")))
    o.init_complete()
    o.step(10)
    assert abs(kes[-1] - o.diag()["ke"]) <= 1e-10 * o.diag()["ke"]


def test_driver_reports_bad_deck(tmp_path):
    deck = tmp_path / "bad.in"
    deck.write_text("SGS_MODEL DSM_GLOBAL\nMESH LATTICE 8\n")
    out = subprocess.run([EXE, "-i", str(deck)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1
    assert "SGS_MODEL" in out.stderr


def test_driver_k4_deck_matches_oracle(tmp_path):
    k4 = ("GAMMA 1.4\nP_REF 0.7142857142857143\nRHO_REF 1.0\nT_REF 1.0\nMU_REF 0.0003\n"
          "MU_POWER_LAW 0.75\nPR_LAM 0.7\nSGS_MODEL VREMAN\nTIMESTEP CFL 0.8\n")
    deck = tmp_path / "k4.in"
    deck.write_text(k4 + "MESH JITTERED 10 0.3 4\nINIT K4 2.0 0.2 9\nNSTEPS 6\nCHECK_INTERVAL 3\n")
    out = subprocess.run([EXE, "-i", str(deck)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    kes = [float(s) for s in re.findall(r"ke: ([0-9.eE+-]+)", out.stdout)]
    m = Mesh.jittered(10, 0.3, seed=4)
    o = OracleSolver(m, parse_params(k4))
    o.set_state(*m.ic_k4(2.0, 0.2, 1.4, seed=9))
    o.init_complete()
    o.step(6)
    assert abs(kes[-1] - o.diag()["ke"]) <= 1e-10 * o.diag()["ke"]


def test_driver_sigma_and_full_sensor_deck_matches_oracle(tmp_path):
    # SGS_MODEL SIGMA and USE_FULL_GR through the C++ host side
    text = tgv_params(sgs="SIGMA") + "SIGMA_COEFF 1.4\nUSE_FULL_GR\n"
    deck = tmp_path / "sigma.in"
    deck.write_text(text + "MESH JITTERED 12 0.3 11\nINIT TGV 0.1\nNSTEPS 8\nCHECK_INTERVAL 4\n")
    out = subprocess.run([EXE, "-i", str(deck)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    kes = [float(s) for s in re.findall(r"ke: ([0-9.eE+-]+)", out.stdout)]
    m = Mesh.jittered(12, 0.3, seed=11)
    o = OracleSolver(m, parse_params(text))
    o.set_state(*tgv_state(m.get("x_cv")))
    o.init_complete()
    o.step(8)
    assert abs(kes[-1] - o.diag()["ke"]) <= 1e-10 * o.diag()["ke"]


def test_driver_restart_from_mesh_and_data_files(tmp_path):
    # RESTART <mesh.mles> and WRITE_SLES, then RESTART <mesh.mles> <data.sles>:
    # the restarted run ends where the uninterrupted run ends
    m = Mesh.jittered(12, 0.3, seed=11)
    mles = str(tmp_path / "box.mles")
    m.write_mles(mles)
    prefix = str(tmp_path / "res")
    full = tmp_path / "full.in"
    full.write_text(tgv_params(sgs="VREMAN") + f"RESTART {mles}\nINIT TGV 0.1\nNSTEPS 10\nCHECK_INTERVAL 5\n"
                    f"WRITE_SLES INTERVAL 4 NAME {prefix}\n")
    a = subprocess.run([EXE, "-i", str(full)], capture_output=True, text=True, timeout=120)
    assert a.returncode == 0, a.stderr
    assert os.path.exists(prefix + ".00000004.sles") and os.path.exists(prefix + ".00000008.sles")
    rest = tmp_path / "rest.in"
    rest.write_text(tgv_params(sgs="VREMAN") + f"RESTART {mles} {prefix}.00000004.sles\nNSTEPS 10\nCHECK_INTERVAL 5\n")
    b = subprocess.run([EXE, "-i", str(rest)], capture_output=True, text=True, timeout=120)
    assert b.returncode == 0, b.stderr
    last = [re.findall(r"> step: 10 .*", o.stdout) for o in (a, b)]
    assert last[0] and last[0] == last[1]
    # the uninterrupted run also matches the oracle on the same mesh
    kes = [float(s) for s in re.findall(r"ke: ([0-9.eE+-]+)", a.stdout)]
    o = OracleSolver(m, parse_params(tgv_params(sgs="VREMAN")))
    o.set_state(*tgv_state(m.get("x_cv")))
    o.init_complete()
    o.step(10)
    assert abs(kes[-1] - o.diag()["ke"]) <= 1e-10 * o.diag()["ke"]
