"""Device layout orders: the lane order of the SELL regions (CG_LANE_ORDER)
and the device cell order (CG_CELL_ORDER) change where data sits and which
thread owns which cell, never a result. A solve must give bit-identical
state, fields and stage RHS in every order, in FP64 and in mixed precision
(the orders are read when a layout is built)."""
import os

import numpy as np
import pytest

from nextgen_amd.ic import tgv_params
from nextgen_amd.ideal_gas import IdealGasSolver, Mesh, parse_params

pytestmark = pytest.mark.gpu

ORDERS = [
    {},                                                      # defaults
    {"CG_LANE_ORDER": "plain"},                              # lane = cell in every slice
    {"CG_CELL_ORDER": "sorted64"},
    {"CG_CELL_ORDER": "sorted256", "CG_LANE_ORDER": "plain"},
]


def run(order, prec):
    saved = {k: os.environ.get(k) for k in ("CG_LANE_ORDER", "CG_CELL_ORDER")}
    try:
        for k in saved:
            os.environ.pop(k, None)
        os.environ.update(order)
        # 12^3 jittered cells: several slices and workgroups, ragged degrees
        m = Mesh.jittered(12, 0.3, seed=11)
        p = parse_params(tgv_params(Re=800.0, Ma=0.2, sgs="VREMAN") + f"PRECISION {prec}\n")
        g = IdealGasSolver(m, p)
    finally:
        for k, v in saved.items():
            os.environ.pop(k, None)
            if v is not None:
                os.environ[k] = v
    g.set_state(*m.ic_tgv(0.2, 1.4))
    g.init_complete()
    rhs = g.calc_rhs(1)
    g.advance_solution(6)
    out = {"rhs": rhs, "state": g.get_state(), "p": g.get_field("p"), "mu_total": g.get_field("mu_total"),
           "fgr": g.get_field("fgr"), "dudx": g.get_field("dudx"), "step": g.report().step}
    g.close()
    return out


@pytest.mark.parametrize("prec", ["FP64", "MIXED"])
def test_results_identical_in_every_layout_order(prec):
    ref = run(ORDERS[0], prec)
    assert ref["step"] == 6
    for order in ORDERS[1:]:
        got = run(order, prec)
        assert got["step"] == ref["step"], order
        np.testing.assert_array_equal(got["rhs"], ref["rhs"], err_msg=str(order))
        for a, b in zip(got["state"], ref["state"]):
            np.testing.assert_array_equal(a, b, err_msg=str(order))
        for name in ("p", "mu_total", "fgr", "dudx"):
            np.testing.assert_array_equal(got[name], ref[name], err_msg=f"{order} {name}")
