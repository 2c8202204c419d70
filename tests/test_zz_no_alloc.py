"""cg_get_diag_static (the diagnostics with the library's static device
scratch, no allocation) against cg_get_diag through the C ABI: bit-identical
on two calls in a row, the second on a smaller solve, in FP64 and mixed
precision."""
import ctypes as C

import pytest

from nextgen_amd.ic import tgv_params
from nextgen_amd.ideal_gas import IdealGasSolver, Mesh, cg_diag, load_library, parse_params

pytestmark = pytest.mark.gpu


def diags(g):
    lib = load_library()
    a, b = cg_diag(), cg_diag()
    assert lib.cg_get_diag(g._h, C.byref(a)) == 0
    assert lib.cg_get_diag_static(g._h, C.byref(b)) == 0
    return a, b


def same(a, b):
    assert (a.step, a.time, a.dt, a.ke, a.enstrophy, a.cfl_max, a.nan_flag) == \
        (b.step, b.time, b.dt, b.ke, b.enstrophy, b.cfl_max, b.nan_flag)
    assert list(a.sums) == list(b.sums)
    assert a.ke > 0


@pytest.mark.parametrize("prec", ["FP64", "MIXED"])
def test_static_diag_equals_diag(prec):
    for n, seed in ((40, 3), (12, 4)):  # 64,000 cells, then 1,728
        m = Mesh.jittered(n, 0.3, seed=seed)
        g = IdealGasSolver(m, parse_params(tgv_params(sgs="VREMAN") + f"PRECISION {prec}\n"))
        g.set_state(*m.ic_tgv(0.1, 1.4))
        g.init_complete()
        print(f"static diag: {prec} ncv {m.ncv}", flush=True)
        same(*diags(g))  # dt is 0 before the first step, so is the CFL
        g.advance_solution(3)
        a, b = diags(g)
        same(a, b)
        assert a.cfl_max > 0.5  # TIMESTEP CFL 0.8
        b2 = cg_diag()
        assert load_library().cg_get_diag_static(g._h, C.byref(b2)) == 0
        same(a, b2)  # the scratch keeps the last call's contents; each call starts afresh
        g.close()
