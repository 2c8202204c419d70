"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Only tests, the smoke check and the CPU-baseline leg of the benchmark import
this module, and only as the checker: it loads ``oracle/liboracle.so`` (the C
restatement in ig_oracle.c) and, when present, ``oracle/_ref/libref_pin.so``
(the reference's own header code, used to pin the restatement).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_pin.so")


class og_mesh(C.Structure):
    _fields_ = [("ncv", C.c_int), ("nfa", C.c_int), ("nef", C.c_int)] + [
        (n, C.c_void_p) for n in ("vol", "inv_vol", "sa_to_vol", "dcmag_hat", "lo", "hi", "n_use", "c_use",
                                  "unit_n_use", "area_use", "aod_use", "n_fa", "cvocv_i", "cvocv_v", "grad_coeff",
                                  "filt_w", "r_vv")
    ] + [("nbf", C.c_int)] + [(n, C.c_void_p) for n in ("cvobf", "zone_bf", "n_bf", "aod_bf", "x_cv")] + [
        ("x_max", C.c_double * 5)]


class og_params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("p_ref", C.c_double), ("rho_ref", C.c_double), ("T_ref", C.c_double),
        ("mu_ref", C.c_double), ("mu_power_law", C.c_double), ("Pr_lam", C.c_double), ("sgs_model", C.c_int),
        ("vreman_coeff", C.c_double), ("a_sgs_coeff", C.c_double), ("a_sgs_coeff2", C.c_double),
        ("lambda_dil_coeff", C.c_double), ("Pr_t", C.c_double), ("limit_asgs", C.c_int), ("coeff_p", C.c_double),
        ("coeff_h", C.c_double), ("gr_nrm", C.c_double), ("diss_mode", C.c_int), ("dt_mode", C.c_int),
        ("dt_data", C.c_double * 2), ("check_interval", C.c_int), ("sigma_coeff", C.c_double),
        ("full_gr", C.c_int), ("delta_r_dsm", C.c_double), ("nsc", C.c_int), ("res_time", C.c_int),
        ("bc_type", C.c_int * 16), ("bc_T", C.c_double * 16), ("bc_u", (C.c_double * 3) * 16),
        ("bc_p", C.c_double * 16), ("bc_rho", C.c_double * 16), ("bc_q", C.c_double * 16),
        ("bc_area", C.c_double * 16), ("bc_sigma", C.c_double * 16), ("bc_L", C.c_double * 16),
        ("sponge_zone", C.c_int), ("sponge_type", C.c_int), ("sponge_strength", C.c_double),
        ("sponge_length", C.c_double), ("sponge_p", C.c_double), ("sponge_trelax", C.c_double),
        ("bc_sc", (C.c_double * 8) * 16), ("bc_mode", C.c_int * 16),
    ]


class og_diag(C.Structure):
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("dt", C.c_double), ("ke", C.c_double),
                ("enstrophy", C.c_double), ("sums", C.c_double * 5), ("cfl_max", C.c_double),
                ("nan_flag", C.c_int)]


# addSourceHook(IdealGasRhs* rhs, time, rk_stage): (user, rhs [ncv][5], ncv, time, stage)
OG_SOURCE_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_double, C.c_int)

_libs = {}
_ref = None


def lib(path=None):
    """The oracle library (default oracle/liboracle.so; another build of
    ig_oracle.c by path, e.g. one compiled for the host CPU)."""
    path = path or ORACLE_SO
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build()")
        L = C.CDLL(path)
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.og_create.restype = vp
        L.og_create.argtypes = [C.POINTER(og_mesh), C.POINTER(og_params)]
        L.og_destroy.argtypes = [vp]
        L.og_set_threads.argtypes = [vp, C.c_int]
        L.og_set_threads.restype = None
        L.og_set_state.argtypes = [vp, dp, dp, dp]
        L.og_get_state.argtypes = [vp, dp, dp, dp]
        L.og_get_field.argtypes = [vp, C.c_char_p, dp]
        L.og_get_field.restype = C.c_int
        for f in ("og_update_state", "og_sgs_props", "og_init_complete", "og_store", "og_restore"):
            getattr(L, f).argtypes = [vp]
        L.og_calc_rhs.argtypes = [vp, dp]
        L.og_calc_cfl.argtypes = [vp, C.c_double]
        L.og_calc_cfl.restype = C.c_double
        L.og_step.argtypes = [vp]
        L.og_step.restype = C.c_int
        L.og_diag_get.argtypes = [vp, C.POINTER(og_diag)]
        for f in ("og_flux_skew", "og_flux_symmetric", "og_flux_compact"):
            getattr(L, f).restype = None
        L.og_flux_skew.argtypes = [dp, dp, dp, dp, dp, dp, dp, dp]
        L.og_flux_symmetric.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.og_flux_compact.argtypes = [dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, dp, dp]
        L.og_fast_pow_posx.restype = C.c_double
        L.og_svd33.restype = C.c_int
        L.og_svd33.argtypes = [dp, dp, dp, dp]
        L.og_sigma_dsigma.restype = C.c_double
        L.og_sigma_dsigma.argtypes = [dp]
        L.og_fast_pow_posx.argtypes = [C.c_double, C.c_double]
        L.og_filter_cv.argtypes = [C.POINTER(og_mesh), dp, dp, C.c_int]
        L.og_filter_cv.restype = None
        L.og_set_time.argtypes = [vp, C.c_int64, C.c_double, C.c_double]
        L.og_set_time.restype = None
        L.og_set_dsm_state.argtypes = [vp, dp, dp]
        L.og_set_dsm_state.restype = None
        for f in ("og_set_scalar", "og_get_scalar"):
            getattr(L, f).argtypes = [vp, C.c_int, dp]
            getattr(L, f).restype = None
        L.og_calc_rhs_scalars.argtypes = [vp, dp]
        L.og_calc_rhs_scalars.restype = None
        L.og_set_source_hook.argtypes = [vp, OG_SOURCE_FN, vp]
        L.og_set_source_hook.restype = None
        L.og_gr_func.restype = C.c_double
        L.og_gr_func.argtypes = [C.c_double, C.c_double]
        L.og_flux_riemann.restype = None
        L.og_flux_riemann.argtypes = [dp, dp, dp, dp, C.c_double, C.c_double]
        L.og_nscbc_outlet_rhs.restype = C.c_double
        L.og_nscbc_outlet_rhs.argtypes = [dp, dp, C.c_double, dp, dp, C.c_double, C.c_double, C.c_double,
                                          C.c_double]
        L.og_nscbc_mdot_rhs.restype = C.c_double
        L.og_nscbc_mdot_rhs.argtypes = [dp, dp, C.c_double, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.og_nscbc_inlet_rhs.restype = C.c_int
        L.og_nscbc_inlet_rhs.argtypes = [dp, dp, C.c_double, dp, dp, C.c_double]
        L.og_wm_solve_tau.restype = C.c_double
        L.og_wm_solve_tau.argtypes = [C.c_double] * 5
        L.og_wm_q_wall.restype = C.c_double
        L.og_wm_eq_solve_qt.restype = C.c_int
        L.og_wm_eq_solve_qt.argtypes = [dp, dp] + [C.c_double] * 7 + [dp] + [C.c_double] * 3 + [C.c_int]
        L.og_wm_q_wall.argtypes = [C.c_double] * 8
        _libs[path] = L
    return _libs[path]


def ref():
    """The reference's own header code (None when it was not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        R = C.CDLL(REF_SO)
        dp = C.POINTER(C.c_double)
        R.ref_flux_skew.argtypes = [dp, dp, dp, dp, dp]
        R.ref_flux_symmetric.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        R.ref_flux_compact.argtypes = [dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, dp, dp]
        R.ref_fast_pow_posx.restype = C.c_double
        R.ref_fast_pow_posx.argtypes = [C.c_double, C.c_double]
        R.ref_svd33.restype = C.c_int
        R.ref_svd33.argtypes = [dp, dp, dp, dp, dp]
        R.ref_erk_wgt3.argtypes = [dp]
        R.ref_flux_riemann.argtypes = [dp, dp, dp, dp, C.c_double, C.c_double]
        R.ref_nscbc_outlet_rhs.restype = C.c_double
        R.ref_nscbc_outlet_rhs.argtypes = [dp, dp, C.c_double, dp, dp, C.c_double, C.c_double, C.c_double,
                                           C.c_double]
        R.ref_nscbc_mdot_rhs.restype = C.c_double
        R.ref_nscbc_mdot_rhs.argtypes = [dp, dp, C.c_double, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double]
        R.ref_nscbc_inlet_rhs.restype = C.c_int
        R.ref_nscbc_inlet_rhs.argtypes = [dp, dp, C.c_double, dp, dp, C.c_double]
        R.ref_wm_solve_tau.restype = C.c_double
        R.ref_wm_solve_tau.argtypes = [C.c_double] * 5
        R.ref_wm_q_wall.restype = C.c_double
        R.ref_wm_eq_solve_qt.restype = C.c_int
        R.ref_wm_eq_solve_qt.argtypes = [dp, dp] + [C.c_double] * 7 + [dp] + [C.c_double] * 3 + [C.c_int]
        R.ref_wm_q_wall.argtypes = [C.c_double] * 8
        _ref = R
    return _ref


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def filter_weights(mesh):
    """Test-filter weight of each compact face, filterCvR1/R2
    (StaticSolver.cpp:9849-9975): |n_fa| |x_vv1 - x_vv0| / 6 with x_vv1 the
    periodic image of the face's second cell (its translation, x + s L, rounded
    once as PeriodicData::periodicTranslate does); MAG and DIST summed left to
    right (Macros.hpp:139,253; pow(x, 2) is taken as x * x)."""
    x_vv = mesh.get("x_vv")
    cvofa = mesh.get("cvofa")
    shift = mesh.get("fa_shift").astype(np.float64)
    L = mesh.get("L")
    n = mesh.get("n_fa")
    x1 = x_vv[cvofa[:, 1]] + shift * L
    d = x1 - x_vv[cvofa[:, 0]]
    dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    area = np.sqrt((n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1]) + n[:, 2] * n[:, 2])
    return np.ascontiguousarray(area * dist / 6.0)


def params_from(cgp, zones=(), zone_area=()) -> og_params:
    """The oracle's parameters; `zones` are the mesh's boundary zone names, each
    given the condition cgp holds under that name; `zone_area` each zone's area
    (its faces' |n_bf| summed in face order)."""
    p = og_params()
    p.sponge_zone = -1
    bcs = {cgp.bc[i].zone.decode(): cgp.bc[i] for i in range(cgp.nbc)}
    for z, name in enumerate(zones):
        b = bcs[name]
        if b.type == 9:
            # SPONGE: RELAX_TIME in T, STRENGTH in q, LENGTH in L, TYPE in mode,
            # SPONGE_PRESSURE in p (negative: P_REF)
            p.sponge_zone, p.sponge_type = z, b.mode
            p.sponge_strength, p.sponge_length, p.sponge_trelax = b.q, b.L, b.T
            p.sponge_p = b.p if b.p >= 0.0 else cgp.p_ref
        p.bc_type[z] = b.type
        p.bc_T[z] = b.T
        for d in range(3):
            p.bc_u[z][d] = b.u[d]
        p.bc_p[z], p.bc_rho[z], p.bc_q[z] = b.p, b.rho, b.q
        p.bc_area[z] = zone_area[z] if len(zone_area) else 0.0
        p.bc_sigma[z], p.bc_L[z] = b.sigma, b.L
        p.bc_mode[z] = b.mode if b.type in (11, 12) else 0
        if 4 <= b.type <= 7 or b.type in (11, 12):  # CBC zones and NSCBC inlets bring scalars in; others 0
            for i in range(min(cgp.nsc, 8)):
                p.bc_sc[z][i] = b.sc[i]
    for name, _ in og_params._fields_:
        if name.startswith("bc_") or name.startswith("sponge_"):
            continue
        if name == "full_gr":
            p.full_gr = cgp.use_full_gr
        elif name == "dt_data":
            p.dt_data[0] = cgp.dt_data[0]
            p.dt_data[1] = cgp.dt_data[1]
        else:
            setattr(p, name, getattr(cgp, name))
    return p


class OracleSolver:
    """The oracle step on the operator arrays of a product mesh. `threads` > 1
    runs its loops on that many threads, bit-identical to one thread."""

    def __init__(self, mesh, cgparams, threads=1, libpath=None):
        self._L = lib(libpath)
        self._keep = {}
        m = og_mesh()
        m.ncv, m.nfa, m.nef = mesh.ncv, mesh.nfa, mesh.nef
        names = {"vol": "vol", "inv_vol": "inv_vol", "sa_to_vol": "sa_to_vol", "dcmag_hat": "dcmag_hat",
                 "lo": "lo", "hi": "hi", "n_use": "n_use", "c_use": "c_use", "unit_n_use": "unit_n_use",
                 "area_use": "area_use", "aod_use": "aod_use", "n_fa": "n_fa", "cvocv_i": "cvocv_i",
                 "cvocv_v": "cvocv_v", "grad_coeff": "grad_coeff"}
        for field, arr in names.items():
            a = np.ascontiguousarray(mesh.get(arr))
            self._keep[field] = a
            setattr(m, field, a.ctypes.data)
        for field, a in (("filt_w", filter_weights(mesh)), ("r_vv", np.ascontiguousarray(mesh.get("r_vv")))):
            self._keep[field] = a
            setattr(m, field, a.ctypes.data)
        m.nbf = getattr(mesh, "nbf", 0)
        if m.nbf:
            for field in ("cvobf", "zone_bf", "n_bf", "aod_bf", "x_cv"):
                a = np.ascontiguousarray(mesh.get(field))
                self._keep[field] = a
                setattr(m, field, a.ctypes.data)
            try:
                for d, v in enumerate(mesh.get("domain_max")):
                    m.x_max[d] = float(v)
            except RuntimeError:
                pass  # a mesh from raw arrays: no extents (a SPONGE cannot run on it)
        self._m = m
        zone_area = ()
        if m.nbf:
            # area_global: each zone's face areas summed in face order (bincount adds in index order)
            zone_area = np.bincount(self._keep["zone_bf"], weights=mesh.get("area_bf"),
                                    minlength=len(mesh.zones))
        self._p = params_from(cgparams, getattr(mesh, "zones", ()), zone_area)
        self.ncv = mesh.ncv
        self._h = self._L.og_create(C.byref(self._m), C.byref(self._p))
        self._L.og_set_threads(self._h, int(threads))

    def set_state(self, rho, u, rhoE):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.ncv, 3)
        rhoE = np.ascontiguousarray(rhoE, dtype=np.float64)
        self._L.og_set_state(self._h, dptr(rho), dptr(u), dptr(rhoE))

    def get_state(self):
        rho, u, rhoE = np.empty(self.ncv), np.empty((self.ncv, 3)), np.empty(self.ncv)
        self._L.og_get_state(self._h, dptr(rho), dptr(u), dptr(rhoE))
        return rho, u, rhoE

    def get_field(self, name):
        out = np.empty({"dudx": (self.ncv, 3, 3), "u_sponge": (self.ncv, 3)}.get(name, (self.ncv,)))
        if self._L.og_get_field(self._h, name.encode(), dptr(out)) != 0:
            raise KeyError(name)
        return out

    def init_complete(self):
        self._L.og_init_complete(self._h)

    def store(self):
        self._L.og_store(self._h)

    def restore(self):
        self._L.og_restore(self._h)

    def set_time(self, step, time, dt):
        self._L.og_set_time(self._h, int(step), float(time), float(dt))

    def set_dsm_state(self, lijmij, mijmij):
        lm = np.ascontiguousarray(lijmij, dtype=np.float64)
        mm = np.ascontiguousarray(mijmij, dtype=np.float64)
        self._L.og_set_dsm_state(self._h, dptr(lm), dptr(mm))

    def set_scalar(self, isc, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        self._L.og_set_scalar(self._h, int(isc), dptr(phi))

    def get_scalar(self, isc):
        out = np.empty(self.ncv)
        self._L.og_get_scalar(self._h, int(isc), dptr(out))
        return out

    def calc_rhs_scalars(self):
        out = np.empty((self.ncv, max(self._p.nsc, 1)))
        self._L.og_calc_rhs_scalars(self._h, dptr(out))
        return out

    def sgs_props(self):
        self._L.og_sgs_props(self._h)

    def sync_post_state(self):
        """syncPostState (IdealGasSolver.hpp:323-330): primitives, gradient,
        SGS and material properties of the current state."""
        self._L.og_update_state(self._h)
        self._L.og_sgs_props(self._h)

    def set_source_hook(self, fn):
        """fn(rhs [ncv][5] numpy view, time, rk_stage) adds to the stage RHS in place
        (addSourceHook, IdealGasSolver.hpp:702,713); None removes it."""
        if fn is None:
            self._hook = None
            self._L.og_set_source_hook(self._h, OG_SOURCE_FN(), None)
            return
        ncv = self.ncv

        def tramp(user, rhs, n, time, stage):
            fn(np.ctypeslib.as_array(rhs, shape=(ncv, 5)), time, stage)

        self._hook = OG_SOURCE_FN(tramp)
        self._L.og_set_source_hook(self._h, self._hook, None)

    def filter_cv(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        ncomp = 1 if phi.ndim == 1 else phi.shape[1]
        out = np.empty_like(phi)
        self._L.og_filter_cv(C.byref(self._m), dptr(phi), dptr(out), ncomp)
        return out

    def calc_rhs(self):
        out = np.empty((self.ncv, 5))
        self._L.og_calc_rhs(self._h, dptr(out))
        return out

    def calc_cfl(self, dt):
        return self._L.og_calc_cfl(self._h, dt)

    def step(self, n=1):
        rc = 0
        for _ in range(n):
            rc = self._L.og_step(self._h)
        return rc

    def diag(self):
        d = og_diag()
        self._L.og_diag_get(self._h, C.byref(d))
        return {"step": d.step, "time": d.time, "dt": d.dt, "ke": d.ke, "enstrophy": d.enstrophy,
                "sums": list(d.sums), "cfl_max": d.cfl_max, "nan_flag": d.nan_flag}

    def close(self):
        if self._h:
            self._L.og_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
