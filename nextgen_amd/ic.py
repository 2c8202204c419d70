"""Initial fields and input decks of the periodic-box workloads.

Taylor-Green vortex on [0, 2*pi]^3 (the TGV deck of the ideal-gas solver,
evaluated at the cell centroids the way an ``INIT`` expression or
``initialHook`` would set it, StaticSolver.hpp:1402-1470, charles_ig.cpp:209-222):

    u =  U0 sin x cos y cos z,  v = -U0 cos x sin y cos z,  w = 0
    p = p0 + rho0 U0^2 / 16 (cos 2x + cos 2y)(cos 2z + 2)
    rho = p / (R T0),  rhoE = p/(gamma-1) + rho |u|^2 / 2   (initComplete, IdealGasSolver.hpp:282-285)

Nondimensionalisation: U0 = 1, rho0 = 1, T0 = T_REF = 1, c0 = U0/Ma,
P_REF = p0 = rho0 c0^2 / gamma, RHO_REF = 1 (so R = p0), MU_REF = rho0 U0 / Re,
MU_POWER_LAW 0.75 (the reference default).

The same TGV field, and the seeded k^4 decaying-turbulence field used at
production sizes, are computed by the library itself (``Mesh.ic_tgv`` and
``Mesh.ic_k4`` over cg_ic_tgv / cg_ic_k4, threaded C++).
"""
from __future__ import annotations

import numpy as np


def tgv_params(Re=1600.0, Ma=0.1, gamma=1.4, sgs="NONE", cfl=0.8):
    """charles.in text of the TGV deck."""
    c0 = 1.0 / Ma
    p0 = c0 * c0 / gamma
    mu = 1.0 / Re
    return (
        f"GAMMA {gamma!r}\nP_REF {p0!r}\nRHO_REF 1.0\nT_REF 1.0\nMU_REF {mu!r}\n"
        f"MU_POWER_LAW 0.75\nPR_LAM 0.7\nSGS_MODEL {sgs}\nTIMESTEP CFL {cfl!r}\n"
    )


def k4_params(Mt=0.3, Re=1600.0, gamma=1.4, sgs="VREMAN", cfl=0.8):
    """Decaying-turbulence deck: uniform rho = 1, p = 1/gamma (c0 = 1), u_rms = Mt,
    MU_REF = Mt / Re (unit length)."""
    return (f"GAMMA {gamma!r}\nP_REF {1.0 / gamma!r}\nRHO_REF 1.0\nT_REF 1.0\nMU_REF {Mt / Re!r}\n"
            f"MU_POWER_LAW 0.75\nPR_LAM 0.7\nSGS_MODEL {sgs}\nTIMESTEP CFL {cfl!r}\n")


def tgv_state(x_cv, Ma=0.1, gamma=1.4):
    x, y, z = x_cv[:, 0], x_cv[:, 1], x_cv[:, 2]
    c0 = 1.0 / Ma
    p0 = c0 * c0 / gamma
    R = p0  # RHO_REF = T_REF = 1
    u = np.empty((len(x), 3))
    u[:, 0] = np.sin(x) * np.cos(y) * np.cos(z)
    u[:, 1] = -np.cos(x) * np.sin(y) * np.cos(z)
    u[:, 2] = 0.0
    p = p0 + (1.0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    rho = p / (R * 1.0)
    rhoE = p / (gamma - 1.0) + 0.5 * rho * (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1] + u[:, 2] * u[:, 2])
    return rho, u, rhoE
