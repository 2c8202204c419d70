"""Ensembles of independent periodic-box solves (one Batch, one fused launch
per kernel) and the decks they are drawn from.

Ensemble: `count` solves; each draws its lattice side n uniformly from
[n_lo, n_hi] (jitter +-0.3h, its own mesh seed), Re uniformly from
[400, 3200], Ma uniformly from [0.05, 0.3], and a Taylor-Green or a k^4
decaying-turbulence initial field by a seeded coin flip.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from .ic import k4_params, tgv_params
from .ideal_gas import CG_PREC_FP64, CG_PREC_MIXED, Mesh, parse_params


@dataclass
class Spec:
    n: int          # lattice side (n^3 cells)
    mesh_seed: int
    ic: str         # "tgv" or "k4"
    Re: float
    Ma: float       # TGV Mach number, or turbulent Mach number of the k^4 field
    ic_seed: int
    steps: int = 0  # horizon (serving)

    @property
    def k0(self):
        # spectrum peak: resolved with kmax = ceil(2.5 k0) below the lattice Nyquist number n/2
        return 2.0 if self.n < 64 else 4.0

    def deck(self, precision=CG_PREC_FP64, check_interval=100):
        text = tgv_params(Re=self.Re, Ma=self.Ma, sgs="VREMAN") if self.ic == "tgv" else k4_params(Mt=self.Ma, Re=self.Re)
        prec = "MIXED" if precision == CG_PREC_MIXED else "FP64"
        return parse_params(text + f"CHECK_INTERVAL {check_interval}\nPRECISION {prec}\n")

    def state(self, mesh: Mesh):
        if self.ic == "tgv":
            return mesh.ic_tgv(self.Ma, 1.4)
        return mesh.ic_k4(self.k0, self.Ma, 1.4, seed=self.ic_seed)

    def mesh(self):
        return Mesh.jittered(self.n, 0.3, seed=self.mesh_seed)


def ensemble_specs(count=256, n_lo=16, n_hi=48, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(Spec(n=int(rng.integers(n_lo, n_hi + 1)), mesh_seed=int(rng.integers(1, 2**31)),
                        ic="tgv" if rng.random() < 0.5 else "k4", Re=float(rng.uniform(400.0, 3200.0)),
                        Ma=float(rng.uniform(0.05, 0.3)), ic_seed=int(rng.integers(1, 2**31))))
    return out


def build_batch(specs, precision, log=None):
    """A Batch holding every spec (meshes are freed once their instance is on the device)."""
    from .batch import Batch

    b = Batch(precision)
    ids = []
    t0 = time.time()
    for k, sp in enumerate(specs):
        m = sp.mesh()
        ids.append(b.add(m, sp.deck(precision), *sp.state(m)))
        m.close()
        if log and (k + 1) % 32 == 0:
            log(f"  {k + 1}/{len(specs)} instances on the device ({time.time() - t0:.1f}s)")
    return b, ids


def run_alone(sp: Spec, precision, steps):
    """One solve alone on the GPU (a batch of one: the same graph replays as a
    batch); returns (seconds of GPU stepping, diag)."""
    from .batch import Batch

    m = sp.mesh()
    b = Batch(precision)
    iid = b.add(m, sp.deck(precision), *sp.state(m))
    m.close()
    ms = b.time_steps(steps)
    d = b.diag(iid)
    b.close()
    return ms * 1e-3, d
