"""Batched solves over the C ABI (cg_batch_* in include/charles_gpu.h).

A Batch holds many independent periodic-box solves of one precision on the
GPU; every kernel launch of a step covers all of them and the step is
replayed from a captured HIP graph. Instances are added (mesh, parameters,
initial state) and removed between steps, so the same object backs the
fixed ensemble and the serving loop (nextgen_amd/serve.py).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .ideal_gas import CG_PREC_FP64, CG_PREC_MIXED, CharlesError, Diag, Mesh, cg_diag, cg_params, load_library

BATCH_EXPORTS = [
    "cg_batch_create", "cg_batch_add", "cg_batch_remove", "cg_batch_step", "cg_batch_get_diag",
    "cg_batch_get_state", "cg_batch_counts", "cg_batch_set_option", "cg_batch_time_steps",
    "cg_batch_last_error", "cg_batch_destroy", "cg_instance_create", "cg_instance_bytes", "cg_instance_destroy",
    "cg_batch_attach", "cg_batch_tune", "cg_layout_create", "cg_layout_bytes", "cg_layout_destroy",
    "cg_instance_create_on",
]

_ready = False


def _lib():
    global _ready
    lib = load_library()
    if not _ready:
        P = C.POINTER
        vp = C.c_void_p
        dp = P(C.c_double)
        sig = {
            "cg_batch_create": (C.c_int, [C.c_int, P(vp)]),
            "cg_batch_add": (C.c_int, [vp, vp, P(cg_params), dp, dp, dp, P(C.c_int)]),
            "cg_batch_remove": (C.c_int, [vp, C.c_int]),
            "cg_batch_step": (C.c_int, [vp, C.c_int]),
            "cg_batch_get_diag": (C.c_int, [vp, C.c_int, P(cg_diag)]),
            "cg_batch_get_state": (C.c_int, [vp, C.c_int, dp, dp, dp]),
            "cg_batch_counts": (C.c_int, [vp, P(C.c_int64)]),
            "cg_batch_set_option": (C.c_int, [vp, C.c_char_p, C.c_int]),
            "cg_batch_time_steps": (C.c_int, [vp, C.c_int, dp]),
            "cg_batch_last_error": (C.c_char_p, [vp]),
            "cg_batch_destroy": (None, [vp]),
            "cg_instance_create": (C.c_int, [vp, P(cg_params), dp, dp, dp, P(vp)]),
            "cg_instance_bytes": (C.c_int, [vp, P(C.c_int64)]),
            "cg_layout_create": (C.c_int, [vp, C.c_int, P(vp)]),
            "cg_layout_bytes": (C.c_int, [vp, P(C.c_int64)]),
            "cg_layout_destroy": (None, [vp]),
            "cg_instance_create_on": (C.c_int, [vp, P(cg_params), dp, dp, dp, P(vp)]),
            "cg_instance_destroy": (None, [vp]),
            "cg_batch_attach": (C.c_int, [vp, vp, P(C.c_int)]),
            "cg_batch_tune": (C.c_int, [vp, C.c_int, C.c_char_p, C.c_int64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _ready = True
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Layout:
    """A mesh's device layout (cg_layout_create), shared by every solve created
    on it with Instance(..., layout=...)."""

    def __init__(self, mesh: Mesh, precision=CG_PREC_FP64):
        self.ncv = mesh.ncv
        self.precision = precision
        h = C.c_void_p()
        rc = _lib().cg_layout_create(mesh._h, int(precision), C.byref(h))
        if rc:
            raise CharlesError(f"error {rc}: {load_library().cg_error_string().decode()}")
        self._h = h

    def bytes(self) -> int:
        b = C.c_int64()
        if _lib().cg_layout_bytes(self._h, C.byref(b)):
            raise CharlesError(load_library().cg_error_string().decode())
        return b.value

    def close(self):
        if self._h:
            _lib().cg_layout_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Instance:
    """A solve prepared outside any batch (cg_instance_create, or
    cg_instance_create_on a shared Layout): device buffers, state and
    initComplete. Safe to build on a worker thread while a batch steps; hand
    it to a batch with Batch.attach."""

    def __init__(self, mesh: Mesh, params: cg_params, rho, u, rhoE, layout: Layout | None = None):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(mesh.ncv, 3)
        rhoE = np.ascontiguousarray(rhoE, dtype=np.float64)
        self.ncv = mesh.ncv
        self.precision = params.precision
        h = C.c_void_p()
        if layout is None:
            rc = _lib().cg_instance_create(mesh._h, C.byref(params), _dptr(rho), _dptr(u), _dptr(rhoE), C.byref(h))
        else:
            rc = _lib().cg_instance_create_on(layout._h, C.byref(params), _dptr(rho), _dptr(u), _dptr(rhoE),
                                              C.byref(h))
        if rc:
            raise CharlesError(f"error {rc}: {load_library().cg_error_string().decode()}")
        self._h = h

    def bytes(self):
        """(device bytes of the solve itself, of its mesh layout)"""
        b = (C.c_int64 * 2)()
        if _lib().cg_instance_bytes(self._h, b):
            raise CharlesError(load_library().cg_error_string().decode())
        return b[0], b[1]

    def close(self):
        if self._h:
            _lib().cg_instance_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent solves advanced together on one GPU."""

    def __init__(self, precision=CG_PREC_FP64):
        self.precision = precision
        h = C.c_void_p()
        rc = _lib().cg_batch_create(int(precision), C.byref(h))
        if rc:
            raise CharlesError(f"error {rc}: {load_library().cg_error_string().decode()}")
        self._h = h
        self.ncv = {}

    def _chk(self, rc):
        if rc != 0:
            msg = _lib().cg_batch_last_error(self._h)
            raise CharlesError(f"error {rc}: {msg.decode() if msg else ''}")

    def add(self, mesh: Mesh, params: cg_params, rho, u, rhoE) -> int:
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(mesh.ncv, 3)
        rhoE = np.ascontiguousarray(rhoE, dtype=np.float64)
        i = C.c_int()
        self._chk(_lib().cg_batch_add(self._h, mesh._h, C.byref(params), _dptr(rho), _dptr(u), _dptr(rhoE),
                                      C.byref(i)))
        self.ncv[i.value] = mesh.ncv
        return i.value

    def attach(self, inst: Instance) -> int:
        """Hand a prepared instance to the batch (which owns it from then on)."""
        i = C.c_int()
        self._chk(_lib().cg_batch_attach(self._h, inst._h, C.byref(i)))
        inst._h = None
        self.ncv[i.value] = inst.ncv
        return i.value

    def remove(self, iid: int):
        self._chk(_lib().cg_batch_remove(self._h, int(iid)))
        self.ncv.pop(iid, None)

    def step(self, nsteps=1):
        self._chk(_lib().cg_batch_step(self._h, int(nsteps)))

    def diag(self, iid: int) -> Diag:
        d = cg_diag()
        self._chk(_lib().cg_batch_get_diag(self._h, int(iid), C.byref(d)))
        return Diag(**d.as_dict())

    def state(self, iid: int):
        n = self.ncv[iid]
        rho, u, rhoE = np.empty(n), np.empty((n, 3)), np.empty(n)
        self._chk(_lib().cg_batch_get_state(self._h, int(iid), _dptr(rho), _dptr(u), _dptr(rhoE)))
        return rho, u, rhoE

    def counts(self):
        c = (C.c_int64 * 4)()
        self._chk(_lib().cg_batch_counts(self._h, c))
        return {"instances": c[0], "cells": c[1], "device_bytes": c[2], "lds_instances": c[3]}

    def set_option(self, name: str, value: int):
        self._chk(_lib().cg_batch_set_option(self._h, name.encode(), int(value)))

    def tune(self, nsteps=4) -> str:
        """Time each launch variant for nsteps steps (the instances advance) and
        keep the fastest; returns the timing table."""
        buf = C.create_string_buffer(1024)
        self._chk(_lib().cg_batch_tune(self._h, int(nsteps), buf, len(buf)))
        return buf.value.decode()

    def time_steps(self, nsteps) -> float:
        ms = C.c_double()
        self._chk(_lib().cg_batch_time_steps(self._h, int(nsteps), C.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            _lib().cg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Batch", "Instance", "Layout", "BATCH_EXPORTS", "CG_PREC_FP64", "CG_PREC_MIXED"]
