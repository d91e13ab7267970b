"""Batched instances: many FluidSystem handles stepped together (mg_batch_* in
include/mg_sph.h). No counterpart in the reference, which steps one FluidSystem at a time."""
from __future__ import annotations

import ctypes as C

from morphogenesis_amd import _ffi as F


class Batch:
    """Steps the given FluidSystem objects together. While the batch exists their work runs on
    the batch's stream; close() (or garbage collection) releases them."""

    def __init__(self, systems):
        self.systems = list(systems)
        if not self.systems:
            raise ValueError("Batch needs at least one FluidSystem")
        self.lib = self.systems[0].lib
        arr = (C.c_void_p * len(self.systems))(*[s.h.value for s in self.systems])
        self.h = C.c_void_p()
        self._call("batch_create", C.byref(self.h), arr, len(self.systems))

    def _call(self, name, *args):
        r = getattr(self.lib, f"mg_{name}")(*args)
        if r != F.MG_OK:
            raise F.MgError(f"mg_{name} -> {F.ERRORS.get(r, r)}: {self.lib.mg_last_error().decode()}")
        return r

    def step(self, nsteps: int, dt: float, flags: int):
        self._call("batch_step", self.h, int(nsteps), float(dt), int(flags))

    def sync(self):
        self._call("batch_sync", self.h)

    def classes(self):
        n = C.c_int()
        self._call("batch_classes", self.h, C.byref(n), None, None)
        members = (C.c_int * n.value)()
        threads = (C.c_uint * n.value)()
        self._call("batch_classes", self.h, C.byref(n), members, threads)
        return [(members[i], threads[i]) for i in range(n.value)]

    def autotune(self, steps: int, dt: float, flags: int):
        n = len(self.classes())
        out = (C.c_uint * n)()
        self._call("batch_autotune", self.h, int(steps), float(dt), int(flags), out)
        return list(out)

    def class_config(self):
        """Per size class: {instances, threads, cand_batch, band_cap} of the dispatch form (block
        size and the pressure sweep's tile shape)."""
        out = []
        for k in range(len(self.classes())):
            inst, thr, kb, cap = C.c_int(), C.c_uint(), C.c_int(), C.c_int()
            self._call("batch_class_config", self.h, k, C.byref(inst), C.byref(thr), C.byref(kb), C.byref(cap))
            out.append({"instances": inst.value, "threads": thr.value, "cand_batch": kb.value,
                        "band_cap": cap.value})
        return out

    def set_tile_shape(self, cls: int, cand_batch: int, band_cap: int):
        """Pins size class cls's pressure sweep shape (cand_batch 8 or 4, band_cap 128 or 64)."""
        self._call("batch_set_tile_shape", self.h, int(cls), int(cand_batch), int(band_cap))

    def launch_config(self):
        """(persistent, workgroups): 1 and the workgroup count for one persistent launch per
        step, 0 for the dispatch form."""
        on, wg = C.c_int(), C.c_uint()
        self._call("batch_launch_config", self.h, C.byref(on), C.byref(wg))
        return int(on.value), int(wg.value)

    def set_persistent(self, on: bool, workgroups_per_cu: int = 0):
        self._call("batch_set_persistent", self.h, int(bool(on)), int(workgroups_per_cu))

    def close(self):
        if self.h:
            self.lib.mg_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
