"""Device-resident HydroPol2D step loop (the HydroPol2D_Main_While mirror).

`Simulation(spec)` uploads a RunSpec once; `run()` advances the loop with no
per-step host round-trip (dt, rainfall integral, hydrograph, FP64 mass
ledgers all live on the device); the getters copy results back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib, abi
from .runspec import RunSpec

_dp = C.POINTER(C.c_double)


class Simulation:
    def __init__(self, spec: RunSpec):
        L = _lib.load()
        _lib.require_device()
        self.spec = spec
        self.ny, self.nx = spec.shape
        cfg, ins, keep = spec.to_c()
        ptr = C.c_void_p()
        status = L.hp2d_sim_create(C.byref(cfg), C.byref(ins), C.byref(ptr))
        del keep
        _lib.check(status)
        self._s = ptr
        self._L = L

    def close(self):
        if getattr(self, "_s", None):
            self._L.hp2d_sim_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, status):
        _lib.check(status, sim=self._s)

    def run(self, max_steps: int = 0) -> bool:
        """Advance up to max_steps accepted steps (0 = to the end). Returns done."""
        done = C.c_int32(0)
        self._check(self._L.hp2d_sim_run(self._s, int(max_steps), C.byref(done)))
        return bool(done.value)

    def run_fixed(self, n_steps: int, graph_steps: int = 0):
        self._check(self._L.hp2d_sim_run_fixed(self._s, int(n_steps), int(graph_steps)))

    def autotune(self):
        """Pick the fastest launch shape on this sim's data (does not advance it)."""
        tj, wpb, ms = C.c_int32(), C.c_int32(), C.c_double()
        self._check(self._L.hp2d_sim_autotune(self._s, C.byref(tj), C.byref(wpb), C.byref(ms)))
        return dict(self.launch_shape(), ms=ms.value)

    def launch_shape(self) -> dict:
        """The step kernel's launch shape: columns per wave task, waves per
        block and grid rows per lane."""
        tj, wpb, rows = C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self._L.hp2d_sim_launch_shape(self._s, C.byref(tj), C.byref(wpb), C.byref(rows)))
        return {"tj": tj.value, "waves_per_block": wpb.value, "rows": rows.value}

    def set_launch_shape(self, tj: int, waves_per_block: int, rows: int):
        """Set the step kernel's launch shape (results do not depend on it)."""
        self._check(self._L.hp2d_sim_set_launch_shape(self._s, int(tj), int(waves_per_block), int(rows)))

    def kernel_time(self):
        li, tot, n = C.c_double(), C.c_double(), C.c_int64()
        self._check(self._L.hp2d_sim_kernel_time(self._s, C.byref(li), C.byref(tot), C.byref(n)))
        return li.value, tot.value, n.value

    def inject_failure(self, step_k: int):
        self._check(self._L.hp2d_debug_inject_failure(self._s, int(step_k), 1))

    def report(self) -> dict:
        rep = abi.RunReport()
        self._check(self._L.hp2d_sim_report(self._s, C.byref(rep)))
        return rep.as_dict()

    def state(self, infiltration: bool = False, water_quality: bool = False,
              groundwater: bool = False) -> dict:
        """Current state: d_t [mm], outflow pages [mm/h], dmax [mm]; optionally I_t
        [mm], the water-quality maps (B_t [kg], P_conc [mg/L], Tot_Washed [kg]) and
        the groundwater head gw_h [m]."""
        ny, nx = self.ny, self.nx
        d = np.zeros((ny, nx), order="F")
        of = np.zeros((ny, nx, 3), order="F")
        dm = np.zeros((ny, nx), order="F")
        It = np.zeros((ny, nx), order="F") if infiltration else None
        Bt = np.zeros((ny, nx), order="F") if water_quality else None
        gh = np.zeros((ny, nx), order="F") if groundwater else None
        p = (lambda a: a.ctypes.data_as(_dp) if a is not None else _dp())
        self._check(self._L.hp2d_sim_get_state(self._s, p(d), p(of), _dp(), p(It), p(dm), p(Bt),
                                               p(gh)))
        out = {"d_t": d, "outflow": of, "dmax": dm}
        if infiltration:
            out["I_t"] = It
        if water_quality:
            pc = np.zeros((ny, nx), order="F")
            tw = np.zeros((ny, nx), order="F")
            self._check(self._L.hp2d_sim_get_wq_maps(self._s, p(pc), p(tw)))
            out.update(B_t=Bt, P_conc=pc, Tot_Washed=tw)
        if groundwater:
            out["gw_h"] = gh
        return out

    def hydrograph(self):
        n = self.spec.n_hydro_records()
        q, t, d = np.zeros(n), np.zeros(n), np.zeros(n)
        got = C.c_int32()
        self._check(self._L.hp2d_sim_get_hydrograph(self._s, q.ctypes.data_as(_dp),
                                                    t.ctypes.data_as(_dp), d.ctypes.data_as(_dp),
                                                    n, C.byref(got)))
        return {"hydro_q": q, "hydro_t": t, "hydro_depth": d, "n_hydro": got.value}

    def record_time_to_peak(self):
        """Start recording, per cell, the model time [min] of the step that last
        raised its maximum depth (a derived output; the reference keeps dmax only)."""
        self._check(self._L.hp2d_sim_record_time_to_peak(self._s))

    def time_to_peak(self):
        t = np.zeros((self.ny, self.nx), order="F")
        self._check(self._L.hp2d_sim_get_time_to_peak(self._s, t.ctypes.data_as(_dp)))
        return t

    def max_risk(self):
        """Human_Instability.max_risk (Human_Instability_Module.m:22)."""
        m = np.zeros((self.ny, self.nx), order="F")
        self._check(self._L.hp2d_sim_get_max_risk(self._s, m.ctypes.data_as(_dp)))
        return m

    def map_records(self):
        """The map stacks taken so far (save_output_maps.m:148-258): a dict of
        field -> ny x nx x n_records arrays (the fields this run records) and
        "t_min" -> the model time of each record."""
        from .abi import MAP_FIELDS
        n = C.c_int32()
        self._check(self._L.hp2d_sim_get_map_records(self._s, 0, 0, 0, None, None, C.byref(n)))
        nr = n.value
        t = np.zeros(nr)
        out = {"t_min": t}
        for f, name in enumerate(MAP_FIELDS):
            buf = np.zeros(nr * self.ny * self.nx)
            st = self._L.hp2d_sim_get_map_records(self._s, f, 0, nr, buf.ctypes.data_as(_dp),
                                                  t.ctypes.data_as(_dp), C.byref(n))
            if st != 0:
                continue  # a field this run does not record
            out[name] = np.asfortranarray(buf.reshape(nr, self.nx, self.ny).transpose(2, 1, 0))
        return out

    def dt_history(self, n_max: int = 4096):
        buf = np.zeros(n_max)
        got = C.c_int32()
        self._check(self._L.hp2d_sim_get_dt_history(self._s, buf.ctypes.data_as(_dp), n_max,
                                                    C.byref(got)))
        return buf[: got.value]


class _Member(Simulation):
    """A member of an Ensemble: the Simulation getters on a borrowed handle."""

    def __init__(self, ens, i, spec, ptr):
        self.spec = spec
        self.ny, self.nx = spec.shape
        self._s = ptr
        self._L = ens._L
        self._ens = ens          # keeps the owner alive

    def close(self):             # the ensemble owns the handle
        self._s = None


class Ensemble:
    """Independent runs (members) advanced together: one batched launch of the
    fused step over every unfinished member's tiles and one batched launch of
    the members' dt / bookkeeping tails per step, each member with its own grid
    size, inputs, clock and adaptive dt (the member loops of
    HydroPol2D_Main_While_Ensemble.m)."""

    def __init__(self, specs, cols_per_wave: int = 0):
        L = _lib.load()
        _lib.require_device()
        self.specs = list(specs)
        n = len(self.specs)
        cfgs = (abi.RunConfig * n)()
        ins = (abi.RunInputs * n)()
        keep = []
        for i, sp in enumerate(self.specs):
            c, inp, k = sp.to_c()
            cfgs[i] = c
            ins[i] = inp
            keep.append(k)
        ptr = C.c_void_p()
        status = L.hp2d_ensemble_create(n, cfgs, ins, int(cols_per_wave), C.byref(ptr))
        del keep
        _lib.check(status)
        self._e = ptr
        self._L = L

    def close(self):
        if getattr(self, "_e", None):
            self._L.hp2d_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.specs)

    def _check(self, status):
        if status != abi.HP2D_OK:
            msg = self._L.hp2d_ensemble_last_error(self._e)
            raise _lib.HP2DError(status, (msg or b"").decode(errors="replace"))

    def run(self, max_steps: int = 1 << 62, graph_steps: int = 50) -> int:
        """Advance unfinished members by up to max_steps steps; returns how many are done."""
        done = C.c_int32(0)
        self._check(self._L.hp2d_ensemble_run(self._e, int(max_steps), int(graph_steps),
                                              C.byref(done)))
        return done.value

    def elapsed_ms(self) -> float:
        ms = C.c_double()
        self._check(self._L.hp2d_ensemble_elapsed(self._e, C.byref(ms)))
        return ms.value

    def record_time_to_peak(self):
        self._check(self._L.hp2d_ensemble_record_time_to_peak(self._e))

    def member(self, i: int) -> Simulation:
        ptr = C.c_void_p()
        self._check(self._L.hp2d_ensemble_member(self._e, int(i), C.byref(ptr)))
        return _Member(self, i, self.specs[i], ptr)
