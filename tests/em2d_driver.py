"""Drive any library that exports the em2d C API through ctypes.

The same code runs the reference build (oracle/_ref/libref_em2d.so,
reference struct layout), the oracle restatement (oracle/_build/liboracle.so)
and the GPU library (zpic_amd/libzpic_em2d.so).  A run is described by a
plain dict so golden fixtures can record it:

    {"nx": [32, 32], "box": [3.2, 3.2], "dt": 0.07, "seed": [12345, 67890],
     "species": [{"name": "electrons", "m_q": -1.0, "ppc": [2, 2],
                  "ufl": [0, 0, 0.6], "uth": [0.1, 0.1, 0.1], "n_sort": 16,
                  "density": {"type": "uniform", "n": 1.0, "start": 0, "end": 0}}],
     "smooth": {"xtype": "binomial", "ytype": "binomial", "xlevel": 1, "ylevel": 1},
     "moving_window": False,
     "laser": {...t_emf_laser fields, "type": "gaussian"...},
     "init_fld": {"E_0": [..], "B_0": [..]},      # uniform initial fields
     "ext_fld": {"E_0": [..], "B_0": [..]}}       # uniform external fields

Construction follows the order of the reference decks and of the Python
module (em2d/input/lwfa.c:12-65; python/source/em2d.pyx:1772-1840):
seed, spec_new for every species, sim_new, then laser, moving window,
smoothing and fields.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from zpic_amd import abi as A


class Em2dRun:
    def __init__(self, lib: C.CDLL, cfg: dict, with_dev: bool = True, sync=None, before_sim=None):
        self.lib = lib
        self.cfg = cfg
        self.abi = A.make_abi(with_dev)
        A.bind_api(lib, self.abi)
        self.sync = sync
        nsp = len(cfg["species"])
        nx = (C.c_int * 2)(*cfg["nx"])
        box = (C.c_float * 2)(*cfg["box"])
        dt = C.c_float(cfg["dt"])
        seed = cfg.get("seed", [12345, 67890])
        # start from an empty deviate cache where the library can clear it
        # (the reference's set_rand_seed keeps it, em2d/random.c:25-29)
        for clear in ("orc_rng_reset", "zpic_rand_clear"):
            if hasattr(lib, clear):
                getattr(lib, clear)()
        lib.set_rand_seed(seed[0], seed[1])
        # the species array is handed to sim_new, which frees it in sim_delete
        libc = C.CDLL(None)
        libc.calloc.restype = C.c_void_p
        libc.calloc.argtypes = [C.c_size_t, C.c_size_t]
        raw = libc.calloc(max(nsp, 1), C.sizeof(self.abi.t_species))
        self.species = C.cast(raw, C.POINTER(self.abi.t_species))
        self._keep = []
        for k, sd in enumerate(cfg["species"]):
            ppc = (C.c_int * 2)(*sd["ppc"])
            ufl = (C.c_float * 3)(*sd.get("ufl", [0, 0, 0]))
            uth = (C.c_float * 3)(*sd.get("uth", [0, 0, 0]))
            dens = None
            if "density" in sd:
                dd = sd["density"]
                d = A.t_density()
                d.n = dd.get("n", 1.0)
                d.type = A.DENSITY[dd.get("type", "uniform")]
                d.start = dd.get("start", 0.0)
                d.end = dd.get("end", 0.0)
                for ax in ("x", "y"):
                    if f"custom_{ax}" in dd:
                        cb = A.DENSITY_FN(lambda v, data, fn=dd[f"custom_{ax}"]: float(fn(v)))
                        self._keep.append(cb)
                        setattr(d, f"custom_{ax}", cb)
                dens = C.pointer(d)
                self._keep.append(d)
            lib.spec_new(C.byref(self.species[k]), sd["name"].encode(), C.c_float(sd["m_q"]),
                         ppc, ufl, uth, nx, box, dt, dens)
            if "n_sort" in sd:
                self.species[k].n_sort = sd["n_sort"]
        if before_sim is not None:
            before_sim(self)
        self.sim = self.abi.t_simulation()
        lib.sim_new(C.byref(self.sim), nx, box, dt, C.c_float(cfg.get("tmax", 0.0)), 0,
                    self.species, nsp)
        if cfg.get("init_fld"):
            f = A.t_emf_init_fld()
            self._field_spec(f, cfg["init_fld"])
            lib.emf_init_fld(C.byref(self.sim.emf), C.byref(f))
        if cfg.get("laser"):
            L = A.t_emf_laser()
            for key, val in cfg["laser"].items():
                setattr(L, key, A.LASER_TYPE[val] if key == "type" else val)
            lib.sim_add_laser(C.byref(self.sim), C.byref(L))
        if cfg.get("moving_window"):
            lib.sim_set_moving_window(C.byref(self.sim))
        if cfg.get("smooth"):
            s = cfg["smooth"]
            sm = A.t_smooth(A.SMOOTH[s.get("xtype", "none")], A.SMOOTH[s.get("ytype", "none")],
                            s.get("xlevel", 0), s.get("ylevel", 0))
            lib.sim_set_smooth(C.byref(self.sim), C.byref(sm))
        if cfg.get("ext_fld"):
            x = A.t_emf_ext_fld()
            self._field_spec(x, cfg["ext_fld"])
            lib.sim_set_ext_fld(C.byref(self.sim), C.byref(x))
        self.n = 0

    def _field_spec(self, f, spec: dict):
        """Fill an initial/external field struct: uniform "E_0"/"B_0" values or
        Python callables "E_custom"/"B_custom" (ix, dx, iy, dy) -> 3 floats."""
        for F in ("E", "B"):
            if f"{F}_custom" in spec:
                from zpic_amd import _lib

                fptr, data = A.field_callback(_lib.load(), spec[f"{F}_custom"], self._keep)
                setattr(f, f"{F}_type", A.FLD_TYPE["custom"])
                setattr(f, f"{F}_custom", fptr)
                setattr(f, f"{F}_custom_data", data)
            elif f"{F}_0" in spec:
                setattr(f, f"{F}_type", A.FLD_TYPE["uniform"])
                setattr(f, f"{F}_0", A.float3(*spec[f"{F}_0"]))

    def step(self, n: int = 1):
        for _ in range(n):
            self.lib.sim_iter(C.byref(self.sim))
            self.n += 1

    def state(self) -> dict:
        """Copies of every array and scalar of the step state."""
        if self.sync is not None:
            self.sync(self)
        nx = tuple(self.cfg["nx"])
        out = {"E": A.grid_view(self.sim.emf.E_buf, nx).copy(),
               "B": A.grid_view(self.sim.emf.B_buf, nx).copy(),
               "J": A.grid_view(self.sim.current.J_buf, nx).copy()}
        e6 = (C.c_double * 6)()
        self.lib.emf_get_energy(C.byref(self.sim.emf), e6)
        out["emf_energy"] = np.array(e6[:], dtype=np.float64)
        kin = []
        for k in range(self.sim.n_species):
            s = self.species[k]
            out[f"part{k}"] = A.particles_view(s).copy()
            kin.append(s.energy)
            out[f"n_move{k}"] = s.n_move
        out["kin_energy"] = np.array(kin, dtype=np.float64)
        out["emf_n_move"] = self.sim.emf.n_move
        return out

    def charge(self, k: int) -> np.ndarray:
        nx = self.cfg["nx"]
        rho = np.zeros((nx[1] + 1) * (nx[0] + 1), dtype=np.float32)
        self.lib.spec_deposit_charge(C.byref(self.species[k]),
                                     rho.ctypes.data_as(C.POINTER(C.c_float)))
        return rho.reshape(nx[1] + 1, nx[0] + 1)

    def close(self):
        if self.sim is not None:
            self.lib.sim_delete(C.byref(self.sim))
            self.sim = None


def weibel(nx=32, ny=32, ppc=2, uz=0.6, uth=0.1, n_sort=16, smooth=None, dx=0.1, dt=0.07,
           seed=(12345, 67890), m_q2=1.0) -> dict:
    """Two counter-streaming species (em2d/input/weibel.c:10-42): electrons and,
    by default, positrons; m_q2=-1 gives two electron populations as in the
    benchmark configurations."""
    return {"nx": [nx, ny], "box": [nx * dx, ny * dx], "dt": dt, "seed": list(seed),
            "species": [
                {"name": "electrons", "m_q": -1.0, "ppc": [ppc, ppc], "ufl": [0, 0, uz],
                 "uth": [uth, uth, uth], "n_sort": n_sort},
                {"name": "positrons" if m_q2 > 0 else "electrons2", "m_q": m_q2, "ppc": [ppc, ppc],
                 "ufl": [0, 0, -uz], "uth": [uth, uth, uth], "n_sort": n_sort}],
            "smooth": smooth}


def lwfa(nx=200, ny=32, dx=0.05, dy=0.2, dt=0.045, start=None, smooth_level=4,
         a0=2.0, omega0=10.0, W0=4.0, fwhm=2.0, focus=None, laser_start=None,
         polarization=float(np.pi / 2)) -> dict:
    """Laser wakefield with a moving window (em2d/input/lwfa.c:12-65)."""
    box = [nx * dx, ny * dy]
    focus = box[0] * 0.6 if focus is None else focus
    laser_start = box[0] * 0.9 if laser_start is None else laser_start
    start = box[0] if start is None else start
    return {"nx": [nx, ny], "box": box, "dt": dt,
            "species": [{"name": "electrons", "m_q": -1.0, "ppc": [4, 2], "n_sort": 16,
                         "density": {"type": "step", "start": start, "n": 1.0}}],
            "laser": {"type": "gaussian", "start": laser_start, "fwhm": fwhm, "a0": a0,
                      "omega0": omega0, "W0": W0, "focus": focus, "axis": box[1] / 2,
                      "polarization": polarization},
            "moving_window": True,
            "smooth": {"xtype": "compensated", "xlevel": smooth_level}}
