"""em2d - Python interface of the GPU em2d time step.

Same classes, constructor arguments, methods and properties as the
reference's Cython module (python/source/em2d.pyx): Density, Species,
ExternalField, InitialField, EMF, Laser, Current, Smooth and Simulation.
It binds libzpic_em2d.so through ctypes; there is no CPU fallback.

Data access differs in one respect: the reference returned zero-copy
views of host buffers (em2d.pyx:305-312, 1044-1150, 1567-1619).  Here the
simulation state lives in GPU memory, so every data property first copies
the device state into the host mirror and returns a numpy view of that
mirror.  Reading a view costs that copy and nothing more.  Writing through
a view of `Species.particles` or of an EMF field (item assignment,
in-place arithmetic, ufuncs with `out=`, `fill`, `np.copyto`, `np.put`,
`np.place`, `np.putmask`) marks the mirror edited, and edited mirrors are
uploaded before the next iteration, as in the reference where such writes
changed the running simulation.  A view read before an iteration shows the
state of that time: read the property again after stepping.  Writes that
bypass numpy's array interface (e.g. through a `np.void` element of the
particle array, or a raw pointer) must be followed by `mark_edited()`.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib
from . import abi as A

_libc = C.CDLL(None)
_libc.calloc.restype = C.c_void_p
_libc.calloc.argtypes = [C.c_size_t, C.c_size_t]
_libc.free.argtypes = [C.c_void_p]

_ABI = A.make_abi(True)


def _lib_():
    return _lib.load()


def _inner(buf_ptr, nx, comp):
    """Interior component view [ny, nx] of a guard-inclusive float3 grid."""
    g = A.grid_view(buf_ptr, nx)
    return g[1:1 + nx[1], 1:1 + nx[0], comp]


class _Mirror(np.ndarray):
    """View of a host mirror that reports writes.

    `_before` (refresh the mirror from the device) runs before a write and
    `_after` (mark the mirror edited) after it; views taken from a _Mirror
    (slices, fields) report to the same callbacks.  Reads call neither."""

    _WRITERS = {np.copyto: "dst", np.put: "a", np.place: "arr", np.putmask: "a"}

    def __array_finalize__(self, obj):
        self._before = getattr(obj, "_before", None)
        self._after = getattr(obj, "_after", None)

    def _write(self, fn):
        if self._before is not None:
            self._before()
        r = fn()
        if self._after is not None:
            self._after()
        return r

    def __setitem__(self, key, value):
        self._write(lambda: super(_Mirror, self).__setitem__(key, value))

    def fill(self, value):
        self._write(lambda: super(_Mirror, self).fill(value))

    def sort(self, *a, **k):
        self._write(lambda: super(_Mirror, self).sort(*a, **k))

    def __array_ufunc__(self, ufunc, method, *inputs, out=None, **kwargs):
        args = tuple(x.view(np.ndarray) if isinstance(x, _Mirror) else x for x in inputs)
        targets = [o for o in (out or ()) if isinstance(o, _Mirror)]
        if method == "at" and isinstance(inputs[0], _Mirror):
            targets.append(inputs[0])  # ufunc.at updates its first operand in place
        if out is not None:
            kwargs["out"] = tuple(o.view(np.ndarray) if isinstance(o, _Mirror) else o for o in out)
        for t in targets:
            if t._before is not None:
                t._before()
        r = getattr(ufunc, method)(*args, **kwargs)
        for t in targets:
            if t._after is not None:
                t._after()
        if out is not None:
            # the caller's own output arrays (in-place operators rebind to them)
            return out[0] if len(out) == 1 else tuple(out)
        return r

    def __array_function__(self, func, types, args, kwargs):
        name = self._WRITERS.get(func)
        plain_types = tuple(np.ndarray if issubclass(t, _Mirror) else t for t in types)
        if name is None:
            return super().__array_function__(func, plain_types, args, kwargs)
        dst = args[0] if args else kwargs.get(name)
        if not isinstance(dst, _Mirror):
            return super().__array_function__(func, plain_types, args, kwargs)
        return dst._write(lambda: super(_Mirror, dst).__array_function__(func, plain_types, args, kwargs))


def _watched(arr: np.ndarray, before, after) -> np.ndarray:
    m = arr.view(_Mirror)
    m._before = before
    m._after = after
    return m


class Density:
    """Density(type='uniform', start=0.0, end=0.0, n=1.0, custom_x=None, custom_y=None)

    Charge density profile of a particle species (em2d.pyx:25-169).
    """

    def __init__(self, *, type: str = "uniform", start: float = 0.0, end: float = 0.0, n: float = 1.0,
                 custom_x=None, custom_y=None):
        self._d = A.t_density()
        self._d.type = A.DENSITY[type]
        self._d.n = n
        self._d.start = start
        self._d.end = end
        self._fx = self._fy = None
        self.custom_func_x = custom_x
        self.custom_func_y = custom_y
        if custom_x is not None:
            self._fx = A.DENSITY_FN(lambda x, data: float(custom_x(x)))
            self._d.custom_x = self._fx
        if custom_y is not None:
            self._fy = A.DENSITY_FN(lambda y, data: float(custom_y(y)))
            self._d.custom_y = self._fy

    def copy(self):
        return Density(type=self.type, start=self.start, end=self.end, n=self.n,
                       custom_x=self.custom_func_x, custom_y=self.custom_func_y)

    _names = {v: k for k, v in A.DENSITY.items()}

    @property
    def n(self):
        return self._d.n

    @n.setter
    def n(self, v):
        self._d.n = v

    @property
    def type(self):
        return self._names[self._d.type]

    @type.setter
    def type(self, v):
        self._d.type = A.DENSITY[v]

    @property
    def start(self):
        return self._d.start

    @start.setter
    def start(self, v):
        self._d.start = v

    @property
    def end(self):
        return self._d.end

    @end.setter
    def end(self, v):
        self._d.end = v


class Species:
    """Species(name, m_q, ppc, ufl=[0,0,0], uth=[0,0,0], density=None, n_sort=16)

    Particle species (em2d.pyx:171-459).  Particle data is read and written
    through the `particles` property.
    """

    _diag_types = {"charge": 0x1000, "pha": 0x2000, "particles": 0x3000}
    _pha_quants = {"x1": 1, "x2": 2, "u1": 4, "u2": 5, "u3": 6}

    def __init__(self, name: str, m_q: float, ppc=(1, 1), *, ufl=(0.0, 0.0, 0.0), uth=(0.0, 0.0, 0.0),
                 density: Density | None = None, n_sort: int = 16):
        self._name = name
        self._m_q = float(m_q)
        self._ppc = [int(v) for v in ppc]
        self._ufl = [float(v) for v in ufl]
        self._uth = [float(v) for v in uth]
        self._n_sort = int(n_sort)
        self._density = density.copy() if density is not None else Density()
        self._ptr = None

    def _new(self, ptr, nx, box, dt):
        lib = _lib_()
        self._ptr = ptr
        lib.spec_new(ptr, self._name.encode(), C.c_float(self._m_q), (C.c_int * 2)(*self._ppc),
                     (C.c_float * 3)(*self._ufl), (C.c_float * 3)(*self._uth), (C.c_int * 2)(*nx),
                     (C.c_float * 2)(*box), C.c_float(dt), C.byref(self._density._d))
        ptr.contents.n_sort = self._n_sort

    @property
    def _s(self):
        if self._ptr is None:
            raise RuntimeError("species is not part of a Simulation yet")
        return self._ptr.contents

    def _sync(self):
        _lib_().spec_sync_host(self._ptr)

    def add(self, ix, x, u):
        """add(ix, x, u) - append one particle (cell index, position in cell, u)."""
        lib = _lib_()
        self._sync()
        s = self._s
        lib.spec_grow_buffer(self._ptr, s.np + 1)
        p = s.part[s.np]
        p.ix, p.iy = int(ix[0]), int(ix[1])
        p.x, p.y = float(x[0]), float(x[1])
        p.ux, p.uy, p.uz = float(u[0]), float(u[1]), float(u[2])
        s.np = s.np + 1
        lib.spec_mark_host_dirty(self._ptr)

    def report(self, type: str, *, quants=(), pha_nx=(), pha_range=()):
        """report(type, quants=[], pha_nx=[], pha_range=[]) - write a ZDF
        diagnostic ("charge", "pha" or "particles") below the current
        directory (em2d.pyx:268-303)."""
        rep = self._diag_types[type]
        if rep == self._diag_types["pha"]:
            nxa = (C.c_int * 2)(*[int(v) for v in pha_nx])
            rng = ((C.c_float * 2) * 2)()
            for a in range(2):
                for b in range(2):
                    rng[a][b] = pha_range[a][b]
            rep = self._pha_quants[quants[0]] + 16 * self._pha_quants[quants[1]] + 0x2000
            _lib_().spec_report(self._ptr, rep, nxa, C.cast(rng, C.POINTER(C.c_float * 2)))
        else:
            _lib_().spec_report(self._ptr, rep, None, None)

    def mark_edited(self):
        """Upload the host particle buffer before the next iteration."""
        _lib_().spec_mark_host_dirty(self._ptr)

    @property
    def particles(self) -> np.ndarray:
        """Structured array (ix, iy, x, y, ux, uy, uz) of the particle buffer.

        Writes to the returned array are uploaded before the next iteration.
        """
        self._sync()
        return _watched(A.particles_view(self._s), self._sync, self.mark_edited)

    def charge(self) -> np.ndarray:
        """Charge density on the grid, [ny, nx] (guard cells dropped)."""
        s = self._s
        rho = np.zeros((s.nx[1] + 1, s.nx[0] + 1), dtype=np.float32)
        _lib_().spec_deposit_charge(self._ptr, rho.ctypes.data_as(C.POINTER(C.c_float)))
        return rho[0:s.nx[1], 0:s.nx[0]]

    def phasespace(self, quants, pha_nx, pha_range) -> np.ndarray:
        """Phase-space density of two of x1, x2, u1, u2, u3, [pha_nx[1], pha_nx[0]]."""
        rep = self._pha_quants[quants[0]] + 16 * self._pha_quants[quants[1]] + 0x2000
        nxa = (C.c_int * 2)(*[int(v) for v in pha_nx])
        rng = ((C.c_float * 2) * 2)()
        for a in range(2):
            for b in range(2):
                rng[a][b] = pha_range[a][b]
        pha = np.zeros((int(pha_nx[1]), int(pha_nx[0])), dtype=np.float32)
        _lib_().spec_deposit_pha(self._ptr, rep, nxa, C.cast(rng, C.POINTER(C.c_float * 2)),
                                 pha.ctypes.data_as(C.POINTER(C.c_float)))
        return pha

    @property
    def energy(self) -> float:
        """Time-centred kinetic energy of the last push (0 before the first)."""
        self._sync()
        return self._s.energy

    @property
    def dx(self):
        return np.array(self._s.dx[:], dtype=np.float32)

    @property
    def dt(self):
        return self._s.dt

    @property
    def iter(self):
        return self._s.iter

    @property
    def ppc(self):
        return np.array(self._s.ppc[:], dtype=np.int32)

    @property
    def n_sort(self):
        return self._s.n_sort if self._ptr is not None else self._n_sort

    @n_sort.setter
    def n_sort(self, value: int):
        if value < 0:
            print("(*error*) Invalid value for n_sort, must be >= 0.", file=sys.stderr)
            return
        self._n_sort = int(value)
        if self._ptr is not None:
            self._s.n_sort = int(value)

    @property
    def n_move(self):
        return self._s.n_move


class _FieldSpec:
    _types = A.FLD_TYPE
    _names = {v: k for k, v in A.FLD_TYPE.items()}

    def _setup(self, struct, E_type, B_type, E_0, B_0, E_custom, B_custom):
        self._f = struct
        self._f.E_type = self._types[E_type]
        self._f.B_type = self._types[B_type]
        self._f.E_0 = A.float3(*np.asarray(E_0, dtype=np.float32))
        self._f.B_0 = A.float3(*np.asarray(B_0, dtype=np.float32))
        self.custom_func_E = E_custom
        self.custom_func_B = B_custom
        self._keep = []
        if E_custom is not None:
            self._f.E_custom, self._f.E_custom_data = A.field_callback(_lib_(), E_custom, self._keep)
        if B_custom is not None:
            self._f.B_custom, self._f.B_custom_data = A.field_callback(_lib_(), B_custom, self._keep)

    @property
    def E_type(self):
        return self._names[self._f.E_type]

    @property
    def B_type(self):
        return self._names[self._f.B_type]

    @property
    def E_0(self):
        return np.array([self._f.E_0.x, self._f.E_0.y, self._f.E_0.z], dtype=np.float32)

    @property
    def B_0(self):
        return np.array([self._f.B_0.x, self._f.B_0.y, self._f.B_0.z], dtype=np.float32)


class ExternalField(_FieldSpec):
    """ExternalField(E_type='none', B_type='none', E_0=[0,0,0], B_0=[0,0,0], E_custom=None, B_custom=None)

    External fields added to the fields seen by particles (em2d.pyx:481-668).
    Custom functions take (ix, dx, iy, dy) and must not depend on time.
    """

    def __init__(self, *, E_type="none", B_type="none", E_0=(0.0, 0.0, 0.0), B_0=(0.0, 0.0, 0.0),
                 E_custom=None, B_custom=None):
        self._setup(A.t_emf_ext_fld(), E_type, B_type, E_0, B_0, E_custom, B_custom)


class InitialField(_FieldSpec):
    """InitialField(E_type='none', B_type='none', E_0=[0,0,0], B_0=[0,0,0], E_custom=None, B_custom=None)

    Initial E and B fields (em2d.pyx:690-877).
    """

    def __init__(self, *, E_type="none", B_type="none", E_0=(0.0, 0.0, 0.0), B_0=(0.0, 0.0, 0.0),
                 E_custom=None, B_custom=None):
        self._setup(A.t_emf_init_fld(), E_type, B_type, E_0, B_0, E_custom, B_custom)


class Laser:
    """Laser(type='plane', start=0, fwhm=0, rise=0, flat=0, fall=0, a0=0, omega0=0,
    polarization=0, W0=0, focus=0, axis=0) - laser pulse parameters (em2d.pyx:1297-1533)."""

    def __init__(self, *, type="plane", start=0.0, fwhm=0.0, rise=0.0, flat=0.0, fall=0.0, a0=0.0, omega0=0.0,
                 polarization=0.0, W0=0.0, focus=0.0, axis=0.0):
        self._l = A.t_emf_laser()
        self._l.type = A.LASER_TYPE[type]
        for k, v in dict(start=start, fwhm=fwhm, rise=rise, flat=flat, fall=fall, a0=a0, omega0=omega0,
                         polarization=polarization, W0=W0, focus=focus, axis=axis).items():
            setattr(self._l, k, v)

    def __getattr__(self, name):
        if name in ("start", "fwhm", "rise", "flat", "fall", "a0", "omega0", "polarization", "W0", "focus",
                    "axis"):
            return getattr(self._l, name)
        raise AttributeError(name)


class Smooth:
    """Smooth(xtype='none', ytype='none', xlevel=0, ylevel=0) - current filter (em2d.pyx:1622-1724).

    The `ylevel` getter returns xlevel, as the reference's does (em2d.pyx:1720).
    """

    _names = {v: k for k, v in A.SMOOTH.items()}

    def __init__(self, *, xtype="none", ytype="none", xlevel=0, ylevel=0):
        self._s = A.t_smooth(A.SMOOTH[xtype], A.SMOOTH[ytype], int(xlevel), int(ylevel))

    @property
    def xtype(self):
        return self._names[self._s.xtype]

    @xtype.setter
    def xtype(self, v):
        self._s.xtype = A.SMOOTH[v]

    @property
    def ytype(self):
        return self._names[self._s.ytype]

    @ytype.setter
    def ytype(self, v):
        self._s.ytype = A.SMOOTH[v]

    @property
    def xlevel(self):
        return self._s.xlevel

    @xlevel.setter
    def xlevel(self, v):
        self._s.xlevel = v

    @property
    def ylevel(self):
        return self._s.xlevel

    @ylevel.setter
    def ylevel(self, v):
        self._s.ylevel = v


class EMF:
    """EM fields of a Simulation (em2d.pyx:879-1294)."""

    _diag_types = {"E": 0, "B": 1, "Ep": 2, "Bp": 3}

    def __init__(self, ptr):
        self._ptr = ptr

    @property
    def _e(self):
        return self._ptr.contents

    def report(self, type: str, fc: int):
        _lib_().emf_report(self._ptr, self._diag_types[type], fc)

    def set_ext_fld(self, ext: ExternalField):
        if self._e.iter == 0:
            _lib_().emf_set_ext_fld(self._ptr, C.byref(ext._f))
        else:
            print("set_ext_fld can only be called before the simulation starts")

    def init_fld(self, init: InitialField):
        if self._e.iter == 0:
            _lib_().emf_init_fld(self._ptr, C.byref(init._f))
        else:
            print("init_fld can only be called before the simulation starts")

    @property
    def nx(self):
        return np.array(self._e.nx[:], dtype=np.int32)

    @property
    def dx(self):
        return np.array(self._e.dx[:], dtype=np.float32)

    @property
    def box(self):
        return np.array(self._e.box[:], dtype=np.float32)

    @property
    def n_move(self):
        return self._e.n_move

    @property
    def energy(self):
        e = (C.c_double * 6)()
        _lib_().emf_get_energy(self._ptr, e)
        return np.array(e[:], dtype=np.float64)

    @property
    def solver_type(self):
        return "Yee"

    def mark_edited(self):
        """Upload the host E and B grids before the next iteration."""
        _lib_().emf_mark_host_dirty(self._ptr)

    def _grid(self, which, comp, part=False):
        lib = _lib_()
        lib.emf_sync_host(self._ptr)
        e = self._e
        if part:
            buf = e.ext_fld.E_part_buf if which == "E" else e.ext_fld.B_part_buf
            if not buf:
                buf = e.E_buf if which == "E" else e.B_buf
        else:
            buf = e.E_buf if which == "E" else e.B_buf
        view = _inner(buf, (e.nx[0], e.nx[1]), comp)
        if part and (e.ext_fld.E_part_buf if which == "E" else e.ext_fld.B_part_buf):
            return view  # fields plus external fields: derived, not state
        return _watched(view, lambda: lib.emf_sync_host(self._ptr), self.mark_edited)

    Ex = property(lambda self: self._grid("E", 0))
    Ey = property(lambda self: self._grid("E", 1))
    Ez = property(lambda self: self._grid("E", 2))
    Bx = property(lambda self: self._grid("B", 0))
    By = property(lambda self: self._grid("B", 1))
    Bz = property(lambda self: self._grid("B", 2))
    Ex_part = property(lambda self: self._grid("E", 0, True))
    Ey_part = property(lambda self: self._grid("E", 1, True))
    Ez_part = property(lambda self: self._grid("E", 2, True))
    Bx_part = property(lambda self: self._grid("B", 0, True))
    By_part = property(lambda self: self._grid("B", 1, True))
    Bz_part = property(lambda self: self._grid("B", 2, True))


class Current:
    """Electric current density of a Simulation (em2d.pyx:1535-1620)."""

    def __init__(self, ptr):
        self._ptr = ptr

    def report(self, jc: int):
        _lib_().current_report(self._ptr, jc)

    def _grid(self, comp):
        lib = _lib_()
        lib.current_sync_host(self._ptr)
        c = self._ptr.contents
        return _inner(c.J_buf, (c.nx[0], c.nx[1]), comp)

    Jx = property(lambda self: self._grid(0))
    Jy = property(lambda self: self._grid(1))
    Jz = property(lambda self: self._grid(2))


class Simulation:
    """Simulation(nx, box, dt, species=None, report=None, mov_window=False, smooth=None,
    init_fld=None, ext_fld=None, open_y=False) - em2d simulation (em2d.pyx:1726-2074).

    open_y=True (no counterpart in the reference) makes the y boundaries
    open: Mur absorbing fields and removal of particles leaving through y
    (zpic_sim_set_open_y)."""

    def __init__(self, nx, box, dt, *, species=None, report=None, mov_window=False, smooth: Smooth | None = None,
                 init_fld: InitialField | None = None, ext_fld: ExternalField | None = None, open_y: bool = False):
        lib = _lib_()
        self._sim = _ABI.t_simulation()
        # seeds of a new C run (em2d.pyx:1781, em2d/random.c:16-17)
        lib.set_rand_seed(12345, 67890)
        if isinstance(species, Species):
            species = [species]
        self._species = list(species) if species else []
        ns = len(self._species)
        nx = [int(v) for v in nx]
        box = [float(v) for v in box]
        arr = None
        if ns:
            raw = _libc.calloc(ns, C.sizeof(_ABI.t_species))
            arr = C.cast(raw, C.POINTER(_ABI.t_species))
            for k, s in enumerate(self._species):
                s._new(C.cast(C.addressof(arr[k]), C.POINTER(_ABI.t_species)), nx, box, dt)
        self._report = report
        lib.sim_new(C.byref(self._sim), (C.c_int * 2)(*nx), (C.c_float * 2)(*box), C.c_float(dt), C.c_float(0.0),
                    0, arr, ns)
        self._n = 0
        self._t = 0.0
        self._emf = EMF(C.cast(C.addressof(self._sim.emf), C.POINTER(_ABI.t_emf)))
        self._current = Current(C.cast(C.addressof(self._sim.current), C.POINTER(_ABI.t_current)))
        self._keep = [smooth, init_fld, ext_fld]
        if mov_window:
            lib.sim_set_moving_window(C.byref(self._sim))
        if open_y:
            lib.zpic_sim_set_open_y(C.byref(self._sim))
        if smooth is not None:
            lib.sim_set_smooth(C.byref(self._sim), C.byref(smooth._s))
        if init_fld is not None:
            lib.emf_init_fld(C.byref(self._sim.emf), C.byref(init_fld._f))
        if ext_fld is not None:
            lib.emf_set_ext_fld(C.byref(self._sim.emf), C.byref(ext_fld._f))
        self._alive = True

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_alive", False):
            _lib_().sim_delete(C.byref(self._sim))
            self._alive = False
            for s in self._species:
                s._ptr = None

    def set_moving_window(self):
        _lib_().sim_set_moving_window(C.byref(self._sim))

    def set_open_y(self):
        """Open (absorbing) y boundaries, before the first iteration."""
        _lib_().zpic_sim_set_open_y(C.byref(self._sim))

    def set_smooth(self, smooth: Smooth):
        self._keep.append(smooth)
        _lib_().sim_set_smooth(C.byref(self._sim), C.byref(smooth._s))

    def add_laser(self, laser: Laser):
        _lib_().sim_add_laser(C.byref(self._sim), C.byref(laser._l))

    def iter(self):
        """Advance one iteration."""
        _lib_().sim_iter(C.byref(self._sim))
        self._n += 1
        self._t = self._n * self._sim.dt

    def run(self, tmax: float):
        """Advance while t <= tmax, calling `report` before each iteration."""
        if tmax < self._t:
            print("Simulation is already at t = {:g}".format(self._t))
            return
        print("\nRunning simulation up to t = {:g} ...".format(tmax))
        lib = _lib_()
        while self._t <= tmax:
            print("n = {:d}, t = {:g}".format(self._n, self._t), end="\r")
            if self._report:
                self._report(self)
            lib.sim_iter(C.byref(self._sim))
            self._n += 1
            self._t = self._n * self._sim.dt
        print("n = {:d}, t = {:g}".format(self._n, self._t), end="\r")
        print("\nDone.")

    def synchronize(self):
        """Wait for all queued GPU work of this simulation."""
        _lib_().zpic_synchronize()

    @property
    def species(self):
        return self._species

    @property
    def emf(self):
        return self._emf

    @property
    def current(self):
        return self._current

    @property
    def n(self):
        return self._n

    @property
    def t(self):
        return self._t

    @property
    def dt(self):
        return self._sim.dt

    @property
    def dx(self):
        return self._emf.dx

    @property
    def nx(self):
        return self._emf.nx

    @property
    def box(self):
        return self._emf.box


class Batch:
    """Independent simulations advanced together on the GPU (parameter
    scans).  No counterpart in the reference module; the reference runs a
    scan's simulations one after another (mods/em1d_param_scan/main.c:50-98).

    `run(steps)` leaves every simulation exactly where `steps` calls of its
    own `iter()` would; each step of all simulations is one replay of a
    captured HIP graph.  Simulations must be periodic, without moving
    window or external fields; their grid, species and filter must not
    change while the batch exists.
    """

    def __init__(self, sims):
        self._sims = list(sims)
        if not self._sims:
            raise ValueError("Batch needs at least one simulation")
        lib = _lib_()
        arr = (C.c_void_p * len(self._sims))(*[C.addressof(s._sim) for s in self._sims])
        self._b = lib.zpic_batch_new(arr, len(self._sims))

    def run(self, steps: int):
        """Advance every simulation by `steps` iterations."""
        if self._b is None:
            raise RuntimeError("batch is closed")
        _lib_().zpic_batch_run(self._b, int(steps))
        for s in self._sims:
            s._n += int(steps)
            s._t = s._n * s._sim.dt

    def close(self):
        if self._b is not None:
            _lib_().zpic_batch_delete(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def simulations(self):
        return self._sims
