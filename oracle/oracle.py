"""ctypes binding of the CPU oracle (oracle/cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to check
its result, and by bench.py's cpu_baseline leg.  The product package never
imports it.

``RefSim`` mirrors the reference's Sim + Flds + Vec<Prtl> (src/lib.rs:204-264,
src/flds/mod.rs:15-65, src/prtls.rs:8-23); ``prec=4`` is the default f32 build
and ``prec=8`` is ``--features dprec``.
"""
import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libcpuref.so")

GRIDS = {"ex": 0, "ey": 1, "ez": 2, "bx": 3, "by": 4, "bz": 5, "jx": 6, "jy": 7, "jz": 8,
         "dsty": 9, "rho": 9, "dens": 10, "bxw": 11, "byw": 12, "bzw": 13}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        vp, i64, i32, d = ct.c_void_p, ct.c_int64, ct.c_int32, ct.c_double
        L.ref_create.restype = vp
        L.ref_create.argtypes = [ct.c_int, i64, i64, i64, i32, i32, d, d, d, i32, i32, i32, i32]
        L.ref_destroy.argtypes = [vp]
        L.ref_last_error.restype = ct.c_char_p
        L.ref_last_error.argtypes = [vp]
        L.ref_set_threads.argtypes = [ct.c_int]
        L.ref_species_add.argtypes = [vp, d, d, i64] + [vp] * 8
        L.ref_get_species.argtypes = [vp, ct.c_int] + [vp] * 8
        L.ref_species_count.restype = i64
        L.ref_species_count.argtypes = [vp, ct.c_int]
        for name in ["ref_get_grid", "ref_set_grid", "ref_get_spectrum", "ref_set_spectrum"]:
            getattr(L, name).argtypes = [vp, ct.c_int, vp]
        L.ref_wave_numbers.argtypes = [vp, vp, vp, vp]
        for name in ["ref_update_position", "ref_apply_bc", "ref_deposit_current", "ref_calc_density",
                     "ref_move_and_deposit", "ref_boris_push", "ref_deposit_ghosts", "ref_update_ghosts",
                     "ref_binomial_pass", "ref_copy_to_spectral", "ref_copy_to_spatial", "ref_step"]:
            getattr(L, name).argtypes = [vp, ct.c_int]
        for name in ["ref_fold_all", "ref_zero_currents", "ref_spectral_update", "ref_fields_update",
                     "ref_prestep"]:
            getattr(L, name).argtypes = [vp]
        L.ref_fft2d.argtypes = [vp, ct.c_int, ct.c_int]
        L.ref_timers.argtypes = [vp, vp]
        L.ref_fft1d.argtypes = [ct.c_int, i64, vp, ct.c_int]
        L.ref_init_positions.restype = i64
        L.ref_init_positions.argtypes = [ct.c_int, i64, i64, i64, i32, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def set_threads(n):
    lib().ref_set_threads(int(n))


def fft1d(x, inverse=False, prec=8):
    """Unnormalised 1D DFT (numpy sign convention) of a complex vector."""
    dt = np.complex128 if prec == 8 else np.complex64
    a = np.ascontiguousarray(x, dtype=dt).copy()
    lib().ref_fft1d(prec, a.size, _p(a), 1 if inverse else 0)
    return a


def init_positions(size_x, size_y, delta, dens, prec=4):
    """src/prtls.rs:191-221: ghost-offset cells and the diagonal offsets."""
    n = dens * (size_x - 2 * delta) * size_y
    ft = np.float64 if prec == 8 else np.float32
    ix = np.zeros(n, np.int64)
    iy = np.zeros(n, np.int64)
    dx = np.zeros(n, ft)
    dy = np.zeros(n, ft)
    lib().ref_init_positions(prec, size_x, size_y, delta, dens, _p(ix), _p(iy), _p(dx), _p(dy))
    return ix, iy, dx, dy


class OracleError(RuntimeError):
    pass


class RefSim:
    def __init__(self, size_x, size_y, delta, dens, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0,
                 prec=4, periodic_x=False, checked=True, left_edge=0, ref_quirks=True):
        self.prec = prec
        self.nx, self.ny = size_x, size_y
        self.ft = np.float64 if prec == 8 else np.float32
        self.ct = np.complex128 if prec == 8 else np.complex64
        self.h = lib().ref_create(prec, size_x, size_y, delta, dens, n_pass, dt, c, gamma_inj,
                                  int(periodic_x), int(checked), int(left_edge), int(ref_quirks))
        self.nspecies = 0

    def close(self):
        if self.h:
            lib().ref_destroy(self.h)
            self.h = None

    __del__ = close

    def _chk(self, rc):
        if rc < 0:
            raise OracleError(lib().ref_last_error(self.h).decode())
        return rc

    def add_species(self, charge, mass, ix, iy, dx, dy, px, py, pz, psa=None):
        ft = self.ft
        arrs = [np.ascontiguousarray(ix, np.int64), np.ascontiguousarray(iy, np.int64)]
        arrs += [np.ascontiguousarray(a, ft) for a in (dx, dy, px, py, pz)]
        ps = None if psa is None else np.ascontiguousarray(psa, ft)
        rc = lib().ref_species_add(self.h, charge, mass, arrs[0].size, *[_p(a) for a in arrs],
                                   None if ps is None else _p(ps))
        self._chk(rc)
        self.nspecies += 1
        return rc

    def species(self, sid):
        n = lib().ref_species_count(self.h, sid)
        ix = np.empty(n, np.int64)
        iy = np.empty(n, np.int64)
        fl = [np.empty(n, self.ft) for _ in range(6)]
        self._chk(lib().ref_get_species(self.h, sid, _p(ix), _p(iy), *[_p(a) for a in fl]))
        return dict(ix=ix, iy=iy, dx=fl[0], dy=fl[1], px=fl[2], py=fl[3], pz=fl[4], psa=fl[5])

    def grid(self, name):
        a = np.empty((self.ny + 2, self.nx + 2), self.ft)
        self._chk(lib().ref_get_grid(self.h, GRIDS[name], _p(a)))
        return a

    def set_grid(self, name, a):
        a = np.ascontiguousarray(a, self.ft).reshape(self.ny + 2, self.nx + 2)
        self._chk(lib().ref_set_grid(self.h, GRIDS[name], _p(a)))

    def spectrum(self, name):
        a = np.empty((self.ny, self.nx), self.ct)
        self._chk(lib().ref_get_spectrum(self.h, GRIDS[name], _p(a)))
        return a

    def set_spectrum(self, name, a):
        a = np.ascontiguousarray(a, self.ct).reshape(self.ny, self.nx)
        self._chk(lib().ref_set_spectrum(self.h, GRIDS[name], _p(a)))

    def wave_numbers(self):
        k = [np.empty((self.ny, self.nx), self.ft) for _ in range(3)]
        self._chk(lib().ref_wave_numbers(self.h, *[_p(a) for a in k]))
        return k

    def timers(self):
        t = np.zeros(3, np.float64)
        self._chk(lib().ref_timers(self.h, _p(t)))
        return dict(dep=t[0], solve=t[1], push=t[2])

    # one call per reference method
    def update_position(self, sid): self._chk(lib().ref_update_position(self.h, sid))
    def apply_bc(self, sid): self._chk(lib().ref_apply_bc(self.h, sid))
    def deposit_current(self, sid): self._chk(lib().ref_deposit_current(self.h, sid))
    def calc_density(self, sid): self._chk(lib().ref_calc_density(self.h, sid))
    def move_and_deposit(self, sid): self._chk(lib().ref_move_and_deposit(self.h, sid))
    def boris_push(self, sid): self._chk(lib().ref_boris_push(self.h, sid))
    def deposit_ghosts(self, name): self._chk(lib().ref_deposit_ghosts(self.h, GRIDS[name]))
    def update_ghosts(self, name): self._chk(lib().ref_update_ghosts(self.h, GRIDS[name]))
    def binomial_pass(self, name): self._chk(lib().ref_binomial_pass(self.h, GRIDS[name]))
    def copy_to_spectral(self, name): self._chk(lib().ref_copy_to_spectral(self.h, GRIDS[name]))
    def copy_to_spatial(self, name): self._chk(lib().ref_copy_to_spatial(self.h, GRIDS[name]))
    def fft2d(self, name, inverse=False): self._chk(lib().ref_fft2d(self.h, GRIDS[name], int(inverse)))
    def fold_all(self): self._chk(lib().ref_fold_all(self.h))
    def zero_currents(self): self._chk(lib().ref_zero_currents(self.h))
    def spectral_update(self): self._chk(lib().ref_spectral_update(self.h))
    def fields_update(self): self._chk(lib().ref_fields_update(self.h))
    def prestep(self): self._chk(lib().ref_prestep(self.h))
    def step(self, n=1): self._chk(lib().ref_step(self.h, n))
