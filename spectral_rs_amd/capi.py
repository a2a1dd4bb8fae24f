"""ctypes binding of the C ABI in include/spectral_rs.h (libspectral_rs_hip.so).

``GpuSim`` is a thin object over one ``sp_ctx``: one method per entry point,
numpy arrays in and out, errors raised as ``SpError`` with the library's
message.  There is no CPU fallback: if the HIP library is not built, or no
GPU is visible, constructing a ``GpuSim`` raises.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libspectral_rs_hip.so")

SP_OK, SP_EINVAL, SP_EHIP, SP_EINVARIANT, SP_ENOMEM = 0, -1, -2, -3, -4
SP_F32, SP_F64 = 0, 1
SP_LEFT_ASSERT, SP_LEFT_REFLECT = 0, 1
SP_STUDY_ABS_POS32, SP_STUDY_TWIDDLE32, SP_STUDY_SUM32 = 1, 2, 4
GRIDS = {"ex": 0, "ey": 1, "ez": 2, "bx": 3, "by": 4, "bz": 5, "jx": 6, "jy": 7, "jz": 8,
         "rho": 9, "dsty": 9, "dens": 10, "bxw": 11, "byw": 12, "bzw": 13}

# exported symbols, in the order include/spectral_rs.h declares them
SYMBOLS = ["sp_create", "sp_destroy", "sp_last_error", "sp_species_add", "sp_species_init",
           "sp_species_count", "sp_get_species", "sp_get_species_raw", "sp_get_grid", "sp_set_grid", "sp_get_spectrum",
           "sp_set_spectrum", "sp_zero_currents", "sp_move_and_deposit", "sp_deposit_ghosts",
           "sp_fields_update", "sp_boris_push", "sp_filter", "sp_filter_noalloc", "sp_fft2d", "sp_spectral_update",
           "sp_prestep", "sp_step", "sp_sync", "sp_set_timing", "sp_get_timers", "sp_set_graph", "sp_set_sort_policy",
           "sp_get_diag", "sp_checkpoint_save", "sp_checkpoint_load", "sp_set_tuning", "sp_get_tuning", "sp_autotune", "sp_ensemble_create",
           "sp_ensemble_destroy", "sp_ensemble_last_error", "sp_ensemble_set_graph", "sp_ensemble_step",
           "sp_ensemble_sync", "sp_ensemble_stats", "sp_dn_create", "sp_dn_destroy", "sp_dn_last_error", "sp_dn_set_params",
           "sp_dn_get_params", "sp_dn_set_loss_weights", "sp_dn_forward", "sp_dn_gradient", "sp_dn_train",
           "sp_set_denoiser", "sp_set_denoiser_input", "sp_set_denoiser_channels",
           "sp_get_denoiser_stats", "sp_set_denoiser_backoff", "sp_get_denoiser_skipped"]
DN_NTRAIN, DN_NPARAMS = 30084, 30088
DN_LAYERS, DN_CH = 5, 32  # the denoiser: 5 convolutions, 32 hidden channels


class SpParams(ct.Structure):
    _fields_ = [("size_x", ct.c_int32), ("size_y", ct.c_int32), ("delta", ct.c_int32),
                ("dens", ct.c_int32), ("n_pass", ct.c_int32), ("dt", ct.c_double), ("c", ct.c_double),
                ("gamma_inj", ct.c_double), ("precision", ct.c_int32), ("periodic_x", ct.c_int32),
                ("checked", ct.c_int32), ("left_edge", ct.c_int32), ("sort_interval", ct.c_int32),
                ("study", ct.c_int32), ("reserved", ct.c_int32 * 2)]


class SpDiag(ct.Structure):
    _fields_ = [("energy_e", ct.c_double), ("energy_b", ct.c_double), ("energy_kin", ct.c_double * 8),
                ("momentum", ct.c_double * 3), ("gauss_pre", ct.c_double), ("gauss_post", ct.c_double),
                ("gauss_post_max", ct.c_double), ("steps", ct.c_int64),
                ("sorts", ct.c_int64)]


class SpTuning(ct.Structure):
    _fields_ = [("chunk", ct.c_int32), ("tile_w_log", ct.c_int32), ("tile_h_log", ct.c_int32),
                ("fft_rows", ct.c_int32), ("fft_cols", ct.c_int32), ("fft_radix_x", ct.c_int32),
                ("fft_radix_y", ct.c_int32), ("reserved", ct.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run python -m spectral_rs_amd.build)")
    L = ct.CDLL(path)
    vp, i32, i64, d = ct.c_void_p, ct.c_int, ct.c_int64, ct.c_double
    L.sp_create.argtypes = [ct.POINTER(SpParams), ct.POINTER(vp)]
    L.sp_destroy.argtypes = [vp]
    L.sp_last_error.restype = ct.c_char_p
    L.sp_last_error.argtypes = [vp]
    L.sp_species_add.argtypes = [vp, d, d, i64] + [vp] * 7 + [ct.POINTER(i32)]
    L.sp_species_init.argtypes = [vp, d, d, d, ct.c_uint64, ct.POINTER(i32)]
    L.sp_species_count.argtypes = [vp, i32, ct.POINTER(i64)]
    L.sp_get_species.argtypes = [vp, i32, i64] + [vp] * 8
    L.sp_get_species_raw.argtypes = [vp, i32, i64] + [vp] * 7
    L.sp_checkpoint_save.argtypes = [vp, ct.c_char_p]
    L.sp_checkpoint_load.argtypes = [vp, ct.c_char_p]
    for n in ["sp_get_grid", "sp_set_grid", "sp_get_spectrum", "sp_set_spectrum"]:
        getattr(L, n).argtypes = [vp, i32, vp, ct.c_size_t]
    for n in ["sp_zero_currents", "sp_deposit_ghosts", "sp_fields_update", "sp_spectral_update",
              "sp_prestep", "sp_sync"]:
        getattr(L, n).argtypes = [vp]
    for n in ["sp_move_and_deposit", "sp_boris_push", "sp_step", "sp_set_timing", "sp_set_graph"]:
        getattr(L, n).argtypes = [vp, i32]
    L.sp_filter.argtypes = [vp, i32, i32]
    L.sp_set_sort_policy.argtypes = [vp, i32, i32, d]
    L.sp_filter_noalloc.argtypes = [vp, ct.c_int32, ct.c_int32, i32, i32, vp]
    L.sp_fft2d.argtypes = [vp, i32, i32]
    L.sp_get_timers.argtypes = [vp, vp, i32]
    L.sp_get_diag.argtypes = [vp, ct.POINTER(SpDiag)]
    L.sp_dn_create.argtypes = [i32, i32, i32, ct.POINTER(vp)]
    L.sp_dn_destroy.argtypes = [vp]
    L.sp_dn_last_error.argtypes = [vp]
    L.sp_dn_last_error.restype = ct.c_char_p
    L.sp_dn_set_params.argtypes = [vp, vp, i32]
    L.sp_dn_get_params.argtypes = [vp, vp, i32]
    L.sp_dn_set_loss_weights.argtypes = [vp, vp]
    L.sp_dn_forward.argtypes = [vp, vp, vp, i32]
    L.sp_dn_gradient.argtypes = [vp, vp, vp, vp, ct.POINTER(ct.c_double)]
    L.sp_dn_train.argtypes = [vp, vp, vp, i32, i32, ct.c_float, ct.c_uint64, vp]
    L.sp_set_denoiser.argtypes = [vp, vp, i32, ct.c_float, ct.c_float]
    L.sp_get_denoiser_stats.argtypes = [vp, ct.POINTER(i32), ct.POINTER(i32)]
    L.sp_set_denoiser_channels.argtypes = [vp, i32]
    L.sp_set_denoiser_input.argtypes = [vp, i32, ct.c_float]
    L.sp_set_denoiser_backoff.argtypes = [vp, i32]
    L.sp_get_denoiser_skipped.argtypes = [vp, ct.POINTER(i32)]
    L.sp_set_tuning.argtypes = [vp, ct.POINTER(SpTuning)]
    L.sp_get_tuning.argtypes = [vp, ct.POINTER(SpTuning)]
    L.sp_autotune.argtypes = [vp, i32, i32, ct.POINTER(SpTuning)]
    L.sp_ensemble_create.argtypes = [ct.POINTER(vp), i32, i32, i32, ct.POINTER(vp)]
    L.sp_ensemble_destroy.argtypes = [vp]
    L.sp_ensemble_last_error.argtypes = [vp]
    L.sp_ensemble_last_error.restype = ct.c_char_p
    L.sp_ensemble_set_graph.argtypes = [vp, i32]
    L.sp_ensemble_step.argtypes = [vp, i32]
    L.sp_ensemble_sync.argtypes = [vp, ct.POINTER(i32)]
    L.sp_ensemble_stats.argtypes = [vp, ct.POINTER(i32), ct.POINTER(i32)]
    for n in SYMBOLS:
        if not n.endswith("_last_error"):
            getattr(L, n).restype = ct.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


class GpuSim:
    """One simulation context on the current GPU (Sim + Flds + Vec<Prtl>)."""

    def __init__(self, size_x, size_y, delta, dens, n_pass=4, dt=0.1, c=3.0, gamma_inj=5.0,
                 prec=4, periodic_x=False, checked=True, left_edge=SP_LEFT_ASSERT, sort_interval=0, study=0):
        L = load()
        self.nx, self.ny = size_x, size_y
        self.prec = prec
        self.ft = np.float64 if prec == 8 else np.float32
        self.ct = np.complex128 if prec == 8 else np.complex64
        p = SpParams(size_x, size_y, delta, dens, n_pass, dt, c, gamma_inj,
                     SP_F64 if prec == 8 else SP_F32, int(periodic_x), int(checked), int(left_edge),
                     int(sort_interval), int(study))
        h = ct.c_void_p()
        rc = L.sp_create(ct.byref(p), ct.byref(h))
        if rc != SP_OK:
            raise SpError(rc, L.sp_last_error(None).decode())
        self.h = h
        self.nspecies = 0

    def close(self):
        if getattr(self, "h", None):
            if load().sp_destroy(self.h) == SP_OK:  # fails while the context is in an ensemble
                self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != SP_OK:
            raise SpError(rc, load().sp_last_error(self.h).decode())
        return rc

    # ---- species ----
    def add_species(self, charge, mass, ix, iy, dx, dy, px, py, pz, psa=None):
        ft = self.ft
        ix = np.ascontiguousarray(ix, np.int32)
        iy = np.ascontiguousarray(iy, np.int32)
        fl = [np.ascontiguousarray(a, ft) for a in (dx, dy, px, py, pz)]
        sid = ct.c_int(-1)
        self._chk(load().sp_species_add(self.h, charge, mass, ix.size, _p(ix), _p(iy),
                                        *[_p(a) for a in fl], ct.byref(sid)))
        self.nspecies += 1
        return sid.value

    def init_species(self, charge, mass, vth, seed):
        sid = ct.c_int(-1)
        self._chk(load().sp_species_init(self.h, charge, mass, vth, seed, ct.byref(sid)))
        self.nspecies += 1
        return sid.value

    def count(self, sid):
        n = ct.c_int64()
        self._chk(load().sp_species_count(self.h, sid, ct.byref(n)))
        return n.value

    def species(self, sid):
        n = self.count(sid)
        ix = np.empty(n, np.int32)
        iy = np.empty(n, np.int32)
        fl = [np.empty(n, self.ft) for _ in range(6)]
        self._chk(load().sp_get_species(self.h, sid, n, _p(ix), _p(iy), *[_p(a) for a in fl]))
        return dict(ix=ix.astype(np.int64), iy=iy.astype(np.int64), dx=fl[0], dy=fl[1], px=fl[2],
                    py=fl[3], pz=fl[4], psa=fl[5])

    def species_raw(self, sid):
        """The species in device order: packed cells (iy << 16 | ix), tags, offsets, momenta."""
        n = self.count(sid)
        cell = np.empty(n, np.int32)
        tag = np.empty(n, np.int32)
        fl = [np.empty(n, self.ft) for _ in range(5)]
        self._chk(load().sp_get_species_raw(self.h, sid, n, _p(cell), _p(tag), *[_p(a) for a in fl]))
        return dict(cell=cell, tag=tag, ix=(cell & 0xFFFF).astype(np.int64), iy=(cell >> 16).astype(np.int64),
                    dx=fl[0], dy=fl[1], px=fl[2], py=fl[3], pz=fl[4])

    def checkpoint_save(self, path):
        self._chk(load().sp_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path):
        """Into a context made with the same parameters and no species."""
        self._chk(load().sp_checkpoint_load(self.h, os.fsencode(path)))
        n = ct.c_int64()
        k = 0
        while load().sp_species_count(self.h, k, ct.byref(n)) == SP_OK:
            k += 1
        self.nspecies = k

    # ---- grids ----
    def grid(self, name):
        a = np.empty((self.ny + 2, self.nx + 2), self.ft)
        self._chk(load().sp_get_grid(self.h, GRIDS[name], _p(a), a.size))
        return a

    def set_grid(self, name, a):
        a = np.ascontiguousarray(a, self.ft).reshape(self.ny + 2, self.nx + 2)
        self._chk(load().sp_set_grid(self.h, GRIDS[name], _p(a), a.size))

    def spectrum(self, name):
        a = np.empty((self.ny, self.nx), self.ct)
        self._chk(load().sp_get_spectrum(self.h, GRIDS[name], _p(a), 2 * a.size))
        return a

    def set_spectrum(self, name, a):
        a = np.ascontiguousarray(a, self.ct).reshape(self.ny, self.nx)
        self._chk(load().sp_set_spectrum(self.h, GRIDS[name], _p(a), 2 * a.size))

    # ---- reference seams ----
    def zero_currents(self): self._chk(load().sp_zero_currents(self.h))
    def move_and_deposit(self, sid): self._chk(load().sp_move_and_deposit(self.h, sid))
    def fold_all(self): self._chk(load().sp_deposit_ghosts(self.h))
    def fields_update(self): self._chk(load().sp_fields_update(self.h))
    def boris_push(self, sid): self._chk(load().sp_boris_push(self.h, sid))
    def filter(self, name, passes): self._chk(load().sp_filter(self.h, GRIDS[name], passes))
    def fft2d(self, name, inverse=False): self._chk(load().sp_fft2d(self.h, GRIDS[name], int(inverse)))
    def spectral_update(self): self._chk(load().sp_spectral_update(self.h))
    def prestep(self): self._chk(load().sp_prestep(self.h))
    def step(self, n=1): self._chk(load().sp_step(self.h, n))
    def sync(self): self._chk(load().sp_sync(self.h))
    def set_timing(self, on=True): self._chk(load().sp_set_timing(self.h, int(on)))

    def set_graph(self, on=True): self._chk(load().sp_set_graph(self.h, int(on)))

    def set_sort_policy(self, min_interval=0, max_interval=0, max_slow_rise=0.0):
        """Adaptive tile-sort schedule (max_interval 0: every sort_interval steps)."""
        self._chk(load().sp_set_sort_policy(self.h, int(min_interval), int(max_interval), float(max_slow_rise)))

    def timers(self, reset=False):
        t = np.zeros(4, np.float64)
        self._chk(load().sp_get_timers(self.h, _p(t), int(reset)))
        return dict(dep=t[0], solve=t[1], push=t[2], particles=t[3])

    def diag(self):
        """FP64 diagnostics as a dict (energies, momentum, Gauss residuals)."""
        d = SpDiag()
        self._chk(load().sp_get_diag(self.h, ct.byref(d)))
        ns = min(self.nspecies, 8)
        return {"energy_e": d.energy_e, "energy_b": d.energy_b,
                "energy_kin": [d.energy_kin[k] for k in range(ns)],
                "momentum": [d.momentum[k] for k in range(3)],
                "gauss_pre": d.gauss_pre, "gauss_post": d.gauss_post, "gauss_post_max": d.gauss_post_max,
                "steps": d.steps, "sorts": d.sorts}

    def set_denoiser(self, params, max_rel=0.5, max_gauss=1.0):
        """Apply a trained denoiser after the binomial filter (params None disables)."""
        if params is None:
            self._chk(load().sp_set_denoiser(self.h, None, 0, 0.0, 0.0))
            return
        p = np.ascontiguousarray(params, np.float32)
        self._chk(load().sp_set_denoiser(self.h, _p(p), p.size, max_rel, max_gauss))

    def set_denoiser_input(self, raw=False, max_hp=0.0):
        """raw: the network reads the unfiltered deposit (standing in for the
        binomial filter, which stays the fallback); max_hp > 0: spectrum check
        of its output against the filter's (high-pass energy ratio bound)."""
        self._chk(load().sp_set_denoiser_input(self.h, int(raw), float(max_hp)))

    def set_denoiser_channels(self, mask):
        """Bit c: write the network's output to Jx, Jy, Jz, rho (c = 0..3)."""
        self._chk(load().sp_set_denoiser_channels(self.h, int(mask)))

    def denoiser_stats(self):
        a, f = ct.c_int32(), ct.c_int32()
        self._chk(load().sp_get_denoiser_stats(self.h, ct.byref(a), ct.byref(f)))
        return a.value, f.value

    def set_denoiser_backoff(self, solves):
        """Solves that keep the filter's result without inference after a failed check."""
        self._chk(load().sp_set_denoiser_backoff(self.h, int(solves)))

    def denoiser_skipped(self):
        k = ct.c_int32()
        self._chk(load().sp_get_denoiser_skipped(self.h, ct.byref(k)))
        return k.value

    # ---- launch shapes ----
    def tuning(self):
        t = SpTuning()
        self._chk(load().sp_get_tuning(self.h, ct.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        """Change some knobs of sp_tuning (the others keep their values)."""
        t = SpTuning()
        self._chk(load().sp_get_tuning(self.h, ct.byref(t)))
        for k, v in kw.items():
            if k not in t.as_dict():
                raise KeyError(k)
            setattr(t, k, int(v))
        self._chk(load().sp_set_tuning(self.h, ct.byref(t)))

    def autotune(self, trials=3, use_cache=True):
        t = SpTuning()
        self._chk(load().sp_autotune(self.h, trials, int(use_cache), ct.byref(t)))
        return t.as_dict()


class Ensemble:
    """Independent runs stepped together (sp_ensemble_*): the members keep
    their own state and getters; they must outlive the ensemble."""

    def __init__(self, sims, branches=4, unit=2):
        L = load()
        self.sims = list(sims)
        arr = (ct.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        e = ct.c_void_p()
        rc = L.sp_ensemble_create(arr, len(self.sims), branches, unit, ct.byref(e))
        if rc != SP_OK:
            raise SpError(rc, L.sp_ensemble_last_error(None).decode())
        self.e = e

    def close(self):
        if getattr(self, "e", None):
            load().sp_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != SP_OK:
            raise SpError(rc, load().sp_ensemble_last_error(self.e).decode())

    def set_graph(self, on=True): self._chk(load().sp_ensemble_set_graph(self.e, int(on)))
    def step(self, n=1): self._chk(load().sp_ensemble_step(self.e, n))

    def sync(self):
        bad = ct.c_int32(-1)
        self._chk(load().sp_ensemble_sync(self.e, ct.byref(bad)))

    def stats(self):
        k, c = ct.c_int32(), ct.c_int32()
        self._chk(load().sp_ensemble_stats(self.e, ct.byref(k), ct.byref(c)))
        return dict(kept=k.value, captures=c.value)


class Denoiser:
    """Training context of the learned current denoiser (sp_dn_*): batches of
    `batch` images of h x w with 4 planes (Jx Jy Jz rho)."""

    def __init__(self, batch, h, w):
        L = load()
        self.batch, self.h, self.w = batch, h, w
        d = ct.c_void_p()
        rc = L.sp_dn_create(batch, h, w, ct.byref(d))
        if rc != SP_OK:
            raise SpError(rc, L.sp_dn_last_error(None).decode())
        self.d = d

    def close(self):
        if getattr(self, "d", None):
            load().sp_dn_destroy(self.d)
            self.d = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != SP_OK:
            raise SpError(rc, load().sp_dn_last_error(self.d).decode())

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._chk(load().sp_dn_set_params(self.d, _p(p), p.size))

    def set_loss_weights(self, w4):
        w = np.ascontiguousarray(w4, np.float32)
        self._chk(load().sp_dn_set_loss_weights(self.d, _p(w)))

    def params(self):
        p = np.zeros(DN_NPARAMS, np.float32)
        self._chk(load().sp_dn_get_params(self.d, _p(p), p.size))
        return p

    def forward(self, x, mode=0):
        """x + r(x); mode 0 zero boundary, 1 periodic rows, 2 periodic rows (fused kernel)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty_like(x)
        self._chk(load().sp_dn_forward(self.d, _p(x), _p(y), int(mode)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(DN_NTRAIN, np.float32)
        loss = ct.c_double()
        self._chk(load().sp_dn_gradient(self.d, _p(x), _p(t), _p(g), ct.byref(loss)))
        return loss.value, g

    def train(self, xs, ts, iters, lr, seed=0):
        """Adam on random batches of the pairs (xs, ts) [n][4][h][w]; xs None reuses the last upload."""
        hist = np.zeros(iters, np.float64)
        if xs is None:
            self._chk(load().sp_dn_train(self.d, None, None, 0, iters, lr, seed, _p(hist)))
        else:
            xs = np.ascontiguousarray(xs, np.float32)
            ts = np.ascontiguousarray(ts, np.float32)
            self._chk(load().sp_dn_train(self.d, _p(xs), _p(ts), xs.shape[0], iters, lr, seed, _p(hist)))
        return hist
