"""ctypes binding of libdnsgpu.so (include/dnsgpu.h) for tests and benchmarks.

The library is the product; this module only marshals arguments.  It never
falls back to anything else: if libdnsgpu.so is missing or the GPU is absent
the calls fail loudly.

Array layouts follow the reference (DNS-solver/src/definitions.h:93):
checkpoint halves are float64 [xdv][ydv][zdv+2][19].
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdnsgpu.so")

FP64, FP32, FP16 = 0, 1, 2
CONST_Q, CONST_F = 0, 1
INIT_REST, INIT_POISEUILLE, INIT_TURBULENT = 0, 1, 2
ENAN = -2

EXPORTS = (
    "dnsgpu_create", "dnsgpu_destroy", "dnsgpu_last_error", "dnsgpu_set_walls",
    "dnsgpu_set_solid_bands", "dnsgpu_init_rest", "dnsgpu_init_field", "dnsgpu_load_aos",
    "dnsgpu_commit_load", "dnsgpu_save_aos", "dnsgpu_get_force", "dnsgpu_step",
    "dnsgpu_statistics", "dnsgpu_fields", "dnsgpu_time_steps", "dnsgpu_device_bytes",
    "dnsgpu_autotune", "dnsgpu_tavg_reset", "dnsgpu_tavg_sample", "dnsgpu_tavg_get",
    "dnsgpu_footprint", "dnsgpu_fields_noalloc", "dnsgpu_step_shape", "dnsgpu_set_step_shape",
    "dnsgpu_step_pieces", "dnsgpu_set_step_pieces", "dnsgpu_set_force",
    "dnsgpu_pair_placement", "dnsgpu_reserve_pair_arenas", "dnsgpu_release_pair_arenas", "dnsgpu_pair_arena_stats",
    "dnsens_create", "dnsens_destroy", "dnsens_member_bytes", "dnsens_footprint", "dnsens_add", "dnsens_remove",
    "dnsens_count", "dnsens_bytes_used", "dnsens_budget", "dnsens_bind_thread", "dnsens_step", "dnsens_info", "dnsens_rows",
    "dnsens_autotune", "dnsens_time_steps", "dnsens_time_kernels", "dnsens_grid",
)

ENS_NO_GRAPH, ENS_AUTOTUNE = 1, 2

TAVG_FIELDS = ("nsteps", "nsamples", "pgrad", "utau", "stress", "stup", "stdown", "Ub", "K", "rho", "Cf_w",
               "Cf_p", "Re_tau", "u_slip_bottom", "u_noslip_bottom", "u_slip_top", "u_noslip_top")


class DnsTavg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in TAVG_FIELDS]


def drag_reduction(base, textured, forcing=CONST_Q):
    """Drag reduction in percent of a textured run against a smooth base run
    (dicts from Channel.tavg()): (Cf0 - Cf)/Cf0 * 100 under constant flow
    rate (TKE-budget/postprocessing/cfriction.c:46,77, from the wall stress
    and from the pressure gradient), 1 - (Ub0/Ub)^2 under constant force."""
    if forcing == CONST_Q:
        return {"DR_w": (base["Cf_w"] - textured["Cf_w"]) / base["Cf_w"] * 100.0,
                "DR_p": (base["Cf_p"] - textured["Cf_p"]) / base["Cf_p"] * 100.0}
    return {"DR_Ub": (1.0 - (base["Ub"] / textured["Ub"]) ** 2) * 100.0}


class DnsCase(ctypes.Structure):
    _fields_ = [
        ("xdv", ctypes.c_int), ("ydv", ctypes.c_int), ("zdv", ctypes.c_int),
        ("Rebs", ctypes.c_double), ("ub", ctypes.c_double), ("CFL", ctypes.c_double),
        ("TSTR", ctypes.c_int), ("forcing", ctypes.c_int), ("force", ctypes.c_double),
        ("precision", ctypes.c_int), ("omega", ctypes.c_double),
    ]


LOG_FIELDS = ("t", "K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "force", "rho")


class DnsStepLog(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in LOG_FIELDS]


class DnsInit(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("uc", ctypes.c_double), ("amp", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("nmodes", ctypes.c_int)]


class DnsEnsInfo(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_double), ("n_rows", ctypes.c_double), ("mean_pgrad", ctypes.c_double),
                ("force", ctypes.c_double), ("last_stress", ctypes.c_double), ("finite", ctypes.c_int),
                ("pad", ctypes.c_int)]


class DnsGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dnsgpu error {code}: {msg}")
        self.code = code


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc")], check=True)


def load():
    """Load libdnsgpu.so (no compute call: safe without a GPU)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        i8 = ctypes.POINTER(ctypes.c_int8)
        i16 = ctypes.POINTER(ctypes.c_int16)
        L.dnsgpu_create.argtypes = [ctypes.POINTER(DnsCase), ctypes.POINTER(vp)]
        L.dnsgpu_destroy.argtypes = [vp]
        L.dnsgpu_destroy.restype = None
        L.dnsgpu_last_error.restype = ctypes.c_char_p
        L.dnsgpu_set_walls.argtypes = [vp, i8, i8]
        L.dnsgpu_set_solid_bands.argtypes = [vp, i16, i16]
        L.dnsgpu_init_rest.argtypes = [vp]
        L.dnsgpu_init_field.argtypes = [vp, ctypes.POINTER(DnsInit)]
        L.dnsgpu_load_aos.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.dnsgpu_commit_load.argtypes = [vp]
        L.dnsgpu_save_aos.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.dnsgpu_get_force.argtypes = [vp, dp]
        L.dnsgpu_step.argtypes = [vp, ctypes.c_long, ctypes.c_int, ctypes.POINTER(DnsStepLog)]
        L.dnsgpu_statistics.argtypes = [vp, dp]
        L.dnsgpu_fields.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.dnsgpu_fields_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.dnsgpu_step_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
        L.dnsgpu_set_step_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.dnsgpu_step_pieces.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.dnsgpu_set_step_pieces.argtypes = [vp, ctypes.c_int]
        L.dnsgpu_set_force.argtypes = [vp, ctypes.c_double]
        L.dnsgpu_time_steps.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.dnsgpu_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp]
        L.dnsgpu_tavg_reset.argtypes = [vp]
        L.dnsgpu_tavg_sample.argtypes = [vp]
        L.dnsgpu_tavg_get.argtypes = [vp, ctypes.POINTER(DnsTavg), dp]
        L.dnsgpu_pair_placement.argtypes = [vp, ctypes.POINTER(ctypes.c_int), dp]
        L.dnsgpu_reserve_pair_arenas.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.dnsgpu_release_pair_arenas.argtypes = []
        L.dnsgpu_pair_arena_stats.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_size_t)]
        L.dnsgpu_device_bytes.argtypes = [vp]
        L.dnsgpu_device_bytes.restype = ctypes.c_size_t
        L.dnsgpu_footprint.argtypes = [ctypes.POINTER(DnsCase)]
        L.dnsgpu_footprint.restype = ctypes.c_size_t
        L.dnsens_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(vp)]
        L.dnsens_destroy.argtypes = [vp]
        L.dnsens_destroy.restype = None
        L.dnsens_member_bytes.argtypes = [vp]
        L.dnsens_member_bytes.restype = ctypes.c_size_t
        L.dnsens_footprint.argtypes = [ctypes.POINTER(DnsCase)]
        L.dnsens_footprint.restype = ctypes.c_size_t
        L.dnsens_add.argtypes = [vp, vp]
        L.dnsens_remove.argtypes = [vp, vp]
        L.dnsens_count.argtypes = [vp]
        L.dnsens_bytes_used.argtypes = [vp]
        L.dnsens_bytes_used.restype = ctypes.c_size_t
        L.dnsens_budget.argtypes = [vp]
        L.dnsens_budget.restype = ctypes.c_size_t
        L.dnsens_step.argtypes = [vp, ctypes.c_int]
        L.dnsens_info.argtypes = [vp, vp, ctypes.POINTER(DnsEnsInfo)]
        L.dnsens_rows.argtypes = [vp, vp, ctypes.POINTER(DnsStepLog), ctypes.c_int]
        L.dnsens_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), dp]
        L.dnsens_time_steps.argtypes = [vp, ctypes.c_int, dp]
        L.dnsens_time_kernels.argtypes = [vp, ctypes.c_int, dp, dp]
        L.dnsens_grid.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _chk(rc):
    if rc != 0:
        raise DnsGpuError(rc, load().dnsgpu_last_error().decode())


def omega_of(Rebs, ub, zdv):
    """dns-const-q-xycut.c:216-217."""
    tau = (3. / Rebs) * ub * ((zdv + 2) - 2.) + 0.5
    return 1. / tau


class Channel:
    """One channel case resident on the GPU (a dnsgpu handle)."""

    def __init__(self, xdv, ydv, zdv, Rebs=100.0, ub=0.1, CFL=0.1, TSTR=0, forcing=CONST_Q,
                 force=0.0, precision=FP64, omega=0.0):
        self.X, self.Y, self.Z = xdv, ydv, zdv
        self.case = DnsCase(xdv, ydv, zdv, Rebs, ub, CFL, TSTR, forcing, force, precision, omega)
        self.h = ctypes.c_void_p()
        _chk(load().dnsgpu_create(ctypes.byref(self.case), ctypes.byref(self.h)))

    def close(self):
        if self.h:
            load().dnsgpu_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def aos_shape(self, nx=None):
        return (self.X if nx is None else nx, self.Y, self.Z + 2, 19)

    def set_walls(self, bottom=None, top=None):
        def p(m):
            if m is None:
                return None
            m = np.ascontiguousarray(m, dtype=np.int8)
            assert m.shape == (self.X, self.Y)
            return m
        b, t = p(bottom), p(top)
        self._walls = (b, t)
        _chk(load().dnsgpu_set_walls(
            self.h,
            b.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)) if b is not None else None,
            t.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)) if t is not None else None))

    def set_solid_bands(self, bottom=None, top=None):
        """Solid rows above the bottom wall / below the top wall per column,
        int16 [xdv][ydv] (riblets); None = no band."""
        def p(m):
            if m is None:
                return None
            m = np.ascontiguousarray(m, dtype=np.int16)
            assert m.shape == (self.X, self.Y)
            return m
        b, t = p(bottom), p(top)
        self._bands = (b, t)
        _chk(load().dnsgpu_set_solid_bands(
            self.h,
            b.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)) if b is not None else None,
            t.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)) if t is not None else None))

    def init_rest(self):
        _chk(load().dnsgpu_init_rest(self.h))

    def init_field(self, kind, uc=0.0, amp=0.0, seed=1, nmodes=4):
        spec = DnsInit(kind, uc, amp, seed, nmodes)
        _chk(load().dnsgpu_init_field(self.h, ctypes.byref(spec)))

    def load_aos(self, ftemp, f, slab=None):
        ftemp = np.ascontiguousarray(ftemp, dtype=np.float64)
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert ftemp.shape == self.aos_shape() and f.shape == self.aos_shape()
        slab = slab or self.X
        for x0 in range(0, self.X, slab):
            nx = min(slab, self.X - x0)
            _chk(load().dnsgpu_load_aos(self.h, x0, nx, _dp(ftemp[x0:x0 + nx]), _dp(f[x0:x0 + nx])))
        _chk(load().dnsgpu_commit_load(self.h))

    def save_aos(self, slab=None):
        ft = np.zeros(self.aos_shape())
        f = np.zeros(self.aos_shape())
        slab = slab or self.X
        for x0 in range(0, self.X, slab):
            nx = min(slab, self.X - x0)
            a = np.zeros(self.aos_shape(nx))
            b = np.zeros(self.aos_shape(nx))
            _chk(load().dnsgpu_save_aos(self.h, x0, nx, _dp(a), _dp(b)))
            ft[x0:x0 + nx] = a
            f[x0:x0 + nx] = b
        return ft, f

    def save_slab(self, x0, nx):
        """Both checkpoint halves of the x-planes [x0, x0 + nx)."""
        a = np.zeros(self.aos_shape(nx))
        b = np.zeros(self.aos_shape(nx))
        _chk(load().dnsgpu_save_aos(self.h, x0, nx, _dp(a), _dp(b)))
        return a, b

    def force(self):
        g = ctypes.c_double()
        _chk(load().dnsgpu_get_force(self.h, ctypes.byref(g)))
        return g.value

    def step(self, n, cntr0=0, logs=True):
        if logs:
            rows = (DnsStepLog * max(n, 1))()
            _chk(load().dnsgpu_step(self.h, cntr0, n, rows))
            return {k: np.array([getattr(rows[i], k) for i in range(n)]) for k in LOG_FIELDS}
        _chk(load().dnsgpu_step(self.h, cntr0, n, None))
        return None

    def statistics(self):
        out = np.zeros((9, self.Z + 2))
        _chk(load().dnsgpu_statistics(self.h, _dp(out)))
        return out

    def fields(self):
        vel = np.zeros((self.X, self.Y, self.Z + 2, 3))
        den = np.zeros((self.X, self.Y, self.Z + 2))
        _chk(load().dnsgpu_fields(self.h, 0, self.X, _dp(vel), _dp(den)))
        return vel, den

    def fields_slab(self, x0, nx, noalloc=False):
        """vel [nx][Y][Z+2][3], den [nx][Y][Z+2] of the x-planes [x0, x0 + nx)
        through dnsgpu_fields, or dnsgpu_fields_noalloc (static staging)."""
        vel = np.zeros((nx, self.Y, self.Z + 2, 3))
        den = np.zeros((nx, self.Y, self.Z + 2))
        fn = load().dnsgpu_fields_noalloc if noalloc else load().dnsgpu_fields
        _chk(fn(self.h, x0, nx, _dp(vel), _dp(den)))
        return vel, den

    def time_steps(self, n, stats=False):
        tk = ctypes.c_double()
        tw = ctypes.c_double()
        _chk(load().dnsgpu_time_steps(self.h, n, 1 if stats else 0, ctypes.byref(tk), ctypes.byref(tw)))
        return tk.value, tw.value

    def autotune(self, trial_steps=5):
        """Pick the step grid by timing candidates on this handle (state kept)."""
        b = ctypes.c_int()
        t = ctypes.c_double()
        _chk(load().dnsgpu_autotune(self.h, trial_steps, ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value

    def step_shape(self, rows=False):
        """(blocks, waves per block) of the step kernel's launch, with
        rows=True (blocks, waves per block, rows per lane)."""
        b, w, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _chk(load().dnsgpu_step_shape(self.h, ctypes.byref(b), ctypes.byref(w), ctypes.byref(r)))
        return (b.value, w.value, r.value) if rows else (b.value, w.value)

    def set_force(self, force):
        _chk(load().dnsgpu_set_force(self.h, force))

    def set_step_shape(self, blocks, waves, rows=1):
        _chk(load().dnsgpu_set_step_shape(self.h, blocks, waves, rows))

    def step_pieces(self):
        """z-pieces per column of the step kernel's launch."""
        p = ctypes.c_int()
        _chk(load().dnsgpu_step_pieces(self.h, ctypes.byref(p)))
        return p.value

    def set_step_pieces(self, pieces):
        _chk(load().dnsgpu_set_step_pieces(self.h, pieces))

    def tavg_reset(self):
        _chk(load().dnsgpu_tavg_reset(self.h))

    def tavg_sample(self):
        _chk(load().dnsgpu_tavg_sample(self.h))

    def tavg(self, profiles=False):
        t = DnsTavg()
        prof = np.zeros((9, self.Z + 2)) if profiles else None
        _chk(load().dnsgpu_tavg_get(self.h, ctypes.byref(t), _dp(prof) if profiles else None))
        out = {k: getattr(t, k) for k in TAVG_FIELDS}
        if profiles:
            out["profiles"] = prof
        return out

    def pair_placement(self):
        """(destination buffers the autotuner's placement tried (0: none),
        step-time gain of the kept pair (0: the original pair kept))."""
        k, g = ctypes.c_int(), ctypes.c_double()
        _chk(load().dnsgpu_pair_placement(self.h, ctypes.byref(k), ctypes.byref(g)))
        return k.value, g.value

    def device_bytes(self):
        return load().dnsgpu_device_bytes(self.h)


def reserve_pair_arenas(bytes_per_half):
    """Pair arenas on the current device (dnsgpu_reserve_pair_arenas);
    returns the bytes each memory half holds."""
    got = ctypes.c_size_t()
    _chk(load().dnsgpu_reserve_pair_arenas(bytes_per_half, ctypes.byref(got)))
    return got.value


def release_pair_arenas():
    _chk(load().dnsgpu_release_pair_arenas())


def pair_arena_stats(device=0):
    """(blocks of half 0, blocks of half 1, bytes of ranges in use)."""
    a, b, u = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    _chk(load().dnsgpu_pair_arena_stats(device, ctypes.byref(a), ctypes.byref(b), ctypes.byref(u)))
    return a.value, b.value, u.value


def footprint(xdv, ydv, zdv, precision=FP64):
    """Device bytes a handle of this grid takes (0: refused)."""
    c = DnsCase(xdv, ydv, zdv, 100.0, 0.1, 0.1, 0, CONST_Q, 0.0, precision, 0.0)
    return load().dnsgpu_footprint(ctypes.byref(c))


def member_footprint(xdv, ydv, zdv, precision=FP64):
    """Device bytes a case of this grid is charged as an ensemble member."""
    c = DnsCase(xdv, ydv, zdv, 100.0, 0.1, 0.1, 0, CONST_Q, 0.0, precision, 0.0)
    return load().dnsens_footprint(ctypes.byref(c))


def _rows_dict(rows, n):
    return {k: np.array([getattr(rows[i], k) for i in range(n)]) for k in LOG_FIELDS}


class Ensemble:
    """Resident cases advancing together (a dnsens pool); members are
    Channel objects."""

    def __init__(self, precision=FP32, budget=0, flags=0):
        self.e = ctypes.c_void_p()
        _chk(load().dnsens_create(precision, budget, flags, ctypes.byref(self.e)))
        self.members = []

    def close(self):
        if self.e:
            load().dnsens_destroy(self.e)
            self.e = ctypes.c_void_p()
            self.members = []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, ch):
        _chk(load().dnsens_add(self.e, ch.h))
        self.members.append(ch)

    def remove(self, ch):
        _chk(load().dnsens_remove(self.e, ch.h))
        self.members.remove(ch)

    def count(self):
        return load().dnsens_count(self.e)

    def bytes_used(self):
        return load().dnsens_bytes_used(self.e)

    def budget(self):
        return load().dnsens_budget(self.e)

    @staticmethod
    def member_bytes(ch):
        return load().dnsens_member_bytes(ch.h)

    def step(self, n):
        _chk(load().dnsens_step(self.e, n))

    def info(self, ch):
        i = DnsEnsInfo()
        _chk(load().dnsens_info(self.e, ch.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in DnsEnsInfo._fields_ if k != "pad"}

    def rows(self, ch, n):
        rows = (DnsStepLog * max(n, 1))()
        _chk(load().dnsens_rows(self.e, ch.h, rows, n))
        return _rows_dict(rows, n)

    def autotune(self, trial_steps=3):
        b = ctypes.c_int()
        t = ctypes.c_double()
        _chk(load().dnsens_autotune(self.e, trial_steps, ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value

    def time_steps(self, n):
        t = ctypes.c_double()
        _chk(load().dnsens_time_steps(self.e, n, ctypes.byref(t)))
        return t.value

    def time_kernels(self, n):
        """(step-kernel ms summed over n steps, total ms) with HIP events."""
        tk = ctypes.c_double()
        tw = ctypes.c_double()
        _chk(load().dnsens_time_kernels(self.e, n, ctypes.byref(tk), ctypes.byref(tw)))
        return tk.value, tw.value

    def grid(self):
        b = ctypes.c_int()
        t = ctypes.c_int()
        _chk(load().dnsens_grid(self.e, ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value
