"""ctypes binding of libdnsserve.so (driver/dnsserve.h): the seeded texture
sweep and its continuous-batching service, for tests and benchmarks."""
import ctypes
import os

import dnsgpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdnsserve.so")

SMOOTH, GROOVES, RIBLETS = 0, 1, 2


class Request(ctypes.Structure):
    _fields_ = [("xdv", ctypes.c_int), ("ydv", ctypes.c_int), ("zdv", ctypes.c_int), ("retau", ctypes.c_double),
                ("texture", ctypes.c_int), ("period", ctypes.c_int), ("param", ctypes.c_double),
                ("arrival", ctypes.c_double), ("seed", ctypes.c_uint64)]


class Config(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("ub", ctypes.c_double), ("transient", ctypes.c_int),
                ("window", ctypes.c_int), ("check", ctypes.c_int), ("tol", ctypes.c_double), ("cap", ctypes.c_int),
                ("budget", ctypes.c_size_t), ("realtime", ctypes.c_int), ("autotune", ctypes.c_int),
                ("init_modes", ctypes.c_int), ("init_amp", ctypes.c_double), ("verbose", ctypes.c_int),
                ("pair_arenas", ctypes.c_int)]


RESULT_FIELDS = ("arrival", "admitted", "finished", "steps", "nodes", "utau", "Re_tau", "Cf_w", "Cf_p", "Ub",
                 "stress", "u_slip")


class Result(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in RESULT_FIELDS] + [("status", ctypes.c_int), ("pad", ctypes.c_int)]


SUMMARY_FIELDS = ("wall", "busy", "node_steps", "mlups", "mlups_busy", "p50", "p99", "mean_ttr", "solves_per_hour",
                  "peak_bytes", "arena_s", "arena_bytes")


class Summary(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in SUMMARY_FIELDS] + [("max_resident", ctypes.c_int),
                                                                 ("completed", ctypes.c_int)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run __graft_entry__.build() first")
        dnsgpu.load()  # the HIP library first (its error strings)
        L = ctypes.CDLL(LIB_PATH)
        L.dnsserve_defaults.argtypes = [ctypes.POINTER(Config)]
        L.dnsserve_defaults.restype = None
        L.dnsserve_generate.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(Request)]
        L.dnsserve_generate.restype = None
        L.dnsserve_case.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(Config), ctypes.POINTER(dnsgpu.DnsCase),
                                    ctypes.POINTER(dnsgpu.DnsInit)]
        L.dnsserve_case.restype = None
        L.dnsserve_sweep.argtypes = [ctypes.POINTER(Request), ctypes.c_int, ctypes.POINTER(Config),
                                     ctypes.POINTER(Result), ctypes.POINTER(Summary)]
        L.dnsserve_open.argtypes = [ctypes.POINTER(Request), ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_void_p)]
        L.dnsserve_sequential.argtypes = [ctypes.POINTER(Request), ctypes.c_int, ctypes.POINTER(Config), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def defaults(**kw):
    c = Config()
    load().dnsserve_defaults(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def generate(n=64, seed=5, mean_gap=2.0):
    reqs = (Request * n)()
    load().dnsserve_generate(n, seed, mean_gap, reqs)
    return reqs


def requests(items):
    """Request array from dicts (xdv, ydv, zdv, retau, texture, period, param, arrival, seed)."""
    reqs = (Request * len(items))()
    for r, d in zip(reqs, items):
        for k, v in d.items():
            setattr(r, k, v)
    return reqs


def case_of(req, cfg):
    c = dnsgpu.DnsCase()
    i = dnsgpu.DnsInit()
    load().dnsserve_case(ctypes.byref(req), ctypes.byref(cfg), ctypes.byref(c), ctypes.byref(i))
    return c, i


def as_dict(s):
    return {k: getattr(s, k) for k, _ in s._fields_ if k != "pad"}


def sweep(reqs, cfg):
    n = len(reqs)
    res = (Result * n)()
    summ = Summary()
    dnsgpu._chk(load().dnsserve_sweep(reqs, n, ctypes.byref(cfg), res, ctypes.byref(summ)))
    return [as_dict(r) for r in res], as_dict(summ)


def sequential(reqs, cfg, precision, steps):
    n = len(reqs)
    st = (ctypes.c_double * n)(*steps)
    secs = (ctypes.c_double * n)()
    dnsgpu._chk(load().dnsserve_sequential(reqs, n, ctypes.byref(cfg), precision, st, secs))
    return list(secs)


class Handle:
    """A handle created, textured and initialised for a request
    (dnsserve_open); usable as an ensemble member."""

    def __init__(self, req, cfg):
        self.h = ctypes.c_void_p()
        dnsgpu._chk(load().dnsserve_open(ctypes.byref(req), ctypes.byref(cfg), ctypes.byref(self.h)))
        self.X, self.Y, self.Z = req.xdv, req.ydv, req.zdv

    def close(self):
        if self.h:
            dnsgpu.load().dnsgpu_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
