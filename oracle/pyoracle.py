"""ctypes view of the CPU restatement in lbm_oracle.c.

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
import this module.

Arrays use the reference's AoS layout without ghost planes,
``[X][Y][Z+2][19]`` float64 (DNS-solver/src/definitions.h:93), which is the
layout of each half of an ``lbf.NNNN`` checkpoint
(DNS-solver/src/dns-const-q-xycut.c:491-503).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class OracleCase(ctypes.Structure):
    _fields_ = [
        ("X", ctypes.c_int), ("Y", ctypes.c_int), ("Z", ctypes.c_int),
        ("omega", ctypes.c_double), ("ub", ctypes.c_double), ("CFL", ctypes.c_double),
        ("TM", ctypes.c_int), ("forcing", ctypes.c_int), ("F", ctypes.c_double),
        ("s", ctypes.POINTER(ctypes.c_int8)),
    ]


LOG_FIELDS = ("t", "K", "flowrate", "massrate", "stup", "stdown", "stress", "pgrad", "force", "rho")


class OracleLog(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in LOG_FIELDS]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        cp = ctypes.POINTER(OracleCase)
        lp = ctypes.POINTER(OracleLog)
        L.oracle_omega.restype = ctypes.c_double
        L.oracle_omega.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.oracle_run.restype = ctypes.c_int
        L.oracle_run.argtypes = [cp, dp, dp, ctypes.c_long, ctypes.c_long, lp]
        L.oracle_step.restype = ctypes.c_int
        L.oracle_step.argtypes = [cp, dp, dp, ctypes.c_long, lp]
        L.oracle_encalc.argtypes = [cp, dp, ctypes.c_long, lp]
        L.oracle_wshearstr.restype = ctypes.c_int
        L.oracle_wshearstr.argtypes = [cp, dp, ctypes.c_long, lp]
        L.oracle_collide_stream.argtypes = [cp, dp, dp, ctypes.c_double]
        L.oracle_statistics.argtypes = [cp, dp, dp]
        L.oracle_moments.argtypes = [cp, dp, dp, dp]
        L.oracle_twoD.argtypes = [cp, dp, dp, dp]
        L.oracle_stream_coverage.argtypes = [cp, ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def omega(Rebs, ub, zdv):
    return lib().oracle_omega(Rebs, ub, zdv)


def wall_mask(X, Y, Z, bottom=None, top=None):
    """Node-type mask [X][Y][Z+2] in the reference codes (solid_init.c:3-8):
    wall rows -1 unless the 2-D bottom/top masks say otherwise."""
    s = np.zeros((X, Y, Z + 2), dtype=np.int8)
    s[:, :, 0] = -1 if bottom is None else bottom
    s[:, :, Z + 1] = -1 if top is None else top
    return s


class Oracle:
    """One channel case on the CPU restatement."""

    def __init__(self, X, Y, Z, omega, ub, CFL, TM=0, forcing=0, F=0.0, mask=None):
        self.X, self.Y, self.Z = X, Y, Z
        self.mask = np.ascontiguousarray(mask if mask is not None else wall_mask(X, Y, Z), dtype=np.int8)
        assert self.mask.shape == (X, Y, Z + 2)
        self.case = OracleCase(X, Y, Z, omega, ub, CFL, TM, forcing, F,
                               self.mask.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))

    def shape(self):
        return (self.X, self.Y, self.Z + 2, 19)

    def run(self, ftemp, f, n, cntr0=0):
        """Advance n steps in place; returns the n log rows as a dict of arrays."""
        rows = (OracleLog * max(n, 1))()
        rc = lib().oracle_run(ctypes.byref(self.case), _dp(ftemp), _dp(f), cntr0, n, rows)
        if rc != 0:
            raise FloatingPointError("Eroor: NAN, found in wall shear stress")
        return {k: np.array([getattr(rows[i], k) for i in range(n)]) for k in LOG_FIELDS}

    def logs_of(self, ftemp, cntr=0):
        r = OracleLog()
        lib().oracle_encalc(ctypes.byref(self.case), _dp(ftemp), cntr, ctypes.byref(r))
        lib().oracle_wshearstr(ctypes.byref(self.case), _dp(ftemp), cntr, ctypes.byref(r))
        return {k: getattr(r, k) for k in LOG_FIELDS}

    def collide_stream(self, ftemp, f, G):
        lib().oracle_collide_stream(ctypes.byref(self.case), _dp(ftemp), _dp(f), G)

    def statistics(self, ftemp):
        out = np.zeros((9, self.Z + 2))
        lib().oracle_statistics(ctypes.byref(self.case), _dp(ftemp), _dp(out))
        return out

    def moments(self, ftemp):
        vel = np.zeros((self.X, self.Y, self.Z + 2, 3))
        den = np.zeros((self.X, self.Y, self.Z + 2))
        lib().oracle_moments(ctypes.byref(self.case), _dp(ftemp), _dp(vel), _dp(den))
        return vel, den

    def twoD(self, ftemp):
        ex = np.zeros(self.Z + 2)
        lb = np.zeros(self.Z + 2)
        lib().oracle_twoD(ctypes.byref(self.case), _dp(ftemp), _dp(ex), _dp(lb))
        return ex, lb

    def coverage(self):
        cnt = np.zeros(self.shape(), dtype=np.int32)
        lib().oracle_stream_coverage(ctypes.byref(self.case), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return cnt
