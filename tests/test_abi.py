"""The C-ABI library loads and exports every entry point include/dnsgpu.h
declares (no compute call: runs without a GPU)."""
import ctypes
import os
import re

import dnsgpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(REPO, "include", "dnsgpu.h")).read()
    return sorted(set(re.findall(r"\b((?:dnsgpu|dnsens)_[a-z_0-9]+)\s*\(", txt)))


def test_header_lists_every_binding():
    assert set(declared_symbols()) == set(dnsgpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dnsgpu.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # dns_case: 3 ints, 3 doubles, 2 ints, double, int, double
    assert ctypes.sizeof(dnsgpu.DnsCase) == 72
    assert ctypes.sizeof(dnsgpu.DnsStepLog) == 80
    assert ctypes.sizeof(dnsgpu.DnsInit) == 40
    assert ctypes.sizeof(dnsgpu.DnsTavg) == 17 * 8
    assert ctypes.sizeof(dnsgpu.DnsEnsInfo) == 5 * 8 + 8
