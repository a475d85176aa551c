"""Host bookkeeping of the pair arenas (dns-solvers_amd/csrc/rangealloc.h),
compiled with g++ and driven by a seeded random sequence of allocations
and releases: live ranges never overlap and stay inside the block, the
bytes in use are the sum of the live rounded sizes, a release that is not
a live range is refused and changes nothing, and after every range is
returned the block is one free range again (neighbours merged)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(os.path.dirname(HERE), "dns-solvers_amd", "csrc")

DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include <vector>
#include "rangealloc.h"

struct Live { size_t off, n; };
static unsigned long long s = 12345;
static size_t rnd(size_t m) { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return (size_t)(s >> 33) % m; }

#define CHECK(c) do { if (!(c)) { printf("FAIL line %d\n", __LINE__); return 1; } } while (0)

int main(int argc, char** argv)
{
    const size_t gran = 2u << 20, cap = (size_t)8 << 30;
    s = (unsigned long long)atoll(argv[1]);
    RangeAlloc ra(cap, gran);
    CHECK(ra.capacity() == cap && ra.empty() && ra.free_ranges() == 1);
    std::vector<Live> live;
    size_t allocs = 0, fails = 0;
    for (int it = 0; it < 20000; it++) {
        if (live.empty() || rnd(100) < 55) {
            const size_t n = 1 + rnd((size_t)1400 << 20);  // up to 1.4 GB, not granule-aligned
            size_t off = 0;
            if (ra.alloc(n, &off)) {
                allocs++;
                const size_t r = ra.rounded(n);
                CHECK(off % gran == 0 && off + r <= cap);
                for (const Live& l : live) CHECK(off + r <= l.off || l.off + ra.rounded(l.n) <= off);
                live.push_back({off, n});
            } else {
                fails++;
            }
        } else {
            const size_t k = rnd(live.size());
            // a release that is not a live range (shifted by a granule) is refused
            const size_t before = ra.used(), fr = ra.free_ranges();
            bool other = false;
            for (const Live& l : live)
                other |= l.off == live[k].off + gran && ra.rounded(l.n) == ra.rounded(live[k].n);
            if (!other) CHECK(!ra.release(live[k].off + gran, live[k].n));
            CHECK(!ra.release(live[k].off, live[k].n + gran));  // another size: refused
            CHECK(ra.used() == before && ra.free_ranges() == fr);
            CHECK(ra.release(live[k].off, live[k].n));
            CHECK(!ra.release(live[k].off, live[k].n));  // twice: refused
            live.erase(live.begin() + (long)k);
        }
        size_t sum = 0;
        for (const Live& l : live) sum += ra.rounded(l.n);
        CHECK(ra.used() == sum);
    }
    for (const Live& l : live) CHECK(ra.release(l.off, l.n));
    CHECK(ra.empty() && ra.free_ranges() == 1);
    size_t off = 1;
    CHECK(ra.alloc(cap, &off) && off == 0 && !ra.alloc(1, &off));
    printf("OK %zu %zu\n", allocs, fails);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("rangealloc")
    src = d / "t.cpp"
    src.write_text(DRIVER)
    exe = d / "t"
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", HDR, "-o", str(exe), str(src)], check=True)
    return str(exe)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_rangealloc_random_sequences(driver, seed):
    out = subprocess.run([driver, str(seed)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout + out.stderr
    allocs, fails = map(int, out.stdout.split()[1:3])
    assert allocs > 1000 and fails > 10  # the block filled up now and then
