"""The scratch pool's range allocator (labelpropcuda_amd/csrc/dlp_pool.h),
host code only: compiled here with g++ against a small driver.

It decides which batches a serving runner admits, so it is checked for the
properties admission relies on: a range is handed out only when it fits,
ranges never overlap, a refused request changes nothing, returned ranges
merge back so an empty pool is one free range again, and the counters
(bytes in use, peak, leases) follow.  A brute-force byte map over random
acquire / release sequences is the reference."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER_DIR = os.path.join(ROOT, "labelpropcuda_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <random>
#include <thread>
#include <vector>
#include "dlp_pool.h"

#define CHECK(c) do { if (!(c)) { std::printf("FAIL line %d: %s\n", __LINE__, #c); return 1; } } while (0)

int main() {
  /* fixed cases */
  {
    RangePool p(1000);                       // rounds down to 768
    CHECK(p.capacity() == 768);
    CHECK(RangePool::rounded(1) == 256 && RangePool::rounded(256) == 256 &&
          RangePool::rounded(257) == 512 && RangePool::rounded(0) == 256);
    int64_t a, b, c, d;
    CHECK(p.acquire(100, &a) && a == 0);
    CHECK(p.acquire(300, &b) && b == 256);   // 512 bytes
    CHECK(!p.acquire(1, &d));                // full
    CHECK(p.inUse() == 768 && p.peak() == 768 && p.leases() == 2);
    CHECK(p.release(a) && !p.release(a));    // second release of the same range fails
    CHECK(p.acquire(200, &c) && c == 0);     // first fit reuses the hole
    CHECK(!p.acquire(257, &d));              // 256 free, 512 needed: refused
    CHECK(p.inUse() == 768 && p.leases() == 3);
    CHECK(p.release(b) && p.release(c));
    CHECK(p.inUse() == 0 && p.freeRanges() == 1 && p.largestFree() == 768);
  }
  {
    /* merges with the free range before and after */
    RangePool p(256 * 5);
    int64_t o[5];
    for (int i = 0; i < 5; ++i) CHECK(p.acquire(256, &o[i]) && o[i] == 256 * i);
    CHECK(p.release(o[1]) && p.release(o[3]));
    CHECK(p.freeRanges() == 2 && p.largestFree() == 256);
    CHECK(p.release(o[2]));                  // joins 1 and 3
    CHECK(p.freeRanges() == 1 && p.largestFree() == 768);
    int64_t big;
    CHECK(p.acquire(700, &big) && big == 256);
    CHECK(p.release(o[0]) && p.release(o[4]) && p.release(big));
    CHECK(p.freeRanges() == 1 && p.largestFree() == 1280 && p.inUse() == 0);
  }
  {
    RangePool p(0);
    int64_t x;
    CHECK(p.capacity() == 0 && !p.acquire(1, &x));
  }
  /* random sequences against a byte map */
  std::mt19937 rng(7);
  for (int trial = 0; trial < 200; ++trial) {
    const int64_t cap = 256 * (1 + rng() % 64);
    RangePool p(cap);
    std::vector<int> owner(cap / 256, -1);
    std::vector<std::pair<int64_t, int64_t>> live;  // offset, bytes
    int64_t peak = 0, inUse = 0, leases = 0;
    for (int step = 0; step < 400; ++step) {
      if (!live.empty() && rng() % 3 == 0) {
        const size_t i = rng() % live.size();
        CHECK(p.release(live[i].first));
        for (int64_t u = live[i].first / 256; u < (live[i].first + live[i].second) / 256; ++u)
          owner[u] = -1;
        inUse -= live[i].second;
        live.erase(live.begin() + i);
        continue;
      }
      const int64_t want = 1 + rng() % (cap / 2 + 1);
      const int64_t need = RangePool::rounded(want);
      /* a fit exists iff some run of free units is long enough */
      bool fits = false;
      int64_t run = 0, firstFit = -1;
      for (size_t u = 0; u < owner.size(); ++u) {
        run = owner[u] < 0 ? run + 256 : 0;
        if (run >= need) { fits = true; firstFit = (int64_t)(u + 1) * 256 - run; break; }
      }
      int64_t off = -1;
      const bool got = p.acquire(want, &off);
      CHECK(got == fits);
      if (!got) continue;
      CHECK(off == firstFit);                // lowest fitting offset
      CHECK(off % 256 == 0 && off + need <= cap);
      for (int64_t u = off / 256; u < (off + need) / 256; ++u) {
        CHECK(owner[u] < 0);                 // no overlap
        owner[u] = step;
      }
      live.push_back({off, need});
      inUse += need;
      ++leases;
      if (inUse > peak) peak = inUse;
      CHECK(p.inUse() == inUse && p.peak() == peak && p.leases() == leases);
    }
    for (auto &l : live) CHECK(p.release(l.first));
    CHECK(p.inUse() == 0 && p.freeRanges() == 1 && p.largestFree() == p.capacity());
  }
  /* two threads leasing and returning concurrently */
  {
    RangePool p(256 * 64);
    auto worker = [&](unsigned seed) {
      std::mt19937 r(seed);
      for (int i = 0; i < 20000; ++i) {
        int64_t o;
        if (p.acquire(256 * (1 + r() % 4), &o)) p.release(o);
      }
    };
    std::thread t1(worker, 1), t2(worker, 2);
    t1.join();
    t2.join();
    CHECK(p.inUse() == 0 && p.freeRanges() == 1);
  }
  std::printf("ok\n");
  return 0;
}
"""


def test_range_pool(tmp_path):
    src = tmp_path / "pool_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pool_driver"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", HEADER_DIR, str(src),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


@pytest.mark.parametrize("flag", ["-fsanitize=address,undefined", "-fsanitize=thread"])
def test_range_pool_sanitized(tmp_path, flag):
    """The same driver under the host sanitizers (memory errors, races)."""
    src = tmp_path / "pool_driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pool_driver"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", flag, "-I", HEADER_DIR,
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f"g++ {flag} not usable here: {r.stderr[:200]}")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
