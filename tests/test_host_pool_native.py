"""The host thread pool of the host-buffer calls (cuspx_amd/csrc/host_pool.hpp)
built natively with g++ under ThreadSanitizer and, separately, under
AddressSanitizer / UndefinedBehaviorSanitizer: every index of a batch runs
exactly once for sizes 0..96, with several threads submitting batches at the
same time and with calls that take uneven time."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cuspx_amd", "csrc")

HARNESS = r"""
#include <atomic>
#include <chrono>
#include <cstdio>
#include <memory>
#include <thread>
#include <vector>
#include "host_pool.hpp"

static int fail(const char* what) { std::printf("FAIL %s\n", what); return 1; }

int main() {
    spx::HostPool& pool = spx::HostPool::get();
    // every index exactly once, batch sizes 0..96
    for (int rep = 0; rep < 3000; rep++) {
        const uint32_t n = (uint32_t)((rep * 7919) % 97);
        std::unique_ptr<std::atomic<int>[]> hit(new std::atomic<int>[n + 1]);
        for (uint32_t k = 0; k <= n; k++) hit[k] = 0;
        pool.run(n, [&](uint32_t k) { hit[k]++; });
        for (uint32_t k = 0; k < n; k++)
            if (hit[k] != 1) return fail("an index ran other than once");
        if (hit[n] != 0) return fail("an index past the batch ran");
    }
    // several submitting threads; calls of uneven length
    std::atomic<long> total{0};
    std::vector<std::thread> th;
    for (int t = 0; t < 4; t++)
        th.emplace_back([&, t] {
            for (int r = 0; r < 300; r++)
                pool.run(33, [&](uint32_t k) {
                    if ((k + r + t) % 11 == 0)
                        std::this_thread::sleep_for(std::chrono::microseconds(20));
                    total += k;
                });
        });
    for (auto& x : th) x.join();
    if (total != 4L * 300 * (32 * 33 / 2)) return fail("sum of indices");
    std::printf("OK\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_host_pool_under_sanitizers(tmp_path, san):
    src = tmp_path / "pool.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "pool"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=" + san, "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-pthread", "-I" + CSRC, str(src), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=180)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr and "runtime error" not in r.stderr
