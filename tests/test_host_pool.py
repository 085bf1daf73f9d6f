"""The batched steps' host worker threads (stormbird_amd/csrc/sb_pool.h), on the CPU.

parallel_for(n, fn) must call fn(i) exactly once for every i in [0, n) whatever n,
rethrow an exception thrown by any call, and stay correct when it is called from two
host threads at once or from inside fn (the kept threads are then busy and the call
starts threads of its own), and in a process forked from the one that started the kept
threads.  A C++ program built with g++ checks each, and a planted
mistake (a chunk boundary off by one) must fail the same checks.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
POOL = os.path.join(ROOT, "stormbird_amd", "csrc", "sb_pool.h")

PROGRAM = r"""
#include <sys/wait.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <thread>
#include <vector>
#include "sb_pool.h"

int main() {
    using sb_pool::parallel_for;
    std::vector<int> hits(5000);
    for (int rep = 0; rep < 600; ++rep) {
        const int n = rep % 7 == 0 ? rep % 40 : 1 + (rep * 37) % 5000;  // below and above the 32-per-thread cut
        std::fill(hits.begin(), hits.end(), 0);
        parallel_for(n, [&](int i) { hits[i] += 1; });
        for (int i = 0; i < 5000; ++i)
            if (hits[i] != (i < n ? 1 : 0)) {
                printf("coverage n %d i %d hits %d\n", n, i, hits[i]);
                return 1;
            }
    }
    int caught = 0;
    for (int rep = 0; rep < 100; ++rep) {
        try {
            parallel_for(1024, [&](int i) {
                if (i == (rep * 13) % 1024) throw std::runtime_error("planted");
            });
        } catch (const std::runtime_error&) {
            ++caught;
        }
    }
    if (caught != 100) {
        printf("exceptions %d\n", caught);
        return 1;
    }
    std::atomic<long> sum{0};
    auto body = [&] {
        for (int rep = 0; rep < 50; ++rep)
            parallel_for(256, [&](int) { parallel_for(64, [&](int j) { sum += j; }); });
    };
    std::thread a(body), b(body);
    a.join();
    b.join();
    const long want = 2L * 50 * 256 * (63 * 64 / 2);
    if (sum != want) {
        printf("nested %ld vs %ld\n", (long)sum, want);
        return 1;
    }
    // a process forked from this one has none of the kept threads: its calls start their own
    fflush(stdout);
    const pid_t child = fork();
    if (child == 0) {
        std::vector<int> h(1024);
        parallel_for(1024, [&](int i) { h[i] = i; });
        for (int i = 0; i < 1024; ++i)
            if (h[i] != i) _exit(3);
        exit(0);  // runs the child's exit handlers
    }
    int status = 0;
    waitpid(child, &status, 0);
    if (!WIFEXITED(status) || WEXITSTATUS(status) != 0) {
        printf("forked child status %d\n", status);
        return 1;
    }
    printf("ok\n");
    return 0;
}
"""


def build_and_run(tmp_path, header_text):
    inc = tmp_path / "inc"
    inc.mkdir(exist_ok=True)
    (inc / "sb_pool.h").write_text(header_text)
    src = tmp_path / "pool.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "pool"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-Wall", "-Werror", "-I", str(inc), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)


def test_parallel_for_covers_every_index_rethrows_and_nests(tmp_path):
    r = build_and_run(tmp_path, open(POOL).read())
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_planted_chunk_boundary_mistake_is_caught(tmp_path):
    text = open(POOL).read()
    good = "i < (int)((long)n * (w + 1) / nt); ++i"
    assert text.count(good) == 1
    r = build_and_run(tmp_path, text.replace(good, "i <= (int)((long)n * (w + 1) / nt) && i < n; ++i"))
    assert r.returncode != 0 and "coverage" in r.stdout, r.stdout
