"""Host-side test of the reconstruction server's admission bookkeeping
(csrc/server_admit.hpp, the loop msr_server_step runs): compiled with g++
into a small driver, no GPU.  A job leaves the queue and counts against the
byte budget only once its start has returned; a start that throws leaves
that job queued and uncounted while the jobs started before it stay
admitted."""
import os
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "mrfingerprintingrecon.jl_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <stdexcept>
#include "server_admit.hpp"
using namespace msr;

static int fails = 0;
#define EXPECT(c) do { if (!(c)) { std::printf("FAIL line %d: %s\n", __LINE__, #c); ++fails; } } while (0)

int main() {
    // bytes of job id: 10 * id
    auto bytes = [](int64_t id) -> size_t { return (size_t)(10 * id); };
    {   // plain admission within the budget and max_active
        std::deque<int64_t> q{1, 2, 3, 4};
        std::vector<int64_t> active;
        size_t ab = 0;
        std::vector<int64_t> started;
        bool any = admit_jobs(q, active, ab, 60, 8, bytes, [&](int64_t id) { started.push_back(id); });
        EXPECT(any);
        EXPECT(active.size() == 3 && ab == 60);           // 10 + 20 + 30 fit, 40 does not
        EXPECT(q.size() == 1 && q.front() == 4);
        EXPECT(started.size() == 3);
    }
    {   // max_active caps the admissions
        std::deque<int64_t> q{1, 1, 1};
        std::vector<int64_t> active;
        size_t ab = 0;
        admit_jobs(q, active, ab, 1000, 2, bytes, [](int64_t) {});
        EXPECT(active.size() == 2 && q.size() == 1 && ab == 20);
    }
    {   // a start that throws: earlier jobs stay admitted, the failing one stays queued
        std::deque<int64_t> q{1, 2, 3};
        std::vector<int64_t> active;
        size_t ab = 0;
        bool threw = false;
        try {
            admit_jobs(q, active, ab, 1000, 8, bytes, [](int64_t id) {
                if (id == 2) throw std::runtime_error("no memory");
            });
        } catch (const std::runtime_error&) {
            threw = true;
        }
        EXPECT(threw);
        EXPECT(active.size() == 1 && active[0] == 1);
        EXPECT(ab == 10);                                 // job 2's bytes not counted
        EXPECT(q.size() == 2 && q.front() == 2);          // job 2 still first in the queue
        // the next step admits it once its start succeeds
        bool any = admit_jobs(q, active, ab, 1000, 8, bytes, [](int64_t) {});
        EXPECT(any && active.size() == 3 && ab == 60 && q.empty());
    }
    {   // nothing fits: nothing changes
        std::deque<int64_t> q{5};
        std::vector<int64_t> active{1};
        size_t ab = 10;
        bool any = admit_jobs(q, active, ab, 40, 8, bytes, [](int64_t) {});
        EXPECT(!any && active.size() == 1 && ab == 10 && q.size() == 1);
    }
    std::printf("%s\n", fails ? "FAILED" : "OK");
    return fails ? 1 : 0;
}
"""


def test_admission_bookkeeping(tmp_path):
    src = tmp_path / "admit.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "admit"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I" + CSRC, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    if r.returncode != 0 and "not found" in r.stderr:
        pytest.skip("g++ unavailable")
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
