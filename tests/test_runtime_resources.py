"""Bounded per-process resources of the product library, checked on the CPU:
- no kernel needs scratch (private) memory, spills registers or uses a dynamic stack, from the
  compiler's per-kernel resource report written by the build (chemtensor_amd/_build/kernel_resources.json);
  a kernel with scratch makes the runtime allocate scratch per queue when it launches;
- the per-thread pool of pinned staging blocks and their events stays within its cap whatever the
  GPU's progress, never hands out a block whose copy is in flight, and frees everything on release
  (chemtensor_amd/csrc/staging.h, driven by tests/staging_check.cpp with simulated buffers and
  events)."""

import json
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "chemtensor_amd", "csrc")
RESOURCES = os.path.join(ROOT, "chemtensor_amd", "_build", "kernel_resources.json")


def demangle(names):
    r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True)
    return r.stdout.splitlines()


def test_kernels_use_no_scratch_and_do_not_spill():
    if not os.path.exists(RESOURCES):
        pytest.skip("library not built here (no kernel resource report)")
    res = json.load(open(RESOURCES))
    assert res
    for k, v in res.items():
        assert v["ScratchSize [bytes/lane]"] == 0, k
        assert v["VGPRs Spill"] == 0 and v["SGPRs Spill"] == 0, k
        assert v["Dynamic Stack"] == "False", k
    # the report covers every kernel defined in the sources
    names = set()
    for f in os.listdir(CSRC):
        if f.endswith(".hip"):
            text = open(os.path.join(CSRC, f)).read()
            names |= set(re.findall(r"__global__[^(]*?\b(k_\w+)\s*\(", text))
    reported = " ".join(demangle(sorted(res)))
    missing = [n for n in names if re.search(r"\b" + n + r"\b", reported) is None]
    assert not missing, missing


@pytest.fixture(scope="module")
def staging_exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("staging") / "staging_check"
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "staging_check.cpp"), "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("scenario,nreq", [("stalled", 3000), ("fast", 3000), ("growing", 400), ("mixed", 5000)])
def test_staging_pool_is_bounded(staging_exe, scenario, nreq):
    out = subprocess.run([str(staging_exe), scenario, str(nreq), "7"], capture_output=True, text=True, check=True, timeout=120).stdout
    v = {k: int(x) for k, x in (line.split() for line in out.splitlines())}
    cap = 32   # STAGE_MAX_BLOCKS
    assert v["max_blocks"] <= cap and v["max_buffers"] <= cap and v["max_events"] <= cap
    assert v["max_bytes"] <= cap * v["cap_of_max_request"]
    assert v["reused_in_flight"] == 0 and v["too_small"] == 0
    assert v["live_buffers_after_release"] == 0 and v["live_events_after_release"] == 0
    if scenario == "stalled":
        # with no copy ever completing on its own the pool fills to its cap and then waits
        assert v["max_blocks"] == cap
    if scenario == "fast":
        assert v["max_blocks"] <= 2
