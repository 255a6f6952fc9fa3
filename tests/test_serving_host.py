"""CPU checks of the serving loop's host side: the workload generator, the
request/result layouts shared with include/hpcsort.h, the checksum and the
summary, and argument rejection by the engine before any device call."""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from hpcsharp_amd import _native as N
from hpcsharp_amd import serving as S

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_trace_matches_the_workload_definition():
    t = S.make_trace()
    gaps = np.diff(np.concatenate([[0.0], t["arrival_s"]]))
    assert t.size == 20000
    assert abs(gaps.mean() - 0.5e-3) < 0.02e-3                 # exponential gaps, mean 1/2000 s
    assert abs(np.std(gaps) / gaps.mean() - 1.0) < 0.05        # exponential: cv = 1
    assert abs((t["op"] == N.HPCS_OP_SORT).mean() - 0.7) < 0.015
    assert set(np.unique(t["dtype"])) == {N.HPCS_DTYPE_U32, N.HPCS_DTYPE_F32, N.HPCS_DTYPE_U64}
    assert np.all(np.abs(np.bincount(t["dtype"], minlength=3)[[0, 1, 2]] / t.size - 1 / 3) < 0.015)
    assert t["n"].min() >= 1 << 10 and t["n"].max() < 1 << 26
    # log-uniform sizes: E[n] = (2^26 - 2^10) / ln(2^16)
    assert abs(t["n"].mean() / ((2 ** 26 - 2 ** 10) / np.log(2 ** 16)) - 1) < 0.05
    tk = t[t["op"] == N.HPCS_OP_TOPK]
    assert tk["k"].min() >= 10 and tk["k"].max() < 10 ** 4 and np.all(tk["k"] <= tk["n"])
    assert np.all(t["k"][t["op"] == N.HPCS_OP_SORT] == 0)
    assert np.array_equal(S.make_trace()["seed"], t["seed"])   # reproducible


def test_layouts_match_the_header():
    src = """
#include <stdio.h>
#include <stddef.h>
#include "hpcsort.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(hpcs_request), offsetof(hpcs_request, seed),
         offsetof(hpcs_request, arrival_s), sizeof(hpcs_request_result), offsetof(hpcs_request_result, status),
         sizeof(hpcs_serve_stats), offsetof(hpcs_serve_stats, coalesced_requests), sizeof(hpcs_serve_config),
         offsetof(hpcs_serve_config, request_timeout_s), offsetof(hpcs_serve_config, coalesce_max_batch_keys));
  return 0; }
"""
    exe = os.path.join("/tmp", f"hpcs_layout_{os.getpid()}")
    r = subprocess.run(["gcc", "-std=c99", "-I", os.path.join(REPO, "include"), "-x", "c", "-", "-o", exe],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    os.remove(exe)
    assert got == [S.REQUEST.itemsize, S.REQUEST.fields["seed"][1], S.REQUEST.fields["arrival_s"][1],
                   S.RESULT.itemsize, S.RESULT.fields["status"][1], ctypes.sizeof(S.ServeStats),
                   S.ServeStats.coalesced_requests.offset, ctypes.sizeof(S.ServeConfig),
                   S.ServeConfig.request_timeout_s.offset, S.ServeConfig.coalesce_max_batch_keys.offset]


def test_checksum_known_answers():
    # mix64(golden) is the first output of splitmix64 seeded with 0
    assert S.checksum(np.zeros(1, dtype=np.uint64)) == 0xE220A8397B1DCDAF
    a = np.arange(5, dtype=np.uint32)
    assert S.checksum(a) != S.checksum(a[::-1].copy())        # position-weighted
    assert S.checksum(np.array([7], dtype=np.uint32)) == S.checksum(np.array([7], dtype=np.uint64))


def test_payload_generator_closed_forms():
    req = np.zeros(1, dtype=S.REQUEST)[0]
    req["n"], req["seed"] = 1000, 12345
    for dt in (N.HPCS_DTYPE_U32, N.HPCS_DTYPE_U64, N.HPCS_DTYPE_F32, N.HPCS_DTYPE_F64):
        req["dtype"] = dt
        x = S.payload(req)
        assert x.dtype == S.DTYPES[dt] and x.size == 1000
        if x.dtype.kind == "f":
            assert np.all(np.isfinite(x))
    req["dtype"] = N.HPCS_DTYPE_U64
    from hpcsharp_amd import datagen
    assert np.array_equal(S.payload(req), datagen.splitmix64(12345, 1000))


def test_summary_percentiles():
    t = np.zeros(4, dtype=S.REQUEST)
    t["arrival_s"] = [0.0, 0.1, 0.2, 0.3]
    t["n"] = 1000
    t["op"] = [0, 0, 1, 1]
    r = np.zeros(4, dtype=S.RESULT)
    r["admit_s"] = t["arrival_s"] + 0.001
    r["done_s"] = t["arrival_s"] + [0.010, 0.020, 0.030, 0.040]
    r["status"] = [0, 0, 0, N.HPCS_E_BUDGET]
    s = S.summarize(t, r)
    assert s["completed"] == 3 and s["rejected"] == 1
    assert s["p50_ms"] == pytest.approx(20.0)
    assert s["admit_wait_p50_ms"] == pytest.approx(1.0)
    assert s["topk_p50_ms"] == pytest.approx(30.0)


@pytest.mark.parametrize("bad", ["op", "n", "k", "order", "slots", "budget", "timeout", "batch_requests",
                                 "batch_keys"])
def test_engine_rejects_bad_requests_before_touching_a_device(bad):
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libhpcsort.so not built")
    t = S.make_trace(nreq=8)
    slots, budget, extra = 2, 10 ** 9, {}
    if bad == "timeout":
        extra = dict(timeout_s=-1.0)
    elif bad == "batch_requests":
        extra = dict(coalesce_max_keys=4096, coalesce_max_requests=1)
    elif bad == "batch_keys":
        extra = dict(coalesce_max_keys=-5)
    elif bad == "op":
        t["op"][3] = 7
    elif bad == "n":
        t["n"][2] = 0
    elif bad == "k":
        t["op"][1], t["k"][1] = N.HPCS_OP_TOPK, t["n"][1] + 1
    elif bad == "order":
        t["arrival_s"][5] = -1.0
    elif bad == "slots":
        slots = 0
    else:
        budget = 0
    with pytest.raises(N.HpcsError) as e:
        S.run(t, slots=slots, hbm_budget=budget, **extra)
    assert e.value.code == N.HPCS_E_INVALID_ARG


@pytest.mark.parametrize("sanitize", ["thread", "address,undefined"])
def test_serving_core_with_a_mock_executor(tmp_path, sanitize):
    """The engine's admission, submitter and reaper threads (the same
    hpcs_serve_core.hpp the library builds) driven by a mock executor:
    results, budget, coalesced batches, injected allocation failures and
    completion deadlines, under the thread sanitizer and, separately, the
    address and undefined-behaviour sanitizers."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "serve_mock_test")
    flags = ["-fsanitize=" + sanitize] + (["-fno-sanitize-recover=undefined"] if "undefined" in sanitize else [])
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", *flags, "-pthread", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(REPO, "include"), os.path.join(REPO, "tests", "host", "serve_mock_test.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66", ASAN_OPTIONS="detect_leaks=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "all checks passed" in run.stdout
