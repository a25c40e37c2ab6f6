"""The product's target-size search replay (tc_sized_search.c) on the CPU:
driven in rounds of speculative candidates like the GPU driver, with byte
counts measured by the oracle encoder, it gives the oracle's sized encode
results (qp, size, status, encodes counted) in every search mode, with and
without the seed, and the seed state evolves the same over a call sequence."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import orc_encode, orc_sized
from test_sized_oracle import FRAMES, _frame

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "topos-codec_amd", "csrc", "host")


@pytest.fixture(scope="module")
def sc(tmp_path_factory):
    t = tmp_path_factory.mktemp("sized")
    out = str(t / "libsized.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra", "-I", HOST, "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "csrc", "sizedcheck.c"), "-o", out,
                    "-lm"], check=True)
    lib = ctypes.CDLL(out)
    return lib


def _table(pkg, abi, syn, oracle, spec):
    by = np.zeros(96, np.uint32)
    rcs = np.zeros(96, np.int32)
    for q in range(96):
        cfg, planes = _frame(pkg, abi, syn, spec, q)
        rc, pkt = orc_encode(oracle, cfg, planes)
        rcs[q] = rc
        by[q] = len(pkt)
    return by, rcs


def _run(sc, by, rcs, cfg, target, qmin, qmax, mode, seed=None, want=9, cap=1 << 30, exact=1):
    res = (ctypes.c_int32 * 13)()
    sd = (ctypes.c_uint32 * 5)(*seed) if seed is not None else None
    st = sc.sc_run(by.ctypes.data_as(ctypes.c_void_p), rcs.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cfg),
                   ctypes.c_uint32(target), ctypes.c_uint32(qmin), ctypes.c_uint32(qmax), ctypes.c_uint64(cap),
                   mode, exact, sd, 1, 1, ctypes.c_uint32(want), res)
    assert st == 0
    return list(res)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("want", [1, 3, 9])
def test_replay_matches_oracle(pkg, abi, syn, oracle, sc, mode, want):
    oracle.orc_sized_set_mode(mode)
    try:
        for spec in FRAMES[:3]:
            by, rcs = _table(pkg, abi, syn, oracle, spec)
            cfg, planes = _frame(pkg, abi, syn, spec, 40)
            lo, hi = int(by[95]), int(by[0])
            for target in (0, lo - 1, lo, (lo + hi) // 2, int(by[30]), int(by[60]) + 1, hi, 2 * hi):
                for (qmin, qmax, qbase) in ((0, 63, 40), (20, 95, 90), (30, 34, 10), (5, 80, 5)):
                    cfg.qp_base = qbase
                    oracle.orc_sized_reset()
                    o = orc_sized(oracle, cfg, planes, target, qmin, qmax)
                    r = _run(sc, by, rcs, cfg, target, qmin, qmax, mode, want=want)
                    assert (r[0], r[1], r[4]) == (o[0], o[1], o[4]), (spec[:3], target, qmin, qmax, r, o[:2], o[3:])
                    if o[0] == 0:
                        assert r[3] == o[3] == len(o[2])
                    if want == 1:
                        assert r[6] == r[5]                 # one measurement per round
    finally:
        oracle.orc_sized_set_mode(0)


def test_replay_seed_sequence(pkg, abi, syn, oracle, sc):
    oracle.orc_sized_set_mode(0)
    oracle.orc_sized_reset()
    seed = [0, 0, 0, 0, 0]
    rounds = []
    for spec_i, targets in ((0, (9000, 9000, 6000, 6000, 12000, 3000)), (1, (5000, 5000, 2500))):
        spec = FRAMES[spec_i]
        by, rcs = _table(pkg, abi, syn, oracle, spec)
        cfg, planes = _frame(pkg, abi, syn, spec, 40)
        for t in targets:
            o = orc_sized(oracle, cfg, planes, t, 0, 95)
            r = _run(sc, by, rcs, cfg, t, 0, 95, 0, seed=seed)
            assert (r[0], r[1], r[3], r[4]) == (o[0], o[1], o[3], o[4]), (spec_i, t, r, o[:2], o[3:])
            if r[7]:
                seed = r[8:13]
            rounds.append(r[5])
    # a repeated target is settled in one launch of candidates
    assert rounds[1] == 1 and rounds[3] == 1
