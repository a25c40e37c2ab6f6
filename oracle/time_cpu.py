#!/usr/bin/env python3
"""CPU timing of the 4K 422 HQ workload (3840x2160 4:2:2 10-bit, qp 62,
qmatrix 0, deblock + refined qp table, 8 block rows per slice; frozen
generator seeds 0-7) on this machine's cores, for the CPU baseline.

  reference  the compiled reference (oracle/_ref/libtopos_ref.so, Release
             flags -O3 -DNDEBUG, built by oracle/ref.mk), timed the way its
             own benchmark does (native/tests/benchmark/tier_decode_matrix.c:
             174-190: 2 warm-ups, then >= 30 timed calls, p50 / p95 / p99) with
             TOPOS_SLICE_THREADS = cores and 1 (its slice thread pool, at most
             16 workers); plus throughput with k concurrent callers and one
             tc_frame_decode_batch of the 8 packets.
  port       the C restatement in oracle/ (single-threaded per frame),
             frame-parallel on `cores` threads (whole frames, encode + decode
             each): the baseline the GPU host can run, since the reference
             stays in this container.

The reference runs only where it was built; the port's ratio to it here is
what relates the GPU host's port number back to the reference. Writes
oracle/reference_timing.json (read by bench.py) and prints a markdown
summary.

Usage: python oracle/time_cpu.py [--iters 30] [--cores N] [--seconds 12]
"""
import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.path.join(HERE, "_ref", "libtopos_ref.so")
ORC = os.path.join(HERE, "liborc.so")
W, H, QP = 3840, 2160, 62


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def setup():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import load_pkg
    pkg = load_pkg()
    abi = pkg.abi
    cfg = abi.make_config(W, H, QP, profile=3, bit_depth=10, qmatrix_id=0, slice_rows=8, entropy_mode=8, features=3)
    syn = pkg.load_synth()
    frames = [pkg.synth_frame422(syn, W, H, s) for s in range(8)]
    return pkg, abi, cfg, frames


def pct(v):
    a = np.asarray(v) * 1e3
    return {"p50": round(float(np.percentile(a, 50)), 2), "p95": round(float(np.percentile(a, 95)), 2),
            "p99": round(float(np.percentile(a, 99)), 2)}


class RefCodec:
    """Reference calls with buffers allocated once (no per-call Python copies)."""

    def __init__(self, abi, cfg):
        self.abi, self.cfg = abi, cfg
        self.lib = ctypes.CDLL(REF)
        self.lib.tc_frame_packet_bound.restype = ctypes.c_size_t
        self.bound = self.lib.tc_frame_packet_bound(ctypes.byref(cfg))
        self.dims = abi.plane_dims(W, H, 0, 0)

    def buffers(self):
        out = np.empty(self.bound, np.uint8)
        dec = [np.empty((h, w), np.uint16) for (w, h) in self.dims]
        return out, dec

    def encode(self, planes, out):
        inp = self.abi.FrameInput()
        inp.struct_size, inp.abi_version = ctypes.sizeof(inp), 2
        for i, p in enumerate(planes):
            inp.planes[i] = p.ctypes.data
        st = self.abi.FrameStats()
        rc = self.lib.tc_frame_encode(ctypes.byref(self.cfg), ctypes.byref(inp), ctypes.c_void_p(out.ctypes.data),
                                      ctypes.c_size_t(self.bound), ctypes.byref(st))
        assert rc == 0
        return st.packet_size

    def decode(self, out, n, dec):
        ptrs = (ctypes.c_void_p * 4)(*([d.ctypes.data for d in dec] + [None]))
        info = self.abi.FrameOutput()
        rc = self.lib.tc_frame_decode(ctypes.c_void_p(out.ctypes.data), ctypes.c_size_t(n), ptrs, None,
                                      ctypes.byref(info))
        assert rc == 0


def ref_child(iters, callers):
    """Runs in a subprocess whose TOPOS_SLICE_THREADS is fixed (the reference
    reads it once per process)."""
    pkg, abi, cfg, frames = setup()
    rc = RefCodec(abi, cfg)
    out, dec = rc.buffers()
    te, td = [], []
    for i in range(iters + 2):
        t0 = time.perf_counter()
        n = rc.encode(frames[i % 8], out)
        t1 = time.perf_counter()
        rc.decode(out, n, dec)
        t2 = time.perf_counter()
        if i >= 2:
            te.append(t1 - t0)
            td.append(t2 - t1)
    res = {"encode_ms": pct(te), "decode_ms": pct(td),
           "frames_per_s_enc_dec": round(1.0 / (np.median(te) + np.median(td)), 2)}
    # k concurrent callers, each with its own buffers (the ABI is thread-safe)
    bufs = [rc.buffers() for _ in range(callers)]
    per = 4

    def work(k):
        o, d = bufs[k]
        for j in range(per):
            n = rc.encode(frames[(k + j) % 8], o)
            rc.decode(o, n, d)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(callers) as ex:
        list(ex.map(work, range(callers)))
    res["concurrent"] = {"callers": callers, "frames_per_s_enc_dec": round(callers * per / (time.perf_counter() - t0), 2)}
    # one tc_frame_decode_batch of the 8 packets
    pkts = []
    for f in frames:
        n = rc.encode(f, out)
        pkts.append(out[:n].copy())

    class BP(ctypes.Structure):
        _fields_ = [("data", ctypes.c_void_p), ("size", ctypes.c_size_t)]
    bp = (BP * 8)(*[BP(p.ctypes.data, p.size) for p in pkts])
    decs = [[np.empty((h, w), np.uint16) for (w, h) in rc.dims] for _ in range(8)]
    ptrs = (ctypes.c_void_p * 32)()
    for i, d in enumerate(decs):
        for p, a in enumerate(d):
            ptrs[4 * i + p] = a.ctypes.data
    infos = (abi.FrameOutput * 8)()
    tb = []
    for _ in range(5):
        t0 = time.perf_counter()
        assert rc.lib.tc_frame_decode_batch(bp, 8, ptrs, None, infos) == 0
        tb.append(time.perf_counter() - t0)
    res["decode_batch_8"] = {"ms": round(float(np.median(tb)) * 1e3, 2),
                             "frames_per_s": round(8 / float(np.median(tb)), 2)}
    print(json.dumps(res))


def port_throughput(abi, cfg, frames, threads, seconds):
    """The oracle restatement, frame-parallel on `threads` threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import orc_decode, orc_encode
    orc = ctypes.CDLL(ORC)
    dims = abi.plane_dims(W, H, 0, 0)
    done = [0] * threads
    t_end = time.perf_counter() + seconds

    def work(k):
        i = k
        while True:
            rc, pkt = orc_encode(orc, cfg, frames[i % len(frames)])
            assert rc == 0
            assert orc_decode(orc, pkt, dims)[0] == 0
            done[k] += 1
            i += threads
            if time.perf_counter() >= t_end:
                return
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(threads)))
    dt = time.perf_counter() - t0
    return {"threads": threads, "frames": sum(done), "seconds": round(dt, 2),
            "frames_per_s_enc_dec": round(sum(done) / dt, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--iters", type=int, default=30)
    ap.add_argument("--cores", type=int, default=len(os.sched_getaffinity(0)))
    ap.add_argument("--seconds", type=float, default=12.0)
    ap.add_argument("--ref-child", type=int, default=0)
    a = ap.parse_args()
    if a.ref_child:
        ref_child(a.iters, a.ref_child)
        return
    pkg, abi, cfg, frames = setup()
    out = {"workload": "3840x2160 YUV 4:2:2 10-bit, 422 HQ (rANS2, qm0, deblock+QPT2, 8 block rows/slice), qp 62, "
                       "frozen generator seeds 0-7",
           "cpu_model": cpu_model(), "cores": a.cores}
    if os.path.exists(REF):
        ref = {}
        for threads in sorted({min(16, a.cores), 1}, reverse=True):
            env = dict(os.environ, TOPOS_SLICE_THREADS=str(threads))
            r = subprocess.run([sys.executable, __file__, "--iters", str(a.iters), "--ref-child", str(a.cores)],
                               env=env, check=True, capture_output=True, text=True)
            ref[f"slice_threads_{threads}"] = json.loads(r.stdout.strip().splitlines()[-1])
        out["reference"] = ref
    out["port"] = {"all_cores": port_throughput(abi, cfg, frames, a.cores, a.seconds),
                   "one_core": port_throughput(abi, cfg, frames, 1, a.seconds)}
    if "reference" in out:
        best = max(v["frames_per_s_enc_dec"] for v in out["reference"].values())
        best = max([best] + [v["concurrent"]["frames_per_s_enc_dec"] for v in out["reference"].values()])
        out["reference_best_frames_per_s"] = best
        out["port_over_reference"] = round(out["port"]["all_cores"]["frames_per_s_enc_dec"] / best, 4)
    with open(os.path.join(HERE, "reference_timing.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
