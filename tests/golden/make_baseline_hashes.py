#!/usr/bin/env python3
"""Writes baseline_ref_hashes.json: for every frame of baseline_cases.cases(),
the SHA-256 of each generated input plane, of the packet the compiled
reference (oracle/_ref/libtopos_ref.so, built from the reference sources by
oracle/ref.mk) encodes from it, and of each plane the reference decodes from
that packet. Run in the build container, where the reference is built; the
JSON is the committed fixture (the reference itself never leaves the
container).

Usage: python tests/golden/make_baseline_hashes.py
"""
import ctypes
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import baseline_cases as bc  # noqa: E402
from conftest import load_pkg  # noqa: E402
from helpers import ref_decode, ref_encode  # noqa: E402


def main():
    os.environ.setdefault("TOPOS_SLICE_THREADS", str(min(8, os.cpu_count() or 1)))
    pkg = load_pkg()
    abi = pkg.abi
    syn = pkg.load_synth()
    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libtopos_ref.so"))
    ref.tc_frame_packet_bound.restype = ctypes.c_size_t
    frames = []
    for c in bc.cases():
        planes = bc.make_frame(pkg, syn, c)
        cfg = bc.make_config(abi, c)
        rc, pkt, _ = ref_encode(ref, abi, cfg, planes)
        assert rc == 0, (bc.case_id(c), rc)
        dims = [(p.shape[1], p.shape[0]) for p in planes]
        drc, dplanes, info = ref_decode(ref, abi, pkt, dims)
        assert drc == 0, (bc.case_id(c), drc)
        group, kind, w, h, seed, kw = c
        frames.append({"id": bc.case_id(c), "group": group, "kind": kind, "width": w, "height": h, "seed": seed,
                       "cfg": kw, "input_sha256": [bc.sha(p) for p in planes], "packet_bytes": len(pkt),
                       "packet_sha256": bc.sha(pkt), "slice_count": int(info.slice_count),
                       "decode_sha256": [bc.sha(p) for p in dplanes]})
        print(frames[-1]["id"], len(pkt), flush=True)
    out = {"what": "compiled reference (oracle/_ref, -O3 -DNDEBUG) encodes and decodes of the frozen generator's "
                   "benchmark frames: SHA-256 of input planes (u16 little-endian, tight rows), packets and "
                   "decoded planes",
           "frames": frames}
    with open(bc.FIXTURE, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
