"""The benchmark workloads as fixture cases: every frame of the CPU case
(1920x1080 4:2:2 10-bit, product and plain settings), the 4K 422 HQ batch
(seeds 0-7), the 4096x2160 4444 XQ + 16-bit alpha batch and the uneven batch
(the distinct frames the benchmark generates), each with its generator
(kind, size, seed) and its codec settings. make_baseline_hashes.py writes the
compiled reference's packet and decode hashes of these frames into
baseline_ref_hashes.json; the tests and the benchmark compare against it."""
import hashlib
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "baseline_ref_hashes.json")

HQ = dict(qp=62, profile=3, qmatrix_id=0, features=3, slice_rows=8)          # 4K 422 HQ product
LT = dict(qp=68, profile=3, qmatrix_id=4, features=1, slice_rows=8)          # 1080p LT product
PROXY = dict(qp=75, profile=3, qmatrix_id=4, features=1, slice_rows=8)       # 720p Proxy product
XQ = dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1, alpha_bit_depth=16)
RAW = dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)


def cases():
    """(group, kind, w, h, seed, cfg kwargs) of every fixture frame."""
    out = []
    for qp in (62, 24):
        out.append(("cpu_1080p", "422", 1920, 1080, 0, dict(qp=qp, profile=3, qmatrix_id=4, features=1,
                                                            slice_rows=8)))
        out.append(("cpu_1080p", "422", 1920, 1080, 0, dict(qp=qp, profile=3)))   # plain C defaults
    for seed in range(8):
        out.append(("hq_4k", "422", 3840, 2160, seed, dict(HQ)))
    # the benchmark's batch frames: seeds 100 + index of the distinct frame
    for i in range(4):
        out.append(("xq_4096", "4444", 4096, 2160, 100 + i, dict(XQ)))
    seed = 100
    for kind, w, h, n, kw in (("422", 1280, 720, 4, PROXY), ("422", 1920, 1080, 4, LT), ("422", 3840, 2160, 2, HQ),
                              ("422", 8192, 4320, 1, HQ), ("raw", 4096, 3072, 2, RAW)):
        for _ in range(n):
            out.append(("uneven", kind, w, h, seed, dict(kw)))
            seed += 1
    return out


def case_id(c):
    group, kind, w, h, seed, kw = c
    extra = "-".join(f"{k}{v}" for k, v in sorted(kw.items()))
    return f"{group}-{kind}-{w}x{h}-s{seed}-{extra}"


def make_config(abi, c):
    _, _, w, h, _, kw = c
    kw = dict(kw)
    qp = kw.pop("qp")
    return abi.make_config(w, h, qp, entropy_mode=8, **kw)


def make_frame(pkg, syn, c):
    _, kind, w, h, seed, _ = c
    if kind == "422":
        return pkg.synth_frame422(syn, w, h, seed)
    if kind == "4444":
        return pkg.synth_frame4444(syn, w, h, seed, alpha=True)
    return pkg.synth_raw12(syn, w, h, seed)


def sha(a):
    return hashlib.sha256(a.tobytes() if hasattr(a, "tobytes") else bytes(a)).hexdigest()


def load_fixture():
    with open(FIXTURE) as f:
        return {e["id"]: e for e in json.load(f)["frames"]}
