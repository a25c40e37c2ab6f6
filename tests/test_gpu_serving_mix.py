"""The serving loop on the serving workload at its stated sizes: 60 % decode
and 40 % encode requests over 1920x1080 4:2:2 LT, 3840x2160 4:2:2 HQ,
4096x2160 4444 XQ + 16-bit alpha and 4096x3072 RAW 12-bit frames (30 / 40 /
20 / 10 %), arriving on a seeded Poisson clock while earlier ones run, under
an HBM budget that holds some back. Every request's result is checked three
ways: alone (a one-frame batch), in the mix, and against the compiled
reference's committed hashes of the same frame (packets byte for byte,
decoded planes sample for sample)."""
import os
import sys
import time

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import baseline_cases as bc  # noqa: E402

# (share, fixture case selector)
MIX = [(0.3, ("uneven", 104)), (0.4, ("hq_4k", 0)), (0.2, ("xq_4096", 100)), (0.1, ("uneven", 111))]


@pytest.mark.gpu
def test_serving_mix_at_stated_sizes(pkg, abi, syn, codec):
    fx = bc.load_fixture()
    kinds = []
    for share, (group, seed) in MIX:
        c = next(c for c in bc.cases() if c[0] == group and c[4] == seed)
        e = fx[bc.case_id(c)]
        planes = bc.make_frame(pkg, syn, c)
        assert [bc.sha(p) for p in planes] == e["input_sha256"]
        cfg = bc.make_config(abi, c)
        dev_in = []
        for p in planes:
            b = codec.alloc(p.nbytes)
            b.upload(p)
            dev_in.append(b)
        # alone: one-frame batch each way, equal to the reference
        out = codec.alloc(sum(p.nbytes for p in planes) * 2 + (1 << 20))
        st, sizes, rcs, _ = codec.encode_batch_dev([cfg], [[b.addr for b in dev_in]], [out])
        assert st == 0, codec.last_error()
        pkt = out.download(sizes[0]).tobytes()
        assert bc.sha(pkt) == e["packet_sha256"]
        dec = [codec.alloc(p.nbytes) for p in planes]
        st, rcs, _ = codec.decode_batch_dev([pkt], [out], [[b.addr for b in dec]])
        assert st == 0, codec.last_error()
        got = [b.download(p.nbytes, np.uint16).reshape(p.shape) for b, p in zip(dec, planes)]
        assert [bc.sha(g) for g in got] == e["decode_sha256"]
        for b in dec:
            b.close()
        kinds.append(dict(share=share, cfg=cfg, planes=planes, dev_in=dev_in, pkt=pkt, dev_pkt=out, e=e))

    rng = np.random.default_rng(2024)
    n = 40
    which = rng.choice(len(kinds), size=n, p=[k["share"] for k in kinds])
    is_dec = rng.random(n) < 0.6
    gaps = rng.exponential(0.012, size=n)          # about 80 requests/s offered
    srv = pkg.Server(codec, 3 << 30, max_batch=8, max_inflight=2)
    try:
        submitted, done = {}, []
        t = time.perf_counter()
        for i in range(n):
            t += float(gaps[i])
            while time.perf_counter() < t:
                for comp in srv.poll(64, timeout_s=0.001):
                    done.append(comp)
            k = kinds[int(which[i])]
            if is_dec[i]:
                st, rid = srv.submit_decode(k["pkt"], dev_packet=k["dev_pkt"].addr if i % 2 else 0, user=i)
            else:
                kw = dict(dev_planes=[b.addr for b in k["dev_in"]]) if i % 2 else dict(host_planes=k["planes"])
                st, rid = srv.submit_encode(k["cfg"], user=i, **kw)
            assert st == 0, codec.last_error()
            submitted[rid] = (bool(is_dec[i]), int(which[i]))
        t_end = time.time() + 120
        while len(done) < n:
            assert time.time() < t_end, f"{len(done)} of {n} completions"
            done += srv.poll(64, timeout_s=1.0)
        for comp in done:
            d, w = submitted[comp.id]
            k = kinds[w]
            assert comp.status == 0, codec.last_error()
            if d:
                dims = [(p.shape[1], p.shape[0]) for p in k["planes"]]
                got = []
                for p, (pw, ph) in enumerate(dims):
                    assert comp.strides[p] == pw
                    a = np.empty((ph, pw), np.uint16)
                    assert codec._lib.tcx_d2h(pkg._ptr(a), comp.dev_planes[p], pw * ph * 2) == 0
                    got.append(a)
                assert [bc.sha(g) for g in got] == k["e"]["decode_sha256"]
            else:
                buf = np.empty(comp.packet_size, np.uint8)
                assert codec._lib.tcx_d2h(pkg._ptr(buf), comp.dev_packet, comp.packet_size) == 0
                assert bc.sha(buf) == k["e"]["packet_sha256"]
            assert srv.release(comp.id) == 0
        assert max(c.batch_frames for c in done) > 1       # requests shared launches
        st = srv.stats()
        assert st.submitted == st.completed == n and st.held_bytes == 0
    finally:
        srv.close()
        for k in kinds:
            for b in k["dev_in"]:
                b.close()
            k["dev_pkt"].close()
