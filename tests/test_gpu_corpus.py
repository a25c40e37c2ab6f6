"""GPU replay of the reference's robustness corpus (256 files of
native/tests/fuzz/seed_corpus_bitstream: valid seeds of every frame class,
bit flips, truncations, forged and over-long fields; plus the complete
packets of golden_bitstream_v1.bin) through the public tc_frame_decode, one
file at a time and all decodable files in one tc_frame_decode_batch: status,
slice status, concealment count and every plane equal the oracle's decode,
which the CPU suite pins to the compiled reference on the same files."""
import ctypes

import numpy as np
import pytest

from helpers import corpus_packets, orc_decode

SENTINEL = 0xA5A5


@pytest.fixture(scope="module")
def corpus(pkg):
    c = pkg.ToposCodec()
    out = []
    for name, pkt in corpus_packets():
        st, info = c.query(pkt)
        out.append((name, pkt, st, info))
    return out


@pytest.mark.gpu
def test_corpus_single_frame_decodes_match_oracle(pkg, abi, oracle, codec, corpus):
    lib = codec._lib
    n_ok = n_conc = 0
    for name, pkt, qst, qinfo in corpus:
        if qst != 0:
            info = abi.FrameOutput()
            planes = [np.full((8, 8), SENTINEL, np.uint16) for _ in range(4)]
            ptrs = (ctypes.c_void_p * 4)(*[p.ctypes.data for p in planes])
            st = lib.tc_frame_decode(pkt, ctypes.c_size_t(len(pkt)), ptrs, None, ctypes.byref(info))
            assert st == qst, name                       # rejected before any output
            assert all((p == SENTINEL).all() for p in planes), name
            continue
        dims = abi.plane_dims(qinfo.visible_width, qinfo.visible_height, qinfo.pixel_format, qinfo.alpha_mode)
        o_rc, o_planes, o_status, o_nc = orc_decode(oracle, pkt, dims)
        st, planes, info = codec.decode(pkt, strict=False)
        assert st == o_rc, name
        assert list(info.slice_status)[:info.slice_count] == o_status[:info.slice_count], name
        assert info.concealed_slices == o_nc, name
        assert all((a == b).all() for a, b in zip(planes, o_planes)), name
        n_ok += 1
        n_conc += st == abi.TC_WARN_CONCEALED
    assert n_ok >= 80 and n_conc >= 60


@pytest.mark.gpu
def test_corpus_batch_decode_matches_oracle(pkg, abi, oracle, codec, corpus):
    good = [(name, pkt) for name, pkt, st, _ in corpus if st == 0]
    st, outs, infos = codec.decode_batch([p for _, p in good])
    assert st == abi.TC_WARN_CONCEALED                   # the first damaged frame's status
    first = None
    for i, (name, pkt) in enumerate(good):
        dims = [(o.shape[1], o.shape[0]) for o in outs[i]]
        o_rc, o_planes, o_status, o_nc = orc_decode(oracle, pkt, dims)
        if first is None and o_rc != 0:
            first = o_rc
        n = infos[i].slice_count
        assert list(infos[i].slice_status)[:n] == o_status[:n], name
        assert infos[i].concealed_slices == o_nc, name
        assert all((a == b).all() for a, b in zip(outs[i], o_planes)), name
    assert first == abi.TC_WARN_CONCEALED
