"""Helpers shared by the tests: oracle / reference calls through ctypes."""
import ctypes

import numpy as np


def orc_encode(oracle, cfg, planes, cap=None):
    ptrs = (ctypes.c_void_p * 4)(*([p.ctypes.data for p in planes] + [None] * (4 - len(planes))))
    n = ctypes.c_size_t(0)
    if cap is None:
        cap = sum(p.nbytes for p in planes) * 4 + 65536
    out = np.empty(cap, np.uint8)
    rc = oracle.orc_frame_encode(ctypes.byref(cfg), ptrs, None, ctypes.c_void_p(out.ctypes.data),
                                 ctypes.c_size_t(cap), ctypes.byref(n))
    return rc, out[: n.value].tobytes() if rc == 0 else b""


def orc_decode(oracle, packet, dims):
    outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    st = (ctypes.c_uint8 * 512)()
    nc = ctypes.c_uint32(0)
    rc = oracle.orc_frame_decode(packet, ctypes.c_size_t(len(packet)), ptrs, None, st, ctypes.byref(nc))
    return rc, outs, list(st), nc.value


def ref_encode(ref, abi, cfg, planes):
    inp = abi.FrameInput()
    inp.struct_size = ctypes.sizeof(inp)
    inp.abi_version = 2
    for i, p in enumerate(planes):
        inp.planes[i] = p.ctypes.data
    bound = ref.tc_frame_packet_bound(ctypes.byref(cfg))
    out = ctypes.create_string_buffer(max(bound, 1))
    st = abi.FrameStats()
    rc = ref.tc_frame_encode(ctypes.byref(cfg), ctypes.byref(inp), out, ctypes.c_size_t(bound), ctypes.byref(st))
    return rc, out.raw[: st.packet_size] if rc == 0 else b"", st


def ref_decode(ref, abi, packet, dims):
    outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
    info = abi.FrameOutput()
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    rc = ref.tc_frame_decode(packet, ctypes.c_size_t(len(packet)), ptrs, None, ctypes.byref(info))
    return rc, outs, info


def frame_for(pkg, syn, kind, w, h, seed):
    if kind == "422":
        return pkg.synth_frame422(syn, w, h, seed)
    if kind == "4444":
        return pkg.synth_frame4444(syn, w, h, seed, alpha=True)
    if kind == "444":
        return pkg.synth_frame4444(syn, w, h, seed, alpha=False)
    if kind == "raw":
        return pkg.synth_raw12(syn, w, h, seed)
    raise ValueError(kind)
