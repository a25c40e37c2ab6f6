"""Python binding of the MI355X intra backend.

Mirrors the frame part of the reference binding (python/topos_codec/
topos_binding.py: ToposCodec.encode :3097, .decode :3308, .decode_batch :3462)
over the C ABI in include/topos_codec.h, plus the device-resident batch calls
of include/topos_gpu.h. The library is loaded from this directory; there is no
fallback: without libtopos_codec_amd.so (or without a GPU) calls fail loudly.
"""
import ctypes
import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load_abi():
    spec = importlib.util.spec_from_file_location("topos_codec_amd_abi", os.path.join(_HERE, "abi.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


abi = _load_abi()
LIB_PATH = os.path.join(_HERE, "libtopos_codec_amd.so")
SYNTH_PATH = os.path.join(_HERE, "libtopos_synth.so")


class ToposCodecError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise ToposCodecError(abi.TC_ERR_NOT_IMPLEMENTED,
                              f"{path} is not built (run topos-codec_amd/build.py)")
    lib = ctypes.CDLL(path)
    u8p, vp = ctypes.POINTER(ctypes.c_uint8), ctypes.c_void_p
    lib.tc_last_error.restype = ctypes.c_char_p
    lib.tc_status_message.restype = ctypes.c_char_p
    lib.tc_frame_packet_bound.restype = ctypes.c_size_t
    lib.tc_frame_packet_bound.argtypes = [vp]
    lib.tc_frame_config_validate.argtypes = [vp]
    lib.tc_frame_encode.argtypes = [vp, vp, vp, ctypes.c_size_t, vp]
    lib.tc_frame_decode.argtypes = [vp, ctypes.c_size_t, vp, vp, vp]
    lib.tc_frame_decode_batch.argtypes = [vp, ctypes.c_uint32, vp, vp, vp]
    lib.tc_frame_plane_geometry.argtypes = [vp, ctypes.c_uint32, vp, vp]
    lib.tcx_dev_alloc.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t]
    lib.tcx_dev_free.argtypes = [ctypes.c_uint64]
    lib.tcx_h2d.argtypes = [ctypes.c_uint64, vp, ctypes.c_size_t]
    lib.tcx_d2h.argtypes = [vp, ctypes.c_uint64, ctypes.c_size_t]
    lib.tcx_encode_batch_dev.argtypes = [vp, ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp]
    lib.tcx_decode_batch_dev.argtypes = [ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp]
    lib.tcx_decode_batch_dev_noalloc.argtypes = [ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp]
    lib.tcx_copy.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t]
    lib.tcx_serve_step.argtypes = [vp, vp]
    lib.tcx_set_timing.argtypes = [ctypes.c_int]
    lib.tcx_get_timing.argtypes = [vp]
    lib.tcx_get_timing_pairs.argtypes = [vp]
    lib.tcx_get_timing_pairs.restype = None
    lib.tcx_device_info.argtypes = [vp, vp]
    lib.tcx_device_info.restype = ctypes.c_int32
    lib.tcx_autotune_batch.argtypes = [vp, ctypes.c_uint32, vp, vp, vp, vp, ctypes.c_uint32, vp]
    lib.tcx_tune_force_tiles.argtypes = [ctypes.c_uint32]
    lib.tcx_tune_force_tiles.restype = None
    lib.tcx_tune_force.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.tcx_tune_force.restype = None
    lib.tcx_tune_reset.argtypes = []
    lib.tcx_tune_reset.restype = None
    lib.tcx_set_graphs.argtypes = [ctypes.c_int]
    lib.tcx_set_graphs.restype = None
    lib.tcx_graph_stats.argtypes = [vp, vp]
    lib.tcx_graph_stats.restype = None
    lib.tcx_now_ns.restype = ctypes.c_int64
    lib.tcx_server_create.argtypes = [vp, vp]
    lib.tcx_server_submit.argtypes = [vp, vp, vp]
    lib.tcx_server_poll.argtypes = [vp, vp, ctypes.c_uint32, vp, ctypes.c_int64]
    lib.tcx_server_release.argtypes = [vp, ctypes.c_uint64]
    lib.tcx_server_get_stats.argtypes = [vp, vp]
    lib.tcx_server_destroy.argtypes = [vp]
    lib.tcx_server_destroy.restype = None
    lib.tc_image_status_message.restype = ctypes.c_char_p
    lib.tc_image_status_message.argtypes = [ctypes.c_int32]
    lib.tc_image_probe.argtypes = [vp, vp]
    lib.tc_image_validate.argtypes = [vp, ctypes.c_uint32, vp]
    lib.tc_image_decode.argtypes = [vp, vp, vp]
    lib.tc_image_decode_preview.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp]
    lib.tc_image_query_decode_buffer.argtypes = [vp, ctypes.c_uint32, vp, vp]
    for name in ("tc_convert_packed_rgb", "tc_convert_yuv444_to_422", "tcx_convert_packed_rgb_dev",
                 "tcx_convert_yuv444_to_422_dev"):
        getattr(lib, name).argtypes = [vp]
        getattr(lib, name).restype = ctypes.c_int32
    lib.tcx_last_convert_ms.restype = ctypes.c_float
    lib.tc_frame_encode_sized.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint8, ctypes.c_uint8, vp, vp,
                                          ctypes.c_size_t, vp]
    lib.tcx_encode_sized_batch_dev.argtypes = [vp, ctypes.c_uint32]
    lib.tc_frame_decode_request.argtypes = [vp, ctypes.c_size_t, vp, vp, vp, vp]
    lib.tc_frame_decode_reduced.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint32, vp, vp, vp]
    lib.tc_frame_decode_scaled.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    lib.tc_frame_decode_batch_reduced.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    lib.tc_decode_scale_dimensions.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, vp, vp]
    lib.tcx_decode_batch_dev_request.argtypes = [ctypes.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.tcx_sized_reset.argtypes = []
    lib.tcx_sized_reset.restype = None
    lib.tcx_sized_set_mode.argtypes = [ctypes.c_int32]
    lib.tcx_sized_set_mode.restype = None
    lib.tcx_sized_set_seed.argtypes = [ctypes.c_int32, ctypes.c_int32]
    lib.tcx_sized_set_seed.restype = None
    movie_signatures(lib)
    lib.tcx_movie_decode_batch_dev.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                               ctypes.c_size_t, vp, vp, vp, vp]
    del u8p
    return lib


def movie_signatures(lib):
    """ctypes signatures of the MOV entry points (the same C ABI as the reference's)."""
    vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    lib.tc_last_error.restype = ctypes.c_char_p
    lib.tc_mux_create.argtypes = [vp, vp, vp]
    lib.tc_mux_add_packet.argtypes = [vp, vp, sz, u64, u32]
    lib.tc_mux_set_alpha_budget.argtypes = [vp, vp]
    lib.tc_mux_set_movie_meta.argtypes = [vp, vp]
    lib.tc_mux_set_raw_meta.argtypes = [vp, vp]
    lib.tc_mux_finish.argtypes = [vp]
    lib.tc_mux_free.argtypes = [vp]
    lib.tc_mux_free.restype = None
    lib.tc_movie_faststart.argtypes = [vp, vp]
    lib.tc_movie_open.argtypes = [vp, vp]
    lib.tc_movie_open_fd.argtypes = [ctypes.c_int, u64, vp]
    lib.tc_movie_info.argtypes = [vp, vp]
    lib.tc_movie_alpha_budget.argtypes = [vp, vp]
    lib.tc_movie_read_raw_meta.argtypes = [vp, vp]
    lib.tc_movie_packet.argtypes = [vp, u32, vp, sz, vp]
    lib.tc_movie_packet_batch.argtypes = [vp, u32, u32, vp, sz, vp, vp, vp]
    lib.tc_movie_packet_pts.argtypes = [vp, u32, vp, vp]
    lib.tc_movie_packet_sync.argtypes = [vp, u32, vp]
    lib.tc_movie_prev_sync.argtypes = [vp, u32, vp]
    lib.tc_movie_close.argtypes = [vp]
    lib.tc_movie_close.restype = None


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class ToposCodec:
    """Frame encode / decode through the GPU backend (host buffers in, host buffers out)."""

    def __init__(self, path=LIB_PATH):
        self._lib = load_library(path)

    def last_error(self):
        return self._lib.tc_last_error().decode(errors="replace")

    def _check(self, status, what):
        if status < 0:
            raise ToposCodecError(status, f"{what}: {self.last_error()}")
        return status

    def validate(self, config):
        return int(self._lib.tc_frame_config_validate(ctypes.byref(config)))

    def packet_bound(self, config):
        return int(self._lib.tc_frame_packet_bound(ctypes.byref(config)))

    def encode(self, config, planes, stats_out=None):
        """Encode one frame; planes are C-contiguous uint16 arrays in ABI plane order."""
        inp = abi.FrameInput()
        inp.struct_size = ctypes.sizeof(abi.FrameInput)
        inp.abi_version = abi.TOPOS_CODEC_ABI_VERSION
        keep = []
        for i, p in enumerate(planes):
            a = np.ascontiguousarray(p, dtype=np.uint16)
            keep.append(a)
            inp.planes[i] = a.ctypes.data
        need = self.packet_bound(config)
        out = np.empty(max(need, 1), np.uint8)
        st = stats_out if stats_out is not None else abi.FrameStats()
        status = int(self._lib.tc_frame_encode(ctypes.byref(config), ctypes.byref(inp), _ptr(out),
                                               ctypes.c_size_t(need), ctypes.byref(st)))
        self._check(status, "tc_frame_encode")
        return out[: st.packet_size].tobytes()

    def encode_status(self, config, planes, out_cap=None):
        """Raw call: returns (status, packet bytes or b'', stats)."""
        inp = abi.FrameInput()
        inp.struct_size = ctypes.sizeof(abi.FrameInput)
        inp.abi_version = abi.TOPOS_CODEC_ABI_VERSION
        keep = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
        for i, a in enumerate(keep):
            inp.planes[i] = a.ctypes.data
        cap = self.packet_bound(config) if out_cap is None else out_cap
        out = np.empty(max(cap, 1), np.uint8)
        st = abi.FrameStats()
        status = int(self._lib.tc_frame_encode(ctypes.byref(config), ctypes.byref(inp), _ptr(out),
                                               ctypes.c_size_t(cap), ctypes.byref(st)))
        pkt = out[: st.packet_size].tobytes() if status == 0 else b""
        return status, pkt, st

    def encode_sized(self, config, planes, target_bytes, qp_min=0, qp_max=abi.TC_QP_MAX, out_cap=None):
        """Target-size encode (reference topos_binding.py:3235 ToposCodec.encode_sized):
        returns (packet, stats, qp_used); raises on a failed status."""
        st, pkt, stats, q = self.encode_sized_status(config, planes, target_bytes, qp_min, qp_max, out_cap)
        self._check(st, "tc_frame_encode_sized")
        return pkt, stats, q

    def encode_sized_status(self, config, planes, target_bytes, qp_min=0, qp_max=abi.TC_QP_MAX, out_cap=None):
        inp = abi.FrameInput()
        inp.struct_size = ctypes.sizeof(abi.FrameInput)
        inp.abi_version = abi.TOPOS_CODEC_ABI_VERSION
        keep = [np.ascontiguousarray(p, dtype=np.uint16) for p in planes]
        for i, p in enumerate(keep):
            inp.planes[i] = p.ctypes.data
        cap = self.packet_bound(config) if out_cap is None else out_cap
        out = np.empty(max(cap, 1), np.uint8)
        stats = abi.FrameStats()
        q = ctypes.c_uint8(0)
        st = int(self._lib.tc_frame_encode_sized(ctypes.byref(config), ctypes.byref(inp), target_bytes, qp_min,
                                                 qp_max, ctypes.byref(q), _ptr(out) if cap else None, cap,
                                                 ctypes.byref(stats)))
        pkt = out[: stats.packet_size].tobytes() if st == 0 else b""
        return st, pkt, stats, q.value

    def sized_reset(self):
        self._lib.tcx_sized_reset()

    def sized_set_mode(self, mode):
        self._lib.tcx_sized_set_mode(mode)

    def query(self, packet):
        info = abi.FrameOutput()
        status = int(self._lib.tc_frame_decode(packet, ctypes.c_size_t(len(packet)), None, None,
                                               ctypes.byref(info)))
        return status, info

    def decode(self, packet, strict=True):
        """Decode one frame; returns (status, [planes], info). Concealment is a warning."""
        status, info = self.query(packet)
        if status != 0:
            if strict:
                self._check(status, "tc_frame_decode(query)")
            return status, [], info
        dims = abi.plane_dims(info.visible_width, info.visible_height, info.pixel_format, info.alpha_mode)
        outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
        ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
        status = int(self._lib.tc_frame_decode(packet, ctypes.c_size_t(len(packet)), ptrs, None,
                                               ctypes.byref(info)))
        if strict:
            self._check(status, "tc_frame_decode")
        return status, outs, info

    def decode_batch(self, packets):
        n = len(packets)
        infos = (abi.FrameOutput * n)()
        pk = (abi.BatchPacket * n)()
        bufs = [np.frombuffer(p, np.uint8) for p in packets]
        for i, b in enumerate(bufs):
            pk[i].data = b.ctypes.data
            pk[i].size = len(packets[i])
        status = int(self._lib.tc_frame_decode_batch(pk, n, None, None, infos))
        self._check(status, "tc_frame_decode_batch(query)")
        outs, ptrs = [], (ctypes.c_void_p * (4 * n))()
        for i in range(n):
            inf = infos[i]
            dims = abi.plane_dims(inf.visible_width, inf.visible_height, inf.pixel_format, inf.alpha_mode)
            o = [np.zeros((h, w), np.uint16) for (w, h) in dims]
            outs.append(o)
            for p, a in enumerate(o):
                ptrs[4 * i + p] = a.ctypes.data
        status = int(self._lib.tc_frame_decode_batch(pk, n, ptrs, None, infos))
        return status, outs, infos


    def convert_packed_rgb(self, src, width, height, *, fmt=abi.CVT_BGR24, out_mode=abi.CVT_OUT_YUV422,
                           bit_depth=10, kr=0.2126, kb=0.0722, full_range=False, alpha_shift=0,
                           alpha_opaque=False, src_stride=0, max_workers=1, planes_out=None):
        """Packed RGB(A) rows -> list of '<u2' planes [Y, U, V(, A)] or [G, B, R(, A)]
        (reference topos_binding.py:3548 ToposCodec.convert_packed_rgb: same
        arguments, plane shapes and checks). src is an (h, w * channels) C-contiguous
        uint8 or uint16 array; 4:2:2 chroma planes are ceil(w / 2) wide."""
        bad = abi.TC_ERR_INVALID_ARGUMENT
        if src is None or not src.flags.c_contiguous:
            raise ToposCodecError(bad, "convert_packed_rgb: src must be a C-contiguous array")
        if src.dtype == np.uint8:
            expect = (abi.CVT_BGR24, abi.CVT_RGB24, abi.CVT_BGRA32, abi.CVT_RGBA32)
        elif src.dtype == np.uint16:
            expect = (abi.CVT_BGR48, abi.CVT_RGB48, abi.CVT_BGRA64, abi.CVT_RGBA64)
        else:
            raise ToposCodecError(bad, f"convert_packed_rgb: dtype {src.dtype} not supported (u8/u16)")
        if fmt not in expect:
            raise ToposCodecError(bad, f"convert_packed_rgb: fmt {fmt} does not match dtype {src.dtype}")
        nch = abi.cvt_channels(fmt)
        if src.ndim != 2 or src.shape[1] != width * nch:
            raise ToposCodecError(bad, f"convert_packed_rgb: src shape {src.shape} != (h, w*{nch})")
        if src.shape[0] != height:
            raise ToposCodecError(bad, f"convert_packed_rgb: src height {src.shape[0]} != {height}")
        has_alpha = bool(alpha_opaque) or nch == 4
        shapes = abi.cvt_plane_shapes(width, height, out_mode, has_alpha)
        if planes_out is not None:
            if len(planes_out) != len(shapes) or any(
                    p.dtype != np.dtype("<u2") or p.shape != s or not p.flags.c_contiguous
                    for p, s in zip(planes_out, shapes)):
                raise ToposCodecError(bad, "convert_packed_rgb: planes_out shape/dtype mismatch")
            planes = list(planes_out)
        else:
            planes = [np.empty(s, dtype="<u2") for s in shapes]
        cp = abi.CvtParams()
        cp.struct_size = ctypes.sizeof(abi.CvtParams)
        cp.format, cp.width, cp.height, cp.out_mode, cp.bit_depth = fmt, width, height, out_mode, bit_depth
        cp.kr, cp.kb = kr, kb
        cp.full_range, cp.alpha_shift, cp.alpha_opaque = int(bool(full_range)), alpha_shift, int(bool(alpha_opaque))
        cp.src = src.ctypes.data
        cp.src_stride = src_stride if src_stride else width * nch * src.itemsize
        cp.y_out, cp.y_stride = planes[0].ctypes.data, width
        cp.u_out, cp.v_out = planes[1].ctypes.data, planes[2].ctypes.data
        if has_alpha:
            cp.a_out, cp.a_stride = planes[3].ctypes.data, width
        cp.max_workers = max_workers
        self._check(int(self._lib.tc_convert_packed_rgb(ctypes.byref(cp))), "tc_convert_packed_rgb")
        return planes

    def convert_yuv444_to_422(self, u_in, v_in, *, bit_depth=16, max_workers=1, planes_out=None):
        """4:4:4 uint16 U/V planes -> 4:2:2 ((a + b + 1) >> 1 of horizontal pairs)
        (reference topos_binding.py:3652 ToposCodec.convert_yuv444_to_422)."""
        bad = abi.TC_ERR_INVALID_ARGUMENT
        u, v = np.asarray(u_in), np.asarray(v_in)
        if u.ndim != 2 or v.ndim != 2 or u.shape != v.shape or u.dtype != np.dtype("<u2") or \
                v.dtype != np.dtype("<u2"):
            raise ToposCodecError(bad, "convert_yuv444_to_422: inputs must be 2-D uint16 planes of one shape")
        height, width = int(u.shape[0]), int(u.shape[1])
        if width <= 0 or height <= 0:
            raise ToposCodecError(bad, "convert_yuv444_to_422: empty planes")
        if int(bit_depth) not in (10, 12, 16):
            raise ToposCodecError(bad, f"convert_yuv444_to_422: bit_depth {bit_depth!r} not supported")
        shape = (height, (width + 1) // 2)
        if planes_out is None:
            u_out, v_out = np.empty(shape, dtype="<u2"), np.empty(shape, dtype="<u2")
        else:
            if len(planes_out) != 2 or any(p.dtype != np.dtype("<u2") or p.shape != shape or
                                           not p.flags.c_contiguous for p in planes_out):
                raise ToposCodecError(bad, "convert_yuv444_to_422: planes_out shape/dtype mismatch")
            u_out, v_out = planes_out
        if not (u.flags.c_contiguous and v.flags.c_contiguous):
            raise ToposCodecError(bad, "convert_yuv444_to_422: planes must be C-contiguous")
        pp = abi.Planar422Params()
        pp.struct_size = ctypes.sizeof(abi.Planar422Params)
        pp.width, pp.height, pp.bit_depth = width, height, int(bit_depth)
        pp.u_in, pp.u_in_stride = u.ctypes.data, u.strides[0] // 2
        pp.v_in, pp.v_in_stride = v.ctypes.data, v.strides[0] // 2
        pp.u_out, pp.u_out_stride = u_out.ctypes.data, u_out.strides[0] // 2
        pp.v_out, pp.v_out_stride = v_out.ctypes.data, v_out.strides[0] // 2
        pp.max_workers = max(1, min(64, int(max_workers)))
        self._check(int(self._lib.tc_convert_yuv444_to_422(ctypes.byref(pp))), "tc_convert_yuv444_to_422")
        return [u_out, v_out]


class ToposImageError(RuntimeError):
    def __init__(self, status, detail=""):
        super().__init__(f"[{status}] {detail}")
        self.status = status


class MemoryIo:
    """topos_io over bytes held in memory (read callback only)."""

    def __init__(self, data):
        self.data = bytes(data)
        self.last_error = ""

        def _read(_ctx, offset, buf, length):
            if offset + length > len(self.data):
                self.last_error = f"short read at {offset}+{length}"
                return abi.TC_ERR_IO
            ctypes.memmove(buf, self.data[offset:offset + length], length)
            return abi.TC_OK

        self._cb = abi.IO_READ_FN(_read)
        self.io = abi.TposIo()
        self.io.struct_size = ctypes.sizeof(abi.TposIo)
        self.io.abi_version = abi.TOPOS_CODEC_ABI_VERSION
        self.io.read = self._cb
        self.io.length = len(self.data)

    def ref(self):
        return ctypes.cast(ctypes.byref(self.io), ctypes.c_void_p)


class MemorySink:
    """topos_io sink appending to a bytearray (write + seek_write)."""

    def __init__(self):
        self.data = bytearray()

        def _write(_ctx, buf, length):
            self.data += ctypes.string_at(buf, length)
            return abi.TC_OK

        def _seek_write(_ctx, offset, buf, length):
            if offset + length > len(self.data):
                return abi.TC_ERR_IO
            self.data[offset:offset + length] = ctypes.string_at(buf, length)
            return abi.TC_OK

        self._w = abi.IO_WRITE_FN(_write)
        self._s = abi.IO_SEEK_WRITE_FN(_seek_write)
        self.io = abi.sized(abi.TposIo, write=self._w, seek_write=self._s)

    def ref(self):
        return ctypes.cast(ctypes.byref(self.io), ctypes.c_void_p)


class ToposMovieFile:
    """MOV reader (reference python/topos_codec/topos_binding.py:1954-2455):
    a file path (opened through tc_movie_open_fd) or the file's bytes."""

    def __init__(self, lib, src):
        self._lib = lib
        self._fd = None
        self._src = None
        h = ctypes.c_void_p()
        if isinstance(src, (bytes, bytearray, memoryview)):
            self._src = MemoryIo(src)
            st = int(lib.tc_movie_open(self._src.ref(), ctypes.byref(h)))
        else:
            self._fd = os.open(src, os.O_RDONLY)
            st = int(lib.tc_movie_open_fd(self._fd, os.fstat(self._fd).st_size, ctypes.byref(h)))
        if st != abi.TC_OK:
            self._close_fd()
            raise ToposCodecError(st, lib.tc_last_error().decode(errors="replace"))
        self.handle = h

    def _close_fd(self):
        if self._fd is not None:
            os.close(self._fd)
            self._fd = None

    def _check(self, st, what):
        if st != abi.TC_OK:
            raise ToposCodecError(st, f"{what}: " + self._lib.tc_last_error().decode(errors="replace"))

    def info(self):
        info = abi.sized(abi.MovieInfo)
        self._check(int(self._lib.tc_movie_info(self.handle, ctypes.byref(info))), "tc_movie_info")
        return info

    def packet(self, index):
        need = ctypes.c_size_t(0)
        st = int(self._lib.tc_movie_packet(self.handle, index, None, 0, ctypes.byref(need)))
        if st != abi.TC_ERR_BUFFER_TOO_SMALL:
            self._check(st, "tc_movie_packet probe")
        buf = ctypes.create_string_buffer(need.value)
        self._check(int(self._lib.tc_movie_packet(self.handle, index, buf, need.value, None)), "tc_movie_packet")
        return buf.raw

    def packets_into_batch(self, start, count):
        """(arena bytes, offsets, sizes) of samples [start, start + count)."""
        need = ctypes.c_size_t(0)
        st = int(self._lib.tc_movie_packet_batch(self.handle, start, count, None, 0, None, None, ctypes.byref(need)))
        if st != abi.TC_ERR_BUFFER_TOO_SMALL:
            self._check(st, "tc_movie_packet_batch probe")
        offs = (ctypes.c_size_t * count)()
        sizes = (ctypes.c_size_t * count)()
        buf = ctypes.create_string_buffer(max(need.value, 1))
        self._check(int(self._lib.tc_movie_packet_batch(self.handle, start, count, buf, need.value, offs, sizes,
                                                         None)), "tc_movie_packet_batch")
        return buf.raw[:need.value], list(offs), list(sizes)

    def pts(self, index):
        pts, dur = ctypes.c_uint64(0), ctypes.c_uint32(0)
        self._check(int(self._lib.tc_movie_packet_pts(self.handle, index, ctypes.byref(pts), ctypes.byref(dur))),
                    "tc_movie_packet_pts")
        return pts.value, dur.value

    def is_sync(self, index):
        v = ctypes.c_uint8(0)
        self._check(int(self._lib.tc_movie_packet_sync(self.handle, index, ctypes.byref(v))), "tc_movie_packet_sync")
        return bool(v.value)

    def prev_sync(self, index):
        v = ctypes.c_uint32(0)
        self._check(int(self._lib.tc_movie_prev_sync(self.handle, index, ctypes.byref(v))), "tc_movie_prev_sync")
        return v.value

    def alpha_budget(self):
        """The tpcB record, or None when the file carries none."""
        bi = abi.sized(abi.AlphaBudgetInfo)
        st = int(self._lib.tc_movie_alpha_budget(self.handle, ctypes.byref(bi)))
        if st == abi.TC_ERR_STATE:
            return None
        self._check(st, "tc_movie_alpha_budget")
        return bi

    def read_raw_meta(self):
        rm = abi.sized(abi.RawMeta)
        self._check(int(self._lib.tc_movie_read_raw_meta(self.handle, ctypes.byref(rm))), "tc_movie_read_raw_meta")
        return bytes(rm.payload[:rm.payload_size])

    def close(self):
        if self.handle:
            self._lib.tc_movie_close(self.handle)
            self.handle = None
        self._close_fd()

    def __enter__(self):
        return self

    def __exit__(self, *_exc):
        self.close()


class ToposMuxFile:
    """MOV writer (reference python/topos_codec/topos_binding.py:2457-2715):
    to a file path, or to memory when path is None (the bytes in .data after
    finish)."""

    def __init__(self, lib, path, config):
        self._lib = lib
        self._path = path
        self._sink = MemorySink()
        h = ctypes.c_void_p()
        st = int(lib.tc_mux_create(ctypes.byref(config), self._sink.ref(), ctypes.byref(h)))
        if st != abi.TC_OK:
            raise ToposCodecError(st, lib.tc_last_error().decode(errors="replace"))
        self.handle = h

    def _check(self, st, what):
        if st != abi.TC_OK:
            raise ToposCodecError(st, f"{what}: " + self._lib.tc_last_error().decode(errors="replace"))

    def add_packet(self, packet, pts, dur=1):
        self._check(int(self._lib.tc_mux_add_packet(self.handle, bytes(packet), len(packet), pts, dur)),
                    "tc_mux_add_packet")

    def set_alpha_budget(self, **kw):
        self._check(int(self._lib.tc_mux_set_alpha_budget(self.handle, ctypes.byref(abi.sized(abi.AlphaBudgetInfo,
                                                                                            **kw)))),
                    "tc_mux_set_alpha_budget")

    def set_movie_meta(self, tier_id=0, vendor="", label=""):
        m = abi.sized(abi.MovieMeta, tier_id=tier_id, vendor=vendor.encode(), label=label.encode())
        self._check(int(self._lib.tc_mux_set_movie_meta(self.handle, ctypes.byref(m))), "tc_mux_set_movie_meta")

    def set_raw_meta(self, payload):
        rm = abi.sized(abi.RawMeta, payload_size=len(payload))
        ctypes.memmove(rm.payload, bytes(payload), len(payload))
        self._check(int(self._lib.tc_mux_set_raw_meta(self.handle, ctypes.byref(rm))), "tc_mux_set_raw_meta")

    @property
    def data(self):
        return bytes(self._sink.data)

    def finish(self):
        self._check(int(self._lib.tc_mux_finish(self.handle)), "tc_mux_finish")
        if self._path is not None:
            with open(self._path, "wb") as f:
                f.write(self._sink.data)

    def close(self):
        if self.handle:
            self._lib.tc_mux_free(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *_exc):
        self.close()


def movie_faststart(lib, data):
    """tc_movie_faststart over bytes: the moov-first file."""
    src = MemoryIo(data)
    dst = MemorySink()
    st = int(lib.tc_movie_faststart(src.ref(), dst.ref()))
    if st != abi.TC_OK:
        raise ToposCodecError(st, lib.tc_last_error().decode(errors="replace"))
    return bytes(dst.data)


def _image_source(src):
    if isinstance(src, (bytes, bytearray, memoryview)):
        return MemoryIo(src)
    with open(src, "rb") as f:
        return MemoryIo(f.read())


class ToposImageCodec:
    """The .toos read path (reference python/topos_codec/topos_image_binding.py:
    ToposImageCodec.probe :463, .validate :476, .decode :523), over
    include/topos_image.h; a source is a file path or the file's bytes."""

    def __init__(self, path=LIB_PATH):
        self._lib = load_library(path)

    def status_message(self, status):
        return self._lib.tc_image_status_message(status).decode(errors="replace")

    def _detail(self, status):
        return f"{self.status_message(status)} ({self._lib.tc_last_error().decode(errors='replace')})"

    def probe_status(self, src):
        io, info = _image_source(src), abi.ImageInfo()
        rc = self._lib.tc_image_probe(io.ref(), ctypes.byref(info))
        return rc, info

    def validate_status(self, src, deep=False):
        io, info = _image_source(src), abi.ImageInfo()
        rc = self._lib.tc_image_validate(io.ref(), abi.TC_IMG_VALIDATE_DEEP if deep else 0, ctypes.byref(info))
        return rc, info

    def probe(self, src):
        rc, info = self.probe_status(src)
        if rc != abi.TC_OK:
            raise ToposImageError(rc, self._detail(rc))
        return info

    def validate(self, src, deep=False):
        rc, info = self.validate_status(src, deep)
        if rc != abi.TC_OK:
            raise ToposImageError(rc, self._detail(rc))
        return info

    def decode_status(self, src):
        """(status, planes, frame output); planes are allocated from the probe."""
        io = _image_source(src)
        info = abi.ImageInfo()
        rc = self._lib.tc_image_probe(io.ref(), ctypes.byref(info))
        if rc != abi.TC_OK:
            return rc, [], None
        views = (abi.PlaneView * abi.TC_FRAME_MAX_PLANES)()
        arrays = []
        for p in range(abi.TC_FRAME_MAX_PLANES):
            views[p].struct_size = ctypes.sizeof(abi.PlaneView)
            views[p].abi_version = abi.TOPOS_CODEC_ABI_VERSION
        for p in range(int(info.plane_count)):
            w, h = ctypes.c_uint32(0), ctypes.c_uint32(0)
            rc = self._lib.tc_image_query_decode_buffer(ctypes.byref(info), p, ctypes.byref(w), ctypes.byref(h))
            if rc != abi.TC_OK:
                return rc, [], None
            arr = np.zeros((h.value, w.value), np.uint16)
            arrays.append(arr)
            views[p].pixels = arr.ctypes.data
            views[p].stride = 0
        out = abi.FrameOutput()
        out.struct_size = ctypes.sizeof(abi.FrameOutput)
        out.abi_version = abi.TOPOS_CODEC_ABI_VERSION
        rc = self._lib.tc_image_decode(io.ref(), ctypes.cast(views, ctypes.c_void_p), ctypes.byref(out))
        return rc, arrays, out

    def decode_preview_status(self, src, target_width, target_height, arrays):
        """tc_image_decode_preview into the given uint16 arrays (one per output
        plane, target geometry): (status, frame output)."""
        io = _image_source(src)
        views = (abi.PlaneView * abi.TC_FRAME_MAX_PLANES)()
        for p in range(abi.TC_FRAME_MAX_PLANES):
            views[p].struct_size = ctypes.sizeof(abi.PlaneView)
            views[p].abi_version = abi.TOPOS_CODEC_ABI_VERSION
        for p, arr in enumerate(arrays):
            views[p].pixels = arr.ctypes.data
            views[p].stride = arr.strides[0] // 2
        out = abi.FrameOutput()
        rc = self._lib.tc_image_decode_preview(io.ref(), target_width, target_height, ctypes.cast(views, ctypes.c_void_p),
                                               ctypes.byref(out))
        return rc, out

    def decode(self, src):
        rc, arrays, out = self.decode_status(src)
        if rc not in (abi.TC_OK, abi.TC_WARN_CONCEALED):
            raise ToposImageError(rc, self._detail(rc))
        return arrays, out


class DeviceBuffer:
    """A block of HBM owned by the backend (freed on close / garbage collection)."""

    def __init__(self, lib, nbytes):
        self._lib = lib
        self.ptr = ctypes.c_uint64(0)
        self.nbytes = int(nbytes)
        st = lib.tcx_dev_alloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes))
        if st != 0:
            raise ToposCodecError(st, lib.tc_last_error().decode(errors="replace"))

    @property
    def addr(self):
        return int(self.ptr.value)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        st = self._lib.tcx_h2d(ctypes.c_uint64(self.addr), _ptr(a), ctypes.c_size_t(a.nbytes))
        if st != 0:
            raise ToposCodecError(st, self._lib.tc_last_error().decode(errors="replace"))

    def download(self, nbytes=None, dtype=np.uint8):
        n = self.nbytes if nbytes is None else int(nbytes)
        out = np.empty(n, np.uint8)
        st = self._lib.tcx_d2h(_ptr(out), ctypes.c_uint64(self.addr), ctypes.c_size_t(n))
        if st != 0:
            raise ToposCodecError(st, self._lib.tc_last_error().decode(errors="replace"))
        return out.view(dtype)

    def close(self):
        if self.ptr.value:
            self._lib.tcx_dev_free(ctypes.c_uint64(self.ptr.value))
            self.ptr = ctypes.c_uint64(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCodec(ToposCodec):
    """Batched encode / decode with inputs and outputs resident in HBM."""

    def alloc(self, nbytes):
        return DeviceBuffer(self._lib, nbytes)

    def copy(self, dst, src, nbytes):
        st = self._lib.tcx_copy(ctypes.c_uint64(dst.addr), ctypes.c_uint64(src.addr), ctypes.c_size_t(nbytes))
        if st != 0:
            raise ToposCodecError(st, self.last_error())

    def encode_sized_batch_dev(self, reqs):
        """tcx_encode_sized_batch_dev over an array of abi.TcxSizedReq (filled in place)."""
        arr = (abi.TcxSizedReq * len(reqs))(*reqs)
        st = int(self._lib.tcx_encode_sized_batch_dev(arr, len(reqs)))
        for i in range(len(reqs)):
            reqs[i] = arr[i]
        return st

    def convert_packed_rgb_dev(self, params):
        """tcx_convert_packed_rgb_dev: params is an abi.CvtParams whose pointers are device addresses."""
        params.struct_size = ctypes.sizeof(abi.CvtParams)
        return int(self._lib.tcx_convert_packed_rgb_dev(ctypes.byref(params)))

    def convert_yuv444_to_422_dev(self, params):
        params.struct_size = ctypes.sizeof(abi.Planar422Params)
        return int(self._lib.tcx_convert_yuv444_to_422_dev(ctypes.byref(params)))

    def set_timing(self, on):
        self._lib.tcx_set_timing(1 if on else 0)

    def set_graphs(self, on):
        self._lib.tcx_set_graphs(1 if on else 0)

    def graph_stats(self):
        """(replays, captures) of launch-sequence graphs so far."""
        h, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        self._lib.tcx_graph_stats(ctypes.byref(h), ctypes.byref(b))
        return h.value, b.value

    def timing(self):
        t = (ctypes.c_float * 7)()
        self._lib.tcx_get_timing(t)
        return dict(zip(["fdct", "slice_enc", "scan", "emit", "slice_dec", "idct", "deblock"], list(t)))

    def device_info(self):
        """(compute units, peak shader clock in kHz) of the device."""
        c, k = ctypes.c_uint32(0), ctypes.c_uint32(0)
        self._lib.tcx_device_info(ctypes.byref(c), ctypes.byref(k))
        return c.value, k.value

    def timing_pairs(self):
        """rANS2 symbol pairs of the last timed encode and decode sequences."""
        n = (ctypes.c_uint64 * 2)()
        self._lib.tcx_get_timing_pairs(n)
        return {"enc_pairs": int(n[0]), "dec_pairs": int(n[1])}

    @staticmethod
    def _enc_args(cfgs, plane_addrs, out_bufs, strides):
        n = len(cfgs)
        a = {"n": n, "cfgs": (abi.FrameConfig * n)(*cfgs), "planes": (ctypes.c_uint64 * (4 * n))(),
             "strides": (ctypes.c_size_t * (4 * n))(), "outs": (ctypes.c_uint64 * n)(*[b.addr for b in out_bufs]),
             "caps": (ctypes.c_size_t * n)(*[b.nbytes for b in out_bufs]), "sizes": (ctypes.c_uint32 * n)(),
             "rcs": (ctypes.c_int32 * n)(), "stats": (abi.FrameStats * n)()}
        for i, pa in enumerate(plane_addrs):
            for p, addr in enumerate(pa):
                a["planes"][4 * i + p] = addr
                if strides is not None:
                    a["strides"][4 * i + p] = strides[i][p]
        return a

    @staticmethod
    def _dec_args(host_packets, dev_packets, plane_addrs, strides):
        n = len(host_packets)
        keep = [np.frombuffer(p, np.uint8) for p in host_packets]
        a = {"n": n, "keep": keep, "hp": (ctypes.c_void_p * n)(*[k.ctypes.data for k in keep]),
             "sizes": (ctypes.c_size_t * n)(*[len(p) for p in host_packets]),
             "dp": (ctypes.c_uint64 * n)(*[b.addr for b in dev_packets]),
             "planes": (ctypes.c_uint64 * (4 * n))(), "strides": (ctypes.c_size_t * (4 * n))(),
             "infos": (abi.FrameOutput * n)(), "rcs": (ctypes.c_int32 * n)()}
        for i, pa in enumerate(plane_addrs):
            for p, addr in enumerate(pa):
                a["planes"][4 * i + p] = addr
                if strides is not None:
                    a["strides"][4 * i + p] = strides[i][p]
        return a

    def encode_batch_dev(self, cfgs, plane_addrs, out_bufs, strides=None):
        """cfgs: list of FrameConfig; plane_addrs: per frame list of 3-4 device addresses."""
        a = self._enc_args(cfgs, plane_addrs, out_bufs, strides)
        st = int(self._lib.tcx_encode_batch_dev(a["cfgs"], a["n"], a["planes"], a["strides"], a["outs"], a["caps"],
                                                a["sizes"], a["rcs"], a["stats"]))
        return st, list(a["sizes"]), list(a["rcs"]), a["stats"]

    def decode_movie_batch_dev(self, movie, start, count, arena, plane_addrs, strides=None):
        """tcx_movie_decode_batch_dev: MOV samples [start, start + count) of a
        ToposMovieFile uploaded into arena (a DeviceBuffer) and decoded into
        plane_addrs (per frame, 3-4 device addresses). Returns (status, rcs, infos)."""
        planes = (ctypes.c_uint64 * (4 * count))()
        strd = (ctypes.c_size_t * (4 * count))() if strides is not None else None
        for i, pa in enumerate(plane_addrs):
            for p, addr in enumerate(pa):
                planes[4 * i + p] = addr
                if strides is not None:
                    strd[4 * i + p] = strides[i][p]
        infos = (abi.FrameOutput * count)()
        rcs = (ctypes.c_int32 * count)()
        st = int(self._lib.tcx_movie_decode_batch_dev(movie.handle, start, count, ctypes.c_uint64(arena.addr),
                                                      arena.nbytes, planes, strd, infos, rcs))
        return st, list(rcs), infos

    def decode_batch_dev(self, host_packets, dev_packets, plane_addrs, strides=None):
        a = self._dec_args(host_packets, dev_packets, plane_addrs, strides)
        st = int(self._lib.tcx_decode_batch_dev(a["n"], a["hp"], a["sizes"], a["dp"], a["planes"], a["strides"],
                                                a["infos"], a["rcs"]))
        return st, list(a["rcs"]), a["infos"]

    def decode_batch_dev_noalloc(self, host_packets, dev_packets, plane_addrs, strides=None):
        """tcx_decode_batch_dev_noalloc: the same decode, with no allocation in the call."""
        a = self._dec_args(host_packets, dev_packets, plane_addrs, strides)
        st = int(self._lib.tcx_decode_batch_dev_noalloc(a["n"], a["hp"], a["sizes"], a["dp"], a["planes"],
                                                        a["strides"], a["infos"], a["rcs"]))
        return st, list(a["rcs"]), a["infos"]

    def autotune_batch(self, cfgs, plane_addrs, out_bufs, reps=2, strides=None):
        """Times the persistent entropy launch shapes on this batch and keeps the
        fastest; returns (status, TcxTuneReport)."""
        a = self._enc_args(cfgs, plane_addrs, out_bufs, strides)
        rep = abi.TcxTuneReport()
        st = int(self._lib.tcx_autotune_batch(a["cfgs"], a["n"], a["planes"], a["strides"], a["outs"], a["caps"],
                                              reps, ctypes.byref(rep)))
        return st, rep

    def tune_reset(self):
        self._lib.tcx_tune_reset()

    def tune_force_tiles(self, tpw):
        """Fixes the calling thread's transform tiles per workgroup (0: back to the tuned tiling)."""
        self._lib.tcx_tune_force_tiles(tpw)

    def tune_force(self, wpg, groups=0):
        """Fixes the calling thread's rANS2 launch shape (wpg 0: back to the tuned shapes)."""
        self._lib.tcx_tune_force(wpg, groups)

    def serve_step(self, enc=None, dec=None):
        """One serving step: enc = (cfgs, plane_addrs, out_bufs[, strides]) and
        dec = (host_packets, dev_packets, plane_addrs[, strides]) run concurrently.
        Returns (status, (sizes, rcs, stats) or None, (rcs, infos) or None)."""
        ea = self._enc_args(*(list(enc) + [None] * (4 - len(enc)))) if enc else None
        da = self._dec_args(*(list(dec) + [None] * (4 - len(dec)))) if dec else None
        eb = TcxEncBatch()
        db = TcxDecBatch()
        if ea:
            eb.cfgs, eb.n = ctypes.cast(ea["cfgs"], ctypes.c_void_p), ea["n"]
            eb.dev_planes, eb.dev_strides = ctypes.cast(ea["planes"], ctypes.c_void_p), ctypes.cast(ea["strides"], ctypes.c_void_p)
            eb.dev_out, eb.out_caps = ctypes.cast(ea["outs"], ctypes.c_void_p), ctypes.cast(ea["caps"], ctypes.c_void_p)
            eb.out_sizes, eb.rcs = ctypes.cast(ea["sizes"], ctypes.c_void_p), ctypes.cast(ea["rcs"], ctypes.c_void_p)
            eb.stats = ctypes.cast(ea["stats"], ctypes.c_void_p)
        if da:
            db.n, db.host_packets = da["n"], ctypes.cast(da["hp"], ctypes.c_void_p)
            db.sizes, db.dev_packets = ctypes.cast(da["sizes"], ctypes.c_void_p), ctypes.cast(da["dp"], ctypes.c_void_p)
            db.dev_planes, db.dev_strides = ctypes.cast(da["planes"], ctypes.c_void_p), ctypes.cast(da["strides"], ctypes.c_void_p)
            db.infos, db.rcs = ctypes.cast(da["infos"], ctypes.c_void_p), ctypes.cast(da["rcs"], ctypes.c_void_p)
        st = int(self._lib.tcx_serve_step(ctypes.byref(eb) if ea else None, ctypes.byref(db) if da else None))
        er = (list(ea["sizes"]), list(ea["rcs"]), ea["stats"]) if ea else None
        dr = (list(da["rcs"]), da["infos"]) if da else None
        return st, er, dr


class Server:
    """The serving loop (tcx_server_*): submit encode / decode requests as they
    arrive, poll completions, release results. Request buffers handed in as
    numpy arrays or bytes are kept alive here until the request is released."""

    def __init__(self, codec, hbm_budget, max_batch=8, max_inflight=4, gather_us=0):
        self._lib = codec._lib
        self._codec = codec
        cfg = abi.TcxServerConfig(hbm_budget, max_batch, max_inflight, gather_us)
        self._h = ctypes.c_void_p()
        st = self._lib.tcx_server_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if st != 0:
            raise ToposCodecError(st, codec.last_error())
        self._keep = {}

    def now_ns(self):
        return int(self._lib.tcx_now_ns())

    def _submit(self, req, keep):
        rid = ctypes.c_uint64(0)
        st = int(self._lib.tcx_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)))
        if st != 0:
            return st, None
        self._keep[rid.value] = keep
        return 0, rid.value

    def submit_encode(self, cfg, dev_planes=None, host_planes=None, strides=None, user=0, stats=None):
        """Returns (status, id). dev_planes: device addresses; host_planes: uint16 arrays."""
        req = abi.TcxRequest()
        req.kind, req.user, req.cfg = abi.TCX_REQ_ENCODE, user, cfg
        keep = [req]
        if host_planes is not None:
            arrs = [np.ascontiguousarray(p, dtype=np.uint16) for p in host_planes]
            keep += arrs
            for i, a in enumerate(arrs):
                req.host_planes[i] = a.ctypes.data
        else:
            for i, a in enumerate(dev_planes):
                req.dev_planes[i] = a
        if strides is not None:
            for i, v in enumerate(strides):
                req.strides[i] = v
        if stats is not None:
            req.stats = ctypes.cast(ctypes.byref(stats), ctypes.c_void_p)
            keep.append(stats)
        return self._submit(req, keep)

    def submit_decode(self, packet, dev_packet=0, user=0, info=None, dev_out_planes=None, out_strides=None):
        """Returns (status, id); packet: the host bytes (kept until release)."""
        req = abi.TcxRequest()
        req.kind, req.user = abi.TCX_REQ_DECODE, user
        buf = np.frombuffer(packet, np.uint8)
        req.packet, req.packet_size, req.dev_packet = buf.ctypes.data, len(packet), dev_packet
        keep = [req, buf, packet]
        if info is not None:
            req.info = ctypes.cast(ctypes.byref(info), ctypes.c_void_p)
            keep.append(info)
        if dev_out_planes is not None:
            for i, a in enumerate(dev_out_planes):
                req.dev_out_planes[i] = a
        if out_strides is not None:
            for i, v in enumerate(out_strides):
                req.out_strides[i] = v
        return self._submit(req, keep)

    def poll(self, max_n=64, timeout_s=None):
        """Completions as TcxCompletion structs; timeout None waits for at least one."""
        out = (abi.TcxCompletion * max_n)()
        n = ctypes.c_uint32(0)
        t = -1 if timeout_s is None else int(timeout_s * 1e9)
        st = self._lib.tcx_server_poll(self._h, out, max_n, ctypes.byref(n), ctypes.c_int64(t))
        if st != 0:
            raise ToposCodecError(st, self._codec.last_error())
        return [out[i] for i in range(n.value)]

    def release(self, rid):
        st = int(self._lib.tcx_server_release(self._h, ctypes.c_uint64(rid)))
        self._keep.pop(rid, None)
        return st

    def stats(self):
        s = abi.TcxServerStats()
        self._lib.tcx_server_get_stats(self._h, ctypes.byref(s))
        return s

    def close(self):
        if self._h:
            self._lib.tcx_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TcxEncBatch(ctypes.Structure):
    """struct tcx_enc_batch (include/topos_gpu.h)"""
    _fields_ = [("cfgs", ctypes.c_void_p), ("n", ctypes.c_uint32), ("dev_planes", ctypes.c_void_p),
                ("dev_strides", ctypes.c_void_p), ("dev_out", ctypes.c_void_p), ("out_caps", ctypes.c_void_p),
                ("out_sizes", ctypes.c_void_p), ("rcs", ctypes.c_void_p), ("stats", ctypes.c_void_p)]


class TcxDecBatch(ctypes.Structure):
    """struct tcx_dec_batch (include/topos_gpu.h)"""
    _fields_ = [("n", ctypes.c_uint32), ("host_packets", ctypes.c_void_p), ("sizes", ctypes.c_void_p),
                ("dev_packets", ctypes.c_void_p), ("dev_planes", ctypes.c_void_p), ("dev_strides", ctypes.c_void_p),
                ("infos", ctypes.c_void_p), ("rcs", ctypes.c_void_p)]


# ---------------------------------------------------------------- workloads

def load_synth(path=SYNTH_PATH):
    lib = ctypes.CDLL(path)
    for name in ("tsyn_frame422", "tsyn_frame4444", "tsyn_raw12", "tsyn_luma10", "tsyn_chroma10",
                 "tsyn_color12", "tsyn_alpha16"):
        getattr(lib, name).restype = ctypes.c_int
    return lib


def synth_frame422(syn, w, h, seed):
    cw = (w + 1) // 2
    y = np.zeros((h, w), np.uint16)
    u = np.zeros((h, cw), np.uint16)
    v = np.zeros((h, cw), np.uint16)
    syn.tsyn_frame422(ctypes.c_uint64(seed), w, h, _ptr(y), _ptr(u), _ptr(v))
    return [y, u, v]


def synth_frame4444(syn, w, h, seed, alpha=True):
    ps = [np.zeros((h, w), np.uint16) for _ in range(4 if alpha else 3)]
    syn.tsyn_frame4444(ctypes.c_uint64(seed), w, h, _ptr(ps[0]), _ptr(ps[1]), _ptr(ps[2]),
                       _ptr(ps[3]) if alpha else None)
    return ps


def synth_raw12(syn, w, h, seed):
    pw, ph = (w + 1) // 2, (h + 1) // 2
    ps = [np.zeros((ph, pw), np.uint16) for _ in range(4)]
    arr = (ctypes.c_void_p * 4)(*[p.ctypes.data for p in ps])
    syn.tsyn_raw12(ctypes.c_uint64(seed), 0, pw, ph, arr)
    return ps
