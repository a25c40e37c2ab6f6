"""ctypes mirrors of the public frame ABI structs and status codes.

The layouts follow the frame section of the Topos C ABI
(native/include/topos_codec.h:33-56 status codes, :175-224 topos_frame_config,
:230-237 topos_frame_input, :240-254 topos_frame_stats, :262-285
topos_frame_output, :374-380 topos_batch_packet, :426-431 topos_plane_view,
:487-495 topos_decoder_config), so one set of Python structs drives both this
library and any other implementation of the same ABI.
"""
import ctypes

TOPOS_CODEC_ABI_VERSION = 2
TC_FRAME_MAX_PLANES = 4
TC_MAX_SLICES = 512

TC_OK = 0
TC_WARN_CONCEALED = 1
TC_ERR_INVALID_ARGUMENT = -1
TC_ERR_OUT_OF_MEMORY = -2
TC_ERR_UNSUPPORTED_VERSION = -3
TC_ERR_UNSUPPORTED_PROFILE = -4
TC_ERR_UNSUPPORTED_PIXEL_FORMAT = -5
TC_ERR_UNSUPPORTED_MATRIX = -6
TC_ERR_UNSUPPORTED_ALPHA_MODE = -7
TC_ERR_LIMIT_EXCEEDED = -8
TC_ERR_MALFORMED = -9
TC_ERR_TRUNCATED = -10
TC_ERR_CHECKSUM_MISMATCH = -11
TC_ERR_STATE = -12
TC_ERR_CANCELLED = -13
TC_ERR_IO = -14
TC_ERR_BUFFER_TOO_SMALL = -15
TC_ERR_NOT_IMPLEMENTED = -16


class FrameConfig(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("visible_width", ctypes.c_uint16),
        ("visible_height", ctypes.c_uint16),
        ("profile", ctypes.c_uint8),
        ("pixel_format", ctypes.c_uint8),
        ("bit_depth", ctypes.c_uint8),
        ("qmatrix_id", ctypes.c_uint8),
        ("qp_base", ctypes.c_uint8),
        ("qp_delta_luma", ctypes.c_int8),
        ("qp_delta_chroma", ctypes.c_int8),
        ("slice_rows", ctypes.c_uint8),
        ("alpha_mode", ctypes.c_uint8),
        ("alpha_bit_depth", ctypes.c_uint8),
        ("alpha_premultiplied", ctypes.c_uint8),
        ("color_range", ctypes.c_uint8),
        ("color_primaries", ctypes.c_uint8),
        ("color_transfer", ctypes.c_uint8),
        ("color_matrix", ctypes.c_uint8),
        ("chroma_siting", ctypes.c_uint8),
        ("sar_num", ctypes.c_uint16),
        ("sar_den", ctypes.c_uint16),
        ("reserved", ctypes.c_uint32 * 8),
    ]


class FrameInput(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("planes", ctypes.c_void_p * TC_FRAME_MAX_PLANES),
        ("strides", ctypes.c_size_t * TC_FRAME_MAX_PLANES),
        ("reserved", ctypes.c_uint32 * 8),
    ]


class FrameStats(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("packet_size", ctypes.c_uint32),
        ("color_payload_bytes", ctypes.c_uint32),
        ("alpha_payload_bytes", ctypes.c_uint32),
        ("color_header_bytes", ctypes.c_uint32),
        ("alpha_header_bytes", ctypes.c_uint32),
        ("slice_count", ctypes.c_uint16),
        ("qp_base", ctypes.c_uint8),
        ("reserved8", ctypes.c_uint8),
        ("alpha_max_abs_error", ctypes.c_uint16),
        ("reserved16", ctypes.c_uint16),
        ("reserved", ctypes.c_uint32 * 8),
    ]


class FrameOutput(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("visible_width", ctypes.c_uint16),
        ("visible_height", ctypes.c_uint16),
        ("coded_width", ctypes.c_uint16),
        ("coded_height", ctypes.c_uint16),
        ("plane_count", ctypes.c_uint8),
        ("bit_depth", ctypes.c_uint8),
        ("profile", ctypes.c_uint8),
        ("pixel_format", ctypes.c_uint8),
        ("alpha_mode", ctypes.c_uint8),
        ("alpha_bit_depth", ctypes.c_uint8),
        ("color_range", ctypes.c_uint8),
        ("color_primaries", ctypes.c_uint8),
        ("color_transfer", ctypes.c_uint8),
        ("color_matrix", ctypes.c_uint8),
        ("chroma_siting", ctypes.c_uint8),
        ("concealed_slices", ctypes.c_uint16),
        ("slice_count", ctypes.c_uint16),
        ("sar_num", ctypes.c_uint16),
        ("sar_den", ctypes.c_uint16),
        ("slice_status", ctypes.c_uint8 * TC_MAX_SLICES),
        ("reserved", ctypes.c_uint32 * 8),
    ]


class BatchPacket(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("size", ctypes.c_size_t)]


class PlaneView(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("pixels", ctypes.c_void_p),
        ("stride", ctypes.c_size_t),
    ]


# ---- .toos envelope (include/topos_image.h) ----
TOPOS_IMAGE_ABI_VERSION = 1
TC_IMG_VALIDATE_DEEP = 1
TC_IMG_ERR_BAD_MAGIC = -100
TC_IMG_ERR_BAD_PREAMBLE = -101
TC_IMG_ERR_BAD_DIRECTORY = -102
TC_IMG_ERR_UNKNOWN_CRITICAL = -103
TC_IMG_ERR_CHUNK_CONFLICT = -104
TC_IMG_ERR_METADATA_CONFLICT = -105
TC_IMG_ERR_LIMIT = -106

IO_READ_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t)
IO_WRITE_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t)
IO_SEEK_WRITE_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                    ctypes.c_size_t)


class TposIo(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("ctx", ctypes.c_void_p),
        ("read", IO_READ_FN),
        ("write", IO_WRITE_FN),
        ("seek_write", IO_SEEK_WRITE_FN),
        ("length", ctypes.c_uint64),
        ("reserved", ctypes.c_uint32 * 4),
    ]


# ---- MOV container (include/topos_codec.h, reference topos_codec.h:670-1014)
TC_MOVIE_MAX_SAMPLES = 67108864
TC_ALPHA_BUDGET_RATIO_UNSET = 0xFFFF
TC_TIER_NONE, TC_TIER_PROXY, TC_TIER_LT, TC_TIER_STANDARD, TC_TIER_HQ = 0, 1, 2, 3, 4
TC_TIER_4444, TC_TIER_4444XQ, TC_TIER_LP, TC_TIER_RAW = 5, 6, 7, 8
TC_RAW_META_MAX = 256

_MOVIE_FIELDS = [
    ("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
    ("visible_width", ctypes.c_uint16), ("visible_height", ctypes.c_uint16),
    ("profile", ctypes.c_uint8), ("pixel_format", ctypes.c_uint8), ("bit_depth", ctypes.c_uint8),
    ("qmatrix_id", ctypes.c_uint8), ("qp_base", ctypes.c_uint8),
    ("qp_delta_luma", ctypes.c_int8), ("qp_delta_chroma", ctypes.c_int8),
    ("alpha_mode", ctypes.c_uint8), ("alpha_bit_depth", ctypes.c_uint8), ("alpha_premultiplied", ctypes.c_uint8),
    ("color_range", ctypes.c_uint8), ("color_primaries", ctypes.c_uint8), ("color_transfer", ctypes.c_uint8),
    ("color_matrix", ctypes.c_uint8), ("chroma_siting", ctypes.c_uint8),
    ("sar_num", ctypes.c_uint16), ("sar_den", ctypes.c_uint16), ("timescale", ctypes.c_uint32),
]


class MovieConfig(ctypes.Structure):
    _fields_ = _MOVIE_FIELDS + [("reserved", ctypes.c_uint32 * 8)]


class MovieInfo(ctypes.Structure):
    _fields_ = _MOVIE_FIELDS + [("sample_count", ctypes.c_uint32), ("faststart", ctypes.c_uint32),
                                ("index_bytes", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 8)]


class AlphaBudgetInfo(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
                ("target_ratio_bp", ctypes.c_uint16), ("actual_ratio_bp", ctypes.c_uint16),
                ("max_abs_error", ctypes.c_uint16), ("flags", ctypes.c_uint16),
                ("frame_count", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 4)]


class MovieMeta(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
                ("tier_id", ctypes.c_uint8), ("reserved", ctypes.c_uint8 * 3),
                ("vendor", ctypes.c_char * 16), ("label", ctypes.c_char * 64)]


class RawMeta(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
                ("payload_size", ctypes.c_uint16), ("reserved", ctypes.c_uint16),
                ("payload", ctypes.c_uint8 * TC_RAW_META_MAX)]


def sized(cls, **kw):
    """A zeroed ABI struct with struct_size / abi_version set and fields from kw."""
    v = cls()
    v.struct_size = ctypes.sizeof(cls)
    v.abi_version = TOPOS_CODEC_ABI_VERSION
    for k, x in kw.items():
        setattr(v, k, x)
    return v


def movie_config(width, height, qp=24, profile=3, qmatrix=1, alpha_mode=0, alpha_bit_depth=16,
                 alpha_premultiplied=False, color_range=1, color_primaries=1, color_transfer=1, color_matrix=1,
                 chroma_siting=0, sar_num=1, sar_den=1, timescale=24000, bit_depth=10, pixel_format=0):
    """topos_movie_config with the reference binding's defaults
    (python/topos_codec/topos_binding.py:4126-4146)."""
    return sized(MovieConfig, visible_width=width, visible_height=height, qp_base=qp, profile=profile,
                 qmatrix_id=qmatrix, alpha_mode=alpha_mode, alpha_bit_depth=alpha_bit_depth if alpha_mode else 0,
                 alpha_premultiplied=int(bool(alpha_premultiplied)), color_range=color_range,
                 color_primaries=color_primaries, color_transfer=color_transfer, color_matrix=color_matrix,
                 chroma_siting=chroma_siting, sar_num=sar_num, sar_den=sar_den, timescale=timescale,
                 bit_depth=bit_depth, pixel_format=pixel_format)


def movie_config_for(frame_cfg, header, timescale=24000):
    """The movie config whose tpcC matches packets encoded with frame_cfg:
    qmatrix, qp and premultiplication from the config, every field the encoder
    normalises (depth, alpha, colour description, siting, sar) from header, the
    FrameOutput of a query decode of one of the packets."""
    names = [f for f, _ in _MOVIE_FIELDS if f not in ("struct_size", "abi_version", "timescale")]
    v = sized(MovieConfig, timescale=timescale, **{k: getattr(frame_cfg, k) for k in names})
    for k in ("visible_width", "visible_height", "profile", "pixel_format", "bit_depth", "alpha_mode",
              "alpha_bit_depth", "color_range", "color_primaries", "color_transfer", "color_matrix",
              "chroma_siting", "sar_num", "sar_den"):
        setattr(v, k, getattr(header, k))
    return v


class ImagePreamble(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
        ("file_version_major", ctypes.c_uint16), ("file_version_minor", ctypes.c_uint16),
        ("flags", ctypes.c_uint32), ("file_size", ctypes.c_uint64), ("directory_offset", ctypes.c_uint64),
        ("directory_entry_size", ctypes.c_uint32), ("directory_count", ctypes.c_uint32),
        ("primary_image_index", ctypes.c_uint32), ("compatibility_flags", ctypes.c_uint32),
        ("header_crc32", ctypes.c_uint32), ("directory_crc32", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 4),
    ]


class ImageIdsc(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
        ("idsc_version_major", ctypes.c_uint16), ("idsc_version_minor", ctypes.c_uint16),
        ("flags", ctypes.c_uint32),
        ("display_x_min", ctypes.c_int32), ("display_y_min", ctypes.c_int32),
        ("display_x_max", ctypes.c_int32), ("display_y_max", ctypes.c_int32),
        ("data_x_min", ctypes.c_int32), ("data_y_min", ctypes.c_int32),
        ("data_x_max", ctypes.c_int32), ("data_y_max", ctypes.c_int32),
        ("orientation", ctypes.c_uint32), ("pixel_aspect_num", ctypes.c_uint32),
        ("pixel_aspect_den", ctypes.c_uint32),
    ] + [(n, ctypes.c_uint8) for n in (
        "channel_model", "channel_count", "sample_kind", "valid_bit_depth", "container_bit_depth",
        "storage_layout", "subsampling", "chroma_siting", "alpha_presence", "alpha_mode", "alpha_bit_depth",
        "alpha_max_abs_err", "codec_id", "payload_major", "payload_minor", "image_profile", "codec_profile",
        "pixel_format", "color_primaries", "color_transfer", "color_matrix", "color_range", "iccp_ref",
        "ocio_ref")] + [("payload_header_crc32", ctypes.c_uint32)]


class ImageInfo(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
        ("preamble", ImagePreamble), ("idsc", ImageIdsc),
        ("pixl_offset", ctypes.c_uint64), ("pixl_size", ctypes.c_uint64),
        ("visible_width", ctypes.c_uint16), ("visible_height", ctypes.c_uint16),
    ] + [(n, ctypes.c_uint8) for n in (
        "plane_count", "has_alpha", "bit_depth", "profile", "pixel_format", "alpha_mode", "alpha_bit_depth",
        "alpha_premultiplied")] + [("reserved", ctypes.c_uint32 * 8)]


class DecoderConfig(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32),
        ("abi_version", ctypes.c_uint32),
        ("max_slice_workers", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 7),
    ]


def make_config(width, height, qp, *, profile=0, pixel_format=0, bit_depth=10,
                qmatrix_id=0, slice_rows=0, entropy_mode=8, features=0,
                alpha_mode=0, alpha_bit_depth=0, qp_delta_luma=0,
                qp_delta_chroma=0, **extra):
    """Build a zero-initialised FrameConfig with the given fields set."""
    c = FrameConfig()
    c.struct_size = ctypes.sizeof(FrameConfig)
    c.abi_version = TOPOS_CODEC_ABI_VERSION
    c.visible_width = width
    c.visible_height = height
    c.profile = profile
    c.pixel_format = pixel_format
    c.bit_depth = bit_depth
    c.qmatrix_id = qmatrix_id
    c.qp_base = qp
    c.slice_rows = slice_rows
    c.alpha_mode = alpha_mode
    c.alpha_bit_depth = alpha_bit_depth
    c.qp_delta_luma = qp_delta_luma
    c.qp_delta_chroma = qp_delta_chroma
    c.reserved[0] = entropy_mode
    c.reserved[5] = features
    for k, v in extra.items():
        setattr(c, k, v)
    return c


def config_fields(cfg):
    """Plain-dict copy of a FrameConfig (for logging / fixtures)."""
    out = {}
    for name, _ in FrameConfig._fields_:
        v = getattr(cfg, name)
        out[name] = list(v) if name == "reserved" else int(v)
    return out


def plane_dims(width, height, pixel_format, alpha_mode):
    """Visible (w, h) of each input plane, in the ABI's plane order."""
    if pixel_format == 3:
        pw, ph = (width + 1) // 2, (height + 1) // 2
        return [(pw, ph)] * 4
    cw = (width + 1) // 2 if pixel_format == 0 else width
    dims = [(width, height), (cw, height), (cw, height)]
    if alpha_mode:
        dims.append((width, height))
    return dims


# ---- serving loop (include/topos_gpu.h; no counterpart in the reference ABI)
TCX_REQ_ENCODE = 1
TCX_REQ_DECODE = 2


class TcxServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_batch", ctypes.c_uint32),
                ("max_inflight", ctypes.c_uint32), ("gather_us", ctypes.c_uint32)]


class TcxRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("user", ctypes.c_uint64),
        ("cfg", FrameConfig),
        ("dev_planes", ctypes.c_uint64 * 4),
        ("host_planes", ctypes.c_void_p * 4),
        ("strides", ctypes.c_size_t * 4),
        ("packet", ctypes.c_void_p),
        ("packet_size", ctypes.c_size_t),
        ("dev_packet", ctypes.c_uint64),
        ("dev_out", ctypes.c_uint64),
        ("dev_out_cap", ctypes.c_size_t),
        ("dev_out_planes", ctypes.c_uint64 * 4),
        ("out_strides", ctypes.c_size_t * 4),
        ("info", ctypes.c_void_p),
        ("stats", ctypes.c_void_p),
    ]


class TcxCompletion(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64), ("user", ctypes.c_uint64),
        ("kind", ctypes.c_uint32), ("status", ctypes.c_int32),
        ("packet_size", ctypes.c_uint32), ("batch_frames", ctypes.c_uint32),
        ("dev_packet", ctypes.c_uint64),
        ("dev_planes", ctypes.c_uint64 * 4),
        ("strides", ctypes.c_size_t * 4),
        ("t_submit_ns", ctypes.c_int64), ("t_admit_ns", ctypes.c_int64),
        ("t_launch_ns", ctypes.c_int64), ("t_done_ns", ctypes.c_int64),
    ]


class TcxServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
        ("batches", ctypes.c_uint64), ("batch_frames", ctypes.c_uint64),
        ("budget_bytes", ctypes.c_uint64), ("held_bytes", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64),
        ("queued", ctypes.c_uint32), ("inflight", ctypes.c_uint32),
    ]


class TcxTuneReport(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint32), ("enc_wpg", ctypes.c_uint32), ("enc_persistent", ctypes.c_uint32),
                ("dec_wpg", ctypes.c_uint32), ("dec_persistent", ctypes.c_uint32),
                ("enc_ms_best", ctypes.c_float), ("enc_ms_default", ctypes.c_float),
                ("dec_ms_best", ctypes.c_float), ("dec_ms_default", ctypes.c_float),
                ("fdct_tpw", ctypes.c_uint32), ("idct_tpw", ctypes.c_uint32),
                ("fdct_ms_best", ctypes.c_float), ("fdct_ms_default", ctypes.c_float),
                ("idct_ms_best", ctypes.c_float), ("idct_ms_default", ctypes.c_float)]


# ---- packed RGB(A) -> planar conversion (include/topos_codec.h) ----
CVT_BGR24, CVT_RGB24, CVT_BGRA32, CVT_RGBA32, CVT_BGR48, CVT_RGB48, CVT_BGRA64, CVT_RGBA64 = range(8)
CVT_OUT_YUV422, CVT_OUT_YUV444, CVT_OUT_GBR = 0, 1, 2


def cvt_channels(fmt):
    return 4 if fmt in (CVT_BGRA32, CVT_RGBA32, CVT_BGRA64, CVT_RGBA64) else 3


def cvt_sample_bytes(fmt):
    return 2 if fmt >= CVT_BGR48 else 1


class CvtParams(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("format", ctypes.c_uint32), ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32), ("out_mode", ctypes.c_uint32), ("bit_depth", ctypes.c_uint32),
        ("kr", ctypes.c_double), ("kb", ctypes.c_double),
        ("full_range", ctypes.c_uint32), ("alpha_shift", ctypes.c_uint32), ("alpha_opaque", ctypes.c_uint32),
        ("src", ctypes.c_void_p), ("src_stride", ctypes.c_uint32),
        ("y_out", ctypes.c_void_p), ("y_stride", ctypes.c_uint32),
        ("u_out", ctypes.c_void_p), ("u_stride", ctypes.c_uint32),
        ("v_out", ctypes.c_void_p), ("v_stride", ctypes.c_uint32),
        ("a_out", ctypes.c_void_p), ("a_stride", ctypes.c_uint32),
        ("max_workers", ctypes.c_uint32),
    ]


class Planar422Params(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
        ("bit_depth", ctypes.c_uint32),
        ("u_in", ctypes.c_void_p), ("u_in_stride", ctypes.c_size_t),
        ("v_in", ctypes.c_void_p), ("v_in_stride", ctypes.c_size_t),
        ("u_out", ctypes.c_void_p), ("u_out_stride", ctypes.c_size_t),
        ("v_out", ctypes.c_void_p), ("v_out_stride", ctypes.c_size_t),
        ("max_workers", ctypes.c_uint32),
    ]


def cvt_plane_shapes(width, height, out_mode, has_alpha):
    """(rows, cols) of the output planes: Y (or G), U (or B), V (or R)[, A]."""
    cw = (width + 1) // 2 if out_mode == CVT_OUT_YUV422 else width
    shapes = [(height, width), (height, cw), (height, cw)]
    if has_alpha:
        shapes.append((height, width))
    return shapes


# ---- target-size encode (include/topos_gpu.h) ----
TC_QP_MAX = 95


class TcxSizedStream(ctypes.Structure):
    _fields_ = [("qp", ctypes.c_uint32), ("valid", ctypes.c_uint32), ("bytes", ctypes.c_uint32),
                ("prev", ctypes.c_uint32), ("prev_bytes", ctypes.c_uint32)]


class TcxSizedReq(ctypes.Structure):
    _fields_ = [
        ("cfg", FrameConfig), ("dev_planes", ctypes.c_uint64 * 4), ("dev_strides", ctypes.c_size_t * 4),
        ("target_bytes", ctypes.c_uint32), ("qp_min", ctypes.c_uint8), ("qp_max", ctypes.c_uint8),
        ("stream", ctypes.POINTER(TcxSizedStream)), ("dev_out", ctypes.c_uint64), ("dev_out_cap", ctypes.c_size_t),
        ("rc", ctypes.c_int32), ("qp_used", ctypes.c_uint8), ("packet_size", ctypes.c_uint32),
        ("encodes", ctypes.c_uint32), ("stats", FrameStats),
    ]


# ---- reduced / scaled preview decode (include/topos_codec.h) ----
DECODE_SCALE_FULL, DECODE_SCALE_HALF, DECODE_SCALE_THIRD, DECODE_SCALE_QUARTER, DECODE_SCALE_EIGHTH = range(5)
DECODE_MODE_FULL, DECODE_MODE_SCALED, DECODE_MODE_REDUCED, DECODE_MODE_AUTO_2K = range(4)
DECODE_FLAG_DROP_ALPHA = 1
SCALE_DIVISOR = {0: 1, 1: 2, 2: 3, 3: 4, 4: 8}
SCALE_LIMIT = {0: 63, 1: 24, 2: 16, 3: 10, 4: 4}


class DecodeRequest(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32), ("mode", ctypes.c_uint32),
                ("scale", ctypes.c_uint32), ("target_width", ctypes.c_uint32), ("target_height", ctypes.c_uint32),
                ("quality", ctypes.c_uint32), ("memory_type", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 4)]


def make_request(mode, scale=0, target=(0, 0), flags=0):
    r = DecodeRequest()
    r.struct_size = ctypes.sizeof(DecodeRequest)
    r.abi_version = TOPOS_CODEC_ABI_VERSION
    r.mode, r.scale, r.flags = mode, scale, flags
    r.target_width, r.target_height = target
    return r


def target_plane_dims(tw, th, pixel_format, plane_count):
    """Output plane sizes of a target-size decode."""
    if pixel_format == 3:
        return [((tw + 1) // 2, (th + 1) // 2)] * plane_count
    cw = (tw + 1) // 2 if pixel_format == 0 else tw
    return [(tw, th), (cw, th), (cw, th), (tw, th)][:plane_count]


# ---- version / capability negotiation and decode surfaces (include/topos_codec.h) ----
class VersionInfo(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
                ("version_major", ctypes.c_uint32), ("version_minor", ctypes.c_uint32),
                ("version_patch", ctypes.c_uint32), ("git_commit", ctypes.c_char_p),
                ("build_target", ctypes.c_char_p), ("reserved", ctypes.c_uint32 * 4)]


class CpuFeatures(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 5)]


class CodecCapabilities(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("max_width", ctypes.c_uint32), ("max_height", ctypes.c_uint32),
                ("auto2k_max_dim", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 4)]


TC_DECODE_MEMORY_CPU = 0
TC_DECODE_MEMORY_HOST_VISIBLE_GPU = 1


class DecodeSurface(ctypes.Structure):
    _fields_ = [("struct_size", ctypes.c_uint32), ("abi_version", ctypes.c_uint32),
                ("memory_type", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("planes", PlaneView * 4)]


def plane_views(arrays):
    """Four topos_plane_view entries over numpy planes (tight strides)."""
    views = (PlaneView * 4)()
    for i, a in enumerate(arrays):
        views[i].struct_size = ctypes.sizeof(PlaneView)
        views[i].abi_version = TOPOS_CODEC_ABI_VERSION
        views[i].pixels = a.ctypes.data
        views[i].stride = a.shape[1]
    return views
