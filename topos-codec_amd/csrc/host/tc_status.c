/* Thread-local last-error text and status strings
 * (reference semantics: native/include/topos_codec.h:143-150,
 * native/src/common/error.c). */
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "tc_host.h"

static _Thread_local char g_last[512];

int32_t tc_fail(int32_t code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last, sizeof(g_last), fmt, ap);
    va_end(ap);
    return code;
}

void tc_clear_error(void) { g_last[0] = '\0'; }

const char* tc_last_error(void) { return g_last; }

int32_t tc_abi_version(void) { return TOPOS_CODEC_ABI_VERSION; }

const char* tc_status_message(int32_t status)
{
    switch (status) {
    case TC_OK: return "ok";
    case TC_WARN_CONCEALED: return "frame delivered with concealed slices";
    case TC_ERR_INVALID_ARGUMENT: return "invalid argument";
    case TC_ERR_OUT_OF_MEMORY: return "out of memory";
    case TC_ERR_UNSUPPORTED_VERSION: return "unsupported bitstream version";
    case TC_ERR_UNSUPPORTED_PROFILE: return "unsupported profile";
    case TC_ERR_UNSUPPORTED_PIXEL_FORMAT: return "unsupported pixel format";
    case TC_ERR_UNSUPPORTED_MATRIX: return "unsupported quantisation matrix";
    case TC_ERR_UNSUPPORTED_ALPHA_MODE: return "unsupported alpha mode";
    case TC_ERR_LIMIT_EXCEEDED: return "limit exceeded";
    case TC_ERR_MALFORMED: return "malformed bitstream";
    case TC_ERR_TRUNCATED: return "truncated bitstream";
    case TC_ERR_CHECKSUM_MISMATCH: return "checksum mismatch";
    case TC_ERR_STATE: return "invalid state";
    case TC_ERR_CANCELLED: return "cancelled";
    case TC_ERR_IO: return "i/o error";
    case TC_ERR_BUFFER_TOO_SMALL: return "buffer too small";
    case TC_ERR_NOT_IMPLEMENTED: return "not implemented";
    case TC_ERR_REFERENCE_INVALID: return "reference invalid";
    default: return "unknown status";
    }
}

/* ---- version and capability negotiation (reference native/src/version.c) ---- */

/* library version the reference reports for the same ABI (0.1.0,
 * native/include/topos_codec_version.h:5-7), so version checks of callers
 * written against it accept this library */
#define TC_LIB_MAJOR 0u
#define TC_LIB_MINOR 1u
#define TC_LIB_PATCH 0u

int32_t tc_version(topos_version_info* out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "tc_version: out is NULL"); }
    out->struct_size = (uint32_t)sizeof(topos_version_info);
    out->abi_version = TOPOS_CODEC_ABI_VERSION;
    out->version_major = TC_LIB_MAJOR;
    out->version_minor = TC_LIB_MINOR;
    out->version_patch = TC_LIB_PATCH;
    out->git_commit = "mi355x-native";
    out->build_target = "x86_64-gcc+hipcc-gfx950";
    return TC_OK;
}

int32_t tc_query_cpu_features(topos_cpu_features* out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "tc_query_cpu_features: out is NULL"); }
    uint32_t f = 0u;
#if defined(__x86_64__) || defined(__i386__)
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx2")) { f |= TOPOS_CPU_X86_AVX2; }
    if (__builtin_cpu_supports("avx512f")) { f |= TOPOS_CPU_X86_AVX512F; }
    if (__builtin_cpu_supports("fma")) { f |= TOPOS_CPU_X86_FMA; }
#elif defined(__aarch64__)
    f |= TOPOS_CPU_ARM_NEON;
#endif
    out->struct_size = (uint32_t)sizeof(topos_cpu_features);
    out->abi_version = TOPOS_CODEC_ABI_VERSION;
    out->flags = f;
    return TC_OK;
}

int32_t tc_query_capabilities(topos_codec_capabilities* out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "tc_query_capabilities: out is NULL"); }
    out->struct_size = (uint32_t)sizeof(topos_codec_capabilities);
    out->abi_version = TOPOS_CODEC_ABI_VERSION;
    /* full, fixed-ratio reduced and AUTO_2K decodes, CPU and mapped-GPU
     * surfaces (tc_decoder_decode_surface), DROP_ALPHA requests; no V7-A /
     * V7-B layouts */
    out->flags = TC_CODEC_CAP_FULL_DECODE | TC_CODEC_CAP_FIXED_REDUCED_DECODE | TC_CODEC_CAP_AUTO_2K_DECODE |
                 TC_CODEC_CAP_CPU_SURFACE_OUTPUT | TC_CODEC_CAP_HOST_VISIBLE_GPU_SURFACE |
                 TC_CODEC_CAP_DECODE_DROP_ALPHA;
    out->max_width = TC_CODEC_MAX_DIM;
    out->max_height = TC_CODEC_MAX_DIM;
    out->auto2k_max_dim = TC_CODEC_AUTO_2K_MAX_DIM;
    return TC_OK;
}

int32_t tc_query_scalable_status(topos_scalable_status* out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "tc_query_scalable_status: out is NULL"); }
    *out = TC_SCALABLE_STATUS_EXPERIMENTAL;
    return TC_OK;
}

const char* tc_scalable_status_name(topos_scalable_status status)
{
    static const char* const names[4] = { "experimental", "reference", "eligible", "default" };
    return (uint32_t)status < 4u ? names[(uint32_t)status] : "unknown";
}

/* Stream classes this library decodes: the rules of the frame header
 * (profiles 3 / 5 / 6 / 7, pixel formats 0..3 at 10 / 12 / 16 bits, RAW =
 * profile 7 = CFA at 12 or 16 bits without alpha, 4:4:4 for profiles 5 / 6,
 * 12 bits for profile 6, alpha modes 0..2), checked in the reference's
 * order so the first failing field names the code (native/src/version.c:90-150). */
int32_t tc_query_support(uint32_t profile, uint32_t pixel_format, uint32_t bit_depth, uint32_t alpha_mode)
{
    const int profile_ok = profile == 3u || profile == 5u || profile == 6u || profile == 7u;
    if (!profile_ok) {
        return tc_fail(TC_ERR_UNSUPPORTED_PROFILE, "profile %u not in {3, 5, 6, 7}", profile);
    }
    const int depth_ok = bit_depth == 10u || bit_depth == 12u || bit_depth == 16u;
    if (pixel_format > 3u || !depth_ok) {
        return tc_fail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "pixel_format %u / bit_depth %u not supported",
                       pixel_format, bit_depth);
    }
    const int raw = profile == 7u;
    if (raw != (pixel_format == 3u)) {
        return tc_fail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "profile 7 and pixel_format 3 (CFA) go together");
    }
    if (raw && bit_depth == 10u) {
        return tc_fail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "profile 7 needs 12 or 16 bits");
    }
    if (raw && alpha_mode != 0u) {
        return tc_fail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "profile 7 carries no alpha");
    }
    if ((profile == 5u || profile == 6u) && pixel_format == 0u) {
        return tc_fail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "profile %u needs a 4:4:4 pixel format", profile);
    }
    if (profile == 6u && bit_depth != 12u) {
        return tc_fail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "profile 6 needs 12 bits");
    }
    if (alpha_mode > 2u) {
        return tc_fail(TC_ERR_UNSUPPORTED_ALPHA_MODE, "alpha_mode %u > 2", alpha_mode);
    }
    return TC_OK;
}
