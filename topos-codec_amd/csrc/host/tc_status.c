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
