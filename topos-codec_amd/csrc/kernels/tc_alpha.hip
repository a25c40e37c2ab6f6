// Alpha plane kernels (MED prediction + Rice run/level pairs).
#include <hip/hip_runtime.h>

#include "tcg.h"
#include "tc_dev.h"

void tcg_alpha_encode_launch(hipStream_t st, const tcg_eplane* planes, const tcg_eslice* slices,
                             uint32_t nslices, const int32_t* unused, uint8_t* scratch,
                             tcg_eresult* results)
{
    (void)st; (void)planes; (void)slices; (void)nslices; (void)unused; (void)scratch; (void)results;
}

void tcg_alpha_decode_launch(hipStream_t st, const tcg_dplane* planes, const tcg_dslice* slices,
                             uint32_t nslices, int32_t* status)
{
    (void)st; (void)planes; (void)slices; (void)nslices; (void)status;
}
