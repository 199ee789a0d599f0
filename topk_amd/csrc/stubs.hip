// Entry points whose subsystems are not built yet: they report
// TOPK_ERR_UNSUPPORTED so the exported symbol set matches include/topk.h.
#include "internal.hpp"

namespace topk {
}  // namespace topk

extern "C" {
int topk_bta(topk_ctx *ctx, uint32_t, uint8_t, const float *, const uint8_t *, uint64_t *, float *, float *) {
    return ctx ? TOPK_ERR_UNSUPPORTED : TOPK_ERR_INVALID;
}
int topk_gpa(topk_ctx *ctx, uint32_t, uint8_t, const float *, const uint8_t *, uint64_t *, float *, float *) {
    return ctx ? TOPK_ERR_UNSUPPORTED : TOPK_ERR_INVALID;
}
int topk_radix_select_u32(topk_ctx *ctx, const uint32_t *, uint64_t, uint64_t, uint32_t *) {
    return ctx ? TOPK_ERR_UNSUPPORTED : TOPK_ERR_INVALID;
}
}
