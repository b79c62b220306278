// triage_internal.h -- launchers of triage.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace mdsm {

using RfiItem = mdsm_rfi_item_t;
enum { RFI_BURST = MDSM_RFI_BURST, RFI_MAINS = MDSM_RFI_MAINS, RFI_TELEGRAPH = MDSM_RFI_TELEGRAPH };

int launch_rfi_inject(const RfiItem *d_items, uint32_t nitems, int64_t max_len, const float *gain,
                      uint64_t seed, float *raw, uint64_t ld, uint64_t n, hipStream_t s);
int launch_snippets(const float *out, uint32_t ndms, uint32_t nsamp, const float *store, uint64_t ld,
                    uint32_t nchans, const int32_t *table, const int32_t *ev_dm, const int32_t *ev_t,
                    uint32_t nev, int SD, int ST, int SB, float scale0, float scale1, float clip,
                    uint16_t *snip, hipStream_t s);
int launch_gather_rows(const uint16_t *src, const int32_t *labels, const uint32_t *idx, uint32_t n,
                       uint64_t row, uint16_t *dst, int32_t *dlab, hipStream_t s);
int launch_probabilities(const float *logits, uint32_t n, float *prob, hipStream_t s);

}  // namespace mdsm
