// ingest.hip -- SIGPROC sample unpack fused with the corner turn.
//
// The reference unpacks each spectrum on the host (read_block,
// src/mdsm-multibeam/file_handler.cpp:193-287) and scatters it into
// [beam][chan][t] float rows one value at a time (readBinaryData,
// src/MDSM.cpp:71-104).  Here the packed bytes are copied to the device
// as they are (1/8 of the float volume for 4-bit data) and one kernel
// decodes and transposes them:
//
//   tile = 64 spectra x 64 channels, 256 threads (4 waves);
//   read:  each wave reads one spectrum's 64 channels (consecutive bytes,
//          shorts or floats) and decodes them into LDS row t;
//   write: each wave writes 64 consecutive samples (256 B) of one channel
//          row from LDS column c (row pitch 65 words: no bank conflicts).
//
// The kernel moves nbits/8 bytes in and 4 bytes out per sample: it is
// bound by HBM writes, and in a streaming reader by the PCIe copy before it.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace mdsm {
namespace {

constexpr int CT_T = 64;   // spectra per tile
constexpr int CT_C = 64;   // channels per tile
constexpr int CT_NT = 256;

template <int NBITS>
__device__ __forceinline__ float decode(const uint8_t *__restrict__ row, uint32_t g) {
  if constexpr (NBITS < 8) {
    const uint32_t bit = g * NBITS;
    return (float)((row[bit >> 3] >> (bit & 7)) & ((1u << NBITS) - 1u));
  } else if constexpr (NBITS == 8) {
    return (float)row[g];
  } else if constexpr (NBITS == 16) {
    return (float)reinterpret_cast<const uint16_t *>(row)[g];
  } else {
    return reinterpret_cast<const float *>(row)[g];
  }
}

template <int NBITS>
__global__ __launch_bounds__(CT_NT) void unpack_ct_kernel(const uint8_t *__restrict__ in,
                                                          uint64_t spec_bytes, uint32_t nspec,
                                                          uint32_t ctot, uint32_t nchans,
                                                          float *__restrict__ out, uint64_t ld,
                                                          uint64_t beam_stride, uint64_t col0) {
  __shared__ float tile[CT_T][CT_C + 1];
  const uint32_t lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  const uint64_t t0 = (uint64_t)blockIdx.x * CT_T;
  const uint32_t g0 = blockIdx.y * CT_C;
  const uint32_t g = g0 + lx;
  for (uint32_t r = ly; r < CT_T; r += CT_NT / 64) {
    const uint64_t t = t0 + r;
    if (t < nspec && g < ctot) tile[r][lx] = decode<NBITS>(in + t * spec_bytes, g);
  }
  __syncthreads();
  const uint64_t t = t0 + lx;
  if (t >= nspec) return;
  for (uint32_t r = ly; r < CT_C; r += CT_NT / 64) {
    const uint32_t gg = g0 + r;
    if (gg >= ctot) break;
    const uint32_t b = gg / nchans, c = gg - b * nchans;
    out[b * beam_stride + (uint64_t)c * ld + col0 + t] = tile[lx][r];
  }
}

}  // namespace

int launch_unpack_corner_turn(const void *in, int nbits, uint32_t nspec, uint32_t nchans,
                              uint32_t nbeams, float *out, uint64_t ld, uint64_t beam_stride,
                              uint64_t col0, hipStream_t s) {
  const uint32_t ctot = nchans * nbeams;
  const uint64_t spec_bytes = (uint64_t)ctot * nbits / 8;
  if (nspec == 0) return MDSM_OK;
  const dim3 grid((nspec + CT_T - 1) / CT_T, (ctot + CT_C - 1) / CT_C);
  const auto *p = static_cast<const uint8_t *>(in);
  switch (nbits) {
    case 1: hipLaunchKernelGGL(unpack_ct_kernel<1>, grid, dim3(CT_NT), 0, s, p, spec_bytes, nspec, ctot, nchans, out, ld, beam_stride, col0); break;
    case 2: hipLaunchKernelGGL(unpack_ct_kernel<2>, grid, dim3(CT_NT), 0, s, p, spec_bytes, nspec, ctot, nchans, out, ld, beam_stride, col0); break;
    case 4: hipLaunchKernelGGL(unpack_ct_kernel<4>, grid, dim3(CT_NT), 0, s, p, spec_bytes, nspec, ctot, nchans, out, ld, beam_stride, col0); break;
    case 8: hipLaunchKernelGGL(unpack_ct_kernel<8>, grid, dim3(CT_NT), 0, s, p, spec_bytes, nspec, ctot, nchans, out, ld, beam_stride, col0); break;
    case 16: hipLaunchKernelGGL(unpack_ct_kernel<16>, grid, dim3(CT_NT), 0, s, p, spec_bytes, nspec, ctot, nchans, out, ld, beam_stride, col0); break;
    case 32: hipLaunchKernelGGL(unpack_ct_kernel<32>, grid, dim3(CT_NT), 0, s, p, spec_bytes, nspec, ctot, nchans, out, ld, beam_stride, col0); break;
    default: set_error("nbits must be 1, 2, 4, 8, 16 or 32 (got %d)", nbits); return MDSM_ERR_ARG;
  }
  MDSM_HIP(hipGetLastError());
  return MDSM_OK;
}

}  // namespace mdsm
