// Integer detector binning of resident uint16 frames (MI355X).
//
// Reference behaviour (paths relative to bobleesj/quantem.gpu):
//   _bin_sum (CPU contract)       src/quantem/gpu/io/backends/cpu/reference.py:27-42
//   dead pixels zeroed first      src/quantem/gpu/io/backends/cpu/reference.py:96-111
//   binned dtype widened to u32   src/quantem/gpu/io/load.py:3009-3026, 3398-3415
//
// out[f][i][j] = sum over the factor x factor block (i*factor.., j*factor..)
// of frame f, dead pixels (mask != 0) counted as 0, rows/cols that do not
// fill a whole block trimmed.  The sum is exact in uint32 (at most
// 65535 * factor^2 < 2^32 for factor <= 256); the 16-bit output keeps its low
// 16 bits, as the reference's uint64 sum cast back to uint16 does.  Generic
// factor: one thread per output pixel, consecutive threads on consecutive
// output columns, so a wave reads factor * 64 consecutive input columns per
// row; factor 2 on rows of a multiple of 8 pixels has a 16-byte-load path.
#include "ssb_internal.h"

namespace qssb {

namespace {
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
}

template <bool WIDE, bool MASK>
__global__ void __launch_bounds__(256)
k_bin_u16(const uint16_t* __restrict__ src, long long frames, int rows, int cols, int factor,
          const uint8_t* __restrict__ bad, void* __restrict__ out) {
  const int orows = rows / factor, ocols = cols / factor;
  const long long per_frame = (long long)orows * ocols;
  const long long total = frames * per_frame;
  for (long long o = (long long)blockIdx.x * 256 + threadIdx.x; o < total; o += (long long)gridDim.x * 256) {
    const long long f = o / per_frame;
    const int p = (int)(o - f * per_frame);
    const int i = p / ocols, j = p - (p / ocols) * ocols;
    const uint16_t* frame = src + f * rows * (long long)cols;
    uint32_t acc = 0u;
    for (int a = 0; a < factor; ++a) {
      const int r = i * factor + a;
      const uint16_t* row = frame + (long long)r * cols + j * factor;
      for (int b = 0; b < factor; ++b) {
        const uint32_t v = row[b];
        acc += (MASK && bad[r * cols + j * factor + b]) ? 0u : v;
      }
    }
    if (WIDE) static_cast<uint32_t*>(out)[o] = acc;
    else static_cast<uint16_t*>(out)[o] = (uint16_t)(acc & 0xffffu);
  }
}

// factor 2 with cols % 8 == 0 (every row 16-byte aligned): a thread takes 4
// output columns, reading 16 bytes from each of its two input rows and
// writing its 4 sums with one store.
template <bool WIDE, bool MASK>
__global__ void __launch_bounds__(256)
k_bin2_u16x8(const uint16_t* __restrict__ src, long long frames, int rows, int cols,
             const uint8_t* __restrict__ bad, void* __restrict__ out) {
  const int orows = rows / 2, groups = cols / 8;          // 4 output columns per group
  const long long per_frame = (long long)orows * groups;
  const long long total = frames * per_frame;
  for (long long o = (long long)blockIdx.x * 256 + threadIdx.x; o < total; o += (long long)gridDim.x * 256) {
    const long long f = o / per_frame;
    const int p = (int)(o - f * per_frame);
    const int i = p / groups, g = p - (p / groups) * groups;
    const uint16_t* r0 = src + (f * rows + 2 * i) * (long long)cols + 8 * g;
    const u32x4 a = *reinterpret_cast<const u32x4*>(r0);
    const u32x4 b = *reinterpret_cast<const u32x4*>(r0 + cols);
    const uint32_t wa[4] = {a.x, a.y, a.z, a.w}, wb[4] = {b.x, b.y, b.z, b.w};
    uint32_t m0[2] = {0u, 0u}, m1[2] = {0u, 0u};
    if (MASK) {
      const uint8_t* q0 = bad + (2 * i) * cols + 8 * g;
      m0[0] = *reinterpret_cast<const uint32_t*>(q0);
      m0[1] = *reinterpret_cast<const uint32_t*>(q0 + 4);
      m1[0] = *reinterpret_cast<const uint32_t*>(q0 + cols);
      m1[1] = *reinterpret_cast<const uint32_t*>(q0 + cols + 4);
    }
    uint32_t acc[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      // word k holds input columns 2k (low half) and 2k+1 (high half) of a row
      uint32_t a0 = wa[k] & 0xffffu, a1 = wa[k] >> 16, b0 = wb[k] & 0xffffu, b1 = wb[k] >> 16;
      if (MASK) {
        const uint32_t ma = m0[k >> 1] >> (16 * (k & 1)), mb = m1[k >> 1] >> (16 * (k & 1));
        a0 = (ma & 0xffu) ? 0u : a0;
        a1 = (ma & 0xff00u) ? 0u : a1;
        b0 = (mb & 0xffu) ? 0u : b0;
        b1 = (mb & 0xff00u) ? 0u : b1;
      }
      acc[k] = a0 + a1 + b0 + b1;
    }
    if (WIDE) {
      u32x4 v;
      v.x = acc[0]; v.y = acc[1]; v.z = acc[2]; v.w = acc[3];
      *reinterpret_cast<u32x4*>(static_cast<uint32_t*>(out) + 4 * o) = v;
    } else {
      uint2 v;
      v.x = (acc[0] & 0xffffu) | (acc[1] << 16);
      v.y = (acc[2] & 0xffffu) | (acc[3] << 16);
      *reinterpret_cast<uint2*>(static_cast<uint16_t*>(out) + 4 * o) = v;
    }
  }
}

hipError_t launch_bin_u16(const uint16_t* src, long long frames, int rows, int cols, int factor,
                          const uint8_t* bad, bool wide, void* out, hipStream_t s) {
  const long long total = frames * (long long)(rows / factor) * (cols / factor);
  if (total <= 0) return hipSuccess;
  const bool aligned16 = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0 &&
                         (!bad || (reinterpret_cast<uintptr_t>(bad) & 3u) == 0);
  if (factor == 2 && cols % 8 == 0 && aligned16) {
    const long long items = frames * (long long)(rows / 2) * (cols / 8);
    long long blocks = (items + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const dim3 grid((unsigned)blocks), block(256);
    if (wide) {
      if (bad) hipLaunchKernelGGL((k_bin2_u16x8<true, true>), grid, block, 0, s, src, frames, rows, cols, bad, out);
      else hipLaunchKernelGGL((k_bin2_u16x8<true, false>), grid, block, 0, s, src, frames, rows, cols, bad, out);
    } else {
      if (bad) hipLaunchKernelGGL((k_bin2_u16x8<false, true>), grid, block, 0, s, src, frames, rows, cols, bad, out);
      else hipLaunchKernelGGL((k_bin2_u16x8<false, false>), grid, block, 0, s, src, frames, rows, cols, bad, out);
    }
    return hipGetLastError();
  }
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  const dim3 grid((unsigned)blocks), block(256);
  if (wide) {
    if (bad) hipLaunchKernelGGL((k_bin_u16<true, true>), grid, block, 0, s, src, frames, rows, cols, factor, bad, out);
    else hipLaunchKernelGGL((k_bin_u16<true, false>), grid, block, 0, s, src, frames, rows, cols, factor, bad, out);
  } else {
    if (bad) hipLaunchKernelGGL((k_bin_u16<false, true>), grid, block, 0, s, src, frames, rows, cols, factor, bad, out);
    else hipLaunchKernelGGL((k_bin_u16<false, false>), grid, block, 0, s, src, frames, rows, cols, factor, bad, out);
  }
  return hipGetLastError();
}

}  // namespace qssb
