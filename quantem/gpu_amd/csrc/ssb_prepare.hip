// Data-preparation kernels of the MI355X SSB engine: the uint16 mean
// diffraction pattern stream, the on-device bright-field selection, the
// per-BF detector geometry, and the BF gather fused with the scan-row FFT.
//
// Reference behaviour (paths relative to bobleesj/quantem.gpu):
//   mean DP / mask     src/quantem/gpu/ssb/compute/cuda/backend.py:420-518
//   geometry           src/quantem/gpu/ssb/compute/cuda/engine.py:784-848
//   G extraction       src/quantem/gpu/ssb/compute/cuda/backend.py:520-589
#include <type_traits>

#include "ssb_internal.h"

namespace qssb {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------
// Mean DP, pass 1: per frame-range partial sums.
// grid = (pixel tiles of 256*VEC, frame ranges).  Every workgroup streams a
// contiguous 256*VEC-pixel slice of `frames_per_range` consecutive frames,
// so each wave issues 16-byte coalesced non-temporal loads.  A range holds
// at most 65536 frames, so the 32-bit partials of 8- and 16-bit samples
// cannot overflow (65536 * 65535 < 2^32, 65536 * 32768 = 2^31); 32-bit
// samples keep 64-bit partials (signed ones sign-extended, modulo 2^64, as
// the reference's uint64 accumulator), float32 samples double-float pairs.
// ------------------------------------------------------------------------
template <int VEC>
__global__ void __launch_bounds__(256)
k_dp_partial(const uint16_t* __restrict__ data, long long frames, int D, int frames_per_range,
             uint32_t* __restrict__ partial) {
  const int pix0 = (blockIdx.x * 256 + threadIdx.x) * VEC;
  const long long f0 = (long long)blockIdx.y * frames_per_range;
  long long f1 = f0 + frames_per_range;
  if (f1 > frames) f1 = frames;
  if (pix0 >= D) return;
  uint32_t acc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) acc[e] = 0u;
  const uint16_t* p = data + f0 * (long long)D + pix0;
  if constexpr (VEC == 8) {
    long long f = f0;
    for (; f + 2 <= f1; f += 2) {
      u32x4 v[2];
#pragma unroll
      for (int u = 0; u < 2; ++u)
        v[u] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p + (long long)u * D));
#pragma unroll
      for (int u = 0; u < 2; ++u) {
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          acc[2 * w] += v[u][w] & 0xFFFFu;
          acc[2 * w + 1] += v[u][w] >> 16;
        }
      }
      p += 2LL * D;
    }
    for (; f < f1; ++f) {
      u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        acc[2 * w] += v[w] & 0xFFFFu;
        acc[2 * w + 1] += v[w] >> 16;
      }
      p += D;
    }
    uint32_t* out = partial + (long long)blockIdx.y * D + pix0;
#pragma unroll
    for (int e = 0; e < 8; ++e) out[e] = acc[e];
  } else {
    for (long long f = f0; f < f1; ++f) {
      acc[0] += p[0];
      p += D;
    }
    partial[(long long)blockIdx.y * D + pix0] = acc[0];
  }
}

// The uint16 stream pass of a reconstruction: the same partial sums, and
// every frame's superset box copied compactly to C[frame][row - r0][col - c0]
// (bh * bw uint16 per frame), so the BF gather reads contiguous box rows
// instead of 128-byte lines under scattered BF pixels.  The box spans whole
// 8-column groups (c0 % 8 == 0, bw % 8 == 0) and det_cols % 8 == 0, so a
// thread's 8 pixels are all inside or all outside it and each inside thread
// copies a frame with one 16-byte store (C rows and frames stay 16-byte
// aligned).  Plain stores: the gather reads C soon after.
__global__ void __launch_bounds__(256)
k_dp_partial_sup(const uint16_t* __restrict__ data, long long frames, int D, int det_cols, int frames_per_range,
                 uint32_t* __restrict__ partial, const SupBox* __restrict__ boxp, uint16_t* __restrict__ C) {
  const int pix0 = (blockIdx.x * 256 + threadIdx.x) * 8;
  const long long f0 = (long long)blockIdx.y * frames_per_range;
  long long f1 = f0 + frames_per_range;
  if (f1 > frames) f1 = frames;
  if (pix0 >= D) return;
  const SupBox box = *boxp;
  const int row = pix0 / det_cols, col0 = pix0 - row * det_cols;
  const long long SB = (long long)box.bh * box.bw;
  const bool copy = box.bw > 0 && row >= box.r0 && row < box.r0 + box.bh && col0 >= box.c0 &&
                    col0 < box.c0 + box.bw;
  uint32_t acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0u;
  u32x4* cp = reinterpret_cast<u32x4*>(
      C + f0 * SB + (copy ? (long long)(row - box.r0) * box.bw + (col0 - box.c0) : 0));
  const long long cstep = SB / 8;   // u32x4 per frame of C
  // Loads and stores share one in-order completion counter: the loads of
  // the next frame pair are issued before the stores of the current one, so
  // waiting for those loads never waits for the stores.  The last pair's
  // prefetch re-reads the range's last frame (in bounds, unused).
  const long long last = f1 - 1;
  auto frame = [&](long long f) {
    return __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(data + (f < last ? f : last) * D + pix0));
  };
  auto add = [&](const u32x4& v) {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      acc[2 * w] += v[w] & 0xFFFFu;
      acc[2 * w + 1] += v[w] >> 16;
    }
  };
  long long f = f0;
  if (f + 2 <= f1) {
    u32x4 cur0 = frame(f), cur1 = frame(f + 1);
    for (; f + 2 <= f1; f += 2) {
      const u32x4 nxt0 = frame(f + 2), nxt1 = frame(f + 3);
      if (copy) {
        cp[0] = cur0;
        cp[cstep] = cur1;
      }
      add(cur0);
      add(cur1);
      cp += 2 * cstep;
      cur0 = nxt0;
      cur1 = nxt1;
    }
  }
  if (f < f1) {
    const u32x4 v = frame(f);
    add(v);
    if (copy) cp[0] = v;
  }
  uint32_t* out = partial + (long long)blockIdx.y * D + pix0;
#pragma unroll
  for (int e = 0; e < 8; ++e) out[e] = acc[e];
}

// Partial-sum arithmetic per sample type: P is the stored partial.
template <typename T> struct DpElem;
template <> struct DpElem<uint8_t> { typedef uint32_t P; };
template <> struct DpElem<int8_t> { typedef int32_t P; };
template <> struct DpElem<uint16_t> { typedef uint32_t P; };
template <> struct DpElem<int16_t> { typedef int32_t P; };
template <> struct DpElem<uint32_t> { typedef unsigned long long P; };
template <> struct DpElem<int32_t> { typedef unsigned long long P; };
template <> struct DpElem<float> { typedef float2 P; };

template <typename P> __device__ __forceinline__ void dp_zero(P& a) { a = 0; }
template <> __device__ __forceinline__ void dp_zero<float2>(float2& a) { a = make_float2(0.f, 0.f); }

template <typename T, typename P>
__device__ __forceinline__ void dp_add(P& a, T v) {
  if constexpr (std::is_same<P, float2>::value) {
    // cascaded summation: the rounding error of every add goes to .y
    const df t = two_sum(a.x, v);
    a.x = t.hi;
    a.y += t.lo;
  } else if constexpr (std::is_same<T, int32_t>::value) {
    a += (unsigned long long)(long long)v;
  } else {
    a += (P)v;
  }
}

// Generic stream pass for the non-uint16 sample types: VEC samples of 16
// bytes per thread and frame when the detector row allows it, else one.
template <typename T, int VEC>
__global__ void __launch_bounds__(256)
k_dp_partial_t(const T* __restrict__ data, long long frames, int D, int frames_per_range,
               typename DpElem<T>::P* __restrict__ partial) {
  typedef typename DpElem<T>::P P;
  const int pix0 = (blockIdx.x * 256 + threadIdx.x) * VEC;
  const long long f0 = (long long)blockIdx.y * frames_per_range;
  long long f1 = f0 + frames_per_range;
  if (f1 > frames) f1 = frames;
  if (pix0 >= D) return;
  P acc[VEC];
#pragma unroll
  for (int e = 0; e < VEC; ++e) dp_zero(acc[e]);
  const T* p = data + f0 * (long long)D + pix0;
  for (long long f = f0; f < f1; ++f) {
    if constexpr (VEC * sizeof(T) == 16) {
      union {
        u32x4 v;
        T e[VEC];
      } u;
      u.v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
#pragma unroll
      for (int e = 0; e < VEC; ++e) dp_add<T, P>(acc[e], u.e[e]);
    } else {
#pragma unroll
      for (int e = 0; e < VEC; ++e) dp_add<T, P>(acc[e], p[e]);
    }
    p += D;
  }
  P* out = partial + (long long)blockIdx.y * D + pix0;
#pragma unroll
  for (int e = 0; e < VEC; ++e) out[e] = acc[e];
}

// Pass 2: fold the frame ranges into exact uint64 sums (integer samples) or
// a double-float sum (float32 samples); mean = f32(sum)/f32(N) with one
// correctly rounded division (backend.py:438-441), times the detector gain
// when one is set (backend.py:442-443; x * 1.0f == x, so no gain leaves the
// mean unchanged).  dcpix[i] = f32(sum) * gain is pixel i's contribution to
// G_b(0,0), used for dc when the integer path cannot be exact.  A workgroup
// owns 64 pixels; its LANES waves fold interleaved shares of the ranges
// (each wave reads 256 contiguous bytes per range) and combine in LDS.
template <typename P> struct DpSum { typedef unsigned long long S; };
template <> struct DpSum<float2> { typedef float2 S; };

template <typename P>
__device__ __forceinline__ void dp_fold(typename DpSum<P>::S& s, P v) {
  if constexpr (std::is_same<P, float2>::value) {
    const df t = df_add(df{s.x, s.y}, df{v.x, v.y});
    s = make_float2(t.hi, t.lo);
  } else if constexpr (std::is_same<P, int32_t>::value) {
    s += (unsigned long long)(long long)v;
  } else {
    s += (unsigned long long)v;
  }
}

template <int LANES, typename P, bool SIGNED>
__global__ void __launch_bounds__(64 * LANES)
k_dp_reduce(const P* __restrict__ partial, int ranges, int D, long long frames, const float* __restrict__ gain,
            unsigned long long* __restrict__ sums, float* __restrict__ mean, float* __restrict__ dcpix) {
  typedef typename DpSum<P>::S S;
  __shared__ S part[LANES][64];
  const int lane = threadIdx.x & 63;
  const int quarter = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  S s;
  dp_zero(s);
  if (i < D) {
    int r = quarter;
    for (; r + 3 * LANES < ranges; r += 4 * LANES) {
      const P a = partial[(long long)r * D + i];
      const P b = partial[(long long)(r + LANES) * D + i];
      const P c = partial[(long long)(r + 2 * LANES) * D + i];
      const P d = partial[(long long)(r + 3 * LANES) * D + i];
      dp_fold<P>(s, a);
      dp_fold<P>(s, b);
      dp_fold<P>(s, c);
      dp_fold<P>(s, d);
    }
    for (; r < ranges; r += LANES) dp_fold<P>(s, partial[(long long)r * D + i]);
  }
  part[quarter][lane] = s;
  __syncthreads();
  if (quarter == 0 && i < D) {
    S t;
    dp_zero(t);
#pragma unroll
    for (int k = 0; k < LANES; ++k) {
      if constexpr (std::is_same<P, float2>::value) {
        const df u = df_add(df{t.x, t.y}, df{part[k][lane].x, part[k][lane].y});
        t = make_float2(u.hi, u.lo);
      } else {
        t += part[k][lane];
      }
    }
    const float g = gain ? gain[i] : 1.0f;
    float total;
    if constexpr (std::is_same<P, float2>::value) {
      total = t.x + t.y;
      sums[i] = 0ull;
    } else {
      total = (float)t;
      sums[i] = t;
    }
    mean[i] = (total / (float)frames) * g;
    float dcv;
    if constexpr (SIGNED) dcv = (float)(long long)t;
    else dcv = total;
    dcpix[i] = dcv * g;
  }
}

// ------------------------------------------------------------------------
// Bright-field selection in one workgroup (1024 threads = 16 waves).
// backend.py:448-518, see select_bf() in the oracle for the float32 rules.
// Reductions run within each wave through cross-lane shuffles, then across
// the 16 wave results in LDS (two barriers per reduction).
// ------------------------------------------------------------------------
#define MASK_THREADS 1024
#define MASK_WAVES (MASK_THREADS / 64)
#define MASK_PER 64   // max pixels per thread (D <= 65536)

__device__ __forceinline__ void two_sum_f(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// compensated (s, c) pair summed over the wave, then over the block
__device__ float block_sum_comp(float s, float c, float* sh) {
#pragma clang fp contract(off)
  for (int off = 32; off > 0; off >>= 1) {
    const float os = __shfl_xor(s, off, 64);
    const float oc = __shfl_xor(c, off, 64);
    float hs, he;
    two_sum_f(s, os, hs, he);
    s = hs;
    c = c + oc + he;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sh[w] = s;
    sh[MASK_WAVES + w] = c;
  }
  __syncthreads();
  float ts = (lane < MASK_WAVES) ? sh[lane] : 0.0f;
  float tc = (lane < MASK_WAVES) ? sh[MASK_WAVES + lane] : 0.0f;
  for (int off = MASK_WAVES / 2; off > 0; off >>= 1) {
    const float os = __shfl_xor(ts, off, 64);
    const float oc = __shfl_xor(tc, off, 64);
    float hs, he;
    two_sum_f(ts, os, hs, he);
    ts = hs;
    tc = tc + oc + he;
  }
  __syncthreads();
  // lanes 0..15 hold the total; broadcast lane 0 to the whole wave
  return __shfl(ts + tc, 0, 64);
}

__device__ float block_max(float v, float* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  float t = (lane < MASK_WAVES) ? sh[lane] : -INFINITY;
  for (int off = MASK_WAVES / 2; off > 0; off >>= 1) t = fmaxf(t, __shfl_xor(t, off, 64));
  __syncthreads();
  return __shfl(t, 0, 64);
}

__device__ __forceinline__ long long shfl_xor_i64(long long v, int off) {
  int lo = (int)(v & 0xffffffffLL), hi = (int)(v >> 32);
  lo = __shfl_xor(lo, off, 64);
  hi = __shfl_xor(hi, off, 64);
  return ((long long)hi << 32) | (unsigned int)lo;
}

// three int32 sums at once (exact: the caller's totals fit in 32 bits)
__device__ void block_sum3_i32(int& a, int& b, int& c, int* sh) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sh[w] = a;
    sh[MASK_WAVES + w] = b;
    sh[2 * MASK_WAVES + w] = c;
  }
  __syncthreads();
  int ta = 0, tb = 0, tc = 0;
  for (int k = 0; k < MASK_WAVES; ++k) {
    ta += sh[k];
    tb += sh[MASK_WAVES + k];
    tc += sh[2 * MASK_WAVES + k];
  }
  __syncthreads();
  a = ta;
  b = tb;
  c = tc;
}

// three int64 sums at once
__device__ void block_sum3_i64(long long& a, long long& b, long long& c, long long* sh) {
  for (int off = 32; off > 0; off >>= 1) {
    a += shfl_xor_i64(a, off);
    b += shfl_xor_i64(b, off);
    c += shfl_xor_i64(c, off);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sh[w] = a;
    sh[MASK_WAVES + w] = b;
    sh[2 * MASK_WAVES + w] = c;
  }
  __syncthreads();
  long long ta = 0, tb = 0, tc = 0;
  for (int k = 0; k < MASK_WAVES; ++k) {
    ta += sh[k];
    tb += sh[MASK_WAVES + k];
    tc += sh[2 * MASK_WAVES + k];
  }
  __syncthreads();
  a = ta;
  b = tb;
  c = tc;
}

// float32 squared distance without contraction into FMA (matches the
// separately rounded CuPy elementwise ops of backend.py:509-511).
__device__ __forceinline__ float dist2_f32(float r, float c, float cr, float cc) {
#pragma clang fp contract(off)
  float a = r - cr;
  float b = c - cc;
  float aa = a * a;
  float bb = b * b;
  return aa + bb;
}

// f32(n / pi) from an exact product with 1/pi split in hi + lo parts.
__device__ __forceinline__ float count_over_pi(float n) {
  const float inv_pi_hi = 0.3183098733425140380859375f;       // f32(1/pi)
  const float inv_pi_lo = 1.284127663345183e-08f;               // 1/pi - hi
  float p = n * inv_pi_hi;
  float e = fmaf(n, inv_pi_hi, -p);
  return p + (e + n * inv_pi_lo);
}

// CACHE: the mean DP (<= 16384 pixels) is staged in LDS once for the passes.
// The passes over a thread's run are unrolled by 8, so their LDS reads are
// issued together rather than one latency per pixel.
#define SELECT_CACHE_PIX 16384
template <bool CACHE>
__global__ void __launch_bounds__(MASK_THREADS)
k_select_bf(const float* __restrict__ mean_g, const unsigned long long* __restrict__ sums,
            const float* __restrict__ dcpix, int exact_dc,
            int det_rows, int det_cols, float threshold, float bf_radius, int max_bf,
            int* __restrict__ bf_rows, int* __restrict__ bf_cols, BfInfo* __restrict__ info,
            const SupBox* __restrict__ boxp) {
  const SupBox box = boxp ? *boxp : SupBox{0, 0, 0, 0};
  __shared__ float sh_f[2 * MASK_WAVES];
  __shared__ long long sh_i[3 * MASK_WAVES];
  __shared__ int sh_i32[3 * MASK_WAVES];
  __shared__ int sh_scan[MASK_WAVES];
  // the LDS copy holds one pad word after each thread's run of `per`
  // pixels, so the lanes of a wave, reading pixel p0 + j of their runs
  // together, hit banks (per + 1) apart instead of `per` apart
  __shared__ float smean[CACHE ? SELECT_CACHE_PIX + MASK_THREADS : 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int D = det_rows * det_cols;
  // each thread owns the contiguous pixels [p0, p0 + per)
  const int per = (D + MASK_THREADS - 1) / MASK_THREADS;   // <= MASK_PER
  if (CACHE) {
    // the padded slot of pixel i is i + i / per; q = i / per by a multiply
    // and shift, exact for i < 2^16 and per <= 64
    const unsigned magic = ((1u << 22) + (unsigned)per - 1u) / (unsigned)per;
#pragma unroll 16
    for (int i = tid; i < D; i += MASK_THREADS) smean[i + (int)(((unsigned)i * magic) >> 22)] = mean_g[i];
    __syncthreads();
  }
  // owner-only view: mean[i] for i in [p0, p1) is pixel i
  const float* mean = CACHE ? smean + tid : mean_g;
#define MV(i) mean[i]
  const int p0 = tid * per;
  const int p1 = min(D, p0 + per);
#define FOR_PIX _Pragma("unroll 8") for (int i = p0; i < p1; ++i)
  // (row, col) of pixel i, advanced incrementally from p0 (no divisions)
  const int r_start = p0 / det_cols, c_start = p0 - (p0 / det_cols) * det_cols;
#define FOR_PIX_RC \
  _Pragma("unroll 8") for (int i = p0, r_ = r_start, c_ = c_start; i < p1; ++i, (++c_ == det_cols) ? (c_ = 0, ++r_) : 0)
  float lmax = -INFINITY;
  FOR_PIX lmax = fmaxf(lmax, MV(i));
  const float mmax = block_max(lmax, sh_f);
  const float cut = mmax * threshold;

  float center_r, center_c, r2;
  if (bf_radius <= 0.0f) {
    Kahan k1;
    k1.init();
    FOR_PIX k1.add(MV(i));
    const float total = block_sum_comp(k1.s, -k1.c, sh_f);
    const float mu = total / (float)D;
    Kahan k2;
    k2.init();
    FOR_PIX {
      const float d = MV(i) - mu;
      k2.add(d * d);
    }
    const float ss = block_sum_comp(k2.s, -k2.c, sh_f);
    const float sd = sqrtf(ss / (float)D);
    const float thr = mu + sd;
    // (counts and row / column sums below 2^31 for D <= 65536)
    int cnt = 0, sr = 0, sc = 0;
    FOR_PIX_RC {
      if (MV(i) > thr) {
        cnt += 1;
        sr += r_;
        sc += c_;
      }
    }
    block_sum3_i32(cnt, sr, sc, sh_i32);
    if (cnt > 0) {
      center_r = (float)sr / (float)cnt;
      center_c = (float)sc / (float)cnt;
      r2 = count_over_pi((float)cnt);
    } else {
      center_r = det_rows * 0.5f;
      center_c = det_cols * 0.5f;
      float rad = fminf((float)det_rows, (float)det_cols) * 0.25f;
      r2 = rad * rad;
    }
  } else {
    Kahan kw, kr, kc;
    kw.init();
    kr.init();
    kc.init();
    long long n = 0, sr = 0, sc = 0;
    FOR_PIX_RC {
      const float wv = MV(i);
      if (wv > cut) {
        const int r = r_, c = c_;
        kw.add(wv);
        kr.add((float)r * wv);
        kc.add((float)c * wv);
        n += 1;
        sr += r;
        sc += c;
      }
    }
    const float wsum = block_sum_comp(kw.s, -kw.c, sh_f);
    const float rsum = block_sum_comp(kr.s, -kr.c, sh_f);
    const float csum = block_sum_comp(kc.s, -kc.c, sh_f);
    block_sum3_i64(n, sr, sc, sh_i);
    if (wsum > 0.0f) {
      center_r = rsum / wsum;
      center_c = csum / wsum;
    } else if (n > 0) {
      center_r = (float)sr / (float)n;
      center_c = (float)sc / (float)n;
    } else {
      center_r = 0.f;
      center_c = 0.f;
    }
    r2 = bf_radius * bf_radius;
  }

  // keep candidates inside the disk, compacted in raster order; also count
  // candidates and sum the exact integer columns of the kept pixels (dc)
  unsigned long long mask_bits = 0ull;
  long long any_cand = 0, dcs = 0, outside = 0;
  int local = 0;
  Kahan kd;
  kd.init();
  FOR_PIX_RC {
    if (MV(i) > cut) {
      any_cand += 1;
      if (dist2_f32((float)r_, (float)c_, center_r, center_c) <= r2) {
        mask_bits |= 1ull << (i - p0);
        local += 1;
        if (r_ < box.r0 || r_ >= box.r0 + box.bh || c_ < box.c0 || c_ >= box.c0 + box.bw) outside += 1;
      }
    }
  }
  // the kept pixels' dc terms, loads issued 8 at a time and added in pixel
  // order (the integer sums are exact in any order anyway)
  for (int i0 = p0; i0 < p1; i0 += 8) {
    if (exact_dc) {
      unsigned long long v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u;
        v[u] = (i < p1 && ((mask_bits >> (i - p0)) & 1ull)) ? sums[i] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) dcs += (long long)v[u];
    } else {
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u;
        v[u] = (i < p1 && ((mask_bits >> (i - p0)) & 1ull)) ? dcpix[i] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = i0 + u;
        if (i < p1 && ((mask_bits >> (i - p0)) & 1ull)) kd.add(v[u]);
      }
    }
  }
  block_sum3_i64(any_cand, dcs, outside, sh_i);
  const float dc_float = exact_dc ? 0.0f : block_sum_comp(kd.s, -kd.c, sh_f);
  // exclusive scan of `local` over the block: wave scan + wave offsets
  int incl = local;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) sh_scan[w] = incl;
  __syncthreads();
  int wave_off = 0, nbf = 0;
  for (int k = 0; k < MASK_WAVES; ++k) {
    const int t = sh_scan[k];
    if (k < w) wave_off += t;
    nbf += t;
  }
  int pos = wave_off + incl - local;
  FOR_PIX_RC {
    if (mask_bits & (1ull << (i - p0))) {
      if (pos < max_bf) {
        bf_rows[pos] = r_;
        bf_cols[pos] = c_;
      }
      ++pos;
    }
  }
#undef FOR_PIX
#undef FOR_PIX_RC
#undef MV
  if (tid == 0) {
    int status = 0;
    if (any_cand == 0) status = QSSB_DEV_NO_CANDIDATES;
    else if (nbf == 0) status = QSSB_DEV_EMPTY_DISK;
    else if (nbf > max_bf) status = QSSB_DEV_CAPACITY;
    const int nb = (status == 0) ? nbf : 0;
    info->nbf = nb;
    info->status = status;
    info->center_r = center_r;
    info->center_c = center_c;
    info->radius2 = r2;
    float dc = 0.0f;
    if (nb > 0 && exact_dc) {
      // G_b(0,0) is the exact integer column sum; dc = mean_b G_b(0,0).
      const unsigned long long tot = (unsigned long long)dcs;
      const unsigned long long qd = tot / (unsigned long long)nb;
      const unsigned long long rm = tot % (unsigned long long)nb;
      dc = (float)qd + (float)rm / (float)nb;
    } else if (nb > 0) {
      // gain-weighted, signed or float samples: compensated float32 sum of
      // the per-pixel G_b(0,0) values
      dc = dc_float / (float)nb;
    }
    info->dc_re = dc;
    info->dc_im = 0.0f;
    info->sup_ok = (box.bw > 0 && outside == 0) ? 1 : 0;
  }
}

// ------------------------------------------------------------------------
// Superset box of the bright field, before the stream pass: per-pixel sums
// of kSupSamples frames spread evenly over the scan (kSupSplits partial
// sums per pixel, kSupSamples / kSupSplits loads per thread, all in flight
// together), then one workgroup takes the box around every pixel above a
// quarter of the brightest, widened by kSupMargin pixels and out to whole
// 8-column groups.  The exact BF selection later checks that it lies inside
// (BfInfo::sup_ok); if not, the gather reads the frames.
// ------------------------------------------------------------------------
constexpr int kSupSamples = 128;
constexpr int kSupSplits = 8;
constexpr int kSupMargin = 2;
constexpr int kSupPer = kSupSamples / kSupSplits;
constexpr int kBoxPix = 8;   // pixels per thread and round of the box pass (64 loads in flight)

__global__ void __launch_bounds__(256)
k_sup_sample(const uint16_t* __restrict__ data, long long frames, int D, uint32_t* __restrict__ part) {
  const int pix = blockIdx.x * 256 + threadIdx.x;
  if (pix >= D) return;
  const long long n = frames < kSupSamples ? frames : kSupSamples;
  const long long stride = frames / n;
  uint32_t v[kSupPer];
#pragma unroll
  for (int k = 0; k < kSupPer; ++k) {
    const long long i = blockIdx.y + (long long)k * kSupSplits;
    v[k] = i < n ? data[i * stride * D + pix] : 0u;
  }
  uint32_t acc = 0u;
#pragma unroll
  for (int k = 0; k < kSupPer; ++k) acc += v[k];
  part[(long long)blockIdx.y * D + pix] = acc;
}

__global__ void __launch_bounds__(MASK_THREADS)
k_sup_box(const uint32_t* __restrict__ part, int det_rows, int det_cols, SupBox* __restrict__ box) {
  __shared__ float sh_f[2 * MASK_WAVES];
  const int tid = threadIdx.x;
  const int D = det_rows * det_cols;
  // totals of the pixels base + tid + k * MASK_THREADS, k < kBoxPix (-1 past D)
  auto totals = [&](int base, float* t) {
    uint32_t v[kBoxPix][kSupSplits];
#pragma unroll
    for (int k = 0; k < kBoxPix; ++k) {
      const int i = base + tid + k * MASK_THREADS;
#pragma unroll
      for (int j = 0; j < kSupSplits; ++j) v[k][j] = i < D ? part[(long long)j * D + i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kBoxPix; ++k) {
      uint32_t s = 0u;
#pragma unroll
      for (int j = 0; j < kSupSplits; ++j) s += v[k][j];
      t[k] = base + tid + k * MASK_THREADS < D ? (float)s : -1.0f;
    }
  };
  constexpr int kChunk = kBoxPix * MASK_THREADS;
  float lmax = 0.0f;
  for (int base = 0; base < D; base += kChunk) {
    float t[kBoxPix];
    totals(base, t);
#pragma unroll
    for (int k = 0; k < kBoxPix; ++k) lmax = fmaxf(lmax, t[k]);
  }
  const float thr = 0.25f * block_max(lmax, sh_f);
  float rmin = -INFINITY, rmax = -INFINITY, cmin = -INFINITY, cmax = -INFINITY;   // rmin/cmin negated
  for (int base = 0; base < D; base += kChunk) {
    float t[kBoxPix];
    totals(base, t);
#pragma unroll
    for (int k = 0; k < kBoxPix; ++k) {
      if (t[k] > thr) {
        const int i = base + tid + k * MASK_THREADS;
        const int r = i / det_cols, c = i - r * det_cols;
        rmin = fmaxf(rmin, (float)-r);
        rmax = fmaxf(rmax, (float)r);
        cmin = fmaxf(cmin, (float)-c);
        cmax = fmaxf(cmax, (float)c);
      }
    }
  }
  rmin = block_max(rmin, sh_f);
  rmax = block_max(rmax, sh_f);
  cmin = block_max(cmin, sh_f);
  cmax = block_max(cmax, sh_f);
  if (tid == 0) {
    SupBox b{0, 0, 0, 0};
    if (thr > 0.0f && rmax >= 0.0f) {
      const int r0 = max(0, (int)(-rmin) - kSupMargin), r1 = min(det_rows - 1, (int)rmax + kSupMargin);
      // whole 8-column groups (det_cols % 8 == 0 on this path)
      const int c0 = (max(0, (int)(-cmin) - kSupMargin) / 8) * 8;
      const int c1 = min(det_cols, (((int)cmax + kSupMargin) / 8 + 1) * 8) - 1;
      const int bh = r1 - r0 + 1, bw = c1 - c0 + 1;
      if (bh > 0 && bw > 0 && bw % 8 == 0 && bh * bw <= kSupMaxPixels) b = SupBox{r0, c0, bh, bw};
    }
    *box = b;
  }
}

int superset_part_count() { return kSupSplits; }

hipError_t launch_superset(const uint16_t* data, long long frames, int det_rows, int det_cols, uint32_t* part,
                           SupBox* box, hipStream_t s) {
  const int D = det_rows * det_cols;
  hipLaunchKernelGGL(k_sup_sample, dim3((D + 255) / 256, kSupSplits), dim3(256), 0, s, data, frames, D, part);
  hipLaunchKernelGGL(k_sup_box, dim3(1), dim3(MASK_THREADS), 0, s, part, det_rows, det_cols, box);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// Per-BF geometry for one scan rotation (engine.py:784-848), evaluated in
// double-float: k = R(-rotation) * (index - centre) * recip, and A(k).
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_bf_geometry(const int* __restrict__ bf_rows, const int* __restrict__ bf_cols,
              const BfInfo* __restrict__ info, int max_bf, df recip_y, df recip_x, df cos_arg, df sin_arg, Optics o,
              float* __restrict__ kx_out, float* __restrict__ kxl_out, float* __restrict__ ky_out,
              float* __restrict__ kyl_out, float* __restrict__ ak_out, const df* __restrict__ rot_dev) {
  const df cos_a = rot_dev ? rot_dev[0] : cos_arg;
  const df sin_a = rot_dev ? rot_dev[1] : sin_arg;
  const int b = blockIdx.x * 256 + threadIdx.x;
  const int nbf = info->nbf;
  if (b >= nbf || b >= max_bf) return;
  const df iy = two_sum((float)bf_rows[b], -info->center_r);
  const df ix = two_sum((float)bf_cols[b], -info->center_c);
  const df kxa = df_mul(iy, recip_y);
  const df kya = df_mul(ix, recip_x);
  const df kx = df_add(df_mul(kxa, cos_a), df_mul(kya, sin_a));
  const df ky = df_sub(df_mul(kya, cos_a), df_mul(kxa, sin_a));
  kx_out[b] = kx.hi;
  kxl_out[b] = kx.lo;
  ky_out[b] = ky.hi;
  kyl_out[b] = ky.lo;
  ak_out[b] = aperture_band(kx, ky, o);
}

// ------------------------------------------------------------------------
// BF gather + scan-row real FFT.
// Work item = (scan row R_r, chunk of 2*PAIRS BF pixels).  The chunk's
// pixels from the N frames of the scan row are read (u16 -> f32), paired as
// re/im of PAIRS complex sequences along R_c, transformed with one complex
// FFT each, and split into the two real-input half spectra.  Output H is
// [b][R_r][slot], slot 0 packing the purely real bins 0 and N/2 as
// (X[0], X[N/2]); slots 1..N/2-1 hold X[k].
// ------------------------------------------------------------------------
template <int N, int PAIRS, typename T>
__global__ void __launch_bounds__(256)
k_gather_rowfft(const T* __restrict__ data, int D, int det_cols,
                const int* __restrict__ bf_rows, const int* __restrict__ bf_cols,
                const BfInfo* __restrict__ info, const float2* __restrict__ tw_g,
                const float* __restrict__ gain, float2* __restrict__ H, const SupBox* __restrict__ boxp,
                const uint16_t* __restrict__ C, int part, int parts, PhasorJob job) {
  // the phasor tables on the side (a few hundred thousand sincos over the
  // whole grid, before the first gather loads)
  if (job.rowph) phasor_work(job, (long long)blockIdx.x * 256 + threadIdx.x, (long long)gridDim.x * 256);
  constexpr int STRIDE = N + 1;
  constexpr int CH = 2 * PAIRS;
  __shared__ float2 z[PAIRS * STRIDE];
  __shared__ float2 tw[N];
  __shared__ int offs[CH];
  __shared__ float gsc[CH];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  // BF pixels [lo, hi) of this part (raster order: the parts read disjoint
  // detector rows, so splitting adds no frame traffic)
  const int nbf_all = info->nbf;
  const int lo = (int)((long long)nbf_all * part / parts);
  const int hi = (int)((long long)nbf_all * (part + 1) / parts);
  const int nchunks = (hi - lo + CH - 1) / CH;
  const long long items = (long long)nchunks * N;
  // compact superset copy: frame stride SB, pixel (r, c) at (r - r0) * bw + (c - c0)
  const bool sup = std::is_same<T, uint16_t>::value && C != nullptr && info->sup_ok != 0;
  SupBox box{0, 0, 0, 0};
  if (sup) box = *boxp;
  const long long fstride = sup ? (long long)box.bh * box.bw : (long long)D;
  const T* src = sup ? reinterpret_cast<const T*>(C) : data;
  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    // XCD-aware order: workgroups go round-robin over the 8 XCDs, so give
    // XCD x the items (row, chunk) of rows x, x+8, ... with the chunks of one
    // scan row consecutive -- they read the same frames, and the 128-byte
    // lines that straddle chunk boundaries then hit in that XCD's L2.
    const long long xcd = item & 7, t = item >> 3;
    const int rr = (int)((t / nchunks) * 8 + xcd);
    const int chunk = (int)(t % nchunks);
    const int b0 = lo + chunk * CH;
    __syncthreads();
    if (tid < CH) {
      const int b = b0 + tid;
      const int o = (b < hi) ? bf_rows[b] * det_cols + bf_cols[b] : -1;
      offs[tid] = (o < 0) ? -1 : (sup ? (bf_rows[b] - box.r0) * box.bw + (bf_cols[b] - box.c0) : o);
      gsc[tid] = (gain && o >= 0) ? gain[o] : 1.0f;
    }
    __syncthreads();
    // gather: lane -> BF pixel within the chunk, lane group -> frame; all
    // of a thread's loads are issued before any is consumed.  Samples become
    // float32 (astype(complex64)) and are scaled by the pixel's gain
    // (backend.py:570-576; a unit gain leaves them unchanged).
    {
      constexpr int FR = 256 / CH;
      constexpr int NL = N / FR;
      const int j = tid % CH;
      const int off = offs[j];
      const float gj = gsc[j];
      const int c0 = tid / CH;
      float* zf = reinterpret_cast<float*>(z);
      const T* base = src + (long long)rr * N * fstride + (off >= 0 ? off : 0);
      T raw[NL];
#pragma unroll
      for (int m = 0; m < NL; ++m) raw[m] = base[(long long)(c0 + m * FR) * fstride];
#pragma unroll
      for (int m = 0; m < NL; ++m) {
        const int c = c0 + m * FR;
        zf[2 * ((j >> 1) * STRIDE + c) + (j & 1)] = (off >= 0) ? (float)raw[m] * gj : 0.0f;
      }
    }
    __syncthreads();
    fft_lds<N, PAIRS, 256, STRIDE, false>(z, tw);
    // split the pairs into half spectra and store
    constexpr int HALF = N / 2;
    for (int it = tid; it < CH * HALF; it += 256) {
      const int jb = it / HALF;           // BF within chunk
      const int k = it - jb * HALF;       // slot
      const int b = b0 + jb;
      if (b >= hi) continue;
      const int p = jb >> 1;
      const float2 zk = z[p * STRIDE + k];
      float2 out;
      if (k == 0) {
        const float2 zh = z[p * STRIDE + HALF];
        out = (jb & 1) ? make_float2(zk.y, zh.y) : make_float2(zk.x, zh.x);
      } else {
        const float2 zm = z[p * STRIDE + (N - k)];
        if ((jb & 1) == 0) {
          out = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        } else {
          // (Z[k] - conj(Z[N-k])) / (2i)
          const float dr = zk.x - zm.x, di = zk.y + zm.y;
          out = make_float2(0.5f * di, -0.5f * dr);
        }
      }
      H[((long long)b * N + rr) * HALF + k] = out;
    }
  }
}

// ------------------------------------------------------------------------
// Launch wrappers
// ------------------------------------------------------------------------
// Frame ranges of the stream pass: about target_wg workgroups in all (the
// default 1536 = 6 per CU was measured best for a 2 GiB stream on MI355X;
// the tuner may choose another), at most max_ranges, and at most 65536
// frames per range so the 32-bit partial sums cannot overflow.  `vec` is
// the pixels per thread of the stream kernel.
static void range_split(long long frames, int D, int vec, int max_ranges, int target_wg, int& ranges, int& fpr) {
  const long long tiles = (D / vec + 255) / 256;
  long long r = (target_wg > 0 ? target_wg : 1536) / (tiles > 0 ? tiles : 1);
  const long long min_r = (frames + 65535) / 65536;
  if (r < min_r) r = min_r;
  if (r > max_ranges) r = max_ranges;
  if (r > frames) r = frames;
  if (r < 1) r = 1;
  fpr = (int)((frames + r - 1) / r);
  ranges = (int)((frames + fpr - 1) / fpr);
}

int dtype_size(int dtype) {
  switch (dtype) {
    case QSSB_DTYPE_U8: case QSSB_DTYPE_I8: return 1;
    case QSSB_DTYPE_U16: case QSSB_DTYPE_I16: return 2;
    case QSSB_DTYPE_U32: case QSSB_DTYPE_I32: case QSSB_DTYPE_F32: return 4;
    default: return 0;
  }
}

bool dtype_wide_partials(int dtype) {
  return dtype == QSSB_DTYPE_U32 || dtype == QSSB_DTYPE_I32 || dtype == QSSB_DTYPE_F32;
}

// pixels per thread of the stream pass for this dtype, detector and base address
static int stream_vec(int dtype, int D, const void* data) {
  const int es = dtype_size(dtype);
  const int v = 16 / es;
  const bool aligned = (reinterpret_cast<uintptr_t>(data) & 15u) == 0;
  return (D % v == 0 && aligned) ? v : 1;
}

template <typename T>
static void launch_partial_t(const void* data, long long frames, int D, int vec, int ranges, int fpr, void* partial,
                             hipStream_t s) {
  typedef typename DpElem<T>::P P;
  constexpr int V = 16 / sizeof(T);
  if (vec == V) {
    dim3 grid((D / V + 255) / 256, ranges);
    hipLaunchKernelGGL((k_dp_partial_t<T, V>), grid, dim3(256), 0, s, reinterpret_cast<const T*>(data), frames, D,
                       fpr, reinterpret_cast<P*>(partial));
  } else {
    dim3 grid((D + 255) / 256, ranges);
    hipLaunchKernelGGL((k_dp_partial_t<T, 1>), grid, dim3(256), 0, s, reinterpret_cast<const T*>(data), frames, D,
                       fpr, reinterpret_cast<P*>(partial));
  }
}

hipError_t launch_mean_dp_stream(int dtype, const void* data, long long frames, int D, void* partial, int max_ranges,
                                 int target_wg, hipStream_t s) {
  const int vec = stream_vec(dtype, D, data);
  int ranges, fpr;
  range_split(frames, D, vec, max_ranges, target_wg, ranges, fpr);
  switch (dtype) {
    case QSSB_DTYPE_U16:
      if (vec == 8) {
        dim3 grid((D / 8 + 255) / 256, ranges);
        hipLaunchKernelGGL(k_dp_partial<8>, grid, dim3(256), 0, s, reinterpret_cast<const uint16_t*>(data), frames,
                           D, fpr, reinterpret_cast<uint32_t*>(partial));
      } else {
        dim3 grid((D + 255) / 256, ranges);
        hipLaunchKernelGGL(k_dp_partial<1>, grid, dim3(256), 0, s, reinterpret_cast<const uint16_t*>(data), frames,
                           D, fpr, reinterpret_cast<uint32_t*>(partial));
      }
      break;
    case QSSB_DTYPE_U8: launch_partial_t<uint8_t>(data, frames, D, vec, ranges, fpr, partial, s); break;
    case QSSB_DTYPE_I8: launch_partial_t<int8_t>(data, frames, D, vec, ranges, fpr, partial, s); break;
    case QSSB_DTYPE_I16: launch_partial_t<int16_t>(data, frames, D, vec, ranges, fpr, partial, s); break;
    case QSSB_DTYPE_U32: launch_partial_t<uint32_t>(data, frames, D, vec, ranges, fpr, partial, s); break;
    case QSSB_DTYPE_I32: launch_partial_t<int32_t>(data, frames, D, vec, ranges, fpr, partial, s); break;
    case QSSB_DTYPE_F32: launch_partial_t<float>(data, frames, D, vec, ranges, fpr, partial, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mean_dp_stream_sup(const uint16_t* data, long long frames, int D, int det_cols, uint32_t* partial,
                                     int max_ranges, int target_wg, const SupBox* box, uint16_t* C, hipStream_t s) {
  if (det_cols % 8 != 0 || stream_vec(QSSB_DTYPE_U16, D, data) != 8) return hipErrorInvalidValue;
  int ranges, fpr;
  range_split(frames, D, 8, max_ranges, target_wg, ranges, fpr);
  dim3 grid((D / 8 + 255) / 256, ranges);
  hipLaunchKernelGGL(k_dp_partial_sup, grid, dim3(256), 0, s, data, frames, D, det_cols, fpr, partial, box, C);
  return hipGetLastError();
}

hipError_t launch_mean_dp_reduce(int dtype, const void* data, long long frames, int D, const void* partial,
                                 int max_ranges, int target_wg, const float* gain, unsigned long long* sums,
                                 float* mean, float* dcpix, hipStream_t s) {
  const int vec = stream_vec(dtype, D, data);
  int ranges, fpr;
  range_split(frames, D, vec, max_ranges, target_wg, ranges, fpr);
  // 16 range lanes per pixel column: 1024-thread workgroups keep enough
  // loads in flight for a reduction that has only D / 64 workgroups
  const dim3 grid((D + 63) / 64), block(1024);
#define QSSB_REDUCE(P, SG)                                                                                  \
  hipLaunchKernelGGL((k_dp_reduce<16, P, SG>), grid, block, 0, s, reinterpret_cast<const P*>(partial), ranges, \
                     D, frames, gain, sums, mean, dcpix)
  switch (dtype) {
    case QSSB_DTYPE_U8: case QSSB_DTYPE_U16: QSSB_REDUCE(uint32_t, false); break;
    case QSSB_DTYPE_I8: case QSSB_DTYPE_I16: QSSB_REDUCE(int32_t, true); break;
    case QSSB_DTYPE_U32: QSSB_REDUCE(unsigned long long, false); break;
    case QSSB_DTYPE_I32: QSSB_REDUCE(unsigned long long, true); break;
    case QSSB_DTYPE_F32: QSSB_REDUCE(float2, false); break;
    default: return hipErrorInvalidValue;
  }
#undef QSSB_REDUCE
  return hipGetLastError();
}

hipError_t launch_mean_dp(int dtype, const void* data, long long frames, int D, void* partial, int max_ranges,
                          int target_wg, const float* gain, unsigned long long* sums, float* mean, float* dcpix,
                          hipStream_t s) {
  hipError_t e = launch_mean_dp_stream(dtype, data, frames, D, partial, max_ranges, target_wg, s);
  if (e != hipSuccess) return e;
  return launch_mean_dp_reduce(dtype, data, frames, D, partial, max_ranges, target_wg, gain, sums, mean, dcpix, s);
}

// ------------------------------------------------------------------------
// Probe disk detection (qg/detector/compute/cuda/probe.py:12-123): centre of
// mass of `mean > ratio*max`, rounded half-to-even; radial profile of
// rint(|r - centre|) bins up to the nearest edge; Gaussian smoothing
// (sigma 2, 13 taps, edge padding); radius = first bin below half of the
// mean of the first five smoothed bins.  One workgroup; every sum runs in a
// fixed order.  out = {status, centre row, centre col, radius}; status 1: no
// positive finite maximum, 2: nothing above the threshold.
// ------------------------------------------------------------------------
#define RADIUS_BINS 128
#define RADIUS_CHUNKS (MASK_THREADS / RADIUS_BINS)

__global__ void __launch_bounds__(MASK_THREADS)
k_detect_bf_radius(const float* __restrict__ mean, int R, int C, float ratio_hi, float ratio_lo,
                   int* __restrict__ out) {
  __shared__ float sh_f[2 * MASK_WAVES];
  __shared__ long long sh_i[3 * MASK_WAVES];
  __shared__ float part_s[RADIUS_BINS * RADIUS_CHUNKS];
  __shared__ int part_n[RADIUS_BINS * RADIUS_CHUNKS];
  __shared__ float prof[RADIUS_BINS];
  __shared__ float smooth[RADIUS_BINS];
  const int tid = threadIdx.x;
  const int D = R * C;
  const int per = (D + MASK_THREADS - 1) / MASK_THREADS;
  const int p0 = tid * per;
  const int p1 = min(D, p0 + per);
  float lmax = -INFINITY;
  for (int i = p0; i < p1; ++i) lmax = fmaxf(lmax, mean[i]);
  const float mmax = block_max(lmax, sh_f);
  if (!(mmax > 0.0f) || isinf(mmax)) {
    if (tid == 0) out[0] = 1;
    return;
  }
  // float32 of (ratio * max) with ratio given as a double-float pair
  const df pr = two_prod(ratio_hi, mmax);
  const float thr = pr.hi + (pr.lo + ratio_lo * mmax);
  long long n = 0, sr = 0, sc = 0;
  for (int i = p0; i < p1; ++i) {
    if (mean[i] > thr) {
      n += 1;
      sr += i / C;
      sc += i % C;
    }
  }
  block_sum3_i64(n, sr, sc, sh_i);
  if (n == 0) {
    if (tid == 0) out[0] = 2;
    return;
  }
  // float32 sums of integer coordinates are exact here (< 2^24)
  const float rcf = (float)sr / (float)n;
  const float ccf = (float)sc / (float)n;
  const int rc = min(max((int)rintf(rcf), 0), R - 1);
  const int cc = min(max((int)rintf(ccf), 0), C - 1);
  const int max_r = min(min(rc, cc), min(R - rc, C - cc));
  int radius;
  if (max_r < 2) {
    radius = max(1, min(R, C) / 4);
  } else {
    const int bins = min(max_r, RADIUS_BINS);
    {
      // thread = (bin, chunk of rows); raster order within the chunk
      const int b = tid % RADIUS_BINS, k = tid / RADIUS_BINS;
      const int ra = (int)(((long long)k * R) / RADIUS_CHUNKS), rb = (int)(((long long)(k + 1) * R) / RADIUS_CHUNKS);
      float s = 0.0f;
      int cnt = 0;
      if (b < bins) {
        for (int r = ra; r < rb; ++r) {
          const float dr = (float)(r - rc);
          for (int c = 0; c < C; ++c) {
            const float dc = (float)(c - cc);
            const int ri = (int)rintf(sqrtf(dr * dr + dc * dc));
            if (ri == b) {
              s += mean[r * C + c];
              cnt += 1;
            }
          }
        }
      }
      part_s[b * RADIUS_CHUNKS + k] = s;
      part_n[b * RADIUS_CHUNKS + k] = cnt;
    }
    __syncthreads();
    if (tid < bins) {
      float s = 0.0f;
      int cnt = 0;
      for (int k = 0; k < RADIUS_CHUNKS; ++k) {
        s += part_s[tid * RADIUS_CHUNKS + k];
        cnt += part_n[tid * RADIUS_CHUNKS + k];
      }
      prof[tid] = cnt > 0 ? s / (float)cnt : s;
    }
    __syncthreads();
    if (bins > 5) {
      if (tid < bins) {
        float wsum = 0.0f, acc = 0.0f;
        for (int t = 0; t < 13; ++t) {
          const float x = (float)(t - 6) / 2.0f;
          wsum += expf(-0.5f * x * x);
        }
        for (int t = 0; t < 13; ++t) {
          const float x = (float)(t - 6) / 2.0f;
          const int src = min(max(tid + t - 6, 0), bins - 1);
          acc += prof[src] * (expf(-0.5f * x * x) / wsum);
        }
        smooth[tid] = acc;
      }
      __syncthreads();
      if (tid == 0) {
        const float centre = (smooth[0] + smooth[1] + smooth[2] + smooth[3] + smooth[4]) / 5.0f;
        const float half = centre * 0.5f;
        int r = -1;
        for (int i = 0; i < bins; ++i) {
          if (smooth[i] < half) {
            r = i;
            break;
          }
        }
        sh_i[0] = (r >= 0) ? r : bins / 2;
      }
      __syncthreads();
      radius = (int)sh_i[0];
    } else {
      radius = min(R, C) / 4;
    }
  }
  if (tid == 0) {
    out[0] = 0;
    out[1] = rc;
    out[2] = cc;
    out[3] = max(1, radius);
  }
}

hipError_t launch_detect_bf_radius(const float* mean, int R, int C, df ratio, int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_detect_bf_radius, dim3(1), dim3(MASK_THREADS), 0, s, mean, R, C, ratio.hi, ratio.lo, out);
  return hipGetLastError();
}

hipError_t launch_select_bf(const float* mean, const unsigned long long* sums, const float* dcpix, bool exact_dc,
                            int det_rows, int det_cols, float threshold, float bf_radius, int max_bf, int* rows,
                            int* cols, BfInfo* info, hipStream_t s, const SupBox* box) {
  if (det_rows * det_cols <= SELECT_CACHE_PIX)
    hipLaunchKernelGGL(k_select_bf<true>, dim3(1), dim3(MASK_THREADS), 0, s, mean, sums, dcpix, (int)exact_dc,
                       det_rows, det_cols, threshold, bf_radius, max_bf, rows, cols, info, box);
  else
    hipLaunchKernelGGL(k_select_bf<false>, dim3(1), dim3(MASK_THREADS), 0, s, mean, sums, dcpix, (int)exact_dc,
                       det_rows, det_cols, threshold, bf_radius, max_bf, rows, cols, info, box);
  return hipGetLastError();
}

hipError_t launch_bf_geometry(const int* rows, const int* cols, const BfInfo* info, int max_bf, df recip_y,
                              df recip_x, df cos_a, df sin_a, const Optics& o, float* kx, float* kxl, float* ky,
                              float* kyl, float* ak, hipStream_t s, const df* rot_dev) {
  hipLaunchKernelGGL(k_bf_geometry, dim3((max_bf + 255) / 256), dim3(256), 0, s, rows, cols, info, max_bf,
                     recip_y, recip_x, cos_a, sin_a, o, kx, kxl, ky, kyl, ak, rot_dev);
  return hipGetLastError();
}

int gather_pairs_default(int N) { return N == 128 ? 16 : (N == 1024 ? 4 : 8); }

int gather_pairs_candidates(int N, int* out) {
  switch (N) {
    case 128: out[0] = 16; out[1] = 8; return 2;
    case 256: out[0] = 8; out[1] = 16; return 2;
    case 512: out[0] = 8; out[1] = 4; return 2;
    case 1024: out[0] = 4; out[1] = 2; return 2;
    default: return 0;
  }
}

hipError_t launch_gather_rowfft(int N, int dtype, const void* data, int D, int det_cols, const int* rows,
                                const int* cols, const BfInfo* info, int nbf_hint, const float2* tw, int pairs,
                                const float* gain, float2* H, hipStream_t s, const SupBox* box, const uint16_t* C,
                                int part, int parts, const PhasorJob* job) {
  if (!box) C = nullptr;
  PhasorJob pj{};
  if (job) pj = *job;
  if (pairs <= 0) pairs = gather_pairs_default(N);
  if (parts < 1 || part < 0 || part >= parts) return hipErrorInvalidValue;
  const int part_hint = (nbf_hint + parts - 1) / parts + 1;   // BF pixels of the largest part
  auto grid_for = [&](int ch) {
    long long items = (long long)((part_hint + ch - 1) / ch) * N;
    if (items < 1) items = 1;
    if (items > 16384) items = 16384;
    return dim3((unsigned)items);
  };
#define QSSB_GATHER(NN, PP, T)                                                                              \
  if (N == NN && pairs == PP) {                                                                             \
    hipLaunchKernelGGL((k_gather_rowfft<NN, PP, T>), grid_for(2 * PP), dim3(256), 0, s,                     \
                       reinterpret_cast<const T*>(data), D, det_cols, rows, cols, info, tw, gain, H, box, C,  \
                       part, parts, pj);                                                                    \
    return hipGetLastError();                                                                               \
  }
  if (dtype == QSSB_DTYPE_U16) {
    QSSB_GATHER(128, 16, uint16_t)
    QSSB_GATHER(128, 8, uint16_t)
    QSSB_GATHER(256, 8, uint16_t)
    QSSB_GATHER(256, 16, uint16_t)
    QSSB_GATHER(512, 8, uint16_t)
    QSSB_GATHER(512, 4, uint16_t)
    QSSB_GATHER(1024, 4, uint16_t)
    QSSB_GATHER(1024, 2, uint16_t)
    return hipErrorInvalidValue;
  }
  // other sample types: the default workgroup shape of each scan size
  pairs = gather_pairs_default(N);
#define QSSB_GATHER_T(T)        \
  QSSB_GATHER(128, 16, T)       \
  QSSB_GATHER(256, 8, T)        \
  QSSB_GATHER(512, 8, T)        \
  QSSB_GATHER(1024, 4, T)       \
  return hipErrorInvalidValue;
  switch (dtype) {
    case QSSB_DTYPE_U8: { QSSB_GATHER_T(uint8_t) }
    case QSSB_DTYPE_I8: { QSSB_GATHER_T(int8_t) }
    case QSSB_DTYPE_I16: { QSSB_GATHER_T(int16_t) }
    case QSSB_DTYPE_U32: { QSSB_GATHER_T(uint32_t) }
    case QSSB_DTYPE_I32: { QSSB_GATHER_T(int32_t) }
    case QSSB_DTYPE_F32: { QSSB_GATHER_T(float) }
    default: return hipErrorInvalidValue;
  }
#undef QSSB_GATHER_T
#undef QSSB_GATHER
}

// ------------------------------------------------------------------------
// Scan normalisation (cuda/backend.py:229-269): centre-crop oversize axes to
// the window (r0, c0, h, w) of the source, pad to T x T with the window's
// mean frame floor(sum / (h*w)) -- the reference's int64 sum, float64 divide
// and truncating cast give exactly the integer floor for these ranges --
// placed at pad_top = (T - h)/2, pad_left = (T - w)/2.
// ------------------------------------------------------------------------
constexpr int kWindowRows = 64;   // window rows per partial sum: 64*1024*65535 < 2^32

// Window rows per partial sum: at most kWindowRows; integer samples split
// further so that a small detector still spreads over ~256 workgroups
// (integer sums are exact in any split); float32 keeps kWindowRows, since
// its double-float sums depend on the split.
static int window_rows(int dtype, int h, int D) {
  if (dtype == QSSB_DTYPE_F32) return kWindowRows;
  const int tiles = (D + 255) / 256;
  const int want = (256 + tiles - 1) / tiles;
  const int rows = (h + want - 1) / want;
  return rows < 1 ? 1 : (rows > kWindowRows ? kWindowRows : rows);
}

__global__ void __launch_bounds__(256)
k_window_partial(const uint16_t* __restrict__ src, int W, int D, int r0, int c0, int h, int w, int rows,
                 uint32_t* __restrict__ partial) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= D) return;
  const int ra = blockIdx.y * rows;
  const int rb = min(h, ra + rows);
  uint32_t acc = 0u;
  for (int r = ra; r < rb; ++r) {
    const uint16_t* row = src + ((long long)(r0 + r) * W + c0) * D + p;
#pragma unroll 8
    for (int c = 0; c < w; ++c) acc += row[(long long)c * D];
  }
  partial[(long long)blockIdx.y * D + p] = acc;
}

__global__ void __launch_bounds__(256)
k_window_mean(const uint32_t* __restrict__ partial, int groups, int D, unsigned long long count,
              uint16_t* __restrict__ mean) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= D) return;
  unsigned long long s = 0ull;
  for (int g = 0; g < groups; ++g) s += partial[(long long)g * D + p];
  mean[p] = (uint16_t)(s / count);
}

template <int VEC>
__global__ void __launch_bounds__(256)
k_pad_copy(const uint16_t* __restrict__ src, int W, int D, int r0, int c0, int h, int w, int T, int pad_top,
           int pad_left, const uint16_t* __restrict__ mean, uint16_t* __restrict__ dst) {
  const long long frame = blockIdx.y + (long long)blockIdx.z * 65535;
  if (frame >= (long long)T * T) return;
  const int i = (int)(frame / T), j = (int)(frame % T);
  const int si = i - pad_top, sj = j - pad_left;
  const bool inside = si >= 0 && si < h && sj >= 0 && sj < w;
  const uint16_t* from = inside ? src + ((long long)(r0 + si) * W + (c0 + sj)) * D : mean;
  uint16_t* to = dst + frame * D;
  const int e = (blockIdx.x * 256 + threadIdx.x) * VEC;
  if (e >= D) return;
  if constexpr (VEC == 8) {
    *reinterpret_cast<u32x4*>(to + e) = *reinterpret_cast<const u32x4*>(from + e);
  } else {
    to[e] = from[e];
  }
}

hipError_t launch_normalize_scan(const uint16_t* src, int H, int W, int D, uint16_t* dst, int T,
                                 uint32_t* partial, uint16_t* mean, hipStream_t s) {
  const int r0 = H > T ? (H - T) / 2 : 0;
  const int c0 = W > T ? (W - T) / 2 : 0;
  const int h = H < T ? H : T;
  const int w = W < T ? W : T;
  const int pad_top = (T - h) / 2, pad_left = (T - w) / 2;
  if (h < T || w < T) {
    const int rows = window_rows(QSSB_DTYPE_U16, h, D);
    const int groups = (h + rows - 1) / rows;
    hipLaunchKernelGGL(k_window_partial, dim3((D + 255) / 256, groups), dim3(256), 0, s, src, W, D, r0, c0, h, w,
                       rows, partial);
    hipLaunchKernelGGL(k_window_mean, dim3((D + 255) / 256), dim3(256), 0, s, partial, groups, D,
                       (unsigned long long)h * (unsigned long long)w, mean);
  }
  const long long frames = (long long)T * T;
  const dim3 grid_f(1, (unsigned)(frames < 65535 ? frames : 65535), (unsigned)((frames + 65534) / 65535));
  const bool vec = (D % 8 == 0) && ((reinterpret_cast<uintptr_t>(src) & 15u) == 0) &&
                   ((reinterpret_cast<uintptr_t>(dst) & 15u) == 0) && ((reinterpret_cast<uintptr_t>(mean) & 15u) == 0);
  if (vec) {
    dim3 g = grid_f;
    g.x = (D / 8 + 255) / 256;
    hipLaunchKernelGGL(k_pad_copy<8>, g, dim3(256), 0, s, src, W, D, r0, c0, h, w, T, pad_top, pad_left, mean, dst);
  } else {
    dim3 g = grid_f;
    g.x = (D + 255) / 256;
    hipLaunchKernelGGL(k_pad_copy<1>, g, dim3(256), 0, s, src, W, D, r0, c0, h, w, T, pad_top, pad_left, mean, dst);
  }
  return hipGetLastError();
}

int normalize_partial_groups(int dtype, int H, int T, int D) {
  const int h = H < T ? H : T;
  const int rows = window_rows(dtype, h, D);
  return (h + rows - 1) / rows;
}

// Scan normalisation of the other sample types: per 64-row group int64
// sums (integers) or double-float sums (float32) of the window, then the
// pad frame sum / count truncated toward zero (the float64 quotient of an
// int64 sum below 2^53 truncates to the same integer) or rounded to float32.
template <typename T> struct WinSum { typedef long long S; };
template <> struct WinSum<float> { typedef float2 S; };

template <typename T>
__global__ void __launch_bounds__(256)
k_window_partial_t(const T* __restrict__ src, int W, int D, int r0, int c0, int h, int w, int rows,
                   typename WinSum<T>::S* __restrict__ partial) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= D) return;
  const int ra = blockIdx.y * rows;
  const int rb = min(h, ra + rows);
  if constexpr (std::is_same<T, float>::value) {
    float hi = 0.0f, lo = 0.0f;
    for (int r = ra; r < rb; ++r) {
      const T* row = src + ((long long)(r0 + r) * W + c0) * D + p;
      for (int c = 0; c < w; ++c) {
        const df t = two_sum(hi, row[(long long)c * D]);
        hi = t.hi;
        lo += t.lo;
      }
    }
    partial[(long long)blockIdx.y * D + p] = make_float2(hi, lo);
  } else {
    long long acc = 0;
    for (int r = ra; r < rb; ++r) {
      const T* row = src + ((long long)(r0 + r) * W + c0) * D + p;
#pragma unroll 8
      for (int c = 0; c < w; ++c) acc += (long long)row[(long long)c * D];
    }
    partial[(long long)blockIdx.y * D + p] = acc;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_window_mean_t(const typename WinSum<T>::S* __restrict__ partial, int groups, int D, long long count,
                T* __restrict__ mean) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= D) return;
  if constexpr (std::is_same<T, float>::value) {
    df s{0.0f, 0.0f};
    for (int g = 0; g < groups; ++g) {
      const float2 v = partial[(long long)g * D + p];
      s = df_add(s, df{v.x, v.y});
    }
    const df q = df_div(s, df{(float)count, (float)(count - (long long)(float)count)});
    mean[p] = q.hi + q.lo;
  } else {
    long long s = 0;
    for (int g = 0; g < groups; ++g) s += partial[(long long)g * D + p];
    mean[p] = (T)(s / count);
  }
}

// Copy (or pad) one frame of `fbytes` bytes per grid row, in 16-byte words
// when every frame is 16-byte aligned, else in elements of `es` bytes.
template <int ES>
__global__ void __launch_bounds__(256)
k_pad_copy_bytes(const uint8_t* __restrict__ src, int W, long long fbytes, int r0, int c0, int h, int w, int T,
                 int pad_top, int pad_left, const uint8_t* __restrict__ mean, uint8_t* __restrict__ dst) {
  const long long frame = blockIdx.y + (long long)blockIdx.z * 65535;
  if (frame >= (long long)T * T) return;
  const int i = (int)(frame / T), j = (int)(frame % T);
  const int si = i - pad_top, sj = j - pad_left;
  const bool inside = si >= 0 && si < h && sj >= 0 && sj < w;
  const uint8_t* from = inside ? src + ((long long)(r0 + si) * W + (c0 + sj)) * fbytes : mean;
  uint8_t* to = dst + frame * fbytes;
  const long long e = ((long long)blockIdx.x * 256 + threadIdx.x) * ES;
  if (e >= fbytes) return;
  if constexpr (ES == 16) {
    *reinterpret_cast<u32x4*>(to + e) = *reinterpret_cast<const u32x4*>(from + e);
  } else if constexpr (ES == 4) {
    *reinterpret_cast<uint32_t*>(to + e) = *reinterpret_cast<const uint32_t*>(from + e);
  } else if constexpr (ES == 2) {
    *reinterpret_cast<uint16_t*>(to + e) = *reinterpret_cast<const uint16_t*>(from + e);
  } else {
    to[e] = from[e];
  }
}

template <typename T>
static hipError_t normalize_typed(const void* src_v, int H, int W, int D, void* dst_v, int T_, void* partial_v,
                                  void* mean_v, hipStream_t s) {
  typedef typename WinSum<T>::S S;
  const int r0 = H > T_ ? (H - T_) / 2 : 0;
  const int c0 = W > T_ ? (W - T_) / 2 : 0;
  const int h = H < T_ ? H : T_;
  const int w = W < T_ ? W : T_;
  const int pad_top = (T_ - h) / 2, pad_left = (T_ - w) / 2;
  const T* src = reinterpret_cast<const T*>(src_v);
  T* mean = reinterpret_cast<T*>(mean_v);
  if (h < T_ || w < T_) {
    const int rows = window_rows(std::is_same<T, float>::value ? QSSB_DTYPE_F32 : QSSB_DTYPE_I32, h, D);
    const int groups = (h + rows - 1) / rows;
    S* partial = reinterpret_cast<S*>(partial_v);
    hipLaunchKernelGGL(k_window_partial_t<T>, dim3((D + 255) / 256, groups), dim3(256), 0, s, src, W, D, r0, c0, h,
                       w, rows, partial);
    hipLaunchKernelGGL(k_window_mean_t<T>, dim3((D + 255) / 256), dim3(256), 0, s, partial, groups, D,
                       (long long)h * w, mean);
  }
  const long long frames = (long long)T_ * T_;
  const long long fbytes = (long long)D * sizeof(T);
  dim3 grid(1, (unsigned)(frames < 65535 ? frames : 65535), (unsigned)((frames + 65534) / 65535));
  auto al = [](const void* q, int a) { return (reinterpret_cast<uintptr_t>(q) % a) == 0; };
  const uint8_t* sb = reinterpret_cast<const uint8_t*>(src_v);
  const uint8_t* mb = reinterpret_cast<const uint8_t*>(mean_v);
  uint8_t* db = reinterpret_cast<uint8_t*>(dst_v);
  if (fbytes % 16 == 0 && al(sb, 16) && al(db, 16) && al(mb, 16)) {
    grid.x = (unsigned)((fbytes / 16 + 255) / 256);
    hipLaunchKernelGGL(k_pad_copy_bytes<16>, grid, dim3(256), 0, s, sb, W, fbytes, r0, c0, h, w, T_, pad_top,
                       pad_left, mb, db);
  } else {
    constexpr int ES = sizeof(T);
    grid.x = (unsigned)((fbytes / ES + 255) / 256);
    hipLaunchKernelGGL(k_pad_copy_bytes<ES>, grid, dim3(256), 0, s, sb, W, fbytes, r0, c0, h, w, T_, pad_top,
                       pad_left, mb, db);
  }
  return hipGetLastError();
}

hipError_t launch_normalize_scan_typed(int dtype, const void* src, int H, int W, int D, void* dst, int T,
                                       void* partial, void* mean, hipStream_t s) {
  switch (dtype) {
    case QSSB_DTYPE_U16:
      return launch_normalize_scan(reinterpret_cast<const uint16_t*>(src), H, W, D, reinterpret_cast<uint16_t*>(dst),
                                   T, reinterpret_cast<uint32_t*>(partial), reinterpret_cast<uint16_t*>(mean), s);
    case QSSB_DTYPE_U8: return normalize_typed<uint8_t>(src, H, W, D, dst, T, partial, mean, s);
    case QSSB_DTYPE_I8: return normalize_typed<int8_t>(src, H, W, D, dst, T, partial, mean, s);
    case QSSB_DTYPE_I16: return normalize_typed<int16_t>(src, H, W, D, dst, T, partial, mean, s);
    case QSSB_DTYPE_U32: return normalize_typed<uint32_t>(src, H, W, D, dst, T, partial, mean, s);
    case QSSB_DTYPE_I32: return normalize_typed<int32_t>(src, H, W, D, dst, T, partial, mean, s);
    case QSSB_DTYPE_F32: return normalize_typed<float>(src, H, W, D, dst, T, partial, mean, s);
    default: return hipErrorInvalidValue;
  }
}

// bytes of the partial sums and of the pad frame of launch_normalize_scan_typed
size_t normalize_workspace_bytes(int dtype, int H, int T, int D, size_t* mean_offset) {
  const int groups = normalize_partial_groups(dtype, H, T, D);
  const size_t per = dtype == QSSB_DTYPE_U16 ? 4 : 8;
  size_t part = (size_t)groups * D * per;
  part = (part + 255) & ~(size_t)255;
  if (mean_offset) *mean_offset = part;
  return part + (size_t)D * dtype_size(dtype) + 256;
}

// ------------------------------------------------------------------------
// Exact BF columns, detector-major (SSBEngine.write_exact_bf_source,
// cuda/engine.py:577-701): out[b][s] = data[s][bf_rows[b]][bf_cols[b]] for
// BF pixels [bf_offset, bf_offset + count), as uint16 or narrowed to uint8
// (the caller narrows only when the whole block fits).  A 16 x 256 tile of
// (BF, scan) passes through LDS so both the frame reads and the column
// writes are contiguous runs.
// ------------------------------------------------------------------------
template <typename IN, typename OUT>
__global__ void __launch_bounds__(256)
k_bf_columns(const IN* __restrict__ data, long long frames, int D, int det_cols, const int* __restrict__ rows,
             const int* __restrict__ cols, int bf_offset, int count, OUT* __restrict__ out) {
  __shared__ IN tile[16][257];
  __shared__ int lin[16];
  const int tid = threadIdx.x;
  const int b0 = blockIdx.y * 16;
  const long long s0 = (long long)blockIdx.x * 256;
  if (tid < 16) {
    const int b = b0 + tid;
    lin[tid] = (b < count) ? rows[bf_offset + b] * det_cols + cols[bf_offset + b] : 0;
  }
  __syncthreads();
  // reads: 16 lanes per frame (one per BF pixel), 16 frames per pass
  const int lb = tid & 15, lf = tid >> 4;
  for (int f = lf; f < 256; f += 16) {
    const long long s = s0 + f;
    if (s < frames && b0 + lb < count) tile[lb][f] = data[s * D + lin[lb]];
  }
  __syncthreads();
  for (int bl = 0; bl < 16; ++bl) {
    const long long s = s0 + tid;
    if (b0 + bl < count && s < frames) out[(long long)(b0 + bl) * frames + s] = (OUT)tile[bl][tid];
  }
}

// max over n uint16 values: per-block maxima, then one block
__global__ void __launch_bounds__(256)
k_max_u16(const uint16_t* __restrict__ data, long long n, unsigned int* __restrict__ partial) {
  unsigned int m = 0u;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    m = max(m, (unsigned int)data[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, off, 64));
  __shared__ unsigned int sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = max(max(sm[0], sm[1]), max(sm[2], sm[3]));
}

__global__ void __launch_bounds__(256)
k_max_final(const unsigned int* __restrict__ partial, int parts, unsigned int* __restrict__ out) {
  unsigned int m = 0u;
  for (int i = threadIdx.x; i < parts; i += 256) m = max(m, partial[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, off, 64));
  __shared__ unsigned int sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = max(max(sm[0], sm[1]), max(sm[2], sm[3]));
}

hipError_t launch_bf_columns(const void* data, bool in_u8, long long frames, int D, int det_cols, const int* rows,
                             const int* cols, int bf_offset, int count, bool narrow, void* out, hipStream_t s) {
  const dim3 grid((unsigned)((frames + 255) / 256), (unsigned)((count + 15) / 16));
  if (in_u8)
    hipLaunchKernelGGL((k_bf_columns<uint8_t, uint8_t>), grid, dim3(256), 0, s,
                       reinterpret_cast<const uint8_t*>(data), frames, D, det_cols, rows, cols, bf_offset, count,
                       reinterpret_cast<uint8_t*>(out));
  else if (narrow)
    hipLaunchKernelGGL((k_bf_columns<uint16_t, uint8_t>), grid, dim3(256), 0, s,
                       reinterpret_cast<const uint16_t*>(data), frames, D, det_cols, rows, cols, bf_offset, count,
                       reinterpret_cast<uint8_t*>(out));
  else
    hipLaunchKernelGGL((k_bf_columns<uint16_t, uint16_t>), grid, dim3(256), 0, s,
                       reinterpret_cast<const uint16_t*>(data), frames, D, det_cols, rows, cols, bf_offset, count,
                       reinterpret_cast<uint16_t*>(out));
  return hipGetLastError();
}

constexpr int kMaxParts = 1024;

hipError_t launch_max_u16(const uint16_t* data, long long n, unsigned int* partial, unsigned int* out,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_max_u16, dim3(kMaxParts), dim3(256), 0, s, data, n, partial);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, partial, kMaxParts, out);
  return hipGetLastError();
}

int max_u16_parts() { return kMaxParts; }

}  // namespace qssb
