// GPU decode of HDF5 bitshuffle-LZ4 ("bslz4") detector chunks straight into
// the resident uint16 frame array the SSB path reads.
//
// Reference behaviour (paths relative to bobleesj/quantem.gpu):
//   CUDA decoder              src/quantem/gpu/io/backends/cuda/decoder.py:258-1146
//   WebGPU decoder (format)   src/quantem/gpu/io/backends/webgpu/bslz4.ts:1-30, 114-167
//   CPU reader (hdf5plugin)   src/quantem/gpu/io/backends/cpu/reference.py:45-113
//
// Chunk layout: big-endian uint64 uncompressed bytes, big-endian uint32
// block bytes, then per block a big-endian uint32 compressed size and an
// LZ4 block; full blocks of block_bytes/2 elements, one last block of the
// remaining elements rounded down to a multiple of 8, and the final n mod 8
// elements raw.  A block decodes to 16 bit-planes of m/8 bytes: bit b of
// element p is bit (p mod 8) of byte b*m/8 + p/8.
//
// Two kernels:
//   k_bslz4_index   one thread per chunk: validates the header and walks the
//                   block sizes into a descriptor table (bounds-checked);
//   k_bslz4_decode  one 64-lane workgroup per block: stages the compressed
//                   block in LDS, decodes it in LDS (every lane parses the
//                   same sequence headers; literal and match bytes are
//                   copied by all lanes, a match of offset d < length
//                   copied in one pass as out[o + t] = out[o - d + t mod d]),
//                   then inverts the bit-shuffle, 8 elements per lane step
//                   (two 8x8 bit-matrix transposes by delta swaps and byte
//                   permutes), with 16-byte stores.
// Malformed input never reads or writes out of bounds: it sets a bit in the
// status word and the block is skipped.
#include <cstring>
#include <vector>

#include "ssb_internal.h"

namespace qssb {

namespace {

constexpr int kMaxBlockBytes = 8192;                                  // 4096 uint16 elements
constexpr int kMaxCompBytes = kMaxBlockBytes + kMaxBlockBytes / 255 + 16;  // LZ4 worst case
// Compressed blocks up to this size are staged in LDS (detector frames
// compress 10-20x); larger ones are parsed from global memory.  The small
// staging buffer leaves room for about twice as many workgroups per CU.
constexpr int kStageBytes = 2048;

struct BlockDesc {
  long long src;   // byte offset of the block's payload in the compressed buffer
  long long dst;   // first output element
  int csize;       // compressed bytes; -1 = raw little-endian tail
  int elems;       // elements (0 = unused slot)
};

enum : int {
  kErrHeader = 1,      // size or block size differs from what the caller declared
  kErrBounds = 2,      // a block or the tail runs past the chunk
  kErrStream = 4,      // a malformed LZ4 sequence
  kErrBlockSize = 8,   // a compressed block larger than the LZ4 bound
};

__device__ __forceinline__ unsigned be32(const uint8_t* p) {
  return ((unsigned)p[0] << 24) | ((unsigned)p[1] << 16) | ((unsigned)p[2] << 8) | (unsigned)p[3];
}

__global__ void __launch_bounds__(256)
k_bslz4_index(const uint8_t* __restrict__ comp, long long comp_bytes, const long long* __restrict__ chunk_off,
              const long long* __restrict__ chunk_len, int n_chunks, long long chunk_elems, int block_elems,
              int slots, BlockDesc* __restrict__ table, int* __restrict__ status) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n_chunks) return;
  BlockDesc* t = table + (long long)c * slots;
  for (int k = 0; k < slots; ++k) t[k].elems = 0;
  const long long base = chunk_off[c];
  const long long len = chunk_len[c];
  if (base < 0 || len < 12 || base > comp_bytes - len) {
    atomicOr(status, kErrBounds);
    return;
  }
  const uint8_t* p = comp + base;
  const unsigned long long total = ((unsigned long long)be32(p) << 32) | be32(p + 4);
  const unsigned bbytes = be32(p + 8);
  if (total != (unsigned long long)chunk_elems * 2ull || bbytes != (unsigned)block_elems * 2u) {
    atomicOr(status, kErrHeader);
    return;
  }
  const long long full = chunk_elems / block_elems;
  const long long rest = chunk_elems - full * block_elems;
  const int last = (int)(rest - rest % 8);
  const int tail = (int)(rest % 8);
  long long pos = 12;
  int k = 0;
  const long long nblk = full + (last ? 1 : 0);
  for (long long i = 0; i < nblk; ++i) {
    if (pos + 4 > len) {
      atomicOr(status, kErrBounds);
      return;
    }
    const unsigned cs = be32(p + pos);
    pos += 4;
    if (cs > (unsigned)kMaxCompBytes) {
      atomicOr(status, kErrBlockSize);
      return;
    }
    if (pos + (long long)cs > len) {
      atomicOr(status, kErrBounds);
      return;
    }
    t[k].src = base + pos;
    t[k].dst = (long long)c * chunk_elems + i * block_elems;
    t[k].csize = (int)cs;
    t[k].elems = (i < full) ? block_elems : last;
    ++k;
    pos += cs;
  }
  if (pos + 2LL * tail != len) {
    atomicOr(status, kErrBounds);
    return;
  }
  if (tail) {
    t[k].src = base + pos;
    t[k].dst = (long long)c * chunk_elems + full * block_elems + last;
    t[k].csize = -1;
    t[k].elems = tail;
  }
}

__global__ void __launch_bounds__(64)
k_bslz4_decode(const uint8_t* __restrict__ comp, const BlockDesc* __restrict__ table, long long n_desc,
               uint16_t* __restrict__ out, int* __restrict__ status) {
  __shared__ uint8_t cin[kStageBytes];
  __shared__ __attribute__((aligned(16))) uint8_t cout[kMaxBlockBytes];
  const int lane = threadIdx.x;
  for (long long d = blockIdx.x; d < n_desc; d += gridDim.x) {
    const BlockDesc bd = table[d];
    if (bd.elems == 0) continue;                       // uniform
    if (bd.csize < 0) {                                // raw tail, little-endian
      if (lane < bd.elems) {
        const uint8_t* s = comp + bd.src + 2 * lane;
        out[bd.dst + lane] = (uint16_t)(s[0] | (s[1] << 8));
      }
      continue;
    }
    const int cs = bd.csize;
    const int osz = bd.elems * 2;
    const bool staged = cs <= kStageBytes;             // uniform
    const uint8_t* gsrc = comp + bd.src;
    __syncthreads();                                   // previous block's readers are done
    if (staged)
      for (int t = lane; t < cs; t += 64) cin[t] = gsrc[t];
    __syncthreads();
    // LZ4 block: every lane walks the same sequence headers.  The workgroup
    // is one wave, whose LDS operations complete in issue order, so the
    // copies need only a compiler ordering fence between them, not a
    // barrier (4% faster on detector data than with barriers).  Holding the
    // block in registers and reading the headers with readlane measured
    // slower (157 vs 215 GB/s decoded).
    auto parse = [&](auto byte_at, int& ip, int& op) -> int {
      while (true) {
        if (ip >= cs) return kErrStream;
        const int token = byte_at(ip++);
        int lit = token >> 4;
        if (lit == 15) {
          int b;
          do {
            if (ip >= cs) return kErrStream;
            b = byte_at(ip++);
            lit += b;
          } while (b == 255);
        }
        if (ip + lit > cs || op + lit > osz) return kErrStream;
        for (int t = lane; t < lit; t += 64) cout[op + t] = byte_at(ip + t);
        ip += lit;
        op += lit;
        if (ip == cs) return 0;                          // the last sequence has no match
        if (ip + 2 > cs) return kErrStream;
        const int off = byte_at(ip) | (byte_at(ip + 1) << 8);
        ip += 2;
        int ml = token & 15;
        if (ml == 15) {
          int b;
          do {
            if (ip >= cs) return kErrStream;
            b = byte_at(ip++);
            ml += b;
          } while (b == 255);
        }
        ml += 4;
        if (off == 0 || off > op || op + ml > osz) return kErrStream;
        asm volatile("" ::: "memory");                   // literals before the match reads them
        if (off >= ml) {                                 // uniform: no overlap, a plain copy
          for (int t = lane; t < ml; t += 64) cout[op + t] = cout[op - off + t];
        } else if (off == 1) {                           // a run of one byte
          const uint8_t r = cout[op - 1];
          for (int t = lane; t < ml; t += 64) cout[op + t] = r;
        } else {
          for (int t = lane; t < ml; t += 64) cout[op + t] = cout[op - off + (t % off)];
        }
        op += ml;
        asm volatile("" ::: "memory");
      }
    };
    int ip = 0, op = 0;
    int err = staged ? parse([&](int i) -> int { return cin[i]; }, ip, op)
                     : parse([&](int i) -> int { return gsrc[i]; }, ip, op);
    if (!err && op != osz) err = kErrStream;
    if (err) {
      if (lane == 0) atomicOr(status, err);
      continue;                                        // uniform: every lane saw the same stream
    }
    __syncthreads();
    // inverse bit-shuffle, 8 elements per step: the 16 plane bytes of the
    // group form two 8x8 bit matrices (planes 0-7 and 8-15, byte = plane,
    // bit = element); each is transposed in a 64-bit word by three
    // delta swaps, after which byte i holds element i's low or high byte,
    // and v_perm interleaves them into the output words.
    const int planeb = bd.elems / 8;
    for (int g = lane; g < planeb; g += 64) {
      unsigned bt[16];
#pragma unroll
      for (int b = 0; b < 16; ++b) bt[b] = cout[b * planeb + g];
      unsigned long long m[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const unsigned lo = bt[8 * h] | (bt[8 * h + 1] << 8) | (bt[8 * h + 2] << 16) | (bt[8 * h + 3] << 24);
        const unsigned hi = bt[8 * h + 4] | (bt[8 * h + 5] << 8) | (bt[8 * h + 6] << 16) | (bt[8 * h + 7] << 24);
        unsigned long long x = ((unsigned long long)hi << 32) | lo;
        unsigned long long t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
        x = x ^ t ^ (t << 7);
        t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
        x = x ^ t ^ (t << 14);
        t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
        x = x ^ t ^ (t << 28);
        m[h] = x;
      }
      const unsigned xl = (unsigned)m[0], xh = (unsigned)(m[0] >> 32);
      const unsigned yl = (unsigned)m[1], yh = (unsigned)(m[1] >> 32);
      // bytes {x.b0, y.b0, x.b1, y.b1} = elements 0 and 1, and so on
      uint4 w;
      w.x = __builtin_amdgcn_perm(yl, xl, 0x05010400u);
      w.y = __builtin_amdgcn_perm(yl, xl, 0x07030602u);
      w.z = __builtin_amdgcn_perm(yh, xh, 0x05010400u);
      w.w = __builtin_amdgcn_perm(yh, xh, 0x07030602u);
      if ((bd.dst & 7) == 0) {                       // uniform: 16-byte aligned groups
        *reinterpret_cast<uint4*>(out + bd.dst + 8LL * g) = w;
      } else {
        const unsigned ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int i = 0; i < 8; ++i) out[bd.dst + 8LL * g + i] = (uint16_t)(ww[i >> 1] >> (16 * (i & 1)));
      }
    }
  }
}

__global__ void k_bslz4_clear(int* __restrict__ status) {
  if (threadIdx.x == 0) *status = 0;
}

}  // namespace

long long bslz4_slots(long long chunk_elems, int block_elems) {
  if (block_elems <= 0) return 0;
  const long long full = chunk_elems / block_elems;
  const long long rest = chunk_elems - full * block_elems;
  return full + ((rest - rest % 8) ? 1 : 0) + ((rest % 8) ? 1 : 0);
}

long long bslz4_workspace_bytes(int n_chunks, long long chunk_elems, int block_elems) {
  return 256 + (long long)n_chunks * bslz4_slots(chunk_elems, block_elems) * (long long)sizeof(BlockDesc);
}

hipError_t launch_bslz4_decode_u16(const uint8_t* comp, long long comp_bytes, const long long* chunk_off,
                                   const long long* chunk_len, int n_chunks, long long chunk_elems, int block_elems,
                                   uint16_t* out, void* workspace, hipStream_t s) {
  int* status = reinterpret_cast<int*>(workspace);
  BlockDesc* table = reinterpret_cast<BlockDesc*>(reinterpret_cast<char*>(workspace) + 256);
  const long long slots = bslz4_slots(chunk_elems, block_elems);
  hipLaunchKernelGGL(k_bslz4_clear, dim3(1), dim3(64), 0, s, status);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bslz4_index, dim3((unsigned)((n_chunks + 255) / 256)), dim3(256), 0, s, comp, comp_bytes,
                     chunk_off, chunk_len, n_chunks, chunk_elems, block_elems, (int)slots, table, status);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long n_desc = (long long)n_chunks * slots;
  const unsigned grid = (unsigned)(n_desc < 65536 ? n_desc : 65536);
  hipLaunchKernelGGL(k_bslz4_decode, dim3(grid > 0 ? grid : 1), dim3(64), 0, s, comp, table, n_desc, out, status);
  return hipGetLastError();
}

}  // namespace qssb

// ------------------------------------------------------------------------
// Host-side encoder (writers of bslz4 companions and test inputs): the
// chunk layout above, bit-shuffled blocks, greedy LZ4 block compression
// (4-byte hash, one candidate per slot; the last 5 bytes are literals and
// no match starts in the last 12, as the block format requires).
// ------------------------------------------------------------------------
namespace qssb {

namespace {

void put_be32(uint8_t* p, unsigned v) {
  p[0] = (uint8_t)(v >> 24);
  p[1] = (uint8_t)(v >> 16);
  p[2] = (uint8_t)(v >> 8);
  p[3] = (uint8_t)v;
}

void bitshuffle_u16(const uint16_t* v, int m, uint8_t* out) {
  const int pb = m / 8;
  for (int i = 0; i < 16 * pb; ++i) out[i] = 0;
  for (int p = 0; p < m; ++p)
    for (int b = 0; b < 16; ++b)
      if ((v[p] >> b) & 1u) out[b * pb + p / 8] |= (uint8_t)(1u << (p % 8));
}

// LZ4 block of src[0, n) into dst; returns its size.
long long lz4_compress_block(const uint8_t* src, int n, uint8_t* dst) {
  constexpr int kHashBits = 12;
  int table[1 << kHashBits];
  for (int& t : table) t = -1;
  auto hash = [&](int i) {
    unsigned x;
    std::memcpy(&x, src + i, 4);
    return (x * 2654435761u) >> (32 - kHashBits);
  };
  long long o = 0;
  auto put_len = [&](int x) {
    while (x >= 255) {
      dst[o++] = 255;
      x -= 255;
    }
    dst[o++] = (uint8_t)x;
  };
  int anchor = 0, i = 0;
  const int limit = n - 12;
  while (i <= limit) {
    const unsigned h = hash(i);
    const int cand = table[h];
    table[h] = i;
    if (cand < 0 || i - cand > 65535 || std::memcmp(src + cand, src + i, 4) != 0) {
      ++i;
      continue;
    }
    int len = 4;
    while (i + len < n - 5 && src[cand + len] == src[i + len]) ++len;
    const int lit = i - anchor;
    const int ml = len - 4;
    dst[o++] = (uint8_t)(((lit < 15 ? lit : 15) << 4) | (ml < 15 ? ml : 15));
    if (lit >= 15) put_len(lit - 15);
    std::memcpy(dst + o, src + anchor, (size_t)lit);
    o += lit;
    const int off = i - cand;
    dst[o++] = (uint8_t)off;
    dst[o++] = (uint8_t)(off >> 8);
    if (ml >= 15) put_len(ml - 15);
    i += len;
    anchor = i;
  }
  const int lit = n - anchor;
  dst[o++] = (uint8_t)((lit < 15 ? lit : 15) << 4);
  if (lit >= 15) put_len(lit - 15);
  std::memcpy(dst + o, src + anchor, (size_t)lit);
  o += lit;
  return o;
}

}  // namespace

long long bslz4_encode_bound(long long n_elems, int block_bytes) {
  const int be = block_bytes / 2;
  const long long blocks = n_elems / be + 1;
  return 12 + blocks * (4 + (long long)block_bytes + block_bytes / 255 + 16) + 16;
}

long long bslz4_encode_u16(const uint16_t* v, long long n, int block_bytes, uint8_t* dst, long long cap) {
  if (cap < bslz4_encode_bound(n, block_bytes)) return -1;
  const int be = block_bytes / 2;
  const unsigned long long total = (unsigned long long)n * 2ull;
  put_be32(dst, (unsigned)(total >> 32));
  put_be32(dst + 4, (unsigned)total);
  put_be32(dst + 8, (unsigned)block_bytes);
  long long o = 12;
  const long long full = n / be;
  const long long rest = n - full * be;
  const int last = (int)(rest - rest % 8);
  const int tail = (int)(rest % 8);
  std::vector<uint8_t> shuf((size_t)block_bytes);
  long long pos = 0;
  for (long long k = 0; k < full + (last ? 1 : 0); ++k) {
    const int m = (k < full) ? be : last;
    bitshuffle_u16(v + pos, m, shuf.data());
    const long long cs = lz4_compress_block(shuf.data(), 2 * m, dst + o + 4);
    put_be32(dst + o, (unsigned)cs);
    o += 4 + cs;
    pos += m;
  }
  for (int t = 0; t < tail; ++t) {
    dst[o++] = (uint8_t)v[pos + t];
    dst[o++] = (uint8_t)(v[pos + t] >> 8);
  }
  return o;
}

}  // namespace qssb
