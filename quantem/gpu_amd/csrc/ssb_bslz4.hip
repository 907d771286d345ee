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
// Three kernels:
//   k_bslz4_index   one thread per chunk: validates the header and walks the
//                   block sizes into a descriptor table (bounds-checked);
//   k_bslz4_parse   one thread per block: walks the block's LZ4 sequence
//                   headers (every length and offset checked) into a table
//                   of {literal start, literal length, offset, match length}
//                   per sequence;
//   k_bslz4_decode  one 64-lane workgroup per block: stages the compressed
//                   block in LDS as whole words and decodes it in LDS in
//                   rounds of up to 64 table entries (one lane per
//                   sequence: all of the round's literal runs at once, then
//                   its matches in order; a match of offset d < length is
//                   out[o + t] = out[o - d + t mod d], a run of one byte is
//                   written with 16-byte stores; a block whose sequences
//                   overflow the table is walked here instead), then inverts
//                   the bit-shuffle, 32 elements per lane step (two 8x8
//                   bit-matrix transposes by delta swaps and byte permutes
//                   per 8 elements), with 16-byte stores.
// Malformed input never reads or writes out of bounds: it sets a bit in the
// status word and the block is skipped.
#include <cstring>
#include <vector>

#include "ssb_internal.h"

namespace qssb {

namespace {

constexpr int kMaxBlockBytes = 8192;                                  // 4096 uint16 elements
constexpr int kMaxCompBytes = kMaxBlockBytes + kMaxBlockBytes / 255 + 16;  // LZ4 worst case
// Compressed blocks up to this size (less the alignment bytes around them)
// are staged in LDS (detector frames compress 10-20x); larger ones are
// parsed from global memory.  The small staging buffer leaves room for
// about twice as many workgroups per CU.
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

// Sequence table entries per block: block_elems / 16 (256 for the default
// 8192-byte blocks: an average of 32 or more decoded bytes per sequence),
// at least 64, and never more than a block can hold (every sequence but
// the last decodes at least 4 bytes).  A block with more sequences is
// parsed by the decode kernel itself.
__host__ __device__ __forceinline__ int seq_cap(int block_elems) {
  const int c = block_elems / 16 > 64 ? block_elems / 16 : 64;
  return c < block_elems / 2 + 1 ? c : block_elems / 2 + 1;
}

// Entry k of block d's sequence table: blocks in groups of 64, entry-major
// inside a group, so that the 64 lanes of a parse wave store one entry each
// to 512 contiguous bytes.
__device__ __forceinline__ long long seq_slot(long long d, int k, int cap) {
  return ((d >> 6) * (long long)cap + k) * 64 + (d & 63);
}

// One thread per block: walks the LZ4 sequence headers of its block (every
// length and offset checked as in the decode) and stores sequence k as
// {literal start | literal length << 16, offset | match length << 16}
// (offset and match length 0 for the block's last sequence).  nseq[d] is
// the count, 0 for unused slots and raw tails, -1 when the block has more
// sequences than the table holds (the decode kernel then parses it), -2
// for a malformed block (its status bit set).  Each thread reads its block
// through a 32-byte window of aligned vector loads.
__global__ void __launch_bounds__(256)
k_bslz4_parse(const uint8_t* __restrict__ comp, long long comp_bytes, const BlockDesc* __restrict__ table,
              long long n_desc, int cap, uint2* __restrict__ seqs, int* __restrict__ nseq, int* __restrict__ status) {
  const long long d = (long long)blockIdx.x * 256 + threadIdx.x;
  if (d >= n_desc) return;
  const BlockDesc bd = table[d];
  if (bd.elems == 0 || bd.csize < 0) {
    nseq[d] = 0;
    return;
  }
  const int cs = bd.csize;
  const int osz = bd.elems * 2;
  const uint8_t* const g = comp + bd.src;
  const uint8_t* const lim = comp + comp_bytes;
  int wb = -64;                                       // the window holds g[wb, wb + 32); g + wb 32-byte aligned
  // the window's words in separate registers (an indexed array would be placed in memory)
  unsigned w0 = 0u, w1 = 0u, w2 = 0u, w3 = 0u, w4 = 0u, w5 = 0u, w6 = 0u, w7 = 0u;
  auto byte = [&](int p) -> int {                     // 0 <= p < cs
    if ((unsigned)(p - wb) >= 32u) {
      const uint8_t* a = reinterpret_cast<const uint8_t*>(reinterpret_cast<uintptr_t>(g + p) & ~(uintptr_t)31);
      wb = (int)(a - g);
      if (a + 32 <= lim) {
        const uint4 v0 = reinterpret_cast<const uint4*>(a)[0], v1 = reinterpret_cast<const uint4*>(a)[1];
        w0 = v0.x, w1 = v0.y, w2 = v0.z, w3 = v0.w;
        w4 = v1.x, w5 = v1.y, w6 = v1.z, w7 = v1.w;
      } else {                                        // the buffer's last bytes: only those inside it
        auto word = [&](int o) {
          unsigned v = 0u;
          for (int i = 0; i < 4; ++i)
            if (a + o + i < lim) v |= (unsigned)a[o + i] << (8 * i);
          return v;
        };
        w0 = word(0), w1 = word(4), w2 = word(8), w3 = word(12);
        w4 = word(16), w5 = word(20), w6 = word(24), w7 = word(28);
      }
    }
    const int k = p - wb;
    const bool b2 = k & 4, b3 = k & 8;
    const unsigned x0 = b2 ? w1 : w0, x1 = b2 ? w3 : w2, x2 = b2 ? w5 : w4, x3 = b2 ? w7 : w6;
    const unsigned y0 = b3 ? x1 : x0, y1 = b3 ? x3 : x2;
    return (int)((((k & 16) ? y1 : y0) >> (8 * (k & 3))) & 0xFFu);
  };
  auto walk = [&](int& n) -> int {
    int ip = 0, op = 0;
    while (true) {
      if (ip >= cs) return kErrStream;
      const int token = byte(ip++);
      int lit = token >> 4;
      if (lit == 15) {
        int b;
        do {
          if (ip >= cs) return kErrStream;
          b = byte(ip++);
          lit += b;
        } while (b == 255);
      }
      if (lit > cs - ip || lit > osz - op) return kErrStream;
      const int lip = ip;
      ip += lit;
      op += lit;
      int off = 0, ml = 0;
      const bool last = ip == cs;
      if (!last) {
        if (ip + 2 > cs) return kErrStream;
        off = byte(ip) | (byte(ip + 1) << 8);
        ip += 2;
        ml = token & 15;
        if (ml == 15) {
          int b;
          do {
            if (ip >= cs) return kErrStream;
            b = byte(ip++);
            ml += b;
          } while (b == 255);
        }
        ml += 4;
        if (off == 0 || off > op || ml > osz - op) return kErrStream;
        op += ml;
      }
      if (n == cap) {
        n = -1;                                       // the decode kernel parses this block
        return 0;
      }
      seqs[seq_slot(d, n, cap)] = make_uint2((unsigned)(lip | (lit << 16)), (unsigned)(off | (ml << 16)));
      ++n;
      if (last) return op == osz ? 0 : kErrStream;
    }
  };
  int n = 0;
  const int err = walk(n);
  if (err) atomicOr(status, err);
  nseq[d] = err ? -2 : n;
}

// Literal runs shorter than this are copied one lane per sequence; longer
// ones by the whole wave.
constexpr int kShortLit = 16;

// The copies of one round of up to 64 sequences, lane j holding sequence j
// ({literal start | length << 16, offset | match length << 16}; 0 beyond n),
// output from op0; returns the round's output bytes.  Literal bytes come
// from src[0, src_end) (LDS or global).  Every literal run of the round is
// copied at once (its destination follows from the lengths before it: a
// wave prefix sum), then the matches in order, each one read of its source
// and one write; a match only reads bytes before its own destination, so
// the literals of later sequences, written first, are never read early.
// The workgroup is one wave, whose LDS operations complete in issue order,
// so the phases need only compiler ordering fences.
template <typename Src>
__device__ __forceinline__ int lz4_round(uint2 e, int n, int op0, int lane, const Src* src, int src_end,
                                         uint8_t* cout) {
  const int lit_l = (int)(e.x >> 16), lip_l = (int)(e.x & 0xFFFFu);
  const int ml_l = (int)(e.y >> 16), off_l = (int)(e.y & 0xFFFFu);
  const bool act = lane < n;
  const int len = act ? lit_l + ml_l : 0;
  int incl = len;
#pragma unroll
  for (int sh = 1; sh < 64; sh <<= 1) {
    const int y = __shfl_up(incl, sh);
    if (lane >= sh) incl += y;
  }
  const int lop = op0 + incl - len;                   // this lane's sequence starts here
  if (act && lit_l < kShortLit) {
    uint8_t b[kShortLit - 1];                         // read together (clamped inside src), then written
#pragma unroll
    for (int i = 0; i < kShortLit - 1; ++i) b[i] = src[min(lip_l + i, src_end - 1)];
#pragma unroll
    for (int i = 0; i < kShortLit - 1; ++i)
      if (i < lit_l) cout[lop + i] = b[i];
  }
  unsigned long long longs = __ballot(act && lit_l >= kShortLit);
  while (longs) {
    const int j = __builtin_ctzll(longs);
    longs &= longs - 1ull;
    const int o = __builtin_amdgcn_readlane(lop, j), s = __builtin_amdgcn_readlane(lip_l, j);
    const int L = __builtin_amdgcn_readlane(lit_l, j);
    for (int t = lane; t < L; t += 64) cout[o + t] = src[s + t];
  }
  asm volatile("" ::: "memory");                      // literals before the matches read them
  const int mdst = lop + lit_l;                       // match destination and source
  const int msrc = mdst - off_l;
  for (int j = 0; j < n; ++j) {
    const int ml = __builtin_amdgcn_readlane(ml_l, j);
    const int dst = __builtin_amdgcn_readlane(mdst, j);
    const int sp = __builtin_amdgcn_readlane(msrc, j);
    const int off = dst - sp;
    if (ml <= 64 && off >= ml) {                      // the common case (and the last sequence, ml = 0)
      if (lane < ml) cout[dst + lane] = cout[sp + lane];
    } else if (off >= ml) {                           // a long plain copy, 4 reads in flight per lane
      for (int t = lane; t < ml; t += 256) {
        uint8_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = cout[sp + min(t + 64 * u, ml - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (t + 64 * u < ml) cout[dst + t + 64 * u] = v[u];
      }
    } else if (off == 1) {                            // a run of one byte: 16-byte stores between the ends
      const unsigned v = (unsigned)cout[dst - 1] * 0x01010101u;
      const int a1 = dst + ml;
      const int h = min(a1, (dst + 15) & ~15);
      const int body = a1 & ~15;
      if (lane < h - dst) cout[dst + lane] = (uint8_t)v;
      for (int q = h + 16 * lane; q + 16 <= body; q += 1024)
        *reinterpret_cast<uint4*>(cout + q) = make_uint4(v, v, v, v);
      const int ts = max(h, body);
      if (lane < a1 - ts) cout[ts + lane] = (uint8_t)v;
    } else {                                          // period off: out[dst + t] = out[sp + t mod off],
      const float inv = __builtin_amdgcn_rcpf((float)off);   // all read from [sp, dst), 4 in flight per lane
      for (int t = lane; t < ml; t += 256) {
        uint8_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int tt = t + 64 * u;
          int r = tt - (int)((float)tt * inv) * off;  // the quotient is within 1 of tt / off
          r += (r < 0) ? off : 0;
          r -= (r >= off) ? off : 0;
          v[u] = cout[sp + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (t + 64 * u < ml) cout[dst + t + 64 * u] = v[u];
      }
    }
    asm volatile("" ::: "memory");
  }
  return __builtin_amdgcn_readlane(incl, 63);
}

__global__ void __launch_bounds__(64)
k_bslz4_decode(const uint8_t* __restrict__ comp, long long comp_bytes, const BlockDesc* __restrict__ table,
               long long n_desc, int cap, const uint2* __restrict__ seqs, const int* __restrict__ nseq,
               uint16_t* __restrict__ out, int* __restrict__ status) {
  __shared__ __attribute__((aligned(16))) unsigned cin32[kStageBytes / 4];
  __shared__ __attribute__((aligned(16))) uint8_t cout[kMaxBlockBytes];
  uint8_t* const cin = reinterpret_cast<uint8_t*>(cin32);
  const int lane = threadIdx.x;
  // Blocks in groups of 16 per XCD (workgroup w runs on XCD w mod 8 and
  // gridDim.x is a multiple of 8), so the 16 blocks whose sequence entries
  // share a 128-byte line read it through one L2.
  const long long npad = (n_desc + 127) / 128 * 128;
  for (long long w = blockIdx.x; w < npad; w += gridDim.x) {
    const long long i = w >> 3;
    const long long d = (i >> 4) * 128 + (w & 7) * 16 + (i & 15);
    if (d >= n_desc) continue;                         // uniform
    const BlockDesc bd = table[d];
    if (bd.elems == 0) continue;
    if (bd.csize < 0) {                                // raw tail, little-endian
      if (lane < bd.elems) {
        const uint8_t* s = comp + bd.src + 2 * lane;
        out[bd.dst + lane] = (uint16_t)(s[0] | (s[1] << 8));
      }
      continue;
    }
    const int n = nseq[d];
    if (n == -2) continue;                             // malformed: flagged by the parse
    const int cs = bd.csize;
    const int osz = bd.elems * 2;
    const uint8_t* gsrc = comp + bd.src;
    const int s0 = (int)(reinterpret_cast<uintptr_t>(gsrc) & 3u);
    const bool staged = s0 + cs <= kStageBytes;        // uniform
    __syncthreads();                                   // previous block's readers are done
    if (staged) {
      // the payload and the bytes around it up to whole 4-byte words, each
      // word read only where it lies inside the compressed buffer
      const uint8_t* const a0 = gsrc - s0;
      const uint8_t* const lim = comp + comp_bytes;
      const int nw = (s0 + cs + 3) >> 2;
      for (int w4 = lane; w4 < nw; w4 += 64) {
        const uint8_t* q = a0 + 4 * w4;
        unsigned v = 0u;
        if (q >= comp && q + 4 <= lim) {
          v = *reinterpret_cast<const unsigned*>(q);
        } else {
          for (int k = 0; k < 4; ++k)
            if (q + k >= comp && q + k < lim) v |= (unsigned)q[k] << (8 * k);
        }
        cin32[w4] = v;
      }
      __syncthreads();
    }
    int err = 0;
    if (n >= 0) {
      int op = 0;
      for (int r0 = 0; r0 < n; r0 += 64) {
        const int m = min(64, n - r0);
        const uint2 e = (lane < m) ? seqs[seq_slot(d, r0 + lane, cap)] : make_uint2(0u, 0u);
        op += staged ? lz4_round(e, m, op, lane, cin + s0, cs, cout) : lz4_round(e, m, op, lane, gsrc, cs, cout);
      }
    } else {
      // more sequences than the table holds: every lane walks the same
      // headers here and the copies follow each header
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
          if (ip == cs) return 0;                        // the last sequence has no match
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
          asm volatile("" ::: "memory");                 // literals before the match reads them
          if (off >= ml) {
            for (int t = lane; t < ml; t += 64) cout[op + t] = cout[op - off + t];
          } else {
            for (int t = lane; t < ml; t += 64) cout[op + t] = cout[op - off + (t % off)];
          }
          op += ml;
          asm volatile("" ::: "memory");
        }
      };
      int ip = 0, op = 0;
      err = staged ? parse([&](int k) -> int { return cin[s0 + k]; }, ip, op)
                   : parse([&](int k) -> int { return gsrc[k]; }, ip, op);
      if (!err && op != osz) err = kErrStream;
    }
    if (err) {
      if (lane == 0) atomicOr(status, err);
      continue;                                        // uniform: every lane saw the same stream
    }
    __syncthreads();
    // inverse bit-shuffle, 8 elements per group: the 16 plane bytes of the
    // group form two 8x8 bit matrices (planes 0-7 and 8-15, byte = plane,
    // bit = element); each is transposed in a 64-bit word by three delta
    // swaps, after which byte i holds element i's low or high byte, and
    // v_perm interleaves them into the output words.  With whole words per
    // plane, a lane takes 4 groups: one LDS word per plane, 64 bytes out.
    const int planeb = bd.elems / 8;
    auto group = [&](const unsigned (&bt)[16]) -> uint4 {
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
      return w;
    };
    if ((planeb & 3) == 0 && (bd.dst & 7) == 0) {      // uniform
      const unsigned* c32 = reinterpret_cast<const unsigned*>(cout);
      const int pw = planeb >> 2;
      for (int g4 = lane; g4 < pw; g4 += 64) {
        unsigned wd[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) wd[b] = c32[b * pw + g4];
        uint4* o = reinterpret_cast<uint4*>(out + bd.dst + 32LL * g4);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          unsigned bt[16];
#pragma unroll
          for (int b = 0; b < 16; ++b) bt[b] = (wd[b] >> (8 * k)) & 0xFFu;
          o[k] = group(bt);
        }
      }
    } else {
      for (int g = lane; g < planeb; g += 64) {
        unsigned bt[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) bt[b] = cout[b * planeb + g];
        const uint4 w = group(bt);
        if ((bd.dst & 7) == 0) {                       // uniform: 16-byte aligned groups
          *reinterpret_cast<uint4*>(out + bd.dst + 8LL * g) = w;
        } else {
          const unsigned ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
          for (int k = 0; k < 8; ++k) out[bd.dst + 8LL * g + k] = (uint16_t)(ww[k >> 1] >> (16 * (k & 1)));
        }
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

// Workspace: the status word, the block table, the per-block sequence
// counts and the sequence table, each from a 256-byte boundary.
namespace {
long long round256(long long x) { return (x + 255) / 256 * 256; }
struct WsLayout {
  long long table, nseq, seqs, total;
};
WsLayout ws_layout(long long n_desc, int block_elems) {
  WsLayout l;
  l.table = 256;
  l.nseq = l.table + round256(n_desc * (long long)sizeof(BlockDesc));
  l.seqs = l.nseq + round256(n_desc * 4);
  l.total = l.seqs + (n_desc + 63) / 64 * 64 * (long long)seq_cap(block_elems) * (long long)sizeof(uint2);
  return l;
}
}  // namespace

long long bslz4_workspace_bytes(int n_chunks, long long chunk_elems, int block_elems) {
  return ws_layout((long long)n_chunks * bslz4_slots(chunk_elems, block_elems), block_elems).total;
}

hipError_t launch_bslz4_decode_u16(const uint8_t* comp, long long comp_bytes, const long long* chunk_off,
                                   const long long* chunk_len, int n_chunks, long long chunk_elems, int block_elems,
                                   uint16_t* out, void* workspace, hipStream_t s) {
  const long long slots = bslz4_slots(chunk_elems, block_elems);
  const long long n_desc = (long long)n_chunks * slots;
  const WsLayout l = ws_layout(n_desc, block_elems);
  char* const ws = reinterpret_cast<char*>(workspace);
  int* status = reinterpret_cast<int*>(ws);
  BlockDesc* table = reinterpret_cast<BlockDesc*>(ws + l.table);
  int* nseq = reinterpret_cast<int*>(ws + l.nseq);
  uint2* seqs = reinterpret_cast<uint2*>(ws + l.seqs);
  const int cap = seq_cap(block_elems);
  hipLaunchKernelGGL(k_bslz4_clear, dim3(1), dim3(64), 0, s, status);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bslz4_index, dim3((unsigned)((n_chunks + 255) / 256)), dim3(256), 0, s, comp, comp_bytes,
                     chunk_off, chunk_len, n_chunks, chunk_elems, block_elems, (int)slots, table, status);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (n_desc == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bslz4_parse, dim3((unsigned)((n_desc + 255) / 256)), dim3(256), 0, s, comp, comp_bytes, table,
                     n_desc, cap, seqs, nseq, status);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long long npad = (n_desc + 127) / 128 * 128;   // a multiple of 8: workgroup w stays on XCD w mod 8
  const unsigned grid = (unsigned)(npad < 262144 ? npad : 262144);
  hipLaunchKernelGGL(k_bslz4_decode, dim3(grid), dim3(64), 0, s, comp, comp_bytes, table, n_desc, cap, seqs, nseq,
                     out, status);
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
