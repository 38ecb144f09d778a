// ZFP streams on the device, the parts around the block coders
// (zfp_lane_kernels.hpp): the serial walk of a variable-size stream without a
// block index (where each block begins), the placement of variable-size
// blocks from their scratch slots into the stream, the block index written
// beside a stream, stream headers and the batch's bookkeeping kernels.
//
// The stream layout is the published ZFP one (Lindstrom, IEEE TVCG 2014;
// LLNL ZFP stream format, codec 5), which the reference reaches through
// zfp-rs 0.1.1 (crates/ndic-zfp/src/lib.rs:249-376).
#include "zfp.hpp"
#include "zfp_dev.hpp"

#include <mutex>

#include <hipcub/hipcub.hpp>

namespace ndic {

namespace {

using namespace zfpdev;

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

// Variable-size modes: moves block g's bits from its scratch slot to its
// place in the stream (header + exclusive prefix sum of the lengths).
__global__ __launch_bounds__(256) void zfp_place_kernel(const ZfpChunk* __restrict__ chunks, uint32_t nchunks,
                                                        uint64_t total, const uint64_t* __restrict__ scan) {
  uint64_t g = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed || !c.offs || !c.dst) return;  // fixed-rate (written in place) or refused
  uint64_t len = c.offs[g];
  if (len == 0) return;
  uint64_t d0 = c.header_bits + scan[g] - scan[c.block_begin];
  const uint64_t* src = c.slots + g * c.slot_words;
  uint64_t* dst = (uint64_t*)c.dst;
  uint64_t first = d0 >> 6, last = (d0 + len - 1) >> 6;
  for (uint64_t dw = first + lane(); dw <= last; dw += 64) {
    uint64_t wb = dw * 64;
    uint64_t lo = wb > d0 ? wb : d0, hi = wb + 64 < d0 + len ? wb + 64 : d0 + len;
    uint64_t sb = lo - d0;
    uint32_t cnt = (uint32_t)(hi - lo);
    uint64_t sw = sb >> 6;
    uint32_t so = (uint32_t)(sb & 63);
    uint64_t v = src[sw] >> so;
    if (so && so + cnt > 64) v |= src[sw + 1] << (64 - so);
    if (cnt < 64) v &= (1ull << cnt) - 1;
    v <<= (uint32_t)(lo - wb);
    if (cnt == 64) dst[dw] = v;
    else atomicOr((unsigned long long*)&dst[dw], (unsigned long long)v);
  }
}

// ---------------------------------------------------------------------------
// The serial walk of a variable-size stream (which has no block index): where
// each block begins, found by parsing the blocks' bits without decoding them.
// One walking wavefront, every value of it but the window wave-uniform, beside
// producer waves that evaluate the group tests at every stream position ahead
// of it (zfp_index_kernel below).
//
// The walking wave reads the stream through a window of 127 words: lane l
// holds words first + l (cur) and first + 64 + l (nxt, loaded while cur is
// parsed).  When a read reaches cur's last word the window moves on by 63
// words, cur taking its own last word and nxt's first 63; reads only move
// forward, 64 stream bits at a time (WalkWindow::peek).
//
// A bit plane's group tests are parsed a word at a time rather than a test at
// a time: a run of significance tokens "1 0^z 1" ends at the first token whose
// closing one is followed by a zero, which a prefix parity of the ones finds.
// Where that cannot decide (a token longer than the word, a bit budget below
// 64) a turn parses one test.  The walk consumes exactly the bits the block
// decoder does (the oracle's decode_float_block / decode_int_block /
// decode_ints), budgets and minimum sizes included.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t i) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)i);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)i);
  return ((uint64_t)hi << 32) | lo;
}

// inclusive prefix parity: bit p of the result is the parity of x's bits 0..p
__device__ __forceinline__ uint64_t prefix_parity(uint64_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return x;
}

struct WalkWindow {
  GlobalSrc g;
  uint64_t first;     // first word of cur (wave-uniform)
  uint64_t cur, nxt;  // lane l: words first + l and first + 64 + l

  // cur's load is waited for at once (the empty asm reads it), so that only
  // nxt's is ever in flight: a wait for cur, where the compiler copies it
  // between registers, then never waits for nxt as well
  __device__ void load_cur(uint64_t w) {
    cur = g.load(w + lane());
    __asm__ volatile("" : : "v"(cur));
  }
  __device__ void init(const uint8_t* base, uint64_t nbytes, uint64_t bit) {
    g.base = base;
    g.nbytes = nbytes;
    first = bit >> 6;
    load_cur(first);
    nxt = g.load(first + 64 + lane());
  }
  // the 64 stream bits from bit `pos` on, the first in bit 0 (pos never
  // moves back, and never by 2^38 bits at once)
  __device__ uint64_t peek(uint64_t pos) {
    const uint64_t w = pos >> 6;
    uint32_t i = (uint32_t)(w - first);
    if (i >= 63) {     // word w + 1 is past cur
      if (i < 126) {   // within the window: move on by 63 words
        const uint64_t last = readlane64(cur, 63);
        const uint64_t up = __shfl_up(nxt, 1);
        cur = lane() == 0 ? last : up;
        first += 63;
      } else {
        first = w;
        load_cur(w);
      }
      nxt = g.load(first + 64 + lane());
      i = (uint32_t)(w - first);
    }
    return peek_at(pos, i);
  }
  // the same where word pos / 64 is word i < 63 of cur (any i < 64 reads
  // within cur)
  __device__ uint64_t peek_at(uint64_t pos, uint32_t i) const {
    const uint64_t lo = readlane64(cur, i);
    const uint64_t hi = readlane64(cur, (i + 1) & 63);
    const uint32_t s = (uint32_t)pos & 63;
    return (lo >> s) | ((hi << 1) << (63 - s));
  }
};

// One word of a bit plane's group tests, from coefficient count n.  Ones at
// odd ranks open tokens "1 0^z 1", ones at even ranks close them; a closing
// one followed by a zero ends the plane, the zero being its last test.  The
// prefix parity of the ones is taken across the wave: lane j holds the count
// of ones in bits 0..j (mlo / mhi are lane j's masks of bits 0..j).
struct TestWord {
  uint64_t closing;      // the closing ones (bits 0..62)
  uint32_t n2, used, p;  // count after the word's tests, bits they take, last closing one taken
  bool z0, end;          // first test negative; the plane ends in the word
  __device__ __forceinline__ TestWord(uint64_t x, uint32_t n, uint32_t mlo, uint32_t mhi) {
    z0 = !(x & 1);
    const uint32_t cnt = __builtin_popcount((uint32_t)x & mlo) + __builtin_popcount((uint32_t)(x >> 32) & mhi);
    closing = x & ~__ballot(cnt & 1) & ~(1ull << 63);
    const uint64_t stop = closing & ~(x >> 1);
    const uint32_t pa = (uint32_t)__builtin_ctzll(stop | (1ull << 63));
    const uint32_t pb = 63 - (uint32_t)__builtin_clzll(closing | 1);
    p = stop ? pa : pb;
    const uint32_t tokens = (uint32_t)__builtin_amdgcn_readlane((int)cnt, (int)p) >> 1;
    n2 = z0 ? n : n + p + 1 - tokens;
    used = z0 ? 1 : (stop ? p + 2 : p + 1);
    end = z0 | (stop != 0);
  }
};

// The test word at every stream position, computed ahead of the walk.  For
// the group tests of a plane starting at bit t, what the word at t decides
// does not depend on the coefficient count n (only its last-coefficient check
// does), so other waves compute it for every t in parallel while the walking
// wave only chains lookups.  A ring slot holds the number of the chunk of
// kSlotBits positions it belongs to (high 32 bits) and the entry (low 32):
//   bits 0..8   C, the coefficients the word's tests make significant
//   bits 9..16  A, the bits the tests take, plus C where E is set
//   bit 17      E, set where the plane's tests end in the word (a negative
//               first test ends them too)
// so that a turn from count n takes A + (E ? n : 0) bits -- the tests, then
// the next plane's n + C verbatim bits -- and ends E planes.  A token longer
// than the word has C = 511 (its zeros may reach the last coefficient), or,
// in 4D blocks, whose tokens reach past the word, no entry: its slot carries
// a chunk number no walk reads.
constexpr uint32_t kEntryEnds = 1u << 17;

template <int D>
__device__ __forceinline__ uint64_t test_word_slot(uint64_t x, uint32_t chunk) {  // x: the 64 bits from t on
  const uint64_t tag = (uint64_t)chunk << 32;
  if (!(x & 1)) return tag | kEntryEnds | (1u << 9);  // C = 0, A = 1
  const uint64_t closing = x & ~prefix_parity(x) & ~(1ull << 63);
  if (!closing) return D < 4 ? tag | 511u : (uint64_t)(chunk ^ 0x80000000u) << 32;
  const uint64_t stop = closing & ~(x >> 1);
  const uint32_t p = stop ? (uint32_t)__builtin_ctzll(stop) : 63 - (uint32_t)__builtin_clzll(closing);
  const uint32_t tokens = (uint32_t)__builtin_popcountll(x & (~0ull >> (63 - p))) >> 1;
  const uint32_t cnt = p + 1 - tokens;
  return tag | (stop ? kEntryEnds | (p + 2 + cnt) << 9 : (p + 1) << 9) | cnt;
}

// The ring of test-word entries in LDS: kRingSlots slots of kSlotBits
// positions; chunk c (positions base + c kSlotBits ..) goes to slot
// c % kRingSlots.  Producers take batches of 64 lanes x kPerLane<D>
// consecutive positions from a shared counter (so all of them work on the
// chunks nearest the walk; a lane's positions share its three stream words)
// and write a batch
// only while its chunk is within kRingSlots of the chunk the walk reads
// (`reading`); when the walk moves on, it raises the counter past the
// batches it no longer needs.  The walk takes an entry only when its chunk
// number is the one the walk reads, so that an entry not yet written, or one
// overwritten by a producer that fell behind, can cost time (that turn parses
// the stream words) but never the result.
constexpr uint32_t kWalkProducers = 15;
constexpr uint32_t kWalkThreads = 64 * (1 + kWalkProducers);
constexpr uint32_t kRingSlots = 16, kSlotBits = 512;
constexpr uint32_t kWalkSpins = 1u << 22;
// positions a producer lane takes per batch: four where blocks are short
// enough that the walk reads most positions; one for 4D blocks, whose long
// verbatim runs the walk jumps over
template <int D> constexpr uint32_t kPerLane = D < 4 ? 4 : 1;

__device__ __forceinline__ uint32_t lds_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_release(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Variable-size modes: one workgroup per field walks its blocks in stream
// order and records where each begins (the format has no block index).  Wave
// 0 walks; waves 1.. fill the ring of test words ahead of it.  Lane j of the
// walking wave keeps the offset of block 64 q + j; it stores 64 at a time.
//
// A block takes three kinds of turn.  Its header is parsed from the stream
// words.  Its group tests are taken from the ring a word at a time, each turn
// reading the entry its predecessor points at before it checks its own (the
// next plane's verbatim bits come with the entry that ends a plane).  When an
// entry says that the last coefficient becomes significant in its word, the
// word is parsed for where (the tests end there and the planes left are
// verbatim) and the block ends.  Everything else -- an entry that is missing,
// a token longer than the word, a bit budget near its end, a block that runs
// out of planes before every coefficient is significant -- goes through the
// general turn, which parses one word (or one test) of the stream.
template <int D, class S>
__global__ __launch_bounds__(kWalkThreads) void zfp_index_kernel(const ZfpChunk* __restrict__ chunks,
                                                                 uint32_t nchunks) {
  typedef __attribute__((address_space(1))) uint64_t* g64;
  constexpr bool fp = ZT<S>::fp;
  using BInt = typename ZT<S>::Int;  // the coded integers
  constexpr uint32_t SZ = 1u << (2 * D);
  constexpr uint32_t intprec = sizeof(BInt) * 8;
  constexpr uint32_t pbits = sizeof(BInt) == 4 ? 5 : 6;
  constexpr uint32_t ebits = ZT<S>::ebits;
  constexpr int ebias = ZT<S>::ebias;
  constexpr uint32_t kRing = kRingSlots * kSlotBits;
  enum { kHeader, kTests, kZeroRun };

  uint32_t ci = blockIdx.x;
  if (ci >= nchunks) return;
  const ZfpChunk& c = chunks[ci];
  if (c.fixed || c.index) return;  // computed addressing, or a block index
  const int minbits = (int)c.minbits, maxbits = (int)c.maxbits, cmaxprec = (int)c.maxprec, minexp = c.minexp;
  const bool rev = minexp < kZfpMinExp;
  const uint64_t nblocks = c.nblocks;
  const g64 offs = (g64)(uintptr_t)(c.offs + c.block_begin);
  const uint32_t l = lane();
  const uint64_t base = c.header_bits;
  const uint64_t stream_bits = c.src_len * 8;
  const uint32_t nring = stream_bits > base ? (uint32_t)((stream_bits - base + kSlotBits - 1) / kSlotBits) : 0u;

  __shared__ uint64_t ring[kRing];
  __shared__ uint32_t reading, next_batch, done;
  for (uint32_t i = threadIdx.x; i < kRing; i += kWalkThreads) ring[i] = ~0ull << 32;  // no chunk's
  if (threadIdx.x == 0) reading = next_batch = done = 0;
  __syncthreads();

  constexpr uint32_t kPL = kPerLane<D>;
  constexpr uint32_t kBatches = kSlotBits / (64 * kPL);  // batches per chunk
  if (threadIdx.x >= 64) {  // a producer
    GlobalSrc g{c.src, c.src_len};
    const uint64_t nbatches = (uint64_t)nring * kBatches;
    for (;;) {
      uint32_t b = 0;
      if (l == 0) b = __hip_atomic_fetch_add(&next_batch, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      b = (uint32_t)__builtin_amdgcn_readfirstlane((int)b);
      if (b >= nbatches || lds_acquire(&done)) break;
      const uint32_t ch = b / kBatches;
      bool go = false;
      for (uint32_t spin = 0; spin < kWalkSpins; ++spin) {  // waits lightly: the walk reads the same LDS
        const uint32_t at = lds_acquire(&reading);
        if (at > ch || ((spin & 7) == 7 && lds_acquire(&done))) break;  // passed by the walk: not needed
        if (at + kRingSlots > ch) {
          go = true;
          break;
        }
        __builtin_amdgcn_s_sleep(4);
      }
      if (!go) continue;
      const uint32_t r = (b % kBatches) * (64 * kPL) + kPL * l;
      const uint64_t t0 = base + (uint64_t)ch * kSlotBits + r;
      const uint64_t w = t0 >> 6;
      const uint32_t sh = (uint32_t)t0 & 63;
      const uint64_t w0 = g.load(w), w1 = g.load(w + 1), w2 = kPL > 1 ? g.load(w + 2) : 0;
      const uint64_t lo = (w0 >> sh) | ((w1 << 1) << (63 - sh));  // bits t0 .. t0 + 63
      const uint64_t hi = (w1 >> sh) | ((w2 << 1) << (63 - sh));  // bits t0 + 64 ..
      uint64_t* const slots = &ring[(ch % kRingSlots) * kSlotBits + r];
#pragma unroll
      for (uint32_t q = 0; q < kPL; ++q)
        __hip_atomic_store(slots + q, test_word_slot<D>(q ? (lo >> q) | (hi << (64 - q)) : lo, ch), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return;
  }

  const uint32_t mlo = l < 32 ? (2u << l) - 1 : ~0u;
  const uint32_t mhi = l < 32 ? 0u : (2u << (l - 32)) - 1;
  WalkWindow win;
  win.init(c.src, c.src_len, c.header_bits);

  uint64_t pos = base, blk_start = pos, kb = 0, mine = 0;
  uint32_t bits = 0, n = 0, k = 0, kmin = 0, zacc = 0, zlim = 0;
  uint32_t cur_ch = ~0u;
  int st = kHeader;
  const auto follow = [&]() {  // let the producers move on to the chunk the walk reads
    const uint32_t ch = (uint32_t)((pos - base) / kSlotBits);
    if (ch != cur_ch) {
      lds_release(&reading, ch);
      __hip_atomic_fetch_max(&next_batch, ch * kBatches, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      cur_ch = ch;
    }
  };
  const auto slot_at = [&](uint32_t off) {  // the ring slot at byte offset off (modulo the ring)
    uint64_t* const at = (uint64_t*)__builtin_assume_aligned((char*)ring + (off & (kRing * 8 - 8)), 8);
    return __hip_atomic_load(at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  const auto end_block = [&]() {
    if (pos - blk_start < (uint64_t)(minbits > 0 ? minbits : 0)) pos = blk_start + (uint32_t)minbits;
    ++kb;
    st = kHeader;
  };
  for (;;) {
    uint64_t te = 0;  // the slot of the turn at pos, when the header read it
    bool have_te = false;
    if (st == kHeader) {
      if (kb == nblocks) break;
      const uint64_t x = win.peek(pos);
      int hdr = 0, prec = cmaxprec;
      if constexpr (fp) {
        if (!(x & 1)) {  // every value zero: one bit (reversible) or minbits
          const uint32_t j = (uint32_t)kb & 63;
          if (l == j) mine = pos;
          if (j == 63) offs[kb - 63 + l] = mine;
          pos += rev || minbits <= 1 ? 1u : (uint32_t)minbits;
          ++kb;
          continue;
        }
        if (rev) {  // 1 1: raw bits; 1 0: a block exponent follows
          hdr = (x & 2) ? 2 : 2 + (int)ebits;
          prec = (int)((x >> hdr) & ((1u << pbits) - 1)) + 1;
          hdr += (int)pbits;
        } else {
          const int emax = (int)((x >> 1) & ((1u << ebits) - 1)) - ebias;
          const int dp = emax - minexp + 2 * (D + 1);
          prec = cmaxprec < (dp > 0 ? dp : 0) ? cmaxprec : (dp > 0 ? dp : 0);
          hdr = 1 + (int)ebits;
        }
      } else if (rev) {
        prec = (int)(x & ((1u << pbits) - 1)) + 1;
        hdr = (int)pbits;
      }
      follow();
      te = slot_at((uint32_t)(pos + (uint32_t)hdr - base) << 3);  // the first plane's tests, read early
      have_te = true;
      const uint32_t j = (uint32_t)kb & 63;
      if (l == j) mine = pos;
      if (j == 63) offs[kb - 63 + l] = mine;
      blk_start = pos;
      pos += (uint32_t)hdr;
      bits = maxbits - hdr > 0 ? (uint32_t)(maxbits - hdr) : 0u;
      n = 0;
      k = intprec;
      kmin = (uint32_t)__builtin_amdgcn_readfirstlane(max((int)intprec - prec, 0));
      if (!bits || k <= kmin) {  // no bit plane
        end_block();
        continue;
      }
      --k;  // the first plane: no verbatim bits, its tests follow
      st = kTests;
    }
    if (st == kTests && n < SZ) {  // turns from the ring
      // Each turn reads the slot its entry points at before it checks the
      // entry.  Where the bits the block can still take stay above the
      // budget less a word and a plane, the budget cannot end it early.
      uint64_t rel = pos - base;
      const uint64_t rel0 = rel;
      uint32_t o = (uint32_t)rel << 3;  // rel's slot, in bytes (modulo the ring)
      uint32_t kl = k - kmin;           // planes after this one
      if (!have_te) te = slot_at(o);    // read into vector registers; made scalar where used
      bool last = false;
      const auto turns = [&](auto bounded) {
        const uint32_t lim = bits > 65 + SZ ? bits - 65 - SZ : 0u;
        for (;;) {
          const uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)te);
          const uint32_t tag = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(te >> 32));
          const uint32_t step = ((e >> 9) & 255) + (n & (uint32_t)((int32_t)(e << 14) >> 31));
          const uint32_t o2 = o + (step << 3);
          const uint64_t te2 = slot_at(o2);
          if (tag != (uint32_t)(rel / kSlotBits)) break;
          if (decltype(bounded)::value && (uint32_t)(rel - rel0) >= lim) break;
          const uint32_t n2 = n + (e & 511);
          if (n2 >= SZ) {  // the last coefficient becomes significant in the word
            last = true;
            break;
          }
          const uint32_t ends = (e >> 17) & 1;
          if (kl < ends) break;  // the block ends with this plane
          n = n2;
          rel += step;
          o = o2;
          kl -= ends;
          te = te2;
        }
      };
      // the rest of the block takes at most kl SZ verbatim bits and 2 SZ + kl + 1 tests
      const uint32_t rest = (kl + 2) * (SZ + 1) + SZ;
      if (bits > rest + 65 + SZ) turns(std::false_type{});
      else turns(std::true_type{});
      pos = base + rel;
      bits -= (uint32_t)(rel - rel0);
      k = kmin + kl;
      if (last) {
        // The word's tests reach the last coefficient: its token is the one at
        // which the count of bits other than closing ones (the even-ranked
        // ones) reaches SZ - n; in a token longer than the word, the zeros
        // reach it (the last one implied).  Where neither is in the word's
        // first 63 bits, the general turn takes the test.  The planes left
        // are SZ verbatim bits each.
        const uint64_t x = win.peek(pos);
        const uint32_t ones = __builtin_popcount((uint32_t)x & mlo) + __builtin_popcount((uint32_t)(x >> 32) & mhi);
        const uint64_t hit = __ballot(l + 1 - (ones >> 1) == SZ - n);
        if (hit) {
          const uint32_t q = (uint32_t)__builtin_ctzll(hit);
          pos += q + 1;
          bits -= q + 1;
          n = SZ;
          if (k > kmin) {
            --k;
            const uint64_t need = (uint64_t)(k + 1 - kmin) * SZ;
            const uint32_t take = need < bits ? (uint32_t)need : bits;
            pos += take;
            bits -= take;
          }
          end_block();
          continue;
        }
      }
    }
    // the general turn
    follow();
    const uint64_t x = win.peek(pos);
    bool plane_end = false;
    if (st == kZeroRun) {  // a test's zeros that ran past the last word
      const uint32_t t = x ? (uint32_t)__builtin_ctzll(x) : 64u;
      if (t < zlim && t < 64) {  // the closing one
        pos += t + 1;
        bits -= t + 1;
        n += zacc + t + 1;
        st = kTests;
      } else if (zlim <= 64) {  // the zeros reach the limit: no closing one
        pos += zlim;
        bits -= zlim;
        n += zacc + zlim + 1;
        st = kTests;
      } else {
        pos += 64;
        bits -= 64;
        zacc += 64;
        zlim -= 64;
      }
    } else if (n >= SZ || !bits) {
      plane_end = true;
    } else {
      const TestWord tw(x, n, mlo, mhi);
      if (bits >= 64 && tw.z0) {
        pos += 1;
        bits -= 1;
        plane_end = true;
      } else if (bits >= 64 && tw.closing && tw.n2 >= SZ) {
        const uint64_t other = ~tw.closing;
        const uint32_t cnt =
            __builtin_popcount((uint32_t)other & mlo) + __builtin_popcount((uint32_t)(other >> 32) & mhi);
        const uint32_t q = (uint32_t)__builtin_ctzll(__ballot(cnt == SZ - n) & other);
        pos += q + 1;
        bits -= q + 1;
        n = SZ;
        plane_end = true;
      } else if (bits >= 64 && tw.closing) {
        n = tw.n2;
        pos += tw.used;
        bits -= tw.used;
        plane_end = tw.end;
      } else {  // one test: a budget below 64, or a token longer than the word
        bits--;
        if (!(x & 1)) {
          pos += 1;
          plane_end = true;
        } else {
          uint32_t lim = SZ - 1 - n;
          if (lim > bits) lim = bits;
          const uint64_t y = x >> 1;
          const uint32_t t = y ? (uint32_t)__builtin_ctzll(y) : 64u;
          if (t < lim && t < 64) {  // the closing one, in view
            pos += t + 2;
            bits -= t + 1;
            n += t + 1;
          } else if (lim <= 63) {  // the zeros reach the limit in view
            pos += lim + 1;
            bits -= lim;
            n += lim + 1;
          } else {  // 63 zeros in view, more to come
            pos += 64;
            bits -= 63;
            zacc = 63;
            zlim = lim - 63;
            st = kZeroRun;
          }
        }
      }
    }
    plane_end |= (st == kTests) & (n >= SZ);  // the tests end with the last coefficient
    if (plane_end) {
      bool block_end = true;
      if (bits && k > kmin) {
        --k;
        if (n == SZ) {  // every coefficient significant: the planes left are SZ verbatim bits each
          const uint64_t need = (uint64_t)(k + 1 - kmin) * SZ;
          const uint32_t take = need < bits ? (uint32_t)need : bits;
          pos += take;
          bits -= take;
        } else {
          const uint32_t m = n < bits ? n : bits;
          bits -= m;
          pos += m;
          st = kTests;
          block_end = false;
        }
      }
      if (block_end) end_block();
    }
  }
  lds_release(&done, 1);
  const uint64_t tail = nblocks & 63;
  if (l < tail) offs[nblocks - tail + l] = mine;
}

// Block index of each variable-size chunk that asks for one: entry k is
// block k's bit offset from the first block, entry nblocks the end.
__global__ __launch_bounds__(256) void zfp_index_write_kernel(const ZfpChunk* __restrict__ chunks, uint32_t nchunks,
                                                              uint64_t total, const uint64_t* __restrict__ scan) {
  const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const ZfpChunk& c = chunks[find_chunk(chunks, nchunks, g)];
  if (c.fixed || !c.index || !c.dst || !c.offs) return;
  const uint64_t k = g - c.block_begin;
  const uint64_t at = scan[g] - scan[c.block_begin];
  c.index[k] = (uint32_t)at;
  if (k + 1 == c.nblocks) c.index[k + 1] = (uint32_t)(at + c.offs[g]);
}

template <int D, class S> void idx(const ZfpChunk* d, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL((zfp_index_kernel<D, S>), dim3(n), dim3(kWalkThreads), 0, s, d, n);
}

#define NDIC_ZFP_DISPATCH(FN, ...)                                                     \
  switch (dims * 8 + scalar) {                                                         \
    case 1 * 8 + kZfpInt32: FN<1, int32_t>(__VA_ARGS__); break;                        \
    case 1 * 8 + kZfpInt64: FN<1, int64_t>(__VA_ARGS__); break;                        \
    case 1 * 8 + kZfpFloat: FN<1, float>(__VA_ARGS__); break;                          \
    case 1 * 8 + kZfpDouble: FN<1, double>(__VA_ARGS__); break;                        \
    case 2 * 8 + kZfpInt32: FN<2, int32_t>(__VA_ARGS__); break;                        \
    case 2 * 8 + kZfpInt64: FN<2, int64_t>(__VA_ARGS__); break;                        \
    case 2 * 8 + kZfpFloat: FN<2, float>(__VA_ARGS__); break;                          \
    case 2 * 8 + kZfpDouble: FN<2, double>(__VA_ARGS__); break;                        \
    case 3 * 8 + kZfpInt32: FN<3, int32_t>(__VA_ARGS__); break;                        \
    case 3 * 8 + kZfpInt64: FN<3, int64_t>(__VA_ARGS__); break;                        \
    case 3 * 8 + kZfpFloat: FN<3, float>(__VA_ARGS__); break;                          \
    case 3 * 8 + kZfpDouble: FN<3, double>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpInt32: FN<4, int32_t>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpInt64: FN<4, int64_t>(__VA_ARGS__); break;                        \
    case 4 * 8 + kZfpFloat: FN<4, float>(__VA_ARGS__); break;                          \
    case 4 * 8 + kZfpDouble: FN<4, double>(__VA_ARGS__); break;                        \
    default: fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");              \
  }

__global__ void zfp_headers_kernel(const ZfpHeader* __restrict__ h, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (uint32_t k = 0; k < h[i].nwords; ++k) h[i].dst[k] = h[i].w[k];
}

__global__ void zfp_totals_kernel(const ZfpChunk* __restrict__ chunks, uint32_t n, const uint64_t* __restrict__ scan,
                                  uint64_t* totals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const ZfpChunk& c = chunks[i];
  if (c.fixed || c.nblocks == 0) {
    totals[i] = 0;
    return;
  }
  uint64_t last = c.block_begin + c.nblocks - 1;
  totals[i] = scan[last] + c.offs[last] - scan[c.block_begin];
}

__global__ void zfp_heads_kernel(const ZfpHead* __restrict__ h, uint32_t n, uint8_t* out) {
  uint32_t i = blockIdx.x, b = threadIdx.x;
  if (i >= n || b >= kZfpHeadBytes) return;
  out[(uint64_t)i * kZfpHeadBytes + b] = b < h[i].len ? h[i].src[b] : 0;
}

}  // namespace

size_t zfp_scan_temp_bytes(uint64_t n) {
  size_t bytes = 0;
  NDIC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)n));
  return bytes;
}

void zfp_exclusive_scan(void* temp, size_t temp_bytes, const uint64_t* d_in, uint64_t* d_out, uint64_t n,
                        hipStream_t stream) {
  if (n == 0) return;
  if (n > 0x7fffffffull) fail(NDIC_INVALID_ARGUMENT, "zfp: too many blocks in one batch");
  NDIC_HIP(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, d_in, d_out, (int)n, stream));
}

void launch_zfp_headers(const ZfpHeader* d_headers, uint32_t n, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(zfp_headers_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, d_headers, n);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_totals(const ZfpChunk* d_chunks, uint32_t nchunks, const uint64_t* d_scan, uint64_t* d_totals,
                       hipStream_t stream) {
  if (!nchunks) return;
  hipLaunchKernelGGL(zfp_totals_kernel, dim3((nchunks + 63) / 64), dim3(64), 0, stream, d_chunks, nchunks, d_scan,
                     d_totals);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_heads(const ZfpHead* d_heads, uint32_t n, uint8_t* d_out, hipStream_t stream) {
  if (!n) return;
  hipLaunchKernelGGL(zfp_heads_kernel, dim3(n), dim3(64), 0, stream, d_heads, n, d_out);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_encode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims, uint32_t scalar,
                       hipStream_t stream) {
  if (total == 0) return;
  if (!launch_zfp_lane_encode(d_chunks, nchunks, total, dims, scalar, stream))
    fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");
}

void launch_zfp_place(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, const uint64_t* d_scan,
                      hipStream_t stream) {
  if (total == 0) return;
  hipLaunchKernelGGL(zfp_place_kernel, dim3((uint32_t)ceil_div(total, 4)), dim3(256), 0, stream, d_chunks, nchunks,
                     total, d_scan);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_index(const ZfpChunk* d_chunks, uint32_t nchunks, uint32_t dims, uint32_t scalar,
                      hipStream_t stream) {
  if (nchunks == 0) return;
  NDIC_ZFP_DISPATCH(idx, d_chunks, nchunks, stream);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_index_write(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, const uint64_t* d_scan,
                            hipStream_t stream) {
  if (total == 0) return;
  hipLaunchKernelGGL(zfp_index_write_kernel, dim3((uint32_t)ceil_div(total, 256)), dim3(256), 0, stream, d_chunks,
                     nchunks, total, d_scan);
  NDIC_HIP(hipGetLastError());
}

void launch_zfp_decode(const ZfpChunk* d_chunks, uint32_t nchunks, uint64_t total, uint32_t dims, uint32_t scalar,
                       const ZfpBrick* d_bricks, uint32_t fixed_maxbits, hipStream_t stream) {
  if (total == 0) return;
  if (!launch_zfp_lane_decode(d_chunks, nchunks, total, dims, scalar, d_bricks, fixed_maxbits, stream))
    fail(NDIC_INVALID_ARGUMENT, "zfp: unsupported field class");
}

}  // namespace ndic
