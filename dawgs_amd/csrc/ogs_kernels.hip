// ogs_kernels.hip -- gfx950 gather-scatter kernels and the operator-layer
// C-ABI (ogs_hip_*) declared in include/ogs_hip.h.
//
// What the kernels compute is fixed by the reference's host operator
// (libs/ogs/ogsOperator.cpp:36-378): for every CSR row, reduce the row's
// gather list in CSR order starting from the op's identity, then write the
// result to the row's scatter list.  How they compute it is MI355X-first:
//
//  * Partition.  A plan cuts the rows into blocks of contiguous rows: for a
//    list without empty rows row r belongs to block floor(rowStarts[r] / W)
//    (at most W + Lmax - 1 nonzeros, so no more rows than that either); for
//    a list with empty rows, and for the scatter kernels, by the merge-path
//    rule floor((rowStarts[r] + r) / W) (at most W rows and W + Lmax - 1
//    nonzeros).  Either way the per-block LDS staging never overflows, and
//    the cut is computed on the device by binary search.  Launches on small
//    lists take narrower blocks until the device is filled.
//  * Packs.  For the partitions a plan serves, each block's list entries are
//    sorted by column at plan creation and stored as 32-bit words (column
//    segment and offset, row-head bit, CSR position) at a fixed stride per
//    block, so a wave's gathers touch few cache lines and a kernel loads a
//    block's words from the block id alone.
//  * Loads.  Index words are read coalesced, each lane holding a fixed number
//    of slots; all slots' loads are issued before the first wait, and the
//    gathers run unpredicated, so a wave keeps all of them in flight.
//  * Reduction.  Gathered values are staged in LDS in CSR order.  With one
//    list (Sym, or an Unsigned operator) the row heads become a bit mask and
//    one lane per (row head, field chunk) folds the row in CSR order; with
//    two lists one lane per row reduces the gather list and writes the
//    scatter list.  Either way fp Add is bit-exact with the host loop.
//    Integer Add/Mul wrap (unsigned arithmetic); Min/Max are the
//    reference's compare-selects, so NaN inputs are skipped like on the host.
//    The precision-form storage types (bfloat16, float with compensated Add)
//    accumulate in float with a two-sum running error.
//  * XCDs.  Workgroups are dispatched round-robin over the 8 XCDs; block ids
//    are remapped so that each XCD walks one contiguous eighth of the rows
//    and neighbouring blocks share that XCD's L2.  Where rows sharing cache
//    lines sit a mesh layer of blocks apart (low degree), a plan may walk
//    each XCD's blocks in tiles of layers instead (kept only where measured
//    faster at plan creation).
//  * K fields.  Fields are interleaved (v[k + id*K]).  For K in {2, 3, 4, 8,
//    16} a work item is up to 4 consecutive fields of one entry (8-byte
//    values: 2, or 1 at K = 3), so one index word feeds contiguous value
//    loads; other K run in field chunks.  The
//    indices are read once per block rather than once per field as in the
//    reference (ogsKernels.okl:41, @outer(1) over K).
//
// No MFMA: nothing here is a dense contraction.  No fast-math: the build
// uses -ffp-contract=off and keeps IEEE semantics.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cstddef>
#include <cstdlib>
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <new>
#include <type_traits>
#include <vector>

#include "ogs_hip.h"
#include "ogs_internal.hpp"

namespace ogs_hip {

// ---------------------------------------------------------------------------
// Reduction ops: ogsOperator.cpp:36-55 (host functors) and ogsUtils.cpp:63-103
// (device inits and OGS_OP macros).  Inits are +-numeric_limits<T>::max(),
// not +-inf and not INT_MIN.
template <typename T> struct Lim;
template <> struct Lim<float> { static __device__ constexpr float max() { return FLT_MAX; } };
template <> struct Lim<double> { static __device__ constexpr double max() { return DBL_MAX; } };
template <> struct Lim<int32_t> { static __device__ constexpr int32_t max() { return INT32_MAX; } };
template <> struct Lim<int64_t> { static __device__ constexpr int64_t max() { return INT64_MAX; } };

template <typename T> __device__ __forceinline__ T wrap_add(T a, T b) { return a + b; }
template <typename T> __device__ __forceinline__ T wrap_mul(T a, T b) { return a * b; }
template <> __device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
template <> __device__ __forceinline__ int32_t wrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
template <> __device__ __forceinline__ int64_t wrap_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
template <> __device__ __forceinline__ int64_t wrap_mul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }

// Red<T, OP>: A is the accumulator, init() its start value, apply() folds
// one member in, fin() gives the value written back.  For the reference's
// four types A is T and fin() is the identity.
template <typename T, int OP> struct Red;
template <typename T> struct Red<T, OGS_ADD> {
  using A = T;
  static __device__ __forceinline__ A init() { return T(0); }
  static __device__ __forceinline__ void apply(A &a, T b) { a = wrap_add(a, b); }
  static __device__ __forceinline__ T fin(A a) { return a; }
};
template <typename T> struct Red<T, OGS_MUL> {
  using A = T;
  static __device__ __forceinline__ A init() { return T(1); }
  static __device__ __forceinline__ void apply(A &a, T b) { a = wrap_mul(a, b); }
  static __device__ __forceinline__ T fin(A a) { return a; }
};
template <typename T> struct Red<T, OGS_MAX> {
  using A = T;
  static __device__ __forceinline__ A init() { return -Lim<T>::max(); }
  static __device__ __forceinline__ void apply(A &a, T b) { a = (b > a) ? b : a; }
  static __device__ __forceinline__ T fin(A a) { return a; }
};
template <typename T> struct Red<T, OGS_MIN> {
  using A = T;
  static __device__ __forceinline__ A init() { return Lim<T>::max(); }
  static __device__ __forceinline__ void apply(A &a, T b) { a = (b < a) ? b : a; }
  static __device__ __forceinline__ T fin(A a) { return a; }
};

// ---------------------------------------------------------------------------
// Precision form (no counterpart in the reference): storage narrower than
// the accumulation.  OGS_FLOAT_C stores float; OGS_BF16 stores bfloat16.
// Add accumulates each row in float, in CSR order, with a two-sum running
// error (the cascaded sum "Sum2" of Ogita, Rump and Oishi, SIAM J. Sci.
// Comput. 26(6), 2005):
//   s = +0, c = 0;  per member x:  t = s + x, z = t - s,
//   e = (s - (t - z)) + (x - z)  (exact: s + x = t + e),  s = t, c += e;
//   result = s + c when s is finite, else s (an overflow or inf/NaN member
//   propagates as in a plain sum).
// Its error is at most u|S| + gamma_{n-1}^2 sum|x| (u = 2^-24), i.e. the
// float rounding of the exact row sum up to a term of order n^2 u^2.  bf16
// results are then rounded to nearest even.  Min/Max compare-select the
// (widened) stored values and are exact.  Mul is not offered for these types.
struct f32c { float v; };       // OGS_FLOAT_C storage
struct bf16s { uint16_t b; };   // OGS_BF16 storage

struct Comp {
  float s, c;
};
__device__ __forceinline__ void comp_add(Comp &a, float x) {
  const float t = a.s + x;
  const float z = t - a.s;
  const float e = (a.s - (t - z)) + (x - z);
  a.s = t;
  a.c += e;
}
__device__ __forceinline__ float comp_fin(const Comp &a) { return __builtin_isfinite(a.s) ? a.s + a.c : a.s; }
// comp_add of one member into {0, 0}: s = 0 + x (-0 becomes +0), c = 0.
// comp_add's c differs only for x inf or NaN, where s stays non-finite from
// then on and comp_fin ignores c, so the result is the same.
__device__ __forceinline__ Comp comp_first(float x) { return Comp{0.0f + x, 0.0f}; }

__device__ __forceinline__ float bf16_widen(bf16s x) { return __uint_as_float((uint32_t)x.b << 16); }
// Round to nearest even; NaN stays NaN (quieted), overflow rounds to inf.
__device__ __forceinline__ bf16s bf16_round(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  return bf16s{(uint16_t)((u & 0x7fffffffu) > 0x7f800000u ? (u >> 16) | 0x40u : r)};
}
constexpr float kBf16Max = 3.38953139e38f;  // 0x7f7f: largest finite bfloat16

template <> struct Red<f32c, OGS_ADD> {
  using A = Comp;
  static __device__ __forceinline__ A init() { return Comp{0.0f, 0.0f}; }
  static __device__ __forceinline__ void apply(A &a, f32c b) { comp_add(a, b.v); }
  static __device__ __forceinline__ f32c fin(const A &a) { return f32c{comp_fin(a)}; }
};
template <> struct Red<f32c, OGS_MAX> {
  using A = float;
  static __device__ __forceinline__ A init() { return -FLT_MAX; }
  static __device__ __forceinline__ void apply(A &a, f32c b) { a = (b.v > a) ? b.v : a; }
  static __device__ __forceinline__ f32c fin(A a) { return f32c{a}; }
};
template <> struct Red<f32c, OGS_MIN> {
  using A = float;
  static __device__ __forceinline__ A init() { return FLT_MAX; }
  static __device__ __forceinline__ void apply(A &a, f32c b) { a = (b.v < a) ? b.v : a; }
  static __device__ __forceinline__ f32c fin(A a) { return f32c{a}; }
};
template <> struct Red<bf16s, OGS_ADD> {
  using A = Comp;
  static __device__ __forceinline__ A init() { return Comp{0.0f, 0.0f}; }
  static __device__ __forceinline__ void apply(A &a, bf16s b) { comp_add(a, bf16_widen(b)); }
  static __device__ __forceinline__ bf16s fin(const A &a) { return bf16_round(comp_fin(a)); }
};
template <> struct Red<bf16s, OGS_MAX> {
  using A = float;
  static __device__ __forceinline__ A init() { return -kBf16Max; }
  static __device__ __forceinline__ void apply(A &a, bf16s b) {
    const float x = bf16_widen(b);
    a = (x > a) ? x : a;
  }
  static __device__ __forceinline__ bf16s fin(A a) { return bf16_round(a); }  // exact: a is a stored value
};
template <> struct Red<bf16s, OGS_MIN> {
  using A = float;
  static __device__ __forceinline__ A init() { return kBf16Max; }
  static __device__ __forceinline__ void apply(A &a, bf16s b) {
    const float x = bf16_widen(b);
    a = (x < a) ? x : a;
  }
  static __device__ __forceinline__ bf16s fin(A a) { return bf16_round(a); }
};

// LDS staging of values in the entry-major kernels: bfloat16 is staged
// widened to float, so that every LDS access is a whole dword (2-byte LDS
// writes to the two halves of one dword collide); other types as stored.
// LRed reduces staged values exactly as Red reduces stored ones.
template <typename T> struct Lds {
  using L = T;
  static __device__ __forceinline__ L in(T x) { return x; }
  static __device__ __forceinline__ T out(L x) { return x; }
};
template <> struct Lds<bf16s> {
  using L = float;
  static __device__ __forceinline__ float in(bf16s x) { return bf16_widen(x); }
  static __device__ __forceinline__ bf16s out(float f) { return bf16s{(uint16_t)(__float_as_uint(f) >> 16)}; }
};
template <typename T, int OP> struct LRed {
  using A = typename Red<T, OP>::A;
  static __device__ __forceinline__ A init() { return Red<T, OP>::init(); }
  static __device__ __forceinline__ void apply(A &a, T x) { Red<T, OP>::apply(a, x); }
  static __device__ __forceinline__ T fin(const A &a) { return Red<T, OP>::fin(a); }
};
template <> struct LRed<bf16s, OGS_ADD> {
  using A = Comp;
  static __device__ __forceinline__ A init() { return Comp{0.0f, 0.0f}; }
  static __device__ __forceinline__ void apply(A &a, float x) { comp_add(a, x); }
  static __device__ __forceinline__ float fin(const A &a) { return bf16_widen(bf16_round(comp_fin(a))); }
};
template <> struct LRed<bf16s, OGS_MAX> {
  using A = float;
  static __device__ __forceinline__ A init() { return -kBf16Max; }
  static __device__ __forceinline__ void apply(A &a, float x) { a = (x > a) ? x : a; }
  static __device__ __forceinline__ float fin(A a) { return a; }  // a stored value: already a bf16
};
template <> struct LRed<bf16s, OGS_MIN> {
  using A = float;
  static __device__ __forceinline__ A init() { return kBf16Max; }
  static __device__ __forceinline__ void apply(A &a, float x) { a = (x < a) ? x : a; }
  static __device__ __forceinline__ float fin(A a) { return a; }
};

// A staged value of the compensated types as float.
__device__ __forceinline__ float lds_float(float x) { return x; }
__device__ __forceinline__ float lds_float(f32c x) { return x.v; }

// Whether (T, OP) has kernels: every op for the reference's types; Add,
// Max and Min for the precision-form storage types.
template <typename T, int OP>
constexpr bool kOffered = !((std::is_same<T, f32c>::value || std::is_same<T, bf16s>::value) && OP == OGS_MUL);

// Workgroup id -> block id such that XCD x (which receives workgroups
// x, x+8, x+16, ...) processes one contiguous run of blocks.
__device__ __forceinline__ int xcd_block(int wg, int nwg) {
  const int xcd = wg % kNumXcd, loc = wg / kNumXcd;
  const int q = nwg / kNumXcd, r = nwg % kNumXcd;
  return (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

// Block extent shared by all kernels.  A partition entry is
// {first row, rowStartsT[row], rowStartsN[row], 0}.
struct Extent {
  bool full;   // rows end at the partition block's end (not cut at nrows)
  int r0, nr;  // rows [r0, r0+nr)
  int g0, nG;  // gather list nonzeros [g0, g0+nG)
  int s0, nS;  // scatter list nonzeros [s0, s0+nS)
};

__device__ __forceinline__ bool block_extent(const int4 *__restrict__ part, int b, int nrows,
                                             const int32_t *__restrict__ gS, int gUseN,
                                             const int32_t *__restrict__ sS, int sUseN, Extent &e) {
  const int4 p0 = part[b];
  const int4 p1 = part[b + 1];
  const int r1 = min(p1.x, nrows);
  e.r0 = p0.x;
  if (e.r0 >= r1) return false;
  e.nr = r1 - e.r0;
  const bool full = (r1 == p1.x);
  e.full = full;
  e.g0 = gUseN ? p0.z : p0.y;
  e.nG = (full ? (gUseN ? p1.z : p1.y) : gS[r1]) - e.g0;
  e.s0 = sUseN ? p0.z : p0.y;
  e.nS = (full ? (sUseN ? p1.z : p1.y) : sS[r1]) - e.s0;
  return true;
}

// ---------------------------------------------------------------------------
// K = 1 kernels.  A block's list entries are visited in the plan's
// column-sorted order when the list has a pack (consecutive lanes then touch
// consecutive addresses of v), and in CSR order otherwise.  Either way each
// gathered value lands in LDS at its CSR position, so the per-row reduction
// below always runs in CSR order.

// One list of a block as the kernel reads it.  pk/base: the plan's pack
// (null: plain colIds).  base[b].x < 0 marks a block the pack could not
// encode; that block reads colIds.
struct ListView {
  const int32_t *starts;
  const int32_t *ids;
  const uint32_t *pk;
  const int4 *base;
  int stride;  // pack words per block
};

// Index words of a block's list slots j (list positions tid + j*kBlock < n):
// pack words when the block is packed, else colIds.  All index loads of a
// block (these and the row starts) are issued before its first gather.
// Slots at or past n get word 0, which decodes to a valid column (the first
// segment base, or column 0), so the gathers can run unpredicated: a load
// under a branch can make the compiler wait for it at the join.  A
// block cut short at nrows (full == false) reads in CSR order: its pack sorts
// the entries of all the block's rows.  Returns whether the block is packed.
__device__ __forceinline__ bool slot_words(const ListView &L, int b, bool full, int g0, int n,
                                           uint32_t (&w)[kSlots], int4 &bb) {
  const int tid = threadIdx.x;
  bb = make_int4(-1, 0, 0, 0);
  if (L.pk && full) bb = L.base[b];
  const bool packed = bb.x >= 0;
  const uint32_t *src = packed ? L.pk + (size_t)b * L.stride : reinterpret_cast<const uint32_t *>(L.ids) + g0;
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int i = tid + j * kBlock;
    w[j] = i < n ? __builtin_nontemporal_load(src + i) : 0u;
  }
  return packed;
}

// Column of a slot's index word.
// (A two-level select on the segment bit pair: a flat compare chain over
// bb's components can be turned into an indexed access of bb, which puts bb
// in scratch memory.)
__device__ __forceinline__ int slot_col(uint32_t w, const int4 &bb, bool packed) {
  if (!packed) return (int)w;
  const uint32_t sg = w >> kPackSegShift;
  const int b0 = bb.x, b1 = bb.y, b2 = bb.z, b3 = bb.w;
  const int lo = (sg & 1u) ? b1 : b0;
  const int hi = (sg & 1u) ? b3 : b2;
  return ((sg & 2u) ? hi : lo) + (int)((w >> kPackOffShift) & kPackOffMask);
}

// LDS (CSR) position of slot j's entry.
__device__ __forceinline__ int slot_pos(uint32_t w, int j, bool packed) {
  return packed ? (int)(w & kPackPosMask) : (int)threadIdx.x + j * kBlock;
}

// Row starts [r0, r0+nr] of a list, loaded into registers.
__device__ __forceinline__ void row_words(const int32_t *__restrict__ starts, int r0, int nr, int (&v)[kRowSlots]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    v[j] = r <= nr ? __builtin_nontemporal_load(starts + r0 + r) : 0;
  }
}

__device__ __forceinline__ void stage_rows(const int (&v)[kRowSlots], int nr, int g0, int32_t *__restrict__ rs) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    if (r <= nr) rs[r] = v[j] - g0;
  }
}


// GatherScatter Sym with KT interleaved fields (KT a compile-time count).
// Lanes walk (entry, field) pairs, entries in the pack's column order, so a
// wave reads whole nodes' fields back to back; values land in LDS at
// pos*KT + k and each head lane reduces one field of its row in CSR order.
// The block holds at most kCap/KT entries (its partition is chosen so).
// KT == 0: the field count is the runtime K, divided by with kmagic =
// floor(2^32/K) + 1, which is exact for quotients of numbers below 2^16 by
// K in [2, 2^16); K = 1 divides by itself.  The batched kernel uses this form
// for every K, K = 1 included.
template <int KT>
struct FieldDiv {
  int K;
  uint32_t magic;
  __device__ __forceinline__ int k() const { return KT ? KT : K; }
  __device__ __forceinline__ int div(int f) const {
    return KT ? f / KT : (K == 1 ? f : (int)__umulhi((uint32_t)f, magic));
  }
};

template <typename T, int OP, int KT>
__device__ __forceinline__ void gs_sym_kv_block(const int4 *__restrict__ part, int b, int nrows, const ListView &G,
                                                int gUseN, T *__restrict__ q, T *__restrict__ val,
                                                uint8_t *__restrict__ head, FieldDiv<KT> fd = {KT, 0}) {
  Extent e;
  if (!block_extent(part, b, nrows, G.starts, gUseN, G.starts, gUseN, e)) return;
  const int n = e.nG;
  if (n == 0) return;
  const int KF = fd.k();
  const int nv = n * KF;
  const int tid = threadIdx.x;
  int4 bb = make_int4(-1, 0, 0, 0);
  if (G.pk && e.full) bb = G.base[b];
  const bool packed = bb.x >= 0;
  const uint32_t *src = packed ? G.pk + (size_t)b * G.stride : reinterpret_cast<const uint32_t *>(G.ids) + e.g0;
  uint32_t w[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int f = tid + j * kBlock;
    w[j] = f < nv ? __builtin_nontemporal_load(src + fd.div(f)) : 0u;
  }
  int rg[kRowSlots];
  if (!packed) row_words(G.starts, e.r0, e.nr, rg);
  T x[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int f = tid + j * kBlock;
    x[j] = q[(size_t)slot_col(w[j], bb, packed) * KF + (f - fd.div(f) * KF)];
  }
  if (packed) {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nv) {
        const int k = f - fd.div(f) * KF;
        const int ps = (int)(w[j] & kPackPosMask);
        val[ps * KF + k] = x[j];
        if (k == 0) head[ps] = (uint8_t)((w[j] >> kPackHeadBit) & 1u);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nv) {
        val[f] = x[j];
        if (f - fd.div(f) * KF == 0) head[fd.div(f)] = 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRowSlots; ++j) {
      const int r = tid + j * kBlock;
      const int a = rg[j] - e.g0;
      if (r < e.nr && a < n) head[a] = 1;
    }
  }
  __syncthreads();
  for (int t = tid; t < nv; t += kBlock) {
    const int i = fd.div(t), k = t - i * KF;
    if (head[i]) {
      typename Red<T, OP>::A red = Red<T, OP>::init();
      int m = i;
      do {
        Red<T, OP>::apply(red, val[m * KF + k]);
        ++m;
      } while (m < n && !head[m]);
      const T acc = Red<T, OP>::fin(red);
      for (int r = i; r < m; ++r) val[r * KF + k] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const int f = tid + j * kBlock;
    if (f < nv) {
      const int k = f - fd.div(f) * KF;
      const int ps = packed ? (int)(w[j] & kPackPosMask) : fd.div(f);
      q[(size_t)slot_col(w[j], bb, packed) * KF + k] = val[ps * KF + k];
    }
  }
}

// Field counts with a compile-time kernel; other K use the field-pass kernel.
__host__ __device__ constexpr bool packed_k(int K) { return K == 2 || K == 3 || K == 4 || K == 8 || K == 16; }

// Entry-major layout of KT interleaved fields: a work item is VW
// consecutive fields of one list entry (KT/VW items per entry), so one index
// word feeds VW value loads from contiguous addresses.  A block stages at most
// kVecBytes of values in LDS.
constexpr int kVecBytes = 16384;
template <typename T> __host__ __device__ constexpr int vec_cap() { return kVecBytes / (int)sizeof(T); }
__host__ __device__ constexpr int vec_width(int KT) { return KT <= 4 ? KT : 4; }

// GatherScatter, K = 1.  SYM: gather and scatter lists are the same array
// (Sym, or an Unsigned operator): the scatter reuses the gather's slots.
// CAPB bounds a block's entries of either list and its rows (two lists:
// launched with kCap / 2 on partitions 1.., whose blocks hold at most that
// many; half the LDS of kCap doubles the resident workgroups, measured 2-18%
// faster).
template <typename T, int OP, bool SYM, int CAPB = kCap>
__global__ __launch_bounds__(kBlock) void k_gs_k1(const int4 *__restrict__ part, int nrows, ListView G, int gUseN,
                                                  ListView S, int sUseN, T *__restrict__ q) {
  __shared__ T gval[CAPB];
  __shared__ T sval[SYM ? 1 : CAPB];
  __shared__ int32_t rsG[CAPB + 1];
  __shared__ int32_t rsS[SYM ? 1 : CAPB + 1];

  const int b = xcd_block(blockIdx.x, gridDim.x);
  Extent e;
  if (!block_extent(part, b, nrows, G.starts, gUseN, S.starts, sUseN, e)) return;
  const int tid = threadIdx.x;

  uint32_t w[kSlots], sw[SYM ? 1 : kSlots];
  int4 gb, sb;
  int rg[kRowSlots], rss[SYM ? 1 : kRowSlots];
  const bool gp = slot_words(G, b, e.full, e.g0, e.nG, w, gb);
  bool sp = false;
  if constexpr (!SYM) sp = slot_words(S, b, e.full, e.s0, e.nS, sw, sb);
  row_words(G.starts, e.r0, e.nr, rg);
  if constexpr (!SYM) row_words(S.starts, e.r0, e.nr, rss);
  T x[kSlots];
  if (e.nG > 0) {  // uniform: an operator may have rows but no columns
#pragma unroll
    for (int j = 0; j < kSlots; ++j) x[j] = q[slot_col(w[j], gb, gp)];
  }
  stage_rows(rg, e.nr, e.g0, rsG);
  if constexpr (!SYM) stage_rows(rss, e.nr, e.s0, rsS);
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    if (tid + j * kBlock < e.nG) gval[slot_pos(w[j], j, gp)] = x[j];
  }
  __syncthreads();

  for (int r = tid; r < e.nr; r += kBlock) {
    const int a = rsG[r], z = rsG[r + 1];
    typename Red<T, OP>::A red = Red<T, OP>::init();
    for (int i = a; i < z; ++i) Red<T, OP>::apply(red, gval[i]);
    const T acc = Red<T, OP>::fin(red);
    if constexpr (SYM) {
      for (int i = a; i < z; ++i) gval[i] = acc;
    } else {
      const int sa = rsS[r], sz = rsS[r + 1];
      for (int i = sa; i < sz; ++i) sval[i] = acc;
    }
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    if constexpr (SYM) {
      if (tid + j * kBlock < e.nG) q[slot_col(w[j], gb, gp)] = gval[slot_pos(w[j], j, gp)];
    } else {
      if (tid + j * kBlock < e.nS) q[slot_col(sw[j], sb, sp)] = sval[slot_pos(sw[j], j, sp)];
    }
  }
}

// Gather, K = 1: gq[row] = OP over the row's list.
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_gather_k1(const int4 *__restrict__ part, int nrows, ListView G, int gUseN,
                                                      const T *__restrict__ q, T *__restrict__ gq) {
  __shared__ T gval[kCap];
  __shared__ int32_t rsG[kWMax + 1];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  Extent e;
  if (!block_extent(part, b, nrows, G.starts, gUseN, G.starts, gUseN, e)) return;
  const int tid = threadIdx.x;
  uint32_t w[kSlots];
  int4 gb;
  int rg[kRowSlots];
  const bool gp = slot_words(G, b, e.full, e.g0, e.nG, w, gb);
  row_words(G.starts, e.r0, e.nr, rg);
  T x[kSlots];
  if (e.nG > 0) {  // uniform: an operator may have rows but no columns
#pragma unroll
    for (int j = 0; j < kSlots; ++j) x[j] = q[slot_col(w[j], gb, gp)];
  }
  stage_rows(rg, e.nr, e.g0, rsG);
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    if (tid + j * kBlock < e.nG) gval[slot_pos(w[j], j, gp)] = x[j];
  }
  __syncthreads();
  for (int r = tid; r < e.nr; r += kBlock) {
    const int a = rsG[r], z = rsG[r + 1];
    typename Red<T, OP>::A red = Red<T, OP>::init();
    for (int i = a; i < z; ++i) Red<T, OP>::apply(red, gval[i]);
    const T acc = Red<T, OP>::fin(red);
    gq[e.r0 + r] = acc;
  }
}

// Scatter, K = 1: q[col] = gq[row] for every member of the row's list.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter_k1(const int4 *__restrict__ part, int nrows, ListView S, int sUseN,
                                                       const T *__restrict__ gq, T *__restrict__ q) {
  __shared__ T sval[kCap];
  __shared__ int32_t rsS[kWMax + 1];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  Extent e;
  if (!block_extent(part, b, nrows, S.starts, sUseN, S.starts, sUseN, e)) return;
  const int tid = threadIdx.x;
  uint32_t w[kSlots];
  int4 sb;
  int rss[kRowSlots];
  const bool sp = slot_words(S, b, e.full, e.s0, e.nS, w, sb);
  row_words(S.starts, e.r0, e.nr, rss);
  T x[kRowSlots];
#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    x[j] = r < e.nr ? gq[e.r0 + r] : T{};
  }
  stage_rows(rss, e.nr, e.s0, rsS);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    if (r < e.nr) {
      const int a = rsS[r], z = rsS[r + 1];
      for (int i = a; i < z; ++i) sval[i] = x[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    if (tid + j * kBlock < e.nS) q[slot_col(w[j], sb, sp)] = sval[slot_pos(w[j], j, sp)];
  }
}

// ---------------------------------------------------------------------------
// GatherScatter, general K: fields [k0, k0+kw) per pass, kw*nnz <= kCap.
// Flat index f = i*kw + k runs over (nonzero, field) pairs, so consecutive
// lanes read consecutive fields of one node.
template <typename T, int OP, bool SYM>
__device__ __forceinline__ void gs_kv_block(const int4 *__restrict__ part, int b, int nrows,
                                            const int32_t *__restrict__ gS, const int32_t *__restrict__ gI, int gUseN,
                                            const int32_t *__restrict__ sS, const int32_t *__restrict__ sI, int sUseN,
                                            int K, int kc, T *__restrict__ q, T *__restrict__ gval, T *__restrict__ sval,
                                            int32_t *__restrict__ rsG, int32_t *__restrict__ rsS) {
  Extent e;
  if (!block_extent(part, b, nrows, gS, gUseN, sS, sUseN, e)) return;
  const int tid = threadIdx.x;

#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    if (r <= e.nr) {
      rsG[r] = gS[e.r0 + r] - e.g0;
      if constexpr (!SYM) rsS[r] = sS[e.r0 + r] - e.s0;
    }
  }

  for (int k0 = 0; k0 < K; k0 += kc) {
    const int kw = min(kc, K - k0);
    const int nvG = e.nG * kw;
    size_t off[kSlots];
    T x[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvG) {
        const int i = f / kw;
        off[j] = (size_t)gI[e.g0 + i] * K + k0 + (f - i * kw);
      }
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvG) x[j] = q[off[j]];
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvG) gval[f] = x[j];
    }
    __syncthreads();

    for (int t = tid; t < e.nr * kw; t += kBlock) {
      const int r = t / kw, k = t - r * kw;
      const int a = rsG[r], b = rsG[r + 1];
      typename Red<T, OP>::A red = Red<T, OP>::init();
      for (int i = a; i < b; ++i) Red<T, OP>::apply(red, gval[i * kw + k]);
      const T acc = Red<T, OP>::fin(red);
      if constexpr (SYM) {
        for (int i = a; i < b; ++i) gval[i * kw + k] = acc;
      } else {
        const int sa = rsS[r], sb = rsS[r + 1];
        for (int i = sa; i < sb; ++i) sval[i * kw + k] = acc;
      }
    }
    __syncthreads();

    if constexpr (SYM) {
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int f = tid + j * kBlock;
        if (f < nvG) q[off[j]] = gval[f];
      }
    } else {
      const int nvS = e.nS * kw;
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int f = tid + j * kBlock;
        if (f < nvS) {
          const int i = f / kw;
          off[j] = (size_t)sI[e.s0 + i] * K + k0 + (f - i * kw);
        }
      }
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int f = tid + j * kBlock;
        if (f < nvS) q[off[j]] = sval[f];
      }
    }
    __syncthreads();
  }
}

template <typename T, int OP, bool SYM>
__global__ __launch_bounds__(kBlock) void k_gs_kv(const int4 *__restrict__ part, int nrows,
                                                  const int32_t *__restrict__ gS, const int32_t *__restrict__ gI, int gUseN,
                                                  const int32_t *__restrict__ sS, const int32_t *__restrict__ sI, int sUseN,
                                                  int K, int kc, T *__restrict__ q) {
  __shared__ T gval[kCap];
  __shared__ T sval[SYM ? 1 : kCap];
  __shared__ int32_t rsG[kWMax + 1];
  __shared__ int32_t rsS[SYM ? 1 : kWMax + 1];
  gs_kv_block<T, OP, SYM>(part, xcd_block(blockIdx.x, gridDim.x), nrows, gS, gI, gUseN, sS, sI, sUseN, K, kc, q,
                          gval, sval, rsG, rsS);
}

// ---------------------------------------------------------------------------
// Batched GatherScatter Sym: one persistent launch over the blocks of many
// plans.  A work item is (descriptor, block); workgroup w takes items w,
// w + grid, ...  K = 1 items run the column-sorted path, K > 1 items the
// field-pass path.  LDS is shared between the two layouts.
struct BatchDev {
  const int4 *part;
  ListView G;
  int nrows;
  int K;
  int kc;
  uint32_t kmagic;  // floor(2^32/K) + 1 for K >= 2
  void *q;
};

// Pointers read from descriptor memory are generic to the compiler, which
// would then access them with flat instructions (counted on the LDS counter
// too, so every wait also drains LDS traffic).  They all point to device
// memory: casting through the global address space lets the compiler use
// global loads and stores.
template <typename P>
__device__ __forceinline__ P *as_global(P *p) {
  return (P *)(__attribute__((address_space(1))) P *)(uintptr_t)p;
}

__device__ __forceinline__ BatchDev load_desc(const BatchDev *descs, int i) {
  BatchDev d = descs[i];
  d.part = as_global(d.part);
  d.G.starts = as_global(d.G.starts);
  d.G.ids = as_global(d.G.ids);
  d.G.pk = as_global(d.G.pk);
  d.G.base = as_global(d.G.base);
  d.q = as_global(d.q);
  return d;
}

// The batched kernel has one path, K fields in one pass for any K (K = 1
// included; the division by K uses the descriptor's magic multiplier), so
// its register budget is that of one path.  Descriptors whose fields take
// several passes run in k_batch_gs_rest, a second node of the same graph.
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_batch_gs(const BatchDev *__restrict__ descs, const int2 *__restrict__ items,
                                                     int nitems) {
  __shared__ T vals[kCap];
  __shared__ uint8_t head[kCap];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int2 w = items[it];
    const BatchDev d = load_desc(descs, w.x);
    gs_sym_kv_block<T, OP, 0>(d.part, w.y, d.nrows, d.G, 0, static_cast<T *>(d.q), vals, head,
                              FieldDiv<0>{d.K, d.kmagic});
    __syncthreads();
  }
}

template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_batch_gs_rest(const BatchDev *__restrict__ descs,
                                                          const int2 *__restrict__ items, int nitems) {
  __shared__ T gval[kCap];
  __shared__ int32_t rsG[kWMax + 1];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const int2 w = items[it];
    const BatchDev d = load_desc(descs, w.x);
    gs_kv_block<T, OP, true>(d.part, w.y, d.nrows, d.G.starts, d.G.ids, 0, d.G.starts, d.G.ids, 0, d.K, d.kc,
                             static_cast<T *>(d.q), gval, nullptr, rsG, nullptr);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// GatherScatter Sym (one list: gather list == scatter list) for KT = 1 and
// KT in {2, 3, 4, 8, 16} interleaved fields, entry-major, one block per
// workgroup.  A block goes through
//   meta    its partition entries and pack base,
//   index   its index words (pack words, or colIds in CSR order),
//   gather  VW contiguous values per item, all slots' loads in flight,
//   finish  LDS staging in CSR order, per-row reduction, scatter stores.
// (Measured on MI355X: a persistent workgroup that walks a run of blocks and
// loads block b+1's metadata and index words during block b's LDS work is
// slower than one block per workgroup -- at 8 resident workgroups per CU
// the hardware already overlaps the blocks' latencies, and its dispatcher
// balances uneven blocks better than static runs.)
//
// Shape per (T, KT): VW fields per work item, SL slots per lane, sized so
// that the values in flight stay within ~24 VGPRs; a block holds at most CAP
// entries (slots, LDS and the pack's 2048 positions all bound it).  8-byte
// fields go two per item (one aligned 16-byte access; measured 3-7% faster
// than four per item at K = 4..16) and, at K = 3, one per item (3-8% faster
// than a 24-byte item); 4- and 2-byte fields go up to four per item (one
// field per item measured ~1.8x slower at K = 3).
template <typename T, int KT>
struct VecShape {
  static constexpr int VW = sizeof(T) == 8 ? (KT == 3 ? 1 : KT >= 2 ? 2 : 1) : vec_width(KT);
  static constexpr int C = KT / VW;
  static constexpr int WORDS = (int)((sizeof(T) + 3) / 4);
  using L = typename Lds<T>::L;
  // slots per lane: by item size, and no more than the LDS capacity fills
  static constexpr int SL0 = VW * WORDS <= 3 ? 8 : VW * WORDS <= 6 ? 4 : 2;
  static constexpr int SLL = (vec_cap<L>() / KT * C + kBlock - 1) / kBlock;
  static constexpr int SL = SL0 < SLL ? SL0 : SLL;
  static constexpr int CAP0 = SL * kBlock / C < vec_cap<L>() / KT ? SL * kBlock / C : vec_cap<L>() / KT;
  static constexpr int CAP = CAP0 < kCap ? CAP0 : kCap;
  static_assert(KT % VW == 0, "field count must be a multiple of the item width");
};

// A block's list and buffer: fixed for a plan launch, per item in a batch.
// These kernels run whole partitions (rows [0, NrowsT) of the T list, or of
// an aliased N list), so no block is cut short at nrows.
struct VecSrc {
  const int4 *part;
  const int32_t *ids;
  const uint32_t *pk;
  const int4 *base;
  void *q;
  int gUseN;
  int stride;  // pack words per block
};

// A block's metadata: its partition entries and pack base.
struct VecMeta {
  int4 p0, p1, bb;
};

template <int SL>
struct VecStage {
  VecSrc src;
  int n;       // entries
  int g0;      // first entry
  int4 bb;     // pack segment bases (x < 0: CSR order)
  uint32_t w[SL];
};

__device__ __forceinline__ void vec_meta(const VecSrc &src, int b, VecMeta &m) {
  m.p0 = src.part[b];
  m.p1 = src.part[b + 1];
  m.bb = src.pk ? src.base[b] : make_int4(-1, 0, 0, 0);
}

// The same for a block whose pack words sit at another slot than its block
// id (a delta pack's fallback block in a pack of fallback blocks only).
__device__ __forceinline__ void vec_meta_at(const VecSrc &src, int b, int slot, VecMeta &m) {
  m.p0 = src.part[b];
  m.p1 = src.part[b + 1];
  m.bb = src.pk ? src.base[slot] : make_int4(-1, 0, 0, 0);
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Index words of block b.  With a pack the words sit at b * stride
// (zero-padded past the block's entries), so their loads go out together
// with the metadata loads instead of after them; otherwise the block's
// colIds are read once its partition entry m has arrived.
template <typename T, int KT, int SL>
__device__ __forceinline__ void vec_index(VecStage<SL> &s, const VecSrc &src, int b, const VecMeta &m) {
  constexpr int C = VecShape<T, KT>::C;
  s.src = src;
  if (src.pk) {
    const uint32_t *w0 = src.pk + (size_t)b * src.stride;
#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int e = ((int)threadIdx.x + j * kBlock) / C;
      s.w[j] = e < src.stride ? __builtin_nontemporal_load(w0 + e) : 0u;
    }
  }
  s.g0 = uniform(src.gUseN ? m.p0.z : m.p0.y);
  s.n = uniform(src.gUseN ? m.p1.z : m.p1.y) - s.g0;
  s.bb = make_int4(uniform(m.bb.x), uniform(m.bb.y), uniform(m.bb.z), uniform(m.bb.w));
  if (!src.pk) {
    const uint32_t *w0 = reinterpret_cast<const uint32_t *>(src.ids) + s.g0;
    const int ni = s.n * C;
#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int it = (int)threadIdx.x + j * kBlock;
      s.w[j] = it < ni ? __builtin_nontemporal_load(w0 + it / C) : 0u;
    }
  }
}

// Gathers run unpredicated: a slot past the block's entries holds word 0,
// which decodes to a valid column (see slot_words); the launcher runs these
// kernels only on operators with at least one column.
template <typename T, int KT, int SL>
__device__ __forceinline__ void vec_gather(const VecStage<SL> &s, T (&x)[SL][VecShape<T, KT>::VW]) {
  constexpr int VW = VecShape<T, KT>::VW, C = VecShape<T, KT>::C;
  const T *q = static_cast<const T *>(s.src.q);
  const bool packed = s.bb.x >= 0;
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = (int)threadIdx.x + j * kBlock;
    const T *a = q + (size_t)slot_col(s.w[j], s.bb, packed) * KT + (it % C) * VW;
#pragma unroll
    for (int v = 0; v < VW; ++v) x[j][v] = a[v];
  }
}

// Row heads as a bit mask, one bit per CSR position of the block, with every
// position from n on marked (so the row of the last head ends at n).  Built
// from the staged head bytes: word wd covers positions [32 wd, 32 wd + 32).
__device__ __forceinline__ void head_mask(const uint8_t *__restrict__ head, int n, uint32_t *__restrict__ hm) {
  const uint32_t *h4 = reinterpret_cast<const uint32_t *>(head);
  for (int wd = threadIdx.x; wd <= (n >> 5); wd += kBlock) {
    uint32_t bits = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t hb = h4[wd * 8 + q];  // four head bytes, each 0 or 1
      bits |= ((hb & 1u) | ((hb >> 7) & 2u) | ((hb >> 14) & 4u) | ((hb >> 21) & 8u)) << (4 * q);
    }
    const int lo = n - wd * 32;  // positions >= n in this word
    if (lo < 32) bits = (lo <= 0 ? 0u : bits) | (~0u << (lo > 0 ? lo : 0));
    hm[wd] = bits;
  }
}

// Members in the row whose head is at position i (the next marked position).
__device__ __forceinline__ int row_len(const uint32_t *__restrict__ hm, int i) {
  int wd = i >> 5;
  uint32_t rest = (hm[wd] >> (i & 31)) >> 1;  // positions i+1 .. 32 wd + 31
  int len = 1;
  if (!rest) {
    len = 32 - (i & 31);
    while ((rest = hm[++wd]) == 0u) len += 32;
    return len + __builtin_ctz(rest);
  }
  return len + __builtin_ctz(rest);
}

// Exclusive prefix counts of the head-mask words (one wave): pc[w] = heads
// in positions [0, 32 w).  An entry's row within its block is then the
// number of heads at or before its CSR position, minus one (head_row).
__device__ __forceinline__ void head_prefix(const uint32_t *__restrict__ hm, int n, int32_t *__restrict__ pc) {
  constexpr int NW = kCap / 32;
  static_assert(NW == 64, "one wave scans the head-mask words");
  const int tid = threadIdx.x;
  if (tid < NW) {
    const int c = tid < ((n + 31) >> 5) ? __popc(hm[tid]) : 0;
    int incl = c;
#pragma unroll
    for (int d = 1; d < NW; d <<= 1) {
      const int y = __shfl_up(incl, d, NW);
      if (tid >= d) incl += y;
    }
    pc[tid] = incl - c;
  }
}
__device__ __forceinline__ int head_row(const uint32_t *__restrict__ hm, const int32_t *__restrict__ pc, int ps) {
  const int wd = ps >> 5;
  const uint32_t upto = (uint32_t)((2ull << (ps & 31)) - 1u);  // positions 32 wd .. ps
  return pc[wd] + __popc(hm[wd] & upto) - 1;
}

// Fold of one (row, field chunk) of a block staged in LDS: the row's len
// members from CSR position i, VW fields from chunk c, in CSR order from the
// op's identity, members read four at a time; the result is written back
// over the row.
// LDS slot of CSR position i in a padded stage: one spare slot per 16, so
// that lanes folding rows of 8 (or 16) members from positions 8 apart read
// 17 dwords apart and 16 lanes cover the 32 banks once, instead of two.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
constexpr int lds_padded(int n) { return n + n / 16; }

// ... with the stage padded (PAD, KT = 1) or not.
template <typename T, int OP, int KT, bool PAD = false>
__device__ __forceinline__ void fold_row(typename Lds<T>::L *__restrict__ val, int i, int c, int len) {
  using L = typename Lds<T>::L;
  using R = LRed<T, OP>;
  constexpr int VW = VecShape<T, KT>::VW;
  static_assert(!PAD || KT == 1, "padded stages hold one field per entry");
  auto at = [&](int m, int v) -> L & { return val[(PAD ? lds_pad(i + m) : i + m) * KT + c * VW + v]; };
  typename R::A red[VW];
  int mb = 0;
  if constexpr (std::is_same<typename R::A, Comp>::value) {
    // compensated sums take the row's first member (len >= 1) in one step
#pragma unroll
    for (int v = 0; v < VW; ++v) red[v] = comp_first(lds_float(at(0, v)));
    mb = 1;
  } else {
#pragma unroll
    for (int v = 0; v < VW; ++v) red[v] = R::init();
  }
  for (; mb + 4 <= len; mb += 4) {
    L y[4][VW];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < VW; ++v) y[u][v] = at(mb + u, v);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < VW; ++v) R::apply(red[v], y[u][v]);
  }
  for (; mb < len; ++mb) {
#pragma unroll
    for (int v = 0; v < VW; ++v) R::apply(red[v], at(mb, v));
  }
  L res[VW];
#pragma unroll
  for (int v = 0; v < VW; ++v) res[v] = R::fin(red[v]);
  for (int r = 0; r < len; ++r)
#pragma unroll
    for (int v = 0; v < VW; ++v) at(r, v) = res[v];
}

// LDS of a row-compacted reduction: the head-mask prefix counts and the row
// count (pc), and (rp) the head position of every row plus the end, n, 16 bits
// each.  rp overlays the block's head bytes, which are dead once the head
// mask is built; the head region is sized for both (head_bytes).
struct RowIndex {
  int32_t *pc;  // kCap / 32 + 1 words, the last the row count
  uint16_t *rp;
};
// Whether a shape's blocks reduce over compacted rows: those below kCap
// entries, whose row list fits the head region with at most 700 more bytes
// (so the kernels keep their resident workgroups per CU); measured on
// bfloat16 K = 3, instruction bound, whose vector instructions fell by a
// quarter.
template <typename T, int KT> __host__ __device__ constexpr bool rows_reduce() { return VecShape<T, KT>::CAP < kCap; }
template <typename T, int KT> __host__ __device__ constexpr int head_bytes() {
  return rows_reduce<T, KT>() && 2 * (VecShape<T, KT>::CAP + 2) > kCap ? 2 * (VecShape<T, KT>::CAP + 2) : kCap;
}

// Row reduction of a block whose n entries' values (KT fields each) are
// staged in LDS at their CSR positions with their row-head bytes (the
// staging barrier already passed), each (row, field chunk) folded by one lane
// (measured faster than every item summing its own row: rows of several
// members would be read once per member).  Without a row index, lanes step
// over the entries and fold at the heads; with one (ri), the heads' positions
// are first compacted in CSR order, so lanes step over rows and none idles on
// an entry that is not a head.  Ends with a barrier.
template <typename T, int OP, int KT>
__device__ __forceinline__ void vec_reduce(int n, typename Lds<T>::L *__restrict__ val, uint8_t *__restrict__ head,
                                           uint32_t *__restrict__ hm, const RowIndex *ri = nullptr) {
  constexpr int C = VecShape<T, KT>::C;
  const int tid = threadIdx.x;
  head_mask(head, n, hm);
  __syncthreads();
  if (!ri) {
    for (int t = tid; t < n * C; t += kBlock) {
      const int i = t / C, c = t - i * C;
      if (!((hm[i >> 5] >> (i & 31)) & 1u)) continue;
      fold_row<T, OP, KT>(val, i, c, row_len(hm, i));
    }
    __syncthreads();
    return;
  }
  head_prefix(hm, n, ri->pc);
  __syncthreads();
  for (int p = tid; p < n; p += kBlock) {
    const int wd = p >> 5;
    const uint32_t bits = hm[wd];
    const int incl = ri->pc[wd] + __popc(bits & (uint32_t)((2ull << (p & 31)) - 1u));  // heads in [0, p]
    if ((bits >> (p & 31)) & 1u) ri->rp[incl - 1] = (uint16_t)p;
    if (p == n - 1) {  // position 0 is a head, so incl here is the block's row count
      ri->rp[incl] = (uint16_t)n;
      ri->pc[kCap / 32] = incl;
    }
  }
  __syncthreads();
  const int nr = n > 0 ? ri->pc[kCap / 32] : 0;
  for (int u = tid; u < nr * C; u += kBlock) {
    const int r = u / C, c = u - r * C;
    const int i = ri->rp[r];
    fold_row<T, OP, KT>(val, i, c, (int)ri->rp[r + 1] - i);
  }
  __syncthreads();
}

// SCATTER_N false: every entry gathered and scattered (one list).  True:
// every entry gathered, the results scattered to the entries whose bit in
// nbits is set (bit j: slot j's entry is an N member) only.
template <typename T, int OP, int KT, int SL, bool SCATTER_N = false>
__device__ __forceinline__ void vec_finish(const VecStage<SL> &s, const VecMeta &m,
                                           const int32_t *__restrict__ starts,
                                           const T (&x)[SL][VecShape<T, KT>::VW],
                                           typename Lds<T>::L *__restrict__ val, uint8_t *__restrict__ head,
                                           uint32_t *__restrict__ hm, uint32_t nbits = 0u,
                                           const RowIndex *ri = nullptr) {
  constexpr int VW = VecShape<T, KT>::VW, C = VecShape<T, KT>::C;
  const int n = s.n;
  if (n == 0) return;
  const int ni = n * C;
  const int tid = threadIdx.x;
  const bool packed = s.bb.x >= 0;
  if (packed) {
#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int it = tid + j * kBlock;
      if (it < ni) {
        const int c = it % C;
        const int ps = (int)(s.w[j] & kPackPosMask);
#pragma unroll
        for (int v = 0; v < VW; ++v) val[ps * KT + c * VW + v] = Lds<T>::in(x[j][v]);
        if (c == 0) head[ps] = (uint8_t)((s.w[j] >> kPackHeadBit) & 1u);
      }
    }
  } else {
    // CSR order (a block the pack could not encode): heads from the row starts
    const int r0 = uniform(m.p0.x), nr = uniform(m.p1.x) - r0;
    int rg[kRowSlots];
    row_words(starts, r0, nr, rg);
#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int it = tid + j * kBlock;
      if (it < ni) {
        const int c = it % C, i = it / C;
#pragma unroll
        for (int v = 0; v < VW; ++v) val[i * KT + c * VW + v] = Lds<T>::in(x[j][v]);
        if (c == 0) head[i] = 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRowSlots; ++j) {
      const int r = tid + j * kBlock;
      const int a = rg[j] - s.g0;
      if (r < nr && a < n) head[a] = 1;  // an empty row marks the next row's head again
    }
  }
  __syncthreads();
  vec_reduce<T, OP, KT>(n, val, head, hm, ri);
  T *q = static_cast<T *>(s.src.q);
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = tid + j * kBlock;
    if (it < ni && (!SCATTER_N || ((nbits >> j) & 1u))) {
      const int c = it % C;
      const int ps = packed ? (int)(s.w[j] & kPackPosMask) : it / C;
      T *d = q + (size_t)slot_col(s.w[j], s.bb, packed) * KT + c * VW;
#pragma unroll
      for (int v = 0; v < VW; ++v) d[v] = Lds<T>::out(val[ps * KT + c * VW + v]);
    }
  }
}

template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gs_sym_vec(const int4 *__restrict__ part,
                                                       const int32_t *__restrict__ starts,
                                                       const int32_t *__restrict__ ids,
                                                       const uint32_t *__restrict__ pk,
                                                       const int4 *__restrict__ base, int stride, int gUseN,
                                                       T *__restrict__ q, const int *__restrict__ order) {
  using V = VecShape<T, KT>;
  using L = typename Lds<T>::L;
  __shared__ L val[vec_cap<L>()];
  __shared__ uint32_t head4[(head_bytes<T, KT>() + 3) / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  uint8_t *head = reinterpret_cast<uint8_t *>(head4);
  const VecSrc src = {part, ids, pk, base, q, gUseN, stride};
  VecStage<V::SL> s;
  VecMeta m;
  T x[V::SL][V::VW];
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  vec_meta(src, b, m);
  vec_index<T, KT>(s, src, b, m);
  vec_gather<T, KT>(s, x);
  if constexpr (rows_reduce<T, KT>()) {
    __shared__ int32_t pc[kCap / 32 + 1];
    const RowIndex ri = {pc, reinterpret_cast<uint16_t *>(head4)};
    vec_finish<T, OP, KT>(s, m, starts, x, val, head, hm, 0u, &ri);
  } else {
    vec_finish<T, OP, KT>(s, m, starts, x, val, head, hm);
  }
}

// ---------------------------------------------------------------------------
// Delta pack: the K = 1 one-list GatherScatter over partition 0 with half the
// index bytes of the 32-bit pack words.
//
// A block's entries in column order (the 32-bit pack's order) form four wave
// ranges of kDWave = 512 entries: wave w takes entries [512 w, 512 w + 512),
// slot j of lane l being entry 512 w + 64 j + l (each wave instruction then
// touches 64 consecutive entries, as in k_gs_sym_vec).  An entry is one
// 16-bit word: CSR position (11 bits) | row-head bit | delta code (4 bits).
// Its column is the previous entry's column (in its wave range) plus a delta:
// code 0 adds 0, codes 1..14 add the plan's table entry (the 14 commonest
// deltas below 256, one byte each), code 15 adds the wave range's next
// exception word (32 bits); the first entry of a range is always an
// exception, holding its column.  The words are stored lane-major within a
// range (lane l's 8 words contiguous), so one 16-byte load per lane fetches a
// range's 512 words, and its E exception words come with them (first E
// lanes, one word each).  A wave recovers its columns with one inclusive scan
// per slot (DPP).  Blocks that do not fit -- a range with more than E
// exceptions, or a block the 32-bit pack left in CSR order -- carry
// kDFallback as every range's first exception word and run from the 32-bit
// pack.  48^3 p=7: 2.13 index bytes per entry instead of 4.
constexpr int kDWave = 512;
constexpr int kDRanges = kCap / kDWave;
constexpr uint32_t kDEscape = 15u;
constexpr uint32_t kDFallback = 0xffffffffu;
constexpr int kDMaxE = 64;
static_assert(kDRanges * 64 == kBlock && kDWave == 8 * 64, "a wave range is 8 slots of one wave");
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Inclusive prefix sum over the 64 lanes of a wave: Hillis-Steele steps
// within each row of 16 lanes (row_shr 1, 2, 4, 8), then row 0's total into
// rows 1 and 3 (row_bcast 15) and rows 0-1's total into rows 2 and 3
// (row_bcast 31).  A lane whose source lies outside its row keeps `old` = 0.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t x) {
  int v = (int)x;
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return (uint32_t)v;
}

// Lanes below this one whose bit is set in the wave mask m.
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Delta of a code from the table (byte c of tb: code c), by byte permutes
// (v_perm_b32 picks byte s of {src0:src1}, src1 the low dword; selector 12
// gives 0).  (A select chain over tb's components was turned by the compiler
// into a per-lane indexed load of the kernel arguments.)
__device__ __forceinline__ uint32_t delta_of(uint32_t code, const uint4 &tb) {
  const uint32_t sel = (code & 7u) | 0x0c0c0c00u;
  const uint32_t lo = __builtin_amdgcn_perm(tb.y, tb.x, sel);
  const uint32_t hi = __builtin_amdgcn_perm(tb.w, tb.z, sel);
  return (code & 8u) ? hi : lo;
}

// The columns of a wave's 8 slots from its 16-bit words (two per dword of
// w2) and its exception word ex (lane l holds the range's exception l): per
// slot the deltas (an exception by its rank among the range's escapes,
// through a lane permute), an inclusive scan, the running carry.  Columns
// are clamped to [0, ncols), so any word content stays in bounds.
__device__ __forceinline__ void delta_columns(const uint32_t (&w2)[4], uint32_t ex, const uint4 &tb, int ncols,
                                              uint32_t (&col)[8]) {
  uint32_t carry = 0u;
  int run = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t code = (w2[j >> 1] >> ((j & 1) * 16 + 12)) & 15u;
    uint32_t d = delta_of(code, tb);
    const bool esc = code == kDEscape;
    const uint64_t mk = __builtin_amdgcn_ballot_w64(esc);
    if (mk) {
      const int rank = (run + lanes_below(mk)) & 63;
      const uint32_t e = (uint32_t)__builtin_amdgcn_ds_bpermute(rank * 4, (int)ex);
      d = esc ? e : d;
      run += __builtin_popcountll(mk);
    }
    const uint32_t sc = wave_scan_incl(d);
    const uint32_t c = carry + sc;
    col[j] = c < (uint32_t)ncols ? c : (uint32_t)(ncols - 1);
    carry += (uint32_t)__builtin_amdgcn_readlane((int)sc, 63);
  }
}

// Row reduction of a delta-pack block staged padded (lds_pad) at its CSR
// positions: a block of at most kCap / 2 - 1 rows lists its row heads in
// the dead head bytes and folds one row per lane (at low degree, rows of 4-8
// members: every lane busy instead of one in eight); a block of more rows
// (mostly single-member rows) folds at the heads of the entries.  Ends with
// a barrier.
template <typename T, int OP>
__device__ __forceinline__ void delta_reduce(int n, typename Lds<T>::L *__restrict__ val, uint8_t *__restrict__ head,
                                             uint32_t *__restrict__ hm, int32_t *__restrict__ pc) {
  const int tid = threadIdx.x;
  head_mask(head, n, hm);
  __syncthreads();
  head_prefix(hm, n, pc);
  __syncthreads();
  const int last = n - 1, lw = last >> 5;
  const int rows = pc[lw] + __popc(hm[lw] & (uint32_t)((2ull << (last & 31)) - 1u));
  if (rows <= kCap / 2 - 1) {
    uint16_t *rp = reinterpret_cast<uint16_t *>(head);  // head bytes are dead once hm is built
    for (int q = tid; q < n; q += kBlock) {
      const uint32_t bits = hm[q >> 5];
      if ((bits >> (q & 31)) & 1u) rp[pc[q >> 5] + __popc(bits & (uint32_t)((1ull << (q & 31)) - 1u))] = (uint16_t)q;
    }
    if (tid == 0) rp[rows] = (uint16_t)n;
    __syncthreads();
    for (int r = tid; r < rows; r += kBlock) {
      const int i = rp[r];
      fold_row<T, OP, 1, true>(val, i, 0, (int)rp[r + 1] - i);
    }
  } else {
    for (int t = tid; t < n; t += kBlock)
      if ((hm[t >> 5] >> (t & 31)) & 1u) fold_row<T, OP, 1, true>(val, t, 0, row_len(hm, t));
  }
  __syncthreads();
}

// Types whose row reduction may run as seg8_reduce (an accumulator of the
// staged type: the reference's four).
template <typename T>
constexpr bool kSeg8 = std::is_same<T, double>::value || std::is_same<T, float>::value ||
                       std::is_same<T, int>::value || std::is_same<T, long long>::value;

// Bits 0..7 of the head bytes held in x (bit k: byte k is 1).  Each 0/1
// byte b_k of a word lands on bit 24 + k of the product with 0x01020408,
// the other partial products on distinct lower bits.
__device__ __forceinline__ uint32_t head_bits8(uint2 x) {
  const uint32_t lo = ((x.x & 0x01010101u) * 0x01020408u) >> 24;
  const uint32_t hi = ((x.y & 0x01010101u) * 0x01020408u) >> 24;
  return lo | (hi << 4);
}

// Stage layouts of a delta-coded block, all in kCap * (sizeof(L) + 1)
// bytes.  kPlain: the values at their CSR positions, then one head byte per
// position (rows folded at their heads, vec_reduce); kRows: the values
// padded (lds_pad), rows folded one per lane (delta_reduce); kSeg: groups of
// 8 positions, each group's 8 values followed by their 8 head bytes
// (seg8_reduce, for rows of at most 8 members: a thread reads a group of
// 8 * sizeof(L) + 8 bytes, which puts 16 consecutive threads' groups on
// distinct banks).
enum { kPlain = 0, kRows = 1, kSeg = 2 };

template <typename L>
__host__ __device__ constexpr int seg_group() { return 8 * (int)sizeof(L) + 8; }
template <typename L>
__device__ __forceinline__ L &seg_val(uint8_t *s, int ps) {
  return *reinterpret_cast<L *>(s + (ps >> 3) * seg_group<L>() + (ps & 7) * (int)sizeof(L));
}
template <typename L>
__device__ __forceinline__ uint8_t &seg_head(uint8_t *s, int ps) {
  return s[(ps >> 3) * seg_group<L>() + 8 * (int)sizeof(L) + (ps & 7)];
}

// Row results of a block staged in the kSeg layout whose rows have at most 8
// members.  Thread t owns positions [8t, 8t + 8) (group t) and folds the
// rows whose heads lie there, in CSR order from the operator's identity, the
// last of them on into group t + 1 up to that group's first head (a row of
// at most 8 members ends there).  res[k]: the result of the row holding
// position 8t + k, for k from the group's first head on (res[7]: of the row
// that continues into group t + 1); bits: head bits of positions 8t .. 8t +
// 15 at bits 8 .. 23 (bit 8 + k: position 8t + k), positions from n on
// counted as heads.  With OPEN the thread also folds the row open at 8t
// again, from group t - 1's last head (bits 0 .. 7: group t - 1's head bits,
// all set for t = 0), so that res covers all 8 positions: the GatherScatter
// writes each position from its own thread that way (one fold per row
// instead measured 1% slower at 96^3 p=3 and 192^3 p=1, equal at 48^3 p=7);
// the Gather needs only the heads' results (one fold per row: 48^3 p=7
// 159.3 -> 156.9 us).
template <typename T, int OP, bool OPEN>
__device__ __forceinline__ void seg8_fold(int n, const uint8_t *__restrict__ s, T (&res)[8], uint32_t &bits) {
  using R = LRed<T, OP>;
  using A = typename R::A;
  constexpr int G = seg_group<T>(), HV = 8 * (int)sizeof(T);
  const int t = (int)threadIdx.x, P = 8 * t;
  const int gp = t > 0 ? t - 1 : t, gn = t + 1 < kCap / 8 ? t + 1 : t;
  const uint8_t *sp = s + gp * G, *so = s + t * G, *sn = s + gn * G;
  if constexpr (OPEN) {
    bits = head_bits8(*reinterpret_cast<const uint2 *>(sp + HV)) |
           (head_bits8(*reinterpret_cast<const uint2 *>(so + HV)) << 8) |
           (head_bits8(*reinterpret_cast<const uint2 *>(sn + HV)) << 16);
    if (t == 0) bits |= 0xffu;  // no previous group
  } else {
    bits = (head_bits8(*reinterpret_cast<const uint2 *>(so + HV)) << 8) |
           (head_bits8(*reinterpret_cast<const uint2 *>(sn + HV)) << 16);
  }
  const int live = n - (P - 8);  // bit k is position P - 8 + k: from bit `live` on, past the block's entries
  if (live < 24) bits |= ~0u << (live > 0 ? live : 0);
  // phases of 8 values each: held apart (sched_barrier) so that no more
  // than 16 are in registers at once
  T x[8];
  A c = R::init();
  if constexpr (OPEN) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = reinterpret_cast<const T *>(sp)[k];
    // the row open at P: from group t - 1's last head
    const int lh = 31 - __builtin_clz(bits & 0xffu);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      A y = c;
      R::apply(y, x[k]);
      c = k >= lh ? y : c;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = reinterpret_cast<const T *>(so)[k];
  A part[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((bits >> (8 + k)) & 1u) c = R::init();
    R::apply(c, x[k]);
    part[k] = c;
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = reinterpret_cast<const T *>(sn)[k];
  // the last own row on to group t + 1's first head
  const int fh = __builtin_ctz(bits >> 16);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    A y = c;
    R::apply(y, x[k]);
    c = k < fh ? y : c;
  }
  res[7] = R::fin(c);
#pragma unroll
  for (int k = 7; k > 0; --k) {
    if ((bits >> (8 + k)) & 1u) c = part[k - 1];
    res[k - 1] = R::fin(c);
  }
}

// Writes a once-per-row group fold's results over the stage (after a
// barrier, since neighbours read the group): own positions from the group's
// first head on, and group t + 1's positions before its first head (the
// last own row's).
template <typename L>
__device__ __forceinline__ void seg8_put(uint8_t *__restrict__ own, uint8_t *__restrict__ next, uint32_t bits,
                                         const L (&res)[8]) {
  const int fo = __builtin_ctz((bits >> 8) | 0x100u), fh = __builtin_ctz(bits >> 16);
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k >= fo) reinterpret_cast<L *>(own)[k] = res[k];
  if ((int)threadIdx.x + 1 < kCap / 8) {
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (k < fh) reinterpret_cast<L *>(next)[k] = res[7];
  }
}

// seg8_fold (OPEN), each own position then taking its row's result (written
// after a barrier, since neighbours read the group).  Ends with a barrier.
template <typename T, int OP>
__device__ __forceinline__ void seg8_reduce(int n, uint8_t *__restrict__ s) {
  T res[8];
  uint32_t bits;
  seg8_fold<T, OP, true>(n, s, res, bits);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) reinterpret_cast<T *>(s + threadIdx.x * seg_group<T>())[k] = res[k];
  __syncthreads();
}

// Gather from a kSeg stage: seg8_fold, then the thread holding a row's head
// puts the row's result at the row's index in the block (the heads before
// it: per thread, summed over the wave by a scan and over the waves before
// through wsum) in the stage, and the block stores its rows' results in
// order, whole lines (stored from the heads directly, rows of neighbouring
// lanes lie ~5 apart at p=7 and each store instruction writes parts of ~20
// lines: 48^3 p=7 164 -> 205 us).
template <typename T, int OP>
__device__ __forceinline__ void seg8_gather(int n, int r0, uint8_t *__restrict__ s, uint32_t *__restrict__ wsum,
                                            T *__restrict__ gq) {
  T res[8];
  uint32_t bits;
  seg8_fold<T, OP, false>(n, s, res, bits);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int own = n - 8 * (int)threadIdx.x;  // own positions below n
  const uint32_t hh = ((bits >> 8) & 0xffu) & (own >= 8 ? 0xffu : own > 0 ? (1u << own) - 1u : 0u);
  const uint32_t hc = (uint32_t)__popc(hh);
  const uint32_t inc = wave_scan_incl(hc);
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();  // (also: every thread's reads of the stage are done)
  uint32_t row = inc - hc;
  for (int w = 0; w < wv; ++w) row += wsum[w];
  const int rows = (int)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
  T *rb = reinterpret_cast<T *>(s);  // rows <= n <= kCap results
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if ((hh >> k) & 1u) rb[row + __popc(hh & ((1u << k) - 1u))] = res[k];
  __syncthreads();
  for (int i = threadIdx.x; i < rows; i += kBlock) gq[(size_t)r0 + i] = rb[i];
}

// The rest of a delta-coded block once a wave range's columns are known:
// slot j of the lane holds entry 512 w + 64 j + lane, its CSR position and
// head bit in half j & 1 of w2[j >> 1] (the 16-bit word's low 12 bits) and
// its column in col[j].  Loads the values, stages them (layout MODE; val is
// the stage, head its head bytes in kPlain and kRows), reduces each row in
// CSR order and stores the results to every member.
template <typename T, int OP, int MODE>
__device__ __forceinline__ void delta_stage_reduce(const uint32_t (&w2)[4], const uint32_t (&col)[8], int n,
                                                   T *__restrict__ q, typename Lds<T>::L *__restrict__ val,
                                                   uint8_t *__restrict__ head, uint32_t *__restrict__ hm,
                                                   int32_t *__restrict__ pc) {
  using L = typename Lds<T>::L;
  static_assert(MODE != kSeg || kSeg8<T>, "kSeg stages the reference's types");
  uint8_t *sg = reinterpret_cast<uint8_t *>(val);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  T x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = q[col[j]];
  const int i0 = wv * kDWave + lane;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (i0 + j * 64 < n) {
      const uint32_t wj = w2[j >> 1] >> ((j & 1) * 16);
      const int ps = (int)(wj & kPackPosMask);
      const uint8_t hb = (uint8_t)((wj >> kPackHeadBit) & 1u);
      if constexpr (MODE == kSeg) {
        seg_val<L>(sg, ps) = Lds<T>::in(x[j]);
        seg_head<L>(sg, ps) = hb;
      } else {
        val[MODE == kRows ? lds_pad(ps) : ps] = Lds<T>::in(x[j]);
        head[ps] = hb;
      }
    }
  }
  __syncthreads();
  if constexpr (MODE == kRows) {
    if (n > 0) delta_reduce<T, OP>(n, val, head, hm, pc);
  } else if constexpr (MODE == kSeg) {
    seg8_reduce<T, OP>(n, sg);
  } else {
    vec_reduce<T, OP, 1>(n, val, head, hm);
  }
  // positions again from w2 (the opaque copy keeps the compiler from holding
  // the staging addresses across the reduction)
  uint32_t w3[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    w3[k] = w2[k];
    asm volatile("" : "+v"(w3[k]));
  }
  L y[8];  // every position is in the stage: read all, then store
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int ps = (int)((w3[j >> 1] >> ((j & 1) * 16)) & kPackPosMask);
    if constexpr (MODE == kSeg) y[j] = seg_val<L>(sg, ps);
    else y[j] = val[MODE == kRows ? lds_pad(ps) : ps];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (i0 + j * 64 < n) q[col[j]] = Lds<T>::out(y[j]);
}

// One block of the K = 1 GatherScatter from the 16-bit delta words, staged
// in layout MODE (val: kCap values, head4: kCap head bytes; kRows: val
// padded).  kRows serves lists of at least two entries per row (short blocks
// of rows; 96^3 p=3 227 -> 219 us), kSeg rows of at most 8 members, kPlain
// the rest (48^3 p=7: the kRows form measured 207 -> 230 us).
template <typename T, int OP, int MODE>
__device__ __forceinline__ void gs_sym_delta_block(int b, const int4 *__restrict__ part,
                                                   const int32_t *__restrict__ starts,
                                                   const int32_t *__restrict__ ids, const uint32_t *__restrict__ pk,
                                                   const int4 *__restrict__ base, int stride,
                                                   const u32x4 *__restrict__ dw, const uint32_t *__restrict__ dx,
                                                   int E, uint4 tb, int ncols, T *__restrict__ q,
                                                   typename Lds<T>::L *__restrict__ val, uint32_t *__restrict__ head4,
                                                   uint32_t *__restrict__ hm, int32_t *__restrict__ pc) {
  using V = VecShape<T, 1>;
  static_assert(V::SL == 8 && V::C == 1 && V::CAP == kCap, "delta pack: one entry per slot, kCap per block");
  uint8_t *head = reinterpret_cast<uint8_t *>(head4);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int4 p0 = part[b], p1 = part[b + 1];
  const u32x4 wd = __builtin_nontemporal_load(dw + (size_t)b * kBlock + tid);
  const uint32_t ex = lane < E ? __builtin_nontemporal_load(dx + ((size_t)b * kDRanges + wv) * E + lane) : 0u;
  if ((uint32_t)uniform((int)ex) == kDFallback) {
    // this block runs from its 32-bit words (every range carries the mark,
    // and the words' slot as exception word 1), staged plain
    const int slot = __builtin_amdgcn_readlane((int)ex, 1);
    const VecSrc src = {part, ids, pk, base, q, 0, stride};
    VecStage<V::SL> s;
    VecMeta m;
    T x[V::SL][1];
    vec_meta_at(src, b, slot, m);
    vec_index<T, 1>(s, src, slot, m);
    vec_gather<T, 1>(s, x);
    vec_finish<T, OP, 1>(s, m, starts, x, val, head, hm);
    return;
  }
  const int n = uniform(p1.y) - uniform(p0.y);
  const uint32_t w2[4] = {wd.x, wd.y, wd.z, wd.w};
  uint32_t col[V::SL];
  delta_columns(w2, ex, tb, ncols, col);
  delta_stage_reduce<T, OP, MODE>(w2, col, n, q, val, head, hm, pc);
}

#define OGS_DELTA_ARGS                                                                                         \
  const int4 *__restrict__ part, const int32_t *__restrict__ starts, const int32_t *__restrict__ ids,          \
      const uint32_t *__restrict__ pk, const int4 *__restrict__ base, int stride, const u32x4 *__restrict__ dw, \
      const uint32_t *__restrict__ dx, int E, uint4 tb, int ncols, T *__restrict__ q, const int *__restrict__ order

// kPlain or kSeg: kCap values and kCap head bytes in one stage (also the
// fallback path's).  kSeg held to 64 VGPRs (8 waves per SIMD; unbounded it
// took 69).
template <typename T, int OP, int MODE>
__global__ __launch_bounds__(kBlock) void k_gs_sym_delta(OGS_DELTA_ARGS) {
  using L = typename Lds<T>::L;
  __shared__ __align__(16) uint8_t stage[kCap * (sizeof(L) + 1)];
  __shared__ uint32_t hm[kCap / 32 + 1];
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  gs_sym_delta_block<T, OP, MODE>(order ? order[xb] : xb, part, starts, ids, pk, base, stride, dw, dx, E, tb, ncols, q,
                                  reinterpret_cast<L *>(stage), reinterpret_cast<uint32_t *>(stage + kCap * sizeof(L)),
                                  hm, nullptr);
}

// The kRows form (lists of >= 2 entries per row), held to 64 VGPRs (8 waves
// per SIMD; unbounded it took 69).
template <typename T, int OP>
__global__ __launch_bounds__(kBlock, 8) void k_gs_sym_delta_rows(OGS_DELTA_ARGS) {
  __shared__ typename Lds<T>::L val[lds_padded(kCap)];
  __shared__ uint32_t head4[kCap / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32];
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  gs_sym_delta_block<T, OP, kRows>(order ? order[xb] : xb, part, starts, ids, pk, base, stride, dw, dx, E, tb, ncols,
                                   q, val, head4, hm, pc);
}
#undef OGS_DELTA_ARGS

// ---------------------------------------------------------------------------
// Byte pack: the delta pack's blocks with one byte per entry.
//
// An entry's column, CSR position and head bit follow from the previous
// entry's (in its wave range, column order) by a triple (column delta,
// position delta mod kCap, head bit).  Code c < 255 names the plan's table
// word c, (column delta << 12) | (position delta << 1) | head bit, the 255
// commonest triples with column deltas below 2^20; code 255 takes the range's
// next exception item, two words (column delta, position delta << 1 | head
// bit).  The first entry of a range is always an exception (deltas from
// column 0, position 0).  A block's 2,048 codes are lane-major within each
// range (lane l's 8 codes contiguous: one 8-byte load per lane); its range
// headers hold each range's first exception item << 7 | item count (at most
// 64: one item per lane).  A block with a longer list, or left in CSR
// order, has kDFallback as header 0 and runs from the 16-bit words.  48^3
// p=7: 1.05 index bytes per entry instead of 2.13.
constexpr uint32_t kBEscape = 255u;
constexpr int kBMaxItems = 64;
constexpr uint32_t kBColLimit = 1u << 20;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Columns, CSR positions and head bits of a wave's 8 slots from its codes
// (cw: 4 per word) and its exception item (lane l holds the range's item l):
// per slot the two deltas (an exception's by its rank among the range's
// escapes, through lane permutes), two inclusive scans, the running carries.
// Positions and head bits come out as the 16-bit words' low 12 bits (two per
// dword of w2); columns are clamped to [0, ncols).
__device__ __forceinline__ void byte_columns(u32x2 cw, u32x2 ex, const uint32_t *__restrict__ tab, int ncols,
                                             uint32_t (&col)[8], uint32_t (&w2)[4]) {
  uint32_t carry = 0u, pcarry = 0u;
  int run = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t code = ((j < 4 ? cw.x : cw.y) >> ((j & 3) * 8)) & 255u;
    const uint32_t t = tab[code];
    uint32_t dc = t >> 12, dp = t & 0xfffu;
    const bool esc = code == kBEscape;
    const uint64_t mk = __builtin_amdgcn_ballot_w64(esc);
    if (mk) {
      const int rank = (run + lanes_below(mk)) & 63;
      const uint32_t e0 = (uint32_t)__builtin_amdgcn_ds_bpermute(rank * 4, (int)ex.x);
      const uint32_t e1 = (uint32_t)__builtin_amdgcn_ds_bpermute(rank * 4, (int)ex.y);
      dc = esc ? e0 : dc;
      dp = esc ? e1 : dp;
      run += __builtin_popcountll(mk);
    }
    const uint32_t sc = wave_scan_incl(dc);
    const uint32_t c = carry + sc;
    col[j] = c < (uint32_t)ncols ? c : (uint32_t)(ncols - 1);
    carry += (uint32_t)__builtin_amdgcn_readlane((int)sc, 63);
    const uint32_t sp = wave_scan_incl(dp >> 1);
    const uint32_t wj = ((pcarry + sp) & kPackPosMask) | ((dp & 1u) << kPackHeadBit);
    pcarry += (uint32_t)__builtin_amdgcn_readlane((int)sp, 63);
    if (j & 1) w2[j >> 1] |= wj << 16;
    else w2[j >> 1] = wj;
  }
}

// MODE kPlain or kSeg, as k_gs_sym_delta.
template <typename T, int OP, int MODE>
__global__ __launch_bounds__(kBlock) void k_gs_sym_byte(
    const int4 *__restrict__ part, const int32_t *__restrict__ starts, const int32_t *__restrict__ ids,
    const uint32_t *__restrict__ pk, const int4 *__restrict__ base, int stride, const u32x4 *__restrict__ dw,
    const uint32_t *__restrict__ dx, int E, uint4 tb, int ncols, T *__restrict__ q, const int *__restrict__ order,
    const u32x2 *__restrict__ bw, const uint4 *__restrict__ bh, const u32x2 *__restrict__ bx,
    const uint32_t *__restrict__ bt) {
  using L = typename Lds<T>::L;
  __shared__ __align__(16) uint8_t stage[kCap * (sizeof(L) + 1)];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ uint32_t tab[256];
  L *val = reinterpret_cast<L *>(stage);
  uint32_t *head4 = reinterpret_cast<uint32_t *>(stage + kCap * sizeof(L));
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  const uint4 h = bh[b];
  if (h.x == kDFallback) {
    gs_sym_delta_block<T, OP, MODE>(b, part, starts, ids, pk, base, stride, dw, dx, E, tb, ncols, q, val, head4, hm,
                                    nullptr);
    return;
  }
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const u32x2 cw = __builtin_nontemporal_load(bw + (size_t)b * kBlock + tid);
  tab[tid] = bt[tid];
  const uint32_t hr = (uint32_t)uniform((int)(wv == 0 ? h.x : wv == 1 ? h.y : wv == 2 ? h.z : h.w));
  const int cnt = (int)(hr & 127u);
  u32x2 ex = {0u, 0u};
  if (lane < cnt) ex = __builtin_nontemporal_load(bx + (size_t)(hr >> 7) + lane);
  const int n = uniform(part[b + 1].y) - uniform(part[b].y);
  __syncthreads();
  uint32_t col[8], w2[4];
  byte_columns(cw, ex, tab, ncols, col, w2);
  delta_stage_reduce<T, OP, MODE>(w2, col, n, q, val, reinterpret_cast<uint8_t *>(head4), hm, nullptr);
}

// ---------------------------------------------------------------------------
// The byte pack's blocks with KT interleaved fields: the one-list
// GatherScatter Sym at K = KT on lists whose rows have at most 8 members
// (the reference's four types and float-compensated storage).
// One decode of a block's codes (the K = 1 kernel's) feeds KT contiguous
// value loads per entry, so the index costs ~1 byte per entry instead of the
// 4-byte pack word of the entry-major kernel, whose blocks hold only as many
// entries as 16 KB of values (682 at K = 3 in fp64).  The stage holds the
// block's kCap / 8 groups of 8 CSR positions, each group's KT x 8 values
// (field-major) followed by its 8 head bytes; thread t folds the rows whose
// heads lie in group t for every field (as seg8_fold), holds the results
// until every thread has read the stage, and writes them back.
// Blocks without codes run from their 16-bit words, and blocks without those
// from their 32-bit words: the launcher runs this kernel only on packs that
// left no block in CSR order.
template <typename L, int KT>
__host__ __device__ constexpr int segk_group() { return KT * 8 * (int)sizeof(L) + 8; }

// The group-t fold of field f (seg8_fold's rows and bits): res[k] is the
// result of the row holding position 8t + k, from the group's first head on;
// with OPEN also before it, the row open at 8t folded again from group
// t - 1's last head (bits 0 .. 7: group t - 1's head bits, all set for
// t = 0).  OPEN serves the plain types' K-field GatherScatter (fp64 K = 3 at
// 48^3 p=7: 560 us, against 572 us folding each row once); the compensated
// sums, whose adds cost more, fold each row once.
template <typename T, int OP, int KT, bool OPEN>
__device__ __forceinline__ void segk_fold(const uint8_t *__restrict__ s, int f, uint32_t bits,
                                          typename Lds<T>::L (&res)[8]) {
  using L = typename Lds<T>::L;
  using R = LRed<T, OP>;
  using A = typename R::A;
  constexpr int G = segk_group<L, KT>();
  const int t = (int)threadIdx.x;
  const int gp = t > 0 ? t - 1 : t, gn = t + 1 < kCap / 8 ? t + 1 : t;
  const L *sp = reinterpret_cast<const L *>(s + gp * G) + 8 * f;
  const L *so = reinterpret_cast<const L *>(s + t * G) + 8 * f;
  const L *sn = reinterpret_cast<const L *>(s + gn * G) + 8 * f;
  L x[8];
  if constexpr (OPEN) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = sp[k];
  }
  A c = R::init();
  if constexpr (OPEN) {
    const int lh = 31 - __builtin_clz(bits & 0xffu);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      A y = c;
      R::apply(y, x[k]);
      c = k >= lh ? y : c;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = so[k];
  A part[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if ((bits >> (8 + k)) & 1u) c = R::init();
    R::apply(c, x[k]);
    part[k] = c;
  }
  __builtin_amdgcn_sched_barrier(0);
  // the last own row on to group t + 1's first head
  const int fh = __builtin_ctz(bits >> 16);
  if constexpr (OPEN) {
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = sn[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      A y = c;
      R::apply(y, x[k]);
      c = k < fh ? y : c;
    }
  } else {
    // only as many of group t + 1's positions as some lane of the wave
    // still needs (a row of at most 8 members continues by at most 7; at
    // p = 7 mostly by none or one)
    int fm = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) fm += __builtin_amdgcn_ballot_w64(fh > k) != 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      if (k < fm) {
        const L xk = sn[k];
        A y = c;
        R::apply(y, xk);
        c = k < fh ? y : c;
      }
    }
  }
  res[7] = R::fin(c);
#pragma unroll
  for (int k = 7; k > 0; --k) {
    if ((bits >> (8 + k)) & 1u) c = part[k - 1];
    res[k - 1] = R::fin(c);
  }
}

// Columns (col) and CSR positions with head bits (w2, as byte_columns) of
// block b's slots, from its byte codes, else its 16-bit words, else its
// 32-bit words (packed: the launchers' condition) at the slot in exception
// word 1.  tab: the block's 256-word LDS table.
__device__ __forceinline__ void byte_kv_decode(int b, int n, const uint32_t *__restrict__ pk,
                                               const int4 *__restrict__ base, int stride,
                                               const u32x4 *__restrict__ dw, const uint32_t *__restrict__ dx, int E,
                                               uint4 tb, int ncols, const u32x2 *__restrict__ bw,
                                               const uint4 *__restrict__ bh, const u32x2 *__restrict__ bx,
                                               const uint32_t *__restrict__ bt, uint32_t *__restrict__ tab,
                                               uint32_t (&col)[8], uint32_t (&w2)[4]) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint4 h = bh[b];
  if (uniform((int)h.x) != (int)kDFallback) {
    const u32x2 cw = __builtin_nontemporal_load(bw + (size_t)b * kBlock + tid);
    tab[tid] = bt[tid];
    const uint32_t hr = (uint32_t)uniform((int)(wv == 0 ? h.x : wv == 1 ? h.y : wv == 2 ? h.z : h.w));
    const int cnt = (int)(hr & 127u);
    u32x2 ex = {0u, 0u};
    if (lane < cnt) ex = __builtin_nontemporal_load(bx + (size_t)(hr >> 7) + lane);
    __syncthreads();
    byte_columns(cw, ex, tab, ncols, col, w2);
  } else {
    const u32x4 wd = __builtin_nontemporal_load(dw + (size_t)b * kBlock + tid);
    const uint32_t ex = lane < E ? __builtin_nontemporal_load(dx + ((size_t)b * kDRanges + wv) * E + lane) : 0u;
    if ((uint32_t)uniform((int)ex) != kDFallback) {
      w2[0] = wd.x; w2[1] = wd.y; w2[2] = wd.z; w2[3] = wd.w;
      delta_columns(w2, ex, tb, ncols, col);
    } else {
      // the block's 32-bit words, column-sorted (a packed block: the
      // launcher's condition), at the slot in exception word 1
      const int slot = __builtin_amdgcn_readlane((int)ex, 1);
      const int4 bq = base[slot];
      const int4 bb = make_int4(uniform(bq.x), uniform(bq.y), uniform(bq.z), uniform(bq.w));
      const uint32_t *w0 = pk + (size_t)slot * stride;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int e = wv * kDWave + j * 64 + lane;
        const uint32_t w = e < n ? __builtin_nontemporal_load(w0 + e) : 0u;
        const uint32_t c = (uint32_t)slot_col(w, bb, true);
        col[j] = c < (uint32_t)ncols ? c : (uint32_t)(ncols - 1);
        const uint32_t lo = w & ((1u << (kPackHeadBit + 1)) - 1u);  // position | head bit
        if (j & 1) w2[j >> 1] |= lo << 16;
        else w2[j >> 1] = lo;
      }
    }
  }
}

// The block's KT fields per slot loaded and staged as groups of 8 positions
// (KT x 8 values, field-major, then the 8 head bytes); ends with a barrier.
template <typename T, int KT>
__device__ __forceinline__ void byte_kv_stage(const uint32_t (&col)[8], const uint32_t (&w2)[4], int n,
                                              const T *__restrict__ q, uint8_t *__restrict__ stage) {
  using L = typename Lds<T>::L;
  constexpr int G = segk_group<L, KT>(), HV = KT * 8 * (int)sizeof(L);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  T x[8][KT];
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int f = 0; f < KT; ++f) x[j][f] = q[(size_t)col[j] * KT + f];
  const int i0 = wv * kDWave + lane;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (i0 + j * 64 < n) {
      const uint32_t wj = w2[j >> 1] >> ((j & 1) * 16);
      const int ps = (int)(wj & kPackPosMask);
      uint8_t *g = stage + (ps >> 3) * G;
#pragma unroll
      for (int f = 0; f < KT; ++f) reinterpret_cast<L *>(g)[8 * f + (ps & 7)] = Lds<T>::in(x[j][f]);
      g[HV + (ps & 7)] = (uint8_t)((wj >> kPackHeadBit) & 1u);
    }
  }
  __syncthreads();
}

template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gs_sym_byte_kv(
    const int4 *__restrict__ part, const uint32_t *__restrict__ pk, const int4 *__restrict__ base, int stride,
    const u32x4 *__restrict__ dw, const uint32_t *__restrict__ dx, int E, uint4 tb, int ncols, T *__restrict__ q,
    const int *__restrict__ order, const u32x2 *__restrict__ bw, const uint4 *__restrict__ bh,
    const u32x2 *__restrict__ bx, const uint32_t *__restrict__ bt) {
  using L = typename Lds<T>::L;
  constexpr int G = segk_group<L, KT>(), HV = KT * 8 * (int)sizeof(L);
  __shared__ __align__(16) uint8_t stage[(kCap / 8) * G];
  __shared__ uint32_t tab[256];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  const int n = uniform(part[b + 1].y) - uniform(part[b].y);
  uint32_t col[8], w2[4];
  byte_kv_decode(b, n, pk, base, stride, dw, dx, E, tb, ncols, bw, bh, bx, bt, tab, col, w2);
  byte_kv_stage<T, KT>(col, w2, n, q, stage);
  const int i0 = wv * kDWave + lane;
  // head bits of positions 8t - 8 .. 8t + 15 (bit k: position 8t - 8 + k),
  // positions from n on (and before 0) counted as heads
  constexpr bool OPEN = !std::is_same<typename LRed<T, OP>::A, Comp>::value;
  const int t = tid, P = 8 * t;
  const int gp = t > 0 ? t - 1 : t, gn = t + 1 < kCap / 8 ? t + 1 : t;
  uint32_t bits = (head_bits8(*reinterpret_cast<const uint2 *>(stage + t * G + HV)) << 8) |
                  (head_bits8(*reinterpret_cast<const uint2 *>(stage + gn * G + HV)) << 16);
  if constexpr (OPEN) {
    bits |= head_bits8(*reinterpret_cast<const uint2 *>(stage + gp * G + HV));
    if (t == 0) bits |= 0xffu;
  }
  const int live = n - (P - 8);
  if (live < 24) bits |= ~0u << (live > 0 ? live : 0);
  L res[KT][8];
#pragma unroll
  for (int f = 0; f < KT; ++f) segk_fold<T, OP, KT, OPEN>(stage, f, bits, res[f]);
  __syncthreads();
#pragma unroll
  for (int f = 0; f < KT; ++f) {
    if constexpr (OPEN) {
#pragma unroll
      for (int k = 0; k < 8; ++k) reinterpret_cast<L *>(stage + t * G)[8 * f + k] = res[f][k];
    } else {
      seg8_put<L>(stage + t * G + f * 8 * (int)sizeof(L), stage + (t + 1) * G + f * 8 * (int)sizeof(L), bits,
                  res[f]);
    }
  }
  __syncthreads();
  // positions again from w2 (the opaque copy keeps the compiler from holding
  // the staging addresses across the reduction)
  uint32_t w3[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    w3[k] = w2[k];
    asm volatile("" : "+v"(w3[k]));
  }
  L y[8][KT];  // every position is in the stage: read all, then store
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int ps = (int)((w3[j >> 1] >> ((j & 1) * 16)) & kPackPosMask);
    const L *g = reinterpret_cast<const L *>(stage + (ps >> 3) * G) + (ps & 7);
#pragma unroll
    for (int f = 0; f < KT; ++f) y[j][f] = g[8 * f];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (i0 + j * 64 < n) {
#pragma unroll
      for (int f = 0; f < KT; ++f) q[(size_t)col[j] * KT + f] = Lds<T>::out(y[j][f]);
    }
}

// Gather with KT fields over the byte pack's blocks (Trans/Sym, and NoTrans
// of an Unsigned operator; rows of at most 8 members): the K-field
// GatherScatter's decode and stage, each row folded once by the thread holding its
// head, the block's rows' results put in row order in the stage (row index
// = heads before it: per thread, a wave scan, the waves before through
// wsum) and stored as one contiguous run, gq[(r0 + row) * KT + f].
template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gather_byte_kv(
    const int4 *__restrict__ part, const uint32_t *__restrict__ pk, const int4 *__restrict__ base, int stride,
    const u32x4 *__restrict__ dw, const uint32_t *__restrict__ dx, int E, uint4 tb, int ncols,
    const T *__restrict__ q, T *__restrict__ gq, const int *__restrict__ order, const u32x2 *__restrict__ bw,
    const uint4 *__restrict__ bh, const u32x2 *__restrict__ bx, const uint32_t *__restrict__ bt) {
  using L = typename Lds<T>::L;
  constexpr int G = segk_group<L, KT>(), HV = KT * 8 * (int)sizeof(L);
  static_assert(kCap * KT * (int)sizeof(L) <= (kCap / 8) * G, "a block's row results fit the stage");
  __shared__ __align__(16) uint8_t stage[(kCap / 8) * G];
  __shared__ uint32_t tab[256];
  __shared__ uint32_t wsum[kDRanges];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  const int r0 = uniform(part[b].x);
  const int n = uniform(part[b + 1].y) - uniform(part[b].y);
  uint32_t col[8], w2[4];
  byte_kv_decode(b, n, pk, base, stride, dw, dx, E, tb, ncols, bw, bh, bx, bt, tab, col, w2);
  byte_kv_stage<T, KT>(col, w2, n, q, stage);
  // head bits of positions 8t .. 8t + 15 at bits 8 .. 23, positions from n
  // on counted as heads
  const int t = tid, P = 8 * t;
  const int gn = t + 1 < kCap / 8 ? t + 1 : t;
  uint32_t bits = (head_bits8(*reinterpret_cast<const uint2 *>(stage + t * G + HV)) << 8) |
                  (head_bits8(*reinterpret_cast<const uint2 *>(stage + gn * G + HV)) << 16);
  const int live = n - (P - 8);
  if (live < 24) bits |= ~0u << (live > 0 ? live : 0);
  L res[KT][8];
#pragma unroll
  for (int f = 0; f < KT; ++f) segk_fold<T, OP, KT, false>(stage, f, bits, res[f]);
  const int own = n - P;  // own positions below n
  const uint32_t hh = ((bits >> 8) & 0xffu) & (own >= 8 ? 0xffu : own > 0 ? (1u << own) - 1u : 0u);
  const uint32_t hc = (uint32_t)__popc(hh);
  const uint32_t inc = wave_scan_incl(hc);
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();  // (also: every thread's reads of the stage are done)
  uint32_t row = inc - hc;
  for (int w = 0; w < wv; ++w) row += wsum[w];
  const int rows = (int)(wsum[0] + wsum[1] + wsum[2] + wsum[3]);
  L *rb = reinterpret_cast<L *>(stage);  // rows <= n <= kCap results of KT fields
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if ((hh >> k) & 1u) {
      const int r = (int)row + __popc(hh & ((1u << k) - 1u));
#pragma unroll
      for (int f = 0; f < KT; ++f) rb[r * KT + f] = res[f][k];
    }
  __syncthreads();
  for (int i = tid; i < rows * KT; i += kBlock) gq[(size_t)r0 * KT + i] = Lds<T>::out(rb[i]);
}

// GatherScatter Trans (gather T, scatter N; ogsOperator.cpp:321-339) over
// the T list alone, on a plan with nMask: each entry's N membership from the
// column mask (a 4-byte word per 32 columns, cache resident), the T pack's
// blocks and reduction as in the one-list kernel, stores to N members only.
// Rows from NrowsN on have no N member, so nothing of theirs is stored.
template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gs_trans_mask(const int4 *__restrict__ part,
                                                          const int32_t *__restrict__ starts,
                                                          const int32_t *__restrict__ ids,
                                                          const uint32_t *__restrict__ pk,
                                                          const int4 *__restrict__ base, int stride,
                                                          T *__restrict__ q, const uint32_t *__restrict__ nMask) {
  using V = VecShape<T, KT>;
  using L = typename Lds<T>::L;
  __shared__ L val[vec_cap<L>()];
  __shared__ uint32_t head4[kCap / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  uint8_t *head = reinterpret_cast<uint8_t *>(head4);
  const VecSrc src = {part, ids, pk, base, q, 0, stride};
  VecStage<V::SL> s;
  VecMeta m;
  T x[V::SL][V::VW];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  vec_meta(src, b, m);
  vec_index<T, KT>(s, src, b, m);
  vec_gather<T, KT>(s, x);
  const bool packed = s.bb.x >= 0;
  const int ni = s.n * V::C;
  uint32_t mw[V::SL];
#pragma unroll
  for (int j = 0; j < V::SL; ++j)
    mw[j] = (int)threadIdx.x + j * kBlock < ni ? nMask[slot_col(s.w[j], s.bb, packed) >> 5] : 0u;
  uint32_t nbits = 0u;
#pragma unroll
  for (int j = 0; j < V::SL; ++j) nbits |= ((mw[j] >> (slot_col(s.w[j], s.bb, packed) & 31)) & 1u) << j;
  vec_finish<T, OP, KT, V::SL, true>(s, m, starts, x, val, head, hm, nbits);
}

// Row heads of a block as bytes at CSR positions: from the pack words, or
// from the row starts for a block kept in CSR order (no empty rows).
template <int SL>
__device__ __forceinline__ void stage_heads(const VecStage<SL> &s, const VecMeta &m, int C,
                                            const int32_t *__restrict__ starts, uint8_t *__restrict__ head) {
  const int tid = threadIdx.x;
  const int ni = s.n * C;
  const bool packed = s.bb.x >= 0;
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = tid + j * kBlock;
    if (it < ni && it % C == 0) {
      if (packed) head[s.w[j] & kPackPosMask] = (uint8_t)((s.w[j] >> kPackHeadBit) & 1u);
      else head[it / C] = 0;
    }
  }
  if (!packed) {
    const int r0 = uniform(m.p0.x), nr = uniform(m.p1.x) - r0;
    int rg[kRowSlots];
    row_words(starts, r0, nr, rg);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRowSlots; ++j) {
      const int r = tid + j * kBlock;
      if (r < nr) head[rg[j] - s.g0] = 1;
    }
  }
}

// The rows of a Gather block whose n entries' values are staged at their CSR
// positions with their head bytes (the staging barrier passed): the head
// bit mask and its per-word prefix counts, then one lane per (row head,
// field chunk) folds the row in CSR order and stores the result at its row,
// r0 + the number of row heads at or before its head.
template <typename T, int OP, int KT>
__device__ __forceinline__ void gather_rows(int n, int r0, const typename Lds<T>::L *__restrict__ val,
                                            const uint8_t *__restrict__ head, uint32_t *__restrict__ hm,
                                            int32_t *__restrict__ pc, T *__restrict__ gq) {
  using V = VecShape<T, KT>;
  using L = typename Lds<T>::L;
  using R = LRed<T, OP>;
  constexpr int VW = V::VW, C = V::C;
  const int ni = n * C;
  const int tid = threadIdx.x;
  head_mask(head, n, hm);
  __syncthreads();
  head_prefix(hm, n, pc);
  __syncthreads();
  for (int t = tid; t < ni; t += kBlock) {
    const int i = t / C, c = t - i * C;
    if (!((hm[i >> 5] >> (i & 31)) & 1u)) continue;
    const int len = row_len(hm, i);
    typename R::A red[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) red[v] = R::init();
    const L *sv = val + i * KT + c * VW;
    // CSR order, one member at a time (reading four at a time, as the
    // GatherScatter does, more than doubles this kernel's registers)
    for (int k = 0; k < len; ++k)
#pragma unroll
      for (int v = 0; v < VW; ++v) R::apply(red[v], sv[k * KT + v]);
    T *dst = gq + (size_t)(r0 + head_row(hm, pc, i)) * KT + c * VW;
#pragma unroll
    for (int v = 0; v < VW; ++v) dst[v] = Lds<T>::out(R::fin(red[v]));
  }
}

// Block b of a Gather whose pack words sit at `slot` (b, or a delta pack's
// fallback block's slot among the fallback blocks' words).
template <typename T, int OP, int KT>
__device__ __forceinline__ void gather_vec_block(const VecSrc &src, int b, int slot,
                                                 const int32_t *__restrict__ starts, T *__restrict__ gq,
                                                 typename Lds<T>::L *__restrict__ val, uint8_t *__restrict__ head,
                                                 uint32_t *__restrict__ hm, int32_t *__restrict__ pc) {
  using V = VecShape<T, KT>;
  constexpr int SL = V::SL, VW = V::VW, C = V::C;
  VecStage<SL> s;
  VecMeta m;
  T x[SL][VW];
  vec_meta_at(src, b, slot, m);
  vec_index<T, KT>(s, src, slot, m);
  vec_gather<T, KT>(s, x);
  const int n = s.n;
  if (n == 0) return;
  const int ni = n * C;
  const int tid = threadIdx.x;
  const int r0 = uniform(m.p0.x);
  const bool packed = s.bb.x >= 0;
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = tid + j * kBlock;
    if (it < ni) {
      const int c = it % C;
      const int ps = packed ? (int)(s.w[j] & kPackPosMask) : it / C;
#pragma unroll
      for (int v = 0; v < VW; ++v) val[ps * KT + c * VW + v] = Lds<T>::in(x[j][v]);
    }
  }
  stage_heads(s, m, C, starts, head);
  __syncthreads();
  gather_rows<T, OP, KT>(n, r0, val, head, hm, pc, gq);
}

// Gather of a T list without empty rows, KT interleaved fields (compile-time
// shape): the GatherScatter body up to the row reduction (values gathered
// through the column-sorted pack and staged at their CSR positions, rows
// found from the head bits, so no row starts are read), each (row, field
// chunk)'s result stored at its row -- the number of row heads at or before
// its head's position -- instead of scattered back.
template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gather_vec(const int4 *__restrict__ part,
                                                       const int32_t *__restrict__ starts,
                                                       const int32_t *__restrict__ ids,
                                                       const uint32_t *__restrict__ pk,
                                                       const int4 *__restrict__ base, int stride,
                                                       const T *__restrict__ q, T *__restrict__ gq,
                                                       const int *__restrict__ order) {
  using L = typename Lds<T>::L;
  __shared__ L val[vec_cap<L>()];
  __shared__ uint32_t head4[kCap / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32];
  const VecSrc src = {part, ids, pk, base, const_cast<T *>(q), 0, stride};
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  gather_vec_block<T, OP, KT>(src, b, b, starts, gq, val, reinterpret_cast<uint8_t *>(head4), hm, pc);
}

// The values of a delta-coded Gather block once a wave range's columns are
// known (as delta_stage_reduce), staged in layout MODE (kPlain: val and
// head; kSeg: the group stage at val) and reduced to gq[r0 + row].
template <typename T, int OP, int MODE>
__device__ __forceinline__ void gather_stage(const uint32_t (&w2)[4], const uint32_t (&col)[8], int n, int r0,
                                             const T *__restrict__ q, T *__restrict__ gq,
                                             typename Lds<T>::L *__restrict__ val, uint8_t *__restrict__ head,
                                             uint32_t *__restrict__ hm, int32_t *__restrict__ pc,
                                             uint32_t *__restrict__ wsum) {
  using L = typename Lds<T>::L;
  static_assert(MODE != kSeg || kSeg8<T>, "kSeg stages the reference's types");
  uint8_t *sg = reinterpret_cast<uint8_t *>(val);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  T x[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = q[col[j]];
  const int i0 = wv * kDWave + lane;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (i0 + j * 64 < n) {
      const uint32_t wj = w2[j >> 1] >> ((j & 1) * 16);
      const int ps = (int)(wj & kPackPosMask);
      const uint8_t hb = (uint8_t)((wj >> kPackHeadBit) & 1u);
      if constexpr (MODE == kSeg) {
        seg_val<L>(sg, ps) = Lds<T>::in(x[j]);
        seg_head<L>(sg, ps) = hb;
      } else {
        val[ps] = Lds<T>::in(x[j]);
        head[ps] = hb;
      }
    }
  }
  __syncthreads();
  if constexpr (MODE == kSeg) seg8_gather<T, OP>(n, r0, sg, wsum, gq);
  else gather_rows<T, OP, 1>(n, r0, val, head, hm, pc, gq);
}

// Gather K = 1 of block b over partition 0 from the 16-bit delta words (see
// k_gs_sym_delta): columns decoded per wave, values staged, rows reduced and
// stored as in k_gather_vec; fallback blocks run k_gather_vec's body from
// the 32-bit pack (staged plain: val holds kCap values, head4 kCap bytes).
template <typename T, int OP, int MODE>
__device__ __forceinline__ void gather_delta_block(int b, const int4 *__restrict__ part,
                                                   const int32_t *__restrict__ starts,
                                                   const int32_t *__restrict__ ids, const uint32_t *__restrict__ pk,
                                                   const int4 *__restrict__ base, int stride,
                                                   const u32x4 *__restrict__ dw, const uint32_t *__restrict__ dx,
                                                   int E, uint4 tb, int ncols, const T *__restrict__ q,
                                                   T *__restrict__ gq, typename Lds<T>::L *__restrict__ val,
                                                   uint32_t *__restrict__ head4, uint32_t *__restrict__ hm,
                                                   int32_t *__restrict__ pc, uint32_t *__restrict__ wsum) {
  using V = VecShape<T, 1>;
  static_assert(V::SL == 8 && V::C == 1 && V::CAP == kCap, "delta pack: one entry per slot, kCap per block");
  uint8_t *head = reinterpret_cast<uint8_t *>(head4);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int4 p0 = part[b], p1 = part[b + 1];
  const u32x4 wd = __builtin_nontemporal_load(dw + (size_t)b * kBlock + tid);
  const uint32_t ex = lane < E ? __builtin_nontemporal_load(dx + ((size_t)b * kDRanges + wv) * E + lane) : 0u;
  if ((uint32_t)uniform((int)ex) == kDFallback) {
    const int slot = __builtin_amdgcn_readlane((int)ex, 1);
    const VecSrc src = {part, ids, pk, base, const_cast<T *>(q), 0, stride};
    gather_vec_block<T, OP, 1>(src, b, slot, starts, gq, val, head, hm, pc);
    return;
  }
  const int r0 = uniform(p0.x);
  const int n = uniform(p1.y) - uniform(p0.y);
  const uint32_t w2[4] = {wd.x, wd.y, wd.z, wd.w};
  uint32_t col[V::SL];
  delta_columns(w2, ex, tb, ncols, col);
  gather_stage<T, OP, MODE>(w2, col, n, r0, q, gq, val, head, hm, pc, wsum);
}

#define OGS_GATHER_DELTA_ARGS                                                                                  \
  const int4 *__restrict__ part, const int32_t *__restrict__ starts, const int32_t *__restrict__ ids,          \
      const uint32_t *__restrict__ pk, const int4 *__restrict__ base, int stride, const u32x4 *__restrict__ dw, \
      const uint32_t *__restrict__ dx, int E, uint4 tb, int ncols, const T *__restrict__ q, T *__restrict__ gq, \
      const int *__restrict__ order

// MODE kPlain or kSeg (rows of at most 8 members).
template <typename T, int OP, int MODE>
__global__ __launch_bounds__(kBlock) void k_gather_delta(OGS_GATHER_DELTA_ARGS) {
  using L = typename Lds<T>::L;
  __shared__ __align__(16) uint8_t stage[kCap * (sizeof(L) + 1)];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32];
  __shared__ uint32_t wsum[kDRanges];
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  gather_delta_block<T, OP, MODE>(order ? order[xb] : xb, part, starts, ids, pk, base, stride, dw, dx, E, tb, ncols, q,
                                  gq, reinterpret_cast<L *>(stage),
                                  reinterpret_cast<uint32_t *>(stage + kCap * sizeof(L)), hm, pc, wsum);
}

// Gather K = 1 from the byte codes (see k_gs_sym_byte); blocks the codes do
// not cover run gather_delta_block.
template <typename T, int OP, int MODE>
__global__ __launch_bounds__(kBlock) void k_gather_byte(OGS_GATHER_DELTA_ARGS, const u32x2 *__restrict__ bw,
                                                        const uint4 *__restrict__ bh, const u32x2 *__restrict__ bx,
                                                        const uint32_t *__restrict__ bt) {
  using L = typename Lds<T>::L;
  __shared__ __align__(16) uint8_t stage[kCap * (sizeof(L) + 1)];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32];
  __shared__ uint32_t wsum[kDRanges];
  __shared__ uint32_t tab[256];
  L *val = reinterpret_cast<L *>(stage);
  uint32_t *head4 = reinterpret_cast<uint32_t *>(stage + kCap * sizeof(L));
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  const uint4 h = bh[b];
  if (h.x == kDFallback) {
    gather_delta_block<T, OP, MODE>(b, part, starts, ids, pk, base, stride, dw, dx, E, tb, ncols, q, gq, val, head4, hm,
                                    pc, wsum);
    return;
  }
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const u32x2 cw = __builtin_nontemporal_load(bw + (size_t)b * kBlock + tid);
  tab[tid] = bt[tid];
  const uint32_t hr = (uint32_t)uniform((int)(wv == 0 ? h.x : wv == 1 ? h.y : wv == 2 ? h.z : h.w));
  const int cnt = (int)(hr & 127u);
  u32x2 ex = {0u, 0u};
  if (lane < cnt) ex = __builtin_nontemporal_load(bx + (size_t)(hr >> 7) + lane);
  const int r0 = uniform(part[b].x);
  const int n = uniform(part[b + 1].y) - uniform(part[b].y);
  __syncthreads();
  uint32_t col[8], w2[4];
  byte_columns(cw, ex, tab, ncols, col, w2);
  gather_stage<T, OP, MODE>(w2, col, n, r0, q, gq, val, reinterpret_cast<uint8_t *>(head4), hm, pc, wsum);
}
#undef OGS_GATHER_DELTA_ARGS

// Scatter of a T list without empty rows, KT interleaved fields: the block's
// rows (KT values each, contiguous) are staged in LDS; an item's row is the
// number of row heads at or before its entry's CSR position (the head bit
// mask and per-word prefix counts), and each item stores its row's fields
// through its pack word (or in CSR order for a block the pack could not
// encode).
template <typename T, int KT>
__global__ __launch_bounds__(kBlock) void k_scatter_vec(const int4 *__restrict__ part,
                                                        const int32_t *__restrict__ starts,
                                                        const int32_t *__restrict__ ids,
                                                        const uint32_t *__restrict__ pk,
                                                        const int4 *__restrict__ base, int stride,
                                                        const T *__restrict__ gq, T *__restrict__ q,
                                                        const int *__restrict__ order) {
  using V = VecShape<T, KT>;
  constexpr int SL = V::SL, VW = V::VW, C = V::C;
  // a block's rows hold at most CAP * KT values (no more rows than entries)
  constexpr int RV = V::CAP * KT, RS = (RV + kBlock - 1) / kBlock;
  static_assert(RV * (int)sizeof(T) <= kVecBytes, "a block's rows fit the LDS stage");
  __shared__ T srow[RV];
  __shared__ uint32_t head4[kCap / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32];
  uint8_t *head = reinterpret_cast<uint8_t *>(head4);
  const VecSrc src = {part, ids, pk, base, q, 0, stride};
  VecStage<SL> s;
  VecMeta m;
  const int xb = xcd_block(blockIdx.x, gridDim.x);
  const int b = order ? order[xb] : xb;
  vec_meta(src, b, m);
  vec_index<T, KT>(s, src, b, m);
  const int n = s.n;
  if (n == 0) return;
  const int ni = n * C;
  const int tid = threadIdx.x;
  const int r0 = uniform(m.p0.x), nv = (uniform(m.p1.x) - r0) * KT;  // nr <= n: no empty rows
  const T *g = gq + (size_t)r0 * KT;
  T y[RS];
#pragma unroll
  for (int j = 0; j < RS; ++j) {
    const int r = tid + j * kBlock;
    y[j] = r < nv ? g[r] : T{};
  }
#pragma unroll
  for (int j = 0; j < RS; ++j) {
    const int r = tid + j * kBlock;
    if (r < nv) srow[r] = y[j];
  }
  stage_heads(s, m, C, starts, head);
  __syncthreads();
  head_mask(head, n, hm);
  __syncthreads();
  head_prefix(hm, n, pc);
  __syncthreads();
  const bool packed = s.bb.x >= 0;
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = tid + j * kBlock;
    if (it < ni) {
      const int c = it % C;
      const int ps = packed ? (int)(s.w[j] & kPackPosMask) : it / C;
      const T *sv = srow + head_row(hm, pc, ps) * KT + c * VW;
      T *d = q + (size_t)slot_col(s.w[j], s.bb, packed) * KT + c * VW;
#pragma unroll
      for (int v = 0; v < VW; ++v) d[v] = sv[v];
    }
  }
}

// Scatter by columns: v[c*K + k] = gv[colRow[c]*K + k] for every column c
// with a row (colRow[c] >= 0: the largest row of the list that holds c; a
// list built by Setup holds each column at most once), columns without a row
// untouched.  Workgroup b covers columns [b*cpb, (b+1)*cpb), whose values
// (at most kColVals) are contiguous, so the stores stream whole lines in
// order; each lane loads its columns' rows (non-temporal), then their values,
// then stores.  KT > 0: K = KT at compile time; KT = 0: any K (cpb = 1 past
// kColVals fields).
constexpr int kColVals = 2048;  // values per workgroup (8 per lane)

template <typename T, int KT>
__global__ __launch_bounds__(kBlock) void k_scatter_cols(int32_t ncols, int32_t K, int32_t cpb,
                                                         const int32_t *__restrict__ colRow,
                                                         const T *__restrict__ gv, T *__restrict__ v) {
  constexpr int U = kColVals / kBlock;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c0 = (int64_t)b * cpb;
  const int64_t left = (int64_t)ncols - c0;
  const int nc = left < cpb ? (int)left : cpb;
  const int tid = threadIdx.x;
  if constexpr (KT > 0) {
    const int nv = nc * KT;
    int r[U], k[U];
    int64_t c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = tid + u * kBlock;
      const int cl = t / KT;
      k[u] = t - cl * KT;
      c[u] = c0 + cl;
      r[u] = t < nv ? __builtin_nontemporal_load(colRow + c[u]) : -1;
    }
    T x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r[u] >= 0) x[u] = gv[(size_t)r[u] * KT + k[u]];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r[u] >= 0) v[(size_t)c[u] * KT + k[u]] = x[u];
  } else {
    const int64_t nv = (int64_t)nc * K;
    for (int64_t t = tid; t < nv; t += kBlock) {
      const int64_t cl = t / K, kk = t - cl * K;
      const int r = colRow[c0 + cl];
      if (r >= 0) v[(size_t)(c0 + cl) * K + kk] = gv[(size_t)r * K + kk];
    }
  }
}

// ---------------------------------------------------------------------------
// Row-delta map: the column -> row map of a list (colRow) in 4-bit codes,
// for the by-column Scatter at K = 1.  Workgroup b covers columns
// [2048 b, 2048 b + 2048) in four wave ranges of 512 (slot j of lane l of
// wave w: column 2048 b + 512 w + 64 j + l, so each wave instruction covers
// 64 consecutive columns).  A column's row is the previous row-bearing
// column's row in its range (0 before the first) plus a delta: code 0 marks
// a column without a row (delta 0, nothing stored), codes 1-14 add the
// plan's table entry (the 14 commonest nonzero deltas, 32 bits each: on a
// structured mesh +1 along an element's new rows, and the jumps back to
// rows a line or a layer of elements earlier), code 15 adds the range's
// next exception word (32 bits, wrapping); the
// first row-bearing column of a range is always an exception (its row, so
// exception word 0 is a row and never the fallback mark).  A lane's 8 codes
// make one dword.  Chunks with more than E exceptions in a range carry the
// fallback mark as every range's exception word 0 and read colRow.
// 48^3 p=7 (56.6 M columns): 226 MB of colRow per Scatter -> ~0.6 B per
// column.
constexpr int kRowChunk = 2048;
constexpr int kRowSample = 1 << 20;  // columns whose deltas pick the table

// Row of column i and whether it is the first row-bearing column of its
// range; prev = the previous row-bearing column's row in the range (or 0).
__device__ __forceinline__ void row_chain(const int32_t *__restrict__ colRow, int64_t ncols, int64_t c0, int i,
                                          int32_t *row, uint32_t *prev, bool *first) {
  const int64_t c = c0 + i;
  *row = c < ncols ? colRow[c] : -1;
  const int r0 = i & ~(kDWave - 1);
  int j = i - 1;
  int32_t pr = -1;
  while (j >= r0) {
    const int32_t x = c0 + j < ncols ? colRow[c0 + j] : -1;
    if (x >= 0) { pr = x; break; }
    --j;
  }
  *first = pr < 0;
  *prev = pr < 0 ? 0u : (uint32_t)pr;
}

// Code of a row under the table (entries 1-14 in LDS; 15: not in it).
__device__ __forceinline__ uint32_t row_code(int32_t row, uint32_t prev, bool first, const uint32_t *tbl) {
  if (row < 0) return 0u;
  if (first) return kDEscape;
  const uint32_t d = (uint32_t)row - prev;
  uint32_t code = kDEscape;
#pragma unroll
  for (int k = 1; k < (int)kDEscape; ++k)
    if (tbl[k] == d && d != 0u) code = (uint32_t)k;
  return code;
}

// Escapes in wave w's range (every lane gets the count).
__device__ __forceinline__ int row_escapes(const int32_t *__restrict__ colRow, int64_t ncols, int64_t c0,
                                           const uint32_t *tbl) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int cnt = 0;
  for (int j = 0; j < kDWave / 64; ++j) {
    int32_t row;
    uint32_t prev;
    bool first;
    row_chain(colRow, ncols, c0, w * kDWave + j * 64 + lane, &row, &prev, &first);
    cnt += __builtin_popcountll(__builtin_amdgcn_ballot_w64(row_code(row, prev, first, tbl) == kDEscape));
  }
  return cnt;
}

// hist[min(m, kDMaxE + 1)] += 1 per chunk, m = its most escapes in one range.
__global__ __launch_bounds__(kBlock) void k_row_delta_count(const int32_t *__restrict__ colRow, int64_t ncols,
                                                            const uint32_t *__restrict__ table,
                                                            uint32_t *__restrict__ hist) {
  __shared__ uint32_t tbl[16];
  __shared__ int most;
  if (threadIdx.x < 16) tbl[threadIdx.x] = table[threadIdx.x];
  if (threadIdx.x == 0) most = 0;
  __syncthreads();
  const int c = row_escapes(colRow, ncols, (int64_t)blockIdx.x * kRowChunk, tbl);
  if ((threadIdx.x & 63) == 0) atomicMax(&most, c);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&hist[most < kDMaxE + 1 ? most : kDMaxE + 1], 1u);
}

// Codes and exceptions (dx zeroed beforehand); a chunk over E escapes in a
// range gets the fallback mark and counts in *fallbacks.
__global__ __launch_bounds__(kBlock) void k_row_delta_build(const int32_t *__restrict__ colRow, int64_t ncols,
                                                            const uint32_t *__restrict__ table, int E,
                                                            uint32_t *__restrict__ codes, uint32_t *__restrict__ dx,
                                                            int32_t *__restrict__ fallbacks) {
  __shared__ uint32_t tbl[16];
  __shared__ int over;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t c0 = (int64_t)b * kRowChunk;
  if (tid < 16) tbl[tid] = table[tid];
  if (tid == 0) over = 0;
  __syncthreads();
  if (row_escapes(colRow, ncols, c0, tbl) > E && lane == 0) over = 1;
  __syncthreads();
  uint32_t *xw = dx + ((size_t)b * kDRanges + w) * E;
  if (over) {
    if (lane == 0) xw[0] = kDFallback;
    if (tid == 0) atomicAdd(fallbacks, 1);
    return;
  }
  uint32_t word = 0u;
  int run = 0;
  for (int j = 0; j < kDWave / 64; ++j) {
    int32_t row;
    uint32_t prev;
    bool first;
    row_chain(colRow, ncols, c0, w * kDWave + j * 64 + lane, &row, &prev, &first);
    const uint32_t code = row_code(row, prev, first, tbl);
    const bool esc = code == kDEscape;
    const uint64_t mk = __builtin_amdgcn_ballot_w64(esc);
    if (esc) xw[run + lanes_below(mk)] = (uint32_t)row - prev;
    run += __builtin_popcountll(mk);
    word |= code << (4 * j);
  }
  codes[(size_t)b * kBlock + tid] = word;
}

// Rows of a lane's 8 slots from its code dword, its range's exception word
// (lane l holds exception l) and the table (lane k < 16 holds entry k; 0 at
// codes 0 and 15): per slot the table delta by a lane permute (code 0: 0,
// no row), an escape by its rank among the range's escapes, an inclusive
// scan, the running carry.  Rows are clamped to [0, nrows).
__device__ __forceinline__ uint32_t row_deltas(uint32_t codes, uint32_t ex, uint32_t tv, int nrows,
                                               int32_t (&row)[kDWave / 64]) {
  uint32_t carry = 0u, has = 0u;
  int run = 0;
#pragma unroll
  for (int j = 0; j < kDWave / 64; ++j) {
    const uint32_t code = (codes >> (4 * j)) & 15u;
    uint32_t d = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(code * 4u), (int)tv);
    const bool esc = code == kDEscape;
    const uint64_t mk = __builtin_amdgcn_ballot_w64(esc);
    if (mk) {
      const int rank = (run + lanes_below(mk)) & 63;
      const uint32_t e = (uint32_t)__builtin_amdgcn_ds_bpermute(rank * 4, (int)ex);
      d = esc ? e : d;
      run += __builtin_popcountll(mk);
    }
    const uint32_t sc = wave_scan_incl(d);
    const int32_t r = (int32_t)(carry + sc);
    row[j] = r < 0 ? 0 : (r < nrows ? r : nrows - 1);
    carry += (uint32_t)__builtin_amdgcn_readlane((int)sc, 63);
    has |= (code != 0u ? 1u : 0u) << j;
  }
  return has;
}

// Scatter K = 1 by columns over the row-delta map (see k_scatter_cols):
// one code dword and (first E lanes) one exception word per lane, the rows
// decoded per wave, their values gathered, then stored in column order.
// Fallback chunks read colRow, or, with fbRows, their slot there (exception
// word 1).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter_rows_delta(int64_t ncols, int32_t nrows,
                                                               const int32_t *__restrict__ colRow,
                                                               const int32_t *__restrict__ fbRows,
                                                               const uint32_t *__restrict__ codes,
                                                               const uint32_t *__restrict__ dx, int E,
                                                               const uint32_t *__restrict__ table,
                                                               const T *__restrict__ gv, T *__restrict__ v) {
  constexpr int U = kDWave / 64;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t c0 = (int64_t)b * kRowChunk + w * kDWave + lane;  // slot j: c0 + 64 j
  const uint32_t cw = __builtin_nontemporal_load(codes + (size_t)b * kBlock + tid);
  const uint32_t ex = lane < E ? __builtin_nontemporal_load(dx + ((size_t)b * kDRanges + w) * E + lane) : 0u;
  const uint32_t tv = lane < 16 ? table[lane] : 0u;
  int32_t row[U];
  uint32_t has = 0u;
  if ((uint32_t)uniform((int)ex) == kDFallback) {
    const int64_t slot = __builtin_amdgcn_readlane((int)ex, 1);
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int64_t c = c0 + 64 * j;
      if (c >= ncols) row[j] = -1;
      else if (fbRows) row[j] = __builtin_nontemporal_load(fbRows + slot * kRowChunk + (c - (int64_t)b * kRowChunk));
      else row[j] = __builtin_nontemporal_load(colRow + c);
      has |= (row[j] >= 0 ? 1u : 0u) << j;
    }
  } else {
    has = row_deltas(cw, ex, tv, nrows, row);
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (c0 + 64 * j >= ncols) has &= ~(1u << j);  // (the map codes no row there)
  }
  T x[U];
#pragma unroll
  for (int j = 0; j < U; ++j)
    if ((has >> j) & 1u) x[j] = gv[row[j]];
#pragma unroll
  for (int j = 0; j < U; ++j)
    if ((has >> j) & 1u) v[c0 + 64 * j] = x[j];
}

// The fallback chunks' rows of a row-delta map in an array of their own:
// each fallback chunk takes the next slot (written as its ranges' exception
// word 1) and copies its kRowChunk entries of colRow there (-1 past ncols).
__global__ __launch_bounds__(kBlock) void k_compact_rows(uint32_t *__restrict__ dx, int E,
                                                         const int32_t *__restrict__ colRow, int64_t ncols,
                                                         int32_t *__restrict__ fbRows, int32_t *__restrict__ count) {
  __shared__ int slot;
  const int b = blockIdx.x;
  if (dx[(size_t)b * kDRanges * E] != kDFallback) return;
  if (threadIdx.x == 0) slot = atomicAdd(count, 1);
  __syncthreads();
  const int sl = slot;
  if (threadIdx.x < kDRanges) dx[((size_t)b * kDRanges + threadIdx.x) * E + 1] = (uint32_t)sl;
  for (int i = threadIdx.x; i < kRowChunk; i += kBlock) {
    const int64_t c = (int64_t)b * kRowChunk + i;
    fbRows[(size_t)sl * kRowChunk + i] = c < ncols ? colRow[c] : -1;
  }
}

// colRow[c] = the largest row of rows [0, nrows) whose list holds column c
// (colRow set to -1 beforehand), one lane per row.
__global__ __launch_bounds__(kBlock) void k_col_rows(const int32_t *__restrict__ rs, int32_t nrows,
                                                     const int32_t *__restrict__ ci, int32_t *__restrict__ colRow) {
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kBlock)
    for (int j = rs[r]; j < rs[r + 1]; ++j) atomicMax(colRow + ci[j], (int32_t)r);
}

// The N list as one bit per column (nMask), checked against the T list:
// bad[0] counts T rows that do not strictly ascend, N rows that do not, and
// N entries outside their T row (colRow[c] != r); with distinct T columns
// (counted by k_count_cols) a zero count makes the N rows ordered subsets of
// their T rows.
__global__ __launch_bounds__(kBlock) void k_n_mask(const int32_t *__restrict__ rsT, const int32_t *__restrict__ ciT,
                                                   const int32_t *__restrict__ rsN, const int32_t *__restrict__ ciN,
                                                   int32_t nrows, const int32_t *__restrict__ colRow,
                                                   uint32_t *__restrict__ mask, int32_t *__restrict__ bad) {
  int nb = 0;
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kBlock) {
    for (int j = rsT[r] + 1; j < rsT[r + 1]; ++j) nb += ciT[j] <= ciT[j - 1];
    for (int j = rsN[r]; j < rsN[r + 1]; ++j) {
      const int c = ciN[j];
      if (j > rsN[r] && c <= ciN[j - 1]) ++nb;
      if (colRow[c] != (int32_t)r) ++nb;
      else atomicOr(mask + (c >> 5), 1u << (c & 31));
    }
  }
  if (nb) atomicAdd(bad, nb);
}

// Columns held by some row (colRow >= 0), added to cnt.
__global__ __launch_bounds__(kBlock) void k_count_cols(const int32_t *__restrict__ colRow, int32_t ncols,
                                                       int32_t *__restrict__ cnt) {
  int n = 0;
  for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < ncols; c += (int64_t)gridDim.x * kBlock)
    n += colRow[c] >= 0;
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) n += __shfl_xor(n, d);
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(cnt, n);
}

// owner[c] for the by-column GatherScatter NoTrans: the N-list member of the
// T row holding column c (-1: no row; -2: a row with no N member), for an N
// list of at most one member per row.
__global__ __launch_bounds__(kBlock) void k_col_owner(int32_t ncols, const int32_t *__restrict__ colRow,
                                                      const int32_t *__restrict__ rsN,
                                                      const int32_t *__restrict__ ciN, int32_t *__restrict__ owner) {
  for (int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x; c < ncols; c += (int64_t)gridDim.x * kBlock) {
    const int r = colRow[c];
    owner[c] = r < 0 ? -1 : (rsN[r + 1] > rsN[r] ? ciN[rsN[r]] : -2);
  }
}

// GatherScatter NoTrans, K = 1, of an operator whose N list holds at most
// one member per row (Signed unique): by columns, every column with a row
// takes f(v[owner]) = fin(apply(init, v[owner])) -- the row's reduction over
// its single N member in the host's arithmetic -- or fin(init) for a row
// without one.  The owner's own lane may rewrite v[owner] meanwhile, but
// f(f(x)) = f(x) for every op and type (x + 0, x * 1, compare-select with the
// identity, and the storage rounding are idempotent), so a column reads the
// same result either way.  Workgroup b covers columns [b*kColVals,
// (b+1)*kColVals), stored in order.
template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gs_owner(int32_t ncols, int32_t K, int32_t cpb,
                                                     const int32_t *__restrict__ owner, T *v) {
  using R = LRed<T, OP>;
  constexpr int U = kColVals / kBlock;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int64_t c0 = (int64_t)b * cpb;
  const int64_t left = (int64_t)ncols - c0;
  const int nc = left < cpb ? (int)left : cpb;
  const int tid = threadIdx.x;
  if constexpr (KT > 0) {
    // item t: field t % KT of column c0 + t / KT
    const int nv = nc * KT;
    int o[U], k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = tid + u * kBlock;
      const int cl = t / KT;
      k[u] = t - cl * KT;
      o[u] = t < nv ? __builtin_nontemporal_load(owner + c0 + cl) : -1;
    }
    T x[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (o[u] >= 0) x[u] = v[(size_t)o[u] * KT + k[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (o[u] == -1) continue;
      typename R::A a = R::init();
      if (o[u] >= 0) R::apply(a, Lds<T>::in(x[u]));
      v[(size_t)c0 * KT + tid + u * kBlock] = Lds<T>::out(R::fin(a));
    }
  } else {
    const int64_t nv = (int64_t)nc * K;
    for (int64_t t = tid; t < nv; t += kBlock) {
      const int64_t cl = t / K, kk = t - cl * K;
      const int o = owner[c0 + cl];
      if (o == -1) continue;
      typename R::A a = R::init();
      if (o >= 0) R::apply(a, Lds<T>::in(v[(size_t)o * K + kk]));
      v[(size_t)(c0 + cl) * K + kk] = Lds<T>::out(R::fin(a));
    }
  }
}

// GatherScatter with two lists (NoTrans / Trans of an operator whose N and T
// lists differ), KT interleaved fields (compile-time shape), over an
// entry-cut T partition whose blocks hold at most CAP/2 entries of either
// list, so that the gathered values, the rows' results and the row starts
// share one list's LDS budget.  Gather side: items gathered through the
// list's pack (or in CSR order) and staged at their CSR positions; one lane
// per (row, field chunk) reduces the row in CSR order from the row starts (an
// empty row gives the op's identity, as on the host).  Scatter side: a map
// from CSR position to row, from the scatter list's row starts; each item
// stores its row's result through its pack word (or in CSR order).
template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_gs2_vec(const int4 *__restrict__ part, ListView G, int gUseN, ListView S,
                                                    int sUseN, T *__restrict__ q) {
  using V = VecShape<T, KT>;
  using L = typename Lds<T>::L;
  using R = LRed<T, OP>;
  constexpr int SL = V::SL, VW = V::VW, C = V::C;
  constexpr int CAP2 = V::CAP / 2;
  constexpr int RSL = (CAP2 + 1 + kBlock - 1) / kBlock;
  __shared__ L val[CAP2 * KT];
  __shared__ L rowv[CAP2 * KT];  // no more rows than T entries
  __shared__ int32_t rsg[CAP2 + 1], rss[CAP2 + 1];
  __shared__ int32_t srow[CAP2];
  const VecSrc gsrc = {part, G.ids, G.pk, G.base, q, gUseN, G.stride};
  const VecSrc ssrc = {part, S.ids, S.pk, S.base, q, sUseN, S.stride};
  VecStage<SL> gs, ss;
  VecMeta m, ms;
  T x[SL][VW];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  vec_meta(gsrc, b, m);
  vec_meta(ssrc, b, ms);
  vec_index<T, KT>(gs, gsrc, b, m);
  vec_index<T, KT>(ss, ssrc, b, ms);
  vec_gather<T, KT>(gs, x);
  const int tid = threadIdx.x;
  const int r0 = uniform(m.p0.x), nr = uniform(m.p1.x) - r0;
  int rg[RSL], rv[RSL];
#pragma unroll
  for (int j = 0; j < RSL; ++j) {
    const int r = tid + j * kBlock;
    rg[j] = r <= nr ? G.starts[r0 + r] - gs.g0 : 0;
    rv[j] = r <= nr ? S.starts[r0 + r] - ss.g0 : 0;
  }
  const int ng = gs.n * C;
  const bool gp = gs.bb.x >= 0;
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = tid + j * kBlock;
    if (it < ng) {
      const int c = it % C;
      const int ps = gp ? (int)(gs.w[j] & kPackPosMask) : it / C;
#pragma unroll
      for (int v = 0; v < VW; ++v) val[ps * KT + c * VW + v] = Lds<T>::in(x[j][v]);
    }
  }
#pragma unroll
  for (int j = 0; j < RSL; ++j) {
    const int r = tid + j * kBlock;
    if (r <= nr) {
      rsg[r] = rg[j];
      rss[r] = rv[j];
    }
  }
  __syncthreads();
  for (int r = tid; r < nr; r += kBlock) {
    const int z = rss[r + 1];
    for (int i = rss[r]; i < z; ++i) srow[i] = r;
  }
  for (int t = tid; t < nr * C; t += kBlock) {
    const int r = t / C, c = t - r * C;
    const int z = rsg[r + 1];
    typename R::A red[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) red[v] = R::init();
    for (int i = rsg[r]; i < z; ++i)
#pragma unroll
      for (int v = 0; v < VW; ++v) R::apply(red[v], val[i * KT + c * VW + v]);
#pragma unroll
    for (int v = 0; v < VW; ++v) rowv[r * KT + c * VW + v] = R::fin(red[v]);
  }
  __syncthreads();
  const int ns = ss.n * C;
  const bool sp = ss.bb.x >= 0;
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int it = tid + j * kBlock;
    if (it < ns) {
      const int c = it % C;
      const int ps = sp ? (int)(ss.w[j] & kPackPosMask) : it / C;
      const L *sv = rowv + srow[ps] * KT + c * VW;
      T *d = q + (size_t)slot_col(ss.w[j], ss.bb, sp) * KT + c * VW;
#pragma unroll
      for (int v = 0; v < VW; ++v) d[v] = Lds<T>::out(sv[v]);
    }
  }
}

// Batched form: one (descriptor, block) item per workgroup.
template <typename T, int OP, int KT>
__device__ __forceinline__ void batch_item(const BatchDev *__restrict__ descs, int2 w,
                                           typename Lds<T>::L *__restrict__ val, uint8_t *__restrict__ head,
                                           uint32_t *__restrict__ hm, int32_t *__restrict__ pc) {
  using V = VecShape<T, KT>;
  const BatchDev d = load_desc(descs, w.x);
  const VecSrc src = {d.part, d.G.ids, d.G.pk, d.G.base, d.q, 0, d.G.stride};
  VecStage<V::SL> s;
  VecMeta m;
  T x[V::SL][V::VW];
  vec_meta(src, w.y, m);
  vec_index<T, KT>(s, src, w.y, m);
  vec_gather<T, KT>(s, x);
  if constexpr (rows_reduce<T, KT>()) {
    const RowIndex ri = {pc, reinterpret_cast<uint16_t *>(head)};
    vec_finish<T, OP, KT>(s, m, d.G.starts, x, val, head, hm, 0u, &ri);
  } else {
    vec_finish<T, OP, KT>(s, m, d.G.starts, x, val, head, hm);
  }
}

// Head-region bytes of a kernel running every batched field-count class.
template <typename T> __host__ __device__ constexpr int batch_head_bytes() {
  constexpr int b[6] = {head_bytes<T, 1>(), head_bytes<T, 2>(), head_bytes<T, 3>(),
                        head_bytes<T, 4>(), head_bytes<T, 8>(), head_bytes<T, 16>()};
  int m = 0;
  for (int i = 0; i < 6; ++i) m = b[i] > m ? b[i] : m;
  return m;
}

// One node per field-count class: the items of each XCD form one contiguous
// run of the class's list.
template <typename T, int OP, int KT>
__global__ __launch_bounds__(kBlock) void k_batch_vec(const BatchDev *__restrict__ descs,
                                                      const int2 *__restrict__ items) {
  using L = typename Lds<T>::L;
  __shared__ L val[vec_cap<L>()];
  __shared__ uint32_t head4[(head_bytes<T, KT>() + 3) / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32 + 1];
  batch_item<T, OP, KT>(descs, items[xcd_block(blockIdx.x, gridDim.x)], val, reinterpret_cast<uint8_t *>(head4), hm,
                        pc);
}

// All classes in one node: workgroup b runs item b of the concatenated class
// lists (class c holds items [end[c-1], end[c])) with that class's body.
struct ClassEnds {
  int end[6];
};

template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_batch_fused(const BatchDev *__restrict__ descs,
                                                        const int2 *__restrict__ items, ClassEnds ce) {
  using L = typename Lds<T>::L;
  __shared__ L val[vec_cap<L>()];
  __shared__ uint32_t head4[(batch_head_bytes<T>() + 3) / 4];
  __shared__ uint32_t hm[kCap / 32 + 1];
  __shared__ int32_t pc[kCap / 32 + 1];
  uint8_t *head = reinterpret_cast<uint8_t *>(head4);
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int2 w = items[b];
  if (b < ce.end[0]) batch_item<T, OP, 1>(descs, w, val, head, hm, pc);
  else if (b < ce.end[1]) batch_item<T, OP, 2>(descs, w, val, head, hm, pc);
  else if (b < ce.end[2]) batch_item<T, OP, 3>(descs, w, val, head, hm, pc);
  else if (b < ce.end[3]) batch_item<T, OP, 4>(descs, w, val, head, hm, pc);
  else if (b < ce.end[4]) batch_item<T, OP, 8>(descs, w, val, head, hm, pc);
  else batch_item<T, OP, 16>(descs, w, val, head, hm, pc);
}

// ---------------------------------------------------------------------------
// Gather: gv[k + row*K] = OP over the row's list.  Rows [0, nrows).
template <typename T, int OP>
__global__ __launch_bounds__(kBlock) void k_gather(const int4 *__restrict__ part, int nrows,
                                                   const int32_t *__restrict__ gS, const int32_t *__restrict__ gI, int gUseN,
                                                   int K, int kc, const T *__restrict__ q, T *__restrict__ gq) {
  __shared__ T gval[kCap];
  __shared__ int32_t rsG[kWMax + 1];

  Extent e;
  if (!block_extent(part, xcd_block(blockIdx.x, gridDim.x), nrows, gS, gUseN, gS, gUseN, e)) return;
  const int tid = threadIdx.x;

#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    if (r <= e.nr) rsG[r] = gS[e.r0 + r] - e.g0;
  }
  for (int k0 = 0; k0 < K; k0 += kc) {
    const int kw = min(kc, K - k0);
    const int nvG = e.nG * kw;
    size_t off[kSlots];
    T x[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvG) {
        const int i = f / kw;
        off[j] = (size_t)gI[e.g0 + i] * K + k0 + (f - i * kw);
      }
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvG) x[j] = q[off[j]];
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvG) gval[f] = x[j];
    }
    __syncthreads();
    for (int t = tid; t < e.nr * kw; t += kBlock) {
      const int r = t / kw, k = t - r * kw;
      const int a = rsG[r], b = rsG[r + 1];
      typename Red<T, OP>::A red = Red<T, OP>::init();
      for (int i = a; i < b; ++i) Red<T, OP>::apply(red, gval[i * kw + k]);
      const T acc = Red<T, OP>::fin(red);
      gq[(size_t)(e.r0 + r) * K + k0 + k] = acc;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Scatter: v[k + col*K] = gv[k + row*K].  Rows [0, nrows).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter(const int4 *__restrict__ part, int nrows,
                                                    const int32_t *__restrict__ sS, const int32_t *__restrict__ sI, int sUseN,
                                                    int K, int kc, const T *__restrict__ gq, T *__restrict__ q) {
  __shared__ T sval[kCap];
  __shared__ int32_t rsS[kWMax + 1];

  Extent e;
  if (!block_extent(part, xcd_block(blockIdx.x, gridDim.x), nrows, sS, sUseN, sS, sUseN, e)) return;
  const int tid = threadIdx.x;

#pragma unroll
  for (int j = 0; j < kRowSlots; ++j) {
    const int r = tid + j * kBlock;
    if (r <= e.nr) rsS[r] = sS[e.r0 + r] - e.s0;
  }
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += kc) {
    const int kw = min(kc, K - k0);
    const int ntask = e.nr * kw;
    for (int t0 = 0; t0 < ntask; t0 += kBlock * kSlots) {
      T x[kSlots];
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int t = t0 + tid + j * kBlock;
        if (t < ntask) {
          const int r = t / kw, k = t - r * kw;
          x[j] = gq[(size_t)(e.r0 + r) * K + k0 + k];
        }
      }
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int t = t0 + tid + j * kBlock;
        if (t < ntask) {
          const int r = t / kw, k = t - r * kw;
          const int a = rsS[r], b = rsS[r + 1];
          for (int i = a; i < b; ++i) sval[i * kw + k] = x[j];
        }
      }
    }
    __syncthreads();
    const int nvS = e.nS * kw;
    size_t off[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvS) {
        const int i = f / kw;
        off[j] = (size_t)sI[e.s0 + i] * K + k0 + (f - i * kw);
      }
    }
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int f = tid + j * kBlock;
      if (f < nvS) q[off[j]] = sval[f];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// extract (ogsKernels.okl:167-177): gq[n] = q[n%K + ids[n/K]*K]
template <typename T>
__global__ __launch_bounds__(kBlock) void k_extract(int64_t total, int K, const int32_t *__restrict__ ids,
                                                    const T *__restrict__ q, T *__restrict__ gq) {
  for (int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x; n < total; n += (int64_t)gridDim.x * kBlock) {
    const int64_t g = n / K;
    const int k = (int)(n - g * K);
    gq[n] = q[(size_t)ids[g] * K + k];
  }
}

// ---------------------------------------------------------------------------
// Plan-time kernels.

// Longest row of a CSR and a validity check of its row starts and column ids.
// out[0] = max row length, out[1] = flags (1: decreasing row start, 2: column
// id out of [0, ncols) -- both errors; 4: some row is empty).
__global__ __launch_bounds__(kBlock) void k_scan_csr(const int32_t *__restrict__ rs, int nrows,
                                                     const int32_t *__restrict__ ci, int nnz, int ncols,
                                                     int32_t *__restrict__ out) {
  __shared__ int32_t smax, sbad;
  if (threadIdx.x == 0) { smax = 0; sbad = 0; }
  __syncthreads();
  int lmax = 0, bad = 0;
  const int stride = gridDim.x * kBlock;
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < nrows; r += stride) {
    const int len = rs[r + 1] - rs[r];
    if (len < 0) bad |= 1;
    if (len == 0) bad |= 4;
    lmax = max(lmax, len);
  }
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < nnz; i += stride) {
    const int c = ci[i];
    if (c < 0 || c >= ncols) bad |= 2;
  }
  atomicMax(&smax, lmax);
  atomicOr(&sbad, bad);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicMax(&out[0], smax);
    atomicOr(&out[1], sbad);
  }
}

// Tiled block order.  Rows run in first-encounter order, so on a structured
// mesh the rows sharing a cache line of values sit in blocks a mesh layer
// apart (D blocks): a line is fetched again when its second block runs after
// it left L2.  Walking each XCD's run of blocks in tiles of kTileL layers x
// kTileP blocks keeps both blocks of such a pair resident together.  D comes
// from the list itself: the most common block distance between the blocks of
// adjacent columns.
constexpr int kTileDMin = 8, kTileDMax = 4096, kTileL = 16, kTileP = 16;

// blk[col] = block (of partition `part`) holding the entry at column col.
__global__ __launch_bounds__(kBlock) void k_col_block(const int4 *__restrict__ part, const int32_t *__restrict__ ids,
                                                      int32_t *__restrict__ blk) {
  const int b = blockIdx.x;
  const int g0 = part[b].y, g1 = part[b + 1].y;
  for (int g = g0 + (int)threadIdx.x; g < g1; g += kBlock) blk[ids[g]] = b;
}

// hist[d] += #columns c with blk[c+1] - blk[c] = d, for d in [kTileDMin, kTileDMax).
__global__ __launch_bounds__(kBlock) void k_block_gaps(const int32_t *__restrict__ blk, int ncols,
                                                       uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[kTileDMax];
  for (int i = threadIdx.x; i < kTileDMax; i += kBlock) h[i] = 0;
  __syncthreads();
  for (int c = blockIdx.x * kBlock + threadIdx.x; c + 1 < ncols; c += gridDim.x * kBlock) {
    const int a = blk[c], z = blk[c + 1];
    const int d = z - a;
    if (a >= 0 && z >= 0 && d >= kTileDMin && d < kTileDMax) atomicAdd(&h[d], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kTileDMax; i += kBlock)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Row partition: entry b = first row r with key(r) >= b*W, where key(r) =
// rowStartsX[r] (X = the list the partition is cut on), plus r when
// rowsInKey (merge path).  Entry nblocks = nrows.  Entries carry both lists'
// starts.  Cut on entries alone, a list without empty rows gets blocks of
// about W entries and never more rows than entries; a list with empty rows
// needs the merge path, which bounds a block's rows and entries together.
__global__ __launch_bounds__(kBlock) void k_partition(const int32_t *__restrict__ rsT, const int32_t *__restrict__ rsN,
                                                      int cutOnN, int rowsInKey, int nrows, int W, int nblocks,
                                                      int4 *__restrict__ part) {
  const int b = blockIdx.x * kBlock + threadIdx.x;
  if (b > nblocks) return;
  const int32_t *__restrict__ rs = cutOnN ? rsN : rsT;
  int row;
  if (b == nblocks) {
    row = nrows;
  } else {
    const int64_t target = (int64_t)b * W;
    int lo = 0, hi = nrows;  // answer in [0, nrows]
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if ((int64_t)(rs[mid] - rs[0]) + (rowsInKey ? mid : 0) < target) lo = mid + 1;
      else hi = mid;
    }
    row = lo;
  }
  part[b] = make_int4(row, rsT[row], rsN[row], 0);
}

// Per-block maxima of a partition: out[0] = max T nonzeros, out[1] = max N
// nonzeros, out[2] = max rows.
__global__ __launch_bounds__(kBlock) void k_part_max(const int4 *__restrict__ part, int nblocks,
                                                     int32_t *__restrict__ out) {
  __shared__ int32_t s[3];
  if (threadIdx.x < 3) s[threadIdx.x] = 0;
  __syncthreads();
  int mT = 0, mN = 0, mR = 0;
  for (int b = blockIdx.x * kBlock + threadIdx.x; b < nblocks; b += gridDim.x * kBlock) {
    const int4 p0 = part[b], p1 = part[b + 1];
    mT = max(mT, p1.y - p0.y);
    mN = max(mN, p1.z - p0.z);
    mR = max(mR, p1.x - p0.x);
  }
  atomicMax(&s[0], mT);
  atomicMax(&s[1], mN);
  atomicMax(&s[2], mR);
  __syncthreads();
  if (threadIdx.x < 3) atomicMax(&out[threadIdx.x], s[threadIdx.x]);
}

// Column-sorted pack of one list over one partition: one workgroup per
// block sorts the block's (column, CSR position) pairs by column (bitonic
// network in LDS), splits the sorted columns into at most kPackSegs runs
// whose offsets from the run's first column fit kPackOffBits, and writes the
// packed words at the block's list positions, each marked with whether its
// entry is the first of its row.  A block needing more runs
// keeps base.x = -1 and is read in CSR order.
__global__ __launch_bounds__(kBlock) void k_build_pack(const int4 *__restrict__ part, int nrows,
                                                       const int32_t *__restrict__ starts,
                                                       const int32_t *__restrict__ ids, int useN, int stride,
                                                       uint32_t *__restrict__ pk, int4 *__restrict__ base) {
  __shared__ int32_t key[kCap];
  __shared__ int16_t val[kCap];
  __shared__ uint8_t head[kCap];
  __shared__ int32_t segStart[kPackSegs], segBase[kPackSegs];
  __shared__ int32_t nseg, fail;
  const int b = blockIdx.x, tid = threadIdx.x;
  const int4 p0 = part[b], p1 = part[b + 1];
  const int r1 = min(p1.x, nrows);
  int n = 0, g0 = 0;
  if (p0.x < r1) {
    g0 = useN ? p0.z : p0.y;
    n = ((r1 == p1.x) ? (useN ? p1.z : p1.y) : starts[r1]) - g0;
  }
  uint32_t *out = pk + (size_t)b * stride;
  if (n <= 0 || n > kCap || n > stride) {
    if (tid == 0) base[b] = make_int4(-1, 0, 0, 0);
    return;
  }
  int n2 = 1;
  while (n2 < n) n2 <<= 1;
  for (int t = tid; t < n2; t += kBlock) {
    key[t] = t < n ? ids[g0 + t] : INT_MAX;
    val[t] = (int16_t)t;
    head[t] = 0;
  }
  __syncthreads();
  for (int r = p0.x + tid; r < r1; r += kBlock) {  // first entry of each non-empty row
    const int a = starts[r] - g0;
    if (starts[r + 1] - g0 > a) head[a] = 1;
  }
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < n2; t += kBlock) {
        const int u = t ^ j;
        if (u > t) {
          const bool up = (t & k) == 0;
          const int a = key[t], c = key[u];
          if ((a > c) == up) {
            key[t] = c;
            key[u] = a;
            const int16_t va = val[t];
            val[t] = val[u];
            val[u] = va;
          }
        }
      }
      __syncthreads();
    }
  }
  // Segments: greedy runs of sorted columns spanning at most kPackOffMask.
  // Segment s+1 starts at the first column past segBase[s] + kPackOffMask
  // (a block-wide min over the sorted keys); at most kPackSegs of them.
  __shared__ int32_t nextStart;
  if (tid == 0) {
    segStart[0] = 0;
    segBase[0] = key[0];
    nseg = 1;
    fail = 0;
  }
  __syncthreads();
  for (int sgi = 0; sgi < kPackSegs; ++sgi) {
    if (tid == 0) nextStart = n;
    __syncthreads();
    const int a = segStart[sgi];
    const uint32_t lim = (uint32_t)segBase[sgi];
    for (int t = a + tid; t < n; t += kBlock)
      if ((uint32_t)key[t] - lim > kPackOffMask) {
        atomicMin(&nextStart, t);
        break;  // keys ascend: this lane's later t are larger
      }
    __syncthreads();
    const int ns = nextStart;
    if (ns >= n) break;
    if (tid == 0) {
      if (sgi + 1 == kPackSegs) {
        fail = 1;
      } else {
        segStart[sgi + 1] = ns;
        segBase[sgi + 1] = key[ns];
        nseg = sgi + 2;
      }
    }
    __syncthreads();
    if (fail) break;
  }
  __syncthreads();
  if (fail) {  // CSR order: the block's column ids
    for (int t = tid; t < n; t += kBlock) out[t] = (uint32_t)ids[g0 + t];
    if (tid == 0) base[b] = make_int4(-1, 0, 0, 0);
    return;
  }
  for (int t = tid; t < n; t += kBlock) {
    int sg = 0;
    for (int s = 1; s < nseg; ++s) sg += (t >= segStart[s]);
    out[t] = ((uint32_t)sg << kPackSegShift) | ((uint32_t)(key[t] - segBase[sg]) << kPackOffShift) |
             ((uint32_t)head[val[t]] << kPackHeadBit) | (uint32_t)val[t];
  }
  if (tid == 0) {
    int bs[kPackSegs] = {0, 0, 0, 0};
    for (int s = 0; s < nseg; ++s) bs[s] = segBase[s];
    base[b] = make_int4(bs[0], bs[1], bs[2], bs[3]);
  }
}

// Count of packed blocks (base.x >= 0) of a pack.
__global__ __launch_bounds__(kBlock) void k_count_packed(const int4 *__restrict__ base, int nblocks,
                                                         int32_t *__restrict__ out) {
  int c = 0;
  for (int b = blockIdx.x * kBlock + threadIdx.x; b < nblocks; b += gridDim.x * kBlock) c += base[b].x >= 0;
  atomicAdd(out, c);
}

// Delta-pack builders, one workgroup per block of partition 0, from the
// block's 32-bit pack words (column order).  Column of sorted entry i:
__device__ __forceinline__ int pack_col(const uint32_t *__restrict__ pk, const int4 &bb, int stride, int b, int i) {
  return slot_col(pk[(size_t)b * stride + i], bb, true);
}

// Code of each delta below 256 under table tb (15: not in the table).
__device__ __forceinline__ void delta_codes(const uint4 &tb, uint8_t *inv) {
  const int t = threadIdx.x;  // kBlock == 256 deltas
  inv[t] = t == 0 ? 0 : (uint8_t)kDEscape;
  __syncthreads();
  if (t >= 1 && t < (int)kDEscape) {
    const uint32_t d = delta_of((uint32_t)t, tb);
    if (d > 0 && d < 256) inv[d] = (uint8_t)t;
  }
  __syncthreads();
}

// Histogram of the deltas in [1, 256) between consecutive entries of a wave
// range, over the packed blocks (the candidates of the delta table).
__global__ __launch_bounds__(kBlock) void k_delta_hist(const int4 *__restrict__ part, const uint32_t *__restrict__ pk,
                                                       const int4 *__restrict__ base, int stride,
                                                       uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  const int b = blockIdx.x, tid = threadIdx.x;
  h[tid] = 0u;
  __syncthreads();
  const int4 bb = base[b];
  if (bb.x >= 0) {
    const int n = part[b + 1].y - part[b].y;
    for (int i = tid; i < n; i += kBlock) {
      if (i % kDWave == 0) continue;
      const int d = pack_col(pk, bb, stride, b, i) - pack_col(pk, bb, stride, b, i - 1);
      if (d > 0 && d < 256) atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  if (h[tid]) atomicAdd(&hist[tid], h[tid]);
}

// Exceptions of wave range w of block b under table inv (the first entry of
// a non-empty range counts): run by wave w, the count in every lane.
__device__ __forceinline__ int delta_escapes(const uint32_t *__restrict__ pk, const int4 &bb, int stride, int b,
                                             int n, const uint8_t *inv) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int cnt = 0;
  for (int j = 0; j < kDWave / 64; ++j) {
    const int i = w * kDWave + j * 64 + lane;
    bool esc = false;
    if (i < n) {
      if (i % kDWave == 0) {
        esc = true;
      } else {
        const int d = pack_col(pk, bb, stride, b, i) - pack_col(pk, bb, stride, b, i - 1);
        esc = d < 0 || d >= 256 || inv[d] == kDEscape;
      }
    }
    cnt += __builtin_popcountll(__builtin_amdgcn_ballot_w64(esc));
  }
  return cnt;
}

// hist[min(m, kDMaxE + 1)] += 1 per packed block, m = the block's most
// exceptions in one range.
__global__ __launch_bounds__(kBlock) void k_delta_count(const int4 *__restrict__ part, const uint32_t *__restrict__ pk,
                                                        const int4 *__restrict__ base, int stride, uint4 tb,
                                                        uint32_t *__restrict__ hist) {
  __shared__ uint8_t inv[256];
  __shared__ int most;
  const int b = blockIdx.x;
  delta_codes(tb, inv);
  const int4 bb = base[b];
  if (bb.x < 0) return;
  if (threadIdx.x == 0) most = 0;
  __syncthreads();
  const int n = part[b + 1].y - part[b].y;
  const int c = delta_escapes(pk, bb, stride, b, n, inv);
  if ((threadIdx.x & 63) == 0) atomicMax(&most, c);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&hist[most < kDMaxE + 1 ? most : kDMaxE + 1], 1u);
}

// The delta pack's words and exceptions (dw and dx zeroed beforehand); a
// block that does not fit gets the fallback mark and counts in *fallbacks.
__global__ __launch_bounds__(kBlock) void k_build_delta(const int4 *__restrict__ part, const uint32_t *__restrict__ pk,
                                                        const int4 *__restrict__ base, int stride, uint4 tb, int E,
                                                        uint16_t *__restrict__ dw, uint32_t *__restrict__ dx,
                                                        int32_t *__restrict__ fallbacks) {
  __shared__ uint8_t inv[256];
  __shared__ int over;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  delta_codes(tb, inv);
  if (tid == 0) over = 0;
  __syncthreads();
  const int4 bb = base[b];
  const int n = part[b + 1].y - part[b].y;
  if (bb.x >= 0 && delta_escapes(pk, bb, stride, b, n, inv) > E && lane == 0) over = 1;
  __syncthreads();
  uint32_t *xw = dx + ((size_t)b * kDRanges + w) * E;
  if (bb.x < 0 || over) {
    if (lane == 0) xw[0] = kDFallback;
    if (lane == 1) xw[1] = (uint32_t)b;  // its words' slot: b in the partition's pack
    if (tid == 0) atomicAdd(fallbacks, 1);
    return;
  }
  int run = 0;
  for (int j = 0; j < kDWave / 64; ++j) {
    const int i = w * kDWave + j * 64 + lane;
    uint32_t word = 0u, ev = 0u;
    bool esc = false;
    if (i < n) {
      const uint32_t pw = pk[(size_t)b * stride + i];
      const int c = slot_col(pw, bb, true);
      uint32_t code = kDEscape;
      if (i % kDWave == 0) {
        ev = (uint32_t)c;
      } else {
        const int d = c - pack_col(pk, bb, stride, b, i - 1);
        if (d >= 0 && d < 256 && inv[d] != kDEscape) code = inv[d];
        else ev = (uint32_t)d;
      }
      esc = code == kDEscape;
      word = (pw & ((1u << (kPackHeadBit + 1)) - 1u)) | (code << 12);  // position, head bit, code
    }
    const uint64_t mk = __builtin_amdgcn_ballot_w64(esc);
    if (esc) xw[run + lanes_below(mk)] = ev;
    run += __builtin_popcountll(mk);
    dw[(size_t)b * kCap + w * kDWave + lane * (kDWave / 64) + j] = (uint16_t)word;
  }
}

// The fallback blocks' 32-bit words in a pack of their own: each fallback
// block of the delta pack takes the next slot (written as every range's
// exception word 1), its words and segment bases copied there.
__global__ __launch_bounds__(kBlock) void k_compact_fallback(uint32_t *__restrict__ dx, int E,
                                                             const uint32_t *__restrict__ pk,
                                                             const int4 *__restrict__ base, int stride,
                                                             uint32_t *__restrict__ fpk, int4 *__restrict__ fbase,
                                                             int32_t *__restrict__ count) {
  __shared__ int slot;
  const int b = blockIdx.x;
  if (dx[(size_t)b * kDRanges * E] != kDFallback) return;
  if (threadIdx.x == 0) slot = atomicAdd(count, 1);
  __syncthreads();
  const int sl = slot;
  if (threadIdx.x < kDRanges) dx[((size_t)b * kDRanges + threadIdx.x) * E + 1] = (uint32_t)sl;
  if (threadIdx.x == 0) fbase[sl] = base[b];
  for (int i = threadIdx.x; i < stride; i += kBlock) fpk[(size_t)sl * stride + i] = pk[(size_t)b * stride + i];
}

// Byte-pack builders, one workgroup per block of partition 0, from the
// block's 32-bit pack words.  The triple key of sorted entry i (i > 0 in its
// wave range): (column delta << 12) | (position delta << 1) | head bit, or
// kDFallback where the column delta is not in [0, 2^20).
__device__ __forceinline__ uint32_t byte_key(const uint32_t *__restrict__ pk, const int4 &bb, int stride, int b,
                                             int i) {
  const uint32_t w = pk[(size_t)b * stride + i], v = pk[(size_t)b * stride + i - 1];
  const uint32_t d = (uint32_t)(slot_col(w, bb, true) - slot_col(v, bb, true));
  if (d >= kBColLimit) return kDFallback;
  const uint32_t dp = ((w & kPackPosMask) - (v & kPackPosMask)) & kPackPosMask;
  return (d << 12) | (dp << 1) | ((w >> kPackHeadBit) & 1u);
}

// keys[s * kCap + i]: the keys of sample block s * every (kDFallback where
// there is none), the candidates of the byte table.
__global__ __launch_bounds__(kBlock) void k_byte_keys(const int4 *__restrict__ part, const uint32_t *__restrict__ pk,
                                                      const int4 *__restrict__ base, int stride, int every,
                                                      uint32_t *__restrict__ keys) {
  const int s = blockIdx.x, b = s * every;
  const int4 bb = base[b];
  const int n = part[b + 1].y - part[b].y;
  for (int i = threadIdx.x; i < kCap; i += kBlock)
    keys[(size_t)s * kCap + i] = (bb.x >= 0 && i < n && i % kDWave != 0) ? byte_key(pk, bb, stride, b, i) : kDFallback;
}

// Code of entry i of a block under the sorted table keys sk (code sc[k] for
// key sk[k], nk keys; kBEscape when absent), and the exception item it
// would need.
__device__ __forceinline__ uint32_t byte_code(const uint32_t *__restrict__ pk, const int4 &bb, int stride, int b,
                                              int i, const uint32_t *sk, const uint8_t *sc, int nk, u32x2 &item) {
  const uint32_t w = pk[(size_t)b * stride + i];
  const uint32_t c = (uint32_t)slot_col(w, bb, true);
  uint32_t pc = 0u, pp = 0u;
  if (i % kDWave != 0) {
    const uint32_t v = pk[(size_t)b * stride + i - 1];
    pc = (uint32_t)slot_col(v, bb, true);
    pp = v & kPackPosMask;
  }
  item.x = c - pc;
  item.y = ((((w & kPackPosMask) - pp) & kPackPosMask) << 1) | ((w >> kPackHeadBit) & 1u);
  if (i % kDWave == 0 || item.x >= kBColLimit) return kBEscape;
  const uint32_t key = (item.x << 12) | item.y;
  int lo = 0, hi = nk;  // first sk[k] >= key
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sk[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < nk && sk[lo] == key) ? (uint32_t)sc[lo] : kBEscape;
}

__device__ __forceinline__ void byte_table(const uint32_t *__restrict__ skey, const uint8_t *__restrict__ scode,
                                           uint32_t *sk, uint8_t *sc) {
  sk[threadIdx.x] = skey[threadIdx.x];
  sc[threadIdx.x] = scode[threadIdx.x];
  __syncthreads();
}

// cnt[4 b + w]: the exception items of wave range w of block b, or 0 for
// every range of a block that does not fit (left in CSR order, or a range
// of more than kBMaxItems items).
__global__ __launch_bounds__(kBlock) void k_byte_count(const int4 *__restrict__ part, const uint32_t *__restrict__ pk,
                                                       const int4 *__restrict__ base, int stride,
                                                       const uint32_t *__restrict__ skey,
                                                       const uint8_t *__restrict__ scode, int nk,
                                                       uint32_t *__restrict__ cnt) {
  __shared__ uint32_t sk[256];
  __shared__ uint8_t sc[256];
  __shared__ int over;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  byte_table(skey, scode, sk, sc);
  if (tid == 0) over = 0;
  const int4 bb = base[b];
  const int n = part[b + 1].y - part[b].y;
  int c = 0;
  if (bb.x >= 0) {
    for (int j = 0; j < kDWave / 64; ++j) {
      const int i = w * kDWave + j * 64 + lane;
      u32x2 item;
      const bool esc = i < n && byte_code(pk, bb, stride, b, i, sk, sc, nk, item) == kBEscape;
      c += __builtin_popcountll(__builtin_amdgcn_ballot_w64(esc));
    }
  }
  __syncthreads();
  if (lane == 0 && (bb.x < 0 || c > kBMaxItems)) over = 1;
  __syncthreads();
  if (lane == 0) cnt[(size_t)b * kDRanges + w] = over ? 0u : (uint32_t)c;
}

// The byte pack's codes, headers and exception items (off: each range's
// first item, the exclusive sum of k_byte_count's counts).
__global__ __launch_bounds__(kBlock) void k_byte_build(const int4 *__restrict__ part, const uint32_t *__restrict__ pk,
                                                       const int4 *__restrict__ base, int stride,
                                                       const uint32_t *__restrict__ skey,
                                                       const uint8_t *__restrict__ scode, int nk,
                                                       const uint32_t *__restrict__ cnt,
                                                       const uint32_t *__restrict__ off, uint8_t *__restrict__ bw,
                                                       uint32_t *__restrict__ bh, u32x2 *__restrict__ bx) {
  __shared__ uint32_t sk[256];
  __shared__ uint8_t sc[256];
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  byte_table(skey, scode, sk, sc);
  const int4 bb = base[b];
  const int n = part[b + 1].y - part[b].y;
  bool fits = bb.x >= 0;
  for (int r = 0; r < kDRanges && fits; ++r) fits = cnt[(size_t)b * kDRanges + r] > 0u || n <= r * kDWave;
  if (!fits) {
    if (tid < kDRanges) bh[(size_t)b * kDRanges + tid] = tid == 0 ? kDFallback : 0u;
    return;
  }
  const uint32_t o = off[(size_t)b * kDRanges + w];
  if (lane == 0) bh[(size_t)b * kDRanges + w] = (o << 7) | cnt[(size_t)b * kDRanges + w];
  int run = 0;
  for (int j = 0; j < kDWave / 64; ++j) {
    const int i = w * kDWave + j * 64 + lane;
    uint32_t code = 0u;
    u32x2 item = {0u, 0u};
    if (i < n) code = byte_code(pk, bb, stride, b, i, sk, sc, nk, item);
    const bool esc = i < n && code == kBEscape;
    const uint64_t mk = __builtin_amdgcn_ballot_w64(esc);
    if (esc) bx[o + run + lanes_below(mk)] = item;
    run += __builtin_popcountll(mk);
    bw[(size_t)b * kCap + w * kDWave + lane * (kDWave / 64) + j] = (uint8_t)code;
  }
}

// ---------------------------------------------------------------------------
// Launch helpers.

static int grid_for(int64_t n, int cap = 65536) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// Choose the partition and field chunk for a launch touching lists with
// per-block maxima read by `maxOf`: prefer one pass over all K fields with
// the widest blocks; otherwise the most fields per pass.
template <class MaxOf>
static const Partition *choose(const Partition *parts, int nparts, int K, MaxOf maxOf, int *kc) {
  const Partition *best = nullptr;
  int bestKc = 0;
  for (int j = 0; j < nparts; ++j) {
    const int m = maxOf(parts[j]);
    const int fit = (m <= 0) ? K : kCap / m;
    const int kcj = fit < K ? fit : K;
    if (kcj <= 0) continue;
    if (kcj == K) { *kc = K; return &parts[j]; }
    if (kcj > bestKc) { bestKc = kcj; best = &parts[j]; }
  }
  *kc = bestKc;
  return best;
}

// The pack of a list over a partition, when the launch uses partition 0.
static ListView list_view(const Plan *p, bool listN, const Partition *pt, bool partIsN) {
  ListView L;
  L.starts = listN ? p->rsN : p->rsT;
  L.ids = listN ? p->ciN : p->ciT;
  L.pk = nullptr;
  L.base = nullptr;
  const Pack *pk = nullptr;
  if (pt == &p->partS[0]) {
    if (!listN) pk = &p->packS;
  } else if (pt == &p->part3[0] || pt == &p->part3[1]) {
    if (!listN || p->aliased) pk = &p->pack3[pt - p->part3];
  } else if (partIsN) {
    if (pt == &p->partN[0] && listN) pk = &p->packNN;
  } else {
    const int j = (int)(pt - p->partT);
    if (j >= 0 && j < kPackParts && j < p->nparts) {
      if (!listN) pk = &p->packT[j];
      else if (j == 0) pk = &p->packNT;
    }
  }
  L.stride = 0;
  if (pk && pk->pk) {
    L.pk = pk->pk;
    L.base = pk->base;
    L.stride = pk->stride;
  }
  return L;
}

// Field counts with a compile-time shape (k_gs_sym_vec, k_batch_vec).
__host__ __device__ constexpr bool vec_k(int K) { return K == 1 || packed_k(K); }

template <typename T>
static int vec_cap_entries(int K) {
  switch (K) {
    case 1: return VecShape<T, 1>::CAP;
    case 2: return VecShape<T, 2>::CAP;
    case 3: return VecShape<T, 3>::CAP;
    case 4: return VecShape<T, 4>::CAP;
    case 8: return VecShape<T, 8>::CAP;
    default: return VecShape<T, 16>::CAP;
  }
}

// Workgroups per CU a launch needs to fill the device: on small meshes each
// step to narrower blocks up to this count measured faster (fewer entries per
// block, shorter dependent chains per CU), past it wider blocks were.
constexpr int kFillPerCu = 4;

// From T partition `vp`, the narrower partitions with packs, step by step,
// until the launch has kFillPerCu workgroups per CU.
static const Partition *fill_device(const Plan *p, const Partition *vp) {
  const Partition *end = p->partT + (p->nparts < kPackParts ? p->nparts : kPackParts);
  while (vp->nblocks < kFillPerCu * p->cus && vp + 1 < end) ++vp;
  return vp;
}

// The widest T partition whose blocks hold at most the shape's entries, made
// `narrow` steps narrower when there is one.
template <typename T>
static const Partition *vec_partition(const Plan *p, int K, int narrow) {
  const int cap = vec_cap_entries<T>(K);
  for (int j = 0; j < p->nparts; ++j) {
    if (p->partT[j].maxT <= cap) {
      const int jj = j + narrow < p->nparts ? j + narrow : p->nparts - 1;
      return &p->partT[jj];
    }
  }
  return nullptr;
}

static_assert(VecShape<float, 3>::CAP == kCap3[0] && VecShape<int32_t, 3>::CAP == kCap3[0] &&
                  VecShape<f32c, 3>::CAP == kCap3[0] && VecShape<bf16s, 3>::CAP == kCap3[0] &&
                  VecShape<double, 3>::CAP == kCap3[1] && VecShape<int64_t, 3>::CAP == kCap3[1],
              "three-field partitions match the three-field shapes");

// The three-field partition of the shape's capacity, when K = 3 and the plan
// built one (a T list cut on entries).
template <typename T>
static const Partition *vec_partition3(const Plan *p, int K) {
  if (K != 3) return nullptr;
  const int cap = vec_cap_entries<T>(3);
  const Partition *p3 = &p->part3[cap == kCap3[0] ? 0 : 1];
  return (p3->nblocks > 0 && p3->maxT <= cap) ? p3 : nullptr;
}

// The entry-cut T partition of a one-list entry-major launch with K fields:
// the three-field partition when K = 3 and it fills the device, else the
// widest power-of-two partition that fits the shape, narrowed to fill the
// device on small lists.
template <typename T>
static const Partition *vec_launch_partition(const Plan *p, int K) {
  const Partition *vp = vec_partition3<T>(p, K);
  if (vp && vp->nblocks >= kFillPerCu * p->cus) return vp;
  vp = vec_partition<T>(p, K, 0);
  return vp ? fill_device(p, vp) : nullptr;
}

template <typename T, int OP, int KT>
static void launch_vec_k(const Partition *pt, const ListView &G, int gN, T *v, hipStream_t st) {
  hipLaunchKernelGGL((k_gs_sym_vec<T, OP, KT>), dim3(pt->nblocks), dim3(kBlock), 0, st, pt->d, G.starts, G.ids, G.pk,
                     G.base, G.stride, gN, v, (pt->tiled & kTiledGS) ? pt->order : nullptr);
}

template <typename T, int OP>
static void launch_vec_op(int K, const Partition *pt, const ListView &G, int gN, T *v, hipStream_t st) {
  switch (K) {
    case 1: launch_vec_k<T, OP, 1>(pt, G, gN, v, st); break;
    case 2: launch_vec_k<T, OP, 2>(pt, G, gN, v, st); break;
    case 3: launch_vec_k<T, OP, 3>(pt, G, gN, v, st); break;
    case 4: launch_vec_k<T, OP, 4>(pt, G, gN, v, st); break;
    case 8: launch_vec_k<T, OP, 8>(pt, G, gN, v, st); break;
    default: launch_vec_k<T, OP, 16>(pt, G, gN, v, st); break;
  }
}

template <typename T>
static int launch_vec(int op, int K, const Partition *pt, const ListView &G, int gN, T *v, hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_vec_op<T, OGS_ADD>(K, pt, G, gN, v, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_vec_op<T, OGS_MUL>(K, pt, G, gN, v, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_vec_op<T, OGS_MAX>(K, pt, G, gN, v, st); break;
    case OGS_MIN: launch_vec_op<T, OGS_MIN>(K, pt, G, gN, v, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

// Resident workgroups per CU of a kernel (queried once per instance).
template <typename F>
static int occupancy_of(F kernel) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(kernel), kBlock, 0) !=
          hipSuccess ||
      n < 1)
    n = 1;
  return n;
}

// The widest entry-cut T partition whose blocks hold at most half the shape's
// entries of either list (the two-list kernel's capacity).
template <typename T>
static const Partition *vec2_partition(const Plan *p, int K) {
  const int cap = vec_cap_entries<T>(K) / 2;
  for (int j = 0; j < p->nparts; ++j) {
    const int m = p->partT[j].maxT > p->partT[j].maxN ? p->partT[j].maxT : p->partT[j].maxN;
    if (m <= cap) return &p->partT[j];
  }
  return nullptr;
}

template <typename T, int OP>
static void launch_gs_trans_mask_op(int K, const Partition *vp, const ListView &G, const Plan *p, T *v,
                                    hipStream_t st) {
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL((k_gs_trans_mask<T, OP, KT>), dim3(vp->nblocks), dim3(kBlock), 0, st, vp->d, p->rsT,
                       p->ciT, G.pk, G.base, G.stride, v, p->nMask);
  };
  switch (K) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
  }
}

template <typename T>
static int launch_gs_trans_mask(int op, int K, const Partition *vp, const ListView &G, const Plan *p, T *v,
                                hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_gs_trans_mask_op<T, OGS_ADD>(K, vp, G, p, v, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_gs_trans_mask_op<T, OGS_MUL>(K, vp, G, p, v, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_gs_trans_mask_op<T, OGS_MAX>(K, vp, G, p, v, st); break;
    case OGS_MIN: launch_gs_trans_mask_op<T, OGS_MIN>(K, vp, G, p, v, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

template <typename T, int OP>
static int launch_gs2(int K, const Partition *vp, const ListView &G, int gN, const ListView &S, int sN, T *v,
                      hipStream_t st) {
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL((k_gs2_vec<T, OP, KT>), dim3(vp->nblocks), dim3(kBlock), 0, st, vp->d, G, gN, S, sN, v);
  };
  switch (K) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
  }
  return (int)hipGetLastError();
}

// Largest field count of the by-column GatherScatter NoTrans.
constexpr int kOwnerMaxK = 16;

template <typename T, int OP>
static void launch_gs_owner_op(int K, const Plan *p, T *v, hipStream_t st) {
  const int32_t cpb = K <= kColVals ? kColVals / K : 1;
  const dim3 grid((unsigned)(((int64_t)p->Ncols + cpb - 1) / cpb)), block(kBlock);
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL((k_gs_owner<T, OP, KT>), grid, block, 0, st, p->Ncols, K, cpb, p->colOwner, v);
  };
  switch (K) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    default: go(std::integral_constant<int, 0>{}); break;
  }
}

template <typename T>
static int launch_gs_owner(int op, int K, const Plan *p, T *v, hipStream_t st) {
  if (p->Ncols == 0) return OGS_SUCCESS;
  switch (op) {
    case OGS_ADD: launch_gs_owner_op<T, OGS_ADD>(K, p, v, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_gs_owner_op<T, OGS_MUL>(K, p, v, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_gs_owner_op<T, OGS_MAX>(K, p, v, st); break;
    case OGS_MIN: launch_gs_owner_op<T, OGS_MIN>(K, p, v, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

// The K = 1 one-list GatherScatter from a delta pack over partition pt
// (CSR (rs, ci) of the list pt indexes: read only by blocks left in CSR
// order, which a bundle has none of).
template <typename T, int OP>
static void launch_delta_on(const Partition &pt, const DeltaPack &d, const int32_t *rs, const int32_t *ci,
                            int ncols, T *v, const int *order, hipStream_t st) {
  const uint4 tb = make_uint4(d.tbl[0], d.tbl[1], d.tbl[2], d.tbl[3]);
  const u32x4 *dw = reinterpret_cast<const u32x4 *>(d.w);
  const dim3 grid(pt.nblocks), block(kBlock);
  auto byte = [&](auto mode) {
    hipLaunchKernelGGL((k_gs_sym_byte<T, OP, decltype(mode)::value>), grid, block, 0, st, pt.d, rs, ci, d.f.pk,
                       d.f.base, d.f.stride, dw, d.x, d.E, tb, ncols, v, order,
                       reinterpret_cast<const u32x2 *>(d.bw), reinterpret_cast<const uint4 *>(d.bh),
                       reinterpret_cast<const u32x2 *>(d.bx), d.bt);
  };
  if constexpr (kSeg8<T>) {
    if (d.seg8) {  // rows of at most 8 members
      if (d.bw) byte(std::integral_constant<int, kSeg>{});
      else
        hipLaunchKernelGGL((k_gs_sym_delta<T, OP, kSeg>), grid, block, 0, st, pt.d, rs, ci, d.f.pk, d.f.base,
                           d.f.stride, dw, d.x, d.E, tb, ncols, v, order);
      return;
    }
  }
  if (d.bw && !d.rows) byte(std::integral_constant<int, kPlain>{});
  else if (d.rows)
    hipLaunchKernelGGL((k_gs_sym_delta_rows<T, OP>), grid, block, 0, st, pt.d, rs, ci, d.f.pk, d.f.base, d.f.stride,
                       dw, d.x, d.E, tb, ncols, v, order);
  else
    hipLaunchKernelGGL((k_gs_sym_delta<T, OP, kPlain>), grid, block, 0, st, pt.d, rs, ci, d.f.pk, d.f.base,
                       d.f.stride, dw, d.x, d.E, tb, ncols, v, order);
}

// ... over the plan's bundled blocks where it has them, else partition 0.
template <typename T, int OP>
static void launch_delta_op(const Plan *p, T *v, hipStream_t st) {
  const Bundle &u = p->bundle;
  if (u.part.d) {
    launch_delta_on<T, OP>(u.part, u.dp, nullptr, nullptr, p->Ncols, v, (u.part.tiled & kTiledGS) ? u.part.order : nullptr,
                           st);
    return;
  }
  const Partition &pt = p->partT[0];
  launch_delta_on<T, OP>(pt, p->dpack, p->rsT, p->ciT, p->Ncols, v, (pt.tiled & kTiledGS) ? pt.order : nullptr, st);
}

template <typename T>
static int launch_delta(int op, const Plan *p, T *v, hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_delta_op<T, OGS_ADD>(p, v, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_delta_op<T, OGS_MUL>(p, v, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_delta_op<T, OGS_MAX>(p, v, st); break;
    case OGS_MIN: launch_delta_op<T, OGS_MIN>(p, v, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

// The one-list GatherScatter with K fields over the byte codes (the
// bundled blocks where the plan has them, else partition 0): for lists whose
// rows have at most 8 members, with no block left in CSR order, and blocks
// enough to fill the device (else null).
static const DeltaPack *byte_kv_pack(const Plan *p, const Partition **pt) {
  const bool bun = p->bundle.part.d != nullptr;
  const Partition &x = bun ? p->bundle.part : p->partT[0];
  const DeltaPack &d = bun ? p->bundle.dp : p->dpack;
  if (!d.w || !d.bw || !d.seg8 || d.csr > 0 || x.nblocks < kFillPerCu * p->cus) return nullptr;
  *pt = &x;
  return &d;
}

// Field counts of the byte-coded K-field kernels: up to 24 bytes of fields
// per entry (against the entry-major kernels at 48^3 p=7: fp64 K = 2
// GatherScatter 389 -> 355 us, Gather 304 -> 295 us; float K = 4 389 ->
// 355 us; fp64 K = 4, 32 bytes, 722 -> 736 us and Gather 570 -> 648 us, not
// taken).
template <typename T>
constexpr bool byte_k(int K) { return K == 2 || K == 3 || (K == 4 && sizeof(T) <= 4); }

template <typename T, int OP, int KT>
static void launch_byte_kv_k(const Partition &pt, const DeltaPack &d, int ncols, T *v, hipStream_t st) {
  const uint4 tb = make_uint4(d.tbl[0], d.tbl[1], d.tbl[2], d.tbl[3]);
  hipLaunchKernelGGL((k_gs_sym_byte_kv<T, OP, KT>), dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, d.f.pk, d.f.base,
                     d.f.stride, reinterpret_cast<const u32x4 *>(d.w), d.x, d.E, tb, ncols, v,
                     (pt.tiled & kTiledGS) ? pt.order : nullptr, reinterpret_cast<const u32x2 *>(d.bw),
                     reinterpret_cast<const uint4 *>(d.bh), reinterpret_cast<const u32x2 *>(d.bx), d.bt);
}

template <typename T, int OP>
static void launch_byte_kv_op(int K, const Partition &pt, const DeltaPack &d, int ncols, T *v, hipStream_t st) {
  if (K == 2) launch_byte_kv_k<T, OP, 2>(pt, d, ncols, v, st);
  else if (K == 3) launch_byte_kv_k<T, OP, 3>(pt, d, ncols, v, st);
  else if constexpr (sizeof(T) <= 4) launch_byte_kv_k<T, OP, 4>(pt, d, ncols, v, st);
}

template <typename T>
static int launch_byte_kv(int op, int K, const Partition &pt, const DeltaPack &d, int ncols, T *v, hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_byte_kv_op<T, OGS_ADD>(K, pt, d, ncols, v, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_byte_kv_op<T, OGS_MUL>(K, pt, d, ncols, v, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_byte_kv_op<T, OGS_MAX>(K, pt, d, ncols, v, st); break;
    case OGS_MIN: launch_byte_kv_op<T, OGS_MIN>(K, pt, d, ncols, v, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

template <typename T>
static int launch_gs_typed(const Plan *p, int op, int trans, int K, void *d_v, hipStream_t st) {
  // list selection: ogsOperator.cpp:321-339
  int nrows;
  bool gListN, sListN;
  if (trans == OGS_TRANS) {
    nrows = p->NrowsN; gListN = false; sListN = true;
  } else if (trans == OGS_SYM) {
    nrows = p->NrowsT; gListN = false; sListN = false;
  } else {
    nrows = p->NrowsT; gListN = true; sListN = false;
  }
  if (nrows == 0 || p->nparts == 0) return OGS_SUCCESS;
  // NoTrans with at most one N member per row: by columns
  if (trans == OGS_NOTRANS && p->colOwner && K <= kOwnerMaxK)
    return launch_gs_owner<T>(op, K, p, static_cast<T *>(d_v), st);
  const int32_t *gS = gListN ? p->rsN : p->rsT, *gI = gListN ? p->ciN : p->ciT;
  const int32_t *sS = sListN ? p->rsN : p->rsT, *sI = sListN ? p->ciN : p->ciT;
  const int gN = gListN ? 1 : 0, sN = sListN ? 1 : 0;
  const bool sym = (gI == sI && gS == sS);
  int kc = 0;
  const Partition *pt = choose(p->partT, p->nparts, K,
                               [&](const Partition &x) { return gN || sN ? (x.maxT > x.maxN ? x.maxT : x.maxN) : x.maxT; }, &kc);
  if (!pt) return OGS_E_ROW_TOO_LONG;
  const dim3 grid(pt->nblocks), block(kBlock);
  T *v = static_cast<T *>(d_v);
  const ListView G = list_view(p, gListN, pt, false), S = list_view(p, sListN, pt, false);
  // one list (rows [0, NrowsT)) and a field count with a compile-time
  // shape: the entry-major kernel on the widest partition whose blocks fit
  // that shape (for K = 1 partition 0 always fits), narrowed to fill the
  // device on small lists.
  if (sym && p->nnzT == 0) return OGS_SUCCESS;  // no entries: nothing to write
  if (sym && vec_k(K)) {
    const Partition *vp = vec_launch_partition<T>(p, K);
    if (vp && K == 1 && vp == &p->partT[0] && p->dpack.w) return launch_delta<T>(op, p, v, st);
    if constexpr (kSeg8<T> || std::is_same<T, f32c>::value) {
      // (float-compensated K = 3 at 48^3 p=7: 314 -> 290 us, 192^3 p=1:
      // 579 -> 328 us; bfloat16, whose values are widened, stays on the
      // entry-major kernel: 253 -> 282 us at 48^3 p=7)
      const Partition *bp = nullptr;
      const DeltaPack *bd = byte_k<T>(K) ? byte_kv_pack(p, &bp) : nullptr;
      if (bd) return launch_byte_kv<T>(op, K, *bp, *bd, p->Ncols, v, st);
    }
    if (vp) return launch_vec<T>(op, K, vp, list_view(p, gListN, vp, false), gN, v, st);
  }
  // Trans over the T list alone, N membership from the column mask: the
  // one-list kernel's blocks and traffic, stores to N members only (48^3
  // p=7 fp64 K = 1: Signed unique 304 -> 236 us, 35 % flagged ids 349 ->
  // 310 us).  NoTrans stays on the kernels above and below: storing to every
  // entry of a flagged operator's T blocks measured slower (437 -> 516 us).
  if (trans == OGS_TRANS && !sym && p->nMask && vec_k(K) && !p->partT[0].rowsInKey && p->nnzT > 0) {
    const Partition *vp = vec_launch_partition<T>(p, K);
    if (vp) return launch_gs_trans_mask<T>(op, K, vp, list_view(p, false, vp, false), p, v, st);
  }
  // two lists, K > 1 with a compile-time shape, T without empty rows: the
  // two-list entry-major kernel (at K = 1 the row-based kernel below, with a
  // whole list's capacity per block, measured 7-13% faster)
  if (!sym && K > 1 && vec_k(K) && !p->partT[0].rowsInKey && p->nnzT > 0) {
    const Partition *vp = vec2_partition<T>(p, K);
    if (vp) {
      vp = fill_device(p, vp);
      const ListView G2 = list_view(p, gListN, vp, false), S2 = list_view(p, sListN, vp, false);
      switch (op) {
        case OGS_ADD: return launch_gs2<T, OGS_ADD>(K, vp, G2, gN, S2, sN, v, st);
        case OGS_MUL:
          if constexpr (kOffered<T, OGS_MUL>) return launch_gs2<T, OGS_MUL>(K, vp, G2, gN, S2, sN, v, st);
          else return OGS_E_NOT_SUPPORTED;
        case OGS_MAX: return launch_gs2<T, OGS_MAX>(K, vp, G2, gN, S2, sN, v, st);
        case OGS_MIN: return launch_gs2<T, OGS_MIN>(K, vp, G2, gN, S2, sN, v, st);
        default: return OGS_E_ARG;
      }
    }
  }
  // two lists at K = 1: partitions 1.. hold at most kCap / 2 entries of
  // either list and rows (checked from their recorded maxima), narrowed to
  // fill the device on small lists
  const Partition *half = (K == 1 && !sym && p->nparts > 1) ? fill_device(p, &p->partT[1]) : nullptr;
  if (half && (half->maxT > kCap / 2 || half->maxN > kCap / 2 || half->maxRows > kCap / 2)) half = nullptr;
#define OGS_GS_LAUNCH(OPV)                                                                                      \
  if (K == 1 && !sym && half) {                                                                                 \
    hipLaunchKernelGGL((k_gs_k1<T, OPV, false, kCap / 2>), dim3(half->nblocks), block, 0, st, half->d, nrows,    \
                       list_view(p, gListN, half, false), gN, list_view(p, sListN, half, false), sN, v);        \
  } else if (K == 1 && !sym) {                                                                                  \
    hipLaunchKernelGGL((k_gs_k1<T, OPV, false>), grid, block, 0, st, pt->d, nrows, G, gN, S, sN, v);            \
  } else {                                                                                                      \
    if (sym) hipLaunchKernelGGL((k_gs_kv<T, OPV, true>), grid, block, 0, st, pt->d, nrows, gS, gI, gN, sS, sI, sN, K, kc, v); \
    else hipLaunchKernelGGL((k_gs_kv<T, OPV, false>), grid, block, 0, st, pt->d, nrows, gS, gI, gN, sS, sI, sN, K, kc, v);   \
  }
  switch (op) {
    case OGS_ADD: OGS_GS_LAUNCH(OGS_ADD) break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) { OGS_GS_LAUNCH(OGS_MUL) } else { return OGS_E_NOT_SUPPORTED; }
      break;
    case OGS_MAX: OGS_GS_LAUNCH(OGS_MAX) break;
    case OGS_MIN: OGS_GS_LAUNCH(OGS_MIN) break;
    default: return OGS_E_ARG;
  }
#undef OGS_GS_LAUNCH
  return (int)hipGetLastError();
}

template <typename T, int OP>
static int launch_gather_vec(int K, const Partition *vp, const ListView &G, const Plan *p, const T *v, T *gv,
                             const int *order, hipStream_t st) {
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL((k_gather_vec<T, OP, KT>), dim3(vp->nblocks), dim3(kBlock), 0, st, vp->d, p->rsT, p->ciT,
                       G.pk, G.base, G.stride, v, gv, order);
  };
  switch (K) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
  }
  return (int)hipGetLastError();
}

template <typename T>
static int launch_scatter_vec(int K, const Partition *vp, const ListView &S, const Plan *p, const T *gv, T *v,
                              hipStream_t st) {
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL((k_scatter_vec<T, KT>), dim3(vp->nblocks), dim3(kBlock), 0, st, vp->d, p->rsT, p->ciT, S.pk,
                       S.base, S.stride, gv, v, nullptr);
  };
  switch (K) {
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
  }
  return (int)hipGetLastError();
}

template <typename T>
static int launch_scatter_cols(int32_t ncols, int K, const int32_t *colRow, const T *gv, T *v, hipStream_t st,
                               const RowMap *rm = nullptr, int32_t nrows = 0) {
  if (ncols == 0) return OGS_SUCCESS;
  if (K == 1 && rm && rm->codes) {
    hipLaunchKernelGGL((k_scatter_rows_delta<T>), dim3(rm->chunks), dim3(kBlock), 0, st, (int64_t)ncols, nrows, colRow,
                       rm->fbRows, rm->codes, rm->x, rm->E, rm->table, gv, v);
    return (int)hipGetLastError();
  }
  const int32_t cpb = K <= kColVals ? kColVals / K : 1;
  const dim3 grid((unsigned)(((int64_t)ncols + cpb - 1) / cpb)), block(kBlock);
  auto go = [&](auto kt) {
    constexpr int KT = decltype(kt)::value;
    hipLaunchKernelGGL((k_scatter_cols<T, KT>), grid, block, 0, st, ncols, K, cpb, colRow, gv, v);
  };
  switch (K) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    default: go(std::integral_constant<int, 0>{}); break;
  }
  return (int)hipGetLastError();
}

// Gather K = 1 from delta pack d over partition pt (CSR rs, ci): from the
// byte codes where the plan has them, staged in groups of 8 for rows of at
// most 8 members.
template <typename T, int OP>
static void launch_gather_delta_on(const Partition &pt, const DeltaPack &d, const int32_t *rs, const int32_t *ci,
                                   int ncols, const T *v, T *gv, const int *order, hipStream_t st) {
  const uint4 tb = make_uint4(d.tbl[0], d.tbl[1], d.tbl[2], d.tbl[3]);
  const u32x4 *dw = reinterpret_cast<const u32x4 *>(d.w);
  const dim3 grid(pt.nblocks), block(kBlock);
  auto go = [&](auto mode) {
    constexpr int M = decltype(mode)::value;
    if (d.bw)
      hipLaunchKernelGGL((k_gather_byte<T, OP, M>), grid, block, 0, st, pt.d, rs, ci, d.f.pk, d.f.base, d.f.stride,
                         dw, d.x, d.E, tb, ncols, v, gv, order, reinterpret_cast<const u32x2 *>(d.bw),
                         reinterpret_cast<const uint4 *>(d.bh), reinterpret_cast<const u32x2 *>(d.bx), d.bt);
    else
      hipLaunchKernelGGL((k_gather_delta<T, OP, M>), grid, block, 0, st, pt.d, rs, ci, d.f.pk, d.f.base, d.f.stride,
                         dw, d.x, d.E, tb, ncols, v, gv, order);
  };
  if constexpr (kSeg8<T>) {
    if (d.seg8) {
      go(std::integral_constant<int, kSeg>{});
      return;
    }
  }
  go(std::integral_constant<int, kPlain>{});
}

// Gather K = 1 of the T list over partition 0 from the delta pack.
template <typename T, int OP>
static void launch_gather_delta_op(const Plan *p, const T *v, T *gv, hipStream_t st) {
  const Partition &pt = p->partT[0];
  launch_gather_delta_on<T, OP>(pt, p->dpack, p->rsT, p->ciT, p->Ncols, v, gv,
                                (pt.tiled & kTiledGather) ? pt.order : nullptr, st);
}

template <typename T>
static int launch_gather_delta(int op, const Plan *p, const T *v, T *gv, hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_gather_delta_op<T, OGS_ADD>(p, v, gv, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_gather_delta_op<T, OGS_MUL>(p, v, gv, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_gather_delta_op<T, OGS_MAX>(p, v, gv, st); break;
    case OGS_MIN: launch_gather_delta_op<T, OGS_MIN>(p, v, gv, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

// Gather with K fields over partition 0's byte codes (rows of at most 8
// members, no block left in CSR order, blocks enough to fill the device).
static bool gather_kv_ready(const Plan *p) {
  const DeltaPack &d = p->dpack;
  return d.w && d.bw && d.seg8 && d.csr == 0 && p->partT[0].nblocks >= kFillPerCu * p->cus;
}

template <typename T, int OP, int KT>
static void launch_gather_kv_k(const Plan *p, const T *v, T *gv, hipStream_t st) {
  const Partition &pt = p->partT[0];
  const DeltaPack &d = p->dpack;
  const uint4 tb = make_uint4(d.tbl[0], d.tbl[1], d.tbl[2], d.tbl[3]);
  hipLaunchKernelGGL((k_gather_byte_kv<T, OP, KT>), dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, d.f.pk, d.f.base,
                     d.f.stride, reinterpret_cast<const u32x4 *>(d.w), d.x, d.E, tb, p->Ncols, v, gv,
                     (pt.tiled & kTiledGather) ? pt.order : nullptr, reinterpret_cast<const u32x2 *>(d.bw),
                     reinterpret_cast<const uint4 *>(d.bh), reinterpret_cast<const u32x2 *>(d.bx), d.bt);
}

template <typename T, int OP>
static void launch_gather_kv_op(int K, const Plan *p, const T *v, T *gv, hipStream_t st) {
  if (K == 2) launch_gather_kv_k<T, OP, 2>(p, v, gv, st);
  else if (K == 3) launch_gather_kv_k<T, OP, 3>(p, v, gv, st);
  else if constexpr (sizeof(T) <= 4) launch_gather_kv_k<T, OP, 4>(p, v, gv, st);
}

template <typename T>
static int launch_gather_kv(int op, int K, const Plan *p, const T *v, T *gv, hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_gather_kv_op<T, OGS_ADD>(K, p, v, gv, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_gather_kv_op<T, OGS_MUL>(K, p, v, gv, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_gather_kv_op<T, OGS_MAX>(K, p, v, gv, st); break;
    case OGS_MIN: launch_gather_kv_op<T, OGS_MIN>(K, p, v, gv, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

template <typename T>
static int launch_gather_typed(const Plan *p, int op, int trans, int K, void *d_gv, const void *d_v, hipStream_t st) {
  // list selection: ogsOperator.cpp:71-79
  const bool useN = (trans == OGS_NOTRANS);
  const int nrows = useN ? p->NrowsN : p->NrowsT;
  if (nrows == 0 || p->nparts == 0) return OGS_SUCCESS;
  // a field count with a compile-time shape from a T list without empty
  // rows: the head-counting kernel on the entry-cut blocks whose blocks fit
  // the shape, narrowed to fill the device on small lists
  if (vec_k(K) && (!useN || p->aliased) && !p->partT[0].rowsInKey && p->nnzT > 0) {
    const Partition *vp = vec_launch_partition<T>(p, K);
    if (vp && K == 1 && vp == &p->partT[0] && p->dpack.w)
      return launch_gather_delta<T>(op, p, static_cast<const T *>(d_v), static_cast<T *>(d_gv), st);
    if constexpr (kSeg8<T> || std::is_same<T, f32c>::value) {
      if (byte_k<T>(K) && gather_kv_ready(p))
        return launch_gather_kv<T>(op, K, p, static_cast<const T *>(d_v), static_cast<T *>(d_gv), st);
    }
    if (vp) {
      const ListView G = list_view(p, false, vp, false);
      const int *order = (K == 1 && (vp->tiled & kTiledGather)) ? vp->order : nullptr;
      switch (op) {
        case OGS_ADD: return launch_gather_vec<T, OGS_ADD>(K, vp, G, p, static_cast<const T *>(d_v),
                                                           static_cast<T *>(d_gv), order, st);
        case OGS_MUL:
          if constexpr (kOffered<T, OGS_MUL>)
            return launch_gather_vec<T, OGS_MUL>(K, vp, G, p, static_cast<const T *>(d_v), static_cast<T *>(d_gv),
                                                 order, st);
          else return OGS_E_NOT_SUPPORTED;
        case OGS_MAX: return launch_gather_vec<T, OGS_MAX>(K, vp, G, p, static_cast<const T *>(d_v),
                                                           static_cast<T *>(d_gv), order, st);
        case OGS_MIN: return launch_gather_vec<T, OGS_MIN>(K, vp, G, p, static_cast<const T *>(d_v),
                                                           static_cast<T *>(d_gv), order, st);
        default: return OGS_E_ARG;
      }
    }
  }
  const bool partN = useN && !p->aliased;
  const Partition *parts = partN ? p->partN : p->partT;
  int kc = 0;
  const Partition *pt = choose(parts, p->nparts, K, [&](const Partition &x) { return useN ? x.maxN : x.maxT; }, &kc);
  if (!pt) return OGS_E_ROW_TOO_LONG;
  const int32_t *gS = useN ? p->rsN : p->rsT;
  const int32_t *gI = useN ? p->ciN : p->ciT;
  const int gN = useN ? 1 : 0;
  const dim3 grid(pt->nblocks), block(kBlock);
  const T *v = static_cast<const T *>(d_v);
  T *gv = static_cast<T *>(d_gv);
  const ListView G = list_view(p, useN && !p->aliased, pt, partN);
#define OGS_G_LAUNCH(OPV)                                                                                          \
  if (K == 1) hipLaunchKernelGGL((k_gather_k1<T, OPV>), grid, block, 0, st, pt->d, nrows, G, gN, v, gv);          \
  else hipLaunchKernelGGL((k_gather<T, OPV>), grid, block, 0, st, pt->d, nrows, gS, gI, gN, K, kc, v, gv);
  switch (op) {
    case OGS_ADD: OGS_G_LAUNCH(OGS_ADD) break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) { OGS_G_LAUNCH(OGS_MUL) } else { return OGS_E_NOT_SUPPORTED; }
      break;
    case OGS_MAX: OGS_G_LAUNCH(OGS_MAX) break;
    case OGS_MIN: OGS_G_LAUNCH(OGS_MIN) break;
    default: return OGS_E_ARG;
  }
#undef OGS_G_LAUNCH
  return (int)hipGetLastError();
}

template <typename T>
static int launch_scatter_typed(const Plan *p, int trans, int K, void *d_v, const void *d_gv, hipStream_t st) {
  // list selection: ogsOperator.cpp:214-222
  const bool useN = (trans == OGS_TRANS);
  const int nrows = useN ? p->NrowsN : p->NrowsT;
  if (nrows == 0 || p->nparts == 0) return OGS_SUCCESS;
  // a dense list (at least half as many entries as columns) scatters by
  // columns where a column's fields take at most 16 bytes: row-wise kernels
  // write such values in parts of lines that blocks far apart share (fp64
  // K = 1 at 48^3 p=7: 198 -> 188 us; 77^3 p=4: 256 -> 190 us; fp64 K = 2 at
  // 48^3 p=7: 61.5 -> 68.7 % of HBM peak).  With wider fields every column
  // re-reads its row's values (fp64 K = 3: 62.4 -> 61.4 %; K = 16: 2.20 ms
  // row-wise -> 2.38 ms by columns), so the row-wise kernels keep those.
  const int32_t *colRow = useN ? p->colRowN : p->colRowT;
  const T *gv = static_cast<const T *>(d_gv);
  T *v = static_cast<T *>(d_v);
  const RowMap *rm = useN ? &p->rowMapN : &p->rowMapT;
  // (Row-wise over partition 0's byte codes -- the block's rows staged in
  // LDS, each entry's row from per-group head counts, ~1 index byte per
  // entry -- measured slower: fp64 K = 1 192^3 p=1 150 -> 272 us, 128^3
  // p=2 171 -> 293 us, 96^3 p=3 183 -> 191 us, where a block's entries
  // write each element's line in parts that other blocks write too (the
  // by-column kernel writes whole lines); fp64 K = 3 48^3 p=7 533 -> 942 us,
  // 96^3 p=3 432 -> 783 us, with 54 KiB of LDS per 2,048-entry block
  // against the row-wise kernel's 512-entry blocks.)
  // Fields of up to 24 bytes go by columns too on lists of at least 3
  // entries per row (low degree), where the row-wise kernels' partial lines
  // cost most: fp64 K = 3 192^3 p=1 756 -> 482 us, 128^3 p=2 646 -> 541 us
  // (96^3 p=3, whose elements fill whole lines, 432 -> 522 us by columns:
  // row-wise there).
  const int64_t colBytes = (int64_t)K * (int64_t)sizeof(T);
  const bool colsFit = colBytes <= 16 || (colBytes <= 24 && (int64_t)(useN ? p->nnzN : p->nnzT) >= 3 * (int64_t)nrows);
  if ((colRow || (K == 1 && rm->codes)) && colsFit)
    return launch_scatter_cols<T>(p->Ncols, K, colRow, gv, v, st, rm, p->NrowsT);
  const bool partN = useN && !p->aliased;
  // K = 1 from a sparse T list without empty rows: rows staged in LDS, each
  // entry's row found by counting heads, on the merge-path blocks (measured
  // faster for this kernel than blocks cut on entries, except at degree 1);
  // on a small list, on the narrower entry-cut blocks (with packs) that fill
  // the device
  if (K == 1 && !colRow && !partN && !p->partT[0].rowsInKey && p->nnzT > 0 && p->partS[0].maxT <= kCap) {
    const Partition *vp = &p->partS[0];
    if (vp->nblocks < kFillPerCu * p->cus && p->nparts > 1) vp = fill_device(p, &p->partT[1]);
    const ListView S = list_view(p, false, vp, false);
    hipLaunchKernelGGL((k_scatter_vec<T, 1>), dim3(vp->nblocks), dim3(kBlock), 0, st, vp->d, p->rsT, p->ciT, S.pk,
                       S.base, S.stride, gv, v, (vp->tiled & kTiledScatter) ? vp->order : nullptr);
    return (int)hipGetLastError();
  }
  // other field counts with a compile-time shape: the same kernel on the
  // entry-cut blocks whose blocks fit the shape (with packs), narrowed to
  // fill the device on small lists
  if (vec_k(K) && !partN && !p->partT[0].rowsInKey && p->nnzT > 0) {
    // (not the three-field partition: 682-entry blocks measured 11 % slower
    // than 512-entry ones for fp64 K = 3 at 48^3 p=7)
    const Partition *vp = vec_partition<T>(p, K, 0);
    if (vp) {
      vp = fill_device(p, vp);
      return launch_scatter_vec<T>(K, vp, list_view(p, false, vp, false), p, gv, v, st);
    }
  }
  if (colRow) return launch_scatter_cols<T>(p->Ncols, K, colRow, gv, v, st, rm, p->NrowsT);
  const Partition *parts = partN ? p->partN : p->partS;
  int kc = 0;
  const Partition *pt = choose(parts, p->nparts, K, [&](const Partition &x) { return useN ? x.maxN : x.maxT; }, &kc);
  if (!pt) return OGS_E_ROW_TOO_LONG;
  const int32_t *sS = useN ? p->rsN : p->rsT;
  const int32_t *sI = useN ? p->ciN : p->ciT;
  const dim3 grid(pt->nblocks), block(kBlock);
  if (K == 1) {
    const ListView S = list_view(p, useN && !p->aliased, pt, partN);
    hipLaunchKernelGGL((k_scatter_k1<T>), grid, block, 0, st, pt->d, nrows, S, useN ? 1 : 0, gv, v);
  } else {
    hipLaunchKernelGGL((k_scatter<T>), grid, block, 0, st, pt->d, nrows, sS, sI, useN ? 1 : 0, K, kc, gv, v);
  }
  return (int)hipGetLastError();
}

template <typename T>
static int launch_extract_typed(int32_t N, int32_t K, const int32_t *ids, const void *q, void *gq, hipStream_t st) {
  const int64_t total = (int64_t)N * K;
  if (total == 0) return OGS_SUCCESS;
  hipLaunchKernelGGL((k_extract<T>), dim3(grid_for(total)), dim3(kBlock), 0, st, total, K, ids,
                     static_cast<const T *>(q), static_cast<T *>(gq));
  return (int)hipGetLastError();
}

static bool bad_type(int t) { return t < OGS_FLOAT || t > OGS_FLOAT_C; }
// Mul is not offered on the precision-form storage types.
static bool not_offered(int type, int op) { return (type == OGS_BF16 || type == OGS_FLOAT_C) && op == OGS_MUL; }
static bool bad_op(int o) { return o < OGS_ADD || o > OGS_MIN; }
static bool bad_trans(int t) { return t < OGS_SYM || t > OGS_TRANS; }

// ---------------------------------------------------------------------------
// Plan construction.

static int read_int(const int32_t *d, hipStream_t st, int32_t *h) {
  hipError_t e = hipMemcpyAsync(h, d, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return (int)e;
}

static int scan_csr(const int32_t *rs, int nrows, const int32_t *ci, int nnz, int ncols, int32_t *scratch,
                    hipStream_t st, int32_t *maxRow, int32_t *flags) {
  hipError_t e = hipMemsetAsync(scratch, 0, 2 * sizeof(int32_t), st);
  if (e != hipSuccess) return (int)e;
  const int64_t n = nrows > nnz ? nrows : nnz;
  hipLaunchKernelGGL(k_scan_csr, dim3(grid_for(n, 4096)), dim3(kBlock), 0, st, rs, nrows, ci, nnz, ncols, scratch);
  int32_t h[2] = {0, 0};
  e = hipMemcpyAsync(h, scratch, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return (int)e;
  *maxRow = h[0];
  *flags = h[1];
  return OGS_SUCCESS;
}

// Block order of a partition of n blocks for layer distance D: each XCD's
// run of blocks (the runs of xcd_block) walked in bands of kTileL layers of D
// blocks, each band in tiles of kTileP blocks, each tile layer by layer;
// order[slot] = block, a permutation of [0, n).
static std::vector<int> tiled_order(int n, int D) {
  std::vector<int> order(n);
  const int q = n / kNumXcd, r = n % kNumXcd;
  int t = 0;
  for (int x = 0; x < kNumXcd; ++x) {
    const int c0 = x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    const int len = x < r ? q + 1 : q;
    const int layers = (len + D - 1) / D;
    for (int l0 = 0; l0 < layers; l0 += kTileL)
      for (int p0 = 0; p0 < D; p0 += kTileP)
        for (int l = l0; l < l0 + kTileL && l < layers; ++l)
          for (int pp = p0; pp < p0 + kTileP && pp < D; ++pp) {
            const int i = l * D + pp;
            if (i < len) order[t++] = c0 + i;
          }
  }
  return order;
}

// Mean time of 8 launches (after one untimed) of `launch`.
template <class F>
static int time_launches(F launch, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float *us) {
  constexpr int reps = 8;
  launch();
  int rc = (int)hipEventRecord(e0, st);
  for (int i = 0; i < reps && !rc; ++i) launch();
  if (!rc) rc = (int)hipEventRecord(e1, st);
  if (!rc) rc = (int)hipEventSynchronize(e1);
  float ms = 0.f;
  if (!rc) rc = (int)hipEventElapsedTime(&ms, e0, e1);
  *us = ms * 1e3f / reps;
  return rc;
}

// Whether `order` makes `launch(order)` at least 5% faster than block id
// order: best of three passes each, the candidates' order alternating (the
// first launches after a setup run slow).
template <class F>
static int order_pays(F launch, const int *order, hipStream_t st, hipEvent_t e0, hipEvent_t e1, bool *pays) {
  float best[2] = {0.f, 0.f};
  int rc = 0;
  for (int pass = 0; pass < 3 && !rc; ++pass) {
    for (int i = 0; i < 2 && !rc; ++i) {
      const int c = (pass & 1) ? 1 - i : i;
      float us = 0.f;
      rc = time_launches([&]() { launch(c ? order : nullptr); }, st, e0, e1, &us);
      if (!rc && (best[c] == 0.f || us < best[c])) best[c] = us;
    }
  }
  *pays = !rc && best[1] < 0.95f * best[0];
  return rc;
}

// The most common block distance D >= kTileDMin between the blocks of
// adjacent columns of the T list over partition `pt` (column ids `ids` in
// the order pt indexes), or 0 when none covers 0.5% of the columns.
static int layer_distance(const Plan *p, const Partition &pt, const int32_t *ids, hipStream_t st, int *D) {
  *D = 0;
  int32_t *blk = nullptr;
  uint32_t *hist = nullptr;
  if (hipMalloc(&blk, sizeof(int32_t) * (size_t)p->Ncols) != hipSuccess ||
      hipMalloc(&hist, sizeof(uint32_t) * kTileDMax) != hipSuccess) {
    (void)hipGetLastError();  // no room for the histogram: keep block id order
    if (blk) (void)hipFree(blk);
    return 0;
  }
  int rc = 0;
  if (!rc) rc = (int)hipMemsetAsync(blk, 0xff, sizeof(int32_t) * (size_t)p->Ncols, st);
  if (!rc) rc = (int)hipMemsetAsync(hist, 0, sizeof(uint32_t) * kTileDMax, st);
  if (!rc) {
    hipLaunchKernelGGL(k_col_block, dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, ids, blk);
    hipLaunchKernelGGL(k_block_gaps, dim3(grid_for(p->Ncols, 1024)), dim3(kBlock), 0, st, blk, p->Ncols, hist);
    rc = (int)hipGetLastError();
  }
  std::vector<uint32_t> h(kTileDMax, 0);
  if (!rc) rc = (int)hipMemcpyAsync(h.data(), hist, sizeof(uint32_t) * kTileDMax, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (blk) (void)hipFree(blk);
  if (hist) (void)hipFree(hist);
  if (rc) return rc;
  int best = 0;
  for (int d = kTileDMin; d < kTileDMax; ++d)
    if (!best || h[d] > h[best]) best = d;
  if (best && (double)h[best] >= 0.005 * p->Ncols) *D = best;
  return 0;
}

// Tiled block orders of the K = 1 kernels of a T list without empty rows:
// for the GatherScatter and Gather over partT[0] and the row-wise Scatter
// over partS[0] (when the list is too sparse to scatter by columns), each
// partition's layer distance, its tiled order, and per kernel whether that
// order measures faster on scratch values (fp64).  Plans whose widest
// partition has fewer than 16 blocks per CU keep block id order.
static int tile_schedule(Plan *p, hipStream_t st) {
  Partition &pt = p->partT[0];
  Partition &ps = p->partS[0];
  if (p->nparts == 0 || pt.nblocks < 16 * p->cus || p->Ncols < 2 || p->nnzT == 0 || pt.maxT > kCap) return 0;
  // the row-wise scatter's own merge-path blocks (a T list too sparse to
  // scatter by columns)
  const bool rowScatter = ps.d != pt.d && ps.maxT <= kCap;
  int DT = 0, DS = 0;
  int rc = layer_distance(p, pt, p->ciT, st, &DT);
  if (!rc && rowScatter) rc = layer_distance(p, ps, p->ciT, st, &DS);
  if (rc || (!DT && !DS)) return rc;
  double *vals = nullptr, *rows = nullptr;
  int *oT = nullptr, *oS = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMalloc(&vals, sizeof(double) * (size_t)p->Ncols) != hipSuccess ||
      hipMalloc(&rows, sizeof(double) * (size_t)p->NrowsT) != hipSuccess) {
    (void)hipGetLastError();  // no room for scratch values: keep block id order
    if (vals) (void)hipFree(vals);
    return 0;
  }
  rc = (int)hipMemsetAsync(vals, 0, sizeof(double) * (size_t)p->Ncols, st);
  if (!rc) rc = (int)hipMemsetAsync(rows, 0, sizeof(double) * (size_t)p->NrowsT, st);
  if (!rc) rc = (int)hipEventCreate(&e0);
  if (!rc) rc = (int)hipEventCreate(&e1);
  auto upload = [&](const Partition &x, int D, int **out) {
    const std::vector<int> order = tiled_order(x.nblocks, D);
    int r = (int)hipMalloc(out, sizeof(int) * order.size());
    if (!r) r = (int)hipMemcpyAsync(*out, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, st);
    if (!r) r = (int)hipStreamSynchronize(st);
    return r;
  };
  bool gs = false, ga = false, sc = false;
  if (!rc && DT) {
    rc = upload(pt, DT, &oT);
    const Pack &k = p->packT[0];
    const DeltaPack &dp = p->dpack;
    // the kernels the K = 1 launches will run: over the delta pack when the plan has one
    if (!rc)
      rc = order_pays([&](const int *o) {
        if (dp.w)
          launch_delta_on<double, OGS_ADD>(pt, dp, p->rsT, p->ciT, p->Ncols, vals, o, st);
        else
          hipLaunchKernelGGL((k_gs_sym_vec<double, OGS_ADD, 1>), dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d,
                             p->rsT, p->ciT, k.pk, k.base, k.stride, 0, vals, o);
      }, oT, st, e0, e1, &gs);
    if (!rc)
      rc = order_pays([&](const int *o) {
        if (dp.w)
          launch_gather_delta_on<double, OGS_ADD>(pt, dp, p->rsT, p->ciT, p->Ncols, vals, rows, o, st);
        else
          hipLaunchKernelGGL((k_gather_vec<double, OGS_ADD, 1>), dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d,
                             p->rsT, p->ciT, k.pk, k.base, k.stride, vals, rows, o);
      }, oT, st, e0, e1, &ga);
  }
  if (!rc && DS) {
    rc = upload(ps, DS, &oS);
    const Pack &k = p->packS;
    if (!rc)
      rc = order_pays([&](const int *o) {
        hipLaunchKernelGGL((k_scatter_vec<double, 1>), dim3(ps.nblocks), dim3(kBlock), 0, st, ps.d, p->rsT, p->ciT,
                           k.pk, k.base, k.stride, rows, vals, o);
      }, oS, st, e0, e1, &sc);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (vals) (void)hipFree(vals);
  if (rows) (void)hipFree(rows);
  if (!rc && (gs || ga)) {
    pt.order = oT;
    pt.tileD = DT;
    pt.tiled = (gs ? kTiledGS : 0) | (ga ? kTiledGather : 0);
  } else if (oT) {
    (void)hipFree(oT);
  }
  if (!rc && sc) {
    ps.order = oS;
    ps.tileD = DS;
    ps.tiled = kTiledScatter;
  } else if (oS) {
    (void)hipFree(oS);
  }
  return rc;
}

// colRow (allocated here) of the list (rs, ci) over rows [0, nrows); left
// null, and the list scattered by rows, when the device has no room for it.
static int build_col_rows(const int32_t *rs, int32_t nrows, const int32_t *ci, int32_t ncols, hipStream_t st,
                          int32_t **colRow) {
  if (hipMalloc(colRow, sizeof(int32_t) * (size_t)ncols) != hipSuccess) {
    (void)hipGetLastError();
    *colRow = nullptr;
    return 0;
  }
  int rc = (int)hipMemsetAsync(*colRow, 0xff, sizeof(int32_t) * (size_t)ncols, st);
  if (!rc && nrows > 0) {
    hipLaunchKernelGGL(k_col_rows, dim3(grid_for(nrows, 16384)), dim3(kBlock), 0, st, rs, nrows, ci, *colRow);
    rc = (int)hipGetLastError();
  }
  return rc;
}

static void free_row_map(RowMap &m) {
  if (m.fbRows) (void)hipFree(m.fbRows);
  if (m.codes) (void)hipFree(m.codes);
  if (m.x) (void)hipFree(m.x);
  if (m.table) (void)hipFree(m.table);
  m = RowMap();
}

// The row-delta map of colRow (none when the device has no room or more
// than a quarter of the chunks would read colRow): the table from a sample
// of 2^20 columns mid-list (the 14 commonest nonzero deltas between
// consecutive row-bearing columns), E from a histogram of each chunk's most
// escapes per range (the fewest map bytes, a fallback chunk counting as its
// 8 KiB of colRow), then the codes.
static int build_row_map(const int32_t *colRow, int32_t ncols, int cus, int32_t *scratch, hipStream_t st, RowMap &m) {
  m = RowMap();
  if (!colRow || ncols <= 0) return 0;
  const int chunks = (int)(((int64_t)ncols + kRowChunk - 1) / kRowChunk);
  if (chunks < kFillPerCu * cus) return 0;  // small lists: the int32 map
  const int32_t w0 = ncols > kRowSample ? ncols / 2 - kRowSample / 2 : 0;
  const int32_t wn = std::min<int32_t>(kRowSample, ncols - w0);
  std::vector<int32_t> cr(wn);
  int rc = (int)hipMemcpyAsync(cr.data(), colRow + w0, sizeof(int32_t) * (size_t)wn, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) return rc;
  std::vector<uint32_t> d;
  d.reserve(wn);
  int32_t prev = -1;
  for (int32_t i = 0; i < wn; ++i) {
    if (cr[i] < 0) continue;
    if (prev >= 0 && cr[i] != prev) d.push_back((uint32_t)cr[i] - (uint32_t)prev);
    prev = cr[i];
  }
  std::sort(d.begin(), d.end());
  std::vector<std::pair<int, uint32_t>> runs;  // (count, delta)
  for (size_t i = 0; i < d.size();) {
    size_t j = i;
    while (j < d.size() && d[j] == d[i]) ++j;
    runs.push_back({(int)(j - i), d[i]});
    i = j;
  }
  std::stable_sort(runs.begin(), runs.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
  uint32_t tbl[16] = {0u};
  for (int c = 1; c < (int)kDEscape && c - 1 < (int)runs.size(); ++c) tbl[c] = runs[c - 1].second;
  uint32_t *hist = nullptr;
  if (hipMalloc(&m.table, sizeof(tbl)) != hipSuccess || hipMalloc(&hist, sizeof(uint32_t) * (kDMaxE + 2)) != hipSuccess) {
    (void)hipGetLastError();  // no room: colRow serves
    free_row_map(m);
    return 0;
  }
  rc = (int)hipMemcpyAsync(m.table, tbl, sizeof(tbl), hipMemcpyHostToDevice, st);
  if (!rc) rc = (int)hipMemsetAsync(hist, 0, sizeof(uint32_t) * (kDMaxE + 2), st);
  if (!rc) {
    hipLaunchKernelGGL(k_row_delta_count, dim3(chunks), dim3(kBlock), 0, st, colRow, (int64_t)ncols, m.table, hist);
    rc = (int)hipGetLastError();
  }
  std::vector<uint32_t> h(kDMaxE + 2, 0u);
  if (!rc) rc = (int)hipMemcpyAsync(h.data(), hist, sizeof(uint32_t) * (kDMaxE + 2), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  (void)hipFree(hist);
  if (rc) { free_row_map(m); return rc; }
  double best = -1.0;
  for (int E = 4; E <= kDMaxE; E += 4) {
    double over = 0.0;
    for (int k = E + 1; k <= kDMaxE + 1; ++k) over += h[k];
    const double bytes = (double)chunks * kDRanges * E * 4.0 + over * 4.0 * kRowChunk;
    if (best < 0.0 || bytes < best) { best = bytes; m.E = E; }
  }
  m.chunks = chunks;
  const size_t nw = (size_t)chunks * kBlock, nx = (size_t)chunks * kDRanges * m.E;
  if (hipMalloc(&m.codes, sizeof(uint32_t) * nw) != hipSuccess || hipMalloc(&m.x, sizeof(uint32_t) * nx) != hipSuccess) {
    (void)hipGetLastError();  // no room: colRow serves
    free_row_map(m);
    return 0;
  }
  rc = (int)hipMemsetAsync(m.x, 0, sizeof(uint32_t) * nx, st);
  if (!rc) rc = (int)hipMemsetAsync(scratch, 0, sizeof(int32_t), st);
  if (!rc) {
    hipLaunchKernelGGL(k_row_delta_build, dim3(chunks), dim3(kBlock), 0, st, colRow, (int64_t)ncols, m.table, m.E,
                       m.codes, m.x, scratch);
    rc = (int)hipGetLastError();
  }
  int32_t fb = 0;
  if (!rc) rc = (int)hipMemcpyAsync(&fb, scratch, sizeof(fb), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  m.fallback = fb;
  if (rc || 4 * (int64_t)fb > chunks) free_row_map(m);
  return rc;
}

// Moves the rows of row map m's fallback chunks out of colRow (ncols
// columns) into an array of their own, so that colRow can go.  False (m
// unchanged) when the device has no room.
static int compact_rows(RowMap &m, const int32_t *colRow, int32_t ncols, int32_t *scratch, hipStream_t st,
                        bool *done) {
  *done = false;
  if (!m.codes || m.fbRows) return 0;
  int32_t *fb = nullptr;
  if (m.fallback > 0 && hipMalloc(&fb, sizeof(int32_t) * (size_t)m.fallback * kRowChunk) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  int rc = (int)hipMemsetAsync(scratch, 0, sizeof(int32_t), st);
  if (!rc && fb) {
    hipLaunchKernelGGL(k_compact_rows, dim3(m.chunks), dim3(kBlock), 0, st, m.x, m.E, colRow, (int64_t)ncols, fb,
                       scratch);
    rc = (int)hipGetLastError();
  }
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) {
    if (fb) (void)hipFree(fb);
    return rc;
  }
  m.fbRows = fb;  // (no fallback chunk: nothing reads colRow)
  *done = true;
  return 0;
}

// nMask of a plan with colRowT (no room or a failed check: none).
static int build_n_mask(Plan *p, int32_t nrows, int32_t *scratch, hipStream_t st) {
  const size_t words = ((size_t)p->Ncols + 31) / 32;
  if (hipMalloc(&p->nMask, sizeof(uint32_t) * words) != hipSuccess) {
    (void)hipGetLastError();
    p->nMask = nullptr;
    return 0;
  }
  hipError_t e = hipMemsetAsync(p->nMask, 0, sizeof(uint32_t) * words, st);
  if (e == hipSuccess) e = hipMemsetAsync(scratch, 0, 2 * sizeof(int32_t), st);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_n_mask, dim3(grid_for(nrows, 16384)), dim3(kBlock), 0, st, p->rsT, p->ciT, p->rsN, p->ciN,
                     nrows, p->colRowT, p->nMask, scratch);
  hipLaunchKernelGGL(k_count_cols, dim3(grid_for(p->Ncols, 16384)), dim3(kBlock), 0, st, p->colRowT, p->Ncols,
                     scratch + 1);
  int32_t h[2] = {0, 0};
  e = hipMemcpyAsync(h, scratch, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return (int)e;
  if (h[0] != 0 || h[1] != p->nnzT) {
    (void)hipFree(p->nMask);
    p->nMask = nullptr;
  }
  return 0;
}

static int build_partition(const Plan *p, const int32_t *rsT, const int32_t *rsN, bool cutOnN, bool rowsInKey,
                           int nrows, int W, int32_t *scratch, hipStream_t st, Partition *out) {
  const int64_t keys = (int64_t)(cutOnN ? p->nnzN : p->nnzT) + (rowsInKey ? nrows : 0);
  const int64_t nb = (keys + W - 1) / W;
  if (nb > INT32_MAX - 1) return OGS_E_ARG;
  out->W = W;
  out->rowsInKey = rowsInKey;
  out->nblocks = (int)nb;
  hipError_t e = hipMalloc(&out->d, sizeof(int4) * (size_t)(nb + 1));
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_partition, dim3(grid_for(nb + 1, 1 << 30)), dim3(kBlock), 0, st, rsT, rsN,
                     cutOnN ? 1 : 0, rowsInKey ? 1 : 0, nrows, W, (int)nb, out->d);
  e = hipMemsetAsync(scratch, 0, 3 * sizeof(int32_t), st);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_part_max, dim3(grid_for(nb, 2048)), dim3(kBlock), 0, st, out->d, (int)nb, scratch);
  int32_t h[3];
  e = hipMemcpyAsync(h, scratch, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return (int)e;
  out->maxT = h[0];
  out->maxN = h[1];
  out->maxRows = h[2];
  return OGS_SUCCESS;
}

static int build_pack(const Partition &pt, int nrows, const int32_t *starts, const int32_t *ids, bool useN,
                      int nnz, int32_t *scratch, hipStream_t st, Pack *out) {
  if (pt.nblocks <= 0 || nnz <= 0) return OGS_SUCCESS;
  // fixed stride: the largest block of this list, rounded up to a wave
  const int mx = useN ? pt.maxN : pt.maxT;
  out->stride = ((mx > 0 ? mx : 1) + 63) / 64 * 64;
  const size_t words = (size_t)pt.nblocks * (size_t)out->stride;
  hipError_t e = hipMalloc(&out->pk, sizeof(uint32_t) * words);
  if (e == hipSuccess) e = hipMalloc(&out->base, sizeof(int4) * (size_t)pt.nblocks);
  if (e == hipSuccess) e = hipMemsetAsync(out->pk, 0, sizeof(uint32_t) * words, st);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_build_pack, dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, nrows, starts, ids, useN ? 1 : 0,
                     out->stride, out->pk, out->base);
  e = hipMemsetAsync(scratch, 0, sizeof(int32_t), st);
  if (e != hipSuccess) return (int)e;
  hipLaunchKernelGGL(k_count_packed, dim3(grid_for(pt.nblocks, 1024)), dim3(kBlock), 0, st, out->base, pt.nblocks,
                     scratch);
  int32_t h = 0;
  e = hipMemcpyAsync(&h, scratch, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return (int)e;
  out->packed = h;
  out->plain = pt.nblocks - h;
  return OGS_SUCCESS;
}

static void free_pack(Pack &k) {
  if (k.pk) (void)hipFree(k.pk);
  if (k.base) (void)hipFree(k.base);
  k = Pack();
}

static void free_byte(DeltaPack &d) {
  for (void *a : {(void *)d.bw, (void *)d.bh, (void *)d.bx, (void *)d.bt})
    if (a) (void)hipFree(a);
  d.bw = nullptr;
  d.bh = d.bx = d.bt = nullptr;
  d.bItems = 0;
  d.bFallback = 0;
}

static void free_delta(DeltaPack &d) {
  free_byte(d);
  if (d.w) (void)hipFree(d.w);
  if (d.x) (void)hipFree(d.x);
  if (d.ownsF) free_pack(d.f);
  d = DeltaPack();
}

// Moves the fallback blocks' 32-bit words of delta pack d (over a partition
// of nblocks blocks, borrowing that partition's pack) into a pack of their
// own, so that the partition's pack can go.  False (d unchanged) when the
// device has no room.
static int compact_fallback(DeltaPack &d, int nblocks, int32_t *scratch, hipStream_t st, bool *done) {
  *done = false;
  if (!d.w || d.ownsF) return 0;
  const Pack k = d.f;
  Pack c;
  c.stride = k.stride;
  c.packed = d.fallback;
  if (d.fallback > 0 &&
      (hipMalloc(&c.pk, sizeof(uint32_t) * (size_t)d.fallback * (size_t)k.stride) != hipSuccess ||
       hipMalloc(&c.base, sizeof(int4) * (size_t)d.fallback) != hipSuccess)) {
    (void)hipGetLastError();
    free_pack(c);
    return 0;
  }
  int rc = 0;
  if (d.fallback > 0) {
    rc = (int)hipMemsetAsync(scratch, 0, sizeof(int32_t), st);
    if (!rc) {
      hipLaunchKernelGGL(k_compact_fallback, dim3(nblocks), dim3(kBlock), 0, st, d.x, d.E, k.pk, k.base, k.stride,
                         c.pk, c.base, scratch);
      rc = (int)hipGetLastError();
    }
    if (!rc) rc = (int)hipStreamSynchronize(st);
  }
  if (rc) {
    free_pack(c);
    return rc;
  }
  d.f = c;
  d.ownsF = true;
  *done = true;
  return 0;
}

// The delta pack of the T list over partition 0, from its 32-bit pack:
// a histogram of the deltas picks the table (the 14 commonest below 256), a
// histogram of each block's most exceptions per range picks E (the fewest
// index bytes: a fallback block reads its 8 KiB of 32-bit words on top of
// its delta words), then the
// words.  Built for entry-cut partitions that fill the device at K = 1 (so
// that the K = 1 launches use partition 0); none when the device has no room.
static int build_delta_pack(const Plan *p, const Partition &pt, const Pack &k, DeltaPack &dp, int32_t *scratch,
                            hipStream_t st) {
  if (!k.pk || k.packed == 0 || pt.rowsInKey || pt.maxT > kCap || pt.nblocks < kFillPerCu * p->cus) return 0;
  uint32_t *hist = nullptr;
  if (hipMalloc(&hist, sizeof(uint32_t) * 256) != hipSuccess) { (void)hipGetLastError(); return 0; }
  int rc = (int)hipMemsetAsync(hist, 0, sizeof(uint32_t) * 256, st);
  if (!rc) {
    hipLaunchKernelGGL(k_delta_hist, dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, k.pk, k.base, k.stride, hist);
    rc = (int)hipGetLastError();
  }
  std::vector<uint32_t> h(256, 0u);
  if (!rc) rc = (int)hipMemcpyAsync(h.data(), hist, sizeof(uint32_t) * 256, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  std::vector<int> cand;
  for (int d = 1; d < 256; ++d)
    if (h[d]) cand.push_back(d);
  std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return h[a] > h[b]; });
  uint8_t tb[16] = {0};
  for (int c = 1; c < (int)kDEscape && c - 1 < (int)cand.size(); ++c) tb[c] = (uint8_t)cand[c - 1];
  for (int i = 0; i < 4; ++i)
    dp.tbl[i] = (uint32_t)tb[4 * i] | ((uint32_t)tb[4 * i + 1] << 8) | ((uint32_t)tb[4 * i + 2] << 16) |
                ((uint32_t)tb[4 * i + 3] << 24);
  const uint4 tbv = make_uint4(dp.tbl[0], dp.tbl[1], dp.tbl[2], dp.tbl[3]);
  // blocks by their most exceptions in one range: [0, kDMaxE], more
  if (!rc) rc = (int)hipMemsetAsync(hist, 0, sizeof(uint32_t) * (kDMaxE + 2), st);
  if (!rc) {
    hipLaunchKernelGGL(k_delta_count, dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, k.pk, k.base, k.stride, tbv, hist);
    rc = (int)hipGetLastError();
  }
  if (!rc) rc = (int)hipMemcpyAsync(h.data(), hist, sizeof(uint32_t) * (kDMaxE + 2), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  (void)hipFree(hist);
  if (rc) return rc;
  double best = -1.0;
  for (int E = 4; E <= kDMaxE; E += 4) {
    double over = k.plain;
    for (int m = E + 1; m <= kDMaxE + 1; ++m) over += h[m];
    const double bytes = (double)pt.nblocks * kDRanges * E * 4.0 + over * 4.0 * kCap;
    if (best < 0.0 || bytes < best) { best = bytes; dp.E = E; }
  }
  const size_t nw = (size_t)pt.nblocks * kCap, nx = (size_t)pt.nblocks * kDRanges * dp.E;
  if (hipMalloc(&dp.w, sizeof(uint16_t) * nw) != hipSuccess || hipMalloc(&dp.x, sizeof(uint32_t) * nx) != hipSuccess) {
    (void)hipGetLastError();  // no room: the 32-bit pack serves
    if (dp.w) (void)hipFree(dp.w);
    dp = DeltaPack();
    return 0;
  }
  rc = (int)hipMemsetAsync(dp.w, 0, sizeof(uint16_t) * nw, st);
  if (!rc) rc = (int)hipMemsetAsync(dp.x, 0, sizeof(uint32_t) * nx, st);
  if (!rc) rc = (int)hipMemsetAsync(scratch, 0, sizeof(int32_t), st);
  if (!rc) {
    hipLaunchKernelGGL(k_build_delta, dim3(pt.nblocks), dim3(kBlock), 0, st, pt.d, k.pk, k.base, k.stride, tbv, dp.E,
                       dp.w, dp.x, scratch);
    rc = (int)hipGetLastError();
  }
  int32_t fb = 0;
  if (!rc) rc = (int)hipMemcpyAsync(&fb, scratch, sizeof(fb), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  dp.fallback = fb;
  dp.csr = k.plain;  // a block left in CSR order is always a fallback block
  dp.f = k;  // fallback blocks read the partition's pack (slot = block id)
  dp.ownsF = false;
  dp.rows = (int64_t)p->nnzT >= 2 * (int64_t)p->NrowsT;
  dp.seg8 = p->maxRowT <= 8;
  // mostly fallback blocks (columns without short runs): the 32-bit pack serves
  if (!rc && 4 * (int64_t)fb > pt.nblocks) free_delta(dp);
  return rc;
}

// The byte pack of delta pack dp's blocks (partition pt, its 32-bit pack k):
// the table from the triples of up to 512 sample blocks spread over the
// partition (the 255 commonest), each range's exception items counted, their
// offsets summed on the host, then the codes.  Built for blocks that fold at
// the heads of their entries or in groups of 8 (not for the kRows form);
// none when the device has no room or more than a quarter of the blocks
// would fall back.
constexpr int kByteSample = 512;

static int build_byte_pack(const Partition &pt, const Pack &k, DeltaPack &dp, hipStream_t st) {
  if (!dp.w || (dp.rows && !dp.seg8) || !k.pk || pt.nblocks == 0) return 0;
  const int nb = pt.nblocks;
  const int every = nb > kByteSample ? nb / kByteSample : 1;
  const int ns = (nb - 1) / every + 1;
  uint32_t *keys = nullptr, *skey = nullptr, *cnt = nullptr, *off = nullptr;
  uint8_t *scode = nullptr;
  auto release = [&]() {
    for (void *a : {(void *)keys, (void *)skey, (void *)cnt, (void *)off, (void *)scode})
      if (a) (void)hipFree(a);
  };
  const size_t nr = (size_t)nb * kDRanges;
  if (hipMalloc(&keys, sizeof(uint32_t) * (size_t)ns * kCap) != hipSuccess ||
      hipMalloc(&skey, sizeof(uint32_t) * 256) != hipSuccess || hipMalloc(&scode, 256) != hipSuccess ||
      hipMalloc(&cnt, sizeof(uint32_t) * nr) != hipSuccess || hipMalloc(&off, sizeof(uint32_t) * nr) != hipSuccess) {
    (void)hipGetLastError();  // no room: the 16-bit words serve
    release();
    return 0;
  }
  hipLaunchKernelGGL(k_byte_keys, dim3(ns), dim3(kBlock), 0, st, pt.d, k.pk, k.base, k.stride, every, keys);
  int rc = (int)hipGetLastError();
  std::vector<uint32_t> h((size_t)ns * kCap);
  if (!rc) rc = (int)hipMemcpyAsync(h.data(), keys, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) { release(); return rc; }
  std::sort(h.begin(), h.end());
  std::vector<std::pair<uint32_t, uint32_t>> runs;  // (count, key)
  for (size_t i = 0; i < h.size() && h[i] != kDFallback;) {
    size_t j = i;
    while (j < h.size() && h[j] == h[i]) ++j;
    runs.push_back({(uint32_t)(j - i), h[i]});
    i = j;
  }
  std::stable_sort(runs.begin(), runs.end(), [](const std::pair<uint32_t, uint32_t> &a,
                                                const std::pair<uint32_t, uint32_t> &b) { return a.first > b.first; });
  const int nk = (int)std::min<size_t>(runs.size(), kBEscape);
  std::vector<uint32_t> tbl(256, 0u), sk(256, kDFallback);
  std::vector<uint8_t> sc(256, (uint8_t)kBEscape);
  std::vector<std::pair<uint32_t, uint8_t>> byKey;
  for (int c = 0; c < nk; ++c) {
    tbl[c] = runs[c].second;
    byKey.push_back({runs[c].second, (uint8_t)c});
  }
  std::sort(byKey.begin(), byKey.end());
  for (int i = 0; i < nk; ++i) {
    sk[i] = byKey[i].first;
    sc[i] = byKey[i].second;
  }
  rc = (int)hipMemcpyAsync(skey, sk.data(), sizeof(uint32_t) * 256, hipMemcpyHostToDevice, st);
  if (!rc) rc = (int)hipMemcpyAsync(scode, sc.data(), 256, hipMemcpyHostToDevice, st);
  if (!rc) {
    hipLaunchKernelGGL(k_byte_count, dim3(nb), dim3(kBlock), 0, st, pt.d, k.pk, k.base, k.stride, skey, scode, nk,
                       cnt);
    rc = (int)hipGetLastError();
  }
  std::vector<uint32_t> c(nr), o(nr);
  if (!rc) rc = (int)hipMemcpyAsync(c.data(), cnt, sizeof(uint32_t) * nr, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) { release(); return rc; }
  int64_t items = 0;
  int32_t fb = 0;
  for (size_t r = 0; r < nr; ++r) {
    o[r] = (uint32_t)items;
    items += c[r];
    if (r % kDRanges == 0 && c[r] == 0u) ++fb;
  }
  if (items >= (int64_t)1 << 25 || 4 * (int64_t)fb > nb) { release(); return 0; }
  DeltaPack b;
  if (hipMalloc(&b.bw, (size_t)nb * kCap) != hipSuccess || hipMalloc(&b.bh, sizeof(uint32_t) * nr) != hipSuccess ||
      hipMalloc(&b.bx, sizeof(uint32_t) * 2 * (size_t)(items > 0 ? items : 1)) != hipSuccess ||
      hipMalloc(&b.bt, sizeof(uint32_t) * 256) != hipSuccess) {
    (void)hipGetLastError();
    free_byte(b);
    release();
    return 0;
  }
  rc = (int)hipMemsetAsync(b.bw, 0, (size_t)nb * kCap, st);
  if (!rc) rc = (int)hipMemcpyAsync(b.bt, tbl.data(), sizeof(uint32_t) * 256, hipMemcpyHostToDevice, st);
  if (!rc) rc = (int)hipMemcpyAsync(off, o.data(), sizeof(uint32_t) * nr, hipMemcpyHostToDevice, st);
  if (!rc) {
    hipLaunchKernelGGL(k_byte_build, dim3(nb), dim3(kBlock), 0, st, pt.d, k.pk, k.base, k.stride, skey, scode, nk,
                       cnt, off, b.bw, b.bh, reinterpret_cast<u32x2 *>(b.bx));
    rc = (int)hipGetLastError();
  }
  if (!rc) rc = (int)hipStreamSynchronize(st);
  release();
  if (rc) { free_byte(b); return rc; }
  dp.bw = b.bw;
  dp.bh = b.bh;
  dp.bx = b.bx;
  dp.bt = b.bt;
  dp.bItems = items;
  dp.bFallback = fb;
  return 0;
}

// ---------------------------------------------------------------------------
// Bundled GatherScatter blocks (Plan::bundle).
//
// On a structured hex mesh the T list's rows, in first-encounter order, run
// along lines of the mesh: row r + Ry is r's neighbour in the next line and
// r + Rz in the next layer.  At low degree most rows are shared vertex and
// edge nodes whose members lie in four (or eight) elements of neighbouring
// lines and layers, so blocks of consecutive rows (one line segment each)
// read and write each element's values in quarters, from four blocks: four
// times the cache lines per wave instruction, and lines written back in
// parts.  A bundle puts ty lines and tz layers of rows side by side (row r
// at (bundle, position along the line, layer, line)), so a block consumes
// whole elements.  48^3-element-equivalent meshes (~56 M entries, fp64):
// 192^3 p=1 281 -> 241 us and 128^3 p=2 306 -> ~270 us per GatherScatter,
// measured with strides from the mesh; no gain at p >= 3.

constexpr int kBundleSample = 4096;      // rows whose shifts are compared
constexpr int kBundleWindow = 1 << 20;   // columns whose row distances are counted
constexpr int kBundleMinLen = 3;         // bundles for lists of >= 3 entries per row

namespace {
// Host copy of rows [r0, r0 + n) of a CSR list.
struct RowSample {
  int32_t r0 = 0, n = 0;
  std::vector<int32_t> rs, ci;  // rs[i] - rs[0] indexes ci
  int first(int i) const { return ci[rs[i] - rs[0]]; }
};
}  // namespace

// Fraction of sample rows r in [0, S) for which row r + L's first column
// minus row r's is the commonest such difference (a lattice shift of a
// structured numbering gives one difference for all but its edge rows).
static double shift_freq(const RowSample &rw, int S, int L) {
  if (L <= 0 || S + L > rw.n) return 0.0;
  std::vector<int64_t> d(64);
  for (int i = 0; i < 64; ++i) {
    const int r = i * (S / 64);
    d[i] = (int64_t)rw.first(r + L) - rw.first(r);
  }
  std::sort(d.begin(), d.end());
  int64_t mode = d[0];
  int run = 1, best = 1;
  for (int i = 1; i < 64; ++i) {
    run = d[i] == d[i - 1] ? run + 1 : 1;
    if (run > best) { best = run; mode = d[i]; }
  }
  if (best < 32) return 0.0;
  int hit = 0;
  for (int r = 0; r < S; ++r) hit += ((int64_t)rw.first(r + L) - rw.first(r)) == mode;
  return (double)hit / S;
}

// Mean number of 64-byte column sectors (8 columns) that rows r and r + L
// share, over every 4th sample row.
static double sector_share(const RowSample &rw, int S, int L) {
  if (L <= 0 || S + L > rw.n) return 0.0;
  int64_t tot = 0;
  for (int r = 0; r < S; r += 4) {
    const int32_t *a = rw.ci.data() + (rw.rs[r] - rw.rs[0]), *ae = rw.ci.data() + (rw.rs[r + 1] - rw.rs[0]);
    const int32_t *b = rw.ci.data() + (rw.rs[r + L] - rw.rs[0]), *be = rw.ci.data() + (rw.rs[r + L + 1] - rw.rs[0]);
    int32_t last = -1;
    while (a < ae && b < be) {  // members ascend, so sectors do
      const int32_t x = *a >> 3, y = *b >> 3;
      if (x < y) ++a;
      else if (y < x) ++b;
      else {
        if (x != last) { ++tot; last = x; }
        ++a;
        ++b;
      }
    }
  }
  return (double)tot / ((S + 3) / 4);
}

// The shift in [c - w, c + w] (>= lo) that is a lattice shift (shift_freq
// >= 0.9 of the window's best and >= 0.5) sharing the most sectors; 0: none.
static int lattice_shift(const RowSample &rw, int S, int c, int w, int lo) {
  const int a = std::max(lo, c - w), b = c + w;
  if (a > b) return 0;
  std::vector<double> fr(b - a + 1);
  double mx = 0.0;
  for (int L = a; L <= b; ++L) mx = std::max(mx, fr[L - a] = shift_freq(rw, S, L));
  int best = 0;
  double bs = -1.0;
  for (int L = a; L <= b; ++L) {
    if (fr[L - a] < 0.5 || fr[L - a] < 0.9 * mx) continue;
    const double sh = sector_share(rw, S, L);
    if (sh > bs) { bs = sh; best = L; }
  }
  return best;
}

// Line and layer row strides of the T list (0 where none is found).
// Candidates: the commonest row distances >= 16 between the rows of
// adjacent columns over a window of the list's columns.  The line stride y
// is the lattice shift near the smallest candidate sharing the most
// sectors.  The layer stride is the lattice shift, within a line of the
// (two commonest) candidates of at least four lines, that maps the line
// shift's breaks -- the rows whose shift by y differs from the common
// column shift, at the edges of each layer -- onto themselves: a layer
// repeats them exactly, a diagonal shift (a line over and a layer up,
// equally a lattice shift) does not.
static int bundle_strides(const Plan *p, hipStream_t st, int32_t *Ry, int32_t *Rz) {
  *Ry = *Rz = 0;
  const int32_t w0 = p->Ncols > kBundleWindow ? p->Ncols / 2 - kBundleWindow / 2 : 0;
  const int32_t wn = std::min<int32_t>(kBundleWindow, p->Ncols - w0);
  std::vector<int32_t> cr(wn);
  int rc = (int)hipMemcpyAsync(cr.data(), p->colRowT + w0, sizeof(int32_t) * (size_t)wn, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) return rc;
  std::vector<int32_t> d;
  d.reserve(wn);
  for (int32_t i = 0; i + 1 < wn; ++i) {
    if (cr[i] < 0 || cr[i + 1] < 0) continue;
    const int32_t x = std::abs(cr[i + 1] - cr[i]);
    if (x >= 16) d.push_back(x);
  }
  if (d.empty()) return 0;
  std::sort(d.begin(), d.end());
  std::vector<std::pair<int, int32_t>> runs;  // (count, distance)
  for (size_t i = 0; i < d.size();) {
    size_t j = i;
    while (j < d.size() && d[j] == d[i]) ++j;
    runs.push_back({(int)(j - i), d[i]});
    i = j;
  }
  std::stable_sort(runs.begin(), runs.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
  if (runs.size() > 6) runs.resize(6);
  int32_t cmin = runs[0].second, cmax = runs[0].second;
  for (const auto &r : runs) { cmin = std::min(cmin, r.second); cmax = std::max(cmax, r.second); }
  // rows sampled mid-list, with room for the largest shift tried
  const int S = kBundleSample;
  const int64_t need = (int64_t)S + 2 * (int64_t)cmax + 256;
  if (p->NrowsT < 4 * (int64_t)S || need > p->NrowsT) return 0;
  RowSample rw;
  rw.r0 = (int32_t)std::max<int64_t>(0, p->NrowsT / 2 - need / 2);
  rw.n = (int32_t)std::min<int64_t>(need, p->NrowsT - rw.r0 - 1);
  rw.rs.resize(rw.n + 1);
  rc = (int)hipMemcpyAsync(rw.rs.data(), p->rsT + rw.r0, sizeof(int32_t) * (rw.n + 1), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) return rc;
  rw.ci.resize(rw.rs[rw.n] - rw.rs[0]);
  rc = (int)hipMemcpyAsync(rw.ci.data(), p->ciT + rw.rs[0], sizeof(int32_t) * rw.ci.size(), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  if (rc) return rc;
  const int y = lattice_shift(rw, S, cmin, 64, 16);
  if (!y) return 0;
  *Ry = y;
  // the line shift's common column shift (a sample's mode) and its breaks
  std::vector<int64_t> dy(S);
  for (int r = 0; r < S; ++r) dy[r] = (int64_t)rw.first(r + y) - rw.first(r);
  std::sort(dy.begin(), dy.end());
  int64_t Dy = dy[0];
  for (int i = 1, run = 1, best = 1; i < S; ++i) {
    run = dy[i] == dy[i - 1] ? run + 1 : 1;
    if (run > best) { best = run; Dy = dy[i]; }
  }
  const int nb = rw.n - y;
  std::vector<uint8_t> brk(nb);
  for (int r = 0; r < nb; ++r) brk[r] = ((int64_t)rw.first(r + y) - rw.first(r)) != Dy;
  int z = 0, tried = 0;
  int64_t zmis = -1;
  for (const auto &r : runs) {  // by count
    if (r.second < 4 * y || tried == 2) continue;
    ++tried;
    const int a = std::max(4 * y, r.second - y - 64), b = r.second + y + 64;
    std::vector<double> fr(b - a + 1);
    double mx = 0.0;
    for (int L = a; L <= b; ++L) mx = std::max(mx, fr[L - a] = shift_freq(rw, S, L));
    for (int L = a; L <= b; ++L) {
      if (fr[L - a] < 0.5 || fr[L - a] < 0.9 * mx) continue;
      const int m = nb - L;
      if (m < S) continue;
      int64_t mis = 0;
      for (int i = 0; i < m; ++i) mis += brk[i] != brk[i + L];
      if (zmis < 0 || mis < zmis) { zmis = mis; z = L; }
    }
  }
  *Rz = z;
  return 0;
}

// Times options 0..n-1 of `launch(i)` against each other: `warm` untimed
// launches of option 0 first (a memory-bound kernel on this board runs its
// first launches after a lull faster than its steady state), then `passes`
// rounds, each timing every option (mean of 8 launches after one untimed)
// with the starting option rotated; best[i] = option i's fastest round.
template <class F>
static int race(F launch, int n, int passes, int warm, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float *best) {
  for (int i = 0; i < n; ++i) best[i] = 0.f;
  for (int i = 0; i < warm; ++i) launch(0);
  int rc = (int)hipGetLastError();
  for (int pass = 0; pass < passes && !rc; ++pass) {
    for (int j = 0; j < n && !rc; ++j) {
      const int c = (pass + j) % n;
      float us = 0.f;
      rc = time_launches([&]() { launch(c); }, st, e0, e1, &us);
      if (!rc && (best[c] == 0.f || us < best[c])) best[c] = us;
    }
  }
  return rc;
}

static void free_bundle(Bundle &u) {
  free_delta(u.dp);
  if (u.part.order) (void)hipFree(u.part.order);
  if (u.part.d) (void)hipFree(u.part.d);
  u = Bundle();
}

// One bundle candidate: its partition, the 32-bit pack its delta pack
// borrows until the choice is made, and its block order (tiled when that
// raced 2% faster than block id order).
struct BundleCand {
  Bundle u;
  Pack k;
  const int *order = nullptr;
};

static void free_cand(BundleCand &c) {
  free_bundle(c.u);
  free_pack(c.k);
  c = BundleCand();
}

// Plan::bundle: for a T list of short rows (>= kBundleMinLen entries per row
// on average) whose partition 0 has a delta pack, the line and layer
// strides, then bundles of 4 lines x 2 layers and of 4 lines x 1 layer, each
// cut on entries like partition 0 with a column-sorted pack (a candidate
// only if every block packs) and a delta pack, in the tiled block order for
// its layer distance; the fastest candidate is kept if it races 5% faster
// than partition 0's launch.  It keeps its 32-bit words for its fallback blocks
// only.
static int build_bundle(Plan *p, int32_t *scratch, hipStream_t st) {
  const Partition &p0 = p->partT[0];
  if (!p->dpack.w || !p->colRowT || p->nparts == 0 || p0.rowsInKey || p->NrowsT == 0 ||
      (int64_t)p->nnzT < (int64_t)kBundleMinLen * p->NrowsT || p0.nblocks < 16 * p->cus)
    return 0;
  int32_t Ry = 0, Rz = 0;
  int rc = bundle_strides(p, st, &Ry, &Rz);
  if (rc || !Ry) return rc;
  double *vals = nullptr;
  if (hipMalloc(&vals, sizeof(double) * (size_t)p->Ncols) != hipSuccess) {
    (void)hipGetLastError();  // no room for scratch values: no bundle
    return 0;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = (int)hipMemsetAsync(vals, 0, sizeof(double) * (size_t)p->Ncols, st);
  if (!rc) rc = (int)hipEventCreate(&e0);
  if (!rc) rc = (int)hipEventCreate(&e1);
  constexpr int kWarm = 96, kPasses = 5;
  const int lmax = p->maxRowT;
  const int shapes[2][2] = {{4, 2}, {4, 1}};
  BundleCand cand[2];
  int ncand = 0;
  for (int c = 0; c < 2 && !rc; ++c) {
    const int ty = shapes[c][0], tz = shapes[c][1];
    if (tz > 1 && !Rz) continue;
    BundleCand &x = cand[ncand];
    Bundle &u = x.u;
    u.Ry = Ry;
    u.Rz = Rz ? Rz : p->NrowsT;
    u.ty = ty;
    u.tz = tz;
    int32_t *rsP = nullptr, *ciP = nullptr;
    rc = bundle_csr(p->rsT, p->ciT, p->NrowsT, p->nnzT, u.Ry, u.Rz, ty, tz, st, &rsP, &ciP);
    if (!rc) rc = build_partition(p, rsP, rsP, false, false, p->NrowsT, kCap - (lmax > 0 ? lmax - 1 : 0), scratch, st,
                                  &u.part);
    if (!rc) rc = build_pack(u.part, p->NrowsT, rsP, ciP, false, p->nnzT, scratch, st, &x.k);
    bool ok = !rc && x.k.pk && x.k.plain == 0;
    if (ok) rc = build_delta_pack(p, u.part, x.k, u.dp, scratch, st);
    ok = ok && !rc && u.dp.w;
    if (ok) rc = build_byte_pack(u.part, x.k, u.dp, st);
    int D = 0;
    if (ok) rc = layer_distance(p, u.part, ciP, st, &D);
    if (rsP) (void)hipFree(rsP);
    if (ciP) (void)hipFree(ciP);
    // tiled whenever the layer distance shows: the element layers between
    // two bundles of layers are read and written by blocks D apart, and in
    // block id order half of those lines leave L2 between them (192^3 p=1:
    // reads 1.41x the minimum, half the 64-byte sectors written in halves)
    if (ok && !rc && D) {
      const std::vector<int> ord = tiled_order(u.part.nblocks, D);
      rc = (int)hipMalloc(&u.part.order, sizeof(int) * ord.size());
      if (!rc) rc = (int)hipMemcpyAsync(u.part.order, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice, st);
      if (!rc) rc = (int)hipStreamSynchronize(st);
      if (!rc) {
        u.part.tileD = D;
        u.part.tiled = kTiledGS;
        x.order = u.part.order;
      }
    }
    if (!rc && ok) ++ncand;
    else free_cand(x);
  }
  // partition 0's launch (option 0) against the candidates
  float t[3] = {0.f, 0.f, 0.f};
  const int *o0 = (p0.tiled & kTiledGS) ? p0.order : nullptr;
  if (!rc && ncand)
    rc = race([&](int i) {
      if (i == 0) launch_delta_on<double, OGS_ADD>(p0, p->dpack, p->rsT, p->ciT, p->Ncols, vals, o0, st);
      else launch_delta_on<double, OGS_ADD>(cand[i - 1].u.part, cand[i - 1].u.dp, nullptr, nullptr, p->Ncols, vals,
                                            cand[i - 1].order, st);
    }, 1 + ncand, kPasses, kWarm, st, e0, e1, t);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(vals);
  int pick = -1;
  for (int i = 0; i < ncand && !rc; ++i)
    if (t[i + 1] < 0.95f * t[0] && (pick < 0 || t[i + 1] < t[pick + 1])) pick = i;
  if (pick >= 0) {
    bool done = false;
    rc = compact_fallback(cand[pick].u.dp, cand[pick].u.part.nblocks, scratch, st, &done);
    if (!rc && done) {
      p->bundle = cand[pick].u;
      cand[pick].u = Bundle();  // moved into the plan
    }
  }
  for (int i = 0; i < ncand; ++i) free_cand(cand[i]);
  return rc;
}

// Device bytes a plan owns (arrays shared between its members counted once).
static int64_t plan_bytes(const Plan *p) {
  std::vector<const void *> seen;
  int64_t total = 0;
  auto add = [&](const void *ptr, int64_t bytes) {
    if (!ptr || std::find(seen.begin(), seen.end(), ptr) != seen.end()) return;
    seen.push_back(ptr);
    total += bytes;
  };
  auto part = [&](const Partition &x) {
    add(x.d, (int64_t)(x.nblocks + 1) * (int64_t)sizeof(int4));
    add(x.order, (int64_t)x.nblocks * (int64_t)sizeof(int));
  };
  auto pack = [&](const Pack &k, const Partition &x) {
    add(k.pk, (int64_t)k.stride * x.nblocks * 4);
    add(k.base, (int64_t)x.nblocks * (int64_t)sizeof(int4));
  };
  for (int j = 0; j < p->nparts; ++j) {
    part(p->partT[j]);
    part(p->partN[j]);
    part(p->partS[j]);
  }
  for (int j = 0; j < kPackParts && j < p->nparts; ++j) pack(p->packT[j], p->partT[j]);
  if (p->nparts > 0) {
    pack(p->packNT, p->partT[0]);
    pack(p->packNN, p->partN[0]);
    pack(p->packS, p->partS[0]);
  }
  for (int i = 0; i < 2; ++i) {
    part(p->part3[i]);
    pack(p->pack3[i], p->part3[i]);
  }
  auto delta = [&](const DeltaPack &d, const Partition &x) {
    if (!d.w) return;
    add(d.w, (int64_t)x.nblocks * kCap * 2);
    add(d.x, (int64_t)x.nblocks * kDRanges * d.E * 4);
    if (d.ownsF) {
      add(d.f.pk, (int64_t)d.f.stride * d.fallback * 4);
      add(d.f.base, (int64_t)d.fallback * (int64_t)sizeof(int4));
    }
    add(d.bw, (int64_t)x.nblocks * kCap);
    add(d.bh, (int64_t)x.nblocks * kDRanges * 4);
    add(d.bx, d.bItems * 8);
    add(d.bt, 256 * 4);
  };
  delta(p->dpack, p->partT[0]);
  if (p->bundle.part.d) {
    part(p->bundle.part);
    delta(p->bundle.dp, p->bundle.part);
  }
  for (const RowMap *m : {&p->rowMapT, &p->rowMapN}) {
    add(m->fbRows, (int64_t)m->fallback * kRowChunk * 4);
    add(m->codes, (int64_t)m->chunks * kBlock * 4);
    add(m->x, (int64_t)m->chunks * kDRanges * m->E * 4);
    add(m->table, 64);
  }
  add(p->colRowT, (int64_t)p->Ncols * 4);
  add(p->colRowN, (int64_t)p->Ncols * 4);
  add(p->colOwner, (int64_t)p->Ncols * 4);
  add(p->nMask, ((int64_t)p->Ncols + 31) / 32 * 4);
  return total;
}

static void free_plan(Plan *p) {
  const bool nt = p->packNT.pk && p->packNT.pk != p->packT[0].pk;
  const bool nn = p->packNN.pk && p->packNN.pk != p->packT[0].pk;
  if (nt) free_pack(p->packNT);
  if (nn) free_pack(p->packNN);
  if (p->packS.pk && p->packS.pk != p->packT[0].pk) free_pack(p->packS);
  for (int j = 0; j < kPackParts; ++j) free_pack(p->packT[j]);
  free_delta(p->dpack);
  free_bundle(p->bundle);
  for (int i = 0; i < 2; ++i) {
    free_pack(p->pack3[i]);
    if (p->part3[i].d) (void)hipFree(p->part3[i].d);
  }
  if (p->colOwner) (void)hipFree(p->colOwner);
  if (p->nMask) (void)hipFree(p->nMask);
  if (p->rowMapN.codes && p->rowMapN.codes != p->rowMapT.codes) free_row_map(p->rowMapN);
  free_row_map(p->rowMapT);
  if (p->colRowN && p->colRowN != p->colRowT) (void)hipFree(p->colRowN);
  if (p->colRowT) (void)hipFree(p->colRowT);
  if (p->partT[0].order) (void)hipFree(p->partT[0].order);
  if (p->partS[0].order && p->partS[0].order != p->partT[0].order) (void)hipFree(p->partS[0].order);
  for (int j = 0; j < p->nparts; ++j) {
    if (p->partT[j].d) (void)hipFree(p->partT[j].d);
    if (p->partN[j].d && p->partN[j].d != p->partT[j].d) (void)hipFree(p->partN[j].d);
    if (p->partS[j].d && p->partS[j].d != p->partT[j].d) (void)hipFree(p->partS[j].d);
  }
  delete p;
}

}  // namespace ogs_hip

using namespace ogs_hip;

extern "C" {

const char *ogs_hip_error_string(int code) {
  switch (code) {
    case OGS_SUCCESS: return "success";
    case OGS_E_ARG: return "invalid argument";
    case OGS_E_ROW_TOO_LONG: return "row multiplicity too large";
    case OGS_E_GATHER_UNDEF: return "gather operation not well-defined";
    case OGS_E_INVALID_SETUP: return "invalid ogs setup requested";
    case OGS_E_HALO_POSITIVE: return "halo group without exactly one positive id";
    case OGS_E_NOT_SUPPORTED: return "type/op combination not supported";
    case OGS_E_CSR: return "CSR arrays inconsistent with the given sizes";
    case OGS_E_NO_DEVICE: return "no HIP device";
    case OGS_E_TOO_LARGE: return "input exceeds the fixed scratch of the no-allocation entry point";
    default: return code > 0 ? hipGetErrorString((hipError_t)code) : "unknown error";
  }
}

int ogs_hip_plan_create(ogs_hip_plan_t *plan, int32_t Ncols, int32_t NrowsN, int32_t NrowsT,
                        const int32_t *d_rowStartsN, const int32_t *d_colIdsN,
                        const int32_t *d_rowStartsT, const int32_t *d_colIdsT, void *stream) {
  return ogs_hip_plan_create_k(plan, Ncols, NrowsN, NrowsT, d_rowStartsN, d_colIdsN, d_rowStartsT, d_colIdsT,
                               OGS_PLAN_ALL_K, stream);
}

int ogs_hip_plan_create_k(ogs_hip_plan_t *plan, int32_t Ncols, int32_t NrowsN, int32_t NrowsT,
                          const int32_t *d_rowStartsN, const int32_t *d_colIdsN,
                          const int32_t *d_rowStartsT, const int32_t *d_colIdsT, int fields, void *stream) {
  if (!plan) return OGS_E_ARG;
  *plan = nullptr;
  if (Ncols < 0 || NrowsN < 0 || NrowsT < 0 || NrowsN > NrowsT) return OGS_E_ARG;
  if (fields != OGS_PLAN_ALL_K && fields != OGS_PLAN_K1) return OGS_E_ARG;
  const bool k1 = fields == OGS_PLAN_K1;
  if (!d_rowStartsN || !d_rowStartsT) return OGS_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Plan *p = new (std::nothrow) Plan();
  if (!p) return OGS_E_ARG;
  p->Ncols = Ncols;
  p->NrowsN = NrowsN;
  p->NrowsT = NrowsT;
  p->rsN = d_rowStartsN;
  p->ciN = d_colIdsN;
  p->rsT = d_rowStartsT;
  p->ciT = d_colIdsT;
  p->aliased = (d_rowStartsN == d_rowStartsT && d_colIdsN == d_colIdsT);
  int rc = (int)hipGetDevice(&p->device);
  if (!rc) rc = (int)hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device);
  if (rc) { delete p; return rc; }
  if (p->cus < 1) p->cus = 1;

  int32_t t0 = 0, tN = 0, n0 = 0, nN = 0;
  if ((rc = read_int(d_rowStartsT, st, &t0)) || (rc = read_int(d_rowStartsT + NrowsT, st, &tN)) ||
      (rc = read_int(d_rowStartsN, st, &n0)) || (rc = read_int(d_rowStartsN + NrowsT, st, &nN))) {
    delete p;
    return rc;
  }
  if (t0 != 0 || n0 != 0 || tN < 0 || nN < 0) { delete p; return OGS_E_CSR; }
  // rows [NrowsN, NrowsT) of the N list are empty (the whole-partition
  // kernels of an aliased operator rely on it)
  if (NrowsN < NrowsT) {
    int32_t nR = 0;
    if ((rc = read_int(d_rowStartsN + NrowsN, st, &nR))) { delete p; return rc; }
    if (nR != nN) { delete p; return OGS_E_CSR; }
  }
  p->nnzT = tN;
  p->nnzN = nN;
  if ((p->nnzT > 0 && !d_colIdsT) || (p->nnzN > 0 && !d_colIdsN)) { delete p; return OGS_E_ARG; }

  int32_t *scratch = nullptr;
  if ((rc = (int)hipMalloc(&scratch, 4 * sizeof(int32_t)))) { delete p; return rc; }
  int32_t flags = 0, fl2 = 0;
  rc = scan_csr(p->rsT, NrowsT, p->ciT, p->nnzT, Ncols, scratch, st, &p->maxRowT, &flags);
  if (!rc) {
    if (p->aliased) { p->maxRowN = p->maxRowT; }
    else rc = scan_csr(p->rsN, NrowsT, p->ciN, p->nnzN, Ncols, scratch, st, &p->maxRowN, &fl2);
  }
  if (!rc && ((flags | fl2) & 3)) rc = OGS_E_CSR;
  // column -> row maps of the dense lists, for the by-column scatter
  if (!rc && Ncols > 0 && 2 * (int64_t)p->nnzT >= Ncols)
    rc = build_col_rows(p->rsT, NrowsT, p->ciT, Ncols, st, &p->colRowT);
  if (!rc) {
    if (p->aliased) p->colRowN = p->colRowT;
    else if (Ncols > 0 && 2 * (int64_t)p->nnzN >= Ncols)
      rc = build_col_rows(p->rsN, NrowsT, p->ciN, Ncols, st, &p->colRowN);
  }
  // their row-delta codes, for Scatter at K = 1, on lists of fewer than two
  // entries per row (mostly single-member rows: the Scatter streams its row
  // values, and the int32 map was a fifth of its bytes: 48^3 p=7 fp64
  // 189 -> 165 us, 55^3 p=6 192 -> 183 us; with rows of 4-8 members the
  // row values are gathered and the decode, in front of those gathers, made
  // 192^3 p=1 149 -> 167 us)
  if (!rc && p->colRowT && (int64_t)p->nnzT < 2 * (int64_t)NrowsT)
    rc = build_row_map(p->colRowT, Ncols, p->cus, scratch, st, p->rowMapT);
  if (!rc && p->colRowN) {
    if (p->colRowN == p->colRowT) p->rowMapN = p->rowMapT;
    else if ((int64_t)p->nnzN < 2 * (int64_t)NrowsN) rc = build_row_map(p->colRowN, Ncols, p->cus, scratch, st, p->rowMapN);
  }
  // the N list as a column mask: GatherScatter Trans over the T list
  if (!rc && !p->aliased && p->colRowT && p->nnzT > 0) rc = build_n_mask(p, NrowsT, scratch, st);
  // the column -> owner map rests on what the mask's check establishes (T
  // columns distinct, every N entry in its own T row): each column of a row
  // then reads the row's one owner, and the owner's own rewrite of its value
  // is idempotent
  if (!rc && !p->aliased && p->nMask && p->maxRowN <= 1) {
    if (hipMalloc(&p->colOwner, sizeof(int32_t) * (size_t)Ncols) != hipSuccess) {
      (void)hipGetLastError();  // no room: the row-based kernels run this GatherScatter
      p->colOwner = nullptr;
    } else {
      hipLaunchKernelGGL(k_col_owner, dim3(grid_for(Ncols, 16384)), dim3(kBlock), 0, st, Ncols, p->colRowT, p->rsN,
                         p->ciN, p->colOwner);
      rc = (int)hipGetLastError();
    }
  }
  // T lists without empty rows (every operator that Setup builds) are cut on
  // entries alone: measured faster than merge-path blocks of the same width
  // for GatherScatter and Gather (more entries per workgroup, fewer blocks).
  // The row-wise scatter kernels (lists too sparse to scatter by columns)
  // keep merge-path partitions (partS), measured faster for them.  N lists
  // may have empty rows and are cut by the merge path.
  const bool cutTOnEntries = (flags & 4) == 0;
  const int lmax = p->maxRowT > p->maxRowN ? p->maxRowT : p->maxRowN;
  if (!rc && lmax > kCap) rc = OGS_E_ROW_TOO_LONG;
  // Partition widths: block nonzeros stay <= kCap >> j.
  for (int j = 0; !rc && j < kMaxParts && NrowsT > 0; ++j) {
    const int W = (kCap >> j) - (lmax > 0 ? lmax - 1 : 0);
    if (W < 1) break;
    rc = build_partition(p, p->rsT, p->rsN, false, !cutTOnEntries, NrowsT, W, scratch, st, &p->partT[j]);
    if (!rc) {
      if (p->aliased) p->partN[j] = p->partT[j];
      else rc = build_partition(p, p->rsT, p->rsN, true, true, NrowsN, W, scratch, st, &p->partN[j]);
    }
    if (!rc) {
      if (cutTOnEntries && !p->colRowT) rc = build_partition(p, p->rsT, p->rsN, false, true, NrowsT, W, scratch, st, &p->partS[j]);
      else p->partS[j] = p->partT[j];
    }
    p->nparts = j + 1;
  }
  // column-sorted packs of the K=1 partitions (a K = 1 plan: only those its
  // launches read -- the partition the one-list kernels run on, and for a
  // Signed operator the half-width one of the two-list kernel)
  const int jK1 = p->nparts > 0 ? (int)(fill_device(p, &p->partT[0]) - p->partT) : 0;
  const int jHalf = p->nparts > 1 ? (int)(fill_device(p, &p->partT[1]) - p->partT) : -1;
  if (!rc && p->nparts > 0) {
    for (int j = 0; j < kPackParts && j < p->nparts && !rc; ++j) {
      if (k1 && j != jK1 && (p->aliased || j != jHalf)) continue;
      rc = build_pack(p->partT[j], NrowsT, p->rsT, p->ciT, false, p->nnzT, scratch, st, &p->packT[j]);
    }
    if (!rc && p->aliased) {
      p->packNT = p->packT[0];
      p->packNN = p->packT[0];
    } else if (!rc) {
      rc = build_pack(p->partT[0], NrowsT, p->rsN, p->ciN, true, p->nnzN, scratch, st, &p->packNT);
      if (!rc) rc = build_pack(p->partN[0], NrowsN, p->rsN, p->ciN, true, p->nnzN, scratch, st, &p->packNN);
    }
    if (!rc) {
      if (p->partS[0].d != p->partT[0].d)
        rc = build_pack(p->partS[0], NrowsT, p->rsT, p->ciT, false, p->nnzT, scratch, st, &p->packS);
      else p->packS = p->packT[0];
    }
    // three-field partitions and packs
    for (int i = 0; i < 2 && !rc && cutTOnEntries && !k1; ++i) {
      const int W = kCap3[i] - (lmax > 0 ? lmax - 1 : 0);
      if (W < 1) continue;
      rc = build_partition(p, p->rsT, p->rsN, false, false, NrowsT, W, scratch, st, &p->part3[i]);
      if (!rc) rc = build_pack(p->part3[i], NrowsT, p->rsT, p->ciT, false, p->nnzT, scratch, st, &p->pack3[i]);
    }
  }
  if (!rc && cutTOnEntries) rc = build_delta_pack(p, p->partT[0], p->packT[0], p->dpack, scratch, st);
  if (!rc && cutTOnEntries) rc = build_byte_pack(p->partT[0], p->packT[0], p->dpack, st);
  if (!rc && cutTOnEntries) rc = tile_schedule(p, st);
  if (!rc && cutTOnEntries) rc = build_bundle(p, scratch, st);
  // a K = 1 Unsigned plan whose K = 1 kernels all read the delta pack
  // (Scatter by columns) keeps only its fallback blocks' 32-bit words of
  // partition 0's pack
  if (!rc && k1 && p->aliased && jK1 == 0 && p->dpack.w && p->colRowT && p->packT[0].pk) {
    bool done = false;
    rc = compact_fallback(p->dpack, p->partT[0].nblocks, scratch, st, &done);
    if (!rc && done) {
      const uint32_t *gone = p->packT[0].pk;
      for (Pack *a : {&p->packNT, &p->packNN, &p->packS})
        if (a->pk == gone) *a = Pack();
      free_pack(p->packT[0]);
    }
  }
  // ... and of its column -> row map only the rows of the row-delta map's
  // fallback chunks (Scatter at K = 1 reads the codes; other field counts
  // run from the CSR)
  if (!rc && k1 && p->aliased && p->rowMapT.codes && p->colRowT) {
    bool done = false;
    rc = compact_rows(p->rowMapT, p->colRowT, p->Ncols, scratch, st, &done);
    if (!rc && done) {
      (void)hipFree(p->colRowT);
      p->colRowT = p->colRowN = nullptr;
      p->rowMapN = p->rowMapT;
    }
  }
  (void)hipFree(scratch);
  if (rc) { free_plan(p); return rc; }
  *plan = p;
  return OGS_SUCCESS;
}

int ogs_hip_plan_retain(ogs_hip_plan_t plan) {
  if (!plan) return OGS_E_ARG;
  plan->refs.fetch_add(1, std::memory_order_relaxed);
  return OGS_SUCCESS;
}

int ogs_hip_plan_destroy(ogs_hip_plan_t plan) {
  if (!plan) return OGS_E_ARG;
  if (plan->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) free_plan(plan);
  return OGS_SUCCESS;
}

int ogs_hip_plan_info(ogs_hip_plan_t p, ogs_hip_plan_info_t *info) {
  if (!p || !info) return OGS_E_ARG;
  info->NrowsN = p->NrowsN;
  info->NrowsT = p->NrowsT;
  info->nnzN = p->nnzN;
  info->nnzT = p->nnzT;
  info->maxRowT = p->maxRowT;
  info->maxRowN = p->maxRowN;
  info->nblocksK1 = p->nparts ? p->partT[0].nblocks : 0;
  info->widthK1 = p->nparts ? p->partT[0].W : 0;
  info->packedK1 = p->packT[0].packed;
  info->plainK1 = p->packT[0].plain;
  info->Ncols = p->Ncols;
  info->tileD = p->nparts ? p->partT[0].tileD : 0;
  info->colsT = p->colRowT != nullptr || p->rowMapT.codes != nullptr;
  info->colsN = p->colRowN != nullptr || p->rowMapN.codes != nullptr;
  info->deltaE = p->dpack.w ? p->dpack.E : 0;
  info->deltaFallback = p->dpack.w ? p->dpack.fallback : 0;
  info->ownerMap = p->colOwner != nullptr;
  info->nMask = p->nMask != nullptr;
  const Bundle &u = p->bundle;
  info->bundleRy = u.part.d ? u.Ry : 0;
  info->bundleRz = u.part.d ? u.Rz : 0;
  info->bundleTy = u.part.d ? u.ty : 0;
  info->bundleTz = u.part.d ? u.tz : 0;
  info->bundleTileD = u.part.d ? u.part.tileD : 0;
  info->bundleFallback = u.part.d ? u.dp.fallback : 0;
  info->rowMapE = p->rowMapT.codes ? p->rowMapT.E : 0;
  info->rowMapFallback = p->rowMapT.codes ? p->rowMapT.fallback : 0;
  info->byteItems = p->dpack.bw ? (int32_t)p->dpack.bItems : 0;
  info->byteFallback = p->dpack.bw ? p->dpack.bFallback : 0;
  const Partition *bp = nullptr;
  info->byteK3 = byte_kv_pack(p, &bp) != nullptr;
  info->bytes = plan_bytes(p);
  return OGS_SUCCESS;
}

int ogs_hip_gather_scatter(ogs_hip_plan_t p, int type, int op, int trans, int32_t K, void *d_v, void *stream) {
  if (!p || bad_type(type) || bad_op(op) || bad_trans(trans) || K < 1) return OGS_E_ARG;
  if (not_offered(type, op)) return OGS_E_NOT_SUPPORTED;
  if (!d_v && p->Ncols > 0) return OGS_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (type) {
    case OGS_FLOAT: return launch_gs_typed<float>(p, op, trans, K, d_v, st);
    case OGS_DOUBLE: return launch_gs_typed<double>(p, op, trans, K, d_v, st);
    case OGS_INT32: return launch_gs_typed<int32_t>(p, op, trans, K, d_v, st);
    case OGS_INT64: return launch_gs_typed<int64_t>(p, op, trans, K, d_v, st);
    case OGS_BF16: return launch_gs_typed<bf16s>(p, op, trans, K, d_v, st);
    default: return launch_gs_typed<f32c>(p, op, trans, K, d_v, st);
  }
}

int ogs_hip_gather(ogs_hip_plan_t p, int type, int op, int trans, int32_t K, void *d_gv, const void *d_v, void *stream) {
  if (!p || bad_type(type) || bad_op(op) || bad_trans(trans) || K < 1) return OGS_E_ARG;
  if (not_offered(type, op)) return OGS_E_NOT_SUPPORTED;
  if ((!d_v && p->Ncols > 0) || (!d_gv && p->NrowsT > 0)) return OGS_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (type) {
    case OGS_FLOAT: return launch_gather_typed<float>(p, op, trans, K, d_gv, d_v, st);
    case OGS_DOUBLE: return launch_gather_typed<double>(p, op, trans, K, d_gv, d_v, st);
    case OGS_INT32: return launch_gather_typed<int32_t>(p, op, trans, K, d_gv, d_v, st);
    case OGS_INT64: return launch_gather_typed<int64_t>(p, op, trans, K, d_gv, d_v, st);
    case OGS_BF16: return launch_gather_typed<bf16s>(p, op, trans, K, d_gv, d_v, st);
    default: return launch_gather_typed<f32c>(p, op, trans, K, d_gv, d_v, st);
  }
}

int ogs_hip_scatter(ogs_hip_plan_t p, int type, int trans, int32_t K, void *d_v, const void *d_gv, void *stream) {
  if (!p || bad_type(type) || bad_trans(trans) || K < 1) return OGS_E_ARG;
  if ((!d_v && p->Ncols > 0) || (!d_gv && p->NrowsT > 0)) return OGS_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (type) {
    case OGS_FLOAT: return launch_scatter_typed<float>(p, trans, K, d_v, d_gv, st);
    case OGS_DOUBLE: return launch_scatter_typed<double>(p, trans, K, d_v, d_gv, st);
    case OGS_INT32: return launch_scatter_typed<int32_t>(p, trans, K, d_v, d_gv, st);
    case OGS_INT64: return launch_scatter_typed<int64_t>(p, trans, K, d_v, d_gv, st);
    case OGS_BF16: return launch_scatter_typed<bf16s>(p, trans, K, d_v, d_gv, st);
    default: return launch_scatter_typed<f32c>(p, trans, K, d_v, d_gv, st);
  }
}

int ogs_hip_extract(int type, int32_t N, int32_t K, const int32_t *d_ids, const void *d_q, void *d_gq, void *stream) {
  if (bad_type(type) || N < 0 || K < 1) return OGS_E_ARG;
  if (N > 0 && (!d_ids || !d_q || !d_gq)) return OGS_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (type) {
    case OGS_FLOAT: return launch_extract_typed<float>(N, K, d_ids, d_q, d_gq, st);
    case OGS_DOUBLE: return launch_extract_typed<double>(N, K, d_ids, d_q, d_gq, st);
    case OGS_INT32: return launch_extract_typed<int32_t>(N, K, d_ids, d_q, d_gq, st);
    case OGS_INT64: return launch_extract_typed<int64_t>(N, K, d_ids, d_q, d_gq, st);
    case OGS_BF16: return launch_extract_typed<bf16s>(N, K, d_ids, d_q, d_gq, st);
    default: return launch_extract_typed<f32c>(N, K, d_ids, d_q, d_gq, st);
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch construction, tuning and graph capture.

namespace ogs_hip {

struct BatchConfig {
  int narrow;  // K > 1 descriptors use the next narrower block width (0: K = 3
               // descriptors on their three-field partition where the plan has one)
  int order;   // 0: items of larger descriptors first, 1: descriptor order
  int fused;   // 1: all field-count classes in one kernel node
  int plans;   // 1: one launch per descriptor through its plan's own kernels
               // (the one-list GatherScatter Sym that ogs_hip_gather_scatter runs)
};

// Field-count classes with a compile-time shape: K = 1, 2, 3, 4, 8, 16.
constexpr int kPipeClasses = 6;
constexpr int kPipeK[kPipeClasses] = {1, 2, 3, 4, 8, 16};
static int pipe_class(int K) {
  for (int c = 0; c < kPipeClasses; ++c)
    if (kPipeK[c] == K) return c;
  return -1;
}

struct BatchTables {
  std::vector<BatchDev> descs;
  std::vector<int2> cls[kPipeClasses];  // blocks of each shaped class (k_batch_vec<KT>)
  std::vector<int2> items;              // other K, one pass (k_batch_gs)
  std::vector<int2> rest;               // other K, several field passes (k_batch_gs_rest)
};

template <typename T>
static int batch_tables(const std::vector<ogs_hip_batch_desc_t> &in, const std::vector<void *> &qs,
                        const BatchConfig &c, BatchTables &t) {
  t.descs.clear();
  for (auto &v : t.cls) v.clear();
  t.items.clear();
  t.rest.clear();
  std::vector<const Partition *> parts(in.size(), nullptr);
  for (size_t i = 0; i < in.size(); ++i) {
    const Plan *p = in[i].plan;
    const int K = in[i].K;
    BatchDev d = {};
    d.q = qs[i];
    d.K = K;
    d.kc = K;
    d.kmagic = K >= 2 ? (uint32_t)((uint64_t(1) << 32) / (uint64_t)K + 1u) : 0u;
    d.nrows = p->NrowsT;
    const Partition *pt = nullptr;
    if (p->NrowsT > 0 && p->nparts > 0 && p->nnzT > 0) {  // no entries: nothing to reduce or write
      int kc = K;
      if (pipe_class(K) >= 0) {
        pt = c.narrow == 0 ? vec_partition3<T>(p, K) : nullptr;
        if (!pt) pt = vec_partition<T>(p, K, K > 1 ? c.narrow : 0);
      }
      if (!pt) {
        pt = choose(p->partT, p->nparts, K, [](const Partition &x) { return x.maxT; }, &kc);
        if (!pt) return OGS_E_ROW_TOO_LONG;
      }
      d.kc = kc;
      d.part = pt->d;
      d.G = list_view(p, false, pt, false);
    }
    t.descs.push_back(d);
    parts[i] = pt;
  }
  // descriptor order of the item lists: larger work first, or as given
  std::vector<size_t> ord(in.size());
  for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
  if (c.order == 0) {
    auto cost = [&](size_t i) { return (int64_t)in[i].plan->nnzT * in[i].K; };
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
  }
  for (size_t i : ord) {
    const Partition *pt = parts[i];
    if (!pt) continue;
    const int K = in[i].K, cl = pipe_class(K);
    const bool shaped = cl >= 0 && t.descs[i].kc == K && pt->maxT <= vec_cap_entries<T>(K);
    std::vector<int2> &dst = shaped ? t.cls[cl] : t.descs[i].kc == K ? t.items : t.rest;
    for (int b = 0; b < pt->nblocks; ++b) dst.push_back(make_int2((int)i, b));
  }
  return OGS_SUCCESS;
}

// Device copies of a batch's tables.
struct BatchDevTables {
  BatchDev *d = nullptr;
  int2 *lists = nullptr;  // the class lists, then items, then rest
  int n[kPipeClasses + 2] = {};
  size_t off[kPipeClasses + 2] = {};
  void release() {
    if (d) (void)hipFree(d);
    if (lists) (void)hipFree(lists);
    d = nullptr;
    lists = nullptr;
    for (int c = 0; c < kPipeClasses + 2; ++c) n[c] = 0, off[c] = 0;
  }
  int total() const {
    int s = 0;
    for (int c = 0; c < kPipeClasses + 2; ++c) s += n[c];
    return s;
  }
};

template <typename T, int OP, int KT>
static void launch_batch_class(const BatchDev *d, const int2 *items, int n, hipStream_t st) {
  hipLaunchKernelGGL((k_batch_vec<T, OP, KT>), dim3(n), dim3(kBlock), 0, st, d, items);
}

template <typename T, int OP>
static void launch_batch_op(const BatchDevTables &t, int cus, int fused, hipStream_t st) {
  const int2 *L = t.lists;
  int nshaped = 0;
  for (int c = 0; c < kPipeClasses; ++c) nshaped += t.n[c];
  if (fused && nshaped) {
    // the class lists are contiguous in t.lists, in class order
    ClassEnds ce;
    int e = 0;
    for (int c = 0; c < kPipeClasses; ++c) ce.end[c] = (e += t.n[c]);
    hipLaunchKernelGGL((k_batch_fused<T, OP>), dim3(nshaped), dim3(kBlock), 0, st, t.d, L, ce);
  }
  if (!fused && t.n[0]) launch_batch_class<T, OP, 1>(t.d, L + t.off[0], t.n[0], st);
  if (!fused && t.n[1]) launch_batch_class<T, OP, 2>(t.d, L + t.off[1], t.n[1], st);
  if (!fused && t.n[2]) launch_batch_class<T, OP, 3>(t.d, L + t.off[2], t.n[2], st);
  if (!fused && t.n[3]) launch_batch_class<T, OP, 4>(t.d, L + t.off[3], t.n[3], st);
  if (!fused && t.n[4]) launch_batch_class<T, OP, 8>(t.d, L + t.off[4], t.n[4], st);
  if (!fused && t.n[5]) launch_batch_class<T, OP, 16>(t.d, L + t.off[5], t.n[5], st);
  // other K: persistent grid-stride kernels over their items
  static const int occ = occupancy_of(k_batch_gs<T, OP>);
  const int g = occ * cus;
  const int n6 = t.n[kPipeClasses], n7 = t.n[kPipeClasses + 1];
  if (n6) hipLaunchKernelGGL((k_batch_gs<T, OP>), dim3(g < n6 ? g : n6), dim3(kBlock), 0, st, t.d,
                             L + t.off[kPipeClasses], n6);
  if (n7) hipLaunchKernelGGL((k_batch_gs_rest<T, OP>), dim3(g < n7 ? g : n7), dim3(kBlock), 0, st, t.d,
                             L + t.off[kPipeClasses + 1], n7);
}

template <typename T>
static int launch_batch_typed(int op, const BatchDevTables &t, int cus, int fused, hipStream_t st) {
  switch (op) {
    case OGS_ADD: launch_batch_op<T, OGS_ADD>(t, cus, fused, st); break;
    case OGS_MUL:
      if constexpr (kOffered<T, OGS_MUL>) launch_batch_op<T, OGS_MUL>(t, cus, fused, st);
      else return OGS_E_NOT_SUPPORTED;
      break;
    case OGS_MAX: launch_batch_op<T, OGS_MAX>(t, cus, fused, st); break;
    case OGS_MIN: launch_batch_op<T, OGS_MIN>(t, cus, fused, st); break;
    default: return OGS_E_ARG;
  }
  return (int)hipGetLastError();
}

template <typename F>
static int by_storage(int type, F &&f) {
  switch (type) {
    case OGS_FLOAT: return f(float{});
    case OGS_DOUBLE: return f(double{});
    case OGS_INT32: return f(int32_t{});
    case OGS_INT64: return f(int64_t{});
    case OGS_BF16: return f(bf16s{});
    case OGS_FLOAT_C: return f(f32c{});
    default: return OGS_E_ARG;
  }
}

static int launch_batch(int type, int op, const BatchDevTables &t, int cus, int fused, hipStream_t st) {
  return by_storage(type, [&](auto tag) { return launch_batch_typed<decltype(tag)>(op, t, cus, fused, st); });
}

static int make_tables(int type, const std::vector<ogs_hip_batch_desc_t> &in, const std::vector<void *> &qs,
                       const BatchConfig &c, BatchTables &t) {
  return by_storage(type, [&](auto tag) { return batch_tables<decltype(tag)>(in, qs, c, t); });
}

static int upload_tables(const BatchTables &t, hipStream_t st, BatchDevTables &out) {
  out.release();
  const std::vector<int2> *src[kPipeClasses + 2];
  for (int c = 0; c < kPipeClasses; ++c) src[c] = &t.cls[c];
  src[kPipeClasses] = &t.items;
  src[kPipeClasses + 1] = &t.rest;
  size_t total = 0;
  for (int c = 0; c < kPipeClasses + 2; ++c) {
    out.off[c] = total;
    out.n[c] = (int)src[c]->size();
    total += src[c]->size();
  }
  hipError_t e = hipMalloc(&out.d, sizeof(BatchDev) * std::max<size_t>(t.descs.size(), 1));
  if (e == hipSuccess) e = hipMalloc(&out.lists, sizeof(int2) * std::max<size_t>(total, 1));
  if (e == hipSuccess && !t.descs.empty())
    e = hipMemcpyAsync(out.d, t.descs.data(), sizeof(BatchDev) * t.descs.size(), hipMemcpyHostToDevice, st);
  for (int c = 0; c < kPipeClasses + 2 && e == hipSuccess; ++c)
    if (out.n[c])
      e = hipMemcpyAsync(out.lists + out.off[c], src[c]->data(), sizeof(int2) * src[c]->size(),
                         hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return (int)e;
}

}  // namespace ogs_hip

struct ogs_hip_batch_s {
  int type = 0, op = 0;
  std::vector<ogs_hip_plan_t> plans;  // retained while the batch lives
  ogs_hip::BatchDevTables tables;
  ogs_hip::BatchConfig cfg = {0, 1, 1, 0};
  int nplans = 0;     // descriptors launched through their plans (cfg.plans)
  int planItems = 0;  // their blocks (partition 0 or the bundles)
  int cus = 1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int candidates = 0;
  float tune_ms = 0.f, best_us = 0.f;
};

namespace ogs_hip {

static size_t type_size(int type) {
  return type == OGS_BF16 ? 2 : type == OGS_FLOAT || type == OGS_INT32 || type == OGS_FLOAT_C ? 4 : 8;
}

static void free_batch(ogs_hip_batch_s *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->tables.release();
  for (ogs_hip_plan_t p : b->plans) (void)ogs_hip_plan_destroy(p);
  delete b;
}

// One GatherScatter Sym launch per descriptor, each through its plan's own
// kernel choice.
static int launch_plans(int type, int op, const std::vector<ogs_hip_batch_desc_t> &in,
                        const std::vector<void *> &qs, hipStream_t st) {
  int rc = 0;
  for (size_t i = 0; i < in.size() && !rc; ++i)
    rc = ogs_hip_gather_scatter(in[i].plan, type, op, OGS_SYM, in[i].K, qs[i], st);
  return rc;
}

// The descriptors a configuration launches through their plans (P) and
// through the batched kernels (B).  (Per-plan launches for the descriptors
// whose plans run byte-coded kernels, the rest batched, was raced too and
// never picked on the 48-mesh batch: 0.98 ms either way.)
struct BatchSplit {
  std::vector<ogs_hip_batch_desc_t> inP, inB;
  std::vector<void *> qsP, qsB;
};
static void split_batch(const BatchConfig &c, const std::vector<ogs_hip_batch_desc_t> &in,
                        const std::vector<void *> &qs, BatchSplit &sp) {
  for (size_t i = 0; i < in.size(); ++i) {
    const bool viaPlan = c.plans == 1;
    (viaPlan ? sp.inP : sp.inB).push_back(in[i]);
    (viaPlan ? sp.qsP : sp.qsB).push_back(qs[i]);
  }
}

// One run of a configuration: the per-plan launches, then the batched kernels.
static int launch_split(int type, int op, const BatchSplit &sp, const BatchDevTables &dt, int cus, int fused,
                        hipStream_t st) {
  int rc = sp.inP.empty() ? 0 : launch_plans(type, op, sp.inP, sp.qsP, st);
  if (!rc && !sp.inB.empty()) rc = launch_batch(type, op, dt, cus, fused, st);
  return rc;
}

// Times a configuration on scratch buffers: 1 warm-up, kTuneLaunches timed launches.
constexpr int kTuneLaunches = 5;  // timed launches per candidate and pass

static int time_config(int type, int op, const std::vector<ogs_hip_batch_desc_t> &in, const std::vector<void *> &qs,
                       const BatchConfig &c, int cus, hipStream_t st, float *us) {
  BatchSplit sp;
  split_batch(c, in, qs, sp);
  BatchTables t;
  int rc = make_tables(type, sp.inB, sp.qsB, c, t);
  if (rc) return rc;
  BatchDevTables dt;
  rc = upload_tables(t, st, dt);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = (int)hipEventCreate(&e0);
  if (!rc) rc = (int)hipEventCreate(&e1);
  auto launch = [&]() { return launch_split(type, op, sp, dt, cus, c.fused, st); };
  if (!rc) rc = launch();
  if (!rc) rc = (int)hipEventRecord(e0, st);
  for (int i = 0; i < kTuneLaunches && !rc; ++i) rc = launch();
  if (!rc) rc = (int)hipEventRecord(e1, st);
  if (!rc) rc = (int)hipEventSynchronize(e1);
  float ms = 0.f;
  if (!rc) rc = (int)hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dt.release();
  *us = ms * 1e3f / kTuneLaunches;
  return rc;
}

}  // namespace ogs_hip

extern "C" {

int ogs_hip_batch_create(ogs_hip_batch_t *batch, int type, int op, const ogs_hip_batch_desc_t *descs, int32_t n,
                         int tune, void *stream) {
  using namespace ogs_hip;
  if (!batch) return OGS_E_ARG;
  *batch = nullptr;
  if (bad_type(type) || bad_op(op) || n < 0 || (n > 0 && !descs)) return OGS_E_ARG;
  if (not_offered(type, op)) return OGS_E_NOT_SUPPORTED;
  std::vector<ogs_hip_batch_desc_t> in(descs, descs + n);
  std::vector<void *> qs(n);
  bool anyKv = false;
  for (int i = 0; i < n; ++i) {
    if (!in[i].plan || in[i].K < 1 || (!in[i].d_v && in[i].plan->Ncols > 0)) return OGS_E_ARG;
    qs[i] = in[i].d_v;
    anyKv |= in[i].K > 1;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  ogs_hip_batch_s *b = new (std::nothrow) ogs_hip_batch_s();
  if (!b) return (int)hipErrorOutOfMemory;
  b->type = type;
  b->op = op;
  int dev = 0, cus = 256;
  int rc = (int)hipGetDevice(&dev);
  if (!rc) rc = (int)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus < 1) cus = 1;
  b->cus = cus;
  BatchConfig best = {0, 1, 1, 0};
  const auto t0 = std::chrono::steady_clock::now();
  if (!rc && tune && n > 0) {
    // scratch copies: tuning launches must not touch the caller's data
    std::vector<void *> scratch(n, nullptr);
    for (int i = 0; i < n && !rc; ++i) {
      const size_t bytes = (size_t)in[i].plan->Ncols * in[i].K * type_size(type);
      if (!bytes) continue;
      rc = (int)hipMalloc(&scratch[i], bytes);
      if (!rc) rc = (int)hipMemcpyAsync(scratch[i], in[i].d_v, bytes, hipMemcpyDeviceToDevice, st);
    }
    // two passes, each candidate keeping its faster time: the first launches
    // after a setup run slow, and one pass of a few launches can mis-rank
    std::vector<BatchConfig> cands;
    for (int fused = 0; fused <= 1; ++fused)
      for (int narrow = 0; narrow <= (anyKv ? 1 : 0); ++narrow)
        for (int order = 0; order <= 1; ++order) cands.push_back({narrow, order, fused, 0});
    cands.push_back({0, 1, 0, 1});
    std::vector<float> times(cands.size(), 0.f);
    for (int pass = 0; pass < 2 && !rc; ++pass) {
      for (size_t i = 0; i < cands.size() && !rc; ++i) {
        float us = 0.f;
        rc = time_config(type, op, in, scratch, cands[i], cus, st, &us);
        if (!rc && (times[i] == 0.f || us < times[i])) times[i] = us;
      }
    }
    b->candidates = (int)cands.size();
    float bestUs = 0.f;
    for (size_t i = 0; i < cands.size() && !rc; ++i) {
      if (bestUs == 0.f || times[i] < bestUs) {
        bestUs = times[i];
        best = cands[i];
      }
    }
    for (void *q : scratch)
      if (q) (void)hipFree(q);
    b->best_us = bestUs;
  }
  b->tune_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  b->cfg = best;
  BatchTables t;
  BatchSplit sp;
  split_batch(best, in, qs, sp);
  b->nplans = (int)sp.inP.size();
  for (const ogs_hip_batch_desc_t &d : sp.inP) {
    const Plan *p = d.plan;
    b->planItems += p->bundle.part.d ? p->bundle.part.nblocks : p->nparts ? p->partT[0].nblocks : 0;
  }
  if (!rc) rc = make_tables(type, sp.inB, sp.qsB, best, t);
  if (!rc) rc = upload_tables(t, st, b->tables);
  // capture the launch in a graph (capture needs a stream of its own)
  hipStream_t cs = nullptr;
  if (!rc) rc = (int)hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (!rc) rc = (int)hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (!rc) {
    const int lr = launch_split(type, op, sp, b->tables, cus, best.fused, cs);
    const int er = (int)hipStreamEndCapture(cs, &b->graph);
    rc = lr ? lr : er;
  }
  if (!rc) rc = (int)hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (cs) (void)hipStreamDestroy(cs);
  if (rc) {
    free_batch(b);
    return rc;
  }
  // the graph reads the plans' arrays: hold each plan once
  for (int i = 0; i < n; ++i) {
    if (std::find(b->plans.begin(), b->plans.end(), in[i].plan) != b->plans.end()) continue;
    (void)ogs_hip_plan_retain(in[i].plan);
    b->plans.push_back(in[i].plan);
  }
  *batch = b;
  return OGS_SUCCESS;
}

int ogs_hip_batch_run(ogs_hip_batch_t batch, void *stream) {
  if (!batch || !batch->exec) return OGS_E_ARG;
  return (int)hipGraphLaunch(batch->exec, static_cast<hipStream_t>(stream));
}

int ogs_hip_batch_destroy(ogs_hip_batch_t batch) {
  if (!batch) return OGS_E_ARG;
  ogs_hip::free_batch(batch);
  return OGS_SUCCESS;
}

int ogs_hip_batch_info(ogs_hip_batch_t batch, ogs_hip_batch_info_t *info) {
  if (!batch || !info) return OGS_E_ARG;
  info->items = batch->tables.total() + batch->planItems;
  int nodes = 0, shaped = 0;
  for (int c = 0; c < ogs_hip::kPipeClasses; ++c) shaped += batch->tables.n[c] > 0;
  for (int c = ogs_hip::kPipeClasses; c < ogs_hip::kPipeClasses + 2; ++c) nodes += batch->tables.n[c] > 0;
  info->nodes = batch->nplans + nodes + (batch->cfg.fused ? (shaped > 0) : shaped);
  info->plans = batch->cfg.plans;
  info->narrow = batch->cfg.narrow;
  info->order = batch->cfg.order;
  info->fused = batch->cfg.fused;
  info->candidates = batch->candidates;
  info->tune_ms = batch->tune_ms;
  info->best_us = batch->best_us;
  return OGS_SUCCESS;
}

}  // extern "C"
