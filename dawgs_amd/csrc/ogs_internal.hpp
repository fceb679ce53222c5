// ogs_internal.hpp -- private layout of an operator plan and the kernel
// tiling constants shared by the gfx950 kernels and their launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>

namespace ogs_hip {

constexpr int kBlock = 256;                                     // lanes per workgroup (4 waves of 64)
constexpr int kCap = 2048;                                      // values staged in LDS per list per pass
constexpr int kSlots = kCap / kBlock;                           // staged values per lane
constexpr int kWMax = kCap;                                     // rows per block: <= W (merge path) or <= nonzeros <= kCap
constexpr int kRowSlots = (kWMax + 1 + kBlock - 1) / kBlock;    // row starts per lane
constexpr int kNumXcd = 8;                                      // MI355X: 8 XCDs x 32 CUs
constexpr int kMaxParts = 7;                                    // partition widths kCap>>0 .. kCap>>6

// Column-sorted pack word: segment (2 bits) | column offset from the
// segment's base (18 bits) | row-head bit (the entry is the first of its row
// in CSR order) | CSR position within the block (11 bits).
constexpr int kPackPosBits = 11;
constexpr int kPackHeadBit = kPackPosBits;
constexpr int kPackOffShift = kPackPosBits + 1;
constexpr int kPackOffBits = 18;
constexpr int kPackSegShift = kPackOffShift + kPackOffBits;
constexpr uint32_t kPackPosMask = (1u << kPackPosBits) - 1;
constexpr uint32_t kPackOffMask = (1u << kPackOffBits) - 1;
static_assert(kPackSegShift == 30, "pack word is 2 + 18 + 1 + 11 bits");
constexpr int kPackSegs = 4;
constexpr int kPackParts = 5;  // T-list packs for partitions 0..4 (K up to 16 in one pass)
// Entries per block of the three-field partitions: a block stages 16 KiB of
// values in LDS, three per entry, of 4 bytes (float, int, compensated float,
// bfloat16 widened to float) or 8 bytes (double, long long).
constexpr int kCap3[2] = {16384 / (4 * 3), 16384 / (8 * 3)};
static_assert(kCap <= (1 << kPackPosBits), "CSR positions of a block must fit the pack word");

// One row partition: nblocks+1 entries {row, rowStartsT[row],
// rowStartsN[row], 0}; block b covers rows [part[b].x, part[b+1].x).
struct Partition {
  int4 *d = nullptr;
  int32_t nblocks = 0;
  int32_t W = 0;        // keys per block: nonzeros, plus rows when rowsInKey
  bool rowsInKey = true;  // merge path (a list with empty rows)
  int32_t maxT = 0;     // largest T-list nonzero count of a block
  int32_t maxN = 0;     // largest N-list nonzero count of a block
  int32_t maxRows = 0;  // largest row count of a block
  int *order = nullptr;  // tiled block run order (null: block id order)
  int32_t tileD = 0;     // layer distance (blocks) the order is tiled for, 0: none
  int32_t tiled = 0;     // kernels that run in `order`: kTiledGS | kTiledGather | kTiledScatter
};

constexpr int kTiledGS = 1, kTiledGather = 2, kTiledScatter = 4;  // Partition::tiled bits

// A list's entries in the blocks of one partition, each block's entries
// sorted by column and packed; base[b] holds the block's segment bases, with
// base[b].x < 0 for a block left in CSR order (more than kPackSegs segments:
// its words are then its column ids in CSR order).  Block b's words start at
// b * stride (stride >= the partition's largest block, zero-padded), so a
// kernel can load them knowing only b.
struct Pack {
  uint32_t *pk = nullptr;
  int4 *base = nullptr;
  int32_t stride = 0;  // words per block
  int32_t packed = 0;  // blocks packed
  int32_t plain = 0;   // blocks left in CSR order
};

// Delta pack of the T list over a partition (K = 1 one-list GatherScatter,
// k_gs_sym_delta): per block kCap 16-bit words (CSR position | row-head bit
// | delta code), and kCap / 512 * E 32-bit exception words.  A fallback
// block (one whose columns the 16-bit words cannot encode) carries the
// fallback mark as every range's exception word 0 and its slot in the
// 32-bit words `f` as word 1: slot b in the partition's full pack, or, once
// compacted, its rank among the fallback blocks in a pack holding only
// theirs (ownsF).
struct DeltaPack {
  uint16_t *w = nullptr;               // nblocks * kCap words
  uint32_t *x = nullptr;               // nblocks * (kCap / 512) * E exception words
  int32_t E = 0;                       // exception words per 512-entry range
  uint32_t tbl[4] = {0u, 0u, 0u, 0u};  // byte c: the delta of code c
  int32_t fallback = 0;                // blocks that run from the 32-bit words
  int32_t csr = 0;                     // of those, blocks left in CSR order (words: column ids)
  Pack f;                              // the fallback blocks' 32-bit words
  bool ownsF = false;                  // f is the compacted copy (else the partition's pack)
  bool rows = false;                   // GatherScatter folds compacted rows (>= 2 entries per row)
  bool seg8 = false;                   // no row has more than 8 members (each thread folds 8 positions)
  // one-byte codes over the same blocks (k_gs_sym_byte; bw null: none): a
  // code per entry into a 255-entry table of (column delta, position delta,
  // head bit), code 255 taking the range's next exception item
  uint8_t *bw = nullptr;               // nblocks * kCap codes
  uint32_t *bh = nullptr;              // nblocks * 4 range headers (exception offset << 7 | count)
  uint32_t *bx = nullptr;              // exception items, two words each
  uint32_t *bt = nullptr;              // 256 table words
  int64_t bItems = 0;                  // exception items
  int32_t bFallback = 0;               // blocks that run from the 16-bit words
};

// A list's column -> row map in 4-bit codes (k_scatter_rows_delta): per
// 2,048-column chunk 256 code dwords (8 columns each) and 4 E exception
// words; chunks that do not fit read the int32 map (colRow).
struct RowMap {
  uint32_t *codes = nullptr;  // chunks * 256 dwords (null: none)
  uint32_t *x = nullptr;      // chunks * 4 * E exception words
  uint32_t *table = nullptr;  // 16 words: the delta of code c (0 at codes 0 and 15)
  int32_t E = 0;
  int32_t chunks = 0;
  int32_t fallback = 0;       // chunks that read colRow
  // the fallback chunks' rows alone (non-null once compacted: colRow is then
  // not kept, and a fallback chunk's exception word 1 holds its slot here)
  int32_t *fbRows = nullptr;
};

// The K = 1 one-list GatherScatter over bundled rows: the T list's rows
// reordered so that a block holds the rows of ty neighbouring lines and tz
// neighbouring layers of a structured mesh side by side (row r's bundle
// position from the line and layer row strides Ry, Rz that the plan detects
// in first-encounter order), cut on entries like partition 0, with its own
// delta pack and block order.  GatherScatter results do not depend on which
// block a row falls in (each row is reduced in its own CSR order), so only
// that launch uses it; the plan keeps it where it measured faster.
struct Bundle {
  Partition part;   // entries index the bundled CSR, which is not kept
  DeltaPack dp;     // compacted fallback words (no block left in CSR order)
  int32_t Ry = 0, Rz = 0, ty = 0, tz = 0;
};

}  // namespace ogs_hip

// The opaque plan behind ogs_hip_plan_t.
struct ogs_hip_plan_s {
  // holders: the creator, ogs_hip_plan_retain callers, the batches over the
  // plan; the last ogs_hip_plan_destroy frees it
  std::atomic<int> refs{1};
  // what the plan keeps alive besides its own arrays (the CSR of a plan made
  // by an ogs_t setup), released with the plan
  std::shared_ptr<void> keep;
  int device = 0;
  int cus = 0;  // compute units of the device (persistent launch sizes)
  int32_t Ncols = 0, NrowsN = 0, NrowsT = 0, nnzN = 0, nnzT = 0;
  int32_t maxRowN = 0, maxRowT = 0;
  const int32_t *rsN = nullptr, *ciN = nullptr, *rsT = nullptr, *ciT = nullptr;
  bool aliased = false;  // N lists are the T lists (Unsigned operator)
  int nparts = 0;
  ogs_hip::Partition partT[ogs_hip::kMaxParts];  // cut on T, rows [0, NrowsT)
  ogs_hip::Partition partN[ogs_hip::kMaxParts];  // cut on N, rows [0, NrowsN)
  // T-list partitions of the row-wise scatter kernels: merge path (measured
  // faster for them than blocks cut on entries); copies of partT when those
  // are merge path (a T list with empty rows) or when the T list scatters
  // by columns (colRowT)
  ogs_hip::Partition partS[ogs_hip::kMaxParts];
  // column-sorted packs: the T list over partT[0..kPackParts) (K=1 and the
  // K-field Sym kernels), the N list over partT[0] and over partN[0] (K=1)
  ogs_hip::Pack packT[ogs_hip::kPackParts];
  ogs_hip::Pack packNT;
  ogs_hip::Pack packNN;
  ogs_hip::Pack packS;  // the T list over partS[0] (a copy of packT[0] when partS is partT)
  ogs_hip::DeltaPack dpack;  // the T list over partT[0] with 16-bit words (w null: none)
  ogs_hip::Bundle bundle;    // bundled K = 1 GatherScatter blocks (part.d null: none)
  // T lists cut on entries into blocks of at most kCap3[i] entries, with
  // their packs: the one-list GatherScatter with three fields runs whole
  // blocks of its shape instead of the next power-of-two width below it
  // (nblocks 0: not built, a T list with empty rows)
  ogs_hip::Partition part3[2];
  ogs_hip::Pack pack3[2];
  // per column, the largest row of the list holding it (-1: none), for lists
  // with at least half as many entries as columns (else null); colRowN is
  // colRowT for an Unsigned operator
  int32_t *colRowT = nullptr;
  int32_t *colRowN = nullptr;
  // the same maps in 4-bit row-delta codes, for Scatter at K = 1 (rowMapN
  // is rowMapT for an Unsigned operator)
  ogs_hip::RowMap rowMapT;
  ogs_hip::RowMap rowMapN;
  // per column, the N-list member of its T row (-1: no row; -2: a row with
  // no N member), for operators whose N list holds at most one member per
  // row and whose T list has colRowT (else null): GatherScatter NoTrans K = 1
  // by columns
  int32_t *colOwner = nullptr;
  // one bit per column, set where the column is an N member of its T row,
  // for operators whose T columns are distinct, whose rows of both lists
  // ascend and whose N rows are subsets of their T rows (else null):
  // GatherScatter Trans over the T list alone
  uint32_t *nMask = nullptr;
};

namespace ogs_hip {
using Plan = ogs_hip_plan_s;

// The single-rank setup on the device (ogs_setup_dev.hip): from host ids to
// the local operator's CSR pair, device arrays handed to the caller (rsN and
// ciN null for the Unsigned kind, whose N lists are the T lists; all null
// for the Halo kind, whose local operator has no rows at one rank).
struct DeviceSetup {
  int32_t G = 0;        // groups (rows)
  int32_t NlocalP = 0;  // groups with a positive member
  int32_t nnzT = 0, nnzN = 0, maxRowT = 0, maxRowN = 0;
  bool gather_defined = false;
  int32_t *rsT = nullptr, *ciT = nullptr, *rsN = nullptr, *ciN = nullptr;
};
// kind: OGS_UNSIGNED / OGS_SIGNED / OGS_HALO; unique writes the signs back
// into h_ids.  Returns OGS_E_HALO_POSITIVE for a Halo group without exactly
// one positive id, or a hipError_t.
int device_setup(int32_t N, int64_t *h_ids, int kind, int unique, hipStream_t st, DeviceSetup *out);
// The T list (rs, ci) with its rows reordered into bundles of ty lines and
// tz layers (row strides Ry, Rz): rsP/ciP allocated here, row p of the
// result being the list's row at the p-th smallest bundle key.
int bundle_csr(const int32_t *rs, const int32_t *ci, int32_t nrows, int32_t nnz, int32_t Ry, int32_t Rz, int ty,
               int tz, hipStream_t st, int32_t **rsP, int32_t **ciP);
// Makes `plan` hold `keep` (device arrays it reads) until it is freed.
inline void plan_keep(ogs_hip_plan_s *plan, std::shared_ptr<void> keep) { plan->keep = std::move(keep); }
}
