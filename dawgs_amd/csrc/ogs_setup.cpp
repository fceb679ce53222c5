// ogs_setup.cpp -- ogs_t::Setup at one rank (the reference's
// libs/ogs/ogsSetup.cpp:44-190 with comm size 1) and the host half of
// ogsOperator_t (setupRowBlocks, device upload).
//
// The result is the reference's: rows numbered by first encounter in local
// order, positive groups first (ogsSetup.cpp:409-433); colIds ascending
// within a row (637-663); N lists holding only the positive members;
// Unsigned operators aliasing N to T (719, 728).  The method differs: with
// one rank there is nothing to route, so instead of three sorts of 32-byte
// nodes the setup makes two linear passes.  Groups are found by a direct
// table over |id| when the ids are dense (the structured-mesh numbering) and
// by an open-addressing hash otherwise, and are numbered in first-encounter
// order, which is already the row order within each sign class.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "ogs.hpp"

namespace libp {
namespace ogs {

namespace {

inline uint64_t uabs(hlong x) { return x < 0 ? (uint64_t)0 - (uint64_t)x : (uint64_t)x; }

// |id| -> group number, first encounter order.  Returns the group count.
dlong number_groups(const dlong N, const hlong *ids, std::vector<dlong> &grp) {
  grp.assign((size_t)N, -1);
  uint64_t maxAbs = 0;
  for (dlong n = 0; n < N; ++n) {
    const uint64_t a = uabs(ids[n]);
    if (a > maxAbs) maxAbs = a;
  }
  dlong G = 0;
  if (maxAbs <= 4 * (uint64_t)N + 1024) {
    std::vector<dlong> table((size_t)maxAbs + 1, -1);
    for (dlong n = 0; n < N; ++n) {
      if (ids[n] == 0) continue;
      dlong &slot = table[uabs(ids[n])];
      if (slot < 0) slot = G++;
      grp[n] = slot;
    }
    return G;
  }
  size_t cap = 16;
  while (cap < 2 * (size_t)N) cap <<= 1;
  std::vector<uint64_t> keys(cap, 0);  // 0 = empty (id 0 never enters)
  std::vector<dlong> vals(cap, -1);
  const size_t mask = cap - 1;
  for (dlong n = 0; n < N; ++n) {
    if (ids[n] == 0) continue;
    const uint64_t key = uabs(ids[n]);
    size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 17) & mask;
    while (keys[h] != 0 && keys[h] != key) h = (h + 1) & mask;
    if (keys[h] == 0) {
      keys[h] = key;
      vals[h] = G++;
    }
    grp[n] = vals[h];
  }
  return G;
}

}  // namespace

void ogs_t::Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Kind _kind, const Method method,
                  const bool _unique, const bool verbose, platform_t &_platform) {
  LIBP_ABORT("ids shorter than N", (size_t)_N > ids.length(), OGS_E_ARG);
  comm = _comm;
  Setup(_N, ids.ptr(), _kind, method, _unique, verbose, _platform);
}

void ogs_t::Setup(const dlong _N, hlong *ids, const Kind _kind, const Method method, const bool _unique,
                  const bool verbose, platform_t &_platform) {
  (void)method;  // one rank: every method is the empty pairwise exchange (ogsAuto.cpp:154-156)
  Free();
  const auto t0 = std::chrono::high_resolution_clock::now();
  platform = _platform;
  N = _N;
  kind = _kind;
  unique = _unique;
  LIBP_ABORT("Invalid ogs setup requested", (kind == Unsigned && unique) || (kind == Halo && unique),
             OGS_E_INVALID_SETUP);                       // ogsSetup.cpp:98-100
  LIBP_ABORT("Invalid ogs kind", kind < Unsigned || kind > Halo, OGS_E_ARG);
  LIBP_ABORT("Negative N", N < 0, OGS_E_ARG);
  LIBP_ABORT("Null ids", N > 0 && ids == nullptr, OGS_E_ARG);

  std::vector<dlong> grp;
  const dlong G = number_groups(N, ids, grp);

  // Group sizes and positive-member counts.  Unsigned treats every id as
  // positive (ogsSetup.cpp:115-116).
  std::vector<dlong> gsize((size_t)G, 0), gpos((size_t)G, 0);
  for (dlong n = 0; n < N; ++n) {
    const dlong g = grp[n];
    if (g < 0) continue;
    gsize[g]++;
    if (kind == Unsigned || ids[n] > 0) gpos[g]++;
  }

  if (unique) {
    // One positive node per group: the first-encountered member.  The
    // reference draws it with rand() (ogsSetup.cpp:264-275); ogs.hpp:96-99
    // only asks for an arbitrary but consistent choice.  Signs are written
    // back into ids (ogsSetup.cpp:139-140).
    std::vector<char> seen((size_t)G, 0);
    for (dlong n = 0; n < N; ++n) {
      const dlong g = grp[n];
      if (g < 0) continue;
      const hlong a = (hlong)uabs(ids[n]);
      ids[n] = seen[g] ? -a : a;
      seen[g] = 1;
      gpos[g] = 1;
    }
    gather_defined = true;
  } else {
    bool one = true;
    for (dlong g = 0; g < G; ++g) {
      if (gpos[g] != 1) one = false;
      LIBP_ABORT("Found " << gpos[g] << " positive Ids for a Halo group.", kind == Halo && gpos[g] != 1,
                 OGS_E_HALO_POSITIVE);                   // ogsSetup.cpp:286-288
    }
    gather_defined = one;                                // ogsSetup.cpp:312-313
  }

  // Rows: positive groups [0, NlocalP) then the others, each class in
  // first-encounter order (ogsSetup.cpp:409-433).  At one rank every group
  // is local: NhaloT = NhaloP = 0.
  dlong nPos = 0;
  for (dlong g = 0; g < G; ++g) nPos += (gpos[g] > 0);
  NlocalP = nPos;
  NlocalT = G;
  NhaloP = NhaloT = 0;
  Ngather = NlocalP + NhaloP;  // ogsSetup.cpp:388
  NgatherGlobal = Ngather;     // ogsSetup.cpp:391-392 at size 1
  std::vector<dlong> row((size_t)G);
  {
    dlong cp = 0, cn = NlocalP;
    for (dlong g = 0; g < G; ++g) row[g] = (gpos[g] > 0) ? cp++ : cn++;
  }

  gatherLocal = std::make_shared<ogsOperator_t>();
  gatherHalo = std::make_shared<ogsOperator_t>();  // empty at one rank
  ogsOperator_t &L = *gatherLocal;
  L.kind = (kind == Unsigned) ? Unsigned : Signed;
  L.Ncols = N;
  gatherHalo->kind = L.kind;
  gatherHalo->Ncols = N;
  gatherHalo->rowStartsT = memory<dlong>(1, 0);
  gatherHalo->rowStartsN = gatherHalo->rowStartsT;
  gatherHalo->aliased = true;

  if (kind == Halo) {
    // LocalHaloSetup (ogsSetup.cpp:779-847) keeps only shared nodes, and at
    // one rank there are none: both operators are empty.
    L.rowStartsT = memory<dlong>(1, 0);
    L.rowStartsN = L.rowStartsT;
    L.aliased = true;
  } else {
    const dlong NrowsT = NlocalT;
    L.NrowsT = NrowsT;
    L.NrowsN = NlocalP;
    L.rowStartsT = memory<dlong>((size_t)NrowsT + 1, 0);
    for (dlong g = 0; g < G; ++g) L.rowStartsT[(size_t)row[g] + 1] = gsize[g];
    for (dlong r = 0; r < NrowsT; ++r) L.rowStartsT[(size_t)r + 1] += L.rowStartsT[r];
    L.nnzT = L.rowStartsT[NrowsT];
    L.colIdsT = memory<dlong>((size_t)L.nnzT);
    std::vector<dlong> fillT(L.rowStartsT.ptr(), L.rowStartsT.ptr() + NrowsT);
    if (kind == Unsigned) {
      L.aliased = true;
      L.rowStartsN = L.rowStartsT;
      L.colIdsN = L.colIdsT;
      L.nnzN = L.nnzT;
      for (dlong n = 0; n < N; ++n)
        if (grp[n] >= 0) L.colIdsT[fillT[row[grp[n]]]++] = n;
    } else {
      L.rowStartsN = memory<dlong>((size_t)NrowsT + 1, 0);
      for (dlong n = 0; n < N; ++n)
        if (grp[n] >= 0 && ids[n] > 0) L.rowStartsN[(size_t)row[grp[n]] + 1]++;
      for (dlong r = 0; r < NrowsT; ++r) L.rowStartsN[(size_t)r + 1] += L.rowStartsN[r];
      L.nnzN = L.rowStartsN[NrowsT];
      L.colIdsN = memory<dlong>((size_t)L.nnzN);
      std::vector<dlong> fillN(L.rowStartsN.ptr(), L.rowStartsN.ptr() + NrowsT);
      for (dlong n = 0; n < N; ++n) {
        const dlong g = grp[n];
        if (g < 0) continue;
        const dlong r = row[g];
        L.colIdsT[fillT[r]++] = n;
        if (ids[n] > 0) L.colIdsN[fillN[r]++] = n;
      }
    }
  }
  grp.clear();

  L.setupRowBlocks();             // ogsSetup.cpp:681, 774
  gatherHalo->setupRowBlocks();   // ogsSetup.cpp:682, 775
  L.upload(platform);
  gatherHalo->upload(platform);

  if (verbose) {
    const double secs = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    std::cout << "ogs Setup: 0 unique labels shared." << std::endl;            // ogsSetup.cpp:317-319
    std::cout << "ogs Setup Time: " << secs << " seconds." << std::endl;       // ogsSetup.cpp:187-189
  }
}

void ogs_t::Free() {
  gatherLocal = nullptr;
  gatherHalo = nullptr;
  N = Ngather = NlocalT = NhaloT = NlocalP = NhaloP = 0;
  NgatherGlobal = 0;
  gather_defined = false;
}

void ogs_t::AssertGatherDefined() const {
  LIBP_ABORT("Gather operation not well-defined.", !gather_defined, OGS_E_GATHER_UNDEF);
}

// setupRowBlocks (ogsOperator.cpp:447-512): greedy packing of consecutive
// rows into blocks of at most gatherNodesPerBlock nonzeros.  The GPU path
// cuts its own partitions; these blocks are kept because they are part of
// the operator's API (NrowBlocks*, blockRowStarts*), and the row-length
// check is the reference's contract.
void ogsOperator_t::setupRowBlocks() {
  const dlong cap = gatherNodesPerBlock;
  dlong sumN = 0, sumT = 0;
  std::vector<dlong> bN, bT;
  if (NrowsN) bN.push_back(0);
  if (NrowsT) bT.push_back(0);
  for (dlong i = 0; i < NrowsT; i++) {
    const dlong sizeN = rowStartsN[(size_t)i + 1] - rowStartsN[i];
    const dlong sizeT = rowStartsT[(size_t)i + 1] - rowStartsT[i];
    LIBP_ABORT("Multiplicity of global node id: " << i << " in ogsOperator_t::setupRowBlocks is too large.",
               sizeN > cap || sizeT > cap, OGS_E_ROW_TOO_LONG);
    if (sumN + sizeN > cap) { bN.push_back(i); sumN = sizeN; } else sumN += sizeN;
    if (sumT + sizeT > cap) { bT.push_back(i); sumT = sizeT; } else sumT += sizeT;
  }
  NrowBlocksN = (dlong)bN.size();
  NrowBlocksT = (dlong)bT.size();
  blockRowStartsN = memory<dlong>(bN.size() + 1);
  blockRowStartsT = memory<dlong>(bT.size() + 1);
  for (size_t b = 0; b < bN.size(); ++b) blockRowStartsN[b] = bN[b];
  for (size_t b = 0; b < bT.size(); ++b) blockRowStartsT[b] = bT[b];
  blockRowStartsN[bN.size()] = NrowsT;  // ogsOperator.cpp:507-508
  blockRowStartsT[bT.size()] = NrowsT;
}

void ogsOperator_t::upload(const platform_t &platform) {
  o_rowStartsT = platform.malloc<dlong>((size_t)NrowsT + 1, rowStartsT.ptr());
  o_colIdsT = platform.malloc<dlong>((size_t)nnzT, colIdsT.ptr());
  if (aliased) {
    o_rowStartsN = o_rowStartsT;
    o_colIdsN = o_colIdsT;
  } else {
    o_rowStartsN = platform.malloc<dlong>((size_t)NrowsT + 1, rowStartsN.ptr());
    o_colIdsN = platform.malloc<dlong>((size_t)nnzN, colIdsN.ptr());
  }
  if (plan_) { ogs_hip_plan_destroy(plan_); plan_ = nullptr; }
  const int rc = ogs_hip_plan_create(&plan_, Ncols, NrowsN, NrowsT, o_rowStartsN.ptr(), o_colIdsN.ptr(),
                                     o_rowStartsT.ptr(), o_colIdsT.ptr(), platform.stream);
  LIBP_ABORT("ogs_hip_plan_create: " << ogs_hip_error_string(rc), rc != 0, rc);
}

ogsOperator_t::~ogsOperator_t() {
  if (plan_) ogs_hip_plan_destroy(plan_);
}

}  // namespace ogs
}  // namespace libp
