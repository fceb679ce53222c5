// ogs_setup.cpp -- ogs_t::Setup at one rank (the reference's
// libs/ogs/ogsSetup.cpp:44-190 with comm size 1) and the host half of
// ogsOperator_t (setupRowBlocks, device upload).
//
// The result is the reference's: rows numbered by first encounter in local
// order, positive groups first (ogsSetup.cpp:409-433); colIds ascending
// within a row (637-663); N lists holding only the positive members;
// Unsigned operators aliasing N to T (719, 728).  The method differs: with
// one rank there is nothing to route, and the grouping, row numbering and
// CSR construction run on the device (ogs_setup_dev.hip: one radix sort by
// |id|, scans and scatters); the host receives the counts and its copies
// of the lists.
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "ogs.hpp"
#include "ogs_internal.hpp"

namespace libp {
namespace ogs {

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

  // groups, signs, row order and the CSR pair, on the device
  ogs_hip::DeviceSetup d;
  const int rc = ogs_hip::device_setup(N, ids, (int)kind, unique ? 1 : 0, platform.stream, &d);
  LIBP_ABORT("Found a Halo group without exactly one positive Id.", rc == OGS_E_HALO_POSITIVE,
             OGS_E_HALO_POSITIVE);                       // ogsSetup.cpp:286-288
  LIBP_ABORT("ogs device setup: " << ogs_hip_error_string(rc), rc != 0, rc);
  gather_defined = d.gather_defined;                     // ogsSetup.cpp:312-313
  const dlong G = d.G;

  // Rows: positive groups [0, NlocalP) then the others, each class in
  // first-encounter order (ogsSetup.cpp:409-433).  At one rank every group
  // is local: NhaloT = NhaloP = 0.
  NlocalP = d.NlocalP;
  NlocalT = G;
  NhaloP = NhaloT = 0;
  Ngather = NlocalP + NhaloP;  // ogsSetup.cpp:388
  NgatherGlobal = Ngather;     // ogsSetup.cpp:391-392 at size 1

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

  const bool onDevice = d.rsT != nullptr;
  if (kind == Halo || !onDevice) {
    // Halo: LocalHaloSetup (ogsSetup.cpp:779-847) keeps only shared nodes,
    // and at one rank there are none.  Otherwise no node has a nonzero id.
    L.rowStartsT = memory<dlong>((size_t)(kind == Halo ? 0 : G) + 1, 0);
    L.rowStartsN = L.rowStartsT;
    L.aliased = true;
  } else {
    const dlong NrowsT = G;
    L.NrowsT = NrowsT;
    L.NrowsN = NlocalP;
    L.nnzT = d.nnzT;
    L.nnzN = d.nnzN;
    L.o_rowStartsT = deviceMemory<dlong>::adopt(d.rsT, (size_t)NrowsT + 1);
    L.o_colIdsT = deviceMemory<dlong>::adopt(d.ciT, (size_t)L.nnzT);
    L.aliased = kind == Unsigned;
    if (L.aliased) {
      L.o_rowStartsN = L.o_rowStartsT;
      L.o_colIdsN = L.o_colIdsT;
    } else {
      L.o_rowStartsN = deviceMemory<dlong>::adopt(d.rsN, (size_t)NrowsT + 1);
      L.o_colIdsN = deviceMemory<dlong>::adopt(d.ciN, (size_t)L.nnzN);
    }
    // the operator's host row starts (row blocks); colIds come down on demand
    auto down = [&](memory<dlong> &h, const deviceMemory<dlong> &dv, size_t n) {
      h = memory<dlong>(n);
      if (n) LIBP_HIP_CHECK(hipMemcpyAsync(h.ptr(), dv.ptr(), n * sizeof(dlong), hipMemcpyDeviceToHost,
                                           platform.stream));
    };
    down(L.rowStartsT, L.o_rowStartsT, (size_t)NrowsT + 1);
    if (L.aliased) L.rowStartsN = L.rowStartsT;
    else down(L.rowStartsN, L.o_rowStartsN, (size_t)NrowsT + 1);
    platform.finish();
  }

  L.setupRowBlocks();             // ogsSetup.cpp:681, 774
  gatherHalo->setupRowBlocks();   // ogsSetup.cpp:682, 775
  if (L.o_rowStartsT.ptr()) L.createPlan(platform, planFields);
  else L.upload(platform, planFields);
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

const memory<dlong> &ogsOperator_t::hostColIdsT() const {
  if (colIdsT.length() != (size_t)nnzT) {
    colIdsT = memory<dlong>((size_t)nnzT);
    if (nnzT) LIBP_HIP_CHECK(hipMemcpy(colIdsT.ptr(), o_colIdsT.ptr(), sizeof(dlong) * (size_t)nnzT,
                                       hipMemcpyDeviceToHost));
  }
  return colIdsT;
}

const memory<dlong> &ogsOperator_t::hostColIdsN() const {
  if (aliased) return hostColIdsT();
  if (colIdsN.length() != (size_t)nnzN) {
    colIdsN = memory<dlong>((size_t)nnzN);
    if (nnzN) LIBP_HIP_CHECK(hipMemcpy(colIdsN.ptr(), o_colIdsN.ptr(), sizeof(dlong) * (size_t)nnzN,
                                       hipMemcpyDeviceToHost));
  }
  return colIdsN;
}

void ogsOperator_t::upload(const platform_t &platform, int fields) {
  o_rowStartsT = platform.malloc<dlong>((size_t)NrowsT + 1, rowStartsT.ptr());
  o_colIdsT = platform.malloc<dlong>((size_t)nnzT, colIdsT.ptr());
  if (aliased) {
    o_rowStartsN = o_rowStartsT;
    o_colIdsN = o_colIdsT;
  } else {
    o_rowStartsN = platform.malloc<dlong>((size_t)NrowsT + 1, rowStartsN.ptr());
    o_colIdsN = platform.malloc<dlong>((size_t)nnzN, colIdsN.ptr());
  }
  createPlan(platform, fields);
}

void ogsOperator_t::createPlan(const platform_t &platform, int fields) {
  if (plan_) { ogs_hip_plan_destroy(plan_); plan_ = nullptr; }
  const int rc = ogs_hip_plan_create_k(&plan_, Ncols, NrowsN, NrowsT, o_rowStartsN.ptr(), o_colIdsN.ptr(),
                                       o_rowStartsT.ptr(), o_colIdsT.ptr(), fields, platform.stream);
  LIBP_ABORT("ogs_hip_plan_create: " << ogs_hip_error_string(rc), rc != 0, rc);
  // the plan reads these lists: it shares their ownership, so that a batch or
  // another holder of the plan keeps them alive past this operator
  struct Lists {
    deviceMemory<dlong> rsN, ciN, rsT, ciT;
  };
  ogs_hip::plan_keep(plan_, std::make_shared<Lists>(Lists{o_rowStartsN, o_colIdsN, o_rowStartsT, o_colIdsT}));
}

ogsOperator_t::~ogsOperator_t() {
  if (plan_) ogs_hip_plan_destroy(plan_);
}

}  // namespace ogs
}  // namespace libp
