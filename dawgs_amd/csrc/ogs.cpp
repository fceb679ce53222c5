// ogs.cpp -- ogs_t's GatherScatter / Gather / Scatter at one rank
// (the reference's libs/ogs/ogs.cpp:39-561) and the handle-layer C-ABI
// (ogs_* in include/ogs_hip.h).
//
// At one rank the halo operator has no rows and the pairwise exchange has
// no neighbours (ogsPairwise.cpp:45-183 with zero sends), so each Start is
// the empty halo gather and each Finish is the local operator's kernel.  The
// reference's device.finish() calls around the host staging of the halo
// buffer (ogs.cpp:74,107,117) guard a copy that is empty here and are not
// issued: every call is one asynchronous launch on the platform's stream.
#include <cstring>
#include <new>

#include "ogs.hpp"

namespace libp {
namespace ogs {

template <typename T>
void ogsOperator_t::GatherScatter(deviceMemory<T> o_v, int k, Op op, Transpose trans, hipStream_t s) const {
  const int rc = ogs_hip_gather_scatter(plan_, (int)ogsType<T>::get(), (int)op, (int)trans, k, o_v.ptr(), s);
  LIBP_ABORT("ogs_hip_gather_scatter: " << ogs_hip_error_string(rc), rc != 0, rc);
}

template <typename T>
void ogsOperator_t::Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, int k, Op op, Transpose trans,
                           hipStream_t s) const {
  const int rc = ogs_hip_gather(plan_, (int)ogsType<T>::get(), (int)op, (int)trans, k, o_gv.ptr(), o_v.ptr(), s);
  LIBP_ABORT("ogs_hip_gather: " << ogs_hip_error_string(rc), rc != 0, rc);
}

template <typename T>
void ogsOperator_t::Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, int k, Transpose trans, hipStream_t s) const {
  const int rc = ogs_hip_scatter(plan_, (int)ogsType<T>::get(), (int)trans, k, o_v.ptr(), o_gv.ptr(), s);
  LIBP_ABORT("ogs_hip_scatter: " << ogs_hip_error_string(rc), rc != 0, rc);
}

// ---- GatherScatter (ogs.cpp:39-123) ----
template <typename T>
void ogs_t::GatherScatter(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans, hipStream_t s) const {
  GatherScatterStart(o_v, k, op, trans, s);
  GatherScatterFinish(o_v, k, op, trans, s);
}

template <typename T>
void ogs_t::GatherScatterStart(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                               hipStream_t s) const {
  LIBP_ABORT("ogs handle is not set up.", !gatherLocal, OGS_E_ARG);
  LIBP_ABORT("k must be positive", k < 1, OGS_E_ARG);
  // halo gather into the exchange workspace: zero rows at one rank (ogs.cpp:58)
  gatherHalo->Gather(deviceMemory<T>(), o_v, k, op, trans, s);
}

template <typename T>
void ogs_t::GatherScatterFinish(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                                hipStream_t s) const {
  // the local operator runs `op`, as the host path does (ogsOperator.cpp:362-378);
  // the reference's device overload always adds (ogsOperator.cpp:399-430)
  gatherLocal->GatherScatter(o_v, k, op, trans, s);  // ogs.cpp:91
  gatherHalo->Scatter(o_v, deviceMemory<T>(), k, trans, s);  // ogs.cpp:122, zero rows
}

template <typename T>
void ogs_t::GatherScatter(memory<T> v, const int k, const Op op, const Transpose trans) {
  deviceMemory<T> o_v = platform.malloc<T>(v.length(), v.ptr());
  GatherScatter(o_v, k, op, trans);
  LIBP_HIP_CHECK(hipMemcpyAsync(v.ptr(), o_v.ptr(), v.length() * sizeof(T), hipMemcpyDeviceToHost, platform.stream));
  platform.finish();
}

// ---- Gather (ogs.cpp:202-374) ----
template <typename T>
void ogs_t::Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                   hipStream_t s) const {
  GatherStart(o_gv, o_v, k, op, trans, s);
  GatherFinish(o_gv, o_v, k, op, trans, s);
}

template <typename T>
void ogs_t::GatherStart(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                        hipStream_t s) const {
  AssertGatherDefined();  // ogs.cpp:218
  LIBP_ABORT("k must be positive", k < 1, OGS_E_ARG);
  // Trans: halo gather into the workspace; else into o_gv + k*NlocalT (ogs.cpp:222-251); zero rows here
  gatherHalo->Gather(o_gv + (size_t)k * NlocalT, o_v, k, op, trans, s);
}

template <typename T>
void ogs_t::GatherFinish(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                         hipStream_t s) const {
  AssertGatherDefined();  // ogs.cpp:260
  gatherLocal->Gather(o_gv, o_v, k, op, trans, s);  // ogs.cpp:265
}

template <typename T>
void ogs_t::Gather(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans) {
  deviceMemory<T> o_v = platform.malloc<T>(v.length(), v.ptr());
  deviceMemory<T> o_gv = platform.malloc<T>(gv.length(), gv.ptr());
  Gather(o_gv, o_v, k, op, trans);
  LIBP_HIP_CHECK(hipMemcpyAsync(gv.ptr(), o_gv.ptr(), gv.length() * sizeof(T), hipMemcpyDeviceToHost, platform.stream));
  platform.finish();
}

// ---- Scatter (ogs.cpp:392-561) ----
template <typename T>
void ogs_t::Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans,
                    hipStream_t s) const {
  ScatterStart(o_v, o_gv, k, trans, s);
  ScatterFinish(o_v, o_gv, k, trans, s);
}

template <typename T>
void ogs_t::ScatterStart(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans,
                         hipStream_t s) const {
  (void)o_v; (void)o_gv; (void)trans; (void)s;
  AssertGatherDefined();  // ogs.cpp:406
  LIBP_ABORT("k must be positive", k < 1, OGS_E_ARG);
  // NoTrans: the halo tail of o_gv goes to the exchange (ogs.cpp:410-441); empty here
}

template <typename T>
void ogs_t::ScatterFinish(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans,
                          hipStream_t s) const {
  AssertGatherDefined();  // ogs.cpp:449
  gatherLocal->Scatter(o_v, o_gv, k, trans, s);  // ogs.cpp:454
  gatherHalo->Scatter(o_v, o_gv + (size_t)k * NlocalT, k, trans, s);  // ogs.cpp:483/486, zero rows
}

template <typename T>
void ogs_t::Scatter(memory<T> v, const memory<T> gv, const int k, const Transpose trans) {
  deviceMemory<T> o_v = platform.malloc<T>(v.length(), v.ptr());
  deviceMemory<T> o_gv = platform.malloc<T>(gv.length(), gv.ptr());
  Scatter(o_v, o_gv, k, trans);
  LIBP_HIP_CHECK(hipMemcpyAsync(v.ptr(), o_v.ptr(), v.length() * sizeof(T), hipMemcpyDeviceToHost, platform.stream));
  platform.finish();
}

// ---- SetupGlobalToLocalMapping (ogsSetup.cpp:867-886) ----
// Scatter (NoTrans) of the row indices 0..NlocalT-1 into an array of -1s:
// each node receives its row; the halo operator has no rows at one rank.
void ogs_t::SetupGlobalToLocalMapping(memory<dlong> GlobalToLocal) {
  LIBP_ABORT("GlobalToLocal shorter than N", GlobalToLocal.length() < (size_t)N, OGS_E_ARG);
  SetupGlobalToLocalMapping(GlobalToLocal.ptr());
}

void ogs_t::SetupGlobalToLocalMapping(dlong *GlobalToLocal) {
  LIBP_ABORT("ogs handle is not set up.", !gatherLocal, OGS_E_ARG);
  LIBP_ABORT("Null GlobalToLocal", N > 0 && !GlobalToLocal, OGS_E_ARG);
  if (N == 0) return;
  std::vector<dlong> rows((size_t)NlocalT + NhaloT);
  for (size_t n = 0; n < rows.size(); ++n) rows[n] = (dlong)n;
  std::vector<dlong> minus((size_t)N, -1);
  deviceMemory<dlong> o_map = platform.malloc<dlong>((size_t)N, minus.data());
  deviceMemory<dlong> o_rows = platform.malloc<dlong>(rows.size(), rows.data());
  gatherLocal->Scatter(o_map, o_rows, 1, NoTrans, platform.stream);
  gatherHalo->Scatter(o_map, o_rows + (size_t)NlocalT, 1, NoTrans, platform.stream);
  LIBP_HIP_CHECK(hipMemcpyAsync(GlobalToLocal, o_map.ptr(), sizeof(dlong) * (size_t)N, hipMemcpyDeviceToHost,
                                platform.stream));
  platform.finish();
}

#define OGS_INSTANTIATE(T)                                                                                       \
  template void ogs_t::GatherScatter(deviceMemory<T>, const int, const Op, const Transpose, hipStream_t) const;   \
  template void ogs_t::GatherScatterStart(deviceMemory<T>, const int, const Op, const Transpose, hipStream_t) const; \
  template void ogs_t::GatherScatterFinish(deviceMemory<T>, const int, const Op, const Transpose, hipStream_t) const; \
  template void ogs_t::GatherScatter(memory<T>, const int, const Op, const Transpose);                          \
  template void ogs_t::Gather(deviceMemory<T>, deviceMemory<T>, const int, const Op, const Transpose, hipStream_t) const; \
  template void ogs_t::GatherStart(deviceMemory<T>, deviceMemory<T>, const int, const Op, const Transpose, hipStream_t) const; \
  template void ogs_t::GatherFinish(deviceMemory<T>, deviceMemory<T>, const int, const Op, const Transpose, hipStream_t) const; \
  template void ogs_t::Gather(memory<T>, const memory<T>, const int, const Op, const Transpose);                 \
  template void ogs_t::Scatter(deviceMemory<T>, deviceMemory<T>, const int, const Transpose, hipStream_t) const; \
  template void ogs_t::ScatterStart(deviceMemory<T>, deviceMemory<T>, const int, const Transpose, hipStream_t) const; \
  template void ogs_t::ScatterFinish(deviceMemory<T>, deviceMemory<T>, const int, const Transpose, hipStream_t) const; \
  template void ogs_t::Scatter(memory<T>, const memory<T>, const int, const Transpose);
OGS_INSTANTIATE(float)
OGS_INSTANTIATE(double)
OGS_INSTANTIATE(int32_t)
OGS_INSTANTIATE(int64_t)
OGS_INSTANTIATE(bfloat16)
OGS_INSTANTIATE(float_c)
#undef OGS_INSTANTIATE

}  // namespace ogs
}  // namespace libp

// ---------------------------------------------------------------------------
// Handle-layer C-ABI.

using namespace libp;


namespace {

template <class F>
int guarded(F &&f) {
  try {
    f();
    return OGS_SUCCESS;
  } catch (const libp::exception &e) {
    return e.code() ? e.code() : OGS_E_ARG;
  } catch (const std::bad_alloc &) {
    return (int)hipErrorOutOfMemory;
  } catch (...) {
    return OGS_E_ARG;
  }
}

template <class F>
int by_type(int type, F &&f) {
  switch (type) {
    case OGS_FLOAT: return f(float{});
    case OGS_DOUBLE: return f(double{});
    case OGS_INT32: return f(int32_t{});
    case OGS_INT64: return f(int64_t{});
    case OGS_BF16: return f(ogs::bfloat16{});
    case OGS_FLOAT_C: return f(ogs::float_c{});
    default: return OGS_E_ARG;
  }
}

bool bad_op(int op) { return op < OGS_ADD || op > OGS_MIN; }
bool bad_trans(int t) { return t < OGS_SYM || t > OGS_TRANS; }

}  // namespace

extern "C" {

int ogs_setup(ogs_handle_t *h, int32_t N, int64_t *ids, int kind, int method, int unique, int verbose) {
  return ogs_setup_k(h, N, ids, kind, method, unique, verbose, OGS_PLAN_ALL_K);
}

int ogs_setup_k(ogs_handle_t *h, int32_t N, int64_t *ids, int kind, int method, int unique, int verbose,
                int fields) {
  if (!h) return OGS_E_ARG;
  *h = nullptr;
  if (fields != OGS_PLAN_ALL_K && fields != OGS_PLAN_K1) return OGS_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OGS_E_NO_DEVICE;
  ogs_handle_s *o = new (std::nothrow) ogs_handle_s();
  if (!o) return (int)hipErrorOutOfMemory;
  o->ogs.planFields = fields;
  const int rc = guarded([&] {
    platform_t platform;
    o->ogs.Setup(N, ids, (ogs::Kind)kind, (ogs::Method)method, unique != 0, verbose != 0, platform);
  });
  if (rc) { delete o; return rc; }
  *h = o;
  return OGS_SUCCESS;
}

int ogs_free(ogs_handle_t h) {
  if (!h) return OGS_E_ARG;
  delete h;
  return OGS_SUCCESS;
}

int ogs_get_info(ogs_handle_t h, ogs_info_t *info) {
  if (!h || !info) return OGS_E_ARG;
  const ogs::ogs_t &o = h->ogs;
  const ogs::ogsOperator_t &L = o.local();
  info->N = o.N;
  info->Ngather = o.Ngather;
  info->NlocalT = o.NlocalT;
  info->NhaloT = o.NhaloT;
  info->NlocalP = o.NlocalP;
  info->NhaloP = o.NhaloP;
  info->NgatherGlobal = o.NgatherGlobal;
  info->kind = (int32_t)o.kind;
  info->unique = o.unique ? 1 : 0;
  info->gather_defined = o.gather_defined ? 1 : 0;
  info->NrowsN = L.NrowsN;
  info->NrowsT = L.NrowsT;
  info->nnzN = L.nnzN;
  info->nnzT = L.nnzT;
  info->NrowBlocksN = L.NrowBlocksN;
  info->NrowBlocksT = L.NrowBlocksT;
  return OGS_SUCCESS;
}

int ogs_get_csr(ogs_handle_t h, int32_t *rowStartsN, int32_t *colIdsN, int32_t *rowStartsT, int32_t *colIdsT,
                int32_t *blockRowStartsN, int32_t *blockRowStartsT) {
  if (!h) return OGS_E_ARG;
  const ogs::ogsOperator_t &L = h->ogs.local();
  auto cp = [](int32_t *dst, const memory<int32_t> &src, size_t n) {
    if (dst && n) std::memcpy(dst, src.ptr(), n * sizeof(int32_t));
  };
  cp(rowStartsN, L.rowStartsN, (size_t)L.NrowsT + 1);
  cp(rowStartsT, L.rowStartsT, (size_t)L.NrowsT + 1);
  if (colIdsN && L.nnzN) cp(colIdsN, L.hostColIdsN(), (size_t)L.nnzN);
  if (colIdsT && L.nnzT) cp(colIdsT, L.hostColIdsT(), (size_t)L.nnzT);
  cp(blockRowStartsN, L.blockRowStartsN, (size_t)L.NrowBlocksN + 1);
  cp(blockRowStartsT, L.blockRowStartsT, (size_t)L.NrowBlocksT + 1);
  return OGS_SUCCESS;
}

ogs_hip_plan_t ogs_get_plan(ogs_handle_t h) { return h ? h->ogs.local().plan() : nullptr; }

int ogs_setup_global_to_local(ogs_handle_t h, int32_t *GlobalToLocal) {
  if (!h) return OGS_E_ARG;
  return guarded([&] { h->ogs.SetupGlobalToLocalMapping(GlobalToLocal); });
}

int ogs_gather_scatter(ogs_handle_t h, int type, int op, int trans, int32_t K, void *d_v, void *stream) {
  if (!h || bad_op(op) || bad_trans(trans) || K < 1) return OGS_E_ARG;
  const ogs::ogs_t &o = h->ogs;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  return by_type(type, [&](auto tag) {
    using T = decltype(tag);
    return guarded([&] {
      o.GatherScatter(deviceMemory<T>(static_cast<T *>(d_v), (size_t)o.N * K), K, (ogs::Op)op, (ogs::Transpose)trans,
                      s);
    });
  });
}

int ogs_gather(ogs_handle_t h, int type, int op, int trans, int32_t K, void *d_gv, const void *d_v, void *stream) {
  if (!h || bad_op(op) || bad_trans(trans) || K < 1) return OGS_E_ARG;
  const ogs::ogs_t &o = h->ogs;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  return by_type(type, [&](auto tag) {
    using T = decltype(tag);
    return guarded([&] {
      o.Gather(deviceMemory<T>(static_cast<T *>(d_gv), (size_t)o.NlocalT * K),
               deviceMemory<T>(const_cast<T *>(static_cast<const T *>(d_v)), (size_t)o.N * K), K, (ogs::Op)op,
               (ogs::Transpose)trans, s);
    });
  });
}

int ogs_scatter(ogs_handle_t h, int type, int trans, int32_t K, void *d_v, const void *d_gv, void *stream) {
  if (!h || bad_trans(trans) || K < 1) return OGS_E_ARG;
  const ogs::ogs_t &o = h->ogs;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  return by_type(type, [&](auto tag) {
    using T = decltype(tag);
    return guarded([&] {
      o.Scatter(deviceMemory<T>(static_cast<T *>(d_v), (size_t)o.N * K),
                deviceMemory<T>(const_cast<T *>(static_cast<const T *>(d_gv)), (size_t)o.NlocalT * K), K,
                (ogs::Transpose)trans, s);
    });
  });
}

}  // extern "C"
