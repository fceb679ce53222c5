// ogs_halo.cpp -- halo_t at one rank (the reference's
// libs/ogs/ogsHalo.cpp:39-359 and ogsSetup.cpp:55-64, 888-914) and its
// handle-layer C-ABI (halo_* in include/ogs_hip.h).
//
// With one rank no node is shared with another rank: the halo operator has
// no rows, the exchange has no neighbours and Nhalo = NhaloT - NhaloP = 0.
// Exchange gathers the halo buffer (zero rows, or the zero-length tail of a
// gathered vector), exchanges it with nobody and scatters it back (zero
// rows): nothing moves and nothing is launched.  What remains is the setup:
// the Halo kind's id check and the counts.
#include <new>

#include "ogs.hpp"

namespace libp {
namespace ogs {

void halo_t::Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Method method, const bool verbose,
                   platform_t &_platform) {
  LIBP_ABORT("ids shorter than N", (size_t)_N > ids.length(), OGS_E_ARG);
  comm = _comm;
  Setup(_N, ids.ptr(), method, verbose, _platform);
}

// ogsSetup.cpp:55-64: an ogsBase_t setup of the Halo kind, not unique.
void halo_t::Setup(const dlong _N, hlong *ids, const Method method, const bool verbose, platform_t &_platform) {
  Free();
  ogs_t base;
  base.Setup(_N, ids, Halo, method, false, verbose, _platform);
  platform = base.platform;
  N = base.N;
  Ngather = base.Ngather;
  NlocalT = base.NlocalT;
  NhaloT = base.NhaloT;
  NlocalP = base.NlocalP;
  NhaloP = base.NhaloP;
  NgatherGlobal = base.NgatherGlobal;
  kind = Halo;
  unique = false;
  gather_defined = base.gather_defined;
  gathered_halo = false;
  Nhalo = NhaloT - NhaloP;
  gatherHalo = base.gatherHalo;
  set_ = true;
}

// ogsSetup.cpp:888-914.  (The reference assigns Ngather to itself there;
// the gathered ogs's Ngather is taken here.)
void halo_t::SetupFromGather(ogs_t &ogs) {
  ogs.AssertGatherDefined();
  Free();
  platform = ogs.platform;
  comm = ogs.comm;
  N = ogs.NlocalT + ogs.NhaloT;
  Ngather = ogs.Ngather;
  Nhalo = ogs.NhaloT - ogs.NhaloP;
  NgatherGlobal = ogs.NgatherGlobal;
  kind = Halo;
  unique = ogs.unique;
  NlocalP = ogs.NlocalP;
  NlocalT = ogs.NlocalT;
  NhaloP = ogs.NhaloP;
  NhaloT = ogs.NhaloT;
  gather_defined = false;
  gathered_halo = true;
  gatherHalo = ogs.gatherHalo;
  set_ = true;
}

void halo_t::Free() {
  set_ = false;
  gatherHalo = nullptr;
  N = Ngather = NlocalT = NhaloT = NlocalP = NhaloP = Nhalo = 0;
  NgatherGlobal = 0;
  gathered_halo = false;
}

void halo_t::Check(const int k) const {
  LIBP_ABORT("halo is not set up.", !set_, OGS_E_ARG);
  LIBP_ABORT("k must be positive", k < 1, OGS_E_ARG);
}

// ---- Exchange (ogsHalo.cpp:39-196) ----
template <typename T>
void halo_t::Exchange(deviceMemory<T> o_v, const int k, hipStream_t s) const {
  ExchangeStart(o_v, k, s);
  ExchangeFinish(o_v, k, s);
}

template <typename T>
void halo_t::ExchangeStart(deviceMemory<T> o_v, const int k, hipStream_t s) const {
  Check(k);
  // gather the NhaloP owned halo values (zero of them), ogsHalo.cpp:47-61
  if (!gathered_halo) gatherHalo->Gather(deviceMemory<T>(), o_v, k, Add, NoTrans, s);
}

template <typename T>
void halo_t::ExchangeFinish(deviceMemory<T> o_v, const int k, hipStream_t s) const {
  Check(k);
  // scatter the received Nhalo values (zero of them), ogsHalo.cpp:99-110
  if (!gathered_halo) gatherHalo->Scatter(o_v, deviceMemory<T>(), k, NoTrans, s);
}

template <typename T>
void halo_t::Exchange(memory<T> v, const int k) {
  Check(k);
  (void)v;  // zero halo rows: the host vector is left as it is (ogsHalo.cpp:148-196)
}

// ---- Combine (ogsHalo.cpp:201-358) ----
template <typename T>
void halo_t::Combine(deviceMemory<T> o_v, const int k, hipStream_t s) const {
  CombineStart(o_v, k, s);
  CombineFinish(o_v, k, s);
}

template <typename T>
void halo_t::CombineStart(deviceMemory<T> o_v, const int k, hipStream_t s) const {
  Check(k);
  // gather (Add, Trans) all NhaloT halo values (zero of them), ogsHalo.cpp:209-222
  if (!gathered_halo) gatherHalo->Gather(deviceMemory<T>(), o_v, k, Add, Trans, s);
}

template <typename T>
void halo_t::CombineFinish(deviceMemory<T> o_v, const int k, hipStream_t s) const {
  Check(k);
  // scatter the combined NhaloP values back (zero of them), ogsHalo.cpp:262-272
  if (!gathered_halo) gatherHalo->Scatter(o_v, deviceMemory<T>(), k, Trans, s);
}

template <typename T>
void halo_t::Combine(memory<T> v, const int k) {
  Check(k);
  (void)v;
}

#define HALO_INSTANTIATE(T)                                                           \
  template void halo_t::Exchange(deviceMemory<T>, const int, hipStream_t) const;       \
  template void halo_t::ExchangeStart(deviceMemory<T>, const int, hipStream_t) const;  \
  template void halo_t::ExchangeFinish(deviceMemory<T>, const int, hipStream_t) const; \
  template void halo_t::Exchange(memory<T>, const int);                               \
  template void halo_t::Combine(deviceMemory<T>, const int, hipStream_t) const;        \
  template void halo_t::CombineStart(deviceMemory<T>, const int, hipStream_t) const;   \
  template void halo_t::CombineFinish(deviceMemory<T>, const int, hipStream_t) const;  \
  template void halo_t::Combine(memory<T>, const int);
HALO_INSTANTIATE(float)
HALO_INSTANTIATE(double)
HALO_INSTANTIATE(int32_t)
HALO_INSTANTIATE(int64_t)
HALO_INSTANTIATE(bfloat16)
HALO_INSTANTIATE(float_c)
#undef HALO_INSTANTIATE

}  // namespace ogs
}  // namespace libp

// ---------------------------------------------------------------------------
// Handle layer.

using namespace libp;

struct halo_handle_s {
  ogs::halo_t halo;
};

namespace {

template <class F>
int halo_guarded(F &&f) {
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
int halo_by_type(int type, F &&f) {
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

template <bool COMBINE>
int halo_move(halo_handle_t h, int type, int32_t K, void *d_v, void *stream) {
  if (!h || K < 1) return OGS_E_ARG;
  const ogs::halo_t &o = h->halo;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  return halo_by_type(type, [&](auto tag) {
    using T = decltype(tag);
    return halo_guarded([&] {
      deviceMemory<T> v(static_cast<T *>(d_v), (size_t)o.N * K);
      if (COMBINE) o.Combine(v, K, s);
      else o.Exchange(v, K, s);
    });
  });
}

}  // namespace

extern "C" {

int halo_setup(halo_handle_t *h, int32_t N, int64_t *ids, int method, int verbose) {
  if (!h) return OGS_E_ARG;
  *h = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return OGS_E_NO_DEVICE;
  halo_handle_s *o = new (std::nothrow) halo_handle_s();
  if (!o) return (int)hipErrorOutOfMemory;
  const int rc = halo_guarded([&] {
    platform_t platform;
    o->halo.Setup(N, ids, (ogs::Method)method, verbose != 0, platform);
  });
  if (rc) { delete o; return rc; }
  *h = o;
  return OGS_SUCCESS;
}

int halo_setup_from_gather(halo_handle_t *h, ogs_handle_t ogs) {
  if (!h || !ogs) return OGS_E_ARG;
  *h = nullptr;
  halo_handle_s *o = new (std::nothrow) halo_handle_s();
  if (!o) return (int)hipErrorOutOfMemory;
  const int rc = halo_guarded([&] { o->halo.SetupFromGather(ogs->ogs); });
  if (rc) { delete o; return rc; }
  *h = o;
  return OGS_SUCCESS;
}

int halo_free(halo_handle_t h) {
  if (!h) return OGS_E_ARG;
  delete h;
  return OGS_SUCCESS;
}

int halo_get_info(halo_handle_t h, halo_info_t *info) {
  if (!h || !info) return OGS_E_ARG;
  const ogs::halo_t &o = h->halo;
  info->N = o.N;
  info->Ngather = o.Ngather;
  info->NlocalT = o.NlocalT;
  info->NhaloT = o.NhaloT;
  info->NlocalP = o.NlocalP;
  info->NhaloP = o.NhaloP;
  info->Nhalo = o.Nhalo;
  info->NgatherGlobal = o.NgatherGlobal;
  info->gathered_halo = o.gathered_halo ? 1 : 0;
  info->gather_defined = o.gather_defined ? 1 : 0;
  return OGS_SUCCESS;
}

int halo_exchange(halo_handle_t h, int type, int32_t K, void *d_v, void *stream) {
  return halo_move<false>(h, type, K, d_v, stream);
}

int halo_combine(halo_handle_t h, int type, int32_t K, void *d_v, void *stream) {
  return halo_move<true>(h, type, K, d_v, stream);
}

}  // extern "C"
