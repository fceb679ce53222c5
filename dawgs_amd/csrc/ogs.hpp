// ogs.hpp -- libp::ogs at one rank on MI355X: the ogs_t object of the
// reference's include/ogs.hpp:216-344, with ogsOperator_t's device half
// served by the gfx950 kernels behind include/ogs_hip.h.
//
// Names, argument meaning and error behaviour follow the reference:
//   Setup(N, ids, comm, kind, method, unique, verbose, platform)  ogs.hpp:221-228
//   GatherScatter / Gather / Scatter (+ Start/Finish)            ogs.hpp:232-341
//   enums Type/Op/Transpose/Method/Kind                          ogs.hpp:179-200
//   LIBP_ABORT -> libp::exception                                utils.hpp:57-98
// Differences: no OCCA (device buffers are plain HIP allocations), comm_t is
// a one-rank communicator, and the host-memory overloads stage through the
// device rather than running an OpenMP loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ogs_hip.h"

namespace libp {

using dlong = int32_t;  // include/types.h:60
using hlong = int64_t;  // include/types.h:52

// libp::exception (libs/core/exception.cpp:75-88), carrying an OGS_E_* or
// hipError_t code so the C-ABI can return it.
class exception : public std::runtime_error {
 public:
  exception(const std::string &what, int code) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

#define LIBP_ABORT(msg, cond, code)                                                   \
  do {                                                                                \
    if (cond) {                                                                       \
      std::ostringstream os_;                                                         \
      os_ << "ABORT: " << msg << " (" << __FILE__ << ":" << __LINE__ << ")";          \
      throw ::libp::exception(os_.str(), (code));                                     \
    }                                                                                 \
  } while (0)

#define LIBP_HIP_CHECK(call)                                                          \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    LIBP_ABORT("HIP error " << hipGetErrorString(e_) << " in " #call, e_ != hipSuccess, (int)e_); \
  } while (0)

// Host array with shared ownership (memory.hpp:34-264, reduced).
template <typename T>
class memory {
 public:
  memory() = default;
  explicit memory(size_t n) : v_(std::make_shared<std::vector<T>>(n)) {}
  memory(size_t n, const T &init) : v_(std::make_shared<std::vector<T>>(n, init)) {}
  T *ptr() { return v_ ? v_->data() : nullptr; }
  const T *ptr() const { return v_ ? v_->data() : nullptr; }
  size_t length() const { return v_ ? v_->size() : 0; }
  T &operator[](size_t i) { return (*v_)[i]; }
  const T &operator[](size_t i) const { return (*v_)[i]; }

 private:
  std::shared_ptr<std::vector<T>> v_;
};

// Device array (memory.hpp:284-454, reduced).  Owning when made by
// platform_t::malloc, non-owning when it wraps a caller's pointer.
template <typename T>
class deviceMemory {
 public:
  deviceMemory() = default;
  deviceMemory(T *p, size_t n) : p_(p), n_(n) {}  // non-owning view
  static deviceMemory owning(size_t n) {
    deviceMemory d;
    T *p = nullptr;
    if (n) LIBP_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    d.own_ = std::shared_ptr<T>(p, [](T *q) { if (q) (void)hipFree(q); });
    d.p_ = p;
    d.n_ = n;
    return d;
  }
  // Takes ownership of a device allocation made elsewhere.
  static deviceMemory adopt(T *p, size_t n) {
    deviceMemory d;
    d.own_ = std::shared_ptr<T>(p, [](T *q) { if (q) (void)hipFree(q); });
    d.p_ = p;
    d.n_ = n;
    return d;
  }
  T *ptr() const { return p_; }
  size_t length() const { return n_; }
  deviceMemory operator+(size_t off) const {
    deviceMemory d = *this;
    d.p_ = p_ + off;
    d.n_ = off <= n_ ? n_ - off : 0;
    return d;
  }

 private:
  std::shared_ptr<T> own_;
  T *p_ = nullptr;
  size_t n_ = 0;
};

// One-rank communicator (comm.hpp:68-565 collapses to identities at size 1).
class comm_t {
 public:
  int rank() const { return 0; }
  int size() const { return 1; }
};

// Device + stream (platform.hpp:56-188, reduced).
class platform_t {
 public:
  platform_t() { LIBP_HIP_CHECK(hipGetDevice(&device)); }
  explicit platform_t(hipStream_t s) : stream(s) { LIBP_HIP_CHECK(hipGetDevice(&device)); }
  int device = 0;
  hipStream_t stream = nullptr;
  void finish() const { LIBP_HIP_CHECK(hipStreamSynchronize(stream)); }
  template <typename T>
  deviceMemory<T> malloc(size_t n) const { return deviceMemory<T>::owning(n); }
  template <typename T>
  deviceMemory<T> malloc(size_t n, const T *src) const {
    deviceMemory<T> d = deviceMemory<T>::owning(n);
    if (n) LIBP_HIP_CHECK(hipMemcpy(d.ptr(), src, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
};

namespace ogs {

// ogs.hpp:179, plus the precision-form storage types (Bf16, FloatC: no
// counterpart in the reference; see ogs_type_t in ogs_hip.h)
typedef enum { Float, Double, Int32, Int64, Bf16, FloatC } Type;
typedef enum { Add, Mul, Max, Min } Op;                        // ogs.hpp:191
typedef enum { Sym, NoTrans, Trans } Transpose;                // ogs.hpp:194
typedef enum { Auto, Pairwise, CrystalRouter, AllToAll } Method;  // ogs.hpp:197
typedef enum { Unsigned, Signed, Halo } Kind;                  // ogs.hpp:200

template <typename T> struct ogsType;                          // ogsUtils.hpp:49-65
template <> struct ogsType<float> { static constexpr Type get() { return Float; } };
template <> struct ogsType<double> { static constexpr Type get() { return Double; } };
template <> struct ogsType<int32_t> { static constexpr Type get() { return Int32; } };
template <> struct ogsType<int64_t> { static constexpr Type get() { return Int64; } };

// Precision-form storage: a bfloat16 bit pattern, and a float whose Add
// reduces with two-sum compensation.
struct bfloat16 {
  uint16_t bits;
};
struct float_c {
  float v;
};
template <> struct ogsType<bfloat16> { static constexpr Type get() { return Bf16; } };
template <> struct ogsType<float_c> { static constexpr Type get() { return FloatC; } };

// The CSR "Z" operator (ogsOperator.hpp:36-132): host lists, their device
// copies and the device plan that runs them.
class ogsOperator_t {
 public:
  dlong Ncols = 0, NrowsN = 0, NrowsT = 0, nnzN = 0, nnzT = 0;
  memory<dlong> rowStartsN, rowStartsT;
  mutable memory<dlong> colIdsN, colIdsT;
  deviceMemory<dlong> o_rowStartsN, o_rowStartsT, o_colIdsN, o_colIdsT;
  dlong NrowBlocksN = 0, NrowBlocksT = 0;
  memory<dlong> blockRowStartsN, blockRowStartsT;
  Kind kind = Unsigned;
  bool aliased = false;  // Unsigned: N lists are the T lists (ogsSetup.cpp:719,728)

  static constexpr int gatherNodesPerBlock = OGS_MAX_ROW;  // ogsOperator.hpp:123

  ogsOperator_t() = default;
  ~ogsOperator_t();
  ogsOperator_t(const ogsOperator_t &) = delete;
  ogsOperator_t &operator=(const ogsOperator_t &) = delete;

  void setupRowBlocks();                 // ogsOperator.cpp:447-512
  void upload(const platform_t &platform, int fields = OGS_PLAN_ALL_K);  // device copies + plan
  void createPlan(const platform_t &platform, int fields = OGS_PLAN_ALL_K);  // plan over the device lists
  // Host colIds, copied from the device on first use when the setup built
  // the lists there (only the CSR query reads them).
  const memory<dlong> &hostColIdsT() const;
  const memory<dlong> &hostColIdsN() const;
  ogs_hip_plan_t plan() const { return plan_; }

  template <typename T>
  void Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, int k, Op op, Transpose trans, hipStream_t s) const;
  template <typename T>
  void Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, int k, Transpose trans, hipStream_t s) const;
  template <typename T>
  void GatherScatter(deviceMemory<T> o_v, int k, Op op, Transpose trans, hipStream_t s) const;

 private:
  ogs_hip_plan_t plan_ = nullptr;
};

// ogsBase_t + ogs_t (ogsBase.hpp:45-106, ogs.hpp:216-344) at one rank.
class ogs_t {
 public:
  platform_t platform;
  comm_t comm;

  dlong N = 0;
  dlong Ngather = 0;
  dlong NlocalT = 0, NhaloT = 0, NlocalP = 0, NhaloP = 0;
  hlong NgatherGlobal = 0;
  Kind kind = Unsigned;
  bool unique = false;
  bool gather_defined = false;
  // field counts the local operator's plan prepares for (OGS_PLAN_ALL_K or
  // OGS_PLAN_K1), read by Setup; no counterpart in the reference
  int planFields = OGS_PLAN_ALL_K;

  void Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Kind _kind, const Method method,
             const bool _unique, const bool verbose, platform_t &_platform);
  // Same, on a raw host array (the C-ABI entry).
  void Setup(const dlong _N, hlong *ids, const Kind _kind, const Method method, const bool _unique,
             const bool verbose, platform_t &_platform);
  void Free();

  // The reference's signatures launch on platform.stream; each has an
  // overload taking the stream explicitly, which never writes the object, so
  // several host threads may share one ogs_t on different streams.
  template <typename T> void GatherScatter(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
    GatherScatter(o_v, k, op, trans, platform.stream);
  }
  template <typename T> void GatherScatterStart(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
    GatherScatterStart(o_v, k, op, trans, platform.stream);
  }
  template <typename T> void GatherScatterFinish(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
    GatherScatterFinish(o_v, k, op, trans, platform.stream);
  }
  template <typename T> void GatherScatter(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                                           hipStream_t s) const;
  template <typename T> void GatherScatterStart(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                                                hipStream_t s) const;
  template <typename T> void GatherScatterFinish(deviceMemory<T> o_v, const int k, const Op op, const Transpose trans,
                                                 hipStream_t s) const;
  template <typename T> void GatherScatter(memory<T> v, const int k, const Op op, const Transpose trans);

  template <typename T> void Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
    Gather(o_gv, o_v, k, op, trans, platform.stream);
  }
  template <typename T> void GatherStart(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
    GatherStart(o_gv, o_v, k, op, trans, platform.stream);
  }
  template <typename T> void GatherFinish(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op, const Transpose trans) {
    GatherFinish(o_gv, o_v, k, op, trans, platform.stream);
  }
  template <typename T> void Gather(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op,
                                    const Transpose trans, hipStream_t s) const;
  template <typename T> void GatherStart(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op,
                                         const Transpose trans, hipStream_t s) const;
  template <typename T> void GatherFinish(deviceMemory<T> o_gv, deviceMemory<T> o_v, const int k, const Op op,
                                          const Transpose trans, hipStream_t s) const;
  template <typename T> void Gather(memory<T> gv, const memory<T> v, const int k, const Op op, const Transpose trans);

  template <typename T> void Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans) {
    Scatter(o_v, o_gv, k, trans, platform.stream);
  }
  template <typename T> void ScatterStart(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans) {
    ScatterStart(o_v, o_gv, k, trans, platform.stream);
  }
  template <typename T> void ScatterFinish(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans) {
    ScatterFinish(o_v, o_gv, k, trans, platform.stream);
  }
  template <typename T> void Scatter(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k, const Transpose trans,
                                     hipStream_t s) const;
  template <typename T> void ScatterStart(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k,
                                          const Transpose trans, hipStream_t s) const;
  template <typename T> void ScatterFinish(deviceMemory<T> o_v, deviceMemory<T> o_gv, const int k,
                                           const Transpose trans, hipStream_t s) const;
  template <typename T> void Scatter(memory<T> v, const memory<T> gv, const int k, const Transpose trans);

  // ogsSetup.cpp:867-886: GlobalToLocal[n] = the gathered index (row) of node
  // n, -1 where no row holds it (id 0); GlobalToLocal has N entries.
  void SetupGlobalToLocalMapping(memory<dlong> GlobalToLocal);
  void SetupGlobalToLocalMapping(dlong *GlobalToLocal);

  const ogsOperator_t &local() const { return *gatherLocal; }
  void AssertGatherDefined() const;  // ogsSetup.cpp:861-864

 private:
  std::shared_ptr<ogsOperator_t> gatherLocal;
  std::shared_ptr<ogsOperator_t> gatherHalo;
  friend class halo_t;
};

// halo_t (ogs.hpp:347-397) at one rank.  A Halo-kind setup validates the ids
// (every group needs exactly one positive id, ogsSetup.cpp:286-288) and
// counts them; with a single rank no node is shared with another rank, so
// the halo operator has no rows and the exchange no neighbours, and
// Exchange/Combine move nothing (ogsHalo.cpp:39-359 with zero halo rows and
// Nhalo = 0).  SetupFromGather (ogsSetup.cpp:888-914) takes the counts of a
// gathered ogs_t, whose halo nodes would follow its NlocalT local rows.
class halo_t {
 public:
  platform_t platform;
  comm_t comm;

  dlong N = 0;
  dlong Ngather = 0;
  dlong NlocalT = 0, NhaloT = 0, NlocalP = 0, NhaloP = 0;
  hlong NgatherGlobal = 0;
  Kind kind = Halo;
  bool unique = false;
  bool gather_defined = false;
  bool gathered_halo = false;
  dlong Nhalo = 0;  // extra received nodes, NhaloT - NhaloP

  void Setup(const dlong _N, memory<hlong> ids, comm_t _comm, const Method method, const bool verbose,
             platform_t &_platform);
  void Setup(const dlong _N, hlong *ids, const Method method, const bool verbose, platform_t &_platform);
  void SetupFromGather(ogs_t &ogs);
  void Free();

  // platform.stream, as in ogs_t; the overloads taking a stream never write
  // the object.
  template <typename T> void Exchange(deviceMemory<T> o_v, const int k) { Exchange(o_v, k, platform.stream); }
  template <typename T> void ExchangeStart(deviceMemory<T> o_v, const int k) { ExchangeStart(o_v, k, platform.stream); }
  template <typename T> void ExchangeFinish(deviceMemory<T> o_v, const int k) { ExchangeFinish(o_v, k, platform.stream); }
  template <typename T> void Exchange(deviceMemory<T> o_v, const int k, hipStream_t s) const;
  template <typename T> void ExchangeStart(deviceMemory<T> o_v, const int k, hipStream_t s) const;
  template <typename T> void ExchangeFinish(deviceMemory<T> o_v, const int k, hipStream_t s) const;
  template <typename T> void Exchange(memory<T> v, const int k);

  template <typename T> void Combine(deviceMemory<T> o_v, const int k) { Combine(o_v, k, platform.stream); }
  template <typename T> void CombineStart(deviceMemory<T> o_v, const int k) { CombineStart(o_v, k, platform.stream); }
  template <typename T> void CombineFinish(deviceMemory<T> o_v, const int k) { CombineFinish(o_v, k, platform.stream); }
  template <typename T> void Combine(deviceMemory<T> o_v, const int k, hipStream_t s) const;
  template <typename T> void CombineStart(deviceMemory<T> o_v, const int k, hipStream_t s) const;
  template <typename T> void CombineFinish(deviceMemory<T> o_v, const int k, hipStream_t s) const;
  template <typename T> void Combine(memory<T> v, const int k);

 private:
  bool set_ = false;
  std::shared_ptr<ogsOperator_t> gatherHalo;  // empty at one rank
  void Check(const int k) const;
};

}  // namespace ogs
}  // namespace libp

// The opaque object behind ogs_handle_t (ogs.cpp, ogs_halo.cpp).
struct ogs_handle_s {
  libp::ogs::ogs_t ogs;
};
