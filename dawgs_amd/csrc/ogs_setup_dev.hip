// ogs_setup_dev.hip -- the single-rank ogs setup on the device: from a
// global numbering `ids` to the local operator's CSR pair, with the same
// result as the reference's setup at one rank (libs/ogs/ogsSetup.cpp:44-776):
//
//  * groups are the nodes with equal |id| (id 0 takes no part);
//  * a group's positive count is its members with id > 0 (every member for
//    the Unsigned kind); a unique setup makes the member with the smallest
//    local index the group's only positive one and writes the signs back;
//  * rows are the groups with a positive member, then the others, each class
//    in first-encounter order -- the order of their smallest local index
//    (ogsSetup.cpp:409-433);
//  * colIdsT lists a row's members ascending, colIdsN its positive members
//    (ogsSetup.cpp:637-663); the Unsigned kind aliases N to T (719, 728);
//  * the Halo kind needs exactly one positive id per group (286-288), and at
//    one rank its local operator has no rows.
//
// Method: one stable radix sort of (|id|, local index) pairs puts each
// group's members together in ascending local order; group heads and
// per-group counts come from scans over the sorted list, the row order from
// one scan over the local indices marked at each group's first member, and
// the CSR from scatters.  No host pass touches the N ids.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <vector>

#include "ogs_hip.h"
#include "ogs_internal.hpp"

namespace ogs_hip {

namespace {

constexpr int kSB = 256;

unsigned sgrid(int64_t n) {
  int64_t g = (n + kSB - 1) / kSB;
  if (g > (1 << 20)) g = 1 << 20;
  return (unsigned)(g < 1 ? 1 : g);
}

// grid of the reductions that end in a few global atomics per wave: small,
// so that the atomics on the same words do not queue up
unsigned rgrid(int64_t n) {
  const unsigned g = sgrid(n);
  return g < 2048 ? g : 2048;
}

__device__ __forceinline__ uint64_t dabs(int64_t x) { return x < 0 ? (uint64_t)0 - (uint64_t)x : (uint64_t)x; }

// max |id| and the count of nonzero ids
__global__ void k_scan_ids(int64_t N, const int64_t *__restrict__ ids, unsigned long long *__restrict__ stat) {
  unsigned long long mx = 0, nz = 0;
  for (int64_t n = blockIdx.x * (int64_t)kSB + threadIdx.x; n < N; n += (int64_t)gridDim.x * kSB) {
    const uint64_t a = dabs(ids[n]);
    mx = a > mx ? a : mx;
    nz += a != 0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mx, off);
    mx = o > mx ? o : mx;
    nz += __shfl_xor(nz, off);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&stat[0], mx);
    atomicAdd(&stat[1], nz);
  }
}

// key = |id| (id 0 sorts after every group as `sentinel`), value = local index
__global__ void k_keys(int64_t N, const int64_t *__restrict__ ids, uint64_t sentinel, uint64_t *__restrict__ key,
                       int32_t *__restrict__ val) {
  for (int64_t n = blockIdx.x * (int64_t)kSB + threadIdx.x; n < N; n += (int64_t)gridDim.x * kSB) {
    const uint64_t a = dabs(ids[n]);
    key[n] = a ? a : sentinel;
    val[n] = (int32_t)n;
  }
}

// head[i] = 1 where a group starts in the sorted list (i < M); pos[i] = 1
// for a positive member (the head for a unique setup)
__global__ void k_heads(int64_t M, const uint64_t *__restrict__ K, const int32_t *__restrict__ S,
                        const int64_t *__restrict__ ids, int kind, int unique, int32_t *__restrict__ head,
                        int32_t *__restrict__ pos) {
  for (int64_t i = blockIdx.x * (int64_t)kSB + threadIdx.x; i < M; i += (int64_t)gridDim.x * kSB) {
    const int h = (i == 0 || K[i] != K[i - 1]) ? 1 : 0;
    head[i] = h;
    pos[i] = unique ? h : (kind == OGS_UNSIGNED ? 1 : (ids[S[i]] > 0 ? 1 : 0));
  }
}

// group starts gs[0..G] (gs[G] = M) from the heads and their inclusive scan
__global__ void k_group_starts(int64_t M, const int32_t *__restrict__ head, const int32_t *__restrict__ gid1,
                               int32_t *__restrict__ gs, int32_t G) {
  for (int64_t i = blockIdx.x * (int64_t)kSB + threadIdx.x; i < M; i += (int64_t)gridDim.x * kSB)
    if (head[i]) gs[gid1[i] - 1] = (int32_t)i;
  if (blockIdx.x == 0 && threadIdx.x == 0) gs[G] = (int32_t)M;
}

// per group: positive count, class mark at its first member (1: has a
// positive member, 2: none), and the statistics
// stat: [0] groups with a positive member, [1] groups with exactly one,
// [2] largest group, [3] largest positive count
__global__ void k_groups(int32_t G, const int32_t *__restrict__ gs, const int32_t *__restrict__ S,
                         const int32_t *__restrict__ pscan, int32_t *__restrict__ gpos, int8_t *__restrict__ mark,
                         int32_t *__restrict__ stat) {
  int32_t np = 0, n1 = 0, mx = 0, mp = 0;
  for (int64_t g = blockIdx.x * (int64_t)kSB + threadIdx.x; g < G; g += (int64_t)gridDim.x * kSB) {
    const int32_t a = gs[g], b = gs[g + 1];
    const int32_t p = pscan[b] - pscan[a];
    gpos[g] = p;
    mark[S[a]] = p > 0 ? 1 : 2;
    np += p > 0;
    n1 += p == 1;
    mx = b - a > mx ? b - a : mx;
    mp = p > mp ? p : mp;
  }
  for (int off = 32; off > 0; off >>= 1) {
    np += __shfl_xor(np, off);
    n1 += __shfl_xor(n1, off);
    const int32_t o = __shfl_xor(mx, off);
    mx = o > mx ? o : mx;
    const int32_t q = __shfl_xor(mp, off);
    mp = q > mp ? q : mp;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&stat[0], np);
    atomicAdd(&stat[1], n1);
    atomicMax(&stat[2], mx);
    atomicMax(&stat[3], mp);
  }
}

// (mark == 1) + ((mark == 2) << 32): one exclusive scan gives both classes'
// first-encounter ranks
__global__ void k_mark_words(int64_t N, const int8_t *__restrict__ mark, uint64_t *__restrict__ w) {
  for (int64_t n = blockIdx.x * (int64_t)kSB + threadIdx.x; n < N; n += (int64_t)gridDim.x * kSB) {
    const int8_t m = mark[n];
    w[n] = (m == 1 ? 1ull : 0ull) | (m == 2 ? (1ull << 32) : 0ull);
  }
}

// row of each group, and the row counts (T and N) placed at the rows
__global__ void k_rows(int32_t G, int32_t NlocalP, const int32_t *__restrict__ gs, const int32_t *__restrict__ S,
                       const int32_t *__restrict__ gpos, const uint64_t *__restrict__ rank,
                       int32_t *__restrict__ row, int32_t *__restrict__ cntT, int32_t *__restrict__ cntN) {
  for (int64_t g = blockIdx.x * (int64_t)kSB + threadIdx.x; g < G; g += (int64_t)gridDim.x * kSB) {
    const uint64_t rk = rank[S[gs[g]]];
    const int32_t r = gpos[g] > 0 ? (int32_t)(rk & 0xffffffffu) : NlocalP + (int32_t)(rk >> 32);
    row[g] = r;
    cntT[r] = gs[g + 1] - gs[g];
    if (cntN) cntN[r] = gpos[g];
  }
}

// colIds: member i of group g lands at its rank within the group
__global__ void k_fill(int64_t M, const int32_t *__restrict__ S, const int32_t *__restrict__ gid1,
                       const int32_t *__restrict__ gs, const int32_t *__restrict__ row,
                       const int32_t *__restrict__ rsT, int32_t *__restrict__ ciT,
                       const int32_t *__restrict__ pos, const int32_t *__restrict__ pscan,
                       const int32_t *__restrict__ rsN, int32_t *__restrict__ ciN) {
  for (int64_t i = blockIdx.x * (int64_t)kSB + threadIdx.x; i < M; i += (int64_t)gridDim.x * kSB) {
    const int32_t g = gid1[i] - 1;
    const int32_t a = gs[g], r = row[g];
    ciT[rsT[r] + (int32_t)(i - a)] = S[i];
    if (ciN && pos[i]) ciN[rsN[r] + (pscan[i] - pscan[a])] = S[i];
  }
}

// unique: the sign each node ends up with
__global__ void k_signs(int64_t M, const int32_t *__restrict__ S, const int32_t *__restrict__ pos,
                        int64_t *__restrict__ ids) {
  for (int64_t i = blockIdx.x * (int64_t)kSB + threadIdx.x; i < M; i += (int64_t)gridDim.x * kSB) {
    const int32_t n = S[i];
    const int64_t a = (int64_t)dabs(ids[n]);
    ids[n] = pos[i] ? a : -a;
  }
}

// Device scratch of one setup, freed on every return path.
struct Arena {
  std::vector<void *> p;
  template <typename T>
  hipError_t get(T **out, size_t n) {
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
    if (e == hipSuccess) p.push_back(q);
    *out = static_cast<T *>(q);
    return e;
  }
  ~Arena() {
    for (void *q : p) (void)hipFree(q);
  }
};

template <typename F>
hipError_t with_temp(Arena &A, F &&f) {
  size_t bytes = 0;
  hipError_t e = f(nullptr, bytes);
  void *tmp = nullptr;
  if (e == hipSuccess) e = A.get(reinterpret_cast<char **>(&tmp), bytes);
  if (e == hipSuccess) e = f(tmp, bytes);
  return e;
}

int bit_width(uint64_t v) {
  int b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

// Bundle key of row r of a list in first-encounter order (lines of Ry rows,
// layers of Rz rows): r's bundle of tz layers zb, its line bundle yb within
// the layer (ty lines each), its position x along the line, its layer l and
// line m within the bundle.  Keys order bundles, then positions, then layers
// and lines, so a bundle's rows at one position sit together.
__global__ void k_bundle_keys(int32_t nrows, int32_t Ry, int32_t Rz, int ty, int tz, uint64_t *__restrict__ key,
                              int32_t *__restrict__ val) {
  const int64_t Z = (int64_t)tz * Rz, Y = (int64_t)ty * Ry, nyb = (Rz + Y - 1) / Y;
  for (int64_t r = blockIdx.x * (int64_t)kSB + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kSB) {
    const int64_t zb = r / Z, q = r - zb * Z, l = q / Rz, s = q - l * Rz;
    const int64_t yb = s / Y, t = s - yb * Y, m = t / Ry, x = t - m * Ry;
    key[r] = (uint64_t)((((zb * nyb + yb) * Ry + x) * tz + l) * ty + m);
    val[r] = (int32_t)r;
  }
}

__global__ void k_bundle_lens(int32_t nrows, const int32_t *__restrict__ rs, const int32_t *__restrict__ perm,
                              int32_t *__restrict__ len) {
  for (int64_t i = blockIdx.x * (int64_t)kSB + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * kSB) {
    const int32_t r = perm[i];
    len[i] = rs[r + 1] - rs[r];
  }
}

__global__ void k_bundle_fill(int32_t nrows, const int32_t *__restrict__ rs, const int32_t *__restrict__ ci,
                              const int32_t *__restrict__ perm, const int32_t *__restrict__ rsP,
                              int32_t *__restrict__ ciP) {
  for (int64_t i = blockIdx.x * (int64_t)kSB + threadIdx.x; i < nrows; i += (int64_t)gridDim.x * kSB) {
    const int32_t r = perm[i], a = rs[r], n = rs[r + 1] - a, o = rsP[i];
    for (int32_t k = 0; k < n; ++k) ciP[o + k] = ci[a + k];
  }
}

}  // namespace

int device_setup(int32_t N, int64_t *h_ids, int kind, int unique, hipStream_t st, DeviceSetup *out) {
  *out = DeviceSetup();
  if (N < 0 || (N > 0 && !h_ids)) return OGS_E_ARG;
  if (N == 0) return OGS_SUCCESS;
  Arena A;
  hipError_t e;
  int64_t *ids = nullptr;
  unsigned long long *ustat = nullptr;
  if ((e = A.get(&ids, N)) || (e = A.get(&ustat, 2))) return (int)e;
  if ((e = hipMemcpyAsync(ids, h_ids, sizeof(int64_t) * N, hipMemcpyHostToDevice, st)) ||
      (e = hipMemsetAsync(ustat, 0, 2 * sizeof(unsigned long long), st)))
    return (int)e;
  hipLaunchKernelGGL(k_scan_ids, dim3(rgrid(N)), dim3(kSB), 0, st, (int64_t)N, ids, ustat);
  unsigned long long hs[2];
  if ((e = hipMemcpyAsync(hs, ustat, sizeof(hs), hipMemcpyDeviceToHost, st)) || (e = hipStreamSynchronize(st)))
    return (int)e;
  const uint64_t maxAbs = hs[0];
  const int64_t M = (int64_t)hs[1];  // members (nonzero ids)
  if (M == 0) return OGS_SUCCESS;

  // stable sort of (|id|, local index): groups contiguous, members ascending
  uint64_t *k0, *k1;
  int32_t *v0, *v1;
  if ((e = A.get(&k0, N)) || (e = A.get(&k1, N)) || (e = A.get(&v0, N)) || (e = A.get(&v1, N))) return (int)e;
  const uint64_t sentinel = maxAbs + 1;
  hipLaunchKernelGGL(k_keys, dim3(sgrid(N)), dim3(kSB), 0, st, (int64_t)N, ids, sentinel, k0, v0);
  const int ebit = bit_width(sentinel);
  if ((e = with_temp(A, [&](void *t, size_t &b) {
         return hipcub::DeviceRadixSort::SortPairs(t, b, k0, k1, v0, v1, N, 0, ebit, st);
       })))
    return (int)e;
  const uint64_t *K = k1;
  const int32_t *S = v1;

  // heads, group ids, positive flags and their scan
  int32_t *head, *gid1, *pos, *pscan;
  if ((e = A.get(&head, M)) || (e = A.get(&gid1, M)) || (e = A.get(&pos, M + 1)) || (e = A.get(&pscan, M + 1)))
    return (int)e;
  hipLaunchKernelGGL(k_heads, dim3(sgrid(M)), dim3(kSB), 0, st, M, K, S, ids, kind, unique, head, pos);
  if ((e = hipMemsetAsync(pos + M, 0, sizeof(int32_t), st))) return (int)e;
  if ((e = with_temp(A, [&](void *t, size_t &b) {
         return hipcub::DeviceScan::InclusiveSum(t, b, head, gid1, (int)M, st);
       })) ||
      (e = with_temp(A, [&](void *t, size_t &b) {
         return hipcub::DeviceScan::ExclusiveSum(t, b, pos, pscan, (int)(M + 1), st);
       })))
    return (int)e;
  int32_t G = 0;
  if ((e = hipMemcpyAsync(&G, gid1 + (M - 1), sizeof(int32_t), hipMemcpyDeviceToHost, st)) ||
      (e = hipStreamSynchronize(st)))
    return (int)e;

  // groups: starts, positive counts, class marks, statistics
  int32_t *gs, *gpos, *istat;
  int8_t *mark;
  if ((e = A.get(&gs, (size_t)G + 1)) || (e = A.get(&gpos, G)) || (e = A.get(&istat, 4)) || (e = A.get(&mark, N)))
    return (int)e;
  if ((e = hipMemsetAsync(mark, 0, (size_t)N, st)) || (e = hipMemsetAsync(istat, 0, 4 * sizeof(int32_t), st)))
    return (int)e;
  hipLaunchKernelGGL(k_group_starts, dim3(sgrid(M)), dim3(kSB), 0, st, M, head, gid1, gs, G);
  hipLaunchKernelGGL(k_groups, dim3(rgrid(G)), dim3(kSB), 0, st, G, gs, S, pscan, gpos, mark, istat);
  int32_t hst[4];
  if ((e = hipMemcpyAsync(hst, istat, sizeof(hst), hipMemcpyDeviceToHost, st)) || (e = hipStreamSynchronize(st)))
    return (int)e;
  out->G = G;
  out->NlocalP = hst[0];
  out->gather_defined = unique || hst[1] == G;
  out->maxRowT = hst[2];
  out->maxRowN = kind == OGS_UNSIGNED ? hst[2] : hst[3];
  if (kind == OGS_HALO && hst[1] != G) return OGS_E_HALO_POSITIVE;

  if (unique) {  // signs back into the caller's ids
    hipLaunchKernelGGL(k_signs, dim3(sgrid(M)), dim3(kSB), 0, st, M, S, pos, ids);
    if ((e = hipMemcpyAsync(h_ids, ids, sizeof(int64_t) * N, hipMemcpyDeviceToHost, st))) return (int)e;
  }
  if (kind == OGS_HALO) {  // no local rows at one rank
    if ((e = hipStreamSynchronize(st))) return (int)e;
    return OGS_SUCCESS;
  }

  // rows by class and first encounter
  uint64_t *mw, *rank;
  int32_t *row;
  if ((e = A.get(&mw, N)) || (e = A.get(&rank, N)) || (e = A.get(&row, G))) return (int)e;
  hipLaunchKernelGGL(k_mark_words, dim3(sgrid(N)), dim3(kSB), 0, st, (int64_t)N, mark, mw);
  if ((e = with_temp(A, [&](void *t, size_t &b) {
         return hipcub::DeviceScan::ExclusiveSum(t, b, mw, rank, N, st);
       })))
    return (int)e;

  // CSR: row counts at the rows, scanned into row starts, then the fill.
  // The result arrays are handed to the caller (not in the arena).
  const bool aliased = kind == OGS_UNSIGNED;
  int32_t *cntT = nullptr, *cntN = nullptr;
  if ((e = A.get(&cntT, (size_t)G + 1)) || (!aliased && (e = A.get(&cntN, (size_t)G + 1)))) return (int)e;
  if ((e = hipMemsetAsync(cntT + G, 0, sizeof(int32_t), st)) ||
      (!aliased && (e = hipMemsetAsync(cntN + G, 0, sizeof(int32_t), st))))
    return (int)e;
  hipLaunchKernelGGL(k_rows, dim3(sgrid(G)), dim3(kSB), 0, st, G, out->NlocalP, gs, S, gpos, rank, row, cntT, cntN);
  int32_t *rsT = nullptr, *ciT = nullptr, *rsN = nullptr, *ciN = nullptr;
  auto release = [&] {
    for (int32_t *q : {rsT, ciT, rsN, ciN})
      if (q) (void)hipFree(q);
  };
  int32_t nnzN = 0;
  e = hipMalloc(&rsT, sizeof(int32_t) * ((size_t)G + 1));
  if (e == hipSuccess) e = hipMalloc(&ciT, sizeof(int32_t) * (size_t)M);
  if (e == hipSuccess)
    e = with_temp(A, [&](void *t, size_t &b) { return hipcub::DeviceScan::ExclusiveSum(t, b, cntT, rsT, G + 1, st); });
  if (e == hipSuccess && !aliased) {
    e = hipMalloc(&rsN, sizeof(int32_t) * ((size_t)G + 1));
    if (e == hipSuccess)
      e = with_temp(A, [&](void *t, size_t &b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, cntN, rsN, G + 1, st);
      });
    if (e == hipSuccess) e = hipMemcpyAsync(&nnzN, rsN + G, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipMalloc(&ciN, sizeof(int32_t) * (size_t)(nnzN ? nnzN : 1));
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fill, dim3(sgrid(M)), dim3(kSB), 0, st, M, S, gid1, gs, row, rsT, ciT, pos, pscan, rsN, ciN);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    release();
    return (int)e;
  }
  out->rsT = rsT;
  out->ciT = ciT;
  out->rsN = rsN;
  out->ciN = ciN;
  out->nnzT = (int32_t)M;
  out->nnzN = aliased ? (int32_t)M : nnzN;
  return OGS_SUCCESS;
}

int bundle_csr(const int32_t *rs, const int32_t *ci, int32_t nrows, int32_t nnz, int32_t Ry, int32_t Rz, int ty,
               int tz, hipStream_t st, int32_t **rsP, int32_t **ciP) {
  *rsP = nullptr;
  *ciP = nullptr;
  if (nrows <= 0 || Ry <= 0 || Rz <= 0 || ty <= 0 || tz <= 0) return OGS_E_ARG;
  Arena A;
  hipError_t e;
  uint64_t *k0, *k1;
  int32_t *v0, *v1, *len;
  if ((e = A.get(&k0, nrows)) || (e = A.get(&k1, nrows)) || (e = A.get(&v0, nrows)) || (e = A.get(&v1, nrows)) ||
      (e = A.get(&len, (size_t)nrows + 1)))
    return (int)e;
  hipLaunchKernelGGL(k_bundle_keys, dim3(sgrid(nrows)), dim3(kSB), 0, st, nrows, Ry, Rz, ty, tz, k0, v0);
  // every key lies below (z bundles) * (line bundles per layer) * Ry * tz * ty
  const int64_t Z = (int64_t)tz * Rz, Y = (int64_t)ty * Ry;
  const uint64_t bound = (uint64_t)((nrows + Z - 1) / Z) * (uint64_t)((Rz + Y - 1) / Y) * (uint64_t)Ry *
                         (uint64_t)tz * (uint64_t)ty;
  const int ebit = bit_width(bound);
  if ((e = with_temp(A, [&](void *t, size_t &b) {
         return hipcub::DeviceRadixSort::SortPairs(t, b, k0, k1, v0, v1, nrows, 0, ebit, st);
       })))
    return (int)e;
  hipLaunchKernelGGL(k_bundle_lens, dim3(sgrid(nrows)), dim3(kSB), 0, st, nrows, rs, v1, len);
  if ((e = hipMemsetAsync(len + nrows, 0, sizeof(int32_t), st))) return (int)e;
  int32_t *r = nullptr, *c = nullptr;
  e = hipMalloc(&r, sizeof(int32_t) * ((size_t)nrows + 1));
  if (e == hipSuccess) e = hipMalloc(&c, sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  if (e == hipSuccess)
    e = with_temp(A, [&](void *t, size_t &b) { return hipcub::DeviceScan::ExclusiveSum(t, b, len, r, nrows + 1, st); });
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_bundle_fill, dim3(sgrid(nrows)), dim3(kSB), 0, st, nrows, rs, ci, v1, r, c);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // the arena is freed on return
  if (e != hipSuccess) {
    if (r) (void)hipFree(r);
    if (c) (void)hipFree(c);
    return (int)e;
  }
  *rsP = r;
  *ciP = c;
  return OGS_SUCCESS;
}

}  // namespace ogs_hip
