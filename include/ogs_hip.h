/*
 * ogs_hip.h -- C-ABI of the MI355X (gfx950) rank-local gather-scatter.
 *
 * Three layers, all in libogs_hip.so:
 *
 *  1. Operator layer (ogs_hip_*): the device half of ogsOperator_t.  A plan
 *     wraps one caller-owned CSR pair (N and T lists) on the device and owns
 *     only what it derives from them (row partitions, column-sorted packs of
 *     the lists, each column's row).  It replaces the OCCA kernel tables and device overloads
 *     of the reference:
 *       ogsOperator_t::Gather  (deviceMemory)  libs/ogs/ogsOperator.cpp:159-187
 *       ogsOperator_t::Scatter (deviceMemory)  libs/ogs/ogsOperator.cpp:266-293
 *       ogsOperator_t::GatherScatter (device)  libs/ogs/ogsOperator.cpp:393-432
 *       OKL gatherScatter/gather/scatter/extract  libs/ogs/okl/ogsKernels.okl:32-177
 *       InitializeKernels (JIT tables)          libs/ogs/ogsUtils.cpp:46-123
 *
 *     With it, the batch entry points (ogs_hip_batch_*): many (plan, K,
 *     buffer) GatherScatters in one HIP graph launch, no counterpart in the
 *     reference.
 *
 *  2. Handle layer (ogs_*): the ogs_t object of include/ogs.hpp:216-344 at
 *     one rank -- Setup from a global numbering, then GatherScatter / Gather
 *     / Scatter on device buffers.
 *
 *  3. Halo layer (halo_*): the halo_t object of include/ogs.hpp:347-397 at
 *     one rank.
 *
 * Conventions (all entry points):
 *  - Device buffers are plain device pointers.  `stream` is a hipStream_t
 *    passed as void* (NULL = the null stream).  Launches are asynchronous.
 *  - Values for K fields are interleaved, field fastest: v[k + id*K]
 *    (ogsOperator.cpp:92-94).
 *  - Return 0 on success; a negative OGS_E_* code for a bad argument or a
 *    broken contract; a positive hipError_t for a HIP failure.  Nothing
 *    throws across the boundary.  ogs_hip_error_string() names a code.
 *  - Plans and handles are immutable after creation; launches from several
 *    host threads on different streams may share them.
 */
#ifndef OGS_HIP_H
#define OGS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Value types.  0..3 follow ogs::Type (include/ogs.hpp:179) and reduce
 * exactly as the reference's host operator does.  4 and 5 are the precision
 * form (no counterpart in the reference): narrow storage, Add accumulated per
 * row in float, in CSR order, with a two-sum running error (the cascaded sum
 * "Sum2" of Ogita, Rump and Oishi, 2005), written back rounded to nearest
 * (even) in the storage type; Max/Min exact; Mul -> OGS_E_NOT_SUPPORTED.
 *   OGS_BF16     bfloat16 storage (2 bytes per value)
 *   OGS_FLOAT_C  float storage (4 bytes), compensated Add; OGS_FLOAT keeps
 *                the reference's plain float sum. */
typedef enum {
  OGS_FLOAT = 0, OGS_DOUBLE = 1, OGS_INT32 = 2, OGS_INT64 = 3, OGS_BF16 = 4, OGS_FLOAT_C = 5
} ogs_type_t;
/* ogs::Op (include/ogs.hpp:191) */
typedef enum { OGS_ADD = 0, OGS_MUL = 1, OGS_MAX = 2, OGS_MIN = 3 } ogs_op_t;
/* ogs::Transpose (include/ogs.hpp:194) */
typedef enum { OGS_SYM = 0, OGS_NOTRANS = 1, OGS_TRANS = 2 } ogs_trans_t;
/* ogs::Kind (include/ogs.hpp:200) */
typedef enum { OGS_UNSIGNED = 0, OGS_SIGNED = 1, OGS_HALO = 2 } ogs_kind_t;
/* ogs::Method (include/ogs.hpp:197); at one rank every method is the
 * empty pairwise exchange (ogsAuto.cpp:154-156). */
typedef enum { OGS_AUTO = 0, OGS_PAIRWISE = 1, OGS_CRYSTAL_ROUTER = 2, OGS_ALLTOALL = 3 } ogs_method_t;

enum {
  OGS_SUCCESS = 0,
  OGS_E_ARG = -1,            /* null pointer, bad enum, K < 1, negative size */
  OGS_E_ROW_TOO_LONG = -2,   /* row longer than the limit (ogsOperator.cpp:460-465) */
  OGS_E_GATHER_UNDEF = -3,   /* Gather/Scatter on a non-unique ogs (ogsSetup.cpp:861-864) */
  OGS_E_INVALID_SETUP = -4,  /* unique with Unsigned or Halo kind (ogsSetup.cpp:98-100) */
  OGS_E_HALO_POSITIVE = -5,  /* Halo group without exactly one positive id (ogsSetup.cpp:286-288) */
  OGS_E_NOT_SUPPORTED = -6,  /* type/op combination not offered (e.g. bf16 Mul) */
  OGS_E_CSR = -7,            /* CSR arrays inconsistent with the sizes given */
  OGS_E_NO_DEVICE = -8,      /* no HIP device visible */
  OGS_E_TOO_LARGE = -9       /* input beyond the fixed scratch of a no-allocation entry point */
};

/* Longest row the setup accepts (ogsOperator.hpp:123 gatherNodesPerBlock). */
#define OGS_MAX_ROW 512

const char *ogs_hip_error_string(int code);

/* ------------------------------------------------------------------ */
/* 1. Operator layer                                                   */

typedef struct ogs_hip_plan_s *ogs_hip_plan_t;

/* Wrap a CSR pair already on the device.  Rows [0,NrowsT) are described by
 * d_rowStartsT/d_colIdsT (all members) and d_rowStartsN/d_colIdsN (positive
 * members only); both row-start arrays have NrowsT+1 entries and rows
 * [NrowsN, NrowsT) of the N list are empty (else OGS_E_CSR, as for
 * decreasing row starts or a column outside [0, Ncols)).  Pass the same
 * pointers for N and T for an Unsigned operator (ogsSetup.cpp:719,728).
 * Synchronises `stream`.  Replaces ogsOperator_t::setupRowBlocks' device
 * upload (ogsOperator.cpp:510-511) with the plan's own partitions. */
int ogs_hip_plan_create(ogs_hip_plan_t *plan, int32_t Ncols, int32_t NrowsN, int32_t NrowsT,
                        const int32_t *d_rowStartsN, const int32_t *d_colIdsN,
                        const int32_t *d_rowStartsT, const int32_t *d_colIdsT, void *stream);
/* ogs_hip_plan_create with a choice of the field counts the plan prepares
 * for (no counterpart in the reference; ogs_hip_plan_create is
 * OGS_PLAN_ALL_K).  OGS_PLAN_K1 keeps only what the K = 1 launches read: the
 * row partitions, the column maps, the pack of the partition the K = 1
 * kernels run on (a Signed operator: and of the two-list kernel's), the
 * 16-bit delta pack, and no 32-bit pack where the delta pack serves every
 * K = 1 kernel of an Unsigned operator.  Any other K still runs, from the
 * CSR (without packs: slower, same results).  48^3 p=7 Unsigned: 0.35 GB
 * instead of 2.06 GB. */
enum { OGS_PLAN_ALL_K = 0, OGS_PLAN_K1 = 1 };
int ogs_hip_plan_create_k(ogs_hip_plan_t *plan, int32_t Ncols, int32_t NrowsN, int32_t NrowsT,
                          const int32_t *d_rowStartsN, const int32_t *d_colIdsN,
                          const int32_t *d_rowStartsT, const int32_t *d_colIdsT, int fields, void *stream);
/* Plans are reference counted (no counterpart in the reference, whose
 * operators are owned by shared_ptr, ogsBase.hpp:62-63): create returns one
 * reference, retain adds one, destroy drops one and frees the plan with the
 * last.  A batch holds a reference to each of its plans until it is
 * destroyed.  A plan from ogs_get_plan also keeps the handle's CSR lists
 * alive; for a plan from ogs_hip_plan_create the caller's CSR arrays must
 * outlive the plan's last holder. */
int ogs_hip_plan_retain(ogs_hip_plan_t plan);
int ogs_hip_plan_destroy(ogs_hip_plan_t plan);

/* ogsOperator_t::GatherScatter (ogsOperator.cpp:311-360 semantics, device
 * path 393-432, but with `op` honoured): in place on d_v (Ncols*K values).
 * Sym: rows NrowsT, gather T, scatter T; NoTrans: rows NrowsT, gather N,
 * scatter T; Trans: rows NrowsN, gather T, scatter N. */
int ogs_hip_gather_scatter(ogs_hip_plan_t plan, int type, int op, int trans, int32_t K,
                           void *d_v, void *stream);
/* ogsOperator_t::Gather (ogsOperator.cpp:60-97, device 159-187):
 * d_gv[k + row*K] = OP over the row's list of d_v.  NoTrans: rows [0,NrowsN)
 * over the N lists; Trans/Sym: rows [0,NrowsT) over the T lists. */
int ogs_hip_gather(ogs_hip_plan_t plan, int type, int op, int trans, int32_t K,
                   void *d_gv, const void *d_v, void *stream);
/* ogsOperator_t::Scatter (ogsOperator.cpp:206-238, device 266-293):
 * d_v[k + col*K] = d_gv[k + row*K].  Trans: rows [0,NrowsN) over the N lists;
 * NoTrans/Sym: rows [0,NrowsT) over the T lists.  Over a list of at least
 * Ncols/2 entries (colsT / colsN of ogs_hip_plan_info) it runs by columns
 * when a column's K values take at most 16 bytes or K is not 2, 3, 4, 8 or
 * 16, else by rows (a plan made for K=1 only: by columns at K=1).  A column held by several rows (never so for Setup-built
 * lists; the reference's row-parallel loops leave any one of their values)
 * gets the largest such row's value by columns, any one of them by rows. */
int ogs_hip_scatter(ogs_hip_plan_t plan, int type, int trans, int32_t K,
                    void *d_v, const void *d_gv, void *stream);
/* OKL extract (ogsKernels.okl:167-177; host ogsOperator.cpp:540-559):
 * d_gq[k + n*K] = d_q[k + d_ids[n]*K] for n < N. */
int ogs_hip_extract(int type, int32_t N, int32_t K, const int32_t *d_ids,
                    const void *d_q, void *d_gq, void *stream);

/* Plan facts, for reporting: block count and width of the K=1 partition
 * used by GatherScatter, the longest row, nnz of both lists, how many K=1
 * blocks of the T list read their columns in the plan's column-sorted order
 * (packedK1) or in CSR order (plainK1), and the layer distance (in blocks)
 * that the K=1 GatherScatter / Gather block order is tiled for (tileD; 0 =
 * block id order).  Plan creation picks a tiled order, per kernel, only where
 * it measures faster on scratch values.  colsT / colsN: 1 where the plan
 * holds each column's row for that list (a list with at least Ncols/2
 * entries), so that Scatter can run by columns (see ogs_hip_scatter); a
 * plan made for K=1 only keeps them as row-delta codes (rowMapE > 0) and
 * the rows of the codes' fallback chunks.
 * deltaE: exception words per 512-entry range of the K=1 GatherScatter's
 * 16-bit delta pack (0 = none; the 32-bit pack serves), deltaFallback: its
 * blocks that run from the 32-bit pack.  ownerMap: 1 where the plan holds
 * each column's owner (GatherScatter NoTrans by columns: a Signed operator
 * with distinct T columns and at most one N member per row, each in its own
 * T row), nMask: 1 where it holds the N-column mask (GatherScatter Trans over
 * the T list).  bundleRy, bundleRz, bundleTy, bundleTz (all 0: none): the
 * K=1 GatherScatter runs over bundled blocks -- rows of bundleTy
 * neighbouring lines and bundleTz layers side by side, from the line and
 * layer row strides bundleRy, bundleRz the plan found in the T list (a
 * structured mesh's shared nodes at low degree) -- kept where that measured
 * faster; bundleTileD: their block order's layer distance (0 = block id
 * order), bundleFallback: their blocks that run from 32-bit words.
 * rowMapE, rowMapFallback: exception words per 512-column range of the T
 * list's column -> row map in 4-bit row-delta codes, which Scatter at K=1
 * by columns reads (0: none, the int32 map serves), and its 2,048-column
 * chunks that read the int32 map.  byteItems, byteFallback: exception items
 * of the K=1 GatherScatter's one-byte codes over the delta pack's blocks
 * (both 0 and deltaE > 0: the 16-bit words serve), and their blocks that run
 * from the 16-bit words.  byteK3: 1 where the one-list GatherScatter and
 * the Gather with 2 or 3 fields of float, double, int32, int64 or
 * float-compensated values (4 fields of the 4-byte ones) run over those
 * codes' blocks too (rows of at most 8 members, no block left in CSR order, blocks
 * enough to fill the device).  bytes:
 * device memory the plan owns (partitions, packs, column maps, block
 * orders), beside the caller's CSR. */
typedef struct {
  int32_t NrowsN, NrowsT, nnzN, nnzT, maxRowT, maxRowN;
  int32_t nblocksK1, widthK1, packedK1, plainK1, Ncols, tileD;
  int32_t colsT, colsN, deltaE, deltaFallback, ownerMap, nMask;
  int32_t bundleRy, bundleRz, bundleTy, bundleTz, bundleTileD, bundleFallback;
  int32_t rowMapE, rowMapFallback, byteItems, byteFallback, byteK3;
  int64_t bytes;
} ogs_hip_plan_info_t;
int ogs_hip_plan_info(ogs_hip_plan_t plan, ogs_hip_plan_info_t *info);

/* Batched GatherScatter: many plans (meshes), each with its own K
 * interleaved fields, in one HIP graph launch.  No counterpart in the
 * reference (it launches one kernel per operator and field set).  Every
 * descriptor runs GatherScatter Sym with the batch's type and op on its
 * buffer d_v (Ncols*K values); buffers must not overlap.  The graph holds
 * one kernel node per field-count class present (K = 1, 2, 3, 4, 8, 16, each
 * with its compiled shape, one workgroup per (descriptor, block) item), or
 * one node running all those classes, and a persistent grid-stride node for
 * any other K.  Buffers are bound at creation.  With tune != 0 the creation
 * times candidate configurations (one node or one per class, block width of
 * K > 1 classes, item order) on scratch copies of the
 * buffers (best of two passes each) and keeps the fastest; the buffers
 * themselves are not touched.
 * Creation allocates and synchronises; ogs_hip_batch_run only launches the
 * graph.  Concurrent runs of one batch race on its buffers.  The batch
 * retains its plans (see ogs_hip_plan_retain), so a plan's other holders
 * may be destroyed first; the value buffers stay the caller's to keep. */
typedef struct {
  ogs_hip_plan_t plan;
  int32_t K;
  void *d_v;
} ogs_hip_batch_desc_t;
typedef struct ogs_hip_batch_s *ogs_hip_batch_t;
int ogs_hip_batch_create(ogs_hip_batch_t *batch, int type, int op, const ogs_hip_batch_desc_t *descs,
                         int32_t n, int tune, void *stream);
int ogs_hip_batch_run(ogs_hip_batch_t batch, void *stream);
int ogs_hip_batch_destroy(ogs_hip_batch_t batch);
/* What the creation chose: work items (blocks over all descriptors),
 * kernel nodes in the graph, whether K>1 classes use the next narrower
 * block width, item order (0 = larger descriptors first, 1 = descriptor
 * order), whether all field-count classes share one kernel node (fused),
 * the tuning time / best run, and whether the graph launches each
 * descriptor through its plan's own GatherScatter kernels instead (plans;
 * a tuning candidate, picked where it ran fastest). */
typedef struct {
  int32_t items, nodes, narrow, order, fused, candidates;
  float tune_ms, best_us;
  int32_t plans;
} ogs_hip_batch_info_t;
int ogs_hip_batch_info(ogs_hip_batch_t batch, ogs_hip_batch_info_t *info);

/* ------------------------------------------------------------------ */
/* 2. Handle layer: ogs_t at one rank                                  */

typedef struct ogs_handle_s *ogs_handle_t;

/* ogs_t::Setup (include/ogs.hpp:221-228; libs/ogs/ogsSetup.cpp:44-190).
 * ids is a host array of N global ids (0 = ignored, negative = flagged);
 * with unique=1 the chosen signs are written back into ids, as the
 * reference does (ogsSetup.cpp:139-140).  The sole positive member of each
 * group is the member with the smallest local index.  |id| must fit an
 * int64 (INT64_MIN is no valid id: the reference takes its abs()).  Builds
 * the CSR on the device (the host keeps the row starts) and creates the
 * operator plan. */
int ogs_setup(ogs_handle_t *h, int32_t N, int64_t *ids, int kind, int method,
              int unique, int verbose);
/* ogs_setup whose operator plan prepares for `fields` (OGS_PLAN_ALL_K, what
 * ogs_setup does, or OGS_PLAN_K1; see ogs_hip_plan_create_k). */
int ogs_setup_k(ogs_handle_t *h, int32_t N, int64_t *ids, int kind, int method,
                int unique, int verbose, int fields);
int ogs_free(ogs_handle_t h);

typedef struct {
  int32_t N, Ngather, NlocalT, NhaloT, NlocalP, NhaloP;
  int64_t NgatherGlobal;
  int32_t kind, unique, gather_defined;
  int32_t NrowsN, NrowsT, nnzN, nnzT;
  int32_t NrowBlocksN, NrowBlocksT; /* the reference's 512-nnz row blocks */
} ogs_info_t;
int ogs_get_info(ogs_handle_t h, ogs_info_t *info);

/* Copy the local operator's host CSR out (arrays sized from ogs_get_info:
 * rowStarts NrowsT+1, colIds nnz, blockRowStarts NrowBlocks+1).  Any
 * pointer may be NULL to skip it. */
int ogs_get_csr(ogs_handle_t h, int32_t *rowStartsN, int32_t *colIdsN,
                int32_t *rowStartsT, int32_t *colIdsT,
                int32_t *blockRowStartsN, int32_t *blockRowStartsT);
/* The operator plan of the local operator (the handle holds one reference;
 * a caller that keeps the plan past ogs_free retains it first). */
ogs_hip_plan_t ogs_get_plan(ogs_handle_t h);

/* ogs_t::GatherScatter on a device buffer (ogs.cpp:39-123). */
/* ogs_t::SetupGlobalToLocalMapping (ogsSetup.cpp:867-886): GlobalToLocal (a
 * host array of N entries) receives each node's gathered index (its row),
 * -1 for nodes with id 0.  Synchronises. */
int ogs_setup_global_to_local(ogs_handle_t h, int32_t *GlobalToLocal);
/* ogs_setup_global_to_local without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The map is staged in
 * a fixed per-GPU scratch array of 2^22 entries (16 MiB, held by each GPU the
 * library runs on until the process ends): N above 2^22 returns
 * OGS_E_TOO_LARGE.  Runs on the handle's stream and synchronises it.  Every
 * call on one GPU shares that GPU's scratch, so calls that overlap on one GPU
 * race. */
int ogs_setup_global_to_local_noalloc(ogs_handle_t h, int32_t *GlobalToLocal);

int ogs_gather_scatter(ogs_handle_t h, int type, int op, int trans, int32_t K,
                       void *d_v, void *stream);
/* ogs_t::Gather on device buffers (ogs.cpp:202-298); needs a gather-defined
 * (unique) setup.  d_gv holds Ngather*K values. */
int ogs_gather(ogs_handle_t h, int type, int op, int trans, int32_t K,
               void *d_gv, const void *d_v, void *stream);
/* ogs_t::Scatter on device buffers (ogs.cpp:392-488). */
int ogs_scatter(ogs_handle_t h, int type, int trans, int32_t K,
                void *d_v, const void *d_gv, void *stream);

/* ------------------------------------------------------------------ */
/* 3. halo_t at one rank (include/ogs.hpp:347-397)                    */
/* halo_setup: halo_t::Setup (ogsSetup.cpp:55-64), a Halo-kind setup: every
 * group of equal |id| needs exactly one positive id (else
 * OGS_E_HALO_POSITIVE, ogsSetup.cpp:286-288).  halo_setup_from_gather:
 * halo_t::SetupFromGather (ogsSetup.cpp:888-914) on a gather-defined
 * (unique) ogs handle (else OGS_E_GATHER_UNDEF), which must outlive the
 * halo.  Exchange / Combine (ogsHalo.cpp:39-359) on d_v (N*K values):
 * at one rank no node is shared with another rank, so there are no halo
 * rows and nothing moves (Nhalo = 0); they validate and return. */
typedef struct halo_handle_s *halo_handle_t;
typedef struct {
  int32_t N, Ngather, NlocalT, NhaloT, NlocalP, NhaloP, Nhalo;
  int64_t NgatherGlobal;
  int32_t gathered_halo, gather_defined;
} halo_info_t;
int halo_setup(halo_handle_t *h, int32_t N, int64_t *ids, int method, int verbose);
int halo_setup_from_gather(halo_handle_t *h, ogs_handle_t ogs);
int halo_free(halo_handle_t h);
int halo_get_info(halo_handle_t h, halo_info_t *info);
int halo_exchange(halo_handle_t h, int type, int32_t K, void *d_v, void *stream);
int halo_combine(halo_handle_t h, int type, int32_t K, void *d_v, void *stream);

#ifdef __cplusplus
}
#endif
#endif
