/*
 * ogs_oracle.c -- CPU restatement of the reference's single-rank host
 * gather-scatter path (paranumal/dawgs).  TEST INFRASTRUCTURE: only the
 * tests, the smoke check and the benchmark's CPU-baseline leg use it, as
 * the checker.  The product never links it.
 *
 * What it restates, and what pins each part (the pins live in tests/):
 *   oracle_make_ids          dawgsMain.cpp:49-143      closed-form id counts and multiplicities
 *   oracle_setup             ogsSetup.cpp:69-190,      brute-force dictionary setup on tiny boxes,
 *                            192-331, 333-566,         CSR invariants, row-length histograms
 *                            569-776
 *   oracle_setup_row_blocks  ogsOperator.cpp:447-512   independent numpy greedy model, block counts
 *   oracle_op_*              ogsOperator.cpp:36-55,    gslib gs_vec (the reference's own -cc
 *                            60-378, 540-559           comparator, built from its sources),
 *                                                      known answers (v=1 -> multiplicity, small
 *                                                      integers, Scatter(Gather) == GS)
 *
 * Departures from the reference, all deliberate:
 *  - Signed+unique picks the group member with the smallest local index as
 *    the sole positive node.  The reference picks rand()%size after an
 *    unstable parallel sort (ogsSetup.cpp:264-275), which no run can
 *    reproduce; ogs.hpp:96-99 only promises "arbitrary but consistent".
 *  - Integer Add/Mul are done in unsigned arithmetic so that overflow wraps
 *    modulo 2^n by definition (the reference's signed overflow is UB in C++
 *    but wraps on the hardware it runs on).
 */
#include "ogs_oracle.h"

#include <float.h>
#include <limits.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* MakeIds (dawgsMain.cpp:49-143) at comm size 1: Factor3 gives 1x1x1 and
 * RankDecomp3 gives rank coordinates (0,0,0), so NX=nx and all offsets are 0. */
void oracle_make_ids(int nx, int ny, int nz, int p, hlong *ids) {
  const hlong NX = nx, NY = ny;
  const int Nq = p + 1, Np = Nq * Nq * Nq;
  for (int K = 0; K < nz; K++)
    for (int J = 0; J < ny; J++)
      for (int I = 0; I < nx; I++) {
        hlong *ids_e = ids + ((hlong)I + (hlong)J * nx + (hlong)K * nx * ny) * Np;
        const hlong baseId = (hlong)I * p + (hlong)J * p * (p * NX + 1) +
                             (hlong)K * p * (p * NX + 1) * (p * NY + 1) + 1;
        for (int k = 0; k < Nq; k++)
          for (int j = 0; j < Nq; j++)
            for (int i = 0; i < Nq; i++)
              ids_e[i + j * Nq + k * Nq * Nq] =
                  i + (hlong)j * (p * NX + 1) + (hlong)k * (p * NX + 1) * (p * NY + 1) + baseId;
      }
}

/* ------------------------------------------------------------------ */
/* parallelNode_t (ogsUtils.hpp:36-47) */
typedef struct {
  dlong localId;
  hlong baseId;
  dlong newId;
  int sign;
  int rank;
  int destRank;
} node_t;

static hlong habs(hlong x) { return x < 0 ? -x : x; }

/* FindSharedNodes sort key (ogsSetup.cpp:245-248) */
static int cmp_abs_base(const void *pa, const void *pb) {
  const hlong a = habs(((const node_t *)pa)->baseId), b = habs(((const node_t *)pb)->baseId);
  return (a > b) - (a < b);
}
/* ConstructSharedNodes sort key (ogsSetup.cpp:341-347): group by |baseId|,
 * positive ids first inside a group. */
static int cmp_abs_then_positive(const void *pa, const void *pb) {
  const node_t *a = (const node_t *)pa, *b = (const node_t *)pb;
  const hlong aa = habs(a->baseId), bb = habs(b->baseId);
  if (aa < bb) return -1;
  if (aa > bb) return 1;
  return (a->baseId < b->baseId) - (a->baseId > b->baseId);
}

/* permute (ogsUtils.hpp:69-81): A[P(n)] <- A[n], with P = newId or localId */
static void permute_by_newId(dlong N, node_t *A) {
  for (dlong n = 0; n < N; ++n) {
    dlong pn = A[n].newId;
    while (pn != n) {
      node_t t = A[n]; A[n] = A[pn]; A[pn] = t;
      pn = A[n].newId;
    }
  }
}
static void permute_by_localId(dlong N, node_t *A) {
  for (dlong n = 0; n < N; ++n) {
    dlong pn = A[n].localId;
    while (pn != n) {
      node_t t = A[n]; A[n] = A[pn]; A[pn] = t;
      pn = A[n].localId;
    }
  }
}

static void op_zero(oracle_op_t *op) { memset(op, 0, sizeof(*op)); }

static void op_free(oracle_op_t *op) {
  free(op->rowStartsT);
  free(op->colIdsT);
  if (!op->aliased) {
    free(op->rowStartsN);
    free(op->colIdsN);
  }
  free(op->blockRowStartsN);
  free(op->blockRowStartsT);
  op_zero(op);
}

void oracle_free(oracle_ogs_t *ogs) {
  op_free(&ogs->local);
  op_free(&ogs->halo);
}

/* setupRowBlocks (ogsOperator.cpp:447-512).  Both block lists iterate the
 * T rows and both end at NrowsT (507-508). */
int oracle_setup_row_blocks(oracle_op_t *op, dlong gatherNodesPerBlock) {
  dlong blockSumN = 0, blockSumT = 0;
  dlong NrowBlocksN = 0, NrowBlocksT = 0;
  if (op->NrowsN) NrowBlocksN++;
  if (op->NrowsT) NrowBlocksT++;
  for (dlong i = 0; i < op->NrowsT; i++) {
    const dlong rowSizeN = op->rowStartsN[i + 1] - op->rowStartsN[i];
    const dlong rowSizeT = op->rowStartsT[i + 1] - op->rowStartsT[i];
    if (rowSizeN > gatherNodesPerBlock || rowSizeT > gatherNodesPerBlock)
      return OR_E_ROW_TOO_LONG; /* 460-465 */
    if (blockSumN + rowSizeN > gatherNodesPerBlock) { NrowBlocksN++; blockSumN = rowSizeN; }
    else blockSumN += rowSizeN;
    if (blockSumT + rowSizeT > gatherNodesPerBlock) { NrowBlocksT++; blockSumT = rowSizeT; }
    else blockSumT += rowSizeT;
  }
  free(op->blockRowStartsN);
  free(op->blockRowStartsT);
  op->blockRowStartsN = (dlong *)calloc((size_t)NrowBlocksN + 1, sizeof(dlong));
  op->blockRowStartsT = (dlong *)calloc((size_t)NrowBlocksT + 1, sizeof(dlong));
  blockSumN = 0; blockSumT = 0;
  NrowBlocksN = 0; NrowBlocksT = 0;
  if (op->NrowsN) NrowBlocksN++;
  if (op->NrowsT) NrowBlocksT++;
  for (dlong i = 0; i < op->NrowsT; i++) {
    const dlong rowSizeN = op->rowStartsN[i + 1] - op->rowStartsN[i];
    const dlong rowSizeT = op->rowStartsT[i + 1] - op->rowStartsT[i];
    if (blockSumN + rowSizeN > gatherNodesPerBlock) { op->blockRowStartsN[NrowBlocksN++] = i; blockSumN = rowSizeN; }
    else blockSumN += rowSizeN;
    if (blockSumT + rowSizeT > gatherNodesPerBlock) { op->blockRowStartsT[NrowBlocksT++] = i; blockSumT = rowSizeT; }
    else blockSumT += rowSizeT;
  }
  op->blockRowStartsN[NrowBlocksN] = op->NrowsT;
  op->blockRowStartsT[NrowBlocksT] = op->NrowsT;
  op->NrowBlocksN = NrowBlocksN;
  op->NrowBlocksT = NrowBlocksT;
  return OR_OK;
}

/* LocalSignedSetup / LocalUnsignedSetup / LocalHaloSetup CSR fill
 * (ogsSetup.cpp:569-847).  `select` picks which nodes belong to this
 * operator: |sign|==1 for the local operator, |sign|==2 for the halo one.
 * posN(node) says whether the node goes into the N list. */
static void fill_operator(oracle_op_t *op, dlong Nids, const node_t *nodes, int wantAbsSign,
                          int haloKindN, int aliasNT, dlong NrowsN, dlong NrowsT, dlong Ncols) {
  op_zero(op);
  op->Ncols = Ncols;
  op->NrowsN = NrowsN;
  op->NrowsT = NrowsT;
  dlong *cntN = (dlong *)calloc((size_t)NrowsT + 1, sizeof(dlong));
  dlong *cntT = (dlong *)calloc((size_t)NrowsT + 1, sizeof(dlong));
  for (dlong i = 0; i < Nids; i++) {
    if (abs(nodes[i].sign) != wantAbsSign) continue;
    const dlong gid = nodes[i].newId;
    /* Signed/Unsigned tally N by baseId>0 (596,599); Halo by sign==2 (798) */
    const int inN = haloKindN ? (nodes[i].sign == 2) : (nodes[i].baseId > 0);
    if (inN) cntN[gid]++;
    cntT[gid]++;
  }
  op->rowStartsT = (dlong *)malloc(((size_t)NrowsT + 1) * sizeof(dlong));
  op->rowStartsT[0] = 0;
  for (dlong r = 0; r < NrowsT; r++) op->rowStartsT[r + 1] = op->rowStartsT[r] + cntT[r];
  op->nnzT = op->rowStartsT[NrowsT];
  op->colIdsT = (dlong *)malloc(((size_t)op->nnzT + 1) * sizeof(dlong));
  if (aliasNT) {
    op->aliased = 1;
    op->rowStartsN = op->rowStartsT;
    op->colIdsN = op->colIdsT;
    op->nnzN = op->nnzT;
  } else {
    op->rowStartsN = (dlong *)malloc(((size_t)NrowsT + 1) * sizeof(dlong));
    op->rowStartsN[0] = 0;
    for (dlong r = 0; r < NrowsT; r++) op->rowStartsN[r + 1] = op->rowStartsN[r] + cntN[r];
    op->nnzN = op->rowStartsN[NrowsT];
    op->colIdsN = (dlong *)malloc(((size_t)op->nnzN + 1) * sizeof(dlong));
  }
  memset(cntN, 0, ((size_t)NrowsT + 1) * sizeof(dlong));
  memset(cntT, 0, ((size_t)NrowsT + 1) * sizeof(dlong));
  /* nodes are in ascending local order, so colIds come out ascending per row (637-663) */
  for (dlong i = 0; i < Nids; i++) {
    if (abs(nodes[i].sign) != wantAbsSign) continue;
    const dlong gid = nodes[i].newId;
    if (!aliasNT) {
      const int inN = haloKindN ? (nodes[i].sign == 2) : (nodes[i].baseId > 0);
      if (inN) op->colIdsN[op->rowStartsN[gid] + cntN[gid]++] = nodes[i].localId;
    }
    op->colIdsT[op->rowStartsT[gid] + cntT[gid]++] = nodes[i].localId;
  }
  free(cntN);
  free(cntT);
}

/* ogsBase_t::Setup (ogsSetup.cpp:69-190) with comm size 1. */
int oracle_setup(dlong N, hlong *ids, int kind, int unique, oracle_ogs_t *ogs) {
  memset(ogs, 0, sizeof(*ogs));
  if ((kind == OR_UNSIGNED && unique) || (kind == OR_HALO && unique)) return OR_E_INVALID_SETUP; /* 98-100 */
  if (kind < OR_UNSIGNED || kind > OR_HALO) return OR_E_ARG;
  ogs->N = N;
  ogs->kind = kind;
  ogs->unique = unique;

  /* squeeze out zero ids (102-121) */
  dlong Nids = 0;
  for (dlong n = 0; n < N; n++) if (ids[n] != 0) Nids++;
  node_t *nodes = (node_t *)malloc(((size_t)Nids + 1) * sizeof(node_t));
  Nids = 0;
  for (dlong n = 0; n < N; n++) {
    if (ids[n] != 0) {
      nodes[Nids].localId = Nids;
      nodes[Nids].baseId = (kind == OR_UNSIGNED) ? habs(ids[n]) : ids[n];
      nodes[Nids].rank = 0;
      nodes[Nids].destRank = 0; /* |id| % size with size 1 */
      nodes[Nids].newId = 0;
      nodes[Nids].sign = 0;
      Nids++;
    }
  }

  /* FindSharedNodes (192-331).  At size 1 the Alltoallv exchanges are the
   * identity, so recvNodes is nodes in send order (= local order). */
  for (dlong n = 0; n < Nids; n++) nodes[n].newId = n; /* 240-242 */
  qsort(nodes, (size_t)Nids, sizeof(node_t), cmp_abs_base);
  int is_unique = 1;
  dlong start = 0;
  for (dlong n = 0; n < Nids; n++) {
    if (n == Nids - 1 || habs(nodes[n].baseId) != habs(nodes[n + 1].baseId)) {
      const dlong end = n + 1;
      int positiveCount = 0;
      if (unique) {
        /* one node per group becomes the sole positive one (264-275);
         * deterministic choice: the smallest local index */
        const hlong baseId = habs(nodes[start].baseId);
        dlong m = start;
        for (dlong i = start; i < end; i++) {
          nodes[i].baseId = -baseId;
          if (nodes[i].localId < nodes[m].localId) m = i;
        }
        nodes[m].baseId = baseId;
        positiveCount = 1;
      } else {
        for (dlong i = start; i < end; i++) if (nodes[i].baseId > 0) positiveCount++;
        if (positiveCount != 1) is_unique = 0;
      }
      if (kind == OR_HALO && positiveCount != 1) { free(nodes); return OR_E_HALO_POSITIVE; } /* 286-288 */
      /* one rank: never shared via MPI (290-304) */
      for (dlong i = start; i < end; i++) nodes[i].sign = 1;
      start = n + 1;
    }
  }
  ogs->gather_defined = (is_unique == 1); /* 312-313 */
  permute_by_newId(Nids, nodes);          /* 326 */

  /* ConstructSharedNodes (333-566) */
  qsort(nodes, (size_t)Nids, sizeof(node_t), cmp_abs_then_positive);
  dlong NbaseIds = 0, NlocalT = 0, NlocalP = 0, NhaloT = 0, NhaloP = 0;
  start = 0;
  for (dlong n = 0; n < Nids; n++) {
    if (n == Nids - 1 || habs(nodes[n].baseId) != habs(nodes[n + 1].baseId)) {
      const dlong end = n + 1;
      int sign = abs(nodes[start].sign);
      if (nodes[start].baseId < 0) { /* no positive id in this group (359-366) */
        sign = -sign;
        for (dlong i = start; i < end; i++) nodes[i].sign = sign;
      }
      if (abs(sign) == 1) { NlocalT++; if (sign == 1) NlocalP++; }
      else { NhaloT++; if (sign == 2) NhaloP++; }
      for (dlong i = start; i < end; i++) nodes[i].newId = NbaseIds;
      NbaseIds++;
      start = end;
    }
  }
  ogs->Ngather = NlocalP + NhaloP;   /* 388 */
  ogs->NgatherGlobal = ogs->Ngather; /* 391-392 at size 1 */
  ogs->NlocalP = NlocalP; ogs->NlocalT = NlocalT;
  ogs->NhaloP = NhaloP;   ogs->NhaloT = NhaloT;

  permute_by_localId(Nids, nodes); /* 407 */

  /* rows numbered by first encounter in local order, positive groups first (409-433) */
  dlong *indexMap = (dlong *)malloc(((size_t)NbaseIds + 1) * sizeof(dlong));
  for (dlong i = 0; i < NbaseIds; i++) indexMap[i] = -1;
  dlong localCntN = 0, localCntT = NlocalP, haloCntN = 0, haloCntT = NhaloP;
  for (dlong n = 0; n < Nids; n++) {
    const dlong newId = nodes[n].newId;
    if (indexMap[newId] == -1) {
      if (nodes[n].sign == 1) indexMap[newId] = localCntN++;
      else if (nodes[n].sign == -1) indexMap[newId] = localCntT++;
      else if (nodes[n].sign == 2) indexMap[newId] = haloCntN++;
      else indexMap[newId] = haloCntT++;
    }
    nodes[n].newId = indexMap[newId];
  }
  free(indexMap);

  /* record real local ids; write chosen signs back when unique (133-144) */
  Nids = 0;
  for (dlong n = 0; n < N; n++) {
    if (ids[n] != 0) {
      nodes[Nids].localId = n;
      if (unique) ids[n] = nodes[Nids].baseId;
      Nids++;
    }
  }

  if (kind == OR_SIGNED) {          /* LocalSignedSetup 569-683 */
    fill_operator(&ogs->local, Nids, nodes, 1, 0, 0, NlocalP, NlocalT, N);
    fill_operator(&ogs->halo, Nids, nodes, 2, 0, 0, NhaloP, NhaloT, N);
  } else if (kind == OR_UNSIGNED) { /* LocalUnsignedSetup 686-776 */
    fill_operator(&ogs->local, Nids, nodes, 1, 0, 1, NlocalP, NlocalT, N);
    fill_operator(&ogs->halo, Nids, nodes, 2, 0, 1, NhaloP, NhaloT, N);
  } else {                          /* LocalHaloSetup 779-847: no local operator */
    op_zero(&ogs->local);
    ogs->local.rowStartsT = (dlong *)calloc(1, sizeof(dlong));
    ogs->local.rowStartsN = (dlong *)calloc(1, sizeof(dlong));
    fill_operator(&ogs->halo, Nids, nodes, 2, 1, 0, NhaloP, NhaloT, N);
  }
  free(nodes);

  int rc = OR_OK;
  if (kind != OR_HALO) rc = oracle_setup_row_blocks(&ogs->local, 512); /* 681, 774 */
  if (rc == OR_OK) rc = oracle_setup_row_blocks(&ogs->halo, 512);      /* 682, 775, 846 */
  if (rc != OR_OK) oracle_free(ogs);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Host operator loops.  Op functors and inits: ogsOperator.cpp:36-55. */
#define OP_ADD_F(a, b) (a) += (b)
#define OP_MUL_F(a, b) (a) *= (b)
#define OP_ADD_I(a, b, U) (a) = (T)((U)(a) + (U)(b))
#define OP_MUL_I(a, b, U) (a) = (T)((U)(a) * (U)(b))
#define OP_MAX(a, b) (a) = ((b) > (a)) ? (b) : (a)
#define OP_MIN(a, b) (a) = ((b) < (a)) ? (b) : (a)

/* Gather (60-97): gv[k+n*K] = OP_{g in row n} v[k+colIds[g]*K], from init */
#define DEF_GATHER(NAME, T, INIT, APPLY)                                               \
  static void NAME(dlong Nrows, const dlong *rowStarts, const dlong *colIds, int K,  \
                   const T *v, T *gv) {                                               \
    _Pragma("omp parallel for")                                                       \
    for (dlong n = 0; n < Nrows; ++n) {                                               \
      const dlong start = rowStarts[n], end = rowStarts[n + 1];                       \
      for (int k = 0; k < K; ++k) {                                                   \
        T val = INIT;                                                                 \
        for (dlong g = start; g < end; ++g) { APPLY(val, v[k + (size_t)colIds[g] * K]); } \
        gv[k + (size_t)n * K] = val;                                                  \
      }                                                                               \
    }                                                                                 \
  }

/* GatherScatter (311-360): per row and field, reduce the gather list then
 * write the scatter list, in place */
#define DEF_GS(NAME, T, INIT, APPLY)                                                   \
  static void NAME(dlong Nrows, const dlong *gS, const dlong *gI, const dlong *sS,   \
                   const dlong *sI, int K, T *v) {                                    \
    _Pragma("omp parallel for")                                                       \
    for (dlong n = 0; n < Nrows; ++n) {                                               \
      const dlong gstart = gS[n], gend = gS[n + 1];                                   \
      const dlong sstart = sS[n], send = sS[n + 1];                                   \
      for (int k = 0; k < K; ++k) {                                                   \
        T val = INIT;                                                                 \
        for (dlong g = gstart; g < gend; ++g) { APPLY(val, v[k + (size_t)gI[g] * K]); } \
        for (dlong s = sstart; s < send; ++s) v[k + (size_t)sI[s] * K] = val;         \
      }                                                                               \
    }                                                                                 \
  }

#define APPLY_ADD_F(a, b) OP_ADD_F(a, b)
#define APPLY_MUL_F(a, b) OP_MUL_F(a, b)
#define APPLY_ADD_I32(a, b) (a) = (int32_t)((uint32_t)(a) + (uint32_t)(b))
#define APPLY_MUL_I32(a, b) (a) = (int32_t)((uint32_t)(a) * (uint32_t)(b))
#define APPLY_ADD_I64(a, b) (a) = (int64_t)((uint64_t)(a) + (uint64_t)(b))
#define APPLY_MUL_I64(a, b) (a) = (int64_t)((uint64_t)(a) * (uint64_t)(b))

#define DEF_ALL(TN, T, ADD, MUL, MAXV)                        \
  DEF_GATHER(gather_##TN##_add, T, (T)0, ADD)                \
  DEF_GATHER(gather_##TN##_mul, T, (T)1, MUL)                \
  DEF_GATHER(gather_##TN##_max, T, (T)(-(MAXV)), OP_MAX)     \
  DEF_GATHER(gather_##TN##_min, T, (T)(MAXV), OP_MIN)        \
  DEF_GS(gs_##TN##_add, T, (T)0, ADD)                        \
  DEF_GS(gs_##TN##_mul, T, (T)1, MUL)                        \
  DEF_GS(gs_##TN##_max, T, (T)(-(MAXV)), OP_MAX)             \
  DEF_GS(gs_##TN##_min, T, (T)(MAXV), OP_MIN)                \
  /* Scatter (206-238): v[k+colIds[g]*K] = gv[k+n*K] */      \
  static void scatter_##TN(dlong Nrows, const dlong *rowStarts, const dlong *colIds, \
                           int K, const T *gv, T *v) {                              \
    _Pragma("omp parallel for")                                                     \
    for (dlong n = 0; n < Nrows; ++n) {                                             \
      const dlong start = rowStarts[n], end = rowStarts[n + 1];                     \
      for (dlong g = start; g < end; ++g)                                           \
        for (int k = 0; k < K; ++k) v[k + (size_t)colIds[g] * K] = gv[k + (size_t)n * K]; \
    }                                                                               \
  }                                                                                 \
  /* extract (540-559): gq[k+n*K] = q[k+ids[n]*K] */                                \
  static void extract_##TN(dlong N, int K, const dlong *ids, const T *q, T *gq) {  \
    for (dlong n = 0; n < N; ++n) {                                                 \
      const dlong gid = ids[n];                                                     \
      for (int k = 0; k < K; ++k) gq[k + (size_t)n * K] = q[k + (size_t)gid * K];   \
    }                                                                               \
  }

DEF_ALL(f32, float, APPLY_ADD_F, APPLY_MUL_F, FLT_MAX)
DEF_ALL(f64, double, APPLY_ADD_F, APPLY_MUL_F, DBL_MAX)
DEF_ALL(i32, int32_t, APPLY_ADD_I32, APPLY_MUL_I32, INT32_MAX)
DEF_ALL(i64, int64_t, APPLY_ADD_I64, APPLY_MUL_I64, INT64_MAX)

/* ------------------------------------------------------------------ */
/* Precision form: OR_BF16 (bfloat16 storage) and OR_FLOAT_C (float
 * storage).  No counterpart in the reference: this restates the library's
 * own specification.  Add reduces each row in float, in CSR order, with a
 * two-sum running error (cascaded sum "Sum2", Ogita, Rump and Oishi, SIAM J.
 * Sci. Comput. 26(6), 2005): s = +0, c = 0; per member x: t = s + x,
 * z = t - s, e = (s - (t - z)) + (x - z), s = t, c += e; result = s + c when
 * s is finite, else s.  bf16 results round to nearest even (NaN quieted).
 * Max/Min compare-select the widened stored values.  Mul: not offered. */
typedef struct { float s, c; } comp_t;

static inline void comp_add(comp_t *a, float x) {
  const float t = a->s + x;
  const float z = t - a->s;
  const float e = (a->s - (t - z)) + (x - z);
  a->s = t;
  a->c += e;
}
static inline float comp_fin(comp_t a) { return isfinite(a.s) ? a.s + a.c : a.s; }

static inline float bf16_widen(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static inline uint16_t bf16_round(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
#define BF16_MAXF 3.38953139e38f /* 0x7f7f */

#define LOAD_F32(x) (x)
#define STORE_F32(f) (f)
#define LOAD_BF16(x) bf16_widen(x)
#define STORE_BF16(f) bf16_round(f)

#define DEF_GATHER_C(NAME, T, LOAD, STORE)                                             \
  static void NAME(dlong Nrows, const dlong *rowStarts, const dlong *colIds, int K,  \
                   const T *v, T *gv) {                                               \
    _Pragma("omp parallel for")                                                       \
    for (dlong n = 0; n < Nrows; ++n) {                                               \
      const dlong start = rowStarts[n], end = rowStarts[n + 1];                       \
      for (int k = 0; k < K; ++k) {                                                   \
        comp_t a = {0.0f, 0.0f};                                                      \
        for (dlong g = start; g < end; ++g) comp_add(&a, LOAD(v[k + (size_t)colIds[g] * K])); \
        gv[k + (size_t)n * K] = STORE(comp_fin(a));                                   \
      }                                                                               \
    }                                                                                 \
  }
#define DEF_GS_C(NAME, T, LOAD, STORE)                                                 \
  static void NAME(dlong Nrows, const dlong *gS, const dlong *gI, const dlong *sS,   \
                   const dlong *sI, int K, T *v) {                                    \
    _Pragma("omp parallel for")                                                       \
    for (dlong n = 0; n < Nrows; ++n) {                                               \
      for (int k = 0; k < K; ++k) {                                                   \
        comp_t a = {0.0f, 0.0f};                                                      \
        for (dlong g = gS[n]; g < gS[n + 1]; ++g) comp_add(&a, LOAD(v[k + (size_t)gI[g] * K])); \
        const T val = STORE(comp_fin(a));                                             \
        for (dlong s = sS[n]; s < sS[n + 1]; ++s) v[k + (size_t)sI[s] * K] = val;      \
      }                                                                               \
    }                                                                                 \
  }
/* bf16 Max/Min: compare-select in float, exact on the stored values */
#define DEF_GATHER_MM16(NAME, INIT, CMP)                                               \
  static void NAME(dlong Nrows, const dlong *rowStarts, const dlong *colIds, int K,  \
                   const uint16_t *v, uint16_t *gv) {                                 \
    _Pragma("omp parallel for")                                                       \
    for (dlong n = 0; n < Nrows; ++n) {                                               \
      for (int k = 0; k < K; ++k) {                                                   \
        float a = INIT;                                                               \
        for (dlong g = rowStarts[n]; g < rowStarts[n + 1]; ++g) {                     \
          const float x = bf16_widen(v[k + (size_t)colIds[g] * K]);                   \
          a = (x CMP a) ? x : a;                                                      \
        }                                                                             \
        gv[k + (size_t)n * K] = bf16_round(a);                                        \
      }                                                                               \
    }                                                                                 \
  }
#define DEF_GS_MM16(NAME, INIT, CMP)                                                   \
  static void NAME(dlong Nrows, const dlong *gS, const dlong *gI, const dlong *sS,   \
                   const dlong *sI, int K, uint16_t *v) {                             \
    _Pragma("omp parallel for")                                                       \
    for (dlong n = 0; n < Nrows; ++n) {                                               \
      for (int k = 0; k < K; ++k) {                                                   \
        float a = INIT;                                                               \
        for (dlong g = gS[n]; g < gS[n + 1]; ++g) {                                   \
          const float x = bf16_widen(v[k + (size_t)gI[g] * K]);                       \
          a = (x CMP a) ? x : a;                                                      \
        }                                                                             \
        const uint16_t val = bf16_round(a);                                           \
        for (dlong s = sS[n]; s < sS[n + 1]; ++s) v[k + (size_t)sI[s] * K] = val;      \
      }                                                                               \
    }                                                                                 \
  }

DEF_GATHER_C(gather_f32c_add, float, LOAD_F32, STORE_F32)
DEF_GS_C(gs_f32c_add, float, LOAD_F32, STORE_F32)
DEF_GATHER_C(gather_bf16_add, uint16_t, LOAD_BF16, STORE_BF16)
DEF_GS_C(gs_bf16_add, uint16_t, LOAD_BF16, STORE_BF16)
DEF_GATHER_MM16(gather_bf16_max, -BF16_MAXF, >)
DEF_GATHER_MM16(gather_bf16_min, BF16_MAXF, <)
DEF_GS_MM16(gs_bf16_max, -BF16_MAXF, >)
DEF_GS_MM16(gs_bf16_min, BF16_MAXF, <)

static void scatter_u16(dlong Nrows, const dlong *rowStarts, const dlong *colIds, int K, const uint16_t *gv,
                        uint16_t *v) {
#pragma omp parallel for
  for (dlong n = 0; n < Nrows; ++n)
    for (dlong g = rowStarts[n]; g < rowStarts[n + 1]; ++g)
      for (int k = 0; k < K; ++k) v[k + (size_t)colIds[g] * K] = gv[k + (size_t)n * K];
}
static void extract_u16(dlong N, int K, const dlong *ids, const uint16_t *q, uint16_t *gq) {
  for (dlong n = 0; n < N; ++n)
    for (int k = 0; k < K; ++k) gq[k + (size_t)n * K] = q[k + (size_t)ids[n] * K];
}

float oracle_bf16_to_float(uint16_t b) { return bf16_widen(b); }
uint16_t oracle_float_to_bf16(float f) { return bf16_round(f); }

typedef void (*gather_fn)(dlong, const dlong *, const dlong *, int, const void *, void *);
typedef void (*gs_fn)(dlong, const dlong *, const dlong *, const dlong *, const dlong *, int, void *);
typedef void (*scatter_fn)(dlong, const dlong *, const dlong *, int, const void *, void *);
typedef void (*extract_fn)(dlong, int, const dlong *, const void *, void *);

/* table order: [type][op] with op order Add, Mul, Max, Min (ogs.hpp:191);
 * types 4 (bf16) and 5 (float, compensated Add) are the precision form,
 * whose Mul entries are empty (not offered). */
#define NTYPES 6
static const gather_fn gather_tab[NTYPES][4] = {
    {(gather_fn)gather_f32_add, (gather_fn)gather_f32_mul, (gather_fn)gather_f32_max, (gather_fn)gather_f32_min},
    {(gather_fn)gather_f64_add, (gather_fn)gather_f64_mul, (gather_fn)gather_f64_max, (gather_fn)gather_f64_min},
    {(gather_fn)gather_i32_add, (gather_fn)gather_i32_mul, (gather_fn)gather_i32_max, (gather_fn)gather_i32_min},
    {(gather_fn)gather_i64_add, (gather_fn)gather_i64_mul, (gather_fn)gather_i64_max, (gather_fn)gather_i64_min},
    {(gather_fn)gather_bf16_add, NULL, (gather_fn)gather_bf16_max, (gather_fn)gather_bf16_min},
    {(gather_fn)gather_f32c_add, NULL, (gather_fn)gather_f32_max, (gather_fn)gather_f32_min}};
static const gs_fn gs_tab[NTYPES][4] = {
    {(gs_fn)gs_f32_add, (gs_fn)gs_f32_mul, (gs_fn)gs_f32_max, (gs_fn)gs_f32_min},
    {(gs_fn)gs_f64_add, (gs_fn)gs_f64_mul, (gs_fn)gs_f64_max, (gs_fn)gs_f64_min},
    {(gs_fn)gs_i32_add, (gs_fn)gs_i32_mul, (gs_fn)gs_i32_max, (gs_fn)gs_i32_min},
    {(gs_fn)gs_i64_add, (gs_fn)gs_i64_mul, (gs_fn)gs_i64_max, (gs_fn)gs_i64_min},
    {(gs_fn)gs_bf16_add, NULL, (gs_fn)gs_bf16_max, (gs_fn)gs_bf16_min},
    {(gs_fn)gs_f32c_add, NULL, (gs_fn)gs_f32_max, (gs_fn)gs_f32_min}};
static const scatter_fn scatter_tab[NTYPES] = {(scatter_fn)scatter_f32, (scatter_fn)scatter_f64,
                                               (scatter_fn)scatter_i32, (scatter_fn)scatter_i64,
                                               (scatter_fn)scatter_u16, (scatter_fn)scatter_f32};
static const extract_fn extract_tab[NTYPES] = {(extract_fn)extract_f32, (extract_fn)extract_f64,
                                               (extract_fn)extract_i32, (extract_fn)extract_i64,
                                               (extract_fn)extract_u16, (extract_fn)extract_f32};

static int bad_args(int type, int opr, int trans, int K) {
  return type < 0 || type >= NTYPES || opr < 0 || opr > 3 || trans < 0 || trans > 2 || K < 1;
}

/* Gather list selection (71-79): NoTrans -> N rows/lists, else T */
int oracle_op_gather(const oracle_op_t *op, int type, int opr, int trans, int K,
                     const void *v, void *gv) {
  if (bad_args(type, opr, trans, K)) return OR_E_ARG;
  if (!gather_tab[type][opr]) return OR_E_NOT_SUPPORTED;
  if (trans == OR_NOTRANS)
    gather_tab[type][opr](op->NrowsN, op->rowStartsN, op->colIdsN, K, v, gv);
  else
    gather_tab[type][opr](op->NrowsT, op->rowStartsT, op->colIdsT, K, v, gv);
  return OR_OK;
}

/* Scatter list selection (214-222): Trans -> N rows/lists, else T */
int oracle_op_scatter(const oracle_op_t *op, int type, int trans, int K, const void *gv, void *v) {
  if (bad_args(type, 0, trans, K)) return OR_E_ARG;
  if (trans == OR_TRANS)
    scatter_tab[type](op->NrowsN, op->rowStartsN, op->colIdsN, K, gv, v);
  else
    scatter_tab[type](op->NrowsT, op->rowStartsT, op->colIdsT, K, gv, v);
  return OR_OK;
}

/* GatherScatter list selection (321-339) */
int oracle_op_gather_scatter(const oracle_op_t *op, int type, int opr, int trans, int K, void *v) {
  if (bad_args(type, opr, trans, K)) return OR_E_ARG;
  if (!gs_tab[type][opr]) return OR_E_NOT_SUPPORTED;
  if (trans == OR_TRANS)
    gs_tab[type][opr](op->NrowsN, op->rowStartsT, op->colIdsT, op->rowStartsN, op->colIdsN, K, v);
  else if (trans == OR_SYM)
    gs_tab[type][opr](op->NrowsT, op->rowStartsT, op->colIdsT, op->rowStartsT, op->colIdsT, K, v);
  else
    gs_tab[type][opr](op->NrowsT, op->rowStartsN, op->colIdsN, op->rowStartsT, op->colIdsT, K, v);
  return OR_OK;
}

int oracle_extract(int type, dlong N, int K, const dlong *ids, const void *q, void *gq) {
  if (bad_args(type, 0, 0, K)) return OR_E_ARG;
  extract_tab[type](N, K, ids, q, gq);
  return OR_OK;
}

/* ogs_t host wrappers at one rank.  The halo operator has no rows, so the
 * halo gather/exchange/scatter steps are no-ops and only gatherLocal runs
 * (ogs.cpp:150-184 GatherScatter; 328-374 Gather; 517-561 Scatter). */
int oracle_gather_scatter(const oracle_ogs_t *ogs, int type, int opr, int trans, int K, void *v) {
  return oracle_op_gather_scatter(&ogs->local, type, opr, trans, K, v);
}

int oracle_gather(const oracle_ogs_t *ogs, int type, int opr, int trans, int K,
                  const void *v, void *gv) {
  if (!ogs->gather_defined) return OR_E_GATHER_UNDEF; /* AssertGatherDefined, ogs.cpp:334,359 */
  return oracle_op_gather(&ogs->local, type, opr, trans, K, v, gv);
}

int oracle_scatter(const oracle_ogs_t *ogs, int type, int trans, int K, const void *gv, void *v) {
  if (!ogs->gather_defined) return OR_E_GATHER_UNDEF; /* ogs.cpp:522,543 */
  return oracle_op_scatter(&ogs->local, type, trans, K, gv, v);
}
