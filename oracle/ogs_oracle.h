/*
 * ogs_oracle.h -- CPU restatement of the reference's single-rank host
 * gather-scatter path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check
 * and the CPU-baseline leg of the benchmark load this as the checker.  The
 * product library never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative
 * to the paranumal/dawgs tree).
 */
#ifndef OGS_ORACLE_H
#define OGS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t dlong;  /* include/types.h:60 */
typedef int64_t hlong;  /* include/types.h:52 */

/* enum orders follow include/ogs.hpp:179-200 */
enum { OR_FLOAT = 0, OR_DOUBLE = 1, OR_INT32 = 2, OR_INT64 = 3,
       /* precision form (the library's own types, no reference counterpart):
        * bfloat16 storage / float storage, Add with two-sum compensation */
       OR_BF16 = 4, OR_FLOAT_C = 5 };
enum { OR_ADD = 0, OR_MUL = 1, OR_MAX = 2, OR_MIN = 3 };
enum { OR_SYM = 0, OR_NOTRANS = 1, OR_TRANS = 2 };
enum { OR_UNSIGNED = 0, OR_SIGNED = 1, OR_HALO = 2 };

/* error codes (the reference throws libp::exception via LIBP_ABORT) */
enum {
  OR_OK = 0,
  OR_E_INVALID_SETUP = -1,   /* ogsSetup.cpp:98-100 */
  OR_E_HALO_POSITIVE = -2,   /* ogsSetup.cpp:286-288 */
  OR_E_ROW_TOO_LONG = -3,    /* ogsOperator.cpp:460-465 */
  OR_E_GATHER_UNDEF = -4,    /* ogsSetup.cpp:861-864 */
  OR_E_ARG = -5,
  OR_E_NOT_SUPPORTED = -6    /* Mul on the precision-form types */
};

/* One CSR gather operator: ogsOperator_t's host members (ogsOperator.hpp:43-64). */
typedef struct {
  dlong Ncols, NrowsN, NrowsT, nnzN, nnzT;
  dlong *rowStartsN, *rowStartsT;   /* NrowsT+1 each */
  dlong *colIdsN, *colIdsT;
  dlong NrowBlocksN, NrowBlocksT;
  dlong *blockRowStartsN, *blockRowStartsT;
  int aliased;                      /* Unsigned: N arrays are the T arrays (ogsSetup.cpp:719,728) */
} oracle_op_t;

/* ogsBase_t's counts (ogsBase.hpp:50-62) plus the two operators. */
typedef struct {
  dlong N, Ngather, NlocalT, NhaloT, NlocalP, NhaloP;
  hlong NgatherGlobal;
  int kind, unique, gather_defined;
  oracle_op_t local, halo;
} oracle_ogs_t;

/* dawgsMain.cpp:49-143 at one rank; ids has nx*ny*nz*(p+1)^3 entries */
void oracle_make_ids(int nx, int ny, int nz, int p, hlong *ids);

/* ogsSetup.cpp:69-190 at comm size 1.  May rewrite ids when unique. */
int oracle_setup(dlong N, hlong *ids, int kind, int unique, oracle_ogs_t *ogs);
void oracle_free(oracle_ogs_t *ogs);

/* ogsOperator.cpp:447-512 (block budget gatherNodesPerBlock = 512) */
int oracle_setup_row_blocks(oracle_op_t *op, dlong gatherNodesPerBlock);

/* ogsOperator_t host templates: Gather 60-117, Scatter 206-238,
 * GatherScatter 311-378, extract 540-559.  Values are interleaved
 * v[k + id*K]. */
int oracle_op_gather(const oracle_op_t *op, int type, int opr, int trans, int K,
                     const void *v, void *gv);
int oracle_op_scatter(const oracle_op_t *op, int type, int trans, int K,
                      const void *gv, void *v);
int oracle_op_gather_scatter(const oracle_op_t *op, int type, int opr, int trans,
                             int K, void *v);
int oracle_extract(int type, dlong N, int K, const dlong *ids, const void *q, void *gq);

/* ogs_t host wrappers at one rank (ogs.cpp:141-184, 318-374, 508-561):
 * the halo operator is empty, so only the local operator does work. */
int oracle_gather_scatter(const oracle_ogs_t *ogs, int type, int opr, int trans, int K, void *v);
int oracle_gather(const oracle_ogs_t *ogs, int type, int opr, int trans, int K,
                  const void *v, void *gv);
int oracle_scatter(const oracle_ogs_t *ogs, int type, int trans, int K,
                   const void *gv, void *v);

/* dawgsMain.cpp:326-330: std::mt19937(seed), uniform_real_distribution(-0.25,0.25) */
void oracle_reference_q(uint32_t seed, dlong n, double *q);

/* bfloat16 <-> float as the precision form defines them (round to nearest
 * even, NaN quieted) */
float oracle_bf16_to_float(uint16_t b);
uint16_t oracle_float_to_bf16(float f);
/* OpenMP thread count actually used by the host loops */
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
