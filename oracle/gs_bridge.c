/*
 * gs_bridge.c -- thin C entry points over the reference's vendored gslib
 * (3rdParty/gslib, Release 1.0.3), built from its own sources into
 * oracle/_ref/.  TEST INFRASTRUCTURE: this is the reference's legacy "-cc"
 * comparator (dawgsGS.cpp:34-71, dawgsMain.cpp:338-352), used only to pin
 * the oracle.  Built without MPI, so the communicator is gslib's single
 * process stand-in (comm.h, comm_init with MPI undefined).
 */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
/* gslib.h would force MPI on; include its pieces in its own order instead */
#include "c99.h"
#include "name.h"
#include "fail.h"
#include "types.h"
#include "comm.h"
#include "mem.h"
#include "gs_defs.h"
#include "gs.h"

/* gsSetup (dawgsGS.cpp:44-71): gs_setup(id, n, comm, unique, gs_auto, verbose) */
void *gsref_setup(int n, const long long *ids, int unique) {
  struct comm com;
  comm_init(&com, 0);
  slong *id = (slong *)malloc(sizeof(slong) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) id[i] = (slong)ids[i];
  struct gs_data *gsh = gs_setup(id, (uint)n, &com, unique, gs_auto, 0);
  free(id);
  comm_free(&com);
  return gsh;
}

/* gsGatherScatterVec (dawgsGS.cpp:39-41) with the domain and op exposed:
 * dom 0..3 = double, float, int, long long; op 0..3 = add, mul, min, max
 * (gs_defs.h:18-31 enum orders) */
void gsref_gs_vec(void *u, int k, int dom, int op, int transpose, void *gsh) {
  static const gs_dom doms[4] = {gs_double, gs_float, gs_int, gs_long_long};
  static const gs_op ops[4] = {gs_add, gs_mul, gs_min, gs_max};
  gs_vec(u, (unsigned)k, doms[dom], ops[op], (unsigned)transpose, (struct gs_data *)gsh, 0);
}

void gsref_free(void *gsh) { gs_free((struct gs_data *)gsh); }
