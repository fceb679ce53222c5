"""dawgs_amd -- MI355X-native rank-local gather-scatter (libp::ogs at one rank).

The compute path is libogs_hip.so (gfx950 HIP kernels behind the C-ABI in
include/ogs_hip.h); this package is its Python host mirror.  Build the
library with ``python -m dawgs_amd.build``.
"""
from .ogs import (Add, AllToAll, Auto, BF16, Batch, CrystalRouter, Double, Float, FloatC, Halo, Int32, Int64, Max, Min,
                  Mul, NoTrans, Pairwise, Plan, PlanAllK, PlanK1, Signed, Sym, Trans, Unsigned, extract, halo_t,
                  ogs_t)
from .mesh import make_ids

__all__ = ["ogs_t", "halo_t", "Plan", "Batch", "extract", "make_ids",
           "Float", "Double", "Int32", "Int64", "BF16", "FloatC", "Add", "Mul", "Max", "Min",
           "Sym", "NoTrans", "Trans", "Auto", "Pairwise", "CrystalRouter", "AllToAll",
           "Unsigned", "Signed", "Halo", "PlanAllK", "PlanK1"]
