"""MI355X (gfx950) implementation of QuEST's retarded-interaction step.

The product is libquest_amd.so (C ABI: include/quest_amd.h, HIP kernels in
csrc/); this package is its Python front end (capi), the seeded system
generators (systems) and the readers of quest_hip's output files (outputs).
"""
from .capi import (ACC_BLOCK, ACC_COMPENSATED, CLAIM_AUTO, CLAIM_DYNAMIC, CLAIM_STATIC, LOOP_AUTO,
                   LOOP_PIPELINED, LOOP_PLAIN, PAIR_ON_THE_FLY, PAIR_TABLE, PAIR_TIME_BLOCKED, PREC_FP64, PREC_MIXED, DYADIC_LAB,
                   DYADIC_ROTATING, Context, QuestError, UnboundedSolution, load, pc_weights)

__all__ = ["ACC_BLOCK", "ACC_COMPENSATED", "CLAIM_AUTO", "CLAIM_DYNAMIC", "CLAIM_STATIC", "LOOP_AUTO",
           "LOOP_PIPELINED", "LOOP_PLAIN", "PAIR_ON_THE_FLY", "PAIR_TABLE", "PAIR_TIME_BLOCKED", "PREC_FP64",
           "PREC_MIXED", "DYADIC_LAB", "DYADIC_ROTATING", "Context", "QuestError", "UnboundedSolution",
           "load", "pc_weights"]
