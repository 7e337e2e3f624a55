"""Algorithmic (compulsory) HBM bytes of the solver kernels.

Each kernel is charged for reading every input array once and writing every
output once; values gathered at neighbour rows are assumed to come from cache
(they are rows of the same vectors).  n = rows, nnz = stored PPE entries.
Per row: the row length (int32), the row's own vector entries read, the
outputs written; per entry: the column (the int32 index in the plain form,
the 16-bit offset j - i in the precision form) and the FP64 value.

  k_bicg_1  t1 = A p, partial (t1 . rh)       reads alen, p, rh; writes t1
  k_bicg_2  t2 = A s, s = r - alpha t1 on the fly, partials (t2 . s), (t2 . t2)
                                             reads alen, r, t1; writes s, t2
  kb_pb_1   t1 = A ph, partial (t1 . rh32)    reads alen, ph (FP32), rh32; writes t1 (FP64)
  kb_pb_2   t2 = A sh, partials (t2 . s), (t2 . t2)
                                             reads alen, sh (FP32), s (FP64); writes t2 (FP64)
"""
from __future__ import annotations

_PER_ROW = {
    "k_bicg_1": (4 + 8 + 8, 8),
    "k_bicg_2": (4 + 8 + 8, 16),
    "kb_pb_1": (4 + 4 + 4, 8),
    "kb_pb_2": (4 + 4 + 8, 8),
}
_PER_NNZ = {"k_bicg_1": 12, "k_bicg_2": 12, "kb_pb_1": 10, "kb_pb_2": 10}


def kernel_bytes(kernel: str, n: int, nnz: float) -> float:
    """Algorithmic bytes of one launch of `kernel` over n rows, nnz entries."""
    rd, wr = _PER_ROW[kernel]
    return n * (rd + wr) + nnz * _PER_NNZ[kernel]


def kernel_read_bytes(kernel: str, n: int, nnz: float) -> float:
    rd, _ = _PER_ROW[kernel]
    return n * rd + nnz * _PER_NNZ[kernel]


def kernels() -> tuple[str, ...]:
    return tuple(_PER_ROW)


# ---------------------------------------------------------------------------
# Whole step.  Algorithmic bytes per particle of the particle kernels of one
# production-loop step (each input array read once, each output written
# once; neighbour fields from cache), with the two rebins, neighbour
# searches, effpor and eddyvis passes counted twice.  The PPE assembly adds
# 12 B per stored entry (its column and value).
_STEP_PER_PARTICLE = {
    "projection": 56,
    "rebin (classify, rank, scatter) x2": 2 * (35 + 8 + 240),
    "neighbour x2": 2 * 81,
    "effpor x2": 2 * 68,
    "eddyvis x2": 2 * 104,
    "int_vel": 253,
    "ppe_assemble": 100,
    "comp_vel": 132,
    "comp_pos": 77,
    "freesurf": 10,
    "opt2_shift (+ apply)": 140 + 12,
    "densityupdate": 48,
    "timestep": 16,
    "scalart + scalarupdate (one sub-step)": 109 + 32,
}

# Per iteration and row.  Plain: the two SpMVs with their dots, the x/r
# update and the direction update.  Precision form: the two FP64-product
# SpMVs, the four level-0 V-cycle passes (FP16 values and 16-bit offsets,
# 4 B per entry, with their vectors), and the s, x/r and direction updates;
# the cell levels are a few percent of the rows and are left out.
_ITER_PER_ROW = {"plain": 64 + 32, "precision": 4 * 16 + 24 + 60 + 24}
_ITER_PER_NNZ = {"plain": 0, "precision": 4 * 4}


def step_bytes(form: str, n: int, nnz: float, iters: float) -> float:
    """Algorithmic HBM bytes of one step of `form` ("plain" or "precision")
    over n particles with nnz stored PPE entries and `iters` iterations."""
    particles = n * sum(_STEP_PER_PARTICLE.values()) + 12 * nnz
    if form == "plain":
        it = kernel_bytes("k_bicg_1", n, nnz) + kernel_bytes("k_bicg_2", n, nnz)
    else:
        it = kernel_bytes("kb_pb_1", n, nnz) + kernel_bytes("kb_pb_2", n, nnz)
    it += n * _ITER_PER_ROW[form] + nnz * _ITER_PER_NNZ[form]
    return particles + iters * it
