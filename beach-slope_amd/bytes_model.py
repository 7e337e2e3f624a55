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
