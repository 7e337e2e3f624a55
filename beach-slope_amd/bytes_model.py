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
  kb_pb_2   t2 = A sh, partials (t2 . s), (t2 . t2), (t2 . rh32), (s . rh32), (s . s)
                                             reads alen, sh (FP32), s (FP64), rh32; writes t2 (FP64)
  kb_mg_pre0   r0 = v - A16 x0 (level 0 of the V-cycle, FP16 values, 16-bit
               offsets)                      reads alen, v (FP32 or FP64: 6 on average), x0 (own);
                                             writes r0
  kb_mg_post0  x = x1 + w (v - A16 x1) d     reads alen, v, x1, d0; writes x
  kb_mg_coarse the coarse cell levels in one workgroup: reads their stencils
               (9 FP32) and 1/diagonal, the first level's right-hand side
               and the dense FP64 inverse; writes the first level's result
               (coarse_bytes)
  kb_mg_restrict0  level-1 right-hand side: reads r0 and the cell starts,
               writes b1
  kb_mg_down   one multi-workgroup cell level (cell_pass_bytes, averaged over
               the levels): reads the stencil (9 FP32), b and 1/diagonal,
               writes x and the parent level's b
  kb_mg_up     one multi-workgroup cell level: reads the stencil, b, x,
               1/diagonal and the parent's correction, writes the result; on
               level 1 also x0 of every row, read and written (prolongation)
  kb_pfinalize one workgroup folds the row blocks' FP64 partials (one array
               for alpha, five for omega, the next rho and ||r||)
"""
from __future__ import annotations

_PER_ROW = {
    "k_bicg_1": (4 + 8 + 8, 8),
    "k_bicg_2": (4 + 8 + 8, 16),
    "kb_pb_1": (4 + 4 + 4, 8),
    "kb_pb_2": (4 + 4 + 8 + 4, 8),
    "kb_mg_pre0": (4 + 6 + 4, 4),
    "kb_mg_post0": (4 + 6 + 4 + 4, 4),
}
_PER_NNZ = {"k_bicg_1": 12, "k_bicg_2": 12, "kb_pb_1": 10, "kb_pb_2": 10,
            "kb_mg_pre0": 4, "kb_mg_post0": 4}

MG_DENSE_MAX = 64
MG_CO_UNKNOWNS = 1024   # unknowns of the coarse workgroup's first level, at most


def cell_levels(cellx: int, celly: int) -> list[int]:
    """Unknowns of the multigrid cell levels 1.. (2x2 aggregation down to at
    most MG_DENSE_MAX unknowns)."""
    ny, nx, out = celly, cellx, []
    while True:
        out.append(ny * nx)
        if ny * nx <= MG_DENSE_MAX:
            return out
        ny, nx = (ny + 1) // 2, (nx + 1) // 2


def coarse_bytes(cellx: int, celly: int) -> float:
    """Algorithmic bytes of one kb_mg_coarse launch: from the first level
    with at most MG_CO_UNKNOWNS unknowns to the dense one, the stencils and
    1/diagonal of every level but the dense one, the first level's
    right-hand side and result, and the dense FP64 inverse."""
    lv = cell_levels(cellx, celly)
    k = next(i for i, n in enumerate(lv) if n <= MG_CO_UNKNOWNS)
    tail = lv[k:]
    return 4.0 * 10 * sum(tail[:-1]) + 4.0 * 2 * tail[0] + 8.0 * tail[-1] ** 2


def cell_pass_bytes(kernel: str, n: int, cellx: int, celly: int) -> float:
    """Algorithmic bytes of one launch of a cell-level kernel of the V-cycle
    at n rows: kb_mg_restrict0, or kb_mg_down / kb_mg_up averaged over the
    multi-workgroup levels (those before the coarse workgroup's first level),
    or kb_pfinalize averaged over its two launches per iteration."""
    lv = cell_levels(cellx, celly)
    k = next(i for i, m in enumerate(lv) if m <= MG_CO_UNKNOWNS)
    if kernel == "kb_mg_restrict0":
        return 4.0 * n + 4.0 * (lv[0] + 1) + 4.0 * lv[0]
    if kernel == "kb_pfinalize":
        return 8.0 * ((n + 255) // 256) * (1 + 5) / 2.0
    if k == 0:
        return 0.0
    if kernel == "kb_mg_down":
        return sum(48.0 * lv[i] + 4.0 * lv[i + 1] for i in range(k)) / k
    if kernel == "kb_mg_up":
        tot = sum(52.0 * lv[i] + 4.0 * lv[i + 1] for i in range(k))
        return (tot + 8.0 * n + 4.0 * (lv[0] + 1)) / k
    raise KeyError(kernel)


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
# 4 B per entry, with their vectors), the s update (reads r, t1, d0; writes
# s, x0: 32 B) and the one x/r/direction pass (reads x, ph, sh, s, t2, p,
# t1, d0; writes x, r, p, x0: 72 B); the cell levels are a few percent of
# the rows and are left out.
_ITER_PER_ROW = {"plain": 64 + 32, "precision": 4 * 16 + 32 + 72}
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
