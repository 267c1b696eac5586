"""Algorithmic bytes and flops of one SpMV per format (what the roofline uses).

Flops are 2*nnz, unpadded, as the reference reports them
(intSpmv_main.c:453).  Bytes are what the format obliges a kernel to move
once: stored arrays (padding included -- the kernel streams it), the source
vector where it is gathered, and the result vector.  sv / sx / sy are the
bytes of one value / x element / y element; k is the number of right-hand
sides.
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def flops(nnz: int, k: int = 1) -> float:
    return 2.0 * nnz * k


def bytes_csr(n: int, nnz: int, sv=8, sx=8, sy=8) -> int:
    return nnz * (sv + 4) + (n + 1) * 4 + n * sx + n * sy


def bytes_axc(n: int, lenAX: int, sv=8, sx=8, sy=8) -> int:
    return (lenAX // 2) * (sv + sx) + (n + 1) * 4 + n * sy


def bytes_k1(n: int, lenVC: int, chunkNum: int, sv=8, sx=8, sy=8) -> int:
    return lenVC * (sv + 4) + 8 * chunkNum + 4 * n + n * sx + n * sy


def bytes_axt(n: int, lenAX: int, lenSEC: int, sv=8, sx=8, sy=8) -> int:
    """AXT with x embedded in the tiles (the reference layout)."""
    return (lenAX // 2) * (sv + sx) + 4 * lenSEC + n * sy


def bytes_axt_gather(n: int, lenAX: int, lenSEC: int, k=1, sv=8, sx=8, sy=8) -> int:
    """AXT with a column map and a separate x (k right-hand sides)."""
    return (lenAX // 2) * (sv + 4) + 4 * lenSEC + k * n * (sx + sy)
