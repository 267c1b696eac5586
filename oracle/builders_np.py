"""numpy restatement of the reference builders -- TEST INFRASTRUCTURE ONLY.

The reference builds each format by walking rows with running counters
(intSpmv_main.c:619-658, 782-837, 1118-1260).  Here every array is written
from closed-form slot positions instead, so that a slip in either the C host
builders or this restatement shows up as a mismatch:

  AXC   element e of row r -> ax[brp[r] + 16*(e//8) + e%8], x at +8
  K1    sorted slot s = 8c + l holds row permi[s]; element e -> chp[c] + 8e + l
  AXT   rows start at rsp_r = sum of padded lengths before r
        NOC th==1 : element e -> tile rsp_r//8 + e//8, lane e%8
        NOC th>1  : element e -> column rsp_r//th + e//th, brick-row e%th
        COM       : element e -> slot row[r] + e
"""
from __future__ import annotations

import numpy as np

HB = 8


def _elem_rows(row):
    """row index and in-row position of every CSR entry"""
    rl = np.diff(row.astype(np.int64))
    r = np.repeat(np.arange(len(rl)), rl)
    e = np.arange(row[-1], dtype=np.int64) - row[:-1].astype(np.int64)[r]
    return r, e, rl


def axc(row, col, val, x):
    r, e, rl = _elem_rows(row)
    brp = np.zeros(len(rl) + 1, np.int64)
    brp[1:] = np.cumsum(2 * HB * ((rl + HB - 1) // HB))
    ax = np.zeros(int(brp[-1]), np.float64)
    pos = brp[r] + 2 * HB * (e // HB) + e % HB
    ax[pos] = val
    ax[pos + HB] = x[col]
    return ax, brp.astype(np.uint32)


def k1(row, col, val):
    r, e, rl = _elem_rows(row)
    n = len(rl)
    chunk = (n + HB - 1) // HB
    order = np.argsort(-rl, kind="stable")
    permi = np.zeros((chunk + 1) * HB, np.uint32)
    permi[:n] = order
    nmc = rl[order[::HB]] if n else np.zeros(0, np.int64)
    chp = np.zeros(chunk, np.int64)
    if chunk > 1:
        chp[1:] = np.cumsum(HB * nmc[:-1])
    lenvc = int(HB * nmc.sum()) if chunk else 0
    kval = np.zeros(lenvc, np.float64)
    kcol = np.zeros(lenvc, np.uint32)
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    s = inv[r]
    pos = chp[s // HB] + HB * e + s % HB
    kval[pos] = val
    kcol[pos] = col
    return dict(permi=permi, nmc=nmc.astype(np.uint32), chp=chp.astype(np.uint32),
                val=kval, col=kcol, chunkNum=chunk, lenVC=lenvc)


def axt(row, col, val, x, th=1, mode="NOC", bs=128, thw=HB):
    r, e, rl = _elem_rows(row)
    ths = thw * th
    grp = thw if th == 1 else th
    pos_len = ((rl + grp - 1) // grp) * grp if mode == "NOC" else rl
    rsp = np.zeros(len(rl), np.int64)
    if len(rl) > 1:
        rsp[1:] = np.cumsum(pos_len[:-1])
    T = rsp // ths
    con = T * 2 * ths + (rsp % th) * 2 * thw + rsp // th - T * thw
    total_cols = int(((pos_len + th - 1) // th).sum())
    tileN = (total_cols + thw - 1) // thw
    lenAX = tileN * 2 * ths
    if mode == "NOC":
        lenSEC = tileN if th == 1 else tileN * thw
    else:
        lenSEC = tileN * ths
    ax = np.zeros(lenAX, np.float64)
    sec = np.zeros(lenSEC, np.uint32)
    if mode == "NOC" and th == 1:
        tile = rsp[r] // thw + e // thw
        p = tile * 2 * thw + e % thw
        sec[tile] = r
    elif mode == "NOC":
        c = rsp[r] // th + e // th
        p = (c // thw) * 2 * ths + (e % th) * 2 * thw + c % thw
        sec[c] = r
    else:
        k = row[:-1].astype(np.int64)[r] + e
        p = (k // thw) * 2 * thw + k % thw
        log = int(np.log2(bs))
        bco = k % bs
        head = (e == 0) | (bco == 0)
        off = np.minimum(rl[r] - e - 1, bs - 1 - bco)
        sec[k[head]] = ((r[head].astype(np.uint64) << np.uint64(log)) | off[head].astype(np.uint64)).astype(np.uint32)
    ax[p] = val
    ax[p + thw] = x[col]
    return dict(ax=ax, sec=sec, con=con.astype(np.uint32), tileN=tileN, lenAX=lenAX,
                lenSEC=lenSEC, colN=total_cols)
