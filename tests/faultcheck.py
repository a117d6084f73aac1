"""Comparison of fault data (GPU against the oracle) with the rupture time checked per face.

The rupture-time component holds 1e10 until a face ruptures and then the time of its rupture
(RSSETBND, RSIBCF.ChF:116-122; SWSETBND).  A per-component relative error against the
largest magnitude would measure it against the 1e10 sentinel, so it is compared per face:
faces the oracle has not ruptured must hold the sentinel exactly, and ruptured faces their
oracle time to the stated relative tolerance of that time.
"""
import numpy as np

SENTINEL = 1e10


def rupture_comp(B):
    """rupture-time component: 9 of the 11 rate-and-state components, 8 of the 9 slip-weakening ones"""
    return {11: 9, 9: 8}[np.asarray(B).shape[0]]


def fault_errors(Bg, Br):
    """Per component: max|gpu - oracle| / max|oracle|, except the rupture time: inf when the
    ruptured sets differ, else the largest |gpu - oracle| / |oracle| over ruptured faces."""
    Bg = np.asarray(Bg).reshape(Bg.shape[0], -1)
    Br = np.asarray(Br).reshape(Br.shape[0], -1)
    scale = np.maximum(np.max(np.abs(Br), axis=1), 1e-300)
    e = np.max(np.abs(Bg - Br), axis=1) / scale
    rt = rupture_comp(Br)
    g, r = Bg[rt], Br[rt]
    sent = r == SENTINEL
    if not np.array_equal(g == SENTINEL, sent):
        e[rt] = np.inf
    else:
        e[rt] = np.max(np.abs(g[~sent] - r[~sent]) / np.abs(r[~sent])) if np.any(~sent) else 0.0
    return e


def new_ruptures(B0, B1):
    """faces that were unruptured in B0 and have a rupture time in B1"""
    rt = rupture_comp(B0)
    return int(np.count_nonzero((np.asarray(B0)[rt] == SENTINEL) & (np.asarray(B1)[rt] != SENTINEL)))
