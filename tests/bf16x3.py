"""numpy model of the 3-term split-BF16 Gram: each FP64 entry a splits into
hi + mid + lo, three values representable in BF16 (round to nearest even from
FP32), the residuals formed exactly in FP64; the Gram keeps the six largest
term products (hi.hi, hi.mid, mid.hi, hi.lo, lo.hi, mid.mid), here summed in
FP64."""
import numpy as np


def bf16_rn(x):
    """round FP64 -> FP32 -> BF16 (nearest even), returned as FP64"""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split3(a):
    hi = bf16_rn(a)
    r = a - hi
    mid = bf16_rn(r)
    lo = bf16_rn(r - mid)
    return hi, mid, lo


def gram3(A, drop=()):
    """upper-triangle-relevant Gram of the split model; `drop` names products
    to leave out (for planted-mistake checks)"""
    h, m, l = split3(np.asarray(A, dtype=np.float64))
    terms = {"hh": (h, h), "hm": (h, m), "mh": (m, h), "hl": (h, l), "lh": (l, h), "mm": (m, m)}
    G = np.zeros((A.shape[1], A.shape[1]))
    for k, (x, y) in terms.items():
        if k not in drop:
            G += x.T @ y
    return G
