"""CPU oracle for IWPP morphological reconstruction by dilation (ctypes wrapper).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the CPU
baseline leg of bench.py may import this package; the product path
(iwpp-mp_amd/) never does.

The C library (oracle/iwpp_oracle.c, built by oracle/Makefile) restates the
reference's update rule (IWPP MP/src/MorphologicalReconstruction.cu:42-47),
zero halos (same file :252-312) and clip (same file :827-836), and computes
the unique fixed point with Vincent's raster / anti-raster / FIFO hybrid.
Pinned by tests/test_oracle.py against a brute-force fixed-point iteration
(scipy grey_dilation), closed forms and invariants.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        sz, vp, i = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
        L.oracle_clip.argtypes = [vp, vp, sz, sz, i, i, i]
        L.oracle_reconstruct.argtypes = [vp, vp, sz, sz, i, i, i, i]
        L.oracle_pass_x.argtypes = [vp, vp, sz, sz, i, i, i]
        L.oracle_pass_y.argtypes = [vp, vp, vp, sz, sz, sz, i, i, i, i, i, i, i]
        L.oracle_pass_raster.argtypes = [vp, vp, vp, sz, sz, sz, i, i, i, i, i, i, i]
        L.oracle_unstable_blocks.argtypes = [vp, vp, sz, sz, i, i, i, i, i, i, vp]
        L.oracle_unstable_blocks.restype = ctypes.c_long
        L.qm_reconstruct.argtypes = [vp, vp, sz, sz, i, i, i, i, i, i, i, ctypes.c_uint64,
                                     ctypes.c_long, i, i, ctypes.POINTER(QMStats)]
        L.qm_dq_fill_swap_kat.argtypes = [i]
        L.qm_dq_fill_swap_kat.restype = ctypes.c_long
        L.qm_uq_unique_kat.argtypes = [i, i, ctypes.c_uint64]
        L.qm_uq_unique_kat.restype = ctypes.c_long
        _LIB = L
    return _LIB


def _dt(a):
    if a.dtype == np.uint8:
        return 1
    if a.dtype == np.uint16:
        return 2
    raise TypeError("uint8 or uint16 images only")


def _pitch(a):
    # C-contiguous rows: the pitch is the row size (strides[0] may be 0 for a
    # single-row view).
    return a.shape[1] * a.itemsize


def _check(rc, what):
    if rc < 0:
        raise ValueError(f"{what} failed with code {rc}")
    return rc


def reconstruct(marker, mask, conn=8):
    """Reconstruction by dilation of `marker` under `mask` (new array)."""
    J = np.ascontiguousarray(marker).copy()
    I = np.ascontiguousarray(mask)
    assert J.shape == I.shape and J.dtype == I.dtype and J.ndim == 2
    h, w = J.shape
    _check(lib().oracle_reconstruct(J.ctypes.data, I.ctypes.data, _pitch(J), _pitch(I),
                                    w, h, _dt(J), conn), "oracle_reconstruct")
    return J


def clip(marker, mask):
    J = np.ascontiguousarray(marker).copy()
    I = np.ascontiguousarray(mask)
    h, w = J.shape
    _check(lib().oracle_clip(J.ctypes.data, I.ctypes.data, _pitch(J), _pitch(I), w, h,
                             _dt(J)), "oracle_clip")
    return J


def pass_x(marker, mask):
    """One row pass, forward then backward (new array)."""
    J = np.ascontiguousarray(marker).copy()
    I = np.ascontiguousarray(mask)
    h, w = J.shape
    _check(lib().oracle_pass_x(J.ctypes.data, I.ctypes.data, _pitch(J), _pitch(I), w, h,
                               _dt(J)), "oracle_pass_x")
    return J


def pass_y(src, mask, conn, direction, block_w, block_h):
    """One blocked, out-of-place column pass (new array)."""
    S = np.ascontiguousarray(src)
    I = np.ascontiguousarray(mask)
    D = np.zeros_like(S)
    h, w = S.shape
    _check(lib().oracle_pass_y(S.ctypes.data, D.ctypes.data, I.ctypes.data, _pitch(S),
                               _pitch(D), _pitch(I), w, h, _dt(S), conn, direction,
                               block_w, block_h), "oracle_pass_y")
    return D


def pass_raster(src, mask, conn, direction, block_w, block_h):
    """One blocked, out-of-place raster pass (new array): each row's column
    step, then its forward and backward scans across the block."""
    S = np.ascontiguousarray(src)
    I = np.ascontiguousarray(mask)
    D = np.zeros_like(S)
    h, w = S.shape
    _check(lib().oracle_pass_raster(S.ctypes.data, D.ctypes.data, I.ctypes.data, _pitch(S),
                                    _pitch(D), _pitch(I), w, h, _dt(S), conn, direction,
                                    block_w, block_h), "oracle_pass_raster")
    return D


class QMStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in
                ("visits", "seeded", "rounds", "restarts", "overlap", "repeat_in_round",
                 "state_left", "steps")]


QUEUE_VARIANTS = {"mutex": 0, "dq": 1, "uq": 2}


QUEUE_FAULTS = {None: 0, "mutex_drop_renotify": 1, "uq_keep_flag": 2}


def queue_model(marker, mask, conn, variant, mp=8, workers=4, seed=1, cap=0, max_sweeps=0,
                fault=None):
    """The CPU queue-phase model (oracle/queue_model.c): reconstruction through
    a queue variant with simulated concurrent workers.  fault plants a
    protocol mistake (tests of the checks).  Returns (new array, stats dict)."""
    J = np.ascontiguousarray(marker).copy()
    I = np.ascontiguousarray(mask)
    h, w = J.shape
    st = QMStats()
    _check(lib().qm_reconstruct(J.ctypes.data, I.ctypes.data, _pitch(J), _pitch(I), w, h, _dt(J),
                                conn, QUEUE_VARIANTS[variant], mp, workers, seed, cap,
                                max_sweeps, QUEUE_FAULTS[fault], ctypes.byref(st)),
           "qm_reconstruct")
    return J, {k: getattr(st, k) for k, _ in QMStats._fields_}


def queue_fill_swap_kat(g):
    """Coordinates not appearing exactly twice after the reference's double
    queue fill/swap sequence, on the model's queue operations (0 = pass)."""
    return _check(lib().qm_dq_fill_swap_kat(g), "qm_dq_fill_swap_kat")


def queue_unique_kat(g, dup, seed=1):
    """Flags left set plus coordinates not exactly once after extracting the
    unique entries of a duplicated write queue (0 = pass)."""
    return _check(lib().qm_uq_unique_kat(g, dup, seed), "qm_uq_unique_kat")


def unstable_blocks(marker, mask, conn, mp=64):
    """uint8 flags per mp x mp block: 1 where a pixel can still be raised."""
    J = np.ascontiguousarray(marker)
    I = np.ascontiguousarray(mask)
    h, w = J.shape
    flags = np.zeros(((h + mp - 1) // mp, (w + mp - 1) // mp), np.uint8)
    _check(lib().oracle_unstable_blocks(J.ctypes.data, I.ctypes.data, _pitch(J), _pitch(I),
                                        w, h, _dt(J), conn, mp, mp, flags.ctypes.data),
           "oracle_unstable_blocks")
    return flags
