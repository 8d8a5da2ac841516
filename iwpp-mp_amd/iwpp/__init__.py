"""Python binding of the MI355X IWPP reconstruction library (ctypes over the C ABI).

The product is the C ABI in include/iwpp.h, implemented by libiwpp.so (HIP
kernels for gfx950 plus the C++ host driver).  This module only binds it:
device memory and streams come from PyTorch, every computation runs in the
library.  There is no fallback: if libiwpp.so is missing, importing the
binding raises.

Reference interface mirrored here (IWPP MP/include/MorphologicalReconstruction.cuh:45-46):
    MorphologicalReconstruction(marker, mask, pitchMarker, pitchMask, size)
with device pointers, row pitches in bytes, a square uint8 image, in place.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
LIB_PATH = os.path.join(_PKG, "libiwpp.so")
SYNTH_PATH = os.path.join(_PKG, "libiwpp_synth.so")

U8, U16 = 1, 2
QUEUES = {"mutex": 0, "dq": 1, "uq": 2, "auto": 3}
SCHEDULES = {"plain": 0, "learned": 1}
MP_SIZE = 64

IWPP_OK = 0
ERRORS = {-1: "IWPP_ERR_ARG", -2: "IWPP_ERR_HIP", -3: "IWPP_ERR_TIMEOUT",
          -4: "IWPP_ERR_UNSUPPORTED", -5: "IWPP_ERR_NOMEM"}


class IWPPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("sweep_rounds", ctypes.c_int), ("col_block_rows", ctypes.c_int),
                ("max_local_iters", ctypes.c_int), ("queue_capacity", ctypes.c_int),
                ("workers_per_cu", ctypes.c_int), ("timeout_s", ctypes.c_double),
                ("trace_phases", ctypes.c_int), ("check_queue", ctypes.c_int),
                ("learned_min_pixels", ctypes.c_longlong), ("split_sweeps", ctypes.c_int),
                ("seam_rows", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("tile_visits", ctypes.c_longlong), ("seeded", ctypes.c_longlong),
                ("sweep_rounds", ctypes.c_int), ("queue_rounds", ctypes.c_int),
                ("restarts", ctypes.c_int), ("queue", ctypes.c_int), ("tile", ctypes.c_int),
                ("sweeps_skipped_tiles", ctypes.c_int), ("schedule", ctypes.c_int),
                ("ms_sweep", ctypes.c_float),
                ("ms_seed", ctypes.c_float), ("ms_queue", ctypes.c_float),
                ("ms_plan", ctypes.c_float),
                ("us_wait", ctypes.c_double), ("us_fetch", ctypes.c_double),
                ("us_converge", ctypes.c_double), ("us_store", ctypes.c_double),
                ("us_notify", ctypes.c_double), ("ms_rows", ctypes.c_float),
                ("ms_cols", ctypes.c_float), ("repeat_visits", ctypes.c_longlong),
                ("state_left", ctypes.c_longlong), ("col_block_rows", ctypes.c_int),
                ("seed_class", ctypes.c_int), ("ms_seams", ctypes.c_float),
                ("seam_rows", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every entry point declared in include/iwpp.h, with its ctypes signature.
_SZ, _VP, _I = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
SIGNATURES = {
    "iwpp_default_options": (None, [ctypes.POINTER(Options)]),
    "iwpp_workspace_create": (_VP, [_SZ, _I]),
    "iwpp_workspace_destroy": (None, [_VP]),
    "iwpp_reconstruct": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _I, _I, _I, _I, _VP, _VP]),
    "iwpp_reconstruct_ex": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _I, _I, _I, _I,
                                 ctypes.POINTER(Options), ctypes.POINTER(Stats), _VP, _VP]),
    "iwpp_clip": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _VP]),
    "iwpp_sweep_rows": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _VP, _VP]),
    "iwpp_sweep_cols": (_I, [_VP, _VP, _VP, _SZ, _SZ, _SZ, _I, _I, _I, _I, _I, _I, _VP, _VP]),
    "iwpp_sweep_raster": (_I, [_VP, _VP, _VP, _SZ, _SZ, _SZ, _I, _I, _I, _I, _I, _I, _VP, _VP]),
    "iwpp_sweep_cols_block_width": (_I, [_VP, _VP, _VP, _SZ, _SZ, _SZ, _I, _I]),
    "iwpp_unstable_megapixels": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _I, _VP, _VP, _VP]),
    "iwpp_reconstruct_batch": (_I, [_VP, _I, _I, _I, _I, _I, _I, ctypes.POINTER(Options),
                                    ctypes.POINTER(Stats), _VP, _VP]),
    "iwpp_batch_graph_create": (_VP, [_VP, _I, _I, _I, _I, _I, _I, ctypes.POINTER(Options),
                                      _VP, _VP]),
    "iwpp_batch_graph_launch": (_I, [_VP, _VP]),
    "iwpp_batch_graph_stats": (_I, [_VP, ctypes.POINTER(Stats), _VP]),
    "iwpp_batch_graph_destroy": (None, [_VP]),
    "iwpp_autotune": (_I, [_VP, _I, _I, _I, _I, ctypes.POINTER(Options), _VP, _VP, _VP]),
    "iwpp_autotune_clear": (None, []),
    "iwpp_sched_set_model": (_I, [_VP, _I]),
    "iwpp_tile_features": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _I, _I, _I, _VP, _VP, _VP, _VP]),
    "iwpp_mlp_forward": (_I, [_VP, _I, _VP, _I, _VP, _VP, _VP]),
    "iwpp_mlp_backward": (_I, [_VP, _I, _VP, _VP, _VP, _VP, _I, ctypes.c_float, ctypes.c_float,
                               _VP, _VP, _VP]),
    "iwpp_mlp_adam": (_I, [_VP, _VP, _VP, _VP, _VP, ctypes.c_float, _I, _VP]),
    "iwpp_sweep_round_seed": (_I, [_VP, _I, _I, _I, _I, _VP, _VP, _VP, _VP]),
    "iwpp_reconstruct_noalloc": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _I,
                                      ctypes.POINTER(Options), _VP]),
    "iwpp_last_error": (ctypes.c_char_p, []),
    "MorphologicalReconstruction": (None, [_VP, _VP, _SZ, _SZ, _I]),
}

SYNTH_SIGNATURES = {
    "iwpp_synth_mask": (_I, [_VP, _SZ, _I, _I, _I, _I, ctypes.c_double, _I, _I, _I,
                             ctypes.c_uint64, _I]),
    "iwpp_synth_marker_hdome": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, _I]),
    "iwpp_synth_marker_seeds": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, _I, ctypes.c_long,
                                     ctypes.c_uint64]),
    "iwpp_synth_batch_shapes": (_I, [_I, ctypes.c_uint64, _VP, _VP, _VP, _VP]),
    "iwpp_synth_random_case": (_I, [_VP, _VP, _SZ, _SZ, _I, _I, ctypes.c_uint64]),
}

# include/iwpp_io.h (PGM files), exported by the same host library.
IO_SIGNATURES = {
    "iwpp_pgm_read": (_I, [ctypes.c_char_p, ctypes.POINTER(_VP), ctypes.POINTER(_I),
                           ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "iwpp_pgm_write": (_I, [ctypes.c_char_p, _VP, _SZ, _I, _I, _I]),
    "iwpp_pgm_free": (None, [_VP]),
}

_LIB = None
_SYNTH = None


def _bind(path, sigs):
    if not os.path.exists(path):
        raise ImportError(f"{os.path.basename(path)} is not built; run "
                          "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def lib():
    """The product library (loads it on first use; raises if it is not built)."""
    global _LIB
    if _LIB is None:
        # PyTorch bundles its own HIP runtime with the same soname
        # (libamdhip64.so.7) as the system one libiwpp.so links against; the
        # first one loaded serves the whole process.  Load torch's first so
        # that torch keeps the runtime it was built with (the other order
        # leaves torch without devices) and both share streams and memory.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _LIB = _bind(LIB_PATH, SIGNATURES)
    return _LIB


def synth_lib():
    global _SYNTH
    if _SYNTH is None:
        _SYNTH = _bind(SYNTH_PATH, {**SYNTH_SIGNATURES, **IO_SIGNATURES})
    return _SYNTH


def _raise(rc):
    if rc != IWPP_OK:
        msg = lib().iwpp_last_error()
        raise IWPPError(rc, msg.decode() if msg else "")
    return rc


def default_options(**kw):
    o = Options()
    lib().iwpp_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


# --------------------------------------------------------------- images -----
def _torch():
    import torch
    return torch


class DeviceImage:
    """A pitched image in device memory: a torch uint8 buffer of H x pitch bytes."""

    def __init__(self, height, width, dtype=U8, pitch=None, device="cuda", buf=None):
        torch = _torch()
        self.h, self.w, self.dtype = int(height), int(width), int(dtype)
        row = self.w * self.dtype
        self.pitch = int(pitch) if pitch else (row + 255) // 256 * 256
        if buf is None:
            buf = torch.zeros((self.h, self.pitch), dtype=torch.uint8, device=device)
        self.buf = buf

    @property
    def ptr(self):
        return self.buf.data_ptr()

    @classmethod
    def from_numpy(cls, a, pitch=None, device="cuda"):
        torch = _torch()
        a = np.ascontiguousarray(a)
        dt = U8 if a.dtype == np.uint8 else U16 if a.dtype == np.uint16 else None
        if dt is None:
            raise TypeError("uint8 or uint16 images only")
        im = cls(a.shape[0], a.shape[1], dt, pitch, device)
        host = torch.zeros((im.h, im.pitch), dtype=torch.uint8)
        host[:, : im.w * dt] = torch.from_numpy(a.view(np.uint8).reshape(im.h, im.w * dt))
        im.buf.copy_(host.to(device))
        return im

    def numpy(self):
        raw = self.buf.cpu().numpy()[:, : self.w * self.dtype]
        return np.ascontiguousarray(raw).view(np.uint8 if self.dtype == U8 else np.uint16)

    def raw_numpy(self):
        return self.buf.cpu().numpy()


def _stream(stream):
    if stream is None:
        return _torch().cuda.current_stream().cuda_stream
    return int(stream)


# Library handles that own device memory, streams, events or graphs.  They
# are released explicitly (close()), when collected, or by the exit hook
# below, which runs while the HIP runtime is still up: graphs first, then
# workspaces.  Finalizers that run during interpreter shutdown do nothing,
# since by then the runtime (torch's, which this library shares) may already
# be unloading.
_LIVE = {}


def _track(obj):
    _LIVE[id(obj)] = weakref.ref(obj)


def _untrack(obj):
    _LIVE.pop(id(obj), None)


def _release_all():
    objs = [r() for r in list(_LIVE.values())]
    objs = [o for o in objs if o is not None]
    for o in sorted(objs, key=lambda o: 0 if isinstance(o, BatchGraph) else 1):
        try:
            o.close()
        except Exception:
            pass
    _LIVE.clear()


atexit.register(_release_all)


class Workspace:
    def __init__(self, max_pixels=0, max_dtype=U8):
        self.handle = lib().iwpp_workspace_create(max_pixels, max_dtype)
        if not self.handle:
            raise IWPPError(-5, "iwpp_workspace_create failed")
        _track(self)

    def close(self):
        if self.handle:
            lib().iwpp_workspace_destroy(self.handle)
            self.handle = None
        _untrack(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if sys.is_finalizing():
                return
            self.close()
        except Exception:
            pass


def reconstruct(marker, mask, connectivity=8, tile=0, queue="uq", schedule="plain",
                options=None, workspace=None, stream=None):
    """In-place reconstruction of DeviceImage `marker` under DeviceImage `mask`.
    Returns the run's counters as a dict."""
    assert marker.h == mask.h and marker.w == mask.w and marker.dtype == mask.dtype
    st = Stats()
    opt = options if options is not None else default_options()
    rc = lib().iwpp_reconstruct_ex(marker.ptr, mask.ptr, marker.pitch, mask.pitch, marker.w,
                                   marker.h, marker.dtype, connectivity, tile,
                                   QUEUES[queue] if isinstance(queue, str) else queue,
                                   SCHEDULES[schedule] if isinstance(schedule, str) else schedule,
                                   ctypes.byref(opt), ctypes.byref(st),
                                   workspace.handle if workspace else None, _stream(stream))
    _raise(rc)
    return st.as_dict()


NOALLOC_MAX_MP = 262144


def reconstruct_noalloc(marker, mask, connectivity=8, options=None, stream=None):
    """In-place reconstruction through the allocation-free entry point (mutex
    queue, plain schedule; scratch in the library's static device array).
    Synchronises the stream."""
    assert marker.h == mask.h and marker.w == mask.w and marker.dtype == mask.dtype
    opt = options if options is not None else default_options()
    _raise(lib().iwpp_reconstruct_noalloc(marker.ptr, mask.ptr, marker.pitch, mask.pitch,
                                          marker.w, marker.h, marker.dtype, connectivity,
                                          ctypes.byref(opt), _stream(stream)))


class ImageDesc(ctypes.Structure):
    _fields_ = [("marker", ctypes.c_void_p), ("mask", ctypes.c_void_p),
                ("pitch_marker", ctypes.c_size_t), ("pitch_mask", ctypes.c_size_t),
                ("width", ctypes.c_int), ("height", ctypes.c_int)]


def _build_array(pairs):
    arr = (ImageDesc * len(pairs))()
    dtype = pairs[0][0].dtype
    for k, (mk, ms) in enumerate(pairs):
        assert mk.h == ms.h and mk.w == ms.w and mk.dtype == ms.dtype == dtype
        arr[k] = ImageDesc(mk.ptr, ms.ptr, mk.pitch, ms.pitch, mk.w, mk.h)
    return arr, dtype


class Batch:
    """(marker, mask) DeviceImage pairs with their iwpp_image array built once.
    Every batch entry point takes one in place of the list of pairs; a caller
    that reconstructs the same buffers again and again saves rebuilding the
    array (about 1 us of host time per pair: 0.2-0.25 ms for 256 images).  The
    images must keep the buffers they had when the Batch was built."""

    def __init__(self, pairs):
        self.pairs = list(pairs)
        self.arr, self.dtype = _build_array(self.pairs)

    def __len__(self):
        return len(self.pairs)

    def __iter__(self):
        return iter(self.pairs)

    def __getitem__(self, k):
        return self.pairs[k]


def _batch_array(pairs):
    if isinstance(pairs, Batch):
        return pairs.arr, pairs.dtype
    return _build_array(pairs)


def reconstruct_batch(pairs, connectivity=8, tile=0, queue="auto", schedule="plain",
                      options=None, workspace=None, stream=None):
    """In-place reconstruction of a list of (marker, mask) DeviceImage pairs in
    one set of launches.  Returns the counters as a dict."""
    arr, dtype = _batch_array(pairs)
    st = Stats()
    opt = options if options is not None else default_options()
    rc = lib().iwpp_reconstruct_batch(arr, len(pairs), dtype, connectivity, tile,
                                      QUEUES[queue] if isinstance(queue, str) else queue,
                                      SCHEDULES[schedule] if isinstance(schedule, str) else schedule,
                                      ctypes.byref(opt), ctypes.byref(st),
                                      workspace.handle if workspace else None, _stream(stream))
    _raise(rc)
    return st.as_dict()


class Tuning(ctypes.Structure):
    _fields_ = [("queue", ctypes.c_int), ("sweep_rounds", ctypes.c_int), ("tile", ctypes.c_int),
                ("ms", ctypes.c_float), ("col_block_rows", ctypes.c_int),
                ("seed_class", ctypes.c_int)]


def autotune(pairs, connectivity=8, trials=2, options=None, workspace=None, stream=None):
    """Measures every (queue, sweep rounds) on these images, restores the
    markers, and caches the winner for queue='auto' calls.  Returns it."""
    arr, dtype = _batch_array(pairs)
    t = Tuning()
    opt = options if options is not None else default_options()
    _raise(lib().iwpp_autotune(arr, len(pairs), dtype, connectivity, trials, ctypes.byref(opt),
                               ctypes.byref(t), workspace.handle if workspace else None,
                               _stream(stream)))
    inv = {v: k for k, v in QUEUES.items()}
    return {"queue": inv[t.queue], "sweep_rounds": t.sweep_rounds, "tile": t.tile, "ms": t.ms,
            "col_block_rows": t.col_block_rows, "seed_class": t.seed_class}


class BatchGraph:
    """A batch reconstruction captured into a hipGraph (mutex queue variant)."""

    def __init__(self, pairs, workspace, connectivity=8, tile=0, schedule="plain", options=None,
                 stream=None):
        arr, dtype = _batch_array(pairs)
        self._arr = arr
        self.workspace = workspace
        # capture needs a stream other than the NULL stream (default priority)
        self.stream = stream if stream is not None else _torch().cuda.Stream()
        opt = options if options is not None else default_options()
        _torch().cuda.synchronize()
        self.handle = lib().iwpp_batch_graph_create(
            arr, len(pairs), dtype, connectivity, tile, QUEUES["mutex"],
            SCHEDULES[schedule] if isinstance(schedule, str) else schedule, ctypes.byref(opt),
            workspace.handle, self._sid())
        if not self.handle:
            msg = lib().iwpp_last_error()
            raise IWPPError(-2, msg.decode() if msg else "graph capture failed")
        _track(self)

    def _sid(self):
        st = self.stream
        return st.cuda_stream if hasattr(st, "cuda_stream") else int(st)

    def launch(self):
        """Replays the graph on the graph's stream (ordered after all work
        already queued on the current stream)."""
        torch = _torch()
        if hasattr(self.stream, "wait_stream"):
            self.stream.wait_stream(torch.cuda.current_stream())
        _raise(lib().iwpp_batch_graph_launch(self.handle, self._sid()))

    def stats(self):
        st = Stats()
        _raise(lib().iwpp_batch_graph_stats(self.handle, ctypes.byref(st), self._sid()))
        return st.as_dict()

    def close(self):
        if self.handle:
            lib().iwpp_batch_graph_destroy(self.handle)
            self.handle = None
        _untrack(self)

    def __del__(self):
        try:
            if sys.is_finalizing():
                return
            self.close()
        except Exception:
            pass


# ------------------------------------------------------ learned schedule ----
MLP_PARAMS = 5378
MLP_FEATURES = 16


def set_model(weights):
    """Install the learned schedule's model (FP32 numpy array of MLP_PARAMS
    values, rounded to BF16 by the library), or remove it with None."""
    if weights is None:
        _raise(lib().iwpp_sched_set_model(None, 0))
        return
    w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
    _raise(lib().iwpp_sched_set_model(w.ctypes.data, w.size))


def tile_features(marker, mask, tile, connectivity=8, sub=32, workspace=None, stream=None):
    """(raw, feat): torch float32 device tensors of shape (tiles, 16)."""
    torch = _torch()
    ntx = (marker.w + tile - 1) // tile
    nty = (marker.h + tile - 1) // tile
    dev = marker.buf.device
    raw = torch.empty((nty * ntx, MLP_FEATURES), dtype=torch.float32, device=dev)
    feat = torch.empty_like(raw)
    _raise(lib().iwpp_tile_features(marker.ptr, mask.ptr, marker.pitch, mask.pitch, marker.w,
                                    marker.h, marker.dtype, connectivity, tile, sub,
                                    raw.data_ptr(), feat.data_ptr(),
                                    workspace.handle if workspace else None, _stream(stream)))
    return raw, feat


def mlp_forward(params, x, bf16, acts=False, stream=None):
    """y (n, 2) [and the hidden activations (n, 128)] of the MLP on device
    tensors: params float32 (FP32 weights) or int16/uint16 BF16 bit patterns."""
    torch = _torch()
    n = x.shape[0]
    y = torch.empty((n, 2), dtype=torch.float32, device=x.device)
    a = torch.empty((n, 128), dtype=torch.float32, device=x.device) if acts else None
    _raise(lib().iwpp_mlp_forward(params.data_ptr(), 1 if bf16 else 0, x.data_ptr(), n,
                                  y.data_ptr(), a.data_ptr() if acts else None,
                                  _stream(stream)))
    return (y, a) if acts else y


def mlp_backward(params, x, acts, y, target, bf16, lw=(1.0, 1.0), stream=None):
    """Gradient (MLP_PARAMS float32 device tensor) of the weighted squared loss."""
    torch = _torch()
    n = x.shape[0]
    work = torch.empty((n * 130,), dtype=torch.float32, device=x.device)
    grad = torch.empty((MLP_PARAMS,), dtype=torch.float32, device=x.device)
    _raise(lib().iwpp_mlp_backward(params.data_ptr(), 1 if bf16 else 0, x.data_ptr(),
                                   acts.data_ptr(), y.data_ptr(), target.data_ptr(), n,
                                   float(lw[0]), float(lw[1]), work.data_ptr(), grad.data_ptr(),
                                   _stream(stream)))
    return grad


def mlp_adam(w, m, v, grad, wbf, lr, step, stream=None):
    _raise(lib().iwpp_mlp_adam(w.data_ptr(), m.data_ptr(), v.data_ptr(), grad.data_ptr(),
                               wbf.data_ptr(), float(lr), int(step), _stream(stream)))


def clip(marker, mask, stream=None):
    _raise(lib().iwpp_clip(marker.ptr, mask.ptr, marker.pitch, mask.pitch, marker.w, marker.h,
                           marker.dtype, _stream(stream)))


def sweep_rows(marker, mask, changed=None, stream=None):
    _raise(lib().iwpp_sweep_rows(marker.ptr, mask.ptr, marker.pitch, mask.pitch, marker.w,
                                 marker.h, marker.dtype,
                                 changed.data_ptr() if changed is not None else None,
                                 _stream(stream)))


def sweep_cols(src, dst, mask, connectivity, direction, block_rows, changed=None, stream=None):
    _raise(lib().iwpp_sweep_cols(src.ptr, dst.ptr, mask.ptr, src.pitch, dst.pitch, mask.pitch,
                                 src.w, src.h, src.dtype, connectivity, direction, block_rows,
                                 changed.data_ptr() if changed is not None else None,
                                 _stream(stream)))


def sweep_raster(src, dst, mask, connectivity, direction, block_rows, changed=None, stream=None):
    _raise(lib().iwpp_sweep_raster(src.ptr, dst.ptr, mask.ptr, src.pitch, dst.pitch, mask.pitch,
                                   src.w, src.h, src.dtype, connectivity, direction, block_rows,
                                   changed.data_ptr() if changed is not None else None,
                                   _stream(stream)))


def sweep_cols_block_width(src, dst, mask):
    return _raise_nonneg(lib().iwpp_sweep_cols_block_width(src.ptr, dst.ptr, mask.ptr, src.pitch,
                                                           dst.pitch, mask.pitch, src.w,
                                                           src.dtype))


def _raise_nonneg(v):
    if v < 0:
        _raise(v)
    return v


def unstable_megapixels(marker, mask, connectivity, stream=None):
    """uint8 numpy flags per 64 x 64 megapixel and their count."""
    torch = _torch()
    nmx = (marker.w + MP_SIZE - 1) // MP_SIZE
    nmy = (marker.h + MP_SIZE - 1) // MP_SIZE
    flags = torch.zeros((nmy, nmx), dtype=torch.uint8, device=marker.buf.device)
    count = torch.zeros((1,), dtype=torch.int32, device=marker.buf.device)
    _raise(lib().iwpp_unstable_megapixels(marker.ptr, mask.ptr, marker.pitch, mask.pitch,
                                          marker.w, marker.h, marker.dtype, connectivity,
                                          flags.data_ptr(), count.data_ptr(), _stream(stream)))
    torch.cuda.synchronize()
    return flags.cpu().numpy(), int(count.item())


def sweep_round_seed(pairs, connectivity=8, block_rows=0, split_sweeps=False, seam_rows=-1,
                     workspace=None, stream=None):
    """One in-place sweep round over a batch of (marker, mask) pairs whose
    last pass (upward raster pass and its fix-up, or with split_sweeps the
    upward column pass) sets the megapixel seeding flags; returns one uint8
    numpy flag grid (ceil(h/64), ceil(w/64)) per image.  seam_rows: rows of
    the raster fix-up passes (0 none, negative the library's choice)."""
    torch = _torch()
    arr, dtype = _batch_array(pairs)
    grids = [((mk.h + MP_SIZE - 1) // MP_SIZE, (mk.w + MP_SIZE - 1) // MP_SIZE) for mk, _ in pairs]
    total = sum(a * b for a, b in grids)
    flags = torch.full((total,), 0xEE, dtype=torch.uint8, device=pairs[0][0].buf.device)
    opt = default_options(split_sweeps=1 if split_sweeps else 0, seam_rows=seam_rows)
    _raise(lib().iwpp_sweep_round_seed(arr, len(pairs), dtype, connectivity, block_rows,
                                       ctypes.byref(opt), flags.data_ptr(),
                                       workspace.handle if workspace else None, _stream(stream)))
    torch.cuda.synchronize()
    host = flags.cpu().numpy()
    out, off = [], 0
    for a, b in grids:
        out.append(host[off:off + a * b].reshape(a, b).copy())
        off += a * b
    return out


def MorphologicalReconstruction(marker, mask, pitchMarker, pitchMask, size):
    """The reference's entry point: device pointers, pitches in bytes, square
    uint8 image of side `size`, 4-connectivity, synchronous, in place."""
    lib().MorphologicalReconstruction(marker, mask, pitchMarker, pitchMask, size)


# ------------------------------------------------------------- synthetic ----
def synth_mask(height, width, dtype=U8, maxval=None, sigma=4.0, ndisks=0, rmin=50, rmax=400,
               seed=1313, nthreads=0):
    np_dt = np.uint8 if dtype == U8 else np.uint16
    maxval = maxval if maxval is not None else (255 if dtype == U8 else 4095)
    out = np.zeros((height, width), np_dt)
    rc = synth_lib().iwpp_synth_mask(out.ctypes.data, out.shape[1] * out.itemsize, width, height, dtype,
                                     maxval, float(sigma), ndisks, rmin, rmax, seed, nthreads)
    if rc:
        raise ValueError(f"iwpp_synth_mask failed ({rc})")
    return out


def synth_marker_hdome(mask, h):
    out = np.zeros_like(mask)
    dt = U8 if mask.dtype == np.uint8 else U16
    rc = synth_lib().iwpp_synth_marker_hdome(out.ctypes.data, mask.ctypes.data, out.shape[1] * out.itemsize,
                                             mask.shape[1] * mask.itemsize, mask.shape[1], mask.shape[0], dt,
                                             int(h))
    if rc:
        raise ValueError(f"iwpp_synth_marker_hdome failed ({rc})")
    return out


def synth_marker_seeds(mask, count, seed):
    out = np.zeros_like(mask)
    dt = U8 if mask.dtype == np.uint8 else U16
    rc = synth_lib().iwpp_synth_marker_seeds(out.ctypes.data, mask.ctypes.data, out.shape[1] * out.itemsize,
                                             mask.shape[1] * mask.itemsize, mask.shape[1], mask.shape[0], dt,
                                             int(count), seed)
    if rc:
        raise ValueError(f"iwpp_synth_marker_seeds failed ({rc})")
    return out


def synth_batch_shapes(n, seed):
    w = np.zeros(n, np.int32)
    h = np.zeros(n, np.int32)
    k = np.zeros(n, np.int32)
    hh = np.zeros(n, np.int32)
    rc = synth_lib().iwpp_synth_batch_shapes(n, seed, w.ctypes.data, h.ctypes.data,
                                             k.ctypes.data, hh.ctypes.data)
    if rc:
        raise ValueError("iwpp_synth_batch_shapes failed")
    return w, h, k, hh


def synth_random_case(height, width, seed):
    j = np.zeros((height, width), np.uint8)
    i = np.zeros((height, width), np.uint8)
    rc = synth_lib().iwpp_synth_random_case(j.ctypes.data, i.ctypes.data, j.shape[1],
                                            i.shape[1], width, height, seed)
    if rc:
        raise ValueError("iwpp_synth_random_case failed")
    return j, i


# ------------------------------------------------------------ PGM files -----
class PGMError(IOError):
    pass


def pgm_read(path):
    """numpy array (uint8 when maxval <= 255, else uint16) and maxval."""
    L = synth_lib()
    data, w, h, mv = ctypes.c_void_p(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = L.iwpp_pgm_read(os.fsencode(path), ctypes.byref(data), ctypes.byref(w),
                         ctypes.byref(h), ctypes.byref(mv))
    if rc != 0:
        raise PGMError(rc, f"iwpp_pgm_read({path}) failed")
    try:
        dt = np.uint8 if mv.value <= 255 else np.uint16
        n = w.value * h.value * np.dtype(dt).itemsize
        buf = (ctypes.c_uint8 * n).from_address(data.value)
        out = np.frombuffer(buf, dtype=dt).reshape(h.value, w.value).copy()
    finally:
        L.iwpp_pgm_free(data)
    return out, mv.value


def pgm_write(path, a, maxval=None):
    a = np.ascontiguousarray(a)
    if a.dtype not in (np.uint8, np.uint16) or a.ndim != 2:
        raise TypeError("2-D uint8 or uint16 array")
    mv = maxval if maxval is not None else (255 if a.dtype == np.uint8 else 65535)
    if (a.dtype == np.uint8) != (mv <= 255):
        raise ValueError("maxval does not match the array type")
    rc = synth_lib().iwpp_pgm_write(os.fsencode(path), a.ctypes.data, a.strides[0], a.shape[1],
                                    a.shape[0], mv)
    if rc != 0:
        raise PGMError(rc, f"iwpp_pgm_write({path}) failed")
