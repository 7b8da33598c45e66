"""ctypes view of the oracle (the C restatement of the reference CPU path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker (or the timed CPU
baseline), never the thing under test.  See mkde_oracle.c for the citations of
the reference functions each entry restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmkde_oracle.so")
REF_KDTREE_PATH = os.path.join(HERE, "_ref", "libref_kdtree.so")

REF32 = 0
FP64 = 1


class Grid(ctypes.Structure):
    """Raster geometry as the reference's AsciiRaster (AsciiRaster.h:9-33)."""

    _fields_ = [
        ("ncols", ctypes.c_int64),
        ("nrows", ctypes.c_int64),
        ("xll", ctypes.c_float),
        ("yll", ctypes.c_float),
        ("cs", ctypes.c_float),
        ("nodata", ctypes.c_float),
    ]


def grid(ncols, nrows, xll, yll, cs, nodata=-9999.0):
    return Grid(int(ncols), int(nrows), float(xll), float(yll), float(cs), float(nodata))


_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_U8P = ctypes.POINTER(ctypes.c_uint8)
_GP = ctypes.POINTER(Grid)

FIXED_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_float)
ADAPT_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_float, ctypes.c_float)

_lib = None


def build():
    """Compile the oracle (and oracle/_ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_native(out_dir):
    """Compile a timing build of the same source with -O3 -march=native for
    the host it runs on (the CPU baseline; the parity build above stays -O2,
    portable).  Returns the library path."""
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "libmkde_oracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
                    "-shared", "-fPIC", "-o", path, os.path.join(HERE, "mkde_oracle.c"), "-lm"], check=True)
    return path


def use_library(path):
    """Load the oracle from another build of mkde_oracle.c (e.g. build_native)."""
    global _lib
    _lib = _load(path)
    return _lib


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _load(LIB_PATH)
    return _lib


def _load(path):
    L = ctypes.CDLL(path)
    sig = {
        "orc_version": (ctypes.c_int, []),
        "orc_col_x": (ctypes.c_float, [_GP, ctypes.c_int64]),
        "orc_row_y": (ctypes.c_float, [_GP, ctypes.c_int64]),
        "orc_distance2": (ctypes.c_double, [ctypes.c_float] * 4),
        "orc_kernel": (ctypes.c_float, [ctypes.c_double, ctypes.c_double]),
        "orc_kernel64": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
        "orc_rule_of_thumb": (ctypes.c_float, [ctypes.c_int64, _FP, _FP, _FP]),
        "orc_mark_boundary": (ctypes.c_int64, [_GP, _FP, _U8P]),
        "orc_dist_to_boundary": (None, [ctypes.c_int64, _FP, _FP, _GP, _U8P, _DP]),
        "orc_narrow_box": (None, [_GP, ctypes.c_float, ctypes.c_float, ctypes.c_float] + [ctypes.POINTER(ctypes.c_int)] * 4),
        "orc_ec_weights": (None, [ctypes.c_int64, _FP, _FP, _DP, _FP, _GP, _FP, _FP, _DP, _I64P]),
        "orc_density_tile": (None, [ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _GP, _FP, _FP, _DP, _I64P]),
        "orc_range_candidates": (ctypes.c_int64, [ctypes.c_int64, _FP, _FP, _FP, _I64P, _I32P]),
        "orc_loo_density": (None, [ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _DP, _FP, _DP]),
        "orc_loglik_ref32": (ctypes.c_float, [ctypes.c_int64, _FP]),
        "orc_loglik_fp64": (ctypes.c_double, [ctypes.c_int64, _DP]),
        "orc_pilot_density": (None, [ctypes.c_int64, _FP, _FP, _FP, _FP, ctypes.c_float, _FP, _DP]),
        "orc_pilot_density_queries": (None, [ctypes.c_int64, _FP, _FP, _FP, _FP, ctypes.c_float, ctypes.c_int64,
                                         _I64P, _FP]),
    "orc_point_density_binned": (None, [ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, ctypes.c_int,
                                            ctypes.c_int64, _I64P, _DP, _I64P]),
        "orc_compgml": (ctypes.c_float, [ctypes.c_int64, _FP]),
        "orc_compgml64": (ctypes.c_double, [ctypes.c_int64, _DP]),
        "orc_abramson": (None, [ctypes.c_int64, ctypes.c_float, ctypes.c_float, ctypes.c_float, _FP, _FP]),
        "orc_abramson64": (None, [ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _DP, _DP]),
        "orc_golden_section": (ctypes.c_float, [ctypes.c_float, FIXED_FN, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _DP]),
        "orc_hooke_jeeves": (ctypes.c_int, [ctypes.c_float] * 4 + [ADAPT_FN, ctypes.c_void_p, _FP]),
        "orc_tile_point_filter": (ctypes.c_int64, [ctypes.c_int64, _FP, _FP, _FP, _GP, _I32P]),
        "orc_msvc_rand_points": (None, [ctypes.c_int64, ctypes.c_uint32, _FP, _FP, _I32P]),
        "orc_num_threads": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def col_x(g, col):
    return lib().orc_col_x(ctypes.byref(g), col)


def row_y(g, row):
    return lib().orc_row_y(ctypes.byref(g), row)


def distance2(x0, y0, x1, y1):
    return lib().orc_distance2(x0, y0, x1, y1)


def kernel(h2, d2):
    return lib().orc_kernel(h2, d2)


def kernel64(h2, d2):
    return lib().orc_kernel64(h2, d2)


def rule_of_thumb(x, y, w):
    x, y, w = _f32(x), _f32(y), _f32(w)
    return lib().orc_rule_of_thumb(len(x), _p(x, _FP), _p(y, _FP), _p(w, _FP))


def mark_boundary(g, mask):
    mask = _f32(mask).reshape(-1)
    flags = np.zeros(mask.size, dtype=np.uint8)
    nb = lib().orc_mark_boundary(ctypes.byref(g), _p(mask, _FP), _p(flags, _U8P))
    return flags.reshape(g.nrows, g.ncols), nb


def dist_to_boundary(x, y, g, flags):
    x, y = _f32(x), _f32(y)
    flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
    out = np.empty(len(x), dtype=np.float64)
    lib().orc_dist_to_boundary(len(x), _p(x, _FP), _p(y, _FP), ctypes.byref(g), _p(flags, _U8P), _p(out, _DP))
    return out


def narrow_box(g, px, py, h):
    v = [ctypes.c_int() for _ in range(4)]
    lib().orc_narrow_box(ctypes.byref(g), px, py, h, *[ctypes.byref(t) for t in v])
    return tuple(t.value for t in v)


def ec_weights(x, y, dist, h, g, mask):
    """Returns (ew float32 as the reference, ew float64 oracle, n_eval)."""
    x, y, h = _f32(x), _f32(y), _f32(h)
    dist = np.ascontiguousarray(dist, dtype=np.float64)
    mask = _f32(mask).reshape(-1)
    n = len(x)
    ew = np.empty(n, dtype=np.float32)
    ew64 = np.empty(n, dtype=np.float64)
    ne = ctypes.c_int64()
    lib().orc_ec_weights(n, _p(x, _FP), _p(y, _FP), _p(dist, _DP), _p(h, _FP), ctypes.byref(g), _p(mask, _FP),
                         _p(ew, _FP), _p(ew64, _DP), ctypes.byref(ne))
    return ew, ew64, ne.value


def density_tile(x, y, w, ew, h, tile, tmask=None, want64=True):
    """Returns (den float32 as the reference, den float64 oracle or None, n_eval)."""
    x, y, w, ew, h = _f32(x), _f32(y), _f32(w), _f32(ew), _f32(h)
    if tmask is not None:
        tmask = _f32(tmask).reshape(-1)
    out = np.empty(tile.nrows * tile.ncols, dtype=np.float32)
    out64 = np.empty(tile.nrows * tile.ncols, dtype=np.float64) if want64 else None
    ne = ctypes.c_int64()
    lib().orc_density_tile(len(x), _p(x, _FP), _p(y, _FP), _p(w, _FP), _p(ew, _FP), _p(h, _FP), ctypes.byref(tile),
                           _p(tmask, _FP), _p(out, _FP), _p(out64, _DP), ctypes.byref(ne))
    shape = (tile.nrows, tile.ncols)
    return out.reshape(shape), (out64.reshape(shape) if want64 else None), ne.value


def range_candidates(x, y, h):
    """CSR (offsets[n+1] int64, idx int32) of N1 candidate sets per query."""
    x, y, h = _f32(x), _f32(y), _f32(h)
    n = len(x)
    total = lib().orc_range_candidates(n, _p(x, _FP), _p(y, _FP), _p(h, _FP), None, None)
    offs = np.zeros(n + 1, dtype=np.int64)
    idx = np.zeros(max(total, 1), dtype=np.int32)
    lib().orc_range_candidates(n, _p(x, _FP), _p(y, _FP), _p(h, _FP), _p(offs, _I64P), _p(idx, _I32P))
    return offs, idx[:total]


def loo_density(x, y, w, ew, h):
    """Returns dict of den1/den0 in float32 (reference order) and float64."""
    x, y, w, ew, h = _f32(x), _f32(y), _f32(w), _f32(ew), _f32(h)
    n = len(x)
    d1 = np.empty(n, np.float32)
    d1_64 = np.empty(n, np.float64)
    d0 = np.empty(n, np.float32)
    d0_64 = np.empty(n, np.float64)
    lib().orc_loo_density(n, _p(x, _FP), _p(y, _FP), _p(w, _FP), _p(ew, _FP), _p(h, _FP),
                          _p(d1, _FP), _p(d1_64, _DP), _p(d0, _FP), _p(d0_64, _DP))
    return {"den1": d1, "den1_64": d1_64, "den0": d0, "den0_64": d0_64}


def loglik_ref32(den1):
    den1 = _f32(den1)
    return lib().orc_loglik_ref32(len(den1), _p(den1, _FP))


def loglik_fp64(den1):
    den1 = np.ascontiguousarray(den1, dtype=np.float64)
    return lib().orc_loglik_fp64(len(den1), _p(den1, _DP))


def pilot_density(x, y, w, ew, h):
    x, y, w, ew = _f32(x), _f32(y), _f32(w), _f32(ew)
    n = len(x)
    d0 = np.empty(n, np.float32)
    d0_64 = np.empty(n, np.float64)
    lib().orc_pilot_density(n, _p(x, _FP), _p(y, _FP), _p(w, _FP), _p(ew, _FP), h, _p(d0, _FP), _p(d0_64, _DP))
    return d0, d0_64


def point_density_binned(x, y, w, ew, h, loo, queries=None):
    """FP64 density at the query points (all points when queries is None) with
    each source's own bandwidth, sources found through a grid of bins; loo
    excludes the query itself.  Returns (den64[len(queries)], contributing pairs)."""
    x, y, w, ew, h = _f32(x), _f32(y), _f32(w), _f32(ew), _f32(h)
    n = len(x)
    q = np.arange(n, dtype=np.int64) if queries is None else np.ascontiguousarray(queries, dtype=np.int64)
    out = np.empty(len(q), np.float64)
    cnt = ctypes.c_int64()
    lib().orc_point_density_binned(n, _p(x, _FP), _p(y, _FP), _p(w, _FP), _p(ew, _FP), _p(h, _FP), int(bool(loo)),
                                   len(q), _p(q, _I64P), _p(out, _DP), ctypes.byref(cnt))
    return out, cnt.value


def pilot_density_queries(x, y, w, ew, h, queries):
    """The reference's float pilot loop (all sources) for the given queries."""
    x, y, w, ew = _f32(x), _f32(y), _f32(w), _f32(ew)
    q = np.ascontiguousarray(queries, dtype=np.int64)
    out = np.empty(len(q), np.float32)
    lib().orc_pilot_density_queries(len(x), _p(x, _FP), _p(y, _FP), _p(w, _FP), _p(ew, _FP), h, len(q), _p(q, _I64P),
                                    _p(out, _FP))
    return out


def compgml(den0):
    den0 = _f32(den0)
    return lib().orc_compgml(len(den0), _p(den0, _FP))


def compgml64(den0):
    den0 = np.ascontiguousarray(den0, dtype=np.float64)
    return lib().orc_compgml64(len(den0), _p(den0, _DP))


def abramson(h0, alpha, gml, den0):
    den0 = _f32(den0)
    out = np.empty(len(den0), np.float32)
    lib().orc_abramson(len(den0), h0, alpha, gml, _p(den0, _FP), _p(out, _FP))
    return out


def abramson64(h0, alpha, gml, den0):
    den0 = np.ascontiguousarray(den0, dtype=np.float64)
    out = np.empty(len(den0), np.float64)
    lib().orc_abramson64(len(den0), h0, alpha, gml, _p(den0, _DP), _p(out, _DP))
    return out


def golden_section(h, fn):
    """fn(h: float) -> log-likelihood.  Returns (hopt, iterations, probes)."""
    cb = FIXED_FN(lambda _u, hh: float(fn(hh)))
    it = ctypes.c_int()
    trace = np.zeros(64, np.float64)
    r = lib().orc_golden_section(h, cb, None, ctypes.byref(it), _p(trace, _DP))
    return r, it.value, trace[: 2 * it.value].reshape(-1, 2)


def hooke_jeeves(h0, alpha0, step_h, step_a, fn):
    """fn(h, alpha) -> log-likelihood.  Returns ((h, alpha, Lmax), iterations)."""
    cb = ADAPT_FN(lambda _u, hh, aa: float(fn(hh, aa)))
    out = np.zeros(3, np.float32)
    it = lib().orc_hooke_jeeves(h0, alpha0, step_h, step_a, cb, None, _p(out, _FP))
    return (float(out[0]), float(out[1]), float(out[2])), it


def tile_point_filter(x, y, h, tile):
    x, y, h = _f32(x), _f32(y), _f32(h)
    idx = np.zeros(len(x), np.int32)
    k = lib().orc_tile_point_filter(len(x), _p(x, _FP), _p(y, _FP), _p(h, _FP), ctypes.byref(tile), _p(idx, _I32P))
    return idx[:k]


def msvc_rand_points(n, seed=100):
    x = np.empty(n, np.float32)
    y = np.empty(n, np.float32)
    raw = np.empty(2 * n, np.int32)
    lib().orc_msvc_rand_points(n, seed, _p(x, _FP), _p(y, _FP), _p(raw, _I32P))
    return x, y, raw


def num_threads():
    return lib().orc_num_threads()


# ------------------------------------------------------------ oracle/_ref ---
_ref = None


def ref_kdtree_available():
    return os.path.exists(REF_KDTREE_PATH)


def ref_kdtree_range(x, y, h):
    """Candidate sets from the reference's own KDtree::SearchRange (CSR)."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_KDTREE_PATH)
        _ref.ref_kdtree_range.restype = ctypes.c_int64
        _ref.ref_kdtree_range.argtypes = [ctypes.c_int, _FP, _FP, _FP, _I64P, _I64P, _I32P]
        _ref.ref_kdtree_num_nodes.restype = ctypes.c_int
        _ref.ref_kdtree_num_nodes.argtypes = [ctypes.c_int, _FP, _FP]
    x, y, h = _f32(x), _f32(y), _f32(h)
    n = len(x)
    counts = np.zeros(n, np.int64)
    total = _ref.ref_kdtree_range(n, _p(x, _FP), _p(y, _FP), _p(h, _FP), _p(counts, _I64P), None, None)
    offs = np.zeros(n + 1, np.int64)
    offs[1:] = np.cumsum(counts)
    idx = np.zeros(max(total, 1), np.int32)
    _ref.ref_kdtree_range(n, _p(x, _FP), _p(y, _FP), _p(h, _FP), _p(counts, _I64P), _p(offs, _I64P), _p(idx, _I32P))
    return offs, idx[:total]
