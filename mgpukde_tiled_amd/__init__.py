"""mgpukde_tiled_amd -- MI355X-native tiled Gaussian KDE (host mirror over libmkde).

Python view of the C ABI in include/mkde.h.  Method names follow the reference
functions of kde_cuda_kdtr_tiled.cu whose role each call takes (MarkBoundary,
CalcDist2Boundary, EdgeCorrectionWeightsExact, LogLikelihood,
MLE_FixedBandWidth, hooke_jeeves, ...), so parity tests read like tests of the
reference's own functions.  Every compute call runs the HIP kernels of
libmkde.so on the GPU; there is no CPU fallback: if the library or a GPU is
missing, the calls raise.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmkde.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "mkde.h")

MKDE_OK = 0
STATUS_NAMES = {1: "MKDE_E_ARG", 2: "MKDE_E_HIP", 3: "MKDE_E_NOMEM", 4: "MKDE_E_STATE", 5: "MKDE_E_IO",
                6: "MKDE_E_RANGE", 7: "MKDE_E_PENDING", 8: "MKDE_E_CANCELLED"}


class MkdeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(code, str(code)), msg))
        self.code = code


class Grid(ctypes.Structure):
    """mkde_grid: geometry of a mask / tile (AsciiRaster.h:9-33 geometry)."""

    _fields_ = [
        ("ncols", ctypes.c_int64),
        ("nrows", ctypes.c_int64),
        ("xll", ctypes.c_float),
        ("yll", ctypes.c_float),
        ("cellsize", ctypes.c_float),
        ("nodata", ctypes.c_float),
    ]

    def __repr__(self):
        return "Grid(ncols=%d, nrows=%d, xll=%r, yll=%r, cellsize=%r, nodata=%r)" % (
            self.ncols, self.nrows, self.xll, self.yll, self.cellsize, self.nodata)


def grid(ncols, nrows, xll, yll, cellsize, nodata=-9999.0):
    return Grid(int(ncols), int(nrows), float(xll), float(yll), float(cellsize), float(nodata))


class JobOpts(ctypes.Structure):
    """mkde_job_opts (see include/mkde.h)."""

    _fields_ = [
        ("hoption", ctypes.c_int),
        ("edge_correction", ctypes.c_int),
        ("optimise", ctypes.c_int),
        ("alpha", ctypes.c_float),
        ("recompute_boundary", ctypes.c_int),
    ]


class JobStats(ctypes.Structure):
    """mkde_job_stats (see include/mkde.h)."""

    _fields_ = [
        ("h0", ctypes.c_float),
        ("h_opt", ctypes.c_float),
        ("alpha_opt", ctypes.c_float),
        ("g", ctypes.c_double),
        ("n_ec_pairs", ctypes.c_int64),
        ("ms_boundary", ctypes.c_double),
        ("ms_ec", ctypes.c_double),
        ("ms_bandwidth", ctypes.c_double),
        ("ms_density", ctypes.c_double),
        ("ms_total", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_U8P = ctypes.POINTER(ctypes.c_uint8)
_IP = ctypes.POINTER(ctypes.c_int)
class ServeJob(ctypes.Structure):
    """mkde_serve_job (see include/mkde.h)."""

    _fields_ = [
        ("n", ctypes.c_int64),
        ("x", ctypes.POINTER(ctypes.c_float)),
        ("y", ctypes.POINTER(ctypes.c_float)),
        ("w", ctypes.POINTER(ctypes.c_float)),
        ("h", ctypes.POINTER(ctypes.c_float)),
        ("grid", Grid),
        ("mask", ctypes.POINTER(ctypes.c_float)),
        ("hoption", ctypes.c_int),
        ("edge_correction", ctypes.c_int),
        ("optimise", ctypes.c_int),
        ("alpha", ctypes.c_float),
        ("n_candidates", ctypes.c_int),
        ("cand_lo", ctypes.c_float),
        ("cand_hi", ctypes.c_float),
        ("out", ctypes.POINTER(ctypes.c_float)),
        ("one_candidate_per_pass", ctypes.c_int),
    ]


class ServeResult(ctypes.Structure):
    """mkde_serve_result (see include/mkde.h)."""

    _fields_ = [
        ("status", ctypes.c_int),
        ("message", ctypes.c_char * 256),
        ("slot", ctypes.c_int),
        ("queue_ms", ctypes.c_double),
        ("run_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("h0", ctypes.c_float),
        ("h_sel", ctypes.c_float),
        ("alpha_sel", ctypes.c_float),
        ("footprint", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["message"] = self.message.decode(errors="replace")
        return d


class ServerInfo(ctypes.Structure):
    """mkde_server_info (see include/mkde.h)."""

    _fields_ = [(k, ctypes.c_int if k == "slots" else ctypes.c_int64) for k in
                ("slots", "budget", "held", "peak_committed", "submitted", "finished", "refused", "releases",
                 "queued", "running")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


MKDE_H_SEARCH = 3
MKDE_E_ARG = 1
MKDE_E_STATE = 4
MKDE_E_RANGE = 6
MKDE_E_NOMEM = 3
MKDE_E_PENDING = 7
MKDE_E_CANCELLED = 8

_GP = ctypes.POINTER(Grid)
_CTX = ctypes.c_void_p

# name -> argtypes (restype is int status unless listed in _RESTYPES)
SIGNATURES = {
    "mkde_last_error": [],
    "mkde_version": [],
    "mkde_ctx_create": [ctypes.c_int, ctypes.POINTER(_CTX)],
    "mkde_ctx_destroy": [_CTX],
    "mkde_synchronize": [_CTX],
    "mkde_points_upload": [_CTX, ctypes.c_int64, _FP, _FP, _FP],
    "mkde_num_points": [_CTX, _I64P],
    "mkde_mask_upload": [_CTX, _GP, _FP],
    "mkde_rule_of_thumb": [_CTX, _FP],
    "mkde_mark_boundary": [_CTX, _U8P, _I64P],
    "mkde_dist_to_boundary": [_CTX, _DP],
    "mkde_set_dist_to_boundary": [_CTX, _DP],
    "mkde_set_bandwidth_fixed": [_CTX, ctypes.c_float],
    "mkde_set_bandwidths": [_CTX, _FP],
    "mkde_get_bandwidths": [_CTX, _FP],
    "mkde_ec_weights": [_CTX, _FP, _I64P],
    "mkde_set_ec_weights": [_CTX, _FP],
    "mkde_get_ec_weights": [_CTX, _FP],
    "mkde_range_candidates": [_CTX, _I64P, _I32P, ctypes.c_int64, _I64P],
    "mkde_density_at_points": [_CTX, ctypes.c_int, _FP],
    "mkde_density_at_points_noalloc": [_CTX, ctypes.c_int, _FP],
    "mkde_loo_loglik": [_CTX, ctypes.c_int, _FP, ctypes.c_int, _DP],
    "mkde_loo_loglik_adaptive": [_CTX, ctypes.c_int, _FP, _FP, ctypes.c_int, _DP],
    "mkde_abramson": [_CTX, ctypes.c_float, ctypes.c_float, ctypes.c_int, _FP, _FP, _DP],
    "mkde_mle_fixed_bandwidth": [_CTX, ctypes.c_float, ctypes.c_int, _FP, _IP],
    "mkde_hooke_jeeves": [_CTX, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int, _FP,
                          _IP],
    "mkde_density_tile": [_CTX, _GP, _FP, _FP],
    "mkde_count_density_pairs": [_CTX, _GP, _FP, _I64P],
    "mkde_tile_upload": [_CTX, _GP, _FP],
    "mkde_density_tile_resident": [_CTX],
    "mkde_tile_download": [_CTX, _FP],
    "mkde_raster_upload": [_CTX, ctypes.c_int64, ctypes.c_int64, _DP, ctypes.c_float, ctypes.c_int64, ctypes.c_int64,
                           _FP],
    "mkde_density_raster": [_CTX, ctypes.c_int64, ctypes.c_int64, _DP, ctypes.c_float, ctypes.c_int64,
                            ctypes.c_int64, _FP, _FP],
    "mkde_last_kernel_ms": [_CTX, ctypes.c_char_p, _DP],
    "mkde_set_option": [_CTX, ctypes.c_char_p, ctypes.c_int64],
    "mkde_autotune_job": [_CTX, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64],
    "mkde_autotune_search": [_CTX, ctypes.c_int, _FP, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int64],
    "mkde_autotune_lookup": [_CTX, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "mkde_get_option": [_CTX, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)],
    "mkde_run_job": [_CTX, ctypes.POINTER(JobOpts), ctypes.POINTER(JobStats)],
    "mkde_density_stats": [_CTX, ctypes.c_int, _I64P],
    "mkde_device_memory": [_CTX, _I64P, _I64P],
    "mkde_tiff_info": [ctypes.c_char_p, _I64P, _I64P, _DP, _DP, _IP, _IP, _IP],
    "mkde_tiff_read": [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _FP],
    "mkde_writer_open": [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, _DP, ctypes.c_float, ctypes.c_char_p,
                         ctypes.POINTER(ctypes.c_void_p)],
    "mkde_writer_put": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _FP],
    "mkde_writer_close": [ctypes.c_void_p, ctypes.c_int],
    "mkde_csv_read_points": [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _I64P, _FP, _FP, _FP, _FP],
    "mkde_csv_write_points": [ctypes.c_char_p, ctypes.c_int64, _FP, _FP, _FP, _FP],
    "mkde_lzw_encode": [_U8P, ctypes.c_int64, _U8P, ctypes.c_int64],
    "mkde_lzw_decode": [_U8P, ctypes.c_int64, _U8P, ctypes.c_int64],
    "mkde_ctx_memory": [_CTX, _I64P],
    "mkde_ctx_release": [_CTX],
    "mkde_tile_from_mask": [_CTX],
    "mkde_job_footprint": [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    "mkde_server_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)],
    "mkde_server_submit": [ctypes.c_void_p, ctypes.POINTER(ServeJob), _I64P],
    "mkde_server_wait": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ServeResult)],
    "mkde_server_cancel": [ctypes.c_void_p, ctypes.c_int64],
    "mkde_server_info_get": [ctypes.c_void_p, ctypes.POINTER(ServerInfo)],
    "mkde_server_destroy": [ctypes.c_void_p],
}
_RESTYPES = {"mkde_last_error": ctypes.c_char_p, "mkde_version": ctypes.c_char_p, "mkde_lzw_encode": ctypes.c_int64,
             "mkde_lzw_decode": ctypes.c_int64, "mkde_job_footprint": ctypes.c_int64}

_lib = None


def load_library(path=LIB_PATH):
    """Load libmkde.so (built in-tree by mgpukde_tiled_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError("libmkde.so is not built (%s); run `python -m mgpukde_tiled_amd.build`" % path)
        L = ctypes.CDLL(path)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def _check(status):
    if status != MKDE_OK:
        raise MkdeError(status, _lib.mkde_last_error().decode(errors="replace"))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _vec(a, n, what, dtype=np.float32):
    """1-D contiguous copy of a with exactly n entries (the native side reads n)."""
    v = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    if v.size != n:
        raise ValueError("%s has %d entries, expected %d" % (what, v.size, n))
    return v


def _mask(mask, g, what="mask"):
    """None, or the grid's ncols*nrows mask values as a contiguous float32 vector."""
    return None if mask is None else _vec(mask, int(g.ncols) * int(g.nrows), what)


class Context:
    """A device context (one HIP device, one stream) with resident state."""

    def __init__(self, device=0):
        load_library()
        h = _CTX()
        _check(_lib.mkde_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            _lib.mkde_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def synchronize(self):
        _check(_lib.mkde_synchronize(self._h))

    # -- inputs ---------------------------------------------------------
    def points_upload(self, x, y, w=None):
        """CopyToDeviceSamplePoints + index build (replaces BuildGPUKDtree)."""
        x = _f32(x).reshape(-1)
        y = _vec(y, x.size, "y")
        w = None if w is None else _vec(w, x.size, "w")
        self._x, self._y, self._w = x, y, w
        _check(_lib.mkde_points_upload(self._h, len(x), _p(x, _FP), _p(y, _FP), _p(w, _FP)))
        self.n = len(x)

    def mask_upload(self, g, values):
        v = _f32(values).reshape(-1)
        if v.size != g.ncols * g.nrows:
            raise ValueError("mask has %d values, grid needs %d" % (v.size, g.ncols * g.nrows))
        _check(_lib.mkde_mask_upload(self._h, ctypes.byref(g), _p(v, _FP)))
        self.mask_grid = g

    # -- reference functions ---------------------------------------------
    def BandWidth_rule_of_thumb(self):
        """h0 = sqrtf(BandWidth2(points)) (tiled:64-68, 481)."""
        h = ctypes.c_float()
        _check(_lib.mkde_rule_of_thumb(self._h, ctypes.byref(h)))
        return h.value

    def MarkBoundary(self):
        g = self.mask_grid
        flags = np.zeros(g.ncols * g.nrows, np.uint8)
        nb = ctypes.c_int64()
        _check(_lib.mkde_mark_boundary(self._h, _p(flags, _U8P), ctypes.byref(nb)))
        return flags.reshape(g.nrows, g.ncols), nb.value

    def CalcDist2Boundary(self):
        d = np.empty(self.n, np.float64)
        _check(_lib.mkde_dist_to_boundary(self._h, _p(d, _DP)))
        return d

    def set_dist_to_boundary(self, d):
        d = _vec(d, self.n, "distances", np.float64)
        _check(_lib.mkde_set_dist_to_boundary(self._h, _p(d, _DP)))

    def set_bandwidth(self, h):
        _check(_lib.mkde_set_bandwidth_fixed(self._h, float(h)))

    def set_bandwidths(self, hs):
        hs = _vec(hs, self.n, "bandwidths")
        _check(_lib.mkde_set_bandwidths(self._h, _p(hs, _FP)))

    def get_bandwidths(self):
        out = np.empty(self.n, np.float32)
        _check(_lib.mkde_get_bandwidths(self._h, _p(out, _FP)))
        return out

    def EdgeCorrectionWeightsExact(self):
        """EC weights with the current bandwidths; returns (ew, n_eval)."""
        ew = np.empty(self.n, np.float32)
        ne = ctypes.c_int64()
        _check(_lib.mkde_ec_weights(self._h, _p(ew, _FP), ctypes.byref(ne)))
        return ew, ne.value

    def set_ec_weights(self, ew=None):
        ew = None if ew is None else _vec(ew, self.n, "EC weights")
        _check(_lib.mkde_set_ec_weights(self._h, _p(ew, _FP)))

    def get_ec_weights(self):
        out = np.empty(self.n, np.float32)
        _check(_lib.mkde_get_ec_weights(self._h, _p(out, _FP)))
        return out

    def SearchRange_all(self):
        """Candidate sets of every point (CSR: offsets, indexes ascending)."""
        tot = ctypes.c_int64()
        offs = np.zeros(self.n + 1, np.int64)
        _check(_lib.mkde_range_candidates(self._h, _p(offs, _I64P), None, 0, ctypes.byref(tot)))
        idx = np.zeros(max(tot.value, 1), np.int32)
        _check(_lib.mkde_range_candidates(self._h, _p(offs, _I64P), _p(idx, _I32P), tot.value, ctypes.byref(tot)))
        return offs, idx[: tot.value]

    def DensityAtPoints(self, loo=False):
        out = np.empty(self.n, np.float32)
        _check(_lib.mkde_density_at_points(self._h, 1 if loo else 0, _p(out, _FP)))
        return out

    def density_at_points_noalloc(self, loo, den_dev):
        """The allocation-free form: den_dev is a device address (int) of n
        floats the caller owns; nothing is allocated on the device."""
        _check(_lib.mkde_density_at_points_noalloc(self._h, 1 if loo else 0, ctypes.cast(den_dev, _FP)))

    def LogLikelihood(self, h_candidates, edge_correction=True):
        """FP64 LOO log-likelihood of each fixed bandwidth (one batched pass)."""
        hc = _f32(np.atleast_1d(h_candidates))
        out = np.empty(len(hc), np.float64)
        _check(_lib.mkde_loo_loglik(self._h, len(hc), _p(hc, _FP), 1 if edge_correction else 0, _p(out, _DP)))
        return out

    def LogLikelihoodAdaptive(self, h, alpha, edge_correction=True):
        h = _f32(np.atleast_1d(h))
        a = _f32(np.broadcast_to(np.atleast_1d(alpha), h.shape))
        out = np.empty(len(h), np.float64)
        _check(_lib.mkde_loo_loglik_adaptive(self._h, len(h), _p(h, _FP), _p(a, _FP), 1 if edge_correction else 0,
                                             _p(out, _DP)))
        return out

    def Abramson(self, h0, alpha=-0.5, edge_correction=True):
        """Pilot at h0, g, h_i = h0*(den0_i/g)^alpha.  Returns (h, den0, g)."""
        hs = np.empty(self.n, np.float32)
        d0 = np.empty(self.n, np.float32)
        g = ctypes.c_double()
        _check(_lib.mkde_abramson(self._h, float(h0), float(alpha), 1 if edge_correction else 0, _p(hs, _FP),
                                  _p(d0, _FP), ctypes.byref(g)))
        return hs, d0, g.value

    def MLE_FixedBandWidth(self, h, edge_correction=True):
        r = ctypes.c_float()
        it = ctypes.c_int()
        _check(_lib.mkde_mle_fixed_bandwidth(self._h, float(h), 1 if edge_correction else 0, ctypes.byref(r),
                                             ctypes.byref(it)))
        return r.value, it.value

    def hooke_jeeves(self, h0, alpha0=-0.5, step_h=None, step_a=0.1, edge_correction=True):
        if step_h is None:
            step_h = np.float32(h0) / np.float32(10)
        out = np.zeros(3, np.float32)
        it = ctypes.c_int()
        _check(_lib.mkde_hooke_jeeves(self._h, float(h0), float(alpha0), float(step_h), float(step_a),
                                      1 if edge_correction else 0, _p(out, _FP), ctypes.byref(it)))
        return (float(out[0]), float(out[1]), float(out[2])), it.value

    def KernelDensityEstimation_tile(self, tile, mask=None):
        """Density raster of one tile (nodata cells keep the mask value)."""
        m = _mask(mask, tile)
        out = np.empty(tile.ncols * tile.nrows, np.float32)
        _check(_lib.mkde_density_tile(self._h, ctypes.byref(tile), _p(m, _FP), _p(out, _FP)))
        return out.reshape(tile.nrows, tile.ncols)

    def count_density_pairs(self, tile, mask=None):
        m = _mask(mask, tile)
        n = ctypes.c_int64()
        _check(_lib.mkde_count_density_pairs(self._h, ctypes.byref(tile), _p(m, _FP), ctypes.byref(n)))
        return n.value

    # -- device-resident tile API (throughput measurement) ----------------
    def tile_upload(self, tile, mask=None):
        m = _mask(mask, tile)
        self._tile = tile
        _check(_lib.mkde_tile_upload(self._h, ctypes.byref(tile), _p(m, _FP)))

    def density_tile_resident(self):
        _check(_lib.mkde_density_tile_resident(self._h))

    def tile_download(self):
        t = self._tile
        out = np.empty(t.ncols * t.nrows, np.float32)
        _check(_lib.mkde_tile_download(self._h, _p(out, _FP)))
        return out.reshape(t.nrows, t.ncols)

    def memory(self):
        """Device bytes held by this context's buffers."""
        v = ctypes.c_int64()
        _check(_lib.mkde_ctx_memory(self._h, ctypes.byref(v)))
        return v.value

    def release(self):
        _check(_lib.mkde_ctx_release(self._h))

    def tile_from_mask(self):
        """The density tile is the resident EC mask raster."""
        self._tile = self.mask_grid
        _check(_lib.mkde_tile_from_mask(self._h))

    def raster_upload(self, ncols, nrows, gt, nodata=-9999.0, tile_cols=4096, tile_rows=4096, mask=None):
        """Make a GeoTIFF-style raster of tiles the current tile (each tile
        with the origin AsciiRasterFromGeoTIFFTile gives it)."""
        m = None if mask is None else _vec(mask, int(ncols) * int(nrows), "mask")
        gtv = _vec(gt, 6, "geotransform", np.float64)
        self._tile = grid(int(ncols), int(nrows), 0.0, 0.0, 1.0, nodata)
        _check(_lib.mkde_raster_upload(self._h, int(ncols), int(nrows), _p(gtv, _DP), float(nodata), int(tile_cols),
                                       int(tile_rows), _p(m, _FP)))

    def KernelDensityEstimation_raster(self, ncols, nrows, gt, nodata=-9999.0, tile_cols=4096, tile_rows=4096,
                                       mask=None, keep_on_device=False):
        """Density of every tile of a raster in one pass (the reference's
        GeoTIFF tile loop); returns the nrows x ncols raster, or None with
        keep_on_device (then tile_download() copies it out)."""
        m = None if mask is None else _vec(mask, int(ncols) * int(nrows), "mask")
        gtv = _vec(gt, 6, "geotransform", np.float64)
        self._tile = grid(int(ncols), int(nrows), 0.0, 0.0, 1.0, nodata)
        out = None if keep_on_device else np.empty(int(ncols) * int(nrows), np.float32)
        _check(_lib.mkde_density_raster(self._h, int(ncols), int(nrows), _p(gtv, _DP), float(nodata),
                                        int(tile_cols), int(tile_rows), _p(m, _FP), _p(out, _FP)))
        return None if out is None else out.reshape(int(nrows), int(ncols))

    def run_job(self, hoption=2, edge_correction=True, optimise=True, alpha=-0.5, recompute_boundary=True):
        """One KDE job on the resident points / mask / tile (the GPU block of
        the reference's main()).  Returns the job statistics as a dict."""
        o = JobOpts(int(hoption), 1 if edge_correction else 0, 1 if optimise else 0, float(alpha),
                    1 if recompute_boundary else 0)
        st = JobStats()
        _check(_lib.mkde_run_job(self._h, ctypes.byref(o), ctypes.byref(st)))
        return st.as_dict()

    def density_stats(self, enable):
        out = np.zeros(8, np.int64)
        _check(_lib.mkde_density_stats(self._h, 1 if enable else 0, _p(out, _I64P)))
        keys = ["candidates", "full_blocks", "partial_blocks", "full_waves", "partial_waves", "empty_waves",
                "skipped_blocks", "unused"]
        return dict(zip(keys, out.tolist()))

    def device_memory(self):
        fr, to = ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.mkde_device_memory(self._h, ctypes.byref(fr), ctypes.byref(to)))
        return fr.value, to.value

    def set_option(self, name, value):
        """mkde_set_option: "search_graph", "density_order", "points_per_bin"."""
        _check(_lib.mkde_set_option(self._h, name.encode(), int(value)))

    def autotune_job(self, hoption=2, edge_correction=True, optimise=False, alpha=-0.5, recompute_boundary=True,
                     reps=3):
        """mkde_autotune_job on the resident points / mask / tile; returns the
        report (dict) and leaves the chosen options set."""
        o = JobOpts(int(hoption), 1 if edge_correction else 0, 1 if optimise else 0, float(alpha),
                    1 if recompute_boundary else 0)
        buf = ctypes.create_string_buffer(1 << 16)
        _check(_lib.mkde_autotune_job(self._h, ctypes.cast(ctypes.byref(o), ctypes.c_void_p), int(reps), buf,
                                      len(buf)))
        return json.loads(buf.value.decode())

    def autotune_search(self, h_candidates, edge_correction=True, reps=3):
        hc = _f32(np.atleast_1d(h_candidates))
        buf = ctypes.create_string_buffer(1 << 16)
        _check(_lib.mkde_autotune_search(self._h, len(hc), _p(hc, _FP), 1 if edge_correction else 0, int(reps), buf,
                                         len(buf)))
        return json.loads(buf.value.decode())

    def autotune_lookup(self, n, cells, hoption, edge_correction=True, optimise=False):
        """True when a recorded choice for the job class was applied."""
        st = _lib.mkde_autotune_lookup(self._h, int(n), int(cells), int(hoption), 1 if edge_correction else 0,
                                       1 if optimise else 0)
        if st == MKDE_E_STATE:
            return False
        _check(st)
        return True

    def get_option(self, name):
        v = ctypes.c_int64()
        _check(_lib.mkde_get_option(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def last_kernel_ms(self, which="density"):
        v = ctypes.c_double()
        _check(_lib.mkde_last_kernel_ms(self._h, which.encode(), ctypes.byref(v)))
        return v.value


# -------------------------------------------------------------- serving --
def job_footprint(n, ncols, nrows, hoption=0, optimise=False, n_candidates=0):
    load_library()
    return int(_lib.mkde_job_footprint(int(n), int(ncols), int(nrows), int(hoption), 1 if optimise else 0,
                                       int(n_candidates)))


class Server:
    """The job server (mkde_server_*): worker slots with their own contexts
    and streams, FIFO start with HBM admission.  submit() returns a ticket;
    the arrays of a job are kept referenced here until wait() returns it."""

    def __init__(self, device=0, slots=4, hbm_budget=0):
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.mkde_server_create(int(device), int(slots), int(hbm_budget), ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, x, y, g, mask=None, out=None, w=None, h=None, hoption=0, edge_correction=True, optimise=False,
               alpha=-0.5, n_candidates=16, cand_lo=0.1, cand_hi=10.0, one_candidate_per_pass=False):
        """Queue one job; returns its ticket.  The density lands in `out`
        (allocated here when None) once wait() reports it finished."""
        x = _f32(x).reshape(-1)
        y = _vec(y, x.size, "y")
        w = None if w is None else _vec(w, x.size, "w")
        h = None if h is None else _vec(h, x.size, "h")
        m = _mask(mask, g)
        if out is None:
            out = np.empty((g.nrows, g.ncols), np.float32)
        if not (out.dtype == np.float32 and out.flags.c_contiguous and out.size == g.ncols * g.nrows):
            raise ValueError("out must be a contiguous float32 array of ncols*nrows values")
        j = ServeJob(len(x), _p(x, _FP), _p(y, _FP), _p(w, _FP), _p(h, _FP), g, _p(m, _FP), int(hoption),
                     1 if edge_correction else 0, 1 if optimise else 0, float(alpha), int(n_candidates),
                     float(cand_lo), float(cand_hi), _p(out, _FP), 1 if one_candidate_per_pass else 0)
        t = ctypes.c_int64()
        _check(_lib.mkde_server_submit(self._h, ctypes.byref(j), ctypes.byref(t)))
        self._keep[t.value] = (x, y, w, h, m, out)
        return t.value

    def wait(self, ticket, timeout_ms=-1.0):
        """(result dict, output raster) once the job has finished; None while
        it is still pending after timeout_ms."""
        r = ServeResult()
        st = _lib.mkde_server_wait(self._h, int(ticket), float(timeout_ms), ctypes.byref(r))
        if st == MKDE_E_PENDING:
            return None
        _check(st)
        keep = self._keep.pop(ticket)
        return r.as_dict(), keep[5]

    def cancel(self, ticket):
        _check(_lib.mkde_server_cancel(self._h, int(ticket)))

    def info(self):
        i = ServerInfo()
        _check(_lib.mkde_server_info_get(self._h, ctypes.byref(i)))
        return i.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            _lib.mkde_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ I/O --
def _b(path):
    return os.fsencode(path)


def tiff_info(path):
    """(ncols, nrows, geotransform, nodata or None, (block_x, block_y))."""
    load_library()
    nc, nr = ctypes.c_int64(), ctypes.c_int64()
    gt = np.zeros(6, np.float64)
    nd = ctypes.c_double()
    has = ctypes.c_int()
    bx, by = ctypes.c_int(), ctypes.c_int()
    _check(_lib.mkde_tiff_info(_b(path), ctypes.byref(nc), ctypes.byref(nr), _p(gt, _DP), ctypes.byref(nd),
                               ctypes.byref(has), ctypes.byref(bx), ctypes.byref(by)))
    return nc.value, nr.value, gt, (nd.value if has.value else None), (bx.value, by.value)


def tiff_read(path, window=None):
    """Band 1 as float32 (optionally a window (xoff, yoff, ncols, nrows))."""
    nc, nr, _, _, _ = tiff_info(path)
    xo, yo, w, h = window if window else (0, 0, nc, nr)
    out = np.empty(w * h, np.float32)
    _check(_lib.mkde_tiff_read(_b(path), xo, yo, w, h, _p(out, _FP)))
    return out.reshape(h, w)


def tiff_write(path, data, geotransform, nodata, georef_from=None, compress=True):
    load_library()
    data = _f32(data)
    nr, nc = data.shape
    gt = np.ascontiguousarray(geotransform, np.float64)
    wr = ctypes.c_void_p()
    _check(_lib.mkde_writer_open(_b(path), nc, nr, _p(gt, _DP), float(nodata),
                                 _b(georef_from) if georef_from else None, ctypes.byref(wr)))
    _check(_lib.mkde_writer_put(wr, 0, 0, nc, nr, _p(data, _FP)))
    _check(_lib.mkde_writer_close(wr, 1 if compress else 0))


def csv_read_points(path, with_h=False, sample_weights=False):
    """ReadSamplePoints: returns (x, y, h, w) float32 arrays."""
    load_library()
    n = ctypes.c_int64()
    _check(_lib.mkde_csv_read_points(_b(path), int(with_h), int(sample_weights), ctypes.byref(n), None, None, None,
                                     None))
    x, y, h, w = (np.empty(n.value, np.float32) for _ in range(4))
    _check(_lib.mkde_csv_read_points(_b(path), int(with_h), int(sample_weights), ctypes.byref(n), _p(x, _FP),
                                     _p(y, _FP), _p(h, _FP), _p(w, _FP)))
    return x, y, h, w


def csv_write_points(path, x, y, h, w):
    load_library()
    x = _f32(x).reshape(-1)
    y, h, w = _vec(y, x.size, "y"), _vec(h, x.size, "h"), _vec(w, x.size, "w")
    _check(_lib.mkde_csv_write_points(_b(path), len(x), _p(x, _FP), _p(y, _FP), _p(h, _FP), _p(w, _FP)))


def lzw_encode(data):
    load_library()
    a = np.frombuffer(bytes(data), np.uint8).copy()
    out = np.empty(2 * len(a) + 64, np.uint8)
    k = _lib.mkde_lzw_encode(_p(a, _U8P), len(a), _p(out, _U8P), len(out))
    if k < 0:
        raise MkdeError(6, "LZW output buffer too small")
    return out[:k].tobytes()


def lzw_decode(data, size):
    load_library()
    a = np.frombuffer(bytes(data), np.uint8).copy()
    out = np.empty(size, np.uint8)
    k = _lib.mkde_lzw_decode(_p(a, _U8P), len(a), _p(out, _U8P), size)
    if k < 0:
        raise MkdeError(5, "corrupt LZW data")
    return out[:k].tobytes()


def declared_symbols(header=HEADER_PATH):
    """Names of the functions declared in include/mkde.h."""
    import re

    txt = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int|int64_t)\s*(mkde_\w+)\s*\(", txt, re.M)))


def version():
    load_library()
    return _lib.mkde_version().decode()
