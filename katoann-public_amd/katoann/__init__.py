"""KatoANN on AMD Instinct MI355X (gfx950): the Python module of the reference
(python/katoann_module.cpp) over the library's C ABI (include/katoann_c.h).

    import katoann
    katoann.build(data, "idx/base", nlist=2048)
    ix = katoann.Index(); ix.load("idx/base")
    ids, dists = ix.search(queries, k=10, nprobe=64)

Same names, defaults, argument meaning and error classes as the reference
module: ValueError for std::invalid_argument, RuntimeError otherwise. The
search runs on the GPU; there is no CPU fallback, and importing fails loudly
when libkatoann.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import builder, formats  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libkatoann.so")

RANKING = {"full": 0, "high": 1, "prefix": 2}
SELECTION = {"threshold": 0, "local": 1}
NSTAGES = 7
STAGES = ("rotate+quantize", "coarse", "select", "plan", "rank", "update", "refine")


class _BuildOptions(C.Structure):
    _fields_ = [("nlist", C.c_uint32), ("train_size", C.c_uint32), ("kmeans_iters", C.c_uint32),
                ("seed", C.c_uint64), ("normalized", C.c_int32), ("rotation_sample", C.c_uint32),
                ("jacobi_sweeps", C.c_uint32)]


class _LoadOptions(C.Structure):
    _fields_ = [("precision", C.c_int32), ("use_amx", C.c_int32), ("prefix_dims", C.c_uint32),
                ("batched_coarse", C.c_int32)]


class _SearchParams(C.Structure):
    _fields_ = [("k", C.c_uint32), ("nprobe", C.c_uint32), ("ranking", C.c_int32),
                ("survivors", C.c_uint32), ("selection", C.c_int32)]


class NprobeModel(C.Structure):
    """katoann_nprobe_model: a trained learned-nprobe MLP (device pointers)."""
    _fields_ = [("nfeat", C.c_uint32), ("hidden", C.c_uint32), ("feat_mean", C.c_void_p),
                ("feat_inv_std", C.c_void_p), ("w1", C.c_void_p), ("b1", C.c_void_p), ("w2", C.c_void_p),
                ("b2", C.c_void_p), ("w3", C.c_void_p), ("b3", C.c_void_p), ("y_mean", C.c_float),
                ("y_std", C.c_float), ("margin", C.c_float), ("nmin", C.c_uint32)]


class _Extras(C.Structure):
    _fields_ = [("nprobe_per_query", C.c_void_p), ("normalize", C.c_int32), ("device_io", C.c_int32),
                ("probes", C.c_void_p), ("probe_dists", C.c_void_p), ("survivors", C.c_void_p),
                ("n_survivors", C.c_void_p), ("rotated", C.c_void_p), ("codes", C.c_void_p),
                ("stream", C.c_void_p), ("nprobe_model", C.POINTER(NprobeModel)), ("nprobe_used", C.c_void_p),
                ("coarse_only", C.c_int32), ("max_bytes", C.c_uint64), ("needed_bytes", C.c_void_p)]


class ServerOptions(C.Structure):
    """katoann_server_options: batches in flight, queries per batch, HBM budget (0: 90% of free)."""
    _fields_ = [("max_inflight", C.c_uint32), ("max_batch", C.c_uint32), ("hbm_budget", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("queries", C.c_uint64), ("batches", C.c_uint64),
                ("p50_ms", C.c_double), ("p95_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("mean_batch", C.c_double), ("achieved_qps", C.c_double), ("budget", C.c_uint64),
                ("peak_reserved", C.c_uint64), ("peak_inflight", C.c_uint32), ("peak_used", C.c_uint64)]


_lib = None


def lib():
    """The loaded libkatoann.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libkatoann.so is not built: run python katoann-public_amd/build.py")
        # One HIP runtime per process. torch ships its own libamdhip64 and asks
        # for it as "libamdhip64.so", which does not match the "libamdhip64.so.7"
        # this library binds; loaded in the other order, the process holds two
        # runtimes and torch's finds no GPU. Loading torch first lets this
        # library bind to torch's runtime (same soname).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        P, u32, u64, i32, f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_double
        L.katoann_last_error.restype = C.c_char_p
        L.katoann_index_create.argtypes = [C.POINTER(C.c_void_p)]
        L.katoann_index_destroy.argtypes = [P]
        L.katoann_load.argtypes = [P, C.c_char_p, C.POINTER(_LoadOptions)]
        L.katoann_load_arrays.argtypes = [P, u32, u32, u32, i32, P, P, P, P, i32, P, u32, C.POINTER(_LoadOptions)]
        L.katoann_search.argtypes = [P, P, u32, C.POINTER(_SearchParams), P, P]
        L.katoann_search_ex.argtypes = [P, P, u32, C.POINTER(_SearchParams), C.POINTER(_Extras), P, P]
        L.katoann_search_noalloc.argtypes = [P, P, u32, C.POINTER(_SearchParams), C.POINTER(_Extras), P, P]
        for name, rt in [("dim", u32), ("nlist", u32), ("size", u32), ("normalized", i32), ("engine", i32),
                         ("precision", i32), ("prefix_dims", u32), ("rotated", i32), ("batched_coarse", i32),
                         ("memory_bytes", u64)]:
            f = getattr(L, "katoann_" + name)
            f.restype = rt
            f.argtypes = [P]
        L.katoann_list_sizes.argtypes = [P, P]
        L.katoann_index_scale.argtypes = [P, P, P, P]
        L.katoann_export_codes.argtypes = [P, P, P, P]
        L.katoann_export_lists.argtypes = [P, P, u32, P, P]
        L.katoann_set_timing.argtypes = [P, i32]
        L.katoann_stage_times.argtypes = [P, P, i32]
        L.katoann_last_work.argtypes = [P, P, i32]
        L.katoann_band_times.argtypes = [P, P, i32]
        L.katoann_set_rank_rows.argtypes = [P, u32]
        L.katoann_set_band_factor.argtypes = [P, u32]
        L.katoann_set_seed_rows.argtypes = [P, u32]
        L.katoann_set_seed_join.argtypes = [P, i32]
        L.katoann_set_pool_allowance.argtypes = [P, u32]
        L.katoann_set_rank_kernel.argtypes = [P, u32, u32, u32]
        L.katoann_set_graphs.argtypes = [P, i32]
        L.katoann_set_load_chunk.argtypes = [u64]
        L.katoann_set_autotune.argtypes = [P, i32]
        L.katoann_tile_config.argtypes = [P, u32, C.POINTER(_SearchParams), P]
        L.katoann_sq16_l2.argtypes = [P, P, P, P, P, P, f64, u32, u32, u32, P]
        L.katoann_sq16_high_l2.argtypes = [P, P, P, P, f64, u32, u32, u32, P, u32]
        L.katoann_normalize_rows.argtypes = [P, u64, u32]
        L.katoann_rotate_rows.argtypes = [P, P, u64, u32, P]
        L.katoann_build.argtypes = [P, u32, u32, C.POINTER(_BuildOptions), C.c_char_p]
        L.katoann_build_device.argtypes = [P, u32, u32, C.POINTER(_BuildOptions), P, P, P, P, C.POINTER(i32)]
        L.katoann_assign_lists.argtypes = [P, u32, u32, P, u32, P]
        L.katoann_kmeans_device.argtypes = [P, u32, u32, P, u32, u32]
        L.katoann_pca_rotation_device.argtypes = [P, u32, u32, P, P, C.POINTER(i32)]
        sz, f32 = C.c_size_t, C.c_float
        L.katoann_mlp_gemm.argtypes = [u32, u32, u32, P, sz, sz, P, sz, sz, P, P, sz, i32, P, P, sz, P]
        L.katoann_mlp_to_bf16.argtypes = [P, P, u64, P]
        L.katoann_mlp_mse.argtypes = [P, P, u32, P, P, P]
        L.katoann_mlp_colsum.argtypes = [P, u32, u32, P, P]
        L.katoann_mlp_pinball.argtypes = [P, P, u32, f32, P, P, P]
        L.katoann_mlp_adam.argtypes = [P, P, P, P, u32, f32, f32, f32, f32, u32, P, P]
        L.katoann_mlp_features.argtypes = [P, u32, u32, u32, P, P, P, P, P]
        L.katoann_search_footprint.argtypes = [P, u32, C.POINTER(_SearchParams), C.POINTER(_Extras),
                                               C.POINTER(u64)]
        L.katoann_default_server_options.argtypes = [C.POINTER(ServerOptions)]
        L.katoann_server_create.argtypes = [P, C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]
        L.katoann_server_destroy.argtypes = [P]
        L.katoann_server_submit.argtypes = [P, P, u32, C.POINTER(_SearchParams), C.POINTER(NprobeModel), i32,
                                            C.POINTER(u64)]
        L.katoann_server_poll.argtypes = [P, u64, C.POINTER(i32)]
        L.katoann_server_wait.argtypes = [P, u64, P, P, C.POINTER(f64)]
        L.katoann_server_drain.argtypes = [P]
        L.katoann_server_stats_get.argtypes = [P, C.POINTER(ServerStats)]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = lib().katoann_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(x, what="data must be a 2-D float32 array"):
    a = np.ascontiguousarray(x, dtype=np.float32)
    if a.ndim != 2:
        raise ValueError(what)
    return a


def set_load_chunk(nbytes: int) -> None:
    """Size of each of the two pinned staging buffers through which index
    loads stream the .invlist vectors to the GPU (default 64 MiB)."""
    _check(lib().katoann_set_load_chunk(int(nbytes)))


def amx_available() -> bool:
    """The reference reports the CPU tile engine; this engine is the GPU."""
    return False


def build(data, prefix: str, nlist: int = 2048, train_size: int = 256000, kmeans_iters: int = 12,
          seed: int = 0, normalized: bool = False, rotation_sample: int = 100000) -> None:
    """Partition data (n x dim float32) on the GPU and write the index files
    under prefix (python/katoann_module.cpp:39-58). The base data is not
    normalised here."""
    data = _f32(data)
    o = _BuildOptions(nlist, train_size, kmeans_iters, seed, int(normalized), rotation_sample, 8)
    _check(lib().katoann_build(_ptr(data), data.shape[0], data.shape[1], C.byref(o), prefix.encode()))


def build_device(data_ptr: int, n: int, dim: int, centroids_ptr: int, order_ptr: int, sizes_ptr: int,
                 rotation_ptr: int | None, nlist: int, train_size: int = 256000, kmeans_iters: int = 12,
                 seed: int = 0, rotation_sample: int = 100000) -> bool:
    """katoann_build_device on device buffers (addresses): data (n x dim f32)
    -> centroids (nlist x dim f32), order (n u32), sizes (nlist u32), rotation
    (dim x dim f32, or None to skip). Returns whether a rotation was computed."""
    o = _BuildOptions(nlist, train_size, kmeans_iters, seed, 0, rotation_sample, 8)
    r = C.c_int32(0)
    _check(lib().katoann_build_device(data_ptr, n, dim, C.byref(o), centroids_ptr, order_ptr, sizes_ptr,
                                      rotation_ptr, C.byref(r)))
    return bool(r.value)


def assign_lists(x_ptr: int, n: int, dim: int, centroids_ptr: int, nlist: int, out_ptr: int) -> None:
    """katoann_assign_lists on device buffers: nearest of nlist centroids for
    each of n vectors, ties to the lowest index, into out (n u32)."""
    _check(lib().katoann_assign_lists(x_ptr, n, dim, centroids_ptr, nlist, out_ptr))


def kmeans_device(train_ptr: int, n: int, dim: int, centroids_ptr: int, nlist: int, iters: int) -> None:
    """katoann_kmeans_device: Lloyd iterations from the centroids in place."""
    _check(lib().katoann_kmeans_device(train_ptr, n, dim, centroids_ptr, nlist, iters))


def pca_rotation_device(sample_ptr: int, m: int, dim: int, rotation_ptr: int):
    """katoann_pca_rotation_device: returns (rotated, eigenvalues decreasing)."""
    eig = np.zeros(dim, np.float64)
    r = C.c_int32(0)
    _check(lib().katoann_pca_rotation_device(sample_ptr, m, dim, rotation_ptr, _ptr(eig), C.byref(r)))
    return bool(r.value), eig


class Index:
    """katoann.Index (python/katoann_module.cpp:60-128)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().katoann_index_create(C.byref(h)))
        self._h = h

    def close(self):
        """Frees the index's device memory now (the object is unusable after)."""
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.katoann_index_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    # ---- reference API
    def load(self, prefix: str, precision: str = "sq16", use_amx: bool = True, prefix_dims: int = 0,
             batched_coarse: bool = True) -> None:
        if precision == "bf16":
            prec = 1
        elif precision == "sq16":
            prec = 0
        else:
            raise ValueError("precision is 'sq16' or 'bf16'")
        o = _LoadOptions(prec, 1 if use_amx else 0, prefix_dims, 1 if batched_coarse else 0)
        _check(lib().katoann_load(self._h, prefix.encode(), C.byref(o)))

    def search(self, queries, k: int = 10, nprobe: int = 64, ranking: str = "prefix", survivors: int = 32,
               selection: str = "threshold"):
        """Search all queries together; returns (ids, squared L2 distances), nearest first.
        Queries are normalized when the index was built with normalized=True."""
        return self.search_ex(queries, k=k, nprobe=nprobe, ranking=ranking, survivors=survivors,
                              selection=selection)[:2]

    @property
    def dim(self):
        return lib().katoann_dim(self._h)

    @property
    def nlist(self):
        return lib().katoann_nlist(self._h)

    @property
    def size(self):
        return lib().katoann_size(self._h)

    @property
    def normalized(self):
        return bool(lib().katoann_normalized(self._h))

    @property
    def prefix_dims(self):
        return lib().katoann_prefix_dims(self._h)

    @property
    def engine(self):
        return {0: "amx", 1: "avx512", 2: "gfx950"}[lib().katoann_engine(self._h)]

    @property
    def memory_bytes(self):
        return lib().katoann_memory_bytes(self._h)

    # ---- extensions
    @property
    def rotated(self):
        return bool(lib().katoann_rotated(self._h))

    @property
    def batched_coarse(self):
        """False when loaded with batched_coarse=False (per-query fp32 coarse search)."""
        return bool(lib().katoann_batched_coarse(self._h))

    def list_sizes(self):
        out = np.zeros(self.nlist, np.uint32)
        _check(lib().katoann_list_sizes(self._h, _ptr(out)))
        return out

    def _params(self, k, nprobe, ranking, survivors, selection):
        if ranking not in RANKING:
            raise ValueError("ranking is 'prefix', 'high' or 'full'")
        if selection not in SELECTION:
            raise ValueError("selection is 'threshold' or 'local'")
        return _SearchParams(k, nprobe, RANKING[ranking], survivors, SELECTION[selection])

    def search_ex(self, queries, k=10, nprobe=64, ranking="prefix", survivors=32, selection="threshold",
                  nprobe_per_query=None, normalize=None, want=()):
        """search() with optional per-query probe counts and extra outputs.
        want: any of "probes", "probe_dists", "survivors", "rotated", "codes"; the extras dict
        also holds "device_bytes", the bytes of the search's device buffers."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim != 2 or q.shape[1] != self.dim:
            raise ValueError("queries must be (n, dim) float32")
        nq = q.shape[0]
        p = self._params(k, nprobe, ranking, survivors, selection)
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        x = _Extras()
        x.normalize = int(self.normalized if normalize is None else normalize)
        extra = {}
        S = k if ranking == "full" else max(survivors, k)
        npq = None
        if nprobe_per_query is not None:
            npq = np.ascontiguousarray(nprobe_per_query, dtype=np.uint32)
            x.nprobe_per_query = _ptr(npq)
        if "probes" in want:
            extra["probes"] = np.empty((nq, nprobe), np.uint32)
            x.probes = _ptr(extra["probes"])
        if "probe_dists" in want:
            extra["probe_dists"] = np.empty((nq, nprobe), np.float32)
            x.probe_dists = _ptr(extra["probe_dists"])
        if "survivors" in want:
            extra["survivors"] = np.empty((nq, S), np.uint64)
            extra["n_survivors"] = np.empty(nq, np.uint32)
            x.survivors = _ptr(extra["survivors"])
            x.n_survivors = _ptr(extra["n_survivors"])
        if "rotated" in want:
            extra["rotated"] = np.empty((nq, self.dim), np.float32)
            x.rotated = _ptr(extra["rotated"])
        if "codes" in want:
            dc = (self.dim + 63) // 64 * 64
            extra["codes"] = np.empty((nq, dc), np.uint16)
            x.codes = _ptr(extra["codes"])
        used = C.c_uint64()
        x.needed_bytes = C.cast(C.byref(used), C.c_void_p)
        _check(lib().katoann_search_ex(self._h, _ptr(q), nq, C.byref(p), C.byref(x), _ptr(ids), _ptr(dists)))
        extra["device_bytes"] = used.value  # the search's device buffers
        return ids, dists, extra

    def search_device(self, q_ptr: int, nq: int, ids_ptr: int, dists_ptr: int, k=10, nprobe=64,
                      ranking="prefix", survivors=32, selection="threshold", npq_ptr: int | None = None,
                      normalize=False, stream: int | None = None, model: NprobeModel | None = None,
                      npq_out_ptr: int | None = None, probes_ptr: int | None = None,
                      probe_dists_ptr: int | None = None, coarse_only: bool = False):
        """Search with queries and outputs in device memory (e.g. torch tensors' data_ptr()).
        model: a learned-nprobe model (per-query nprobe in [model.nmin, nprobe]);
        coarse_only: stop after the probe selection (probes / probe dists only)."""
        p = self._params(k, nprobe, ranking, survivors, selection)
        x = _Extras()
        x.device_io = 1
        x.normalize = int(normalize)
        x.nprobe_per_query = npq_ptr
        x.stream = stream
        x.nprobe_used = npq_out_ptr
        x.probes = probes_ptr
        x.probe_dists = probe_dists_ptr
        x.coarse_only = int(coarse_only)
        if model is not None:
            x.nprobe_model = C.pointer(model)
        _check(lib().katoann_search_ex(self._h, C.c_void_p(q_ptr), nq, C.byref(p), C.byref(x),
                                       C.c_void_p(ids_ptr or 0), C.c_void_p(dists_ptr or 0)))

    def search_noalloc(self, q_ptr: int, nq: int, ids_ptr: int, dists_ptr: int, k=10, nprobe=64,
                       ranking="prefix", survivors=32, selection="threshold", npq_ptr: int | None = None,
                       normalize=False, stream: int | None = None):
        """search_device() through the allocation-free entry point: the search runs in
        the library's static per-GPU scratch (KATOANN_NOALLOC_SCRATCH_BYTES) and
        allocates no device memory."""
        p = self._params(k, nprobe, ranking, survivors, selection)
        x = _Extras()
        x.device_io = 1
        x.normalize = int(normalize)
        x.nprobe_per_query = npq_ptr
        x.stream = stream
        _check(lib().katoann_search_noalloc(self._h, C.c_void_p(q_ptr), nq, C.byref(p), C.byref(x),
                                            C.c_void_p(ids_ptr or 0), C.c_void_p(dists_ptr or 0)))

    def load_arrays(self, dim, nlist, n, normalized, centroids, list_sizes, ids, vectors, rotation,
                    prefix_dims, vectors_device_ptr: int | None = None):
        """Load from arrays (vectors grouped by list, list order). With
        vectors_device_ptr the vectors are read from device memory."""
        cen = np.ascontiguousarray(centroids, np.float32)
        sz = np.ascontiguousarray(list_sizes, np.uint32)
        idv = np.ascontiguousarray(ids, np.uint32)
        rot = None if rotation is None else np.ascontiguousarray(rotation, np.float32)
        o = _LoadOptions(0, 1, 0, 1)
        if vectors_device_ptr is not None:
            vp, on_dev = C.c_void_p(vectors_device_ptr), 1
        else:
            vec = np.ascontiguousarray(vectors, np.float32)
            vp, on_dev = _ptr(vec), 0
        _check(lib().katoann_load_arrays(self._h, dim, nlist, n, int(normalized), _ptr(cen), _ptr(sz), _ptr(idv),
                                         vp, on_dev, _ptr(rot), prefix_dims, C.byref(o)))

    def scale(self):
        lo, inv, ssq = C.c_float(), C.c_double(), C.c_double()
        _check(lib().katoann_index_scale(self._h, C.byref(lo), C.byref(inv), C.byref(ssq)))
        return lo.value, inv.value, ssq.value

    def export_codes(self):
        sizes = self.list_sizes()
        n = int(sizes.sum())
        dc = (self.dim + 63) // 64 * 64
        codes = np.empty((n, dc), np.uint16)
        cc = np.empty((self.nlist, dc), np.uint16)
        ids = np.empty(n, np.uint32)
        _check(lib().katoann_export_codes(self._h, _ptr(codes), _ptr(cc), _ptr(ids)))
        return codes, cc, ids

    def centroid_codes(self):
        dc = (self.dim + 63) // 64 * 64
        cc = np.empty((self.nlist, dc), np.uint16)
        _check(lib().katoann_export_codes(self._h, None, _ptr(cc), None))
        return cc

    def export_lists(self, lists):
        """Codes (rows x Dc u16) and ids of the rows of the given lists, list
        after list in the order given."""
        lists = np.ascontiguousarray(lists, np.uint32)
        sizes = self.list_sizes()
        n = int(sizes[lists].astype(np.int64).sum())
        dc = (self.dim + 63) // 64 * 64
        codes = np.empty((n, dc), np.uint16)
        ids = np.empty(n, np.uint32)
        _check(lib().katoann_export_lists(self._h, _ptr(lists), lists.size, _ptr(codes), _ptr(ids)))
        return codes, ids

    def set_timing(self, on: bool = True):
        _check(lib().katoann_set_timing(self._h, int(on)))

    def stage_times(self):
        out = np.zeros(NSTAGES, np.float32)
        _check(lib().katoann_stage_times(self._h, _ptr(out), NSTAGES))
        return dict(zip(STAGES, out.tolist()))

    def band_times(self):
        """Ranking-launch ms of each launched probe-rank band of the last
        timed search (a joined seed remainder's band is not launched)."""
        out = np.zeros(16, np.float32)
        _check(lib().katoann_band_times(self._h, _ptr(out), 16))
        return [float(t) for t in out if t > 0]

    def last_work(self):
        out = np.zeros(4, np.uint64)
        _check(lib().katoann_last_work(self._h, _ptr(out), 4))
        return dict(pairs=int(out[0]), items=int(out[1]), kept=int(out[2]), bands=int(out[3]))

    def set_rank_rows(self, rows: int):
        _check(lib().katoann_set_rank_rows(self._h, rows))

    def set_band_factor(self, log2_factor: int):
        """Probe-rank bands of the ranking pass grow by 2**log2_factor (1..12)."""
        _check(lib().katoann_set_band_factor(self._h, log2_factor))

    def set_seed_rows(self, rows: int):
        """Seed band of the ranking pass: each query's nearest lists up to
        `rows` rows (at most `rows` of each) before the probe-rank bands, so
        every query has a threshold early; 0 = no seed band."""
        _check(lib().katoann_set_seed_rows(self._h, rows))

    def set_seed_join(self, on: bool = True):
        """With a seed band and nprobe <= nlist / 16: rank the rest of each
        query's seed lists in the first probe-rank band, with that band's
        queries of the same lists (rows read once, one pool update less)."""
        _check(lib().katoann_set_seed_join(self._h, int(on)))

    def set_pool_allowance(self, slots: int):
        """Survivor pool room per query beyond S (0 = sized per search); a band
        that keeps more reruns the batch with 4x the room."""
        _check(lib().katoann_set_pool_allowance(self._h, slots))

    def set_rank_kernel(self, kernel: int, variant: int = 0, queries_per_item: int = 1024):
        """Ranking kernel: 0 = block items (queries staged into LDS), 1 = wave
        items (rows held in registers, queries streamed), 2 = block items with a
        shared query ring, 3 = query-stationary wave items (up to 64 of a list's
        queries held in registers, 32 with variant 1, rows streamed;
        queries_per_item is then the rows per item); variant 0/1 = larger or
        smaller row tile per wave, 2 (kernel 1) = query codes loaded two
        fragments ahead, 3 (kernel 1) = larger row tile at one wave per SIMD."""
        _check(lib().katoann_set_rank_kernel(self._h, kernel, variant, queries_per_item))

    def set_autotune(self, on: bool = True):
        """Time the candidate ranking tiles on the first search of each shape
        and keep the fastest (results do not depend on the tile)."""
        _check(lib().katoann_set_autotune(self._h, int(on)))

    def tile_config(self, nq, k=10, nprobe=64, ranking="prefix", survivors=32, selection="threshold"):
        """The tile (kernel, variant, queries per item, log2 band factor, seed
        rows, joined seed remainder) a search of this shape runs with."""
        out = np.zeros(6, np.uint32)
        prm = self._params(k, nprobe, ranking, survivors, selection)
        _check(lib().katoann_tile_config(self._h, nq, C.byref(prm), _ptr(out)))
        return dict(kernel=int(out[0]), variant=int(out[1]), queries_per_item=int(out[2]), log2_band_factor=int(out[3]),
                    seed_rows=int(out[4]), seed_join=int(out[5]))

    def set_graphs(self, on: bool = True):
        """Replay each search's kernels from a captured hipGraph (default on)."""
        _check(lib().katoann_set_graphs(self._h, int(on)))


# ---- kernel-level entry points (reference include/katoann/kernels.h:34-52)

def sq16_l2(qhi, qlo, xhi, xlo, qnorm, xnorm, step_sq):
    qhi, qlo, xhi, xlo = (np.ascontiguousarray(a, np.uint8) for a in (qhi, qlo, xhi, xlo))
    qn, xn = np.ascontiguousarray(qnorm, np.int64), np.ascontiguousarray(xnorm, np.int64)
    m, dc = qhi.shape
    n = xhi.shape[0]
    out = np.empty((m, n), np.float32)
    _check(lib().katoann_sq16_l2(_ptr(qhi), _ptr(qlo), _ptr(xhi), _ptr(xlo), _ptr(qn), _ptr(xn), step_sq, m, n,
                                 dc, _ptr(out)))
    return out


def sq16_high_l2(qhi, xhi, qnorm, xnorm, step_sq, padded_dim=None):
    """qhi: m x padded_dim; xhi: n x plane_dim (the leading padded_dim bytes are read)."""
    qhi, xhi = np.ascontiguousarray(qhi, np.uint8), np.ascontiguousarray(xhi, np.uint8)
    qn, xn = np.ascontiguousarray(qnorm, np.int64), np.ascontiguousarray(xnorm, np.int64)
    m, pd = qhi.shape
    n, plane = xhi.shape
    out = np.empty((m, n), np.float32)
    _check(lib().katoann_sq16_high_l2(_ptr(qhi), _ptr(xhi), _ptr(qn), _ptr(xn), step_sq, m, n, pd, _ptr(out), plane))
    return out


def normalize_rows(x):
    y = np.array(x, dtype=np.float32, copy=True, order="C")
    _check(lib().katoann_normalize_rows(_ptr(y), y.shape[0], y.shape[1]))
    return y


def rotate_rows(x, rotation):
    x = np.ascontiguousarray(x, np.float32)
    r = np.ascontiguousarray(rotation, np.float32)
    out = np.empty_like(x)
    _check(lib().katoann_rotate_rows(_ptr(x), _ptr(out), x.shape[0], x.shape[1], _ptr(r)))
    return out


class Server:
    """Continuous-batching request loop over a loaded Index (no reference
    counterpart; replaces the CLI's synchronous `load` batching,
    tools/katoann.cpp:267-337). submit() queues host queries and returns a
    ticket at once; wait(ticket) returns (ids, dists, latency_ms)."""

    def __init__(self, index: Index, max_inflight: int = 4, max_batch: int = 4096, hbm_budget: int = 0):
        o = ServerOptions(max_inflight, max_batch, hbm_budget)
        h = C.c_void_p()
        _check(lib().katoann_server_create(index._h, C.byref(o), C.byref(h)))
        self._h, self._index, self._k = h, index, {}

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.katoann_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, queries, k=10, nprobe=64, ranking="prefix", survivors=32, selection="threshold",
               model: NprobeModel | None = None, normalize=None) -> int:
        q = np.ascontiguousarray(queries, dtype=np.float32)
        if q.ndim != 2 or q.shape[1] != self._index.dim:
            raise ValueError("queries must be (n, dim) float32")
        p = self._index._params(k, nprobe, ranking, survivors, selection)
        t = C.c_uint64()
        norm = int(self._index.normalized if normalize is None else normalize)
        _check(lib().katoann_server_submit(self._h, _ptr(q), q.shape[0], C.byref(p),
                                           C.byref(model) if model is not None else None, norm, C.byref(t)))
        self._k[t.value] = (q.shape[0], k)
        return t.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int32()
        _check(lib().katoann_server_poll(self._h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int):
        nq, k = self._k.pop(ticket)
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        lat = C.c_double()
        _check(lib().katoann_server_wait(self._h, ticket, _ptr(ids), _ptr(dists), C.byref(lat)))
        return ids, dists, lat.value

    def drain(self):
        _check(lib().katoann_server_drain(self._h))

    def stats(self) -> dict:
        s = ServerStats()
        _check(lib().katoann_server_stats_get(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def footprint(self, nq, k=10, nprobe=64, ranking="prefix", survivors=32, selection="threshold",
                  model: NprobeModel | None = None) -> int:
        """Device bytes one search of nq queries with these parameters allocates."""
        p = self._index._params(k, nprobe, ranking, survivors, selection)
        x = _Extras()
        if model is not None:
            x.nprobe_model = C.pointer(model)
        b = C.c_uint64()
        _check(lib().katoann_search_footprint(self._index._h, nq, C.byref(p), C.byref(x), C.byref(b)))
        return b.value
