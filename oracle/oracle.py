"""ctypes wrapper of the CPU oracle (oracle/liboracle.so). Test infrastructure:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.

Also reads the dumps of the reference driver (oracle/_ref/katoann_ref_search).
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SEARCH = os.path.join(REF_DIR, "katoann_ref_search")
REF_CLI = os.path.join(REF_DIR, "katoann")

_lib = None


def build(force: bool = False) -> None:
    """Compiles liboracle.so (and the reference under _ref/ when its sources exist)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args + ["liboracle.so"], check=True, stdout=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(args + ["ref"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u32, u64, i32, f64, f32 = C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_float
        L.oracle_index_from_raw.restype = P
        L.oracle_index_from_raw.argtypes = [u32, u32, u32, P, P, P, P, P, u32]
        L.oracle_index_from_codes.restype = P
        L.oracle_index_from_codes.argtypes = [u32, u32, u32, P, P, P, P, P, f32, f64, f64, u32]
        L.oracle_index_free.argtypes = [P]
        L.oracle_index_scale.argtypes = [P, P, P, P]
        L.oracle_index_rows.restype = u64
        L.oracle_index_rows.argtypes = [P]
        L.oracle_index_export.argtypes = [P, P, P, P, P, P]
        L.oracle_queries.argtypes = [P, P, u32, P, P]
        L.oracle_search.restype = i32
        L.oracle_search.argtypes = [P, P, u32, u32, u32, u32, i32, i32, P, P, P, P, P, P, P, P]
        L.oracle_rotate_rows.argtypes = [P, P, u64, u32, P]
        L.oracle_sq16_l2.argtypes = [P, P, P, P, P, P, f64, u32, u32, u32, P]
        L.oracle_sq16_high_l2.argtypes = [P, u32, P, u32, P, P, f64, u32, u32, u32, P]
        L.oracle_normalize_rows.argtypes = [P, u64, u32]
        L.oracle_index_set_flat.restype = i32
        L.oracle_index_set_flat.argtypes = [P, i32]
        L.oracle_index_set_bf16.restype = i32
        L.oracle_index_set_bf16.argtypes = [P, i32]
        L.oracle_flat_dists.restype = i32
        L.oracle_flat_dists.argtypes = [P, P, u32, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def rotate_rows(x: np.ndarray, rot: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().oracle_rotate_rows(_p(x), _p(out), x.shape[0], x.shape[1], _p(np.ascontiguousarray(rot, np.float32)))
    return out


def normalize_rows(x: np.ndarray) -> np.ndarray:
    y = np.array(x, dtype=np.float32, copy=True, order="C")
    lib().oracle_normalize_rows(_p(y), y.shape[0], y.shape[1])
    return y


def sq16_l2(qh, ql, xh, xl, qn, xn, step_sq):
    m, dc = qh.shape
    n = xh.shape[0]
    out = np.empty((m, n), np.float32)
    lib().oracle_sq16_l2(_p(qh), _p(ql), _p(xh), _p(xl), _p(qn), _p(xn), step_sq, m, n, dc, _p(out))
    return out


def sq16_high_l2(qh, xh, qn, xn, step_sq, pd):
    m, qs = qh.shape
    n, plane = xh.shape
    out = np.empty((m, n), np.float32)
    lib().oracle_sq16_high_l2(_p(qh), qs, _p(xh), plane, _p(qn), _p(xn), step_sq, m, n, pd, _p(out))
    return out


RANKING = {"full": 0, "high": 1, "prefix": 2}


class OracleIndex:
    """The reference load + search restated on the CPU."""

    def __init__(self, handle, d, nlist, npts, prefix_dims, keep=()):
        self.h = handle
        self.d, self.nlist, self.npts, self.prefix_dims = d, nlist, npts, prefix_dims
        self.dc = (d + 63) // 64 * 64
        self._keep = keep

    @classmethod
    def from_raw(cls, d, nlist, npts, sizes, ids, vecs, centroids, rotation, prefix_dims):
        sizes = np.ascontiguousarray(sizes, np.uint32)
        ids = np.ascontiguousarray(ids, np.uint32)
        vecs = np.ascontiguousarray(vecs, np.float32)
        cen = np.ascontiguousarray(centroids, np.float32)
        rot = None if rotation is None else np.ascontiguousarray(rotation, np.float32)
        h = lib().oracle_index_from_raw(d, nlist, npts, _p(sizes), _p(ids), _p(vecs), _p(cen), _p(rot),
                                        prefix_dims)
        return cls(h, d, nlist, npts, prefix_dims)

    @classmethod
    def from_files(cls, prefix, prefix_dims=0):
        from katoann import formats
        ix = formats.read_index(prefix)
        pd = prefix_dims or ix["prefix_dims"]
        return cls.from_raw(ix["dim"], ix["nlist"], ix["n"], ix["sizes"], ix["ids"], ix["vecs"],
                            ix["centroids"], ix["rotation"], pd)

    @classmethod
    def from_codes(cls, d, nlist, npts, sizes, ids, codes, ccodes, rotation, lo, inv_step, step_sq,
                   prefix_dims):
        sizes = np.ascontiguousarray(sizes, np.uint32)
        ids = np.ascontiguousarray(ids, np.uint32)
        codes = np.ascontiguousarray(codes, np.uint16)
        ccodes = np.ascontiguousarray(ccodes, np.uint16)
        rot = None if rotation is None else np.ascontiguousarray(rotation, np.float32)
        h = lib().oracle_index_from_codes(d, nlist, npts, _p(sizes), _p(ids), _p(codes), _p(ccodes),
                                          _p(rot), lo, inv_step, step_sq, prefix_dims)
        return cls(h, d, nlist, npts, prefix_dims)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_index_free(self.h)
            self.h = None

    def set_flat(self, flat=True):
        """Per-query fp32 coarse search (the reference's batched_coarse = false)."""
        if lib().oracle_index_set_flat(self.h, 1 if flat else 0) != 0:
            raise ValueError("the per-query coarse search needs the float centroids")
        return self

    def set_bf16(self, on=True):
        """bf16 precision (the reference's LoadOptions::precision = kBf16):
        full ranking over bf16 rows after the per-query coarse search."""
        if lib().oracle_index_set_bf16(self.h, 1 if on else 0) != 0:
            raise ValueError("bf16 precision needs an index from raw vectors")
        return self

    def flat_dists(self, q):
        """Per-query coarse distances of every list, nq x nlist (unrotated queries)."""
        q = np.ascontiguousarray(q, np.float32)
        out = np.empty((q.shape[0], self.nlist), np.float32)
        if lib().oracle_flat_dists(self.h, _p(q), q.shape[0], _p(out)) != 0:
            raise ValueError("the per-query coarse search needs the float centroids")
        return out

    def scale(self):
        lo, inv, ssq = C.c_float(), C.c_double(), C.c_double()
        lib().oracle_index_scale(self.h, C.byref(lo), C.byref(inv), C.byref(ssq))
        return lo.value, inv.value, ssq.value

    def export(self):
        n = lib().oracle_index_rows(self.h)
        codes = np.empty((n, self.dc), np.uint16)
        ccodes = np.empty((self.nlist, self.dc), np.uint16)
        norm = np.empty(n, np.int64)
        head = np.empty(n, np.int64)
        cnorm = np.empty(self.nlist, np.int64)
        lib().oracle_index_export(self.h, _p(codes), _p(ccodes), _p(norm), _p(head), _p(cnorm))
        return dict(codes=codes, ccodes=ccodes, norm=norm, headnorm=head, cnorm=cnorm)

    def queries(self, q):
        q = np.ascontiguousarray(q, np.float32)
        qrot = np.empty((q.shape[0], self.d), np.float32)
        codes = np.empty((q.shape[0], self.dc), np.uint16)
        lib().oracle_queries(self.h, _p(q), q.shape[0], _p(qrot), _p(codes))
        return qrot, codes

    def search(self, q, k=10, nprobe=64, ranking="prefix", survivors=32, libstdc_ties=False,
               nprobe_per_query=None):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        S = max(survivors, k)
        ids = np.empty((nq, k), np.uint32)
        dists = np.empty((nq, k), np.float32)
        probes = np.empty((nq, nprobe), np.uint32)
        pdist = np.empty((nq, nprobe), np.float32)
        surv = np.empty((nq, S), np.uint64)
        nsurv = np.empty(nq, np.uint32)
        flags = np.empty(nq, np.uint8)
        npq = None if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.uint32)
        rc = lib().oracle_search(self.h, _p(q), nq, k, nprobe, survivors, RANKING[ranking],
                                 1 if libstdc_ties else 0, _p(npq), _p(ids), _p(dists), _p(probes),
                                 _p(pdist), _p(surv), _p(nsurv), _p(flags))
        if rc != 0:
            raise ValueError("oracle search: bad arguments")
        return dict(ids=ids, dists=dists, probes=probes, probe_dists=pdist, surv=surv, nsurv=nsurv,
                    flags=flags)


# ------------------------------------------------------------ reference dumps

def read_ref_dump(path: str):
    """Parses the `katoann_ref_search dump` output."""
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    def take(n):
        nonlocal off
        b = buf[off:off + n]
        off += n
        return b
    hdr = struct.unpack("<8I", take(32))
    d, dc, nlist, npts, head, rotated, batched, _ = hdr
    lo = struct.unpack("<f", take(4))[0]
    inv_step, step, step_sq = struct.unpack("<3d", take(24))
    ids, hi, lo_, norm, hnorm, sizes = [], [], [], [], [], []
    for _ in range(nlist):
        (n,) = struct.unpack("<I", take(4))
        sizes.append(n)
        ids.append(np.frombuffer(take(4 * n), np.uint32))
        hi.append(np.frombuffer(take(n * dc), np.uint8).reshape(n, dc))
        lo_.append(np.frombuffer(take(n * dc), np.uint8).reshape(n, dc))
        norm.append(np.frombuffer(take(8 * n), np.int64))
        hnorm.append(np.frombuffer(take(8 * n), np.int64))
    ch = np.frombuffer(take(nlist * dc), np.uint8).reshape(nlist, dc)
    cl = np.frombuffer(take(nlist * dc), np.uint8).reshape(nlist, dc)
    cn = np.frombuffer(take(8 * nlist), np.int64)
    cat = lambda xs, shape: np.concatenate(xs) if xs else np.zeros(shape)
    return dict(d=d, dc=dc, nlist=nlist, npts=npts, prefix_dims=head, rotated=rotated, lo=lo,
                inv_step=inv_step, step=step, step_sq=step_sq, sizes=np.array(sizes, np.uint32),
                ids=cat(ids, 0), hi=cat(hi, (0, dc)), lo_plane=cat(lo_, (0, dc)), norm=cat(norm, 0),
                headnorm=cat(hnorm, 0), cen_hi=ch, cen_lo=cl, cen_norm=cn)


def ref_search(prefix, queries_fbin, k, nprobe, survivors, ranking="prefix", selection="local",
               batch=4096, out=None, coarse="batched"):
    """Runs the reference Index::search (AVX-512 engine) through katoann_ref_search.
    nprobe: an int (returns ids, dists) or a list of them, searched after one
    load (returns a list of (ids, dists))."""
    import tempfile
    many = not isinstance(nprobe, (int, np.integer))
    nps = list(nprobe) if many else [int(nprobe)]
    tmp = out or tempfile.mktemp(suffix=".bin")
    subprocess.run([REF_SEARCH, "search", prefix, queries_fbin, tmp, str(k), ",".join(str(x) for x in nps),
                    str(survivors), ranking, selection, str(batch), "0", coarse], check=True)
    raw = np.fromfile(tmp, dtype=np.uint8)
    if out is None:
        os.remove(tmp)
    res = []
    block = raw.size // len(nps)
    for i in range(len(nps)):
        r = raw[i * block:(i + 1) * block]
        nqk = r.size // 8
        res.append((r[:4 * nqk].view(np.uint32), r[4 * nqk:].view(np.float32)))
    return res if many else res[0]


def ref_flat_order(prefix, queries_fbin, nq, nlist):
    """The reference's per-query coarse order of every list (Index::coarse_flat_
    with nprobe = nlist) through katoann_ref_search, nq x nlist."""
    import tempfile
    tmp = tempfile.mktemp(suffix=".bin")
    subprocess.run([REF_SEARCH, "flat", prefix, queries_fbin, tmp], check=True)
    out = np.fromfile(tmp, dtype=np.uint32).reshape(nq, nlist)
    os.remove(tmp)
    return out
