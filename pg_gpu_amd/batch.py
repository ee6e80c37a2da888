"""Ragged region batches and HIP graph capture (libpgld pgld_batch_* /
pgld_graph_*).

The reference computes LD decay of independent regions one
moments_ld.compute_ld_statistics call at a time (moments_ld.py:40-203), each
its own chunk loop.  A RegionBatch holds the tile schedule of every region at
once -- regions of any SNP count, with their own populations and sample
counts -- and runs counts -> statistics -> per-(region, bin) sums as one
persistent launch whose schedule is picked by the library's autotuner.  A
Graph captures a sequence of libpgld calls (e.g. packing every region's
populations followed by the batch pass) and replays it as one submission.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .band import PackedPop


class RegionBatch:
    """Distance-bin (or window) keyed LD sums over many independent regions.

    positions: list of per-region sorted position arrays.  Keys are per region:
    ``edges`` (bins of |distance|, [e_k, e_{k+1})) or ``window`` (window of
    the left SNP).  Results are (n_regions, n_keys, n_stats)."""

    def __init__(self, positions, max_dist, edges=None, *, window=None, n_keys=None, origin=0.0, abs_dist=False):
        L.lib()
        pos = [np.ascontiguousarray(L.to_numpy(p), dtype=np.float64) for p in positions]
        self.n_regions = len(pos)
        if self.n_regions == 0:
            raise ValueError("a region batch needs at least one region")
        n_var = np.array([len(p) for p in pos], dtype=np.int64)
        cat = np.ascontiguousarray(np.concatenate(pos)) if n_var.sum() else np.zeros(1)
        if window is not None:
            if n_keys is None:
                span = max((p[-1] - origin) for p in pos if len(p)) if n_var.sum() else 0.0
                n_keys = int(np.floor(span / window)) + 1
            mode, e, width = L.KEY_WINDOW, None, float(window)
        else:
            if edges is None:
                raise ValueError("RegionBatch needs bin edges or a window width")
            e = np.ascontiguousarray(edges, dtype=np.float64)
            n_keys = len(e) - 1
            mode, width = (L.KEY_BIN_ABS if abs_dist else L.KEY_BIN), 1.0
        self.n_keys = int(n_keys)
        self.n_var = n_var
        self._keep = (cat, e, n_var)
        h = C.c_void_p()
        L.call("pgld_batch_plan_create", self.n_regions, n_var.ctypes.data_as(C.c_void_p),
               cat.ctypes.data_as(C.c_void_p), float(max_dist), mode, None,
               None if e is None else e.ctypes.data_as(C.c_void_p), self.n_keys, float(origin), width, C.byref(h))
        self._h = h
        nt, npairs, ns = C.c_int64(), C.c_int64(), C.c_int64()
        L.call("pgld_band_plan_info", h, C.byref(nt), C.byref(npairs), C.byref(ns))
        self.n_tiles, self.n_pairs, self.n_slots = nt.value, npairs.value, ns.value
        self.n_stats = None
        self._bound = None

    def bind(self, pops, stat_set="self"):
        """pops[r] = list of PackedPop of region r (same population count and
        family for every region)."""
        if len(pops) != self.n_regions:
            raise ValueError("one population list per region")
        P = len(pops[0])
        fam = pops[0][0].family
        fp4 = pops[0][0].fp4
        for r, pl in enumerate(pops):
            if len(pl) != P or any(p.family != fam for p in pl):
                raise ValueError(f"region {r}: every region needs {P} populations of one family")
            for p in pl:
                if p.fp4 != fp4:
                    raise ValueError(f"region {r}: every population must be packed in one layout (int8 or FP4)")
                if p.planes.shape[1] != self.n_var[r]:
                    raise ValueError(f"region {r}: planes have {p.planes.shape[1]} SNPs, plan has {self.n_var[r]}")
        flat = [p for pl in pops for p in pl]
        sset = L.SET_MOMENTS if stat_set == "moments" else L.SET_SELF
        planes, k1 = L.ptr_array([p.planes for p in flat])
        margs, k2 = L.ptr_array([p.marg for p in flat])
        kpad = (C.c_int64 * len(flat))(*[p.k_pad for p in flat])
        ns = (C.c_int32 * len(flat))(*[p.n_samples for p in flat])
        L.call("pgld_batch_bind_fp4" if fp4 else "pgld_batch_bind", self._h, fam, P, sset, planes, margs, kpad, ns)
        self.n_stats = (3 * P) if (sset == L.SET_SELF and P > 1) else {1: 3, 2: 15, 3: 45, 4: 105}[P]
        self._bound = (flat, k1, k2)  # keep the planes alive while bound
        return self

    def _out(self, out, key_pairs, want_pairs):
        if self.n_stats is None:
            raise RuntimeError("bind() the populations first")
        if out is None:
            out = L.empty((self.n_regions, self.n_keys, self.n_stats), np.float64)
        if want_pairs and key_pairs is None:
            key_pairs = L.empty((self.n_regions, self.n_keys), np.float64)
        return out, key_pairs

    def sums(self, with_missing, out=None, key_pairs=None, want_pairs=False, stream=None):
        """One persistent pass over every region.  Returns the (n_regions,
        n_keys, n_stats) FP64 sums (and pair counts per key if asked)."""
        out, key_pairs = self._out(out, key_pairs, want_pairs)
        L.call("pgld_batch_sums", self._h, int(bool(with_missing)), L.ptr(out), L.ptr(key_pairs),
               stream if stream is not None else L.stream_handle())
        return (out, key_pairs) if (want_pairs or key_pairs is not None) else out

    _ORDERS = ("cost", "natural", "blocked")

    def schedule(self, grid=0, order="cost", xcd_split=False, kernel_variant=0):
        """grid 0 = one workgroup per tile; order "cost" (descending work),
        "natural" or "blocked" (8 x 8 tile blocks per XCD for L2 reuse);
        kernel_variant 0 / 1 / 2 / 3 (FP4 within-population haplotype sets: K
        loop streamed across populations / per population / streamed with
        rolled epilogue rows / lean: per population with one load stage,
        three workgroups per CU; the sums are identical)."""
        if order not in self._ORDERS:
            raise ValueError(f"order must be one of {self._ORDERS}")
        L.call("pgld_batch_schedule", self._h, int(grid), self._ORDERS.index(order), int(bool(xcd_split)),
               int(kernel_variant))

    def autotune(self, with_missing):
        """Time the schedule and kernel-variant candidates, keep the fastest;
        returns the choice ("ms" is None when the in-process or on-disk
        cache answered)."""
        out, _ = self._out(None, None, False)
        g, o, x, kv, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        L.call("pgld_batch_autotune", self._h, int(bool(with_missing)), L.ptr(out), None, L.stream_handle(),
               C.byref(g), C.byref(o), C.byref(x), C.byref(kv), C.byref(ms))
        return {"grid": g.value, "order": self._ORDERS[o.value], "xcd_split": bool(x.value),
                "kernel_variant": kv.value, "ms": None if ms.value < 0 else float(ms.value)}

    def set_timing(self, enable=True):
        L.call("pgld_band_plan_set_timing", self._h, int(bool(enable)))

    def kernel_ms(self):
        ms = C.c_float()
        L.call("pgld_band_plan_kernel_ms", self._h, C.byref(ms))
        return float(ms.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().pgld_band_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured sequence of libpgld calls.  ``capture(fn)`` runs fn with
    torch's current stream set to a dedicated capture stream; every libpgld
    call fn makes on it is recorded (calls must not allocate or synchronize:
    pack into existing PackedPop buffers with validate=False, RegionBatch.sums
    into existing outputs).  ``launch()`` replays the whole sequence."""

    def __init__(self):
        T = L.torch()
        L.require_gpu()
        self.stream = T.cuda.Stream()
        self._g = None

    def capture(self, fn):
        T = L.torch()
        s = C.c_void_p(self.stream.cuda_stream)
        T.cuda.synchronize()
        with T.cuda.stream(self.stream):
            L.call("pgld_graph_begin", s)
            try:
                fn()
            except Exception:
                g = C.c_void_p()
                L.lib().pgld_graph_end(s, C.byref(g))
                if g.value:
                    L.lib().pgld_graph_destroy(g)
                raise
            g = C.c_void_p()
            L.call("pgld_graph_end", s, C.byref(g))
        self.close()
        self._g = g
        return self

    def launch(self, stream=None):
        if self._g is None:
            raise RuntimeError("nothing captured")
        L.call("pgld_graph_launch", self._g, stream if stream is not None else L.stream_handle())

    def close(self):
        if self._g is not None and self._g.value:
            L.lib().pgld_graph_destroy(self._g)
        self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackBatch:
    """Repacks the populations of many regions into existing PackedPop
    buffers with one launch (pgld_pack_batch_*): the same planes, margins
    and flags as pack_population on each, without thousands of small
    launches.  datas[r], rows_per_region[r] and packed[r] as pack_regions;
    the matrices, row lists and buffers must stay alive while it is used."""

    def __init__(self, datas, rows_per_region, packed):
        T = L.torch()
        jobs = []
        self._keep = []
        for d, rows, pk in zip(datas, rows_per_region, packed):
            g = L.to_device(d, np.int8)
            self._keep.append(g)
            for rw, p in zip(rows, pk):
                if rw is None:
                    rt, ns = None, g.shape[0]
                else:
                    rt = rw if (L.is_tensor(rw) and rw.is_cuda and rw.dtype == T.int32) else \
                        L.to_device(np.asarray(L.to_numpy(rw), dtype=np.int64), np.int32)
                    ns = int(rt.shape[0])
                    self._keep.append(rt)
                jobs.append((g, rt, ns, p))
        if not jobs:
            raise ValueError("nothing to pack")
        fam = jobs[0][3].family
        fp4 = jobs[0][3].fp4
        if any(p.family != fam or p.fp4 != fp4 for _, _, _, p in jobs):
            raise ValueError("all populations must share family and plane layout")
        n = len(jobs)
        data, _k1 = L.ptr_array([g for g, _, _, _ in jobs])
        rows_p, _k2 = L.ptr_array([rt for _, rt, _, _ in jobs])
        planes, _k3 = L.ptr_array([p.planes for _, _, _, p in jobs])
        margs, _k4 = L.ptr_array([p.marg for _, _, _, p in jobs])
        flags, _k5 = L.ptr_array([p.flags for _, _, _, p in jobs])
        n_rows = (C.c_int64 * n)(*[g.shape[0] for g, _, _, _ in jobs])
        n_var = (C.c_int64 * n)(*[g.shape[1] for g, _, _, _ in jobs])
        n_sel = (C.c_int64 * n)(*[ns for _, _, ns, _ in jobs])
        k_pad = (C.c_int64 * n)(*[p.k_pad for _, _, _, p in jobs])
        for _, _, ns, p in jobs:
            if p.n_samples != ns:
                raise ValueError("packed buffers do not match the row lists")
        h = C.c_void_p()
        L.call("pgld_pack_batch_create", fam, int(fp4), n, data, n_rows, n_var, rows_p, n_sel, k_pad, planes, margs,
               flags, C.byref(h))
        self._h = h

    def run(self, stream=None):
        L.call("pgld_pack_batch_run", self._h, stream if stream is not None else L.stream_handle())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().pgld_pack_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pack_regions(datas, rows_per_region, family="hap", out=None, validate=True, fp4=False):
    """PackedPop lists for every region: datas[r] is region r's int8 matrix,
    rows_per_region[r] its population row lists (fp4: FP4 planes)."""
    from .band import pack_population
    res = []
    for r, (d, rows) in enumerate(zip(datas, rows_per_region)):
        prev = out[r] if out is not None else [None] * len(rows)
        res.append([pack_population(d, rw, family, validate=validate, out=o, fp4=fp4) for rw, o in zip(rows, prev)])
    return res


__all__ = ["RegionBatch", "Graph", "PackBatch", "pack_regions", "PackedPop"]
