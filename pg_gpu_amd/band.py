"""Tiled int8-MFMA band path: packed indicator planes, band plans and fused
per-key LD sums (libpgld pgld_pack_planes / pgld_band_*).

This replaces the chunk loop of moments_ld._compute_ld_sums
(moments_ld.py:248-343) and HaplotypeMatrix.compute_ld_statistics_gpu_*
(haplotype_matrix.py:1592-1866): one launch counts every pair of the distance
band for every population with int8 MFMA, evaluates the statistics from the
exact counts and reduces them per key (distance bin or window) in a fixed
order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def k_pad_for(n_samples, fp4=False):
    """K padding of a population: a multiple of 128 samples (256 for FP4
    planes, whose rows hold two samples per byte), at least one multiple."""
    q = 256 if fp4 else 128
    return max(q, ((int(n_samples) + q - 1) // q) * q)


@dataclass
class PackedPop:
    """Indicator planes of one population (SNP-major, K padded)."""
    family: int
    planes: object       # int8 tensor (n_planes, n_var, k_pad), or (n_planes, n_var, k_pad // 2) for FP4
    marg: object         # int32 tensor (n_planes, n_var)
    k_pad: int
    n_samples: int
    has_missing: object  # bool, or None when the flags were not read back
    flags: object = None  # int32 tensor [any missing, any out of domain]
    bits: object = None   # uint32-as-int32 tensor of bit planes (popcount variant), made on demand
    bits_fresh: bool = False
    fp4: bool = False     # planes are FP4 (pgld_pack_planes_fp4): fused sums and r^2 passes; counts and
    #                       the popcount variant take int8 planes

    def bit_planes(self):
        """Bit planes of this population (32 samples per word), rebuilt from
        the int8 planes when they were repacked since."""
        if self.fp4:
            raise ValueError("bit planes are made from int8 planes, not FP4 ones")
        n_planes, n_var, kp = self.planes.shape
        if self.bits is None:
            self.bits = L.empty((n_planes, n_var, 32 * ((kp + 1023) // 1024)), np.int32)
            self.bits_fresh = False
        if not self.bits_fresh:
            L.call("pgld_pack_bits", L.ptr(self.planes), n_planes, n_var, kp, L.ptr(self.bits), L.stream_handle())
            self.bits_fresh = True
        return self.bits


def pack_population(data, rows=None, family="hap", k_pad=None, validate=True, out=None, fp4=False):
    """Pack one population's rows of a (rows, n_var) int8 matrix.

    fp4=True packs the indicators as FP4 (two samples per byte) for the
    FP4-MFMA passes (BandPlan.sums and BandPlan.r2, every estimator, with or
    without missing data); BandPlan.counts and the popcount variant take the
    int8 planes.

    validate=True reads the kernel's value flags back (one small sync) to
    reject out-of-domain values and report missing data; validate=False skips
    that read for inputs already checked (has_missing is then None).  ``out``
    may pass a previous PackedPop of the same shape whose buffers are reused."""
    fam = L.FAMILY_HAP if family == "hap" else L.FAMILY_GENO
    g = L.to_device(data, np.int8)
    n_rows, n_var = g.shape
    rows_t = None
    n_sel = n_rows
    if rows is not None:
        T = L.torch()
        if not validate and L.is_tensor(rows) and rows.is_cuda and rows.dtype == T.int32 and rows.dim() == 1:
            rows_t, n_sel = rows, int(rows.shape[0])  # trusted device row list: no host round trip
        else:
            r = np.asarray(L.to_numpy(rows), dtype=np.int64)
            if r.size and (r.min() < 0 or r.max() >= n_rows):
                raise IndexError("population row index out of range")
            rows_t = L.to_device(r, np.int32)
            n_sel = int(r.size)
    kp = k_pad_for(n_sel, fp4) if k_pad is None else int(k_pad)
    npl = 2 if fam == L.FAMILY_HAP else 3
    row = kp // 2 if fp4 else kp
    if out is not None and tuple(out.planes.shape) == (npl, n_var, row) and out.fp4 == fp4:
        planes, marg, flags = out.planes, out.marg, out.flags
    else:
        planes = L.empty((npl, n_var, row), np.int8)
        marg = L.empty((npl, n_var), np.int32)
        flags = L.empty((2,), np.int32)
    L.call("pgld_pack_planes_fp4" if fp4 else "pgld_pack_planes", fam, L.ptr(g), n_rows, n_var, L.ptr(rows_t),
           n_sel, kp, L.ptr(planes), L.ptr(marg), L.ptr(flags), L.stream_handle())
    bits = out.bits if (out is not None and out.planes is planes and not fp4) else None
    if not validate:
        return PackedPop(fam, planes, marg, kp, n_sel, None, flags, bits, fp4=fp4)
    f = flags.cpu().numpy()
    if f[1]:
        raise ValueError("input values outside {-1, 0, 1}" if fam == L.FAMILY_HAP
                         else "genotype values outside {-1, 0, 1, 2}")
    return PackedPop(fam, planes, marg, kp, n_sel, bool(f[0]), flags, bits, fp4=fp4)


class BandPlan:
    """Tile schedule of the band {(i, j): i < j, pos[j] - pos[i] <= max_dist}
    keyed by distance bins (``edges``) or by windows of the left SNP
    (``origin``, ``width``, ``n_keys``)."""

    def __init__(self, positions, max_dist, edges=None, *, window=None, n_keys=None, origin=0.0,
                 key_coord=None, abs_dist=False):
        L.lib()
        pos = np.ascontiguousarray(L.to_numpy(positions), dtype=np.float64)
        self.n_var = int(pos.shape[0])
        self.max_dist = float(max_dist)
        kc = None if key_coord is None else np.ascontiguousarray(L.to_numpy(key_coord), dtype=np.float64)
        if window is not None:
            mode = L.KEY_WINDOW
            if n_keys is None:
                span = (pos[-1] - origin) if self.n_var else 0.0
                n_keys = int(np.floor(span / window)) + 1 if self.n_var else 1
            e = None
            width = float(window)
        else:
            if edges is None:
                raise ValueError("BandPlan needs bin edges or a window width")
            e = np.ascontiguousarray(edges, dtype=np.float64)
            n_keys = len(e) - 1
            mode = L.KEY_BIN_ABS if abs_dist else L.KEY_BIN
            width = 1.0
        self.n_keys = int(n_keys)
        self.key_mode = mode
        self._keep = (pos, kc, e)
        h = C.c_void_p()
        L.call("pgld_band_plan_create", pos.ctypes.data_as(C.c_void_p), self.n_var, self.max_dist, mode,
               None if kc is None else kc.ctypes.data_as(C.c_void_p),
               None if e is None else e.ctypes.data_as(C.c_void_p), self.n_keys, float(origin), width,
               C.byref(h))
        self._h = h
        nt, npairs, ns = C.c_int64(), C.c_int64(), C.c_int64()
        L.call("pgld_band_plan_info", h, C.byref(nt), C.byref(npairs), C.byref(ns))
        self.n_tiles, self.n_pairs, self.n_slots = nt.value, npairs.value, ns.value

    def set_timing(self, enable=True):
        """Record HIP events around each fused-kernel launch of sums()."""
        L.call("pgld_band_plan_set_timing", self._h, int(bool(enable)))

    def kernel_ms(self):
        """Duration of the last timed fused-kernel launch (waits for it)."""
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

    # ------------------------------------------------------------------
    def sums(self, pops, stat_set="moments", with_missing=None, want_pairs=False):
        """Fused band pass over packed populations.  Returns float64 tensors
        (n_keys, n_stats) and, if asked, the pair count per key."""
        P = len(pops)
        fam = pops[0].family
        if any(p.family != fam for p in pops):
            raise ValueError("all populations must be of one family")
        fp4 = pops[0].fp4
        if any(p.fp4 != fp4 for p in pops):
            raise ValueError("all populations must be packed in one layout (int8 or FP4)")
        if with_missing is None:
            if any(p.has_missing is None for p in pops):
                raise ValueError("with_missing must be given for populations packed with validate=False")
            with_missing = any(p.has_missing for p in pops)
        sset = L.SET_MOMENTS if stat_set == "moments" else L.SET_SELF
        if sset == L.SET_SELF and P > 1:
            n_stats = 3 * P
        else:
            n_stats = {1: 3, 2: 15, 3: 45, 4: 105}[P]
        sums = L.empty((self.n_keys, n_stats), np.float64)
        kp = L.empty((self.n_keys,), np.float64) if want_pairs else None
        planes, _k1 = L.ptr_array([p.planes for p in pops])
        margs, _k2 = L.ptr_array([p.marg for p in pops])
        kpad = (C.c_int64 * P)(*[p.k_pad for p in pops])
        ns = (C.c_int32 * P)(*[p.n_samples for p in pops])
        L.call("pgld_band_sums_fp4" if fp4 else "pgld_band_sums", self._h, fam, P, planes, margs, kpad, ns,
               int(bool(with_missing)), sset, L.ptr(sums), L.ptr(kp), L.stream_handle())
        return (sums, kp) if want_pairs else sums

    def r2(self, pop, n_hist=100, thresh=0.2, sparse_cap=1 << 20, per_pair=False, histogram=True,
           with_missing=None, method="mfma", estimator="r2", snp_sums=False):
        """r^2-family values of every band pair of one haplotype population
        through the tiled counts (pgld_band_ld_pairs / pgld_band_r2_bits).

        estimator: "r2" (from the pair's counts, NaN where undefined),
        "naive" (per-site frequencies, 0 where undefined, as pairwise_r2),
        "sigma_d2" (unbiased D^2 / pi_2), "d_naive" (signed D = p_AB - p_i p_j
        on the frequencies of "naive", as pairwise_LD_v); for a genotype population
        "rogers_huff_r" (signed dosage r, no missing data) or "dosage_r2"
        (dosage r^2 about site means).  Returns a dict with
          hist      int64 (n_keys, n_hist) pairs per bin [b/n, (b+1)/n),
          nan_pairs int64 (n_keys,) pairs whose value is NaN,
          sparse_i, sparse_j, sparse_r2  pairs with value > thresh, sorted by
                    (i, j) (at most sparse_cap are kept),
          n_sparse  number of such pairs (may exceed sparse_cap),
          r2        FP64 per band pair in (i, j) order if per_pair,
          row_sums, col_sums, row_counts, col_counts  per-SNP sums over the
                    pairs (v, j > v) and (i < v, v) if snp_sums."""
        if pop.fp4 and method != "mfma":
            raise ValueError("the popcount variant takes int8 planes (pack without fp4)")
        geno_est = estimator in ("rogers_huff_r", "dosage_r2")
        if (pop.family == L.FAMILY_GENO) != geno_est:
            raise ValueError("estimators 'r2', 'naive', 'sigma_d2', 'd_naive' take haplotype planes; "
                             "'rogers_huff_r' and 'dosage_r2' take genotype planes")
        if with_missing is None:
            with_missing = pop.has_missing
            if with_missing is None:
                raise ValueError("with_missing must be given for populations packed with validate=False")
        est = {"r2": L.EST_R2, "naive": L.EST_R2_NAIVE, "sigma_d2": L.EST_SIGMA_D2, "rogers_huff_r": L.EST_RH_R,
               "dosage_r2": L.EST_DIP_R2, "d_naive": L.EST_D_NAIVE}.get(estimator)
        if est is None:
            raise ValueError("estimator is 'r2', 'naive', 'sigma_d2', 'd_naive', 'rogers_huff_r' or 'dosage_r2'")
        if method not in ("mfma", "popcount"):
            raise ValueError("method is 'mfma' or 'popcount'")
        if method == "popcount" and (est != L.EST_R2 or snp_sums):
            raise ValueError("the popcount variant computes the 'r2' estimator only")
        T = L.torch()
        hist = L.empty((self.n_keys, n_hist + 1), np.int64) if histogram else None
        cap = int(sparse_cap) if sparse_cap else 0
        sp_i = L.empty((max(cap, 1),), np.int32)
        sp_j = L.empty((max(cap, 1),), np.int32)
        sp_r = L.empty((max(cap, 1),), np.float32)
        n_sp = L.empty((1,), np.int64)
        r2 = L.empty((self.n_pairs,), np.float64) if per_pair else None
        sums = L.empty((2 * self.n_var,), np.float64) if snp_sums else None
        cnts = L.empty((2 * self.n_var,), np.int64) if snp_sums else None
        if method == "mfma":
            L.call("pgld_band_ld_pairs_fp4" if pop.fp4 else "pgld_band_ld_pairs", self._h, L.ptr(pop.planes),
                   L.ptr(pop.marg), pop.k_pad, pop.n_samples,
                   int(bool(with_missing)), est, int(n_hist), float(thresh), L.ptr(hist), L.ptr(sp_i), L.ptr(sp_j),
                   L.ptr(sp_r), cap, L.ptr(n_sp), L.ptr(r2), L.ptr(sums), L.ptr(cnts), L.stream_handle())
        else:
            L.call("pgld_band_r2_bits", self._h, L.ptr(pop.bit_planes()), L.ptr(pop.marg), pop.k_pad,
                   pop.n_samples, int(bool(with_missing)), int(n_hist), float(thresh), L.ptr(hist), L.ptr(sp_i),
                   L.ptr(sp_j), L.ptr(sp_r), cap, L.ptr(n_sp), L.ptr(r2), L.stream_handle())
        n_total = int(n_sp.cpu().item())
        k = min(n_total, cap)
        si, sj, sr = sp_i[:k], sp_j[:k], sp_r[:k]
        if k:
            order = T.argsort(si.to(T.int64) * self.n_var + sj.to(T.int64))
            si, sj, sr = si[order], sj[order], sr[order]
        out = {"sparse_i": si, "sparse_j": sj, "sparse_r2": sr, "n_sparse": n_total}
        if histogram:
            out["hist"] = hist[:, :n_hist]
            out["nan_pairs"] = hist[:, n_hist]
        if per_pair:
            out["r2"] = r2
        if snp_sums:
            out["row_sums"], out["col_sums"] = sums[:self.n_var], sums[self.n_var:]
            out["row_counts"], out["col_counts"] = cnts[:self.n_var], cnts[self.n_var:]
        return out

    def counts(self, pop, with_missing=None, method="mfma"):
        """Per-pair counts of one population through the tiled path (int8
        MFMA, or AND + popcount on bit planes), in the band's (i, j) pair
        order: int32 (n_pairs, 4) or (n_pairs, 9)."""
        if pop.fp4:
            raise ValueError("the counts pass takes int8 planes (pack without fp4)")
        if with_missing is None:
            with_missing = pop.has_missing
        if method not in ("mfma", "popcount"):
            raise ValueError("method is 'mfma' or 'popcount'")
        nc = 4 if pop.family == L.FAMILY_HAP else 9
        out = L.empty((self.n_pairs, nc), np.int32)
        src = pop.planes if method == "mfma" else pop.bit_planes()
        L.call("pgld_band_counts" if method == "mfma" else "pgld_band_counts_bits", self._h, pop.family, L.ptr(src),
               L.ptr(pop.marg), pop.k_pad, pop.n_samples, int(bool(with_missing)), None, L.ptr(out),
               L.stream_handle())
        return out


# ---------------------------------------------------------------------------
# Fused band pass where a fused kernel exists, the per-pair path otherwise
# ---------------------------------------------------------------------------

FUSED_MAX_SAMPLES = 32767  # int16 count stash and 24-bit products of the fused kernels


def fused_supported(family, n_pops, stat_set, pop_sizes):
    """Whether the fused band kernels cover this statistic set: haplotype and
    genotype sets for 1-4 populations (the 105 four-population moments in
    three column groups; the three- and four-population genotype moments park
    their counts in a global workspace); at most FUSED_MAX_SAMPLES samples per
    population."""
    if max(pop_sizes) > FUSED_MAX_SAMPLES:
        return False
    return 1 <= n_pops <= 4


def ld_sums(data, family, positions, rows, max_dist, *, edges=None, window=None, n_keys=None, origin=0.0,
            key_coord=None, abs_dist=False, stat_set="moments", method="auto"):
    """Per-key FP64 sums of the LD statistics of every pair within max_dist
    (keys: distance bins ``edges`` of ``key_coord`` or the positions, or
    windows of the left SNP).  method "auto" runs the fused band pass and
    falls back to the per-pair path (plain counts -> statistics -> keyed
    sums, as moments_ld.py:294-343) when no fused kernel covers the set, a
    population exceeds the fused kernels' sample limit, or a tile's keys x
    statistics exceed the fused kernel's LDS; "tiled" / "plain" force one.
    Returns (sums (n_keys, n_stats), pairs per key) as device tensors."""
    pos = np.ascontiguousarray(L.to_numpy(positions), dtype=np.float64)
    n_rows = data.shape[0]
    sizes = [n_rows if r is None else len(r) for r in rows]
    P = len(rows)
    if window is not None:
        mode = L.KEY_WINDOW
        if n_keys is None:
            n_keys = int(np.floor((pos[-1] - origin) / window)) + 1 if len(pos) else 1
    else:
        mode = L.KEY_BIN_ABS if abs_dist else L.KEY_BIN
        n_keys = len(edges) - 1
    if method == "tiled" or (method == "auto" and fused_supported(family, P, stat_set, sizes)):
        if window is not None:
            plan = BandPlan(pos, max_dist, window=window, n_keys=n_keys, origin=origin)
        else:
            plan = BandPlan(pos, max_dist, edges=edges, key_coord=key_coord, abs_dist=abs_dist)
        packed = [pack_population(data, r, family, fp4=True) for r in rows]
        try:
            return plan.sums(packed, stat_set=stat_set, want_pairs=True)
        except L.PgldError as e:
            if method == "tiled" or e.status != L.EUNSUPPORTED:
                raise
    from . import ld_pipeline as LP
    return LP.plain_keyed_sums(data, family, pos, rows, max_dist, mode, n_keys, edges=edges, origin=origin,
                               width=1.0 if window is None else window, key_coord=key_coord, stat_set=stat_set)
