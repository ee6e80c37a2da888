"""Minimal VCF reader for the LD path (host side, NumPy).

Replaces the scikit-allel ``read_vcf`` call that the reference's loaders use
(haplotype_matrix.py:317, genotype_matrix.py:368): it returns the GT calls
as an int8 (n_variants, n_samples, 2) array with -1 for missing alleles, the
1-based POS column, the CHROM column and the sample names.  Plain or gzipped
text; the FORMAT column must carry GT (any position).
"""
from __future__ import annotations

import gzip

import numpy as np


def _open(path):
    with open(path, "rb") as f:
        magic = f.read(2)
    if magic == b"\x1f\x8b":
        return gzip.open(path, "rt")
    return open(path, "rt")


def _parse_gt(field, gt_index):
    parts = field.split(":")
    if gt_index >= len(parts):
        return -1, -1
    gt = parts[gt_index]
    sep = "|" if "|" in gt else "/"
    alleles = gt.split(sep)
    a0 = -1 if alleles[0] in (".", "") else int(alleles[0])
    if len(alleles) > 1:
        a1 = -1 if alleles[1] in (".", "") else int(alleles[1])
    else:
        a1 = -1
    return a0, a1


def read_vcf(path, samples=None):
    """Read a VCF; returns dict(gt, positions, chrom, samples).

    ``samples`` optionally restricts (and orders) the loaded sample columns
    by name, as scikit-allel's ``samples=`` argument does."""
    sample_names = None
    pos, chrom, rows = [], [], []
    sel = None
    with _open(path) as f:
        for line in f:
            if line.startswith("##"):
                continue
            if line.startswith("#CHROM"):
                hdr = line.rstrip("\n").split("\t")
                sample_names = hdr[9:]
                if samples is not None:
                    idx = {s: k for k, s in enumerate(sample_names)}
                    sel = [idx[s] for s in samples]
                    sample_names = list(samples)
                continue
            if not line.strip():
                continue
            fields = line.rstrip("\n").split("\t")
            fmt = fields[8].split(":")
            if "GT" not in fmt:
                continue
            gi = fmt.index("GT")
            calls = fields[9:]
            if sel is not None:
                calls = [calls[k] for k in sel]
            rows.append([_parse_gt(c, gi) for c in calls])
            pos.append(int(fields[1]))
            chrom.append(fields[0])
    if sample_names is None:
        raise ValueError(f"{path}: no #CHROM header line")
    if not rows:
        raise ValueError(f"No variants found in {path}")
    gt = np.asarray(rows, dtype=np.int16)
    if gt.max() > 127:
        raise ValueError("allele index above 127 cannot be stored as int8")
    return {
        "gt": gt.astype(np.int8),
        "positions": np.asarray(pos, dtype=np.int64),
        "chrom": np.asarray(chrom),
        "samples": list(sample_names),
    }


def read_pop_file(pop_file):
    """sample -> population map from a whitespace-delimited file; a header
    line whose first field is 'sample' is skipped (haplotype_matrix.py:
    494-500)."""
    pop_map = {}
    with open(pop_file) as f:
        for line in f:
            parts = line.strip().split()
            if len(parts) >= 2 and parts[0] != "sample":
                pop_map[parts[0]] = parts[1]
    return pop_map
