"""Write tests/golden/moments_im_4bins.json from the reference's frozen moments
output, without unpickling it.

The reference caches moments.LD.Parsing.compute_ld_statistics on its IM
example (tests/cache/moments_ld_test_im_4bins_deme0-deme1.pkl, written by
tests/test_ld_validation_full.py:39-73).  This script walks the pickle's
opcode stream with pickletools.genops (which executes nothing), collects the
float64 payloads of the numpy scalars (bin edges) and arrays (per-bin sums,
heterozygosity) and the statistic/population name strings, and stores them
as JSON.  Run once in a container that has /root/reference:

    python tests/golden/make_moments_fixture.py
"""
import json
import os
import pickletools
import struct
import sys

SRC = "/root/reference/tests/cache/moments_ld_test_im_4bins_deme0-deme1.pkl"
DST = os.path.join(os.path.dirname(__file__), "moments_im_4bins.json")


def main(src=SRC, dst=DST):
    data = open(src, "rb").read()
    blobs, strings = [], []
    for op, arg, _ in pickletools.genops(data):
        if op.name in ("SHORT_BINBYTES", "BINBYTES") and isinstance(arg, bytes):
            if len(arg) % 8 == 0 and len(arg) >= 8:
                blobs.append(struct.unpack("<%dd" % (len(arg) // 8), arg))
        elif op.name in ("SHORT_BINUNICODE", "BINUNICODE"):
            strings.append(arg)
    scalars = [b[0] for b in blobs if len(b) == 1]
    arrays = [list(b) for b in blobs if len(b) > 1]
    edges = []
    for i in range(0, len(scalars), 2):
        edges.append([scalars[i], scalars[i + 1]])
    ld_names = [s for s in strings if s.startswith(("DD_", "Dz_", "pi2_"))]
    het_names = [s for s in strings if s.startswith("H_")]
    pops = [s for s in strings if s.startswith("deme")]
    assert len(arrays) == len(edges) + 1, (len(arrays), len(edges))
    out = {
        "source": "moments.LD.Parsing.compute_ld_statistics(examples/data/im-parsing-example.vcf, "
                  "pop_file=examples/data/im_pop.txt, pops=[deme0, deme1], "
                  "bp_bins=logspace(2,5,4), use_genotypes=False)",
        "bins": edges,
        "sums": arrays[:-1],
        "het": arrays[-1],
        "ld_names": ld_names,
        "het_names": het_names,
        "pops": pops,
    }
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst, "bins", len(edges), "stats", len(ld_names))


if __name__ == "__main__":
    main(*sys.argv[1:])
