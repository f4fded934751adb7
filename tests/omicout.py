"""Readers of the scan's output files, following OmicOUT's parser of
_sto.ibin / _sto.dbin (OmicOUT/main.cpp:280-461) and the _dis.txt layout
(src/AIOwrapper.cpp:945,1497-1511).  Test helpers.

Two _sto.dbin layouts (src/AIOwrapper.cpp:1454-1531):
- default: per (trait block, SNP block) a header (first trait, record
  count, AR offset), then records of n_used, R2 (float16), size_p and
  size_p results (AL index when covariates are displayed, AR index, B, SE);
- -f (storePInd): no header, no size_p and no indices; every (trait, SNP)
  record holds n_used, R2 and h_max = r (+ l - 1 with -i) results (B, SE),
  h descending, so a record's position says what it is
  (OmicOUT/main.cpp:347-420)."""
import struct

import numpy as np


def read_ibin(path, namelength=32):
    with open(path, "rb") as f:
        n, l, r, t, m, nl = struct.unpack("<6i", f.read(24))
        disp_cov, store_all = struct.unpack("<2?", f.read(2))
        def names(k):
            raw = f.read(nl * k)
            return [raw[i * nl:(i + 1) * nl].split(b"\0", 1)[0].decode() for i in range(k)]
        al = names(l - 1)
        ar = names(r * m)
        y = names(t)
    return dict(n=n, l=l, r=r, t=t, m=m, namelength=nl, disp_cov=disp_cov, store_all=store_all,
                al_names=al, ar_names=ar, y_names=y)


def float32_from_bits16(b):
    return struct.unpack("<f", struct.pack("<I", (b & 0xFFFF) << 16))[0]


def read_dbin(path, disp_cov, store_all, l, r, max_blocks=None):
    """Blocks of (trait block, SNP block) lists: returns a list of records
    dict(y, ar_offset, n_used, r2, results=[(al_idx, ar_idx, B, SE)])."""
    recs = []
    data = open(path, "rb").read()
    pos = 0
    blocks = 0
    while pos < len(data) and (max_blocks is None or blocks < max_blocks):
        if store_all:
            raise ValueError("the -f layout is positional: read it with read_dbin_pind")
        y_idx, count, ar_off = struct.unpack_from("<3i", data, pos)
        pos += 12
        for _ in range(count):
            n_used, r2b = struct.unpack_from("<HH", data, pos)
            pos += 4
            size_p = data[pos]
            pos += 1
            res = []
            for _ in range(size_p):
                al = 255
                if disp_cov:
                    al = data[pos]
                    pos += 1
                ar = struct.unpack_from("<H", data, pos)[0]
                pos += 2
                B, SE = struct.unpack_from("<ff", data, pos)
                pos += 8
                res.append((al, ar, B, SE))
            recs.append(dict(y=y_idx, ar_offset=ar_off, n_used=n_used, r2=float32_from_bits16(r2b), results=res))
        blocks += 1
    return recs


def read_dis(path):
    lines = open(path).read().splitlines()
    assert lines[0] == "Phe\tSNP\tn\tnPct\tB\tR2\tSE\tT\tP", lines[0]
    rows = []
    for ln in lines[1:]:
        f = ln.split("\t")
        rows.append(dict(phe=f[0], snp=f[1], n=int(f[2]), npct=float(f[3]), B=float(f[4]), R2=float(f[5]),
                         SE=float(f[6]), T=float(f[7]), P=float(f[8]), raw=ln))
    return rows


def expected_dis_rows(o, names, l, r, min_p_disp, min_r2_disp, real_n, tb=1000, mb=500, store_all=False):
    """The _dis.txt rows the reference writes for a dense oracle result, in
    its order: trait block, SNP block, trait, SNP, h descending."""
    t, m, p = o["beta"].shape
    rows = []
    for y0 in range(0, t, tb):
        for s0 in range(0, m, mb):
            for y in range(y0, min(t, y0 + tb)):
                for s in range(s0, min(m, s0 + mb)):
                    if not o["emitted"][y, s]:
                        continue
                    for h in range(p - 1, -1, -1):
                        if not o["kept"][y, s, h]:
                            continue
                        P = o["pval"][y, s, h]
                        R2 = np.float32(o["r2"][y, s])
                        if not ((P <= np.float32(min_p_disp) and R2 >= np.float32(min_r2_disp)) or store_all):
                            continue
                        ii = s - s0
                        ar = ii * r + (h - l if h >= l else 0)
                        aname = (names["al"][h] + ":") if h < l else " "
                        aname += names["ar"][s0 * r + ar]
                        rows.append(dict(phe=names["y"][y], snp=aname, n=int(o["n_used"][y]),
                                         B=o["beta"][y, s, h], SE=o["se"][y, s, h],
                                         T=o["tstat"][y, s, h], P=P, R2=float(R2),
                                         npct=o["n_used"][y] / real_n))
    return rows


def expected_records(o, l, r, tb=1000, mb=500):
    """The _sto.dbin records of a dense oracle result, in file order, each
    as (block's first trait, AR offset, n_used, [(AL idx, AR idx, B, SE)])."""
    t, m, p = o["beta"].shape
    out = []
    for y0 in range(0, t, tb):
        for s0 in range(0, m, mb):
            first = None
            for y in range(y0, min(t, y0 + tb)):
                for s in range(s0, min(m, s0 + mb)):
                    if not o["emitted"][y, s]:
                        continue
                    if first is None:
                        first = y
                    res = []
                    for h in range(p - 1, -1, -1):
                        if o["kept"][y, s, h]:
                            ii = s - s0
                            res.append((255 if h >= l else h, ii * r + (h - l if h >= l else 0),
                                        o["beta"][y, s, h], o["se"][y, s, h]))
                    out.append((first, s0 * r, int(o["n_used"][y]), res))
    return out


def read_dbin_pind(path, h_max):
    """The -f (storePInd) layout: fixed-size records in file order, each
    dict(n_used, r2, results=[(B, SE)] * h_max)."""
    data = open(path, "rb").read()
    size = 4 + 8 * h_max
    assert len(data) % size == 0, (len(data), size)
    recs = []
    for pos in range(0, len(data), size):
        n_used, r2b = struct.unpack_from("<HH", data, pos)
        vals = struct.unpack_from("<%df" % (2 * h_max), data, pos + 4)
        recs.append(dict(n_used=n_used, r2=float32_from_bits16(r2b),
                         results=[(vals[2 * k], vals[2 * k + 1]) for k in range(h_max)]))
    return recs


def pind_order(t, m, tb=1000, mb=500):
    """(trait, SNP) of each -f record in the writer's order: trait block, SNP
    block, trait, SNP.  OmicOUT reads them as trait-major over every SNP,
    which is the same order when t <= tb and m <= mb."""
    out = []
    for y0 in range(0, t, tb):
        for s0 in range(0, m, mb):
            for y in range(y0, min(t, y0 + tb)):
                for s in range(s0, min(m, s0 + mb)):
                    out.append((y, s))
    return out
