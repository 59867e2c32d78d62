"""An independent, deliberately naive Python writer of the 2D tile image.

Written from the byte layout (matrix/sparse_matrix_format.h:130-240) and the
writer's rules (matrix/fm_utils.cpp:267-433: runs of one non-zero go to the
COO region, longer runs to SCSR row parts; empty block rows get one empty
tile), with dictionaries and sorting instead of the oracle's two-pass scan.
Agreement of two independent writers on random inputs is the check the
reference itself offers through its twin writers (fm_utils.cpp vs
flash-graph/fg_utils.cpp:1097-1280).  Small inputs only.
"""
import struct
from collections import defaultdict


def normalise(src, dst, n=0, dedup=True, keep_self=False, symmetrize=False):
    edges = []
    for s, d in zip(src, dst):
        s, d = int(s), int(d)
        if s == d and not keep_self:
            continue
        edges.append((s, d))
        if symmetrize and s != d:
            edges.append((d, s))
    if n == 0:
        n = max(max(int(x) for x in src), max(int(x) for x in dst)) + 1
    rows = defaultdict(list)
    for s, d in edges:
        rows[s].append(d)
    for r in rows:
        rows[r].sort()
        if dedup:
            rows[r] = sorted(set(rows[r]))
    return rows, n


def write(src, dst, log2h, log2w, n=0, dedup=True, keep_self=False, symmetrize=False):
    rows, n = normalise(src, dst, n, dedup, keep_self, symmetrize)
    H, W = 1 << log2h, 1 << log2w
    nbr = (n + H - 1) // H
    hdr = struct.pack("<qiiQQQiiQQ", 0x0123456789FFFFFF, 2, 2, 0, n, n, 2, 0, H, W)
    out = bytearray(hdr + bytes(4096 - len(hdr)))
    offs = []
    for br in range(nbr):
        offs.append(len(out))
        tiles = defaultdict(lambda: ([], []))  # bc -> (scsr runs, coo entries)
        for r in range(br * H, min(n, (br + 1) * H)):
            runs = defaultdict(list)
            for c in rows.get(r, []):
                runs[c >> log2w].append(c)
            for bc, run in runs.items():
                if len(run) > W:  # fm_utils.cpp:403-406
                    raise ValueError("more neighbours than the block size")
                if len(run) == 1:
                    tiles[bc][1].append((r & (H - 1), run[0] & (W - 1)))
                else:
                    tiles[bc][0].append((r & (H - 1), [c & (W - 1) for c in run]))
        if not tiles:
            out += struct.pack("<IIIHH", br, 0, 0, 0, 0)
            continue
        for bc in sorted(tiles):
            scsr, coo = tiles[bc]
            scsr.sort()
            coo.sort()
            nnz = sum(len(c) for _, c in scsr) + len(coo)
            nrow = len(scsr) + len(coo)
            out += struct.pack("<IIIHH", br, bc, nnz, nrow, len(coo))
            for r, cols in scsr:
                out += struct.pack("<H", 0x8000 | r)
                out += struct.pack("<%dH" % len(cols), *cols)
            out += struct.pack("<H", 0xFFFF)
            for r, c in coo:
                out += struct.pack("<HH", 0x8000 | r, c)
    offs.append(len(out))
    return bytes(out), offs, n
