"""Model of the SpMM kernel's lane-sequential decode (flashx_amd/csrc/
spmm.hip, decode_chunk and the chunk/window bookkeeping of spmm_lane and
spmm_pipe), run on the CPU against oracle-built images: the (row, column)
stream it produces for whole block rows and for items split at arbitrary
words (with the carried row) equals a plain sequential decode of the tile
format (matrix/sparse_matrix.h:219-331).  The model exists to check the
algorithm's invariants -- 64*R-word aligned chunks, a directory window of
DWIN entries, at most NB tiles per lane -- on real tile streams; the GPU
tests check the kernel itself."""
import numpy as np
import pytest

def tiles_of(img_words, bstart, bend):
    # directory of one block row: (start_rel, bc) per tile
    d = []
    p = bstart
    while p < bend:
        br = img_words[p] | (img_words[p+1] << 16)
        bc = img_words[p+2] | (img_words[p+3] << 16)
        nnz = img_words[p+4] | (img_words[p+5] << 16)
        nrow, ncoo = img_words[p+6], img_words[p+7]
        d.append((p - bstart, bc))
        if nnz == 0:
            size = 8
        else:
            size = 8 + (nrow - ncoo) + (nnz - ncoo) + 1 + 2 * ncoo
        p += size
    assert p == bend, (p, bend)
    return d

def seq_decode(words, d, begin, pend, carry, log2w):
    out = []
    row = carry
    ti = 0
    for pos in range(begin, pend):
        while ti + 1 < len(d) and d[ti+1][0] <= pos: ti += 1
        ts, bc = d[ti]
        if pos - ts < 8: continue
        w = words[pos]
        if w & 0x8000: row = w & 0x7FFF
        else: out.append((row, (bc << log2w) | w))
    return out, row

def lane_decode(words_abs, bw, d, begin, pend, carry, log2w, R):
    CW = 64 * R
    NB = 1 + (R + 10) // 11
    DWIN = ((CW - 1) // 11 + 2 + 63) // 64 * 64
    abs_first = (bw + begin) & ~7
    abs_end = bw + pend
    # dir_begin: tile containing begin
    ct = max(i for i, e in enumerate(d) if e[0] <= begin)
    dlast = len(d) - 1
    out = []
    c = abs_first
    while c < abs_end:
        win = [d[ct + j] if ct + j <= dlast else (0xFFFFFFFF, 0) for j in range(DWIN)] + [(0xFFFFFFFF, 0)]
        lanes = []
        for l in range(64):
            p0 = (c - bw) + l * R
            q = max(p0, begin)
            j = 0
            st = DWIN // 2
            while st >= 1:
                if win[j + st][0] <= q: j += st
                st //= 2
            ts0 = [win[min(j + b, DWIN)][0] for b in range(NB)]
            tb0 = [win[min(j + b, DWIN)][1] for b in range(NB)]
            lanes.append((p0, ts0, tb0))
        # pass 1
        cnt = []; lrow = []; hh = []
        for l in range(64):
            p0, ts0, tb0 = lanes[l]
            ts, tb = ts0[:], tb0[:]
            n = 0; lr = 0; h = False
            for i in range(R):
                pos = p0 + i
                a = c + l * R + i
                w = words_abs[a] if a < len(words_abs) else 0
                valid = begin <= pos < pend
                if NB > 1 and valid and pos >= ts[1]:
                    ts = ts[1:] + [0xFFFFFFFF]; tb = tb[1:] + [0]
                body = valid and (pos - ts[0]) >= 8
                if body and (w & 0x8000): lr = w & 0x7FFF; h = True
                if body and not (w & 0x8000): n += 1
            cnt.append(n); lrow.append(lr); hh.append(h)
        rows_in = []
        for l in range(64):
            below = [m for m in range(l) if hh[m]]
            rows_in.append(lrow[below[-1]] if below else carry)
        anyh = [m for m in range(64) if hh[m]]
        if anyh: carry = lrow[anyh[-1]]
        # pass 2
        q = []
        for l in range(64):
            p0, ts0, tb0 = lanes[l]
            ts, tb = ts0[:], tb0[:]
            r = rows_in[l]
            for i in range(R):
                pos = p0 + i
                a = c + l * R + i
                w = words_abs[a] if a < len(words_abs) else 0
                valid = begin <= pos < pend
                if NB > 1 and valid and pos >= ts[1]:
                    ts = ts[1:] + [0xFFFFFFFF]; tb = tb[1:] + [0]
                body = valid and (pos - ts[0]) >= 8
                if body and (w & 0x8000): r = w & 0x7FFF
                if body and not (w & 0x8000): q.append((r, (tb[0] << log2w) | w))
        assert len(q) == sum(cnt)
        out += q
        cn = c + CW
        pn = cn - bw
        cw = sum(1 for j in range(DWIN) if win[j][0] <= pn)
        if cn < abs_end: ct += cw - 1
        c = cn
    return out



@pytest.mark.parametrize("scale,lh,lw,R", [(10, 5, 6, 8), (10, 7, 10, 16), (11, 4, 4, 8), (9, 9, 9, 16)])
def test_lane_decode_model(orc, scale, lh, lw, R):
    rng = np.random.default_rng(scale * 100 + lh)
    s, dd = orc.rmat_edges(scale, 16 << scale, scale)
    img = orc.build_2d(s, dd, n=1 << scale, log2h=lh, log2w=lw, flags=orc.DEDUP)
    words = np.frombuffer(img.data, dtype=np.uint16)
    offs = img.offs
    for br in range(0, len(offs) - 1, max(1, (len(offs) - 1) // 24)):
        bs, be = offs[br] // 2, offs[br + 1] // 2
        d = tiles_of(words, bs, be)
        nw = be - bs
        cuts = sorted(set([0, nw] + [int(x) for x in rng.integers(0, nw, size=3)]))
        for a0, a1 in zip(cuts[:-1], cuts[1:]):
            _, carry = seq_decode(words[bs:be], d, 0, a0, -1, lw)
            ref, _ = seq_decode(words[bs:be], d, a0, a1, carry, lw)
            got = lane_decode(words, bs, d, a0, a1, carry, lw, R)
            assert got == ref, (br, a0, a1)


# ---------------------------------------------------------------------------
# Weighted images: decode_chunk_w (values region skipped; value index of an
# entry = its chunk index minus the chunk index at which its tile starts).

def tiles_of_w(img_words, bstart, bend, vw):
    """(start_rel, bc, values_rel) per tile of one weighted block row."""
    d = []
    p = bstart
    while p < bend:
        bc = img_words[p + 2] | (img_words[p + 3] << 16)
        nnz = img_words[p + 4] | (img_words[p + 5] << 16)
        nrow = img_words[p + 6]
        if nnz == 0:
            d.append((p - bstart, bc, p - bstart + 8))
            p += 8
            continue
        vb = p - bstart + 8 + nrow + nnz + 1
        d.append((p - bstart, bc, vb))
        p += 8 + nrow + nnz + 1 + vw * nnz
    assert p == bend
    return d


def seq_decode_w(words, d, begin, pend, log2w, vw):
    out = []
    row = -1
    ti = 0
    nz_in_tile = 0
    for pos in range(begin, pend):
        while ti + 1 < len(d) and d[ti + 1][0] <= pos:
            ti += 1
            nz_in_tile = 0
        ts, bc, vb = d[ti]
        if pos - ts < 8 or pos >= vb:
            continue
        w = words[pos]
        if w & 0x8000:
            row = w & 0x7FFF
        else:
            out.append((row, (bc << log2w) | w, vb + nz_in_tile * vw))
            nz_in_tile += 1
    return out


def lane_decode_w(words_abs, bw, d, begin, pend, log2w, R, vw):
    CW = 64 * R
    NB = 1 + (R + 10) // 11
    DWIN = ((CW - 1) // 11 + 2 + 63) // 64 * 64
    SENT = (0xFFFFFFFF, 0, 0xFFFFFFFF)
    ct = max(i for i, e in enumerate(d) if e[0] <= begin)
    dlast = len(d) - 1
    out = []
    carry, c0 = -1, 0
    c = (bw + begin) & ~7
    while c < bw + pend:
        win = [d[ct + j] if ct + j <= dlast else SENT for j in range(DWIN)] + [SENT]
        lanes = []
        for l in range(64):
            p0 = (c - bw) + l * R
            q = max(p0, begin)
            j = 0
            st = DWIN // 2
            while st >= 1:
                if win[j + st][0] <= q:
                    j += st
                st //= 2
            lanes.append((p0, q, j, [win[min(j + b, DWIN)] for b in range(NB)]))
        wq = {0: -c0}
        cnt, lrow, hh, xs = [], [], [], []
        for l, (p0, q, j, t0) in enumerate(lanes):
            t = list(t0)
            n, lr, h, jc = 0, 0, False, j
            x = [(0, j)] if (j > 0 and t0[0][0] == q and q < pend) else []
            for i in range(R):
                pos, a = p0 + i, c + l * R + i
                w = words_abs[a] if a < len(words_abs) else 0
                valid = begin <= pos < pend
                if NB > 1 and valid and pos >= t[1][0]:
                    t = t[1:] + [SENT]
                    jc += 1
                    x.append((n, jc))
                body = valid and pos - t[0][0] >= 8 and pos < t[0][2]
                if body and w & 0x8000:
                    lr, h = w & 0x7FFF, True
                if body and not w & 0x8000:
                    n += 1
            cnt.append(n); lrow.append(lr); hh.append(h); xs.append(x)
        base = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(int)
        for l in range(64):
            for n_before, jj in xs[l]:
                wq[jj] = int(base[l]) + n_before
        rows_in = []
        for l in range(64):
            below = [m for m in range(l) if hh[m]]
            rows_in.append(lrow[below[-1]] if below else carry)
        anyh = [m for m in range(64) if hh[m]]
        if anyh:
            carry = lrow[anyh[-1]]
        for l, (p0, q, j, t0) in enumerate(lanes):
            t = list(t0)
            r, o, qt = rows_in[l], int(base[l]), wq.get(j)  # unused by lanes without entries
            for i in range(R):
                pos, a = p0 + i, c + l * R + i
                w = words_abs[a] if a < len(words_abs) else 0
                valid = begin <= pos < pend
                if NB > 1 and valid and pos >= t[1][0]:
                    t = t[1:] + [SENT]
                    qt = o
                body = valid and pos - t[0][0] >= 8 and pos < t[0][2]
                if body and w & 0x8000:
                    r = w & 0x7FFF
                if body and not w & 0x8000:
                    assert qt is not None, "value base of this tile was never published"
                    out.append((r, (t[0][1] << log2w) | w, t[0][2] + (o - qt) * vw))
                    o += 1
        qn = int(sum(cnt))
        cn = c + CW
        adv = sum(1 for e in win[:DWIN] if e[0] <= cn - bw)
        if cn < bw + pend:
            # window tile adv-1 holds word cn; one that starts exactly there
            # has no columns in this chunk and no published value base
            c0 = 0 if win[adv - 1][0] == cn - bw else qn - wq[adv - 1]
            ct += adv - 1
        c = cn
    return out


@pytest.mark.parametrize("dt,lh,lw,R", [(np.float32, 5, 6, 8), (np.float64, 7, 10, 16), (np.int32, 4, 4, 8),
                                        (np.int64, 9, 9, 16)])
def test_weighted_lane_decode_model(orc, dt, lh, lw, R):
    rng = np.random.default_rng(lh * 7 + lw)
    s, dd = orc.rmat_edges(10, 16 << 10, lh)
    v = (rng.random(len(s)) * 10).astype(dt)
    img = orc.build_2d_w(s, dd, v, n=1 << 10, log2h=lh, log2w=lw, flags=orc.DEDUP)
    vw = np.dtype(dt).itemsize // 2
    words = np.frombuffer(img.data, dtype=np.uint16)
    offs = img.offs
    for br in range(0, len(offs) - 1, max(1, (len(offs) - 1) // 16)):
        bs, be = offs[br] // 2, offs[br + 1] // 2
        d = tiles_of_w(words, bs, be, vw)
        # items start at tile starts (weighted splits and rotations)
        starts = [e[0] for e in d]
        cuts = sorted(set([0, be - bs] + [starts[int(x)] for x in rng.integers(0, len(starts), size=3)]))
        for a0, a1 in zip(cuts[:-1], cuts[1:]):
            ref = seq_decode_w(words[bs:be], d, a0, a1, lw, vw)
            got = lane_decode_w(words, bs, d, a0, a1, lw, R, vw)
            assert got == ref, (br, a0, a1)


def weighted_image(orc, dt, lh, lw, scale=10):
    rng = np.random.default_rng(lh * 7 + lw)
    s, dd = orc.rmat_edges(scale, 16 << scale, lh)
    v = (rng.random(len(s)) * 10).astype(dt)
    return orc.build_2d_w(s, dd, v, n=1 << scale, log2h=lh, log2w=lw, flags=orc.DEDUP)


@pytest.mark.parametrize("dt,lh,lw", [(np.float32, 4, 4), (np.int64, 6, 5)])
def test_weighted_decode_model_tile_at_chunk_edge(orc, dt, lh, lw):
    """Whole block rows of images in which tiles start exactly on a chunk
    edge right after a chunk that held entries: the next chunk's value
    positions restart at the new tile (no columns carried into it)."""
    img = weighted_image(orc, dt, lh, lw)
    vw = np.dtype(dt).itemsize // 2
    words = np.frombuffer(img.data, dtype=np.uint16)
    offs = img.offs
    for R in (8, 16):
        edges_after_entries = 0
        for br in range(len(offs) - 1):
            bs, be = offs[br] // 2, offs[br + 1] // 2
            d = tiles_of_w(words, bs, be, vw)
            starts = set(e[0] for e in d[1:])
            ref = seq_decode_w(words[bs:be], d, 0, be - bs, lw, vw)
            entry = np.zeros(be - bs, bool)  # column words of the block row
            for ts, _, vb in d:
                body = words[bs + ts + 8:bs + vb]
                entry[ts + 8:vb] = (body & 0x8000) == 0
            for cn in range((bs & ~7) + 64 * R, be, 64 * R):
                if cn - bs in starts and entry[max(0, cn - 64 * R - bs):cn - bs].any():
                    edges_after_entries += 1
            got = lane_decode_w(words, bs, d, 0, be - bs, lw, R, vw)
            assert got == ref, (R, br)
        assert edges_after_entries > 0, R
