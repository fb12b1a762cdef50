"""A CPU model of the tiled LLP parser (alpacc_amd/csrc/parse_seg.hip),
checked against the oracle with small tiles.

The model runs the kernels' algorithm step by step -- pass A tile profiles,
the tile scans, pass B's lane-sequential window effects with lane stacks,
the cross-lane sparse-table search and the open/unmatched lists, the far
search over tiles of the same input, and the per-input verdicts -- with
assertions on every stack and list index.  Small tiles (a few positions per
lane, a handful of lanes) make every cross-lane and cross-tile path run on
short inputs.  It pins the algorithm; the GPU tests pin the kernels."""
import random
import struct
import zlib

import pytest
from alpacc_amd.gen import tables, emit_hip
import pyoracle

PT, APL, RWG, RLV = 8, 2, 4, 2


def configure(pt, apl):
    global PT, APL, RWG, RLV
    PT, APL, RWG = pt, apl, pt // apl
    RLV = max(1, (RWG - 1).bit_length())
BAD_PAIR, BAD_BDEPTH, BAD_MATCH, BAD_COUNT, BAD_ROOT, BAD_SDEPTH = 1, 2, 4, 8, 16, 32
TT = 1 << 62

def load(name):
    g = tables.load(f'grammars/{name}.alp')
    d = emit_hip.device_tables(g)
    return g, d

def rec_of(d, i):
    r = d['pair_rec'][64 * i:64 * i + 64]
    return dict(bpop=r[0], bpush=r[1], np=r[2], nt=r[3], spop=r[4], spush=r[5], tmask=r[6],
                br=[r[8 + 2 * j] | (r[9 + 2 * j] << 8) for j in range(8)], prods=r[24:32], par=r[32:40], own=r[40:64])

def run(g, d, strings, tree):
    p = g.parser
    base = p.end_t + 1; q, k = p.q, p.k
    count = len(strings)
    tok = [t for s in strings for t in s]
    tok_off = [0]
    for s in strings: tok_off.append(tok_off[-1] + len(s))
    tile_first = []; nt_ = 0
    for s in strings:
        tile_first.append(nt_); nt_ += (len(s) + 2 + PT - 1) // PT
    tile_first.append(nt_); ntiles = nt_
    tile_seg = []
    for t in range(ntiles):
        i = max(j for j in range(count) if tile_first[j] <= t)
        tile_seg.append(i)
    flags = [0] * count
    def tile_of(t):
        sg = tile_seg[t]; lt = t - tile_first[sg]
        m = tok_off[sg + 1] - tok_off[sg] + 2
        return sg, lt, m, tok[tok_off[sg]:tok_off[sg + 1]]
    def pair_at(m, tk, pp):
        code = 0
        for j in range(q + k):
            f = pp - q + j
            v = p.empty_t if (f < 0 or f >= m) else (p.start_t if f == 0 else (p.end_t if f == m - 1 else tk[f - 1]))
            if v >= base: return None
            code = code * base + v
        w = d['window'][code]
        return None if w == 0xFFFF else w
    def sm_cat(a, b): return (a[0] + b[0], min(a[1], a[0] + b[1]))
    # pass A
    trec = []
    for t in range(ntiles):
        sg, lt, m, tk = tile_of(t)
        np_ = nt = 0; B = (0, 0); S = (0, 0)
        for lane in range(PT):
            pp = lt * PT + lane
            if pp >= m: continue
            i = pair_at(m, tk, pp)
            if i is None: flags[sg] |= BAD_PAIR; continue
            r = rec_of(d, i)
            np_ += r['np']; nt += r['nt']
            B = sm_cat(B, (r['bpush'] - r['bpop'], -r['bpop'])); S = sm_cat(S, (r['spush'] - r['spop'], -r['spop']))
        trec.append(dict(np=np_, nt=nt, bs=B[0], bm=B[1], ss=S[0], sm=S[1]))
    # pass S
    info = []; sb = {}
    acc = [0] * 8
    for t in range(ntiles):
        r = trec[t]; first = tile_first[tile_seg[t]] == t
        ti = dict(p_off=acc[0], t_off=acc[1], bstart=acc[2], sstart=acc[3] + (1 if first else 0), bmin=r['bm'], smin=r['sm'],
                  ob_off=acc[4], ub_off=acc[5], os_off=acc[6], us_off=acc[7])
        info.append(ti)
        if first: sb[tile_seg[t]] = dict(p_off=ti['p_off'], t_off=ti['t_off'], bstart=ti['bstart'], sstart=ti['sstart'])
        x = [r['np'], r['nt'], r['bs'], r['ss'] + (1 if first else 0), r['bs'] - r['bm'], -r['bm'], r['ss'] - r['sm'], -r['sm']]
        acc = [a + b for a, b in zip(acc, x)]
    so = dict(np=acc[0], nt=acc[1], bend=acc[2], cend=acc[3], ob=acc[4], ub=acc[5], os=acc[6], us=acc[7])
    mb = [info[t]['bstart'] + trec[t]['bm'] for t in range(ntiles)]
    ms = [info[t]['sstart'] + trec[t]['sm'] for t in range(ntiles)]
    ob = [None] * so['ob']; ub = [None] * so['ub']; os_ = [None] * so['os']; us = [None] * so['us']
    nodes = so['np'] - so['nt']
    tree_prod = [None] * nodes; tree_par = [None] * nodes; tok_par = [None] * so['nt']; deriv = [None] * so['np']
    # pass B (lane-sequential, APL positions per lane)
    bcap = APL * (d['max_bpop'] + d['max_bpush']); scap = APL * (d['max_spop'] + d['max_spush'])
    for t in range(ntiles):
        sg, lt, m, tk = tile_of(t)
        ti = info[t]; inb = sb[sg]
        R = []
        for pos in range(PT):
            pp = lt * PT + pos
            i = pair_at(m, tk, pp) if pp < m else None
            R.append(rec_of(d, i) if i is not None else dict(bpop=0, bpush=0, np=0, nt=0, spop=0, spush=0, tmask=0, br=[0]*8, prods=[0]*8, par=[0]*8, own=[0]*24))
        lane_v = []
        for ln in range(RWG):
            rs = R[ln * APL:(ln + 1) * APL]
            lane_v.append([sum(r['np'] for r in rs), sum(r['nt'] for r in rs), sum(r['bpush'] - r['bpop'] for r in rs), sum(r['spush'] - r['spop'] for r in rs)])
        ex = []; run_ = [0, 0, 0, 0]
        for x in lane_v: ex.append(list(run_)); run_ = [a_ + b_ for a_, b_ in zip(run_, x)]
        tile_np, tile_nt = run_[0], run_[1]
        c_off = ti['p_off'] - ti['t_off']; c_in = inb['p_off'] - inb['t_off']
        s_par = {}; s_tpar = {}; s_prod = {}
        BST = [[None] * bcap for _ in range(RWG)]; SST = [[None] * scap for _ in range(RWG)]
        L = []
        for ln in range(RWG):
            slot0, term0, bS, sS = ex[ln]; cb0 = slot0 - term0
            bsp = bun = ssp = sun = 0; bdep = bmin = sdep = smin = 0; slot = slot0; tinc = term0
            for pk in range(APL):
                r = R[ln * APL + pk]
                for e in range(r['bpop']):
                    sym = r['br'][e]
                    if bsp > 0:
                        bsp -= 1
                        if BST[ln][bsp] != sym: flags[sg] |= BAD_MATCH
                    else:
                        assert bcap - 1 - bun >= bsp, 'bstack collide'
                        BST[ln][bcap - 1 - bun] = sym; bun += 1
                bdep -= r['bpop']; bmin = min(bmin, bdep)
                for e in range(r['bpush']):
                    assert bsp < bcap - bun, 'bstack full'
                    BST[ln][bsp] = r['br'][4 + e]; bsp += 1
                bdep += r['bpush']
                cbk = slot - tinc; tseen = 0
                for j in range(r['np']):
                    pr = r['prods'][j]; term = (r['tmask'] >> j) & 1
                    c = cbk + j - tseen; tl = tinc + tseen; tseen += term
                    if not tree: s_prod[slot + j] = pr
                    elif not term: s_prod[c] = pr
                    code = r['par'][j]; par = -1
                    if code & 0x80:
                        if ssp > 0:
                            ssp -= 1; par = cb0 + SST[ln][ssp]
                        else:
                            v_ = (0x80 | (tl - term0)) if term else (c - cb0)
                            assert 0 <= (tl - term0 if term else c - cb0) < 0x80
                            assert scap - 1 - sun >= ssp, 'sstack collide'
                            SST[ln][scap - 1 - sun] = v_; sun += 1
                        sdep -= 1; smin = min(smin, sdep)
                    else:
                        par = cbk + code
                    if tree and par >= 0:
                        if term: s_tpar[tl] = par
                        else: s_par[c] = par
                for e in range(r['spush']):
                    assert ssp < scap - sun, 'sstack full'
                    o = cbk - cb0 + r['own'][e]; assert 0 <= o < 256
                    SST[ln][ssp] = o; ssp += 1
                sdep += r['spush']; slot += r['np']; tinc += r['nt']
            L.append(dict(bS=bS, sS=sS, cb0=cb0, term0=term0, bsp=bsp, bun=bun, ssp=ssp, sun=sun, lbm=bS + bmin, lsm=sS + smin))
        mbl = [x['lbm'] for x in L]; msl = [x['lsm'] for x in L]
        def tab(a):
            tb = [list(a)]
            for kk in range(1, RLV):
                h = 1 << (kk - 1)
                tb.append([min(tb[kk - 1][x], tb[kk - 1][x - h]) if x >= h else tb[kk - 1][x] for x in range(RWG)])
            return tb
        tbb, tbs = tab(mbl), tab(msl)
        def prev_below(tb, fr, Lv):
            l = fr
            if l < 0 or tb[0][l] < Lv: return l
            for kk in range(RLV - 1, -1, -1):
                if l >= 0 and tb[kk][l] >= Lv: l -= 1 << kk
            return l
        def sufmin(a, l): return min(a[l + 1:]) if l + 1 < RWG else 0x7FFFFFFF
        # previous strictly smaller lane minimum (-1: none), as the kernel finds
        # it: in-wave sparse tables (tile = 4 waves of WM lanes here), in-wave
        # prefix minima and wave minima
        WM = max(1, RWG // 4); WL = (WM - 1).bit_length(); assert RWG % WM == 0
        def wtab(a):
            tb = [list(a)]
            for kk in range(1, WL):
                h = 1 << (kk - 1)
                tb.append([min(tb[kk - 1][x], tb[kk - 1][x - h]) if x % WM >= h else tb[kk - 1][x] for x in range(RWG)])
            pw = [min(a[x - x % WM:x + 1]) for x in range(RWG)]
            wm = [min(a[v * WM:(v + 1) * WM]) for v in range(RWG // WM)]
            return tb, pw, wm
        def wprev(T, fr, Lv):
            tb, pw, wm = T
            def descend(l):
                for kk in range(WL - 1, -1, -1):
                    assert 0 <= l < RWG
                    if tb[kk][l] >= Lv: l -= 1 << kk
                assert 0 <= l and tb[0][l] < Lv
                return l
            if fr < 0: return -1
            if tb[0][fr] < Lv: return fr
            if pw[fr] < Lv: return descend(fr)
            for v in range(fr // WM - 1, -1, -1):
                if wm[v] < Lv: return descend(v * WM + WM - 1)
            return -1
        wtb, wts = wtab(mbl), wtab(msl)
        psb = [wprev(wtb, x - 1, mbl[x]) for x in range(RWG)]
        pss = [wprev(wts, x - 1, msl[x]) for x in range(RWG)]
        for x in range(RWG):
            for ps_, m_ in ((psb, mbl), (pss, msl)):
                assert ps_[x] < x and (ps_[x] < 0 or m_[ps_[x]] < m_[x])
                assert all(m_[y] >= m_[x] for y in range(max(ps_[x], -1) + 1, x))
        def jump(ps_, m_, l, D):
            # last lane <= l whose minimum is below D: lanes between l and its
            # previous smaller lane have minima >= the minimum at l
            while l >= 0 and m_[l] >= D:
                l = ps_[l]
            return l
        for ln in range(RWG):
            X = L[ln]
            l = ln - 1
            for u in range(X['bun']):
                D = X['bS'] - u; sym = BST[ln][bcap - 1 - u]
                l = jump(psb, mbl, l, D)
                assert l == max(-1, prev_below(tbb, ln - 1, D))
                if l >= 0:
                    idx = D - tbb[0][l] - 1; assert 0 <= idx < L[l]['bsp'], ('bst idx', idx)
                    if BST[l][idx] != sym: flags[sg] |= BAD_MATCH
                else:
                    j = ti['ub_off'] + (-D); assert ti['ub_off'] <= j < ti['ub_off'] - ti['bmin'], 'ub'
                    ub[j] = sym
            bsuf = sufmin(mbl, ln)
            for e in range(X['bsp']):
                lv = X['lbm'] + 1 + e
                if lv <= bsuf:
                    j = ti['ob_off'] + (lv - ti['bmin'] - 1); assert ti['ob_off'] <= j < ti['ob_off'] + trec[t]['bs'] - trec[t]['bm'], 'ob'
                    ob[j] = BST[ln][e]
            if tree:
                l = ln - 1
                for u in range(X['sun']):
                    D = X['sS'] - u; tg = SST[ln][scap - 1 - u]
                    l = jump(pss, msl, l, D)
                    assert l == max(-1, prev_below(tbs, ln - 1, D))
                    term = tg & 0x80; idx = (X['term0'] if term else X['cb0']) + (tg & 0x7F)
                    if l >= 0:
                        e_ = D - tbs[0][l] - 1; assert 0 <= e_ < L[l]['ssp'], ('sst idx', e_)
                        pc = L[l]['cb0'] + SST[l][e_]
                        if term: s_tpar[idx] = pc
                        else: s_par[idx] = pc
                    else:
                        jj = ti['us_off'] + (-D); assert ti['us_off'] <= jj < ti['us_off'] - ti['smin'], 'us'
                        us[jj] = (TT | (ti['t_off'] + idx)) if term else (c_off + idx)
                ssuf = sufmin(msl, ln)
                for e in range(X['ssp']):
                    lv = X['lsm'] + 1 + e
                    if lv <= ssuf:
                        j = ti['os_off'] + (lv - ti['smin'] - 1); assert ti['os_off'] <= j < ti['os_off'] + trec[t]['ss'] - trec[t]['sm'], 'os'
                        os_[j] = c_off + X['cb0'] + SST[ln][e]
        if ti['bstart'] + ti['bmin'] < inb['bstart']: flags[sg] |= BAD_BDEPTH
        if ti['sstart'] + ti['smin'] < inb['sstart'] - 1: flags[sg] |= BAD_SDEPTH
        if tree:
            for i in range(tile_np - tile_nt):
                tree_prod[c_off + i] = s_prod.get(i); tree_par[c_off + i] = (c_off + s_par[i] - c_in) if i in s_par else 'far'
            for i in range(tile_nt):
                tok_par[ti['t_off'] + i] = (c_off + s_tpar[i] - c_in) if i in s_tpar else 'far'
        else:
            for i in range(tile_np): deriv[ti['p_off'] + i] = s_prod[i]
    # pass C
    def prev_less_tile(arr, t, v):
        for t2 in range(t - 1, -1, -1):
            if arr[t2] < v: return t2
        return -1
    for t in range(ntiles):
        ti = info[t]; sg = tile_seg[t]; first = tile_first[sg]; inb = sb[sg]; c_in = inb['p_off'] - inb['t_off']
        for u in range(-ti['bmin']):
            D = ti['bstart'] - u; b = ub[ti['ub_off'] + u]
            t2 = prev_less_tile(mb, t, D)
            if t2 < first: flags[sg] |= BAD_MATCH; continue
            o = info[t2]; j = o['ob_off'] + (D - o['bstart'] - o['bmin'] - 1)
            assert o['ob_off'] <= j < o['ob_off'] + trec[t2]['bs'] - trec[t2]['bm'], 'far ob'
            if ob[j] != b: flags[sg] |= BAD_MATCH
        if tree:
            for u in range(-ti['smin']):
                D = ti['sstart'] - u; target = us[ti['us_off'] + u]
                t2 = prev_less_tile(ms, t, D)
                if t2 < first:
                    if D != inb['sstart']: flags[sg] |= BAD_ROOT
                    pc = 0
                else:
                    o = info[t2]; j = o['os_off'] + (D - o['sstart'] - o['smin'] - 1)
                    assert o['os_off'] <= j < o['os_off'] + trec[t2]['ss'] - trec[t2]['sm'], 'far os'
                    pc = os_[j] - c_in
                if target & TT: tok_par[target & ~TT] = pc
                else: tree_par[target] = pc
    # pass F
    out = []
    for i in range(count):
        f0, f1 = tile_first[i], tile_first[i + 1]
        inb = sb[i]
        if f1 < ntiles:
            e = info[f1]; p_end, t_end, b_end, c_end = e['p_off'], e['t_off'], e['bstart'], e['sstart'] - 1
        else:
            p_end, t_end, b_end, c_end = so['np'], so['nt'], so['bend'], so['cend']
        fl = flags[i]
        if b_end != inb['bstart']: fl |= BAD_BDEPTH
        if c_end != inb['sstart'] - 1: fl |= BAD_SDEPTH
        if t_end - inb['t_off'] != tok_off[i + 1] - tok_off[i]: fl |= BAD_COUNT
        if not tree: fl &= BAD_PAIR | BAD_BDEPTH | BAD_MATCH
        if fl:
            out.append(None); continue
        if tree:
            c0 = inb['p_off'] - inb['t_off']; N = (p_end - inb['p_off']) - (t_end - inb['t_off'])
            out.append((tree_prod[c0:c0 + N], tree_par[c0:c0 + N], tok_par[inb['t_off']:t_end]))
        else:
            out.append(deriv[inb['p_off']:p_end])
    return out

def oracle_parser(gname, toks):
    rec = pyoracle.record(gname, 'parser', bytes(toks))
    if rec[0] == 0: return None
    n = struct.unpack_from('<Q', rec, 1)[0]
    return list(rec[9:9 + n])

def oracle_tree(gname, text):
    rec = pyoracle.record(gname, 'both', text)
    if rec[0] == 0: return None, None
    T = struct.unpack_from('<Q', rec, 1)[0]; o = 9
    ids = list(rec[o:o + T]); o += T
    o += 8 * T  # starts
    g = tables.load(f'grammars/{gname}.alp'); lb = g.length_bytes
    o += lb * T
    N = struct.unpack_from('<Q', rec, o)[0]; o += 8
    prods = list(rec[o:o + N]); o += N
    par = list(struct.unpack_from(f'<{N}q', rec, o)); o += 8 * N
    tp = list(struct.unpack_from(f'<{T}q', rec, o))
    return ids, (prods, par, tp)


GRAMMARS = ["arithmetic", "json", "sexp", "paper_grammar"]


def _cases(gname, g, rng):
    from alpacc_amd import datagen
    from test_gpu_parity import nested_text
    p = g.parser
    ids = sorted(set(g.lex_both.term_id.values()) - {g.lex_both.ignore_id})
    strs = [[rng.choice(ids + [p.empty_t, p.start_t]) for _ in range(rng.randrange(0, 12))] for _ in range(30)]
    if gname == "paper_grammar":
        texts = [b"1+[2+3]", b"[[1]+2]", b"1+", b""]
    else:
        texts = [datagen.generate(gname, n, seed=n, conforming=True).tobytes() for n in (1, 5, 40, 300)]
        for _ in range(3):
            t = nested_text(gname, rng, rng.choice([50, 400]), rng.choice([5, 60]))
            texts.append({"json": b"[" + t + b"]", "sexp": b"(" + t + b")"}.get(gname, t))
    lexed = [oracle_tree(gname, tx) for tx in texts]
    return strs, [x for x in lexed if x[0] is not None]


@pytest.mark.parametrize("pt,apl", [(8, 2), (16, 4), (12, 1), (64, 2)])
@pytest.mark.parametrize("gname", GRAMMARS)
def test_tile_model_matches_oracle(gname, pt, apl):
    configure(pt, apl)
    rng = random.Random(zlib.crc32(f"{gname} {pt} {apl}".encode()))
    g, d = load(gname)
    strs, good = _cases(gname, g, rng)
    tok_strs = strs + [x[0] for x in good]
    got = run(g, d, tok_strs, tree=False)  # one batch: flat derivations (parse_test)
    for s, gg in zip(tok_strs, got):
        assert gg == oracle_parser(gname, s), (gname, s[:20])
    got = run(g, d, [x[0] for x in good], tree=True)  # one batch: compact trees
    for (_, tr), gg in zip(good, got):
        assert gg is not None and tuple(gg) == tuple(tr)
    mixed = strs[:10] + [x[0] for x in good]  # rejected and accepted strings in one batch
    for s, gg in zip(mixed, run(g, d, mixed, tree=True)):
        if pyoracle.record(gname, "parser", bytes(s))[0] == 0:
            assert gg is None
