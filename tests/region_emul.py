"""CPU emulation of the region executor (sympiler_amd/csrc/region_exec.hip,
k_regions) on the schedule the library builds (symamd_region_schedule_host:
the records, headers and values the device path uploads).

The emulation runs the tasks one after another in rank order and, for every
band of every region, does what the kernel does -- wait, load the band's
halo rows into the LDS image, solve the band's slices from LDS, store the
rows, publish -- while checking the protocol that makes the concurrent
kernel correct:
  * every wait entry names a region of lower rank (tickets are taken in rank
    order, so the lowest unfinished task can always proceed);
  * every halo row a band loads was produced in a band of its region that the
    waits of this band or an earlier one cover (published before the load);
  * every LDS slot a slice reads is final when read: a row of the region
    solved in an earlier slice, a halo row already loaded, or the zero slot;
  * rows, halo ids, slice, band, halo-band and wait bounds are consistent.
The arithmetic is the kernel's: the four packed products in order (padding
products are +0.0 * +0.0), the overflow entries in order, then the division,
without fused multiply-adds -- so the result must equal the serial
sptrsv_csr (sparse_blas/sptrsv.cpp:17-25), or for the transposed schedule
ltsolve (sparse_blas/sptrsv_lt.cpp:11-20), bit for bit.
"""
import ctypes
from ctypes import POINTER, c_int, c_void_p

import numpy as np

PK = 4
RANK_BITS = 26


def rec_layout(nr, nh, ne, ns, nd, nb, vs=8):
    """Byte offsets of a region record's sections (rec_layout, common.h)."""
    L, o = {}, 0
    for name, size in (("pval", PK * nr * vs), ("diag", nr * vs), ("lval", ne * vs),
                       ("code", PK * nr * 2), ("cnt", nr), ("rows", nr * 4), ("halo", nh * 4),
                       ("slices", (ns + 1) * 4), ("bands", (nb + 1) * 4),
                       ("hband", (nb + 1) * 4), ("wptr", (nb + 1) * 4), ("deps", nd * 4),
                       ("lep", (nr + 1) * 4 if ne > 0 else 0), ("lcode", ne * 2)):
        L[name] = o
        o = (o + size + 15) & ~15
    L["bytes"] = o
    return L


def schedule_host(Lp, Li, Lx, cap, band, transposed=False):
    import sympiler_amd as sa
    fn = sa.lib().symamd_region_schedule_host
    fn.argtypes = [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                   c_void_p, c_void_p]
    fn.restype = c_int
    n = len(Lp) - 1
    Lp, Li, Lx = (np.ascontiguousarray(a) for a in (Lp, Li, Lx))
    info = np.zeros(8, np.int32)
    P = lambda a: a.ctypes.data_as(c_void_p)
    rc = fn(n, P(Lp), P(Li), P(Lx), cap, band, int(transposed), P(info), None, None)
    if rc != 0:
        raise RuntimeError(sa.lib().symamd_last_error().decode())
    rec = np.zeros(int(info[4]), np.uint8)
    rhead = np.zeros(8 * int(info[0]), np.int32)
    rc = fn(n, P(Lp), P(Li), P(Lx), cap, band, int(transposed), P(info), P(rec), P(rhead))
    assert rc == 0
    return {"R": int(info[0]), "smax": int(info[1]), "hmax": int(info[2]),
            "recmax": int(info[3]), "slmax": int(info[5]), "levels": int(info[6]),
            "rec": rec, "rhead": rhead}


def _records(sch):
    out = []
    rec = sch["rec"]
    for r in range(sch["R"]):
        nr, nh, ne, ns, nd, unit, units, nb = (int(v) for v in sch["rhead"][8 * r:8 * r + 8])
        L = rec_layout(nr, nh, ne, ns, nd, nb)
        assert L["bytes"] == 16 * units
        blk = rec[16 * unit:16 * (unit + units)]
        f64 = lambda off, cnt: blk[off:off + 8 * cnt].view(np.float64)
        i32 = lambda off, cnt: blk[off:off + 4 * cnt].view(np.int32)
        u16 = lambda off, cnt: blk[off:off + 2 * cnt].view(np.uint16).astype(np.int64)
        out.append(dict(
            nr=nr, nh=nh, ne=ne, ns=ns, nd=nd, nb=nb,
            pval=f64(L["pval"], PK * nr).reshape(nr, PK), diag=f64(L["diag"], nr),
            lval=f64(L["lval"], ne), code=u16(L["code"], PK * nr).reshape(nr, PK),
            cnt=blk[L["cnt"]:L["cnt"] + nr].astype(np.int64), rows=i32(L["rows"], nr),
            halo=i32(L["halo"], nh), slices=i32(L["slices"], ns + 1),
            bands=i32(L["bands"], nb + 1), hband=i32(L["hband"], nb + 1),
            wptr=i32(L["wptr"], nb + 1), deps=i32(L["deps"], nd),
            lep=i32(L["lep"], nr + 1) if ne > 0 else None, lcode=u16(L["lcode"], ne)))
    return out


def emulate(sch, n, rhs):
    """Runs the schedule on right-hand side `rhs` (length n), checking the
    protocol; returns the solution."""
    R, smax, hmax = sch["R"], sch["smax"], sch["hmax"]
    zero = smax + hmax
    recs = _records(sch)
    # every row once; each row's region rank and band
    row_rank = np.full(n, -1, np.int64)
    row_band = np.full(n, -1, np.int64)
    for r, c in enumerate(recs):
        assert 0 < c["nr"] <= smax and c["nh"] <= hmax and c["nb"] >= 1
        sl, bd = c["slices"], c["bands"]
        assert sl[0] == 0 and sl[-1] == c["nr"] and np.all(np.diff(sl) > 0)
        assert bd[0] == 0 and bd[-1] == c["ns"] and np.all(np.diff(bd) > 0)
        assert c["hband"][0] == 0 and c["hband"][-1] == c["nh"] and np.all(np.diff(c["hband"]) >= 0)
        assert c["wptr"][0] == 0 and c["wptr"][-1] == c["nd"] and np.all(np.diff(c["wptr"]) >= 0)
        assert np.all(row_rank[c["rows"]] == -1)
        row_rank[c["rows"]] = r
        for b in range(c["nb"]):
            row_band[c["rows"][sl[bd[b]]:sl[bd[b + 1]]]] = b
    assert np.all(row_rank >= 0)
    X = np.array(rhs, np.float64, copy=True)
    published = np.zeros(R, np.int64)
    for r, c in enumerate(recs):
        nr = c["nr"]
        xs = np.full(zero + 1, np.nan)
        xs[zero] = 0.0
        ready = np.zeros(zero + 1, bool)
        ready[zero] = True
        solved_in = np.full(nr, -1)
        waited = {}
        B = X[c["rows"]].copy()  # the right-hand sides of its rows
        for b in range(c["nb"]):
            for e in c["deps"][c["wptr"][b]:c["wptr"][b + 1]]:
                u, need = int(e) & ((1 << RANK_BITS) - 1), int(e) >> RANK_BITS
                assert u < r, "a wait on a region of equal or higher rank"
                assert 1 <= need <= recs[u]["nb"]
                assert published[u] >= need
                waited[u] = max(waited.get(u, 0), need)
            h0, h1 = c["hband"][b], c["hband"][b + 1]
            for q in range(h0, h1):
                h = int(c["halo"][q])
                u = int(row_rank[h])
                assert u != r, "a halo row of the region itself"
                assert waited.get(u, 0) >= row_band[h] + 1, "halo row loaded before its band is waited for"
                xs[smax + q] = X[h]
                ready[smax + q] = True
            sl = c["slices"]
            for s in range(c["bands"][b], c["bands"][b + 1]):
                t = np.arange(sl[s], sl[s + 1])
                codes = c["code"][t]
                own = codes < smax
                assert np.all(ready[codes] | own), "a slot read before it is loaded"
                if own.any():
                    assert np.all(codes[own] < nr)
                    assert np.all(solved_in[codes[own]] >= 0) and np.all(solved_in[codes[own]] < s)
                acc = B[t].copy()
                for u in range(PK):
                    acc = acc - c["pval"][t, u] * xs[codes[:, u]]
                for k, tt in enumerate(t):
                    if c["cnt"][tt] > PK:
                        for e in range(c["lep"][tt], c["lep"][tt + 1]):
                            cd = int(c["lcode"][e])
                            assert ready[cd] or (cd < nr and 0 <= solved_in[cd] < s)
                            acc[k] = acc[k] - c["lval"][e] * xs[cd]
                xs[t] = acc / c["diag"][t]
                solved_in[t] = s
                ready[t] = True
            r0, r1 = sl[c["bands"][b]], sl[c["bands"][b + 1]]
            X[c["rows"][r0:r1]] = xs[r0:r1]
            published[r] = b + 1
    return X
