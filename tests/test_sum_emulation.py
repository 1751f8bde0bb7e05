"""CPU emulation of the exact ordered reduction (cloth_sim_self_collision_amd/
csrc/cs_ordered_sum.hpp: k_sum_tiles, k_sum_exact, wave_walk, wave_burst),
step for step and lane by lane, recording every index the kernels read or
write.  It checks, on the adversarial inputs of the GPU test of the same
reduction:

  - every global index stays inside its buffer: addends t[0, n), tile
    summaries [0, sum_tiles(n)); every LDS index inside its array: the staged
    tile buf[0, kSumBuf), the summary window [0, kChainWin);
  - the emulated result equals the left-to-right loop of the reference's
    InnerProduct (matrix_square_sparse.cpp:313-325) bit for bit, which checks
    the emulation's reading of the algorithm.

Integer arithmetic wraps at 64 bits as on the device.  The approximate tile
sums only steer the binade guesses (a wrong guess costs time, never the
result), so they are formed here with numpy's sum.
"""
import math

import numpy as np
import pytest

# constants of cs_ordered_sum.hpp
K_SUM_BLOCK = 256
K_SUM_PER = 4
K_SUM_TILE = K_SUM_BLOCK * K_SUM_PER
K_WALK_PER = 4
K_WALK_CHUNK = 64 * K_WALK_PER
K_SUM_BUF = K_SUM_TILE + K_WALK_CHUNK
K_CHAIN_WIN = 1024
K_BURST_MIN = 32
K_BURST_QUIET = 32
K_BURST_FINISH = 64
DBL_MIN = 2.2250738585072014e-308
DBL_MAX = 1.7976931348623157e308
TWO52 = 4503599627370496.0
TWO62 = 4611686018427387904.0
M64 = (1 << 64) - 1


def wrap64(x: int) -> int:
    x &= M64
    return x - (1 << 64) if x >> 63 else x


def sum_tiles(n):
    return (n + K_SUM_TILE - 1) // K_SUM_TILE if n > 0 else 1


def sign_exp(x: float) -> int:
    return int(np.float64(x).view(np.uint64)) >> 52


def ilogb(x: float) -> int:
    return math.frexp(x)[1] - 1


def sum_normal(s: float) -> bool:
    return DBL_MIN <= abs(s) <= DBL_MAX


def rint(q: float) -> float:
    return float(np.rint(q))  # round half to even, as the device rint


class Bounds:
    """Largest and smallest index seen per buffer."""

    def __init__(self):
        self.lo, self.hi = {}, {}

    def see(self, name, idx):
        idx = np.asarray(idx)
        if idx.size == 0:
            return
        self.lo[name] = min(self.lo.get(name, 1 << 62), int(idx.min()))
        self.hi[name] = max(self.hi.get(name, -(1 << 62)), int(idx.max()))

    def check(self, name, size):
        if name in self.lo:
            assert self.lo[name] >= 0 and self.hi[name] < size, (name, self.lo[name], self.hi[name], size)


def tile_summary(t, n, b, tsum, bnd):
    """k_sum_exact's per-block part: binade guess and integer summary."""
    bnd.see("tsum", np.arange(b))
    p = float(np.sum(tsum[:b])) if b else 0.0
    code, scale = -1, 0
    if DBL_MIN <= abs(p) <= DBL_MAX:
        code, scale = sign_exp(p), 52 - ilogb(p)
    ok = code >= 0
    runs, prefs = [], []
    for tid in range(K_SUM_BLOCK):
        run, pref, tok = 0, [], ok
        for j in range(K_SUM_PER):
            idx = b * K_SUM_TILE + tid * K_SUM_PER + j
            if idx < n:
                bnd.see("t", idx)
            x = float(t[idx]) if idx < n else 0.0
            if tok:
                q = math.ldexp(x, scale) if math.isfinite(x) else x
                r = rint(q) if math.isfinite(q) else q
                if not (abs(q) < TWO62) or abs(q - r) == 0.5:
                    tok = False
                else:
                    run = wrap64(run + int(r))
            pref.append(run)
        runs.append(run)
        prefs.append(pref)
        ok = ok and tok
    total, off = 0, []
    for r in runs:
        off.append(total)
        total = wrap64(total + r)
    allp = [wrap64(off[tid] + prefs[tid][j]) for tid in range(K_SUM_BLOCK) for j in range(K_SUM_PER)]
    return (code if ok else -1), total, min(allp), max(allp)


def wave_burst(buf, m, len_, s, bnd):
    k, quiet = m, m
    bnd.see("buf", np.arange(k, k + 16))
    w = [buf[k + j] for j in range(16)]
    while True:
        hb = sign_exp(s)
        s = s + w[0]
        bnd.see("buf", np.arange(k + 16, k + 32))
        wn = [buf[k + 16 + j] for j in range(16)]
        if len_ - k >= 16:
            for j in range(1, 16):
                s = s + w[j]
            k += 16
        else:
            for j in range(1, 16):
                if k + j < len_:
                    s = s + w[j]
            k = len_
        if sign_exp(s) != hb:
            quiet = k
        if k >= len_ or (k - m >= K_BURST_MIN and k - quiet >= K_BURST_QUIET and len_ - k > K_BURST_FINISH):
            return k, s
        w = wn


def wave_walk(buf, len_, s, bnd):
    lo52, hi53 = TWO52 + 1.0, 2.0 * TWO52 - 1.0
    lanes = np.arange(64)
    pos = 0
    while pos < len_:
        if math.isnan(s):
            break
        m = len_
        if sum_normal(s):
            scale = 52 - ilogb(s)
            sg = -1.0 if s < 0.0 else 1.0
            S = math.ldexp(s, scale)
            Sm = 0.0
            c0 = pos & ~(K_WALK_PER - 1)
            while c0 < len_:
                i0 = c0 + K_WALK_PER * lanes
                idx = i0[:, None] + np.arange(K_WALK_PER)[None, :]
                bnd.see("buf", idx)
                x = buf[idx]
                use = (idx >= pos) & (idx < len_)
                with np.errstate(over="ignore", invalid="ignore"):
                    q = np.ldexp(x, scale)
                    r = np.rint(q)
                    conv = (np.abs(q) < TWO52) & (np.abs(q - r) != 0.5)
                add = np.where(use & conv, r, 0.0)
                inc = np.cumsum(add, axis=1)  # exact where it is used (see the header)
                run = inc[:, -1]
                excl = np.concatenate([[0.0], np.cumsum(run)[:-1]])
                off = S + excl
                fail = use & ~conv
                with np.errstate(invalid="ignore"):
                    a = sg * (off[:, None] + inc)
                fail |= use & ((a < lo52) | (a > hi53))
                has = fail.any(axis=1)
                if has.any():
                    f = int(np.argmax(has))
                    jf = int(np.argmax(fail[f]))
                    before = off[f] if jf == 0 else off[f] + inc[f, jf - 1]
                    Sm = float(before)
                    m = int(i0[f] + jf)
                    break
                S = float(off[63] + run[63])
                c0 += K_WALK_CHUNK
            s = math.ldexp(Sm if m < len_ else S, -scale)
        elif s == 0.0 or math.isinf(s):
            c0 = pos & ~3
            while c0 < len_:
                i0 = c0 + 4 * lanes
                idx = i0[:, None] + np.arange(4)[None, :]
                bnd.see("buf", idx)
                x = buf[idx]
                if s == 0.0:
                    keep = x == 0.0
                else:
                    keep = ~(np.isnan(x) | (np.isinf(x) & ((x < 0.0) != (s < 0.0))))
                bad = (idx >= pos) & (idx < len_) & ~keep
                has = bad.any(axis=1)
                if has.any():
                    f = int(np.argmax(has))
                    m = int(i0[f] + int(np.argmax(bad[f])))
                    break
                c0 += 4 * 64
        else:
            m = pos
        if m >= len_:
            break
        pos, s = wave_burst(buf, m, len_, s, bnd)
    return s


def emulate(t):
    """The two kernels of one reduction; returns (sum, index bounds)."""
    t = np.asarray(t, dtype=np.float64)
    n = len(t)
    nt = sum_tiles(n)
    bnd = Bounds()
    # k_sum_tiles: approximate tile sums (any order)
    tsum = np.array([float(np.sum(t[b * K_SUM_TILE:(b + 1) * K_SUM_TILE])) if n else 0.0
                     for b in range(nt)])
    summ = [tile_summary(t, n, b, tsum, bnd) for b in range(nt)]
    bnd.see("summary", np.arange(nt))
    # chain on one wave of the last block
    s = 0.0
    buf = np.zeros(K_SUM_BUF)
    b0 = 0
    while b0 < nt and not math.isnan(s):
        wend = min(nt, b0 + K_CHAIN_WIN)
        for j0 in range(b0, wend, 256):
            j = np.minimum(j0 + np.arange(256), wend - 1)
            bnd.see("summary", j)
            bnd.see("win", j - b0)
        bb = b0
        while bb < wend and not math.isnan(s):
            while bb < wend and sum_normal(s):
                e = ilogb(s)
                S = int(math.ldexp(s, 52 - e))
                k = bb + np.arange(64)
                inn = k < wend
                kk = np.minimum(k, wend - 1) - b0
                bnd.see("win", kk)
                tot = [summ[b0 + q][1] if i else 0 for q, i in zip(kk, inn)]
                incl, acc = [], 0
                for v in tot:
                    acc = wrap64(acc + v)
                    incl.append(acc)
                f = 64
                for lane in range(64):
                    base = wrap64(S + incl[lane] - tot[lane])
                    lo = wrap64(base + summ[b0 + kk[lane]][2])
                    hi = wrap64(base + summ[b0 + kk[lane]][3])
                    inside = (lo >= (1 << 52) + 1 and hi <= (1 << 53) - 1) if s > 0.0 else \
                             (lo >= -((1 << 53) - 1) and hi <= -((1 << 52) + 1))
                    if not (inn[lane] and summ[b0 + kk[lane]][0] == sign_exp(s) and inside):
                        f = lane
                        break
                if f > 0:
                    s = math.ldexp(float(wrap64(S + incl[f - 1])), e - 52)
                bb += f
                if f < 64:
                    break
            if bb >= wend or math.isnan(s):
                break
            begin = bb * K_SUM_TILE
            len_ = min(n, begin + K_SUM_TILE) - begin
            if len_ > 0:
                src = begin + np.minimum(np.arange(K_SUM_TILE), len_ - 1)
                bnd.see("t", src)
                buf[:K_SUM_TILE] = t[src]
                s = wave_walk(buf, len_, s, bnd)
            bb += 1
        b0 += K_CHAIN_WIN
    bnd.check("t", n)
    bnd.check("tsum", nt)
    bnd.check("summary", nt)
    bnd.check("win", K_CHAIN_WIN)
    bnd.check("buf", K_SUM_BUF)
    return s, bnd


def _sequential(t):
    s = 0.0
    for x in t:
        s += float(x)
    return s


def _cases():
    """The inputs of the device test of this reduction (smaller where the
    emulation would be slow), plus short vectors whose last tile is shorter
    than the staging width."""
    rng = np.random.default_rng(11)
    out = []
    for n in (1, 2, 17, 255, 1000, 1023, 1024, 1025, 16385, 70000):
        out.append(rng.standard_normal(n))
        out.append(rng.standard_normal(n) ** 2 * np.exp2(rng.integers(-30, 30, n)))
        out.append(np.round(rng.standard_normal(n) * 8) / 4)
        out.append(np.concatenate([np.zeros(n // 2), rng.standard_normal(n - n // 2)]))
        t = rng.standard_normal(n)
        t[::7] *= -1e10
        out.append(t)
        out.append(np.exp2(rng.integers(-60, 60, n)).astype(float) * rng.choice([-1, 1], n))
        out.append(rng.standard_normal(n) * np.exp2(np.linspace(300, -300, n)))
    n = 20000
    out.append((1.0 + 1e-3 * rng.standard_normal(n)) * np.where(np.arange(n) % 2, -1.0, 1.0))
    a = rng.standard_normal(70000) * np.exp2(rng.integers(-20, 20, 70000))
    out.append(a * (a + 0.5 * rng.standard_normal(70000)))
    out.append(rng.standard_normal(3000) * np.exp2(rng.integers(-1074, -1000, 3000)))
    t = rng.standard_normal(5000)
    t[100] = np.inf
    out.append(t)
    t = t.copy()
    t[4000] = -np.inf
    out.append(t)
    t = rng.standard_normal(9000)
    t[3000] = np.nan
    out.append(t)
    out.append(np.zeros(0))
    # addends of 2^61..2^62 ulps: per-thread integer runs wrap at 64 bits
    out.append(np.full(4096, 1.0) * np.exp2(61.5) * np.where(np.arange(4096) % 3, 1.0, -1.0))
    return out


@pytest.mark.parametrize("k", range(len(_cases())))
def test_emulated_reduction_in_bounds_and_exact(k):
    t = _cases()[k]
    got, _ = emulate(t)
    want = _sequential(t)
    assert got == want or (math.isnan(got) and math.isnan(want)), (len(t), got, want)


def test_more_than_one_summary_window():
    """1.1 M addends: 1075 tiles, two windows of tile summaries."""
    rng = np.random.default_rng(5)
    t = rng.standard_normal(1_100_000)
    got, bnd = emulate(t)
    assert got == _sequential(t)
    assert bnd.hi["summary"] == sum_tiles(len(t)) - 1
