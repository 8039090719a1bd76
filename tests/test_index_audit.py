"""CPU audit of the device kernels' memory indexing.

Every index a kernel forms from its lane number, pass, butterfly and
element is enumerated here for every variant the library builds, and
checked against the size of the buffer it addresses:

* the register-FFT passes of every compiled radix plan (the plans are read
  from fft_wave.hpp): LDS reads and writes inside the row buffer, the
  writers' address shortcuts equal to the readers' layout, the reads' split
  into a lane base plus compile-time offsets, twiddle entries (float pairs
  where the plan keeps them) inside the plan's table and equal to the entry
  the pass means, and no LDS bank conflicts (64-bit accesses are served 16
  lanes at a time over 32 banks);
* the runtime-planned passes of every FFT length 2..2048 the planner
  accepts: LDS indices of both row buffers and twiddle / direct-DFT entries
  inside the staged table;
* the per-lane row and depth-table offsets of the sweep, step, scatterer
  and LinDown/LinUp kernels (row_sample / row_store, phase table, reference
  ids, split-step and taper tables, scratch row) for full, padded and
  runtime-planned rows: model samples inside the row's buffer descriptor,
  padding samples outside it (the hardware bounds check reads 0 and drops
  the store, so no address outside the row is formed), and the phase-table
  entry equal to the mirrored wavenumber index;
* the depth rows of LinDown / LinUp items for every (dlo, dcount) that
  launch_lin accepts;
* the LDS tiles and image / scatter writes of the uneven offset imaging
  and scatter kernels for every lag count they are built for.
"""

import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wem-fwi_amd", "csrc")
LANES = np.arange(64)


def _src(name):
    with open(os.path.join(CSRC, name)) as f:
        return f.read()


def compiled_plans():
    """{(N, K): radices} of the FftPlan specialisations in fft_wave.hpp and
    the number of plans the tuner uses per length (FftPlans)."""
    s = _src("fft_wave.hpp")
    plans = {}
    for m in re.finditer(r"struct FftPlan<(\d+), (\d+)> \{\s*static constexpr int P = (\d+);\s*"
                         r"static constexpr int R\[P\] = \{([^}]*)\};", s):
        N, K, P = int(m.group(1)), int(m.group(2)), int(m.group(3))
        R = tuple(int(v) for v in m.group(4).split(","))
        assert len(R) == P and math.prod(R) == N
        plans[(N, K)] = R
    m = re.search(r"value = N == 1024 \? (\d+) : \(N == 2048 \? (\d+) : 1\)", s)
    count = {128: 1, 256: 1, 512: 1, 1024: int(m.group(1)), 2048: int(m.group(2))}
    used = {k: v for k, v in plans.items() if k[1] < count[k[0]]}
    assert sorted(used) == sorted((N, K) for N in count for K in range(count[N]))
    return used


PLANS = compiled_plans()


def plan_ns(R, p):
    return math.prod(R[:p])


def lay_am(R, p):
    """LDS layout between pass p and p + 1 (fft_wave.hpp lay_a / lay_m)."""
    ns = plan_ns(R, p)
    if 1 < ns < 16:
        w = ns * R[p]
        return (int(math.log2(w)) if w % 16 == 0 else 4), ns
    return 4, 1


def lay(R, p, i):
    a, m = lay_am(R, p)
    return i + (i >> a) * m


def lidx(i):
    return i + (i >> 4)


def lds_size(N):
    """LdsSize<N>: the largest layout over the length's plans."""
    m = N + N // 16
    for (n, _), R in PLANS.items():
        if n == N:
            for p in range(len(R) - 1):
                m = max(m, int(lay(R, p, N - 1)) + 1)
    return m


def tw_pair(R, p):
    ns = plan_ns(R, p)
    return ns > 1 and ns * R[p] <= 16


def tw_off(R, p):
    return sum((R[q] - 1) * plan_ns(R, q) * (2 if tw_pair(R, q) else 1) for q in range(p))


def bank_excess(addrs):
    """Extra LDS cycles of one 64-bit access (float2 addresses per lane):
    16 lanes at a time over 32 four-byte banks."""
    ex = 0
    for g0 in range(0, 64, 16):
        banks = np.concatenate([(2 * addrs[g0:g0 + 16]) % 32, (2 * addrs[g0:g0 + 16] + 1) % 32])
        ex += int(np.bincount(banks, minlength=32).max()) - 1
    return ex


@pytest.mark.parametrize("N,K", sorted(PLANS))
def test_compiled_plan_lds_and_twiddle_indices(N, K):
    R = PLANS[(N, K)]
    E, P, size = N // 64, len(R), lds_size(N)
    t = LANES
    tw_size = tw_off(R, P)
    assert size <= N + N // 16 + 8 * (N // 64) + 64  # the row buffers stay near N + N/16
    for p, Rp in enumerate(R):
        NS, NB, BPL = plan_ns(R, p), N // Rp, E // Rp
        if p > 0:
            for b in range(BPL):
                for r in range(Rp):
                    c = 64 * b + r * NB
                    a = lay(R, p - 1, t + c)
                    assert a.min() >= 0 and a.max() < size, (p, b, r)
                    # lane base + compile-time offset, as the reads address it
                    assert np.array_equal(a, lay(R, p - 1, t) + lay(R, p - 1, c))
                    assert bank_excess(a) == 0, ("read", p, b, r)
        if p < P - 1:
            for b in range(BPL):
                j = t + 64 * b
                for r in range(Rp):
                    i = (j // NS) * NS * Rp + j % NS + r * NS
                    a = lay(R, p, i)
                    assert a.min() >= 0 and a.max() < size
                    # the writers' shortcuts address the same layout
                    if NS == 1:
                        assert np.array_equal(a, (Rp * t + ((Rp * t) >> 4)) + 68 * Rp * b + r)
                    elif NS >= 64:
                        q = (64 * b) // NS
                        cc = q * NS * Rp + (64 * b - q * NS) + r * NS
                        assert np.array_equal(a, (t + (t >> 4)) + cc + (cc >> 4))
                    elif NS % 16 == 0:
                        u = (t // NS) * NS * Rp + (t & (NS - 1))
                        cc = (64 * b // NS) * NS * Rp + r * NS
                        assert np.array_equal(a, (u + (u >> 4)) + cc + (cc >> 4))
                    assert bank_excess(a) == 0, ("write", p, b, r)
        if NS > 1:
            pair = tw_pair(R, p)
            base = tw_off(R, p) + (t if NS >= 64 else (t & (NS - 1))) * (2 if pair else 1)
            for b in range(BPL):
                j = t + 64 * b
                for r in range(1, Rp):
                    if pair:
                        e = base + 2 * (r - 1) * NS
                        assert np.array_equal(e, tw_off(R, p) + 2 * ((r - 1) * NS + j % NS))
                        assert (e + 1).max() < tw_size
                    else:
                        e = base + (r - 1) * NS + (b * 64 if NS > 64 else 0)
                        assert np.array_equal(e, tw_off(R, p) + (r - 1) * NS + j % NS)
                        assert e.max() < tw_size


# ---------------------------------------------------------------- runtime plans

GEN_MAX_RADIX = 61


def gen_plan(n, maxr):
    """fft_gen_plan (host.cpp): powers of maxr, one smaller power of two,
    then the odd primes in increasing order; None beyond the supported set."""
    rad, m = [], n
    while m % maxr == 0:
        rad.append(maxr)
        m //= maxr
    for r in (8, 4, 2):
        if r < maxr and m % r == 0:
            rad.append(r)
            m //= r
    p = 3
    while m > 1:
        while m % p == 0:
            if p > GEN_MAX_RADIX:
                return None
            rad.append(p)
            m //= p
        p += 2
    return rad if len(rad) <= 10 else None


def gen_direct(R):
    return R > 5 and R not in (8, 16)


def gen_tw_pair(R, NS):
    return NS > 1 and NS * R <= 64 and not gen_direct(R)


def test_runtime_plans_every_length():
    checked = 0
    for n in range(2, 2049):
        N = max(128, 1 << (n - 1).bit_length())
        if n == N:
            continue  # compiled plans
        rad = gen_plan(n, 16 if N >= 1024 else 4)
        if rad is None:
            continue
        size = lds_size(N)
        cap = N + 3 * GEN_MAX_RADIX + 64
        total = sum((R - 1) * math.prod(rad[:p]) * (2 if gen_tw_pair(R, math.prod(rad[:p])) else 1) +
                    (R if gen_direct(R) else 0) for p, R in enumerate(rad))
        assert total <= cap, n
        # the row's samples in and out of buffer A
        assert lidx(n - 1) < size
        NS, off = 1, 0
        for R in rad:
            NB = n // R
            j = np.arange(NB)
            q, m = j // NS, j % NS
            for r in range(R):
                assert lidx(j + r * NB).max() < size
                assert lidx(q * NS * R + m + r * NS).max() < size
            if NS > 1:
                for r in range(1, R):
                    e = (r - 1) * NS + m
                    hi = off + (2 * e + 1 if gen_tw_pair(R, NS) else e)
                    assert hi.max() < total, (n, R, NS)
            if gen_direct(R):
                assert off + (R - 1) * NS + R <= total
            off += (R - 1) * NS * (2 if gen_tw_pair(R, NS) else 1) + (R if gen_direct(R) else 0)
            NS *= R
        assert NS == n and off == total
        checked += 1
    assert checked > 900   # lengths whose prime factors are at most 61


# ---------------------------------------------------------------- sweep rows

def _row_cases():
    out = []
    for N in (128, 256, 512, 1024, 2048):
        out.append((N, N, 0, False))                         # full row
        for tap in (1, 5, 28, 64, N // 4 - 1):
            if N - 2 * tap >= 2:
                out.append((N, N - 2 * tap, tap, False))     # padded row, FFT length N
        for nk in (N // 2 + 3, N - 7, N - 1):
            for tap in (0, 3):
                if nk - 2 * tap >= 2 and nk <= N and gen_plan(nk, 16 if N >= 1024 else 4):
                    out.append((N, nk - 2 * tap, tap, True))  # runtime-planned length nk < N
    return out


@pytest.mark.parametrize("N,nx,tap,gen", _row_cases())
def test_sweep_lane_offsets(N, nx, tap, gen):
    E = N // 64
    t = LANES[:, None]
    e = np.arange(E)[None, :]
    full = not gen and nx == N and tap == 0
    x = t + 64 * e - tap
    valid = (x >= 0) & (x < nx)
    # row_sample / row_store: byte offset in a descriptor of nx * 8 bytes
    off = (t * 8 + 512 * e) if full else ((x.astype(np.int64) % 2 ** 32) * 8) % 2 ** 32
    if full:
        assert valid.all() and (off < nx * 8).all()
    else:
        assert (off[valid] < nx * 8).all() and (off[~valid] >= nx * 8).all()
        assert valid.sum() == nx
    k = t + 64 * e
    nk = nx + 2 * tap
    if not gen:
        # ptab_at: entries min(k, N - k) of an (N/2 + 1)-entry table
        idx = np.where(e < N // 128, t + 64 * e, (64 - t) + 64 * (E - 1 - e))
        assert idx.min() >= 0 and idx.max() <= N // 2
        assert np.array_equal(idx, np.minimum(k, N - k))
    else:
        kh = np.where(k <= nk // 2, k, np.where(k < nk, nk - k, 0))
        assert kh.max() <= nk // 2 < N // 2 + 1
    # lane tables [nz][N]: refid bytes (4 per load), phi / taper entries,
    # the wave's scratch row [N]
    assert (t * E + E - 1 < N).all()
    if E >= 4:
        assert (t * E + 4 * (E // 4 - 1) + 3 < N).all()
    assert ((t + 64 * e) < N).all()


@pytest.mark.parametrize("nz", [2, 3, 17, 300])
@pytest.mark.parametrize("direction", [+1, -1])
def test_lin_items_stay_inside_the_depth_rows(nz, direction):
    """launch_lin accepts dlo >= (dir > 0 ? 0 : 1) and dlo + dcount <=
    (dir > 0 ? nz - 1 : nz); every item's rows d and d + dir then lie in
    [0, nz) (zbase = 0), for the whole range and every sub-range."""
    lo_min = 0 if direction > 0 else 1
    hi_max = nz - 1 if direction > 0 else nz
    for dlo in range(lo_min, hi_max):
        for dcount in range(1, hi_max - dlo + 1):
            d = np.arange(dlo, dlo + dcount)
            assert d.min() >= 0 and d.max() < nz
            assert (d + direction).min() >= 0 and (d + direction).max() < nz


# ---------------------------------------------------------------- uneven kernels

K_TX, K_HALO, K_ROWS = 256, 20, 4


@pytest.mark.parametrize("nhm", [1, 3, 11, 21, 41])
def test_uneven_offset_kernels_tiles(nhm):
    src = _src("uneven.hip")
    assert f"constexpr int kTX = {K_TX};" in src and f"constexpr int kHalo = {K_HALO};" in src
    assert f"constexpr int kRows = {K_ROWS};" in src
    t = np.arange(K_TX)
    kseg = K_TX + 2 * K_HALO
    hm = nhm // 2
    wd_cap = K_TX + 4 * hm
    for nh in range(1, nhm + 1, 2):
        hh = nh // 2
        # imaging: D / R segments [kRows][kSeg], lags h in [-hh, hh]
        seg = K_TX + 2 * hh
        assert seg <= kseg
        for h in range(-hh, hh + 1):
            for idx in (t + hh + h, t + hh - h):
                assert idx.min() >= 0 and idx.max() < seg
        # scatter: image tile [nh][wi] in si[NHM * (kTX + 2 HM)], D segments
        # [kRows][WD]
        wi, wd = K_TX + 2 * hh, K_TX + 4 * hh
        assert nh * wi <= nhm * (K_TX + 2 * hm) and wd <= wd_cap
        for j in range(nh):
            h = j - hh
            assert (j * wi + t - h + hh).min() >= j * wi and (j * wi + t - h + hh).max() < (j + 1) * wi
            b = t - 2 * h + 2 * hh
            assert b.min() >= 0 and b.max() < wd
        # the branch-free lag loops run all NHM lags of the template with
        # h = min(j - hh, hh) (lags past nh read a valid pair, never used;
        # the scatter reads no image value for them)
        assert "const int h = min(j0 + c - hh, hh);" in src or "const int h = min(j - hh, hh);" in src
        for j in range(nhm):
            h = min(j - hh, hh)
            for idx in (t + hh + h, t + hh - h):
                assert idx.min() >= 0 and idx.max() < seg
            b = t - 2 * h + 2 * hh
            assert b.min() >= 0 and b.max() < wd


def test_uneven_image_columns_inside_the_image():
    """Image writes go to column X + t = x0 + x for aperture columns x in
    [0, nx); the host rejects shots with x0 < 0 or x0 + nx > nx_image, so
    every written column is inside the image for every tile X."""
    src = _src("uneven.hip")
    assert "S.x0 < 0 || S.x0 + m.d.nx > nx_image" in src
    rng = np.random.default_rng(0)
    for _ in range(200):
        nxi = int(rng.integers(300, 5000))
        nx = int(rng.integers(2, nxi + 1))
        x0 = int(rng.integers(0, nxi - nx + 1))
        for X in range(0, nxi + K_TX, K_TX):
            x = X - x0 + np.arange(K_TX)
            ok = (x >= 0) & (x < nx)
            col = X + np.arange(K_TX)
            assert (col[ok] < nxi).all() and (col[ok] >= 0).all()
