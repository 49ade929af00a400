"""CPU oracle of the encoding wizard -- TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker; never by the
product path).

What it restates, and how each part is pinned:
  * candidate_sizes(x): exact encoded size of one 65536-value int64 row-group under every candidate expression of the
    wizard's pool, in closed form. The reference finds the best expression by encoding the column with every
    expression and comparing sizes (ExpColEncoder<PT>::encode_all, AllExpEncodedCol::GetBest,
    src/encoder/exp_col_encoder.cpp:19-28, 117-129; pool src/cor/exp_pool.cpp:23-144). Page layout
    (src/cor/lyt/page/page.cpp:120-130: header 8 B, arr_c u32 offsets, (vec_c-1)*arr_c u32 entry points, buffers),
    per-vector buffers (include/fls/cor/exp/*/..._exp.hpp RPN tables), bit widths from ffor_analyze
    (src/cor/prm/ffor/ffor_analyze.cpp:78-102, exact bit length), RLE split (src/cor/lyt/vec.cpp:50-71), dictionary
    insertion order (src/cor/lyt/dic/dic_fac.cpp:259-330) and the HZL dictionary payload
    (src/cor/eng/hzl_cmp.cpp:43-176). PINNED by tests/test_wizard_oracle.py against the byte length of the C oracle's
    own encoder (oracle/fls_oracle.c, itself pinned to the reference's local_result size CSVs) for every candidate on
    every data family; brute-force restatements of the delta/segment rules pin the vectorized forms.
  * features(x): the 64 sample statistics the learned selector reads (a definition of this build, not of the
    reference). PINNED by brute-force Python loops on small samples (tests/test_wizard_oracle.py).
  * mlp_forward_backward(...): softmax cross-entropy MLP 64-256-256-K with the GPU kernels' rounding points
    (BF16 operands, FP32 accumulation). PINNED by float64 finite differences of the loss (tests/test_wizard_oracle.py).
  * spell(columns): the structural wizard Wizard::Spell (src/wizard/wizard.cpp:19-151) as literal loops over Python
    values, whose == / != and dict semantics match the reference's types (ints; floats: -0.0 == 0.0, a NaN equals
    nothing and each NaN is its own dict key, as in the reference's unordered_map; strings as bytes). PINNED by
    tables built to a known plan: the columns of the reference's own generated example
    (data/generated/example_1/schema.json: a constant int and string, an equal group of three, a one-to-one pair),
    the 25% distinct-value boundary, NaN and signed-zero cases, type mismatches, a first with two seconds, and the
    order effects of the reference's loops (tests/test_wizard_oracle.py).
"""
from __future__ import annotations

import numpy as np

CANDIDATES = (0, 1, 11, 6, 7, 16, 17, 22, 2, 3, 4, 5, 27, 12, 13, 14)  # expression ids (i64 pool)
EXC_LIMIT = 100  # CFG::CMPR::EXC_LIMIT_C (include/fls/cfg/cfg.hpp:67)
_ORDER8 = np.array([0, 4, 2, 6, 1, 5, 3, 7])
_K = np.arange(1024)
ORIG_OF = (_K >> 7) + 8 * _ORDER8[(_K >> 4) & 7] + 64 * (_K & 15)  # transposed index k -> original position
K = len(CANDIDATES)
RG = 65536
V = RG // 1024
SAMPLE_VECTORS = (0, 21, 42, 63)
N_FEATURES = 64
H = 256
KPAD = 32

_U64 = np.uint64


def bitlen(d: np.ndarray) -> np.ndarray:
    """Exact bit length of unsigned 64-bit values (0 -> 0)."""
    d = np.asarray(d, dtype=np.uint64)
    out = np.zeros(d.shape, dtype=np.int64)
    for b in range(64):
        out = np.where((d >> _U64(b)) != 0, b + 1, out)
    return out


def _seg_deltas(x: np.ndarray, seg: int, mask=None) -> np.ndarray:
    """Differences of consecutive values (mod 2^bits) inside aligned segments of `seg` values, 0 at each segment
    start -- the transposed lane delta (unrsum) order expressed in the original order."""
    d = np.zeros_like(x)
    d[1:] = x[1:] - x[:-1]
    d[::seg] = 0
    if mask is not None:
        d &= mask
    return d


def _vec_bw(x2: np.ndarray) -> np.ndarray:
    """Per-row FFOR bit width of a (n_vec, 1024) uint64 array holding int64 values: bitlen(max - min) with min and max
    taken as signed values, as ANALYZE does for int64 columns (ffor_analyze.cpp:145-160, ffor_analyze<i64_pt>)."""
    b = np.asarray(x2, dtype=np.uint64) ^ _U64(1 << 63)
    mn = b.min(axis=1)
    mx = b.max(axis=1)
    return bitlen((mx - mn).astype(np.uint64))


def hzl_size(dvals: np.ndarray, width: int = 64) -> int:
    """Bytes of an HZL chain of `width`-bit values (no exceptions; a full-width chunk ends the FFOR part):
    12 + 128 bw per full chunk of 1024, then 12 + the raw rest (src/cor/eng/hzl_cmp.cpp:43-176)."""
    n = dvals.size
    pos, size = 0, 0
    while n - pos >= 1024:
        c = dvals[pos: pos + 1024].astype(np.uint64)
        bw = int(bitlen(np.array([c.max() - c.min()], dtype=np.uint64))[0])
        if bw >= width:
            break
        size += 12 + 128 * bw
        pos += 1024
    return size + 12 + (width // 8) * (n - pos)


def patch_plan(v: np.ndarray):
    """find_best_params (src/cor/prm/ffor/ffor_analyze.cpp:41-75) on one vector of uint64 values: the (bw, base)
    minimising 128 bw + 8 (values outside [base, base + 2^bw - 1]) with at most EXC_LIMIT values outside, bw < 64,
    base = the first occurrence of one of the 101 smallest values (bw = 0: only base 0, window {0}); ties to the
    smaller start index, then the larger bw; none below 128 * 64 bytes -> (64, 0). Every (bw, start) pair is
    evaluated (one search per pair)."""
    sv = np.sort(v.astype(np.uint64))
    M = np.uint64((1 << 64) - 1)
    starts = np.array([i for i in range(EXC_LIMIT + 1) if i == 0 or sv[i] != sv[i - 1]])
    si = sv[starts]                                               # (S,)
    bws = np.arange(1, 64, dtype=np.uint64)
    rng_ = (np.uint64(1) << bws) - np.uint64(1)                   # (63,)
    upper = np.where(si[None, :] > M - rng_[:, None], M, si[None, :] + rng_[:, None])  # (63, S), saturating
    inside = np.searchsorted(sv, upper.ravel(), side="right").reshape(upper.shape) - starts[None, :]
    exc = 1024 - inside
    best = None
    for b in range(63):
        for j in range(starts.size):
            if exc[b, j] <= EXC_LIMIT:
                key = (128 * (b + 1) + 8 * int(exc[b, j]), int(starts[j]), -(b + 1))
                best = key if best is None or key < best else best
    zeros = int(np.count_nonzero(sv == 0))  # bw = 0: the window {0}, from base 0 only
    if sv[0] == 0 and 1024 - zeros <= EXC_LIMIT:
        key = (8 * (1024 - zeros), 0, 0)
        best = key if best is None or key < best else best
    if best is None or best[0] >= 128 * 64:
        return 64, 0
    return -best[2], int(sv[best[1]])


def patch_exceptions(v: np.ndarray, bw: int, base: int) -> np.ndarray:
    """Transposed indices of a vector's exceptions under plan (bw, base), ascending (the encoder's order)."""
    t = v.astype(np.uint64)[ORIG_OF]
    if bw >= 64:
        return np.zeros(0, dtype=np.int64)
    if bw == 0:
        return np.nonzero(t != 0)[0]
    upper = min(base + (1 << bw) - 1, (1 << 64) - 1)
    return np.nonzero((t < np.uint64(base)) | (t > np.uint64(upper)))[0]


def dictionary(x: np.ndarray):
    """Insertion-order dictionary: (values in code order, code of every element)."""
    uniq, first, inv = np.unique(x, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")  # code c -> unique index
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return uniq[order], rank[inv].astype(np.uint32)


def candidate_sizes(x: np.ndarray) -> np.ndarray:
    """Exact encoded bytes of a 65536-value int64 row-group under each candidate (order of CANDIDATES)."""
    assert x.size == RG
    u = np.ascontiguousarray(x).view(np.uint64)
    u2 = u.reshape(V, 1024)
    bwf = _vec_bw(u2)
    bwd = _vec_bw(_seg_deltas(u, 64).reshape(V, 1024))
    chg = np.ones(RG, dtype=bool)
    chg[1:] = u[1:] != u[:-1]
    chg[::1024] = True  # runs restart in every vector
    runs = chg.reshape(V, 1024).sum(axis=1)
    inner = chg.copy()
    inner[::16] = False  # position deltas are 0 at every 16-value segment start
    bwp = inner.reshape(V, 1024).any(axis=1).astype(np.int64)
    dvals, codes = dictionary(u)
    c2 = codes.reshape(V, 1024).astype(np.uint32)
    bwc = bitlen((c2.max(axis=1) - c2.min(axis=1)).astype(np.uint64))
    cd = _seg_deltas(codes.astype(np.uint32), 32).reshape(V, 1024)
    bwcd = bitlen((cd.max(axis=1) - cd.min(axis=1)).astype(np.uint64))
    dict_page = 8 + hzl_size(dvals)

    def page(arr_c, bufs):
        return 8 + 4 * arr_c * V + int(np.sum(bufs))

    sizes = [
        page(1, np.full(V, 8192)),                              # 0  UNCOMPRESSED
        page(3, 9 + 128 * bwf),                                  # 1  FFOR
        page(4, 137 + 128 * bwd),                                # 11 DELTA
        page(5, 8 * runs + 131 + 128 * bwp),                     # 6  RLE, FFOR'd delta positions
        page(2, 8 * runs + 2048),                                # 7  RLE, plain positions
        dict_page + page(1, np.full(V, 4096)),                   # 16 DICT, plain codes
        dict_page + page(3, 5 + 128 * bwc),                      # 17 DICT, FFOR codes
        dict_page + page(4, 133 + 128 * bwcd),                   # 22 DICT, delta codes
    ]
    # patched FFOR (2 S_PATCH, 3 S_PATCH_REC, 4 B_PATCH, 5 LL_PATCH): per vector [exceptions][positions][bw][base][packed]
    s_b, srec_b, b_b, ll_b, chain = [], [], [], [], []
    for vv in range(V):
        bw, base = patch_plan(u2[vv])
        pos = patch_exceptions(u2[vv], bw, base)
        n = int(pos.size)
        hop = 1 << min(bw, 11)
        m = n + int(sum((int(g) - 1) // hop for g in np.diff(pos)))  # the LL chain's hop links
        pk = 9 + 128 * bw
        s_b.append(8 * n + 2 + 2 * n + pk)
        srec_b.append(2 + 2 * n + pk)
        b_b.append(8 * n + 128 + pk)
        ll_b.append(8 * m + 4 + pk)
        chain.append(u2[vv][ORIG_OF][pos])
    srec_chain = hzl_size(np.concatenate(chain) if chain else np.zeros(0, np.uint64))
    # DICT_RLE (27): per vector [run codes: the page's HZL chain of u32][pos lane bases][bw][base u16][packed]
    run_codes = codes[chg]
    sizes += [page(5, s_b), page(5, srec_b) + srec_chain, page(5, b_b), page(5, ll_b),
              dict_page + page(5, 131 + 128 * bwp) + hzl_size(run_codes.astype(np.uint64), 32)]
    # DELTA patches (12 S_PATCH, 13 B_PATCH, 14 LL_PATCH): the patch plan of the lane differences, per vector
    # [lane bases][exceptions][positions | bitmap | count + first][bw][base][packed]
    dl = _seg_deltas(u, 64).reshape(V, 1024)
    ds_b, db_b, dll_b = [], [], []
    for vv in range(V):
        bw, base = patch_plan(dl[vv])
        pos = patch_exceptions(dl[vv], bw, base)
        n = int(pos.size)
        hop = 1 << min(bw, 11)
        m = n + int(sum((int(g) - 1) // hop for g in np.diff(pos)))
        pk = 9 + 128 * bw
        ds_b.append(128 + 8 * n + 2 + 2 * n + pk)
        db_b.append(128 + 8 * n + 128 + pk)
        dll_b.append(128 + 8 * m + 4 + pk)
    sizes += [page(6, ds_b), page(6, db_b), page(6, dll_b)]
    return np.array(sizes, dtype=np.uint64)


def best(sizes: np.ndarray) -> np.ndarray:
    return np.argmin(sizes, axis=-1).astype(np.int32)


# ---- features ---------------------------------------------------------------------------------------------------
def _log2f(v):
    return np.log2(np.float32(v)).astype(np.float32)


def features(x: np.ndarray) -> np.ndarray:
    """64 float32 statistics of a 65536-value row-group, read from SAMPLE_VECTORS only."""
    u = np.ascontiguousarray(x).view(np.uint64).reshape(V, 1024)
    s = np.ascontiguousarray(u[list(SAMPLE_VECTORS)])  # (4, 1024)
    si = s.view(np.int64)
    S = s.shape[0]
    flat = s.ravel()
    bwf = _vec_bw(s)
    bwd = np.array([_vec_bw(_seg_deltas(s[i], 64)[None])[0] for i in range(S)])
    chg = np.ones_like(s, dtype=bool)
    chg[:, 1:] = s[:, 1:] != s[:, :-1]
    runs = chg.sum(axis=1)
    inner = chg.copy()
    inner[:, ::16] = False
    bwp = inner.any(axis=1).astype(np.int64)
    eq = int((s[:, 1:] == s[:, :-1]).sum())
    nondec = int((si[:, 1:] >= si[:, :-1]).sum())
    inc = int((si[:, 1:] > si[:, :-1]).sum())
    uniq, cnt = np.unique(flat, return_counts=True)
    D = int(uniq.size)
    top = int(cnt.max())
    _, codes = dictionary(flat)
    c2 = codes.reshape(S, 1024).astype(np.uint32)
    bwc = bitlen((c2.max(axis=1) - c2.min(axis=1)).astype(np.uint64))
    cdl = np.stack([_seg_deltas(c2[i], 32) for i in range(S)])
    bwcd = bitlen((cdl.max(axis=1) - cdl.min(axis=1)).astype(np.uint64))
    umax, umin = int(flat.max()), int(flat.min())
    smax, smin = int(si.max()), int(si.min())
    bws = int(bitlen(np.array([(smax - smin) % (1 << 64)], dtype=np.uint64))[0])
    plans = [patch_plan(s[i]) for i in range(S)]
    pbw = np.array([bw for bw, _ in plans])
    pexc = np.array([patch_exceptions(s[i], bw, base).size for i, (bw, base) in enumerate(plans)])
    sd = [_seg_deltas(s[i], 64) for i in range(S)]
    dplans = [patch_plan(sd[i]) for i in range(S)]
    dpbw = np.array([bw for bw, _ in dplans])
    dpexc = np.array([patch_exceptions(sd[i], bw, base).size for i, (bw, base) in enumerate(dplans)])

    f32 = np.float32
    mbwf, mbwd, mruns, mbwp = (f32(np.sum(a)) / f32(S) for a in (bwf, bwd, runs, bwp))
    mbwc, mbwcd = f32(np.sum(bwc)) / f32(S), f32(np.sum(bwcd)) / f32(S)
    vv = f32(V)
    est = np.array([
        f32(8) + f32(4) * vv + f32(8192) * vv,
        f32(8) + f32(12) * vv + vv * (f32(9) + f32(128) * mbwf),
        f32(8) + f32(16) * vv + vv * (f32(137) + f32(128) * mbwd),
        f32(8) + f32(20) * vv + vv * (f32(8) * mruns + f32(131) + f32(128) * mbwp),
        f32(8) + f32(8) * vv + vv * (f32(8) * mruns + f32(2048)),
        f32(28) + f32(8) * f32(D) + f32(4) * vv + f32(4096) * vv,
        f32(28) + f32(8) * f32(D) + f32(12) * vv + vv * (f32(5) + f32(128) * mbwc),
        f32(28) + f32(8) * f32(D) + f32(16) * vv + vv * (f32(133) + f32(128) * mbwcd),
    ], dtype=np.float32)
    mpbw, mexc = f32(np.sum(pbw)) / f32(S), f32(np.sum(pexc)) / f32(S)
    pk = f32(9) + f32(128) * mpbw
    base20 = f32(8) + f32(20) * vv
    est = np.concatenate([est, np.array([
        base20 + vv * (f32(10) * mexc + f32(2) + pk),
        base20 + vv * (f32(2) * mexc + f32(2) + pk) + f32(12) + f32(8) * mexc * vv,
        base20 + vv * (f32(8) * mexc + f32(128) + pk),
        base20 + vv * (f32(8) * mexc + f32(4) + pk),
        f32(28) + f32(8) * f32(D) + f32(20) * vv + vv * (f32(131) + f32(128) * mbwp) + f32(12) + f32(4) * mruns * vv,
    ], dtype=np.float32)])
    mdpbw, mdexc = f32(np.sum(dpbw)) / f32(S), f32(np.sum(dpexc)) / f32(S)
    pkd = f32(9) + f32(128) * mdpbw
    base24 = f32(8) + f32(24) * vv
    est = np.concatenate([est, np.array([
        base24 + vv * (f32(10) * mdexc + f32(130) + pkd),
        base24 + vv * (f32(8) * mdexc + f32(256) + pkd),
        base24 + vv * (f32(8) * mdexc + f32(132) + pkd),
    ], dtype=np.float32)])
    K = len(CANDIDATES)
    F = np.zeros(N_FEATURES, dtype=np.float32)
    F[0:K] = est / f32(RG * 8)
    F[K:2 * K] = _log2f(est) / f32(24)
    G = 2 * K
    F[G + 0], F[G + 1], F[G + 2] = mbwf / f32(64), f32(bwf.max()) / f32(64), f32(bwf.min()) / f32(64)
    F[G + 3], F[G + 4], F[G + 5] = mbwd / f32(64), f32(bwd.max()) / f32(64), f32(bwd.min()) / f32(64)
    F[G + 6], F[G + 7], F[G + 8] = mruns / f32(1024), f32(runs.max()) / f32(1024), f32(runs.min()) / f32(1024)
    F[G + 9] = _log2f(mruns) / f32(10)
    F[G + 10] = mbwp
    F[G + 11] = f32(D) / f32(S * 1024)
    F[G + 12] = _log2f(D) / f32(12)
    F[G + 13] = f32(int(bitlen(np.array([D - 1], dtype=np.uint64))[0])) / f32(16)
    F[G + 14] = f32(top) / f32(S * 1024)
    F[G + 15] = f32(eq) / f32(S * 1023)
    F[G + 16] = f32(nondec) / f32(S * 1023)
    F[G + 17] = f32(inc) / f32(S * 1023)
    F[G + 18] = f32(int(bitlen(np.array([umax], dtype=np.uint64))[0])) / f32(64)
    F[G + 19] = f32(int(bitlen(np.array([umin], dtype=np.uint64))[0])) / f32(64)
    F[G + 20] = f32(bws) / f32(64)
    F[G + 21] = mbwc / f32(32)
    F[G + 22] = mbwcd / f32(32)
    for i in range(S):
        F[G + 23 + i] = f32(bwf[i]) / f32(64)
    F[59] = mpbw / f32(64)
    F[60] = mexc / f32(100)
    F[61] = mdpbw / f32(64)
    F[62] = mdexc / f32(100)
    F[63] = f32(1)
    return F


# ---- MLP ---------------------------------------------------------------------------------------------------------
def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack_params(p: np.ndarray):
    """Parameter block -> (W1t[256,64], b1, W2t[256,256], b2, W3t[32,256], b3)."""
    o = 0
    out = []
    for shape in ((H, N_FEATURES), (H,), (H, H), (H,), (KPAD, H), (KPAD,)):
        n = int(np.prod(shape))
        out.append(p[o: o + n].reshape(shape))
        o += n
    return out


N_PARAMS = H * N_FEATURES + H + H * H + H + KPAD * H + KPAD


def mlp_forward_backward(p: np.ndarray, X: np.ndarray, y: np.ndarray, n_classes: int, rounding: bool = True):
    """Loss (mean softmax cross-entropy over the first n_classes logits) and its gradient block. With rounding,
    operands are rounded to bf16 where the GPU kernels store bf16 (inputs, weights, activations, upstream
    gradients); sums are float32 (float64 without rounding)."""
    r = bf16 if rounding else (lambda a: a)
    dt = np.float32 if rounding else np.float64
    W1t, b1, W2t, b2, W3t, b3 = (a.astype(dt) for a in unpack_params(p.astype(dt)))
    B = X.shape[0]
    Xb = r(X.astype(dt))
    A1 = Xb @ r(W1t).T + b1
    H1 = r(np.maximum(A1, 0))
    A2 = H1 @ r(W2t).T + b2
    H2 = r(np.maximum(A2, 0))
    Z = H2 @ r(W3t).T + b3
    Zk = Z[:, :n_classes]
    m = Zk.max(axis=1, keepdims=True)
    e = np.exp(Zk - m)
    P = e / e.sum(axis=1, keepdims=True)
    loss = float(np.mean(-np.log(P[np.arange(B), y])))
    dZ = np.zeros((B, KPAD), dtype=dt)
    dZ[:, :n_classes] = P
    dZ[np.arange(B), y] -= 1
    dZ /= B
    dZb = r(dZ)
    gW3t = dZb.T @ H2
    gb3 = dZ.sum(axis=0)
    dH2 = r((dZb @ r(W3t)) * (H2 > 0))
    gW2t = dH2.T @ H1
    gb2 = dH2.sum(axis=0)
    dH1 = r((dH2 @ r(W2t)) * (H1 > 0))
    gW1t = dH1.T @ Xb
    gb1 = dH1.sum(axis=0)
    g = np.concatenate([gW1t.ravel(), gb1, gW2t.ravel(), gb2, gW3t.ravel(), gb3]).astype(dt)
    return loss, g, Z


def mlp_loss(p, X, y, n_classes):
    return mlp_forward_backward(p, X, y, n_classes, rounding=False)[0]


# -- structural wizard ------------------------------------------------------------------------------------------

def spell(columns):
    """columns: list of (type_name, values) with values a Python list (ints, floats or bytes). Returns (kinds, refs):
    kinds in EXPRESSION / CONSTANT / EQUAL / MAP_SOURCE / MAP_TARGET, refs the column each kind refers to (-1)."""
    n = len(columns)
    rows = len(columns[0][1]) if n else 0
    kinds, refs, det = ["EXPRESSION"] * n, [-1] * n, [False] * n

    # constant_check (wizard.cpp:88-93): stats.is_constant, cleared when a value differs from the previous one
    # (src/table/attribute.cpp:136-138)
    for c, (_, v) in enumerate(columns):
        if not any(v[k] != v[k - 1] for k in range(1, len(v))):
            kinds[c], det[c] = "CONSTANT", True

    # equality_check (wizard.cpp:96-111), Rowgroup::Equal (src/table/rowgroup.cpp:113-156)
    for i in range(n):
        for j in range(i + 1, n):
            if det[j]:
                continue
            if columns[i][0] == columns[j][0] and not any(a != b for a, b in zip(columns[i][1], columns[j][1])):
                kinds[j], refs[j], det[j] = "EQUAL", i, True

    def good_for_dictionary(c):  # src/table/rowgroup.cpp:192-217
        distinct = {}
        for x in columns[c][1]:
            distinct.setdefault(x, None)
        return not (len(distinct) / rows > 0.25) if rows else True

    def is_map_1t1(f, s):  # src/table/rowgroup.cpp:159-175
        m = {}
        for a, b in zip(columns[f][1], columns[s][1]):
            if a in m:
                if m[a] != b:
                    return False
            else:
                m[a] = b
        return True

    # map_1t1_check (wizard.cpp:125-151): the first's own flag is tested only in the outer loop
    for f in range(n):
        if det[f]:
            continue
        for s in range(f + 1, n):
            if det[s]:
                continue
            if not good_for_dictionary(f):
                continue
            if is_map_1t1(f, s):
                if kinds[f] != "MAP_SOURCE":
                    kinds[f], refs[f] = "MAP_SOURCE", s
                det[f] = True
                kinds[s], refs[s], det[s] = "MAP_TARGET", f, True
    return kinds, refs
