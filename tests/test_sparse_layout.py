"""The sparse kernel's nonzero lists (bnm_sparse_layout: the host builder that
bnm_ctx_set_sc runs), decoded on the CPU: walking each lane's steps must
visit exactly the nonzeros of its slice of the row, in increasing k, with the
SC's weights (rounded to FP32 for the FP32 layout), and never leave the
slice; padding and gap-bridging slots carry zero weights.  The column-paced
placement is what keeps a half-wave's gathers in distinct LDS banks: on the
40 %-dense N = 400 SC no slot of any 32-lane group reads two addresses in one
bank.  Also: where the kernel does not apply (too dense, FP64 beyond its 64
register slots, N beyond 512 / lanes), the slot count is 0."""
import numpy as np
import pytest

import bnm_cuda_amd as bnm
from bnm_cuda_amd import workloads as W


def _block_sc(N=300, seed=21):
    rng = np.random.default_rng(seed)
    A = np.zeros((N, N))
    for lo in (0, 140):
        blk = (rng.random((140, 140)) < 0.3) * rng.lognormal(0, 1, (140, 140))
        A[lo:lo + 140, lo:lo + 140] = np.triu(blk, 1)
    for i in range(10):
        A[i, 270 - 2 * i] = 0.5
    A = A + A.T
    return A / A.sum(1).mean()


def _signed_sc(N=160, seed=5):
    """Non-symmetric, signed weights, some rows empty."""
    rng = np.random.default_rng(seed)
    A = (rng.random((N, N)) < 0.15) * rng.normal(0, 1, (N, N))
    A[::17] = 0.0
    return A


def _decode(P, Wt, D, N, LS, es):
    """Per lane column (slice sl of row j): the (k, weight) of every
    nonzero-weight slot, walking the byte steps; asserts the walk stays in
    the slice."""
    KH = (N + LS - 1) // LS
    out = {}
    for sl in range(LS):
        for j in range(N):
            col = sl * N + j
            addr, items = 0, []
            for p in range(P):
                step = (int(D[p // 4, col]) >> (8 * (p % 4))) & 0xFF
                assert step % es == 0
                addr += step
                k = sl * KH + addr // es
                assert sl * KH <= k < max(sl * KH + 1, min(N, (sl + 1) * KH))
                if Wt[p, col] != 0.0:
                    items.append((k, Wt[p, col]))
            out[(sl, j)] = items
    return out


CASES = [("N150-20%", lambda: W.synthetic_sc(150, 0.20, 3)), ("N200-33%", lambda: W.synthetic_sc(200, 0.33, 12)),
         ("N400-40%", lambda: W.synthetic_sc(400, 0.40, 13)), ("block-gaps", _block_sc), ("signed", _signed_sc),
         ("two-blocks", lambda: _two_blocks())]


@pytest.mark.parametrize("name,make", CASES)
@pytest.mark.parametrize("LS", [1, 2])
@pytest.mark.parametrize("fp64", [False, True])
def test_layout_decodes_to_the_sc_rows(name, make, LS, fp64):
    SC = make()
    N = SC.shape[0]
    P, Wt, D = bnm.sparse_layout(SC, LS, fp64)
    if P == 0:
        # only where the kernel does not apply: N beyond 512 / lanes, or (FP64,
        # 64 register slots, steps of at most 31 nodes) a row the FP32 layout
        # already needs many slots for
        assert N > 512 // LS or (fp64 and bnm.sparse_layout(SC, LS, False)[0] > 48)
        return
    es = 8 if fp64 else 4
    assert P % 8 == 0 and P <= (64 if fp64 else 208)
    ref = SC if fp64 else SC.astype(np.float32).astype(np.float64)
    KH = (N + LS - 1) // LS
    got = _decode(P, Wt, D, N, LS, es)
    for (sl, j), items in got.items():
        ks = [k for k in range(sl * KH, min(N, (sl + 1) * KH)) if ref[j, k] != 0.0]
        assert [k for k, _ in items] == ks, (name, sl, j)
        assert [w for _, w in items] == [ref[j, k] for k in ks], (name, sl, j)
        assert P >= len(ks)


def _two_blocks(N=200, B=90, seed=23):
    """Two 30 %-dense blocks of B nodes, isolated nodes after them, and links
    from rows 0-4 to the far end of the second block (gaps over 63 nodes)."""
    rng = np.random.default_rng(seed)
    A = np.zeros((N, N))
    for lo in (0, B):
        A[lo:lo + B, lo:lo + B] = np.triu((rng.random((B, B)) < 0.3) * rng.lognormal(0, 1, (B, B)), 1)
    for i in range(5):
        A[i, 2 * B - 5 - i] = 0.5
    return A + A.T


def test_layout_bridges_long_gaps():
    """Rows with gaps of more than 63 nodes: FP32 steps hold at most 63
    nodes, FP64 steps 31, so the walk takes zero-weight bridging slots."""
    SC = _two_blocks()
    assert np.diff(np.nonzero(SC[0])[0]).max() > 63
    for fp64, es in ((False, 4), (True, 8)):
        P, Wt, D = bnm.sparse_layout(SC, 1, fp64)
        assert P > 0
        steps = [(int(D[p // 4, 0]) >> (8 * (p % 4))) & 0xFF for p in range(P)]
        assert max(steps) == 255 // es * es
        zero_moves = sum(1 for p in range(P) if steps[p] and Wt[p, 0] == 0.0)
        assert zero_moves >= 1


def test_paced_slots_keep_half_waves_in_distinct_banks():
    SC = W.synthetic_sc(400, 0.40, 13)
    P, Wt, D = bnm.sparse_layout(SC, 1, False)
    N = 400
    cols = np.zeros((P, N), np.int64)
    for j in range(N):
        addr = 0
        for p in range(P):
            addr += (int(D[p // 4, j]) >> (8 * (p % 4))) & 0xFF
            cols[p, j] = addr // 4
    conflicts = 0
    for g0 in range(0, N, 32):          # 32-lane groups (half-waves)
        for p in range(P):
            ks = np.unique(cols[p, g0:g0 + 32])
            banks = ks % 32
            conflicts += len(banks) - len(np.unique(banks))
    print("bank conflicts over", P * ((N + 31) // 32), "half-wave slots:", conflicts)
    assert conflicts == 0


def test_layout_not_offered():
    assert bnm.sparse_layout(W.synthetic_sc(300, 0.9, 5), 1, False)[0] == 0       # too dense (P > 208)
    assert bnm.sparse_layout(W.synthetic_sc(400, 0.40, 13), 1, True)[0] == 0      # FP64: P > 64
    assert bnm.sparse_layout(W.synthetic_sc(300, 0.2, 5), 2, False)[0] == 0      # 2 lanes: N > 256
    assert bnm.sparse_layout(np.zeros((50, 50)), 1, False)[0] == 0                 # no nonzeros
