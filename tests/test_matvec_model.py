"""CPU model of the fused local-operator kernel (torchtt_amd/csrc/kernels_matvec.h, launch plan in
contract.hip: mv_plan): the same launch plan, operator packing and buffer indices, with every LDS
and global access checked against its buffer's extent, and the result checked against the
contraction 'lsr,smnS,LSR,vrnR->vlmL' (reference cpp/matvecs.h:196-212). The matrix-core steps are
modelled as the products they compute on the fragments the kernel loads."""
import numpy as np
import pytest

PITCH, COLS, MAX_LDS = 17, 16, 160 * 1024 - 1024


def plan(l1, s, r1, n, S, l2, r2, elem):
    """contract.hip: mv_plan."""
    if s < 1 or s > 16 or S < 1 or S > 32:
        return None
    BL = max(1, min(8 // s if s <= 8 else 1, l1))
    NLS = 4 if BL * s <= 4 else (8 if BL * s <= 8 else 16)
    BR = max(1, min(COLS // BL, 32 // S, r2))
    nlb = (l1 + BL - 1) // BL
    P = (r2 + BR - 1) // BR
    if P > 18:
        return None
    KS, MS = s * n, n * S
    KST, RB = (KS + 3) // 4, (MS + 15) // 16
    lds = elem * (r1 * NLS + (4 * KST + 16 * RB) * PITCH)
    if lds > MAX_LDS:
        return None
    return dict(BL=BL, NLS=NLS, BR=BR, nlb=nlb, P=P, KS=KS, MS=MS, KST=KST, RB=RB, lds=lds)


def pack(Apm, MS, KS, KST, RB):
    """KPackFragA: element lane of fragment (rb, ks) = Apm[16 rb + lane % 16][4 ks + lane / 16]."""
    out = np.zeros(RB * KST * 64)
    i = np.arange(out.size)
    lane, f = i & 63, i >> 6
    ks, rb = f % KST, f // KST
    row, k = rb * 16 + (lane & 15), ks * 4 + (lane >> 4)
    ok = (row < MS) & (k < KS)
    out[ok] = Apm[row[ok], k[ok]]
    return out


def unpack_block(Apk, rb, ks, KST):
    """The 16 x 4 A fragment a wave's lanes load for row block rb, k step ks."""
    lane = np.arange(64)
    idx = ((rb * KST + ks) << 6) + lane
    assert idx.max() < Apk.size
    blk = np.zeros((16, 4))
    blk[lane & 15, lane >> 4] = Apk[idx]
    return blk


def inb(idx, size, what):
    idx = np.asarray(idx)
    assert idx.size == 0 or (idx.min() >= 0 and idx.max() < size), (what, int(idx.min()), int(idx.max()), size)


def model(PhiL, A, PhiR, x):
    nvec, r1, n, r2 = x.shape
    l1, s, _ = PhiL.shape
    l2, S, _ = PhiR.shape
    pl = plan(l1, s, r1, n, S, l2, r2, 8)
    assert pl is not None
    BL, BR, NLS, P, KS, MS, KST, RB = (pl[k] for k in ("BL", "BR", "NLS", "P", "KS", "MS", "KST", "RB"))
    # permute_core_A: Apm[(m, S)][(s, n)] = A[s, m, n, S]
    Apm = A.transpose(1, 3, 0, 2).reshape(MS, KS)
    Apk = pack(Apm, MS, KS, KST, RB)
    # the stage-2 operand as the waves' fragment loads see it (every fragment index checked)
    Afrag = np.zeros((16 * RB, 4 * KST))
    for rb in range(RB):
        for ks in range(KST):
            Afrag[rb * 16:rb * 16 + 16, ks * 4:ks * 4 + 4] = unpack_block(Apk, rb, ks, KST)
    y = np.full((nvec, l1, n, l2), np.nan)
    part = np.full((P, nvec, l1, n, l2), np.nan) if P > 1 else None
    lds_elems = r1 * NLS + (4 * KST + 16 * RB) * PITCH
    assert pl["lds"] == 8 * lds_elems
    counts = np.zeros((nvec, pl["nlb"]), dtype=int)
    for v in range(nvec):
        for p in range(P):
            for lb in range(pl["nlb"]):
                l0, R0 = lb * BL, p * BR
                BLv, BRv = min(BL, l1 - l0), min(BR, r2 - R0)
                nls = BLv * s
                # PhiL slice
                PL = np.zeros((r1, NLS))
                for j in range(NLS):
                    if j < nls:
                        inb([(l0 + j // s) * s + j % s], l1 * s, "PhiL row")
                        PL[:, j] = PhiL[l0 + j // s, j % s, :]
                # stage 1
                T1 = np.zeros((4 * KST, PITCH))
                u = np.arange(n * BR)
                ni, R = u // BR, u % BR
                live = R < BRv
                xi = v * r1 * n * r2 + np.arange(r1)[:, None] * n * r2 + (ni * r2 + R0 + R)[None, :]
                inb(xi[:, live], x.size, "x")
                xs = np.where(live[None, :], x.reshape(-1)[np.where(live[None, :], xi, 0)], 0.0)
                acc = PL.T @ xs  # acc[j][u] = sum_r PL[r][j] x[r][u]
                for j in range(NLS):
                    l, si = j // s, j % s
                    if l < BL:
                        row, col = si * n + ni, l * BR + R
                        assert row.max() < KS and col.max() < COLS
                        T1[row, col] = acc[j]
                # stage 2: T2 = Apm T1 blockwise from the packed fragments
                T2 = np.zeros((16 * RB, PITCH))
                inb([(KST * 4 - 1) * PITCH + 15], T1.size, "T1 read")
                T2[:, :COLS] = Afrag @ T1[:, :COLS]
                # stage 3
                M3, K3 = BLv * n, S * BR
                K3S = (K3 + 3) // 4
                assert K3S <= 8
                Am = np.zeros((M3, 4 * K3S))
                Bm = np.zeros((4 * K3S, l2))
                for k in range(K3):
                    Si, R = k // BR, k % BR
                    i = np.arange(M3)
                    li, mi = i // n, i % n
                    t2i = (mi * S + Si) * PITCH + li * BR + R
                    inb(t2i, T2.size, "T2 read")
                    assert (li * BR + R).max() < BL * BR <= COLS
                    Am[:, k] = T2.reshape(-1)[t2i]
                    if R < BRv:
                        pri = (np.arange(l2) * S + Si) * r2 + R0 + R
                        inb(pri, PhiR.size, "PhiR")
                        Bm[k, :] = PhiR.reshape(-1)[pri]
                Y = Am @ Bm  # rows (l, m), columns L
                dst = part[p, v] if P > 1 else y[v]
                dst[l0:l0 + BLv] = Y.reshape(BLv, n, l2)
                counts[v, lb] += 1
    if P > 1:
        assert (counts == P).all()
        y = part[0].copy()
        for q in range(1, P):
            y += part[q]
    return y, pl


SHAPES = [
    # l1, s, r1, n, S, l2, r2, nvec
    (44, 4, 44, 64, 4, 44, 44, 1),
    (84, 4, 84, 64, 4, 84, 84, 1),
    (1, 1, 1, 16, 2, 5, 5, 1),       # first core
    (7, 2, 7, 9, 1, 1, 1, 2),        # last core, two vectors
    (4, 4, 30, 33, 4, 4, 29, 1),     # residual-proxy product (rz = 4)
    (30, 4, 30, 33, 4, 4, 29, 1),    # enrichment product
    (13, 3, 11, 17, 3, 9, 23, 1),    # rank-3 operator, ragged blocks
    (5, 4, 6, 96, 4, 3, 8, 1),       # long mode
    (140, 4, 140, 64, 4, 140, 140, 1),
    (132, 4, 132, 64, 4, 132, 132, 1),
]


def test_plan_declines_what_the_kernel_does_not_take():
    """Shapes whose tiles exceed the LDS (s n + n S rows of 17 columns) or whose R needs more than
    18 partial sums run as three GEMMs instead."""
    assert plan(5, 4, 6, 160, 4, 3, 8, 8) is None          # 174 KB of LDS
    assert plan(5, 4, 6, 160, 4, 3, 8, 4) is not None      # FP32 tiles fit
    assert plan(2, 1, 2, 16, 1, 2, 300, 8) is None          # 300 / 16 > 18 partial sums
    assert plan(8, 17, 8, 16, 4, 8, 8, 8) is None           # operator rank above 16


@pytest.mark.parametrize("shape", SHAPES)
def test_fused_matvec_model(shape):
    l1, s, r1, n, S, l2, r2, nvec = shape
    rng = np.random.default_rng(sum(shape))
    PhiL = rng.standard_normal((l1, s, r1))
    A = rng.standard_normal((s, n, n, S))
    PhiR = rng.standard_normal((l2, S, r2))
    x = rng.standard_normal((nvec, r1, n, r2))
    y, pl = model(PhiL, A, PhiR, x)
    # 'lsr,smnS,LSR,vrnR->vlmL' as three tensordots (numpy's own path search for the four-operand
    # einsum takes minutes here)
    t1 = np.tensordot(PhiL, x, axes=([2], [1]))            # l s v n R
    t2 = np.tensordot(A, t1, axes=([0, 2], [1, 3]))          # m S l v R
    ref = np.tensordot(t2, PhiR, axes=([1, 4], [1, 2]))      # m l v L
    ref = ref.transpose(2, 1, 0, 3)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref), pl
