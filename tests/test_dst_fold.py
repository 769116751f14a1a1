"""The double even/odd fold of the FP64 sine passes, restated with numpy and checked
against scipy's DST-I (FFTW's RODFT00 convention: X_j = 2 sum_k x_k sin(pi j k / M)).

M = N + 1 = 4Q. Odd rows pair as (2m+1, M-1-2m) through two Q x Q products of the odd-
and even-indexed P_k = x_k + x_{M-k}; even rows are a DST-I of modulus H = 2Q in
Q_k = x_k - x_{M-k} and fold the same way into rows 4m+2 and 4m+4.

The third fold (dst_pass_f64t_kernel) splits each T2 product by the parity of its fold
index: T2[i][Q-1-m] = (-1)^i T2[i][m], so rows m and Q-1-m (m < Q/2) come from the
sums over even and odd i; the other product of each block is read at columns m and
Q-1-m.
"""
import numpy as np
import pytest
import scipy.fft


def double_fold_dst(x):
    N = x.shape[-1]
    M = N + 1
    H, Q = M // 2, M // 4
    xx = np.concatenate([np.zeros(x.shape[:-1] + (1,)), x, np.zeros(x.shape[:-1] + (1,))], -1)
    k = np.arange(1, H)
    P = np.concatenate([xx[..., k] + xx[..., M - k], xx[..., H:H + 1]], -1)  # P_1..P_H
    Qk = np.concatenate([xx[..., k] - xx[..., M - k], np.zeros(x.shape[:-1] + (1,))], -1)
    i = np.arange(Q)
    m = np.arange(Q)
    T1 = 2 * np.sin(np.pi * np.outer(2 * i + 1, 2 * m + 1) / M)
    T2 = 2 * np.sin(np.pi * np.outer(2 * i + 2, 2 * m + 1) / M)
    T4 = 2 * np.sin(np.pi * 4 * np.outer(i[:-1] + 1, m[:-1] + 1) / M)
    A = P[..., 0::2] @ T1          # P_{2i+1}
    B = P[..., 1::2] @ T2          # P_{2i+2}
    q = Qk[..., :Q]                # Q_1..Q_Q
    qh = Qk[..., H - 2 - i]        # Q_{H-1-i}
    Pp = np.where(i == Q - 1, q, q + qh)
    Qp = (q - qh)[..., :Q - 1]
    out = np.zeros(x.shape[:-1] + (N,))
    out[..., 2 * m] = A + B                  # rows 2m+1
    out[..., M - 2 - 2 * m] = A - B          # rows M-1-2m
    out[..., 4 * m + 1] = Pp @ T2            # rows 4m+2
    out[..., 4 * m[:-1] + 3] = Qp @ T4       # rows 4m+4
    return out


def triple_fold_dst(x, slip=False):
    """The third fold with the kernel's row maps: per m < Q/2 the odd block writes rows
    2m+1, M-1-2m, 2m~+1, M-1-2m~ (m~ = Q-1-m) and the even block rows 4m+2, 4m~+2, 4m+4,
    4m~+4 (m~ < Q-1). slip: a planted mistake (E added to row m~ instead of subtracted)."""
    N = x.shape[-1]
    M = N + 1
    H, Q = M // 2, M // 4
    xx = np.concatenate([np.zeros(x.shape[:-1] + (1,)), x, np.zeros(x.shape[:-1] + (1,))], -1)
    k = np.arange(1, H)
    P = np.concatenate([xx[..., k] + xx[..., M - k], xx[..., H:H + 1]], -1)
    Qk = np.concatenate([xx[..., k] - xx[..., M - k], np.zeros(x.shape[:-1] + (1,))], -1)
    i = np.arange(Q)
    m = np.arange(Q // 2)
    mb = Q - 1 - m
    T1 = 2 * np.sin(np.pi * np.outer(2 * i + 1, 2 * np.arange(Q) + 1) / M)
    T2 = 2 * np.sin(np.pi * np.outer(2 * i + 2, 2 * np.arange(Q) + 1) / M)
    T4 = np.zeros((Q, Q))
    T4[:Q - 1, :Q - 1] = 2 * np.sin(np.pi * 4 * np.outer(i[:-1] + 1, i[:-1] + 1) / M)
    q = Qk[..., :Q]
    qh = Qk[..., H - 2 - i]
    Pp = np.where(i == Q - 1, q, q + qh)
    Qp = np.where(i == Q - 1, 0.0, q - qh)
    ev, od = i[0::2], i[1::2]
    sgn = 1.0 if slip else -1.0
    out = np.zeros(x.shape[:-1] + (N,))
    # odd block: T1 on P_{2i+1} at columns m and m~; T2 on P_{2i+2} split by parity of i
    Xf, Xh = P[..., 0::2], P[..., 1::2]
    f, g = Xf @ T1[:, m], Xf @ T1[:, mb]
    o, e = Xh[..., ev] @ T2[ev][:, m], Xh[..., od] @ T2[od][:, m]
    tp, tm = o + e, o + sgn * e
    out[..., 2 * m] = f + tp
    out[..., M - 2 - 2 * m] = f - tp
    out[..., 2 * mb] = g + tm
    out[..., M - 2 - 2 * mb] = g - tm
    # even block: T4 on Q' at columns m and m~; T2 on P' split by parity of i
    f, g = Qp @ T4[:, m], Qp @ T4[:, mb]
    o, e = Pp[..., ev] @ T2[ev][:, m], Pp[..., od] @ T2[od][:, m]
    out[..., 4 * m + 1] = o + e
    out[..., 4 * mb + 1] = o + sgn * e
    out[..., 4 * m + 3] = f
    keep = mb < Q - 1
    out[..., 4 * mb[keep] + 3] = g[..., keep]
    return out


@pytest.mark.parametrize("M", [8, 16, 64, 256, 1024, 2048])
def test_triple_fold_equals_dst1(M):
    rng = np.random.default_rng(M + 1)
    x = rng.standard_normal((3, M - 1))
    ref = scipy.fft.dst(x, type=1, axis=-1)
    got = triple_fold_dst(x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(triple_fold_dst(x, slip=True) - ref).max() > 1e-3 * np.abs(ref).max()


@pytest.mark.parametrize("M", [8, 12, 16, 200, 1024, 2048])
def test_double_fold_equals_dst1(M):
    rng = np.random.default_rng(M)
    x = rng.standard_normal((3, M - 1))
    ref = scipy.fft.dst(x, type=1, axis=-1)
    got = double_fold_dst(x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_double_fold_catches_a_sign_slip():
    """A plausible slip (the even-k half added to the mirrored rows instead of
    subtracted) must fail the comparison."""
    M = 64
    x = np.random.default_rng(1).standard_normal(M - 1)
    ref = scipy.fft.dst(x, type=1)
    got = double_fold_dst(x)
    m = np.arange(M // 4)
    bad = got.copy()
    bad[M - 2 - 2 * m] = got[2 * m]  # A + B where A - B belongs
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(bad - ref).max() > 1e-3 * np.abs(ref).max()


def single_fold_dst(x):
    """The single fold for any M = N + 1: H = floor(M/2) pairs (k, M - k); for even M the
    middle element x_H has no partner. Odd rows from P_k, even rows from Q_k."""
    N = x.shape[-1]
    M = N + 1
    H = M // 2
    E = H if M % 2 else H - 1
    xx = np.concatenate([np.zeros(x.shape[:-1] + (1,)), x, np.zeros(x.shape[:-1] + (1,))], -1)
    k = np.arange(1, H + 1)
    partner = np.where((M % 2 == 1) | (k < H), xx[..., M - k], 0.0)
    P = xx[..., k] + partner
    Q = (xx[..., k] - partner)[..., :E]
    m = np.arange(H)
    So = 2 * np.sin(np.pi * np.outer(k, 2 * m + 1) / M)
    Se = 2 * np.sin(np.pi * np.outer(k[:E], 2 * (m[:E] + 1)) / M)
    out = np.zeros(x.shape[:-1] + (N,))
    out[..., 2 * m] = P @ So            # rows 2m+1
    out[..., 2 * m[:E] + 1] = Q @ Se    # rows 2m+2
    return out


@pytest.mark.parametrize("M", [8, 9, 10, 15, 16, 61, 150, 151])
def test_single_fold_equals_dst1_any_length(M):
    """Odd lengths and lengths not divisible by 4 take the single fold (the double fold
    needs M divisible by 4)."""
    rng = np.random.default_rng(M)
    x = rng.standard_normal((2, M - 1))
    ref = scipy.fft.dst(x, type=1, axis=-1)
    assert np.abs(single_fold_dst(x) - ref).max() <= 1e-12 * np.abs(ref).max()
