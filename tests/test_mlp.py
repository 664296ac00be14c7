"""Warm-start MLP kernels against the CPU oracle (oracle/mlp_oracle.py)."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO

gpu = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def test_oracle_gradients_by_finite_differences():
    """The oracle's analytic gradients (without BF16 rounding) against central
    differences of its own loss on a tiny network."""
    rng = np.random.default_rng(0)
    sizes = [5, 7, 6, 4]
    W = [rng.standard_normal((a, c)) * 0.5 for a, c in zip(sizes[:-1], sizes[1:])]
    b = [rng.standard_normal(c) * 0.1 for c in sizes[1:]]
    X = rng.standard_normal((3, 5)); T = rng.standard_normal((3, 4))
    ident = lambda x: np.asarray(x, dtype=np.float64)
    saved = MO.bf16
    MO.bf16 = ident
    try:
        loss, _, gW, gb = MO.loss_and_grads(W, b, X, T)
        h = 1e-6
        for l in range(3):
            for (i, j) in [(0, 0), (1, 2), (sizes[l] - 1, sizes[l + 1] - 1)]:
                Wp = [w.copy() for w in W]; Wp[l][i, j] += h
                Wm = [w.copy() for w in W]; Wm[l][i, j] -= h
                fd = (MO.loss_and_grads(Wp, b, X, T)[0] - MO.loss_and_grads(Wm, b, X, T)[0]) / (2 * h)
                assert abs(fd - gW[l][i, j]) <= 1e-6 * max(1.0, abs(fd))
            bp = [x.copy() for x in b]; bp[l][0] += h
            bm = [x.copy() for x in b]; bm[l][0] -= h
            fd = (MO.loss_and_grads(W, bp, X, T)[0] - MO.loss_and_grads(W, bm, X, T)[0]) / (2 * h)
            assert abs(fd - gb[l][0]) <= 1e-6 * max(1.0, abs(fd))
    finally:
        MO.bf16 = saved


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5e-3], dtype=np.float32)
    r = MO.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 4 * 2 ** -8 and r[3] == 1.0
    assert abs(r[4] - (-2.5e-3)) <= 2 ** -8 * 2.5e-3


@gpu
@pytest.mark.parametrize('M,N,K', [(70, 45, 33), (64, 64, 64), (5, 130, 300)])
def test_gemm_layouts(M, N, K):
    torch = _torch()
    from vulcan_amd.mlp import MLP, to_bf16, bf16_to_f32
    from vulcan_amd import _lib
    from vulcan_amd.device import dptr, stream_handle
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    ref = MO.bf16(A) @ MO.bf16(B) + bias
    lib = _lib.load()
    dev = 'cuda'
    for a_t in (False, True):
        for b_t in (False, True):
            Ad = to_bf16(torch.as_tensor(A.T.copy() if a_t else A, device=dev))
            Bd = to_bf16(torch.as_tensor(B.T.copy() if b_t else B, device=dev))
            sam, sak = (1, M) if a_t else (K, 1)
            sbk, sbn = (1, K) if b_t else (N, 1)
            Cd = torch.empty((M, N), dtype=torch.float32, device=dev)
            Y = torch.empty((M, N), dtype=torch.int16, device=dev)
            _lib.check(lib.vk_mlp_gemm(M, N, K, dptr(Ad), sam, sak, dptr(Bd), sbk, sbn, dptr(Cd), N,
                                       dptr(torch.as_tensor(bias, device=dev)), 0, dptr(Y), N, stream_handle()),
                       'gemm')
            c = Cd.cpu().numpy()
            assert np.max(np.abs(c - ref)) <= 1e-5 * np.abs(ref).max() * np.sqrt(K), (a_t, b_t)
            y = bf16_to_f32(Y).cpu().numpy()
            assert np.max(np.abs(y - MO.silu(c))) <= 2 ** -7 * (np.abs(MO.silu(c)).max() + 1e-3)


@gpu
def test_mlp_forward_backward_adam_against_oracle():
    torch = _torch()
    from vulcan_amd.mlp import MLP
    sizes = [37, 64, 96, 50]
    net = MLP(sizes, seed=3)
    rng = np.random.default_rng(7)
    X = rng.standard_normal((20, 37)).astype(np.float32)
    T = rng.standard_normal((20, 50)).astype(np.float32)
    mask = (rng.random((20, 50)) > 0.2).astype(np.float32)
    W = [w.cpu().numpy().astype(np.float64) for w in net.W]
    b = [x.cpu().numpy().astype(np.float64) for x in net.b]
    loss_ref, Y_ref, gW_ref, gb_ref = MO.loss_and_grads(W, b, X.astype(np.float64), T.astype(np.float64), mask)
    dev = 'cuda'
    loss, Y = net.loss_and_backward(torch.as_tensor(X, device=dev), torch.as_tensor(T, device=dev),
                                    torch.as_tensor(mask, device=dev))
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    assert np.max(np.abs(Y.cpu().numpy() - Y_ref)) <= 1e-4 * np.abs(Y_ref).max()
    L = net.nlayers
    for l in range(L):
        g = net.grads[l].cpu().numpy()
        assert np.max(np.abs(g - gW_ref[l])) <= 5e-3 * np.abs(gW_ref[l]).max(), l
        gb = net.grads[L + l].cpu().numpy()
        assert np.max(np.abs(gb - gb_ref[l])) <= 5e-3 * np.abs(gb_ref[l]).max(), l
    # one Adam step
    w0 = net.W[0].cpu().numpy().astype(np.float64)
    g0 = net.grads[0].cpu().numpy().astype(np.float64)
    net.adam_step(lr=1e-3, wd=1e-4)
    w1_ref, _, _ = MO.adam(w0, g0, np.zeros_like(w0), np.zeros_like(w0), 1, lr=1e-3, wd=1e-4)
    assert np.max(np.abs(net.W[0].cpu().numpy() - w1_ref)) <= 1e-6
