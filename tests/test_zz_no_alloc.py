"""vk_block_solve_system_noalloc (workspace in a static per-device scratch)
against vk_block_solve_system (workspace allocated per call), through the C
ABI: identical results on two calls in a row, the second smaller, for a
chain factorisation and for cyclic reduction; a system beyond the scratch is
refused with VK_ETOOBIG before anything runs."""
import numpy as np
import pytest
import scipy.linalg

from oracle import ros2_oracle as O

from vulcan_amd import _lib
from vulcan_amd import network as nw

gpu = pytest.mark.gpu
VK_ETOOBIG = -1003


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def _sized_network(ni):
    return nw.Network(species=['s%d' % i for i in range(ni)], elements=['H'], compo=np.ones((ni, 1), np.int32),
                      mass=np.full(ni, 2.0), nasa9=np.zeros((ni, 20)),
                      reac=np.array([[0, -2, -2, -2]], np.int32), prod=np.array([[1, -2, -2, -2]], np.int32),
                      kind=np.zeros(1, np.int32), par=np.zeros((1, 8)), rev=np.zeros(1, np.int32))


def _system(rng, nzs, ni):
    L = int(sum(nzs))
    D = rng.standard_normal((L, ni, ni))
    D += np.eye(ni) * (np.abs(D).sum(axis=2).max() * 0.3)
    D[:, 0, 0] = 0.0                        # force pivoting
    Ls = rng.standard_normal((L, ni))
    Us = rng.standard_normal((L, ni))
    o = 0
    for nz in nzs:
        Ls[o] = 0.0
        Us[o + nz - 1] = 0.0
        o += nz
    b = rng.standard_normal((L, ni))
    return D, Ls, Us, b


@gpu
@pytest.mark.parametrize('variant', [1, 4])
def test_noalloc_solve_matches_allocating_solve(variant):
    torch = _torch()
    from vulcan_amd.device import DeviceColumns, DevicePlan, dptr, stream_handle
    lib = _lib.load()
    ni = 93
    plan = DevicePlan(_sized_network(ni))
    _lib.check(lib.vk_plan_set_factor_variant(plan.handle, variant), 'vk_plan_set_factor_variant')
    rng = np.random.default_rng(50 + variant)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device='cuda')
    # two calls in a row on the same scratch, the second smaller
    for nzs in ([150], [40, 23]):
        cols = DeviceColumns(nzs)
        D, Ls, Us, b = _system(rng, nzs, ni)
        Dd, Ld, Ud, bd = T(D), T(Ls), T(Us), T(b)
        x_ref = torch.empty_like(bd)
        x_new = torch.full_like(bd, np.nan)
        info_ref = torch.zeros(len(nzs), dtype=torch.int32, device='cuda')
        info_new = torch.zeros(len(nzs), dtype=torch.int32, device='cuda')
        _lib.check(lib.vk_block_solve_system(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(bd),
                                             dptr(x_ref), 1, dptr(info_ref), stream_handle()), 'vk_block_solve_system')
        _lib.check(lib.vk_block_solve_system_noalloc(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud),
                                                     dptr(bd), dptr(x_new), 1, dptr(info_new), stream_handle()),
                   'vk_block_solve_system_noalloc')
        torch.cuda.synchronize()
        xr, xn = x_ref.cpu().numpy(), x_new.cpu().numpy()
        assert np.array_equal(xr, xn), (variant, nzs)
        assert np.array_equal(info_ref.cpu().numpy(), info_new.cpu().numpy())
        # and it is the solution: column 0 against the banded LAPACK solve
        nz0 = nzs[0]
        ab, bw = O.band_from_blocks(D[:nz0], Ls[:nz0], Us[:nz0])
        xd = scipy.linalg.solve_banded((bw, bw), ab, b[:nz0].ravel())
        assert np.max(np.abs(xn[:nz0].ravel() - xd)) <= 1e-10 * np.max(np.abs(xd))
        cols.close()


@gpu
def test_noalloc_refuses_a_system_beyond_the_scratch():
    torch = _torch()
    from vulcan_amd.device import DeviceColumns, DevicePlan, dptr, stream_handle
    lib = _lib.load()
    ni = 93
    plan = DevicePlan(_sized_network(ni))
    _lib.check(lib.vk_plan_set_factor_variant(plan.handle, 4), 'vk_plan_set_factor_variant')
    cols = DeviceColumns([150, 150])           # cyclic reduction of two columns needs ~166 MB
    D = torch.zeros((300, ni, ni), dtype=torch.float64, device='cuda')
    Ls, Us, b, x = (torch.zeros((300, ni), dtype=torch.float64, device='cuda') for _ in range(4))
    info = torch.zeros(2, dtype=torch.int32, device='cuda')
    rc = lib.vk_block_solve_system_noalloc(plan.handle, cols.handle, dptr(D), dptr(Ls), dptr(Us), dptr(b),
                                           dptr(x), 1, dptr(info), stream_handle())
    assert rc == VK_ETOOBIG
    torch.cuda.synchronize()
    assert not x.any()                         # nothing ran
