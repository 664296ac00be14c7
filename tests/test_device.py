"""HIP kernels against the CPU oracle and the reference's golden vectors.

Every GPU test calls the library through its C ABI (ctypes) on device
buffers.  Bars:

* bit-exact where the kernel follows the reference's expression tree and the
  inputs are identical: chemdf, diffdf, Jacobian blocks, lhs blocks;
* rate constants to 1e-12 relative (device exp/log/pow differ from glibc in
  the last bit; exp(-E/T) with E/T up to 60 amplifies that ~60-fold);
* linear solves and Ros2 attempts to 1e-10 relative per step, relative to
  max(|y|, 1e-10 x layer total) (entries below a mixing ratio of 1e-10 are
  held to 1e-20 of the layer total); LAPACK's banded LU itself is only
  accurate to ~6e-11 on the stiffest early steps, so the device solves are
  refined once with a double-double residual to stay well inside the bar;
* batching: a column's result does not depend on the other columns of the
  batch (bitwise).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from _golden_util import load, network_from, column_from, traj_config
from oracle import ros2_oracle as O
from vulcan_amd import _lib
from vulcan_amd import network as nw
from vulcan_amd import workloads as W
from vulcan_amd.config import Config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, 'include', 'vulcan_amd.h')).read()
    declared = set(re.findall(r'^int\s+(vk_\w+)\s*\(', hdr, flags=re.M))
    assert declared, 'no declarations parsed'
    assert declared == set(_lib.EXPORTS)
    lib = _lib.load()                      # loading needs no device
    for name in declared:
        assert hasattr(lib, name), name


def test_plan_refuses_networks_beyond_the_compiled_limits():
    """Refused before anything touches the device: more species than the
    kernels are compiled for (VK_NI_LIMIT = 160), more pair rates or
    derivative slots than a layer's 64 KiB of LDS holds, empty networks."""
    lib = _lib.load()

    def create(**kw):
        d = _lib.NetworkDesc()
        d.ni, d.npairs, d.natom = 10, 10, 1
        for k, v in kw.items():
            setattr(d, k, v)
        h = C.c_void_p()
        return lib.vk_plan_create(C.byref(d), C.byref(h))
    assert create(ni=161) == -1003          # VK_ETOOBIG
    assert create(npairs=9000) == -1003
    assert create(nslot=9000) == -1003
    assert create(ni=0) == -1001            # VK_EINVAL
    assert create(npairs=0) == -1001


gpu = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def step_err(a, b, ysum):
    """(max |a-b| / max(|b|, 1e-10 * layer total): relative error with a floor
        at mixing ratio 1e-10,  max |a-b| / layer total)"""
    a = np.asarray(a); b = np.asarray(b)
    den = np.maximum(np.abs(b), 1e-10 * ysum[:, None])
    rel = np.max(np.abs(a - b) / den)
    absn = np.max(np.abs(a - b) / ysum[:, None])
    return rel, absn


def _sized_network(ni):
    """Smallest valid network with ni species (factor/solve only use the size)."""
    return nw.Network(species=['s%d' % i for i in range(ni)], elements=['H'], compo=np.ones((ni, 1), np.int32),
                      mass=np.full(ni, 2.0), nasa9=np.zeros((ni, 20)),
                      reac=np.array([[0, -2, -2, -2]], np.int32), prod=np.array([[min(1, ni - 1), -2, -2, -2]], np.int32),
                      kind=np.zeros(1, np.int32), par=np.zeros((1, 8)), rev=np.zeros(1, np.int32))


@pytest.fixture(scope='module')
def golden():
    return load()


def _batch(net, cols, y0, cfg, history=0, k=None, alpha=None):
    from vulcan_amd.device import DevicePlan, ColumnBatch
    plan = DevicePlan(net, alpha=alpha)
    return ColumnBatch(plan, cols, y0, cfg, history=history, k=k)


@gpu
def test_rates_match_oracle():
    torch = _torch()
    c = W.full_case()
    b = _batch(c.net, [c.col], [c.y0], c.cfg)
    torch.cuda.synchronize()
    kd = b.k.cpu().numpy()
    ko = O.rates(c.net, c.col.Tco, c.col.M, c.col.tau)
    nzm = ko != 0
    assert np.array_equal(kd == 0, ko == 0)
    rel = np.max(np.abs(kd[nzm] - ko[nzm]) / np.abs(ko[nzm]))
    print('rates: max rel %.3e' % rel)
    assert rel < 1e-12


@gpu
def test_chemdf_diffdf_bitwise_against_reference(golden):
    torch = _torch()
    d = golden
    net, col = network_from(d), column_from(d)
    b = _batch(net, [col], [d['y0']], Config(), k=d['k'], alpha=d['alpha'])
    lib = _lib.load()
    out = torch.empty_like(b.y)
    from vulcan_amd.device import dptr, stream_handle
    _lib.check(lib.vk_chemdf(b.plan.handle, b.layout.handle, dptr(b.y), dptr(b.M), dptr(b.k), dptr(out),
                             stream_handle()), 'vk_chemdf')
    assert np.array_equal(out.cpu().numpy(), d['chemdf'])
    for mol, key in ((1, 'diffdf_mol'), (0, 'diffdf_nomol')):
        _lib.check(lib.vk_diffdf(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y), mol, dptr(out),
                                 stream_handle()), 'vk_diffdf')
        assert np.array_equal(out.cpu().numpy(), d[key]), key


@gpu
@pytest.mark.parametrize('which', ['golden', 'full'])
def test_jac_and_lhs_bitwise_against_oracle(golden, which):
    torch = _torch()
    from vulcan_amd.device import dptr, stream_handle
    if which == 'golden':
        d = golden
        net, col, y, k, alpha = network_from(d), column_from(d), d['y0'], d['k'], d['alpha']
    else:
        c = W.full_case(nz=40)
        net, col, y, alpha = c.net, c.col, c.y0, None
        k = O.rates(net, col.Tco, col.M, col.tau)
    b = _batch(net, [col], [y], Config(), k=k, alpha=alpha)
    lib = _lib.load()
    nz, ni = y.shape
    J = torch.empty((nz, ni, ni), dtype=torch.float64, device='cuda')
    _lib.check(lib.vk_jac(b.plan.handle, b.layout.handle, dptr(b.y), dptr(b.M), dptr(b.k), dptr(J),
                          stream_handle()), 'vk_jac')
    Jo = O.jac_blocks(net, y, col.M, k)
    assert np.array_equal(J.cpu().numpy(), Jo)
    dt = 3.0e-9
    dtc = torch.full((1,), dt, dtype=torch.float64, device='cuda')
    Ls = torch.empty((nz, ni), dtype=torch.float64, device='cuda')
    Us = torch.empty_like(Ls)
    for mol in (1, 0):
        _lib.check(lib.vk_lhs(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y), dptr(b.k), mol,
                              dptr(dtc), dptr(J), dptr(Ls), dptr(Us), stream_handle()), 'vk_lhs')
        D, L, U = O.lhs_blocks(col, y, dt, Jo, bool(mol))
        assert np.array_equal(J.cpu().numpy(), D)
        assert np.array_equal(Ls.cpu().numpy(), L)
        assert np.array_equal(Us.cpu().numpy(), U)
        if which == 'golden':
            key = 'tot' if mol else 'nomol'
            assert np.array_equal(Ls.cpu().numpy(), d['lhs_%s_L' % key])


@gpu
def test_mfma_f64_operand_layout():
    """The FP64 MFMA lane layout the blocked inversion is written for."""
    torch = _torch()
    from vulcan_amd.device import dptr, stream_handle
    out = torch.zeros(256, dtype=torch.float64, device='cuda')
    _lib.check(_lib.load().vk_selftest_mfma_f64(dptr(out), stream_handle()), 'vk_selftest_mfma_f64')
    i = np.arange(16)[:, None]; k = np.arange(4)[None, :]
    A = i + 0.25 * k + 1.0
    B = 1.0 - 0.5 * np.arange(4)[:, None] + 0.125 * np.arange(16)[None, :]
    assert np.array_equal(out.cpu().numpy().reshape(16, 16), A @ B)


@gpu
def test_wave_max_u32():
    """The DPP + readlane wave maximum of the blocked pivot search."""
    torch = _torch()
    from vulcan_amd.device import dptr, stream_handle
    rng = np.random.default_rng(3)
    for trial in range(4):
        v = rng.integers(0, 2 ** 31, 64, dtype=np.int64).astype(np.uint32)
        if trial == 1:
            v[:] = 0; v[37] = 5
        if trial == 2:
            v[63] = 2 ** 32 - 7                 # top bit set: compared unsigned
        vin = torch.as_tensor(v.view(np.int32), device='cuda')
        out = torch.zeros_like(vin)
        _lib.check(_lib.load().vk_selftest_wave_max_u32(dptr(vin), dptr(out), stream_handle()),
                   'vk_selftest_wave_max_u32')
        assert (out.cpu().numpy().view(np.uint32) == v.max()).all(), trial


@gpu
def test_pivot_reciprocal_matches_ieee_division():
    """The pivot step's reciprocal (hardware estimate and two Newton-Raphson
    steps) against IEEE 1/d computed on the host, over magnitudes 2^-600 to
    2^600 of both signs and the powers of two; it is allowed one ulp, and
    in practice is exact."""
    torch = _torch()
    from vulcan_amd.device import dptr, stream_handle
    rng = np.random.default_rng(11)
    n = 1 << 20
    d = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-600, 601, n)) * rng.choice([-1.0, 1.0], n)
    d[:1201] = np.ldexp(1.0, np.arange(-600, 601))
    vin = torch.as_tensor(d, device='cuda')
    out = torch.empty_like(vin)
    _lib.check(_lib.load().vk_selftest_recip_f64(n, dptr(vin), dptr(out), stream_handle()), 'vk_selftest_recip_f64')
    got = out.cpu().numpy()
    ref = 1.0 / d
    ulp = np.abs(got.view(np.int64) - ref.view(np.int64))
    print('max ulp %d, %d of %d differ' % (ulp.max(), int((ulp != 0).sum()), n))
    assert ulp.max() <= 1


FACTOR_VARIANTS = [0, 1, 2, 3, 4]   # register GJ, blocked MFMA (1 or 2 WGs per CU), two-wave, cyclic reduction


@gpu
@pytest.mark.parametrize('variant', FACTOR_VARIANTS)
@pytest.mark.parametrize('ni,nz', [(30, 60), (93, 150), (150, 12), (7, 3), (16, 2), (40, 3), (17, 5), (112, 9),
                                   (131, 7), (64, 4), (65, 3), (128, 3), (145, 4), (160, 3)])
def test_block_factor_solve_against_dense(ni, nz, variant):
    torch = _torch()
    if variant == 3 and ni > 112:
        pytest.skip('the two-wave variant holds blocks up to ni = 112')
    from vulcan_amd.device import DeviceColumns, DevicePlan, dptr, stream_handle
    rng = np.random.default_rng(ni * 1000 + nz)
    D = rng.standard_normal((nz, ni, ni))
    D += np.eye(ni) * (np.abs(D).sum(axis=2).max() * 0.3)
    D[:, 0, 0] = 0.0                        # force pivoting
    L = rng.standard_normal((nz, ni)); L[0] = 0
    U = rng.standard_normal((nz, ni)); U[-1] = 0
    b = rng.standard_normal((nz, ni))
    A = O.lhs_dense(D, L, U)
    x = np.linalg.solve(A, b.ravel()).reshape(nz, ni)
    net = _sized_network(ni)
    plan = DevicePlan(net)
    lib = _lib.load()
    _lib.check(lib.vk_plan_set_factor_variant(plan.handle, variant), 'vk_plan_set_factor_variant')
    cols = DeviceColumns([nz])
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device='cuda')
    Dd, Ld, Ud, bd = T(D), T(L), T(U), T(b)
    Al = A.astype(np.longdouble)
    res = lambda v: (b.ravel().astype(np.longdouble) - Al @ v.ravel().astype(np.longdouble))
    info = torch.zeros(1, dtype=torch.int32, device='cuda')
    # the whole solve, with and without one refinement step (every variant)
    xs = []
    for refine in (0, 1):
        xo = torch.empty_like(bd)
        _lib.check(lib.vk_block_solve_system(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(bd),
                                             dptr(xo), refine, dptr(info), stream_handle()), 'vk_block_solve_system')
        xs.append(xo.cpu().numpy())
    assert int(info.cpu()[0]) == 0
    assert np.array_equal(Dd.cpu().numpy(), D)          # the blocks are not modified
    assert np.max(np.abs(xs[0] - x)) <= 1e-12 * np.abs(x).max()
    assert np.max(np.abs(res(xs[1]))) <= np.max(np.abs(res(xs[0]))) + 1e-15 * np.abs(b).max()
    if variant == 4:
        # the factor/solve pair of the public ABI needs a workspace the cyclic reduction keeps itself
        Di = torch.empty_like(Dd)
        assert lib.vk_block_factor(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(Di), None,
                                   stream_handle()) == -1001
        return
    Di = torch.empty_like(Dd)
    _lib.check(lib.vk_block_factor(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(Di), dptr(info),
                                   stream_handle()), 'vk_block_factor')
    xd_ = bd.clone()
    _lib.check(lib.vk_block_solve(plan.handle, cols.handle, dptr(Di), dptr(Ld), dptr(Ud), dptr(xd_),
                                  stream_handle()), 'vk_block_solve')
    xd = xd_.cpu().numpy()
    assert int(info.cpu()[0]) == 0
    assert np.max(np.abs(xd - x)) <= 1e-12 * np.abs(x).max()
    # one refinement step with the double-double residual
    r = torch.empty_like(bd)
    _lib.check(lib.vk_block_residual(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(xd_), dptr(bd),
                                     dptr(r), stream_handle()), 'vk_block_residual')
    rd = r.cpu().numpy()
    r_ref = res(xd).astype(np.float64)
    assert np.max(np.abs(rd.ravel() - r_ref)) <= 1e-3 * np.abs(r_ref).max() + 1e-300
    _lib.check(lib.vk_block_solve(plan.handle, cols.handle, dptr(Di), dptr(Ld), dptr(Ud), dptr(r),
                                  stream_handle()), 'vk_block_solve')
    x2 = xd + r.cpu().numpy()
    # the refined solution has the smaller residual (numpy's own solution is
    # no exact reference at this conditioning)
    assert np.max(np.abs(res(x2))) <= np.max(np.abs(res(xd))) + 1e-15 * np.abs(b).max()


@gpu
@pytest.mark.parametrize('ni,nzs', [(21, [2, 3, 5, 8, 13, 150, 4]),    # 1 to 7 levels
                                    (21, [6, 150, 9, 2, 3, 5]),       # odd counts at several levels
                                    (93, [37, 40]),                   # three row groups per block
                                    (93, [75, 2])])
def test_cyclic_reduction_ragged_batch_matches_dense(ni, nzs):
    """Cyclic reduction over columns of different heights."""
    torch = _torch()
    from vulcan_amd.device import DeviceColumns, DevicePlan, dptr, stream_handle
    rng = np.random.default_rng(5 + len(nzs) + ni)
    plan = DevicePlan(_sized_network(ni))
    lib = _lib.load()
    _lib.check(lib.vk_plan_set_factor_variant(plan.handle, 4), 'vk_plan_set_factor_variant')
    cols = DeviceColumns(nzs)
    Ds, Ls, Us, bs, xs = [], [], [], [], []
    for nz in nzs:
        # the shift keeps the blocks' eigenvalues (a disc of radius ~sqrt(ni)) off zero
        D = rng.standard_normal((nz, ni, ni)) + 6 * max(1.0, np.sqrt(ni / 21)) * np.eye(ni)
        L = rng.standard_normal((nz, ni)); L[0] = 0
        U = rng.standard_normal((nz, ni)); U[-1] = 0
        b = rng.standard_normal((nz, ni))
        xs.append(np.linalg.solve(O.lhs_dense(D, L, U), b.ravel()).reshape(nz, ni))
        Ds.append(D); Ls.append(L); Us.append(U); bs.append(b)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(np.concatenate(a)), dtype=torch.float64, device='cuda')
    Dd, Ld, Ud, bd = T(Ds), T(Ls), T(Us), T(bs)
    xo = torch.empty_like(bd)
    info = torch.zeros(len(nzs), dtype=torch.int32, device='cuda')
    _lib.check(lib.vk_block_solve_system(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(bd), dptr(xo),
                                         0, dptr(info), stream_handle()), 'vk_block_solve_system')
    assert not info.cpu().numpy().any()
    x = np.concatenate(xs)
    assert np.max(np.abs(xo.cpu().numpy() - x)) <= 1e-12 * np.abs(x).max()


@gpu
def test_integration_with_cyclic_reduction_matches_chain():
    """The same batch integrated with cyclic reduction and with the chained
    block Thomas factorisation agrees to rounding over 20 attempts."""
    torch = _torch()
    cases = W.batched_cases(3, nz_range=(12, 30))
    net = cases[0].net
    runs = []
    for variant in (0, 4):
        b = _batch(net, [c.col for c in cases], [c.y0 for c in cases], cases[0].cfg)
        _lib.check(_lib.load().vk_plan_set_factor_variant(b.plan.handle, variant), 'vk_plan_set_factor_variant')
        b.set_graph(True)
        b.integrate(20, use_conv=False)
        torch.cuda.synchronize()
        runs.append((b.y.cpu().numpy(), b.host_state()))
    (y0, s0), (y4, s4) = runs
    assert np.array_equal(s0['count'], s4['count'])
    rel, absn = step_err(y4, y0, y0.sum(axis=1))
    print('cyclic reduction vs chain after 20 attempts: rel %.3e abs %.3e' % (rel, absn))
    assert rel < 1e-9


@gpu
def test_zero_pivot_is_reported():
    """A singular block yields info = its layer index + 1 (LAPACK-style), for every variant."""
    torch = _torch()
    from vulcan_amd.device import DeviceColumns, DevicePlan, dptr, stream_handle
    ni, nz = 20, 6
    D = np.tile(np.eye(ni), (nz, 1, 1)) * 3.0
    D[1, 5, :] = 0.0                         # layer 1 singular (row 5 zero), lower chain
    L = np.zeros((nz, ni)); U = np.zeros((nz, ni))
    lib = _lib.load()
    for variant in FACTOR_VARIANTS:
        plan = DevicePlan(_sized_network(ni))   # ni = 20: every variant applies
        _lib.check(lib.vk_plan_set_factor_variant(plan.handle, variant), 'vk_plan_set_factor_variant')
        cols = DeviceColumns([nz])
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device='cuda')
        Dd, Ld, Ud = T(D), T(L), T(U)
        Di = torch.empty_like(Dd)
        info = torch.zeros(1, dtype=torch.int32, device='cuda')
        if variant == 4:
            b = torch.ones((nz, ni), dtype=torch.float64, device='cuda')
            xo = torch.empty_like(b)
            _lib.check(lib.vk_block_solve_system(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(b),
                                                 dptr(xo), 0, dptr(info), stream_handle()), 'vk_block_solve_system')
        else:
            _lib.check(lib.vk_block_factor(plan.handle, cols.handle, dptr(Dd), dptr(Ld), dptr(Ud), dptr(Di),
                                           dptr(info), stream_handle()), 'vk_block_factor')
        assert int(info.cpu()[0]) == 2, variant


@gpu
def test_factor_tuner_picks_the_fastest_variant():
    torch = _torch()
    import ctypes as C
    from vulcan_amd.device import DeviceColumns, DevicePlan, stream_handle
    lib = _lib.load()
    plan = DevicePlan(_sized_network(93))
    cols = DeviceColumns([40, 23, 60])
    ms = (C.c_float * len(FACTOR_VARIANTS))()
    _lib.check(lib.vk_tune_factor(plan.handle, cols.handle, 2, ms, stream_handle()), 'vk_tune_factor')
    t = list(ms)
    print('factor variants ms:', t)
    assert all(v > 0 for v in t)                       # ni = 93, 3 columns: every variant applies
    assert lib.vk_plan_factor_variant(plan.handle) == int(np.argmin(t))
    big = DevicePlan(_sized_network(150))
    big_cols = DeviceColumns([7])
    ms2 = (C.c_float * len(FACTOR_VARIANTS))()
    _lib.check(lib.vk_tune_factor(big.handle, big_cols.handle, 1, ms2, stream_handle()), 'vk_tune_factor')
    assert ms2[3] == -1.0 and all(ms2[v] > 0 for v in (0, 1, 2, 4))
    assert lib.vk_plan_set_factor_variant(big.handle, 3) == -1001        # VK_EINVAL


@gpu
def test_solver_attempt_against_reference(golden):
    _torch()
    d = golden
    net, col = network_from(d), column_from(d)
    b = _batch(net, [col], [d['y0']], Config(), k=d['k'], alpha=d['alpha'])
    st = b.host_state()
    st['dt'] = float(d['solver_dt'])
    b.set_host_state(st)
    b.attempt()
    sol = b.y.cpu().numpy()
    ref = d['solver_sol']
    rel, absn = step_err(sol, ref, ref.sum(axis=1))
    print('golden attempt: rel %.3e  abs/total %.3e' % (rel, absn))
    assert rel < 1e-10
    dl = float(b.host_state()['delta'][0])
    assert abs(dl - float(d['solver_delta'])) <= 1e-9 * float(d['solver_delta'])


EXACT_BERR_MAX = 1e-15   # backward error of the refined ("exact") solves: FP64 rounding


def assert_step_parity(dev, ref, exact, label, exact_berr=0.0):
    """The device step must agree with the reference step (the oracle's banded
    LAPACK solve) to 1e-10 per entry (relative, floored at mixing ratio
    1e-10) and 1e-13 of the layer total.  Where the linear systems are so
    ill-conditioned that the reference's own forward error exceeds that, the
    device step must be at least as close to the exact step (the same step
    with refined solves, whose backward error must be at FP64 rounding) as
    the reference is."""
    assert exact_berr < EXACT_BERR_MAX, (label, exact_berr)
    ysum = exact.sum(axis=1)
    rel, absn = step_err(dev, ref, ysum)
    rel_d, abs_d = step_err(dev, exact, ysum)
    rel_r, abs_r = step_err(ref, exact, ysum)
    print('%s: device-reference rel %.3e abs %.3e | device-exact rel %.3e abs %.3e | reference-exact rel %.3e abs %.3e'
          % (label, rel, absn, rel_d, abs_d, rel_r, abs_r))
    if rel < 1e-10 and absn < 1e-13:
        return max(rel, 0.0), absn
    assert rel_d <= max(1e-10, rel_r), label
    assert abs_d <= max(1e-13, abs_r), label
    return rel, absn


@gpu
@pytest.mark.parametrize('which', ['small', 'full'])
def test_solver_attempt_against_oracle(which):
    _torch()
    import copy
    c = W.small_case() if which == 'small' else W.full_case()
    k = O.rates(c.net, c.col.Tco, c.col.M, c.col.tau)
    b = _batch(c.net, [c.col], [c.y0], c.cfg, k=k)
    for dt in (1e-10, 1e-6):
        st0 = O.new_state(c.net, c.y0, k, c.cfg)
        st0.dt = dt
        stx = copy.deepcopy(st0)
        O.ros2_solver(c.net, c.col, st0, c.cfg)
        O.ros2_solver(c.net, c.col, stx, c.cfg, exact=True)
        b.y.copy_(__import__('torch').as_tensor(c.y0))
        b.ymix.copy_(__import__('torch').as_tensor(c.y0 / c.y0.sum(axis=1)[:, None]))
        st = b.host_state(); st['dt'] = dt; b.set_host_state(st)
        b.attempt()
        assert_step_parity(b.y.cpu().numpy(), st0.y, stx.y, '%s dt=%g' % (which, dt), stx.exact_berr)


def _push_state(b, st, col, cfg):
    """Copy an oracle state (and the column geometry) into the device batch."""
    import torch
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64)
    b.y.copy_(T(st.y)); b.ymix.copy_(T(st.ymix)); b.y_prev.copy_(T(st.y))
    b.atom_ini.copy_(T(st.atom_ini[None])); b.atom_loss.copy_(T(st.atom_loss[None]))
    b.atom_loss_prev.copy_(T(st.atom_loss[None]))
    nz = col.nz
    b.g.copy_(T(col.g)); b.Hp.copy_(T(col.Hp)); b.dz.copy_(T(col.dz)); b.zco.copy_(T(col.zco))
    pad = lambda v: np.append(v, 0.0)
    b.dzi.copy_(T(pad(col.dzi))); b.Hpi.copy_(T(pad(col.Hpi)))
    h = b.host_state()
    h['t'], h['dt'], h['count'], h['done'] = st.t, st.dt, st.count, 0
    b.set_host_state(h)


@gpu
def test_per_step_parity_small_200_steps():
    """The small configuration: 200 Ros2 steps from dt = 1e-10; each device
    step starts from the oracle's state and must match the oracle's step
    (to the reference's own accuracy where its linear solves are
    ill-conditioned, see assert_step_parity)."""
    _torch()
    import copy
    c = W.small_case()
    cfg = c.cfg
    k = O.rates(c.net, c.col.Tco, c.col.M, c.col.tau)
    b = _batch(c.net, [c.col], [c.y0], cfg, k=k)
    st = O.new_state(c.net, c.y0, k, cfg)
    tb = nw.build_tables(c.net)
    worst_rel = worst_abs = 0.0
    loose = 0

    def oracle_step(s_, exact):
        s_.y_prev = s_.y.copy(); s_.atom_loss_prev = s_.atom_loss.copy()
        O.one_step(c.net, c.col, s_, cfg, tb, exact=exact)
        if s_.count % cfg.update_frq == 0:
            O.update_mu_dz(c.col, s_)
        s_.y = np.vstack(c.col.M) * s_.ymix
        s_.t += s_.dt; s_.count += 1
        O.step_size(s_, cfg)

    for step in range(200):
        _push_state(b, st, c.col, cfg)
        col0 = copy.deepcopy(c.col)
        stx = copy.deepcopy(st)
        oracle_step(st, False)
        col_ref = c.col
        c.col = col0                       # the exact step starts from the same geometry
        oracle_step(stx, True)
        c.col = col_ref
        # device step: attempts until the count advances
        for _ in range(64):
            b.integrate(1, use_conv=False)
            if int(b.host_state()['count'][0]) == st.count:
                break
        h = b.host_state()[0]
        assert int(h['count']) == st.count
        rel, absn = assert_step_parity(b.y.cpu().numpy(), st.y, stx.y, 'step %d' % step, stx.exact_berr)
        loose += rel >= 1e-10 or absn >= 1e-13
        worst_rel, worst_abs = max(worst_rel, rel), max(worst_abs, absn)
        ddt = abs(h['dt'] - st.dt) / st.dt
        assert ddt <= 1e-9 or abs(h['dt'] - stx.dt) <= max(1e-9 * stx.dt, 2.0 * abs(st.dt - stx.dt)), step
    print('per-step parity over 200 steps: worst device-reference rel %.3e abs/total %.3e; %d steps decided '
          'against the exact step' % (worst_rel, worst_abs, loose))


@gpu
def test_free_running_trajectory_against_reference(golden):
    """The reference's 61-step Integration trajectory (rejections,
    update_mu_dz every 10 steps, conv) on the device."""
    _torch()
    d = golden
    net, col = network_from(d), column_from(d)
    for key in ('g', 'Hp', 'dz', 'zco', 'dzi', 'Ti', 'Hpi'):
        setattr(col, key, d['trajg0_' + key].copy())
    cfg = traj_config(d)
    b = _batch(net, [col], [d['y0']], cfg, history=cfg.conv_step + 1, k=d['k'], alpha=d['alpha'])
    for _ in range(400):
        b.integrate(8)
        if int(b.host_state()['done'][0]):
            break
    h = b.host_state()[0]
    fin = d['traj_final']
    assert int(h['end_case']) == int(fin[4]) and int(h['count']) == int(fin[5])
    assert [int(h['delta_count']), int(h['nega_count']), int(h['loss_count'])] == [int(x) for x in fin[6:9]]
    yref = d['traj_y'][-1]
    rel, absn = step_err(b.y.cpu().numpy(), yref, yref.sum(axis=1))
    print('61-step trajectory: rel %.3e abs/total %.3e, longdy %.6e vs %.6e' % (rel, absn, h['longdy'], fin[2]))
    assert rel < 1e-7 and absn < 1e-12
    assert abs(h['longdy'] - fin[2]) <= 1e-6 * abs(fin[2])


@gpu
def test_batch_is_column_independent():
    torch = _torch()
    cases = W.batched_cases(3, nz_range=(12, 30))
    net = cases[0].net
    k_all = None
    b = _batch(net, [c.col for c in cases], [c.y0 for c in cases], cases[0].cfg)
    b.integrate(6, use_conv=False)
    for i, c in enumerate(cases):
        bi = _batch(net, [c.col], [c.y0], c.cfg)
        bi.integrate(6, use_conv=False)
        assert np.array_equal(b.column('y', i), bi.y.cpu().numpy())
        assert b.host_state()[i]['count'] == bi.host_state()[0]['count']
    del k_all


@gpu
def test_graph_replay_is_bitwise_eager():
    """The captured per-attempt graph, replayed many times with host polls
    and eager work in between, gives exactly the eager launches' results;
    a parameter change forces a re-capture."""
    torch = _torch()
    cases = W.batched_cases(3, nz_range=(12, 30))
    net, cfg = cases[0].net, cases[0].cfg
    cfg.conv_step = 20
    runs = []
    for graph in (False, True):
        b = _batch(net, [c.col for c in cases], [c.y0 for c in cases], cfg, history=cfg.conv_step + 1)
        b.set_graph(graph)
        for rep in range(6):
            b.integrate(7)
            b.host_state()                      # synchronising poll between replays
            b.atom_loss_prev.add_(0.0)          # eager torch work on the same stream
        b.integrate(5, use_update_mu=False)     # different scalar arguments: new capture
        b.integrate(3)
        torch.cuda.synchronize()
        runs.append((b.y.cpu().numpy(), b.ymix.cpu().numpy(), b.y_prev.cpu().numpy(),
                     b.host_state(), b.atom_loss.cpu().numpy()))
    e, g = runs
    for a_, b_ in zip(e, g):
        assert np.array_equal(np.asarray(a_).view(np.uint8), np.asarray(b_).view(np.uint8))
    assert e[3]['attempts'].min() == 6 * 7 + 8 or e[3]['done'].any()


@gpu
def test_mixed_networks_batch_matches_separate_batches():
    """Columns of three networks (ni = 40 / 93 / 150) in one MixedBatch, each
    network on its own stream, give exactly the results of one ColumnBatch
    per network (same block-inversion variant)."""
    torch = _torch()
    from vulcan_amd.device import ColumnBatch, DevicePlan, MixedBatch
    cases = W.mixed_cases(6)
    for c in cases:
        c.cfg.conv_step = 20
    mb = MixedBatch(cases, history=21, graph=True, tune=False)
    mb.integrate(12)
    st = mb.host_state()
    assert len(mb.groups) == 3
    for idx, b, _ in mb.groups:
        net = cases[int(idx[0])].net
        ref = ColumnBatch(DevicePlan(net), [cases[i].col for i in idx], [cases[i].y0 for i in idx], cases[int(idx[0])].cfg,
                          history=21, graph=False, tune=False)
        ref.integrate(12)
        torch.cuda.synchronize()
        assert np.array_equal(ref.y.cpu().numpy(), b.y.cpu().numpy()), net.ni
        rs = ref.host_state()
        assert np.array_equal(rs['count'], st['count'][idx]) and np.array_equal(rs['t'], st['t'][idx])
        ref.close()
    mb.close()
