"""p3m_solve_noalloc through the C ABI.

The allocation-free solve carves every working buffer from the library's
fixed device scratch array; its results must equal p3m_solve's on the same
device arrays to the FP64 bar of the device path (forces rel-rms <= 1e-12,
energy and virial rel <= 1e-12: the two differ only in the order of the
floating-point atomic additions into the mesh; FP32 mesh 1e-5).  Two calls in
a row, the second with fewer particles (the array keeps the first call's
contents), both force index rules, and the size limit, which is rejected
before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from tests import p3m_systems as S

pytestmark = pytest.mark.gpu


def _dptr(t):
    return C.c_void_p(t.data_ptr())


def _context(L, _lib, box, pos_t, q_t, mesh, cao, alpha, r_cut, precision, epsilon=0.):
    ctx = C.c_void_p()
    _lib.raise_for(L.p3m_create(0, C.byref(ctx)))
    p = _lib.p3m_params()
    L.p3m_params_default(C.byref(p))
    for d in range(3):
        p.mesh[d] = mesh
        p.box_l[d] = box[d]
    p.cao, p.alpha, p.r_cut, p.accuracy = cao, alpha, r_cut, 1e-3
    p.prefactor, p.epsilon, p.tune, p.verbose = 1.7, epsilon, 0, 0
    p.precision = precision
    _lib.raise_for(L.p3m_set_params(ctx, C.byref(p)), ctx)
    n = q_t.shape[0]
    _lib.raise_for(L.p3m_activate(ctx, n, _dptr(pos_t), _dptr(q_t), _lib.P3M_DEVICE_POINTERS),
                   ctx)
    return ctx


def _solve(L, _lib, fn, ctx, pos_t, q_t, upos_t, n, flags):
    import torch
    f = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    e = C.c_double()
    v = (C.c_double * 9)()
    torch.cuda.synchronize()
    _lib.raise_for(fn(ctx, n, _dptr(pos_t), _dptr(q_t), _dptr(upos_t), _dptr(f),
                      flags | _lib.P3M_DEVICE_POINTERS, C.byref(e), v), ctx)
    return f.cpu().numpy(), e.value, np.array(v[:])


@pytest.mark.parametrize("precision,conv,epsilon", [(0, 0, 0.), (0, 1, 0.), (0, 0, 3.),
                                                     (1, 0, 0.)])
def test_noalloc_equals_solve_twice_second_smaller(precision, conv, epsilon):
    import torch
    from espresso_amd import _lib
    L = _lib.load()
    rng = np.random.default_rng(97)
    Lb = 17.
    n1, n2 = 6000, 2500
    pos = rng.uniform(0., Lb, (n1, 3))
    q = np.where(np.arange(n1) % 2 == 0, 1., -1.)
    q[::7] = 0.
    upos = pos + Lb * rng.integers(-1, 2, (n1, 3))
    pos_t = torch.tensor(pos, device="cuda")
    q_t = torch.tensor(q, device="cuda")
    upos_t = torch.tensor(upos, device="cuda")
    ctx = _context(L, _lib, [Lb] * 3, pos_t, q_t, 24, 5, 1.1, 3.2, precision, epsilon)
    f_tol = 1e-12 if precision == 0 else 1e-5
    e_tol = 1e-12 if precision == 0 else 1e-5
    flags = _lib.P3M_FORCES | _lib.P3M_ENERGY | _lib.P3M_PRESSURE
    try:
        _lib.raise_for(L.p3m_set_force_convention(ctx, conv), ctx)
        for n in (n1, n2):
            # the second call: a prefix of the same arrays, fewer particles
            print(f"[noalloc] precision={precision} conv={conv} eps={epsilon} n={n}", flush=True)
            f_ref, e_ref, v_ref = _solve(L, _lib, L.p3m_solve, ctx, pos_t, q_t, upos_t, n, flags)
            f, e, v = _solve(L, _lib, L.p3m_solve_noalloc, ctx, pos_t, q_t, upos_t, n, flags)
            assert np.all(np.isfinite(f))
            err_f = S.rel_rms(f, f_ref)
            print(f"[noalloc] force rel-rms {err_f:.3e}, energy {e:.12g} vs {e_ref:.12g}",
                  flush=True)
            assert err_f <= f_tol
            assert abs(e - e_ref) <= e_tol * abs(e_ref)
            assert np.max(np.abs(v - v_ref)) <= e_tol * np.max(np.abs(v_ref))
    finally:
        L.p3m_destroy(ctx)


def test_noalloc_two_contexts_alternate():
    """Two contexts on one device share the fixed array: alternating calls
    each give their own context's results."""
    import torch
    from espresso_amd import _lib
    L = _lib.load()
    flags = _lib.P3M_FORCES | _lib.P3M_ENERGY
    ctxs, data = [], []
    try:
        for k, (n, Lb, mesh) in enumerate(((3000, 14., 20), (1200, 9., 12))):
            rng = np.random.default_rng(300 + k)
            pos_t = torch.tensor(rng.uniform(0., Lb, (n, 3)), device="cuda")
            q_t = torch.tensor(np.where(np.arange(n) % 2 == 0, 1., -1.), device="cuda")
            ctx = _context(L, _lib, [Lb] * 3, pos_t, q_t, mesh, 4, 1.0, 3.0, 0)
            ctxs.append(ctx)
            ref = _solve(L, _lib, L.p3m_solve, ctx, pos_t, q_t, pos_t, n, flags)
            data.append((pos_t, q_t, n, ref))
        for rep in range(2):
            for ctx, (pos_t, q_t, n, (f_ref, e_ref, _)) in zip(ctxs, data):
                f, e, _ = _solve(L, _lib, L.p3m_solve_noalloc, ctx, pos_t, q_t, pos_t, n, flags)
                assert S.rel_rms(f, f_ref) <= 1e-12
                assert abs(e - e_ref) <= 1e-12 * abs(e_ref)
    finally:
        for ctx in ctxs:
            L.p3m_destroy(ctx)


def test_noalloc_rejects_oversize_and_missing_setup():
    import torch
    from espresso_amd import _lib
    L = _lib.load()
    cap = C.c_int64()
    assert L.p3m_noalloc_scratch_bytes(C.byref(cap)) == _lib.P3M_OK
    assert cap.value == 48 << 20
    n_big = cap.value // 44 + 1000  # more particles than the array holds
    Lb = 40.
    rng = np.random.default_rng(11)
    pos_t = torch.tensor(rng.uniform(0., Lb, (n_big, 3)), device="cuda")
    q_t = torch.tensor(np.where(np.arange(n_big) % 2 == 0, 1., -1.), device="cuda")
    f = torch.zeros((n_big, 3), dtype=torch.float64, device="cuda")
    # a context that was never activated: rejected
    ctx0 = C.c_void_p()
    _lib.raise_for(L.p3m_create(0, C.byref(ctx0)))
    try:
        st = L.p3m_solve_noalloc(ctx0, 10, _dptr(pos_t), _dptr(q_t), None, _dptr(f),
                                 _lib.P3M_FORCES | _lib.P3M_DEVICE_POINTERS, None, None)
        assert st != _lib.P3M_OK
    finally:
        L.p3m_destroy(ctx0)
    ctx = _context(L, _lib, [Lb] * 3, pos_t[:1000], q_t[:1000], 32, 3, 0.8, 3.0, 0)
    try:
        torch.cuda.synchronize()
        st = L.p3m_solve_noalloc(ctx, n_big, _dptr(pos_t), _dptr(q_t), None, _dptr(f),
                                 _lib.P3M_FORCES | _lib.P3M_DEVICE_POINTERS, None, None)
        assert st == _lib.P3M_ERR_INVALID
        assert b"scratch" in L.p3m_last_error(ctx)
        # host pointers are rejected too
        h = np.zeros((10, 3))
        st = L.p3m_solve_noalloc(ctx, 10, h.ctypes.data_as(C.c_void_p),
                                 h[:, 0].copy().ctypes.data_as(C.c_void_p), None,
                                 h.ctypes.data_as(C.c_void_p), _lib.P3M_FORCES, None, None)
        assert st == _lib.P3M_ERR_INVALID
        # the context still solves within the limit afterwards
        n = 1000
        flags = _lib.P3M_FORCES | _lib.P3M_ENERGY
        f_ref, e_ref, _ = _solve(L, _lib, L.p3m_solve, ctx, pos_t, q_t, pos_t, n, flags)
        f1, e1, _ = _solve(L, _lib, L.p3m_solve_noalloc, ctx, pos_t, q_t, pos_t, n, flags)
        assert S.rel_rms(f1, f_ref) <= 1e-12 and abs(e1 - e_ref) <= 1e-12 * abs(e_ref)
    finally:
        L.p3m_destroy(ctx)
