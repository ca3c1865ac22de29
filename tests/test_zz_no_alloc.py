"""The allocation-free Hamiltonian (mag_hamiltonian_noalloc) against
mag_hamiltonian_async: the same kernels in the same order, so the three
results (kinetic, potential, H) agree bit for bit. Two calls in a row, the
second on a smaller lattice, check that nothing of the first call's scratch
content leaks into the second. Device memory for the results comes straight
from the HIP runtime."""
import ctypes

import numpy as np
import pytest

import mag_amd as M
from mag_amd import _lib

pytestmark = pytest.mark.gpu


def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def case(ctx, orc, dims, prec, arg, seed):
    U = M.Field(ctx, dims, M.LINK18 if prec == M.F64 else M.LINK12, prec)
    U.upload(orc.haar_links(dims, seed))
    P = M.Field(ctx, dims, M.ALG8, prec)
    M.init_momentum(P, seed + 1)
    return U, P


def read(h, ptr, n):
    out = np.zeros(n)
    assert h.hipMemcpy(out.ctypes.data, ptr, n * 8, 2) == 0  # device to host
    return out


def test_noalloc_matches_async_on_two_calls(ctx, orc):
    h = hip()
    ptr = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(ptr), 64) == 0
    try:
        base = ptr.value
        iwa = M.Arg_chmc(beta=2.6, gauge=M.Gauge(M.IWASAKI, -0.331))
        wil = M.Arg_chmc(beta=5.9)
        # a larger FP64 Iwasaki lattice, then a smaller FP32 Wilson one and a
        # smaller FP64 one: each call against the synchronous and _async paths
        for dims, prec, arg, seed in [((8, 8, 8, 16), M.F64, iwa, 3), ((4, 4, 4, 8), M.F32, wil, 5),
                                      ((4, 4, 6, 4), M.F64, iwa, 7)]:
            print(f"noalloc case dims={dims} prec={prec}", flush=True)
            U, P = case(ctx, orc, dims, prec, arg, seed)
            M.get_hamiltonian_async(U, P, arg, base)
            M.get_hamiltonian_noalloc(U, P, arg, base + 32)
            ctx.synchronize()
            v = read(h, base, 7)
            assert np.array_equal(v[0:3], v[4:7]), (dims, v)
            H, parts = M.get_hamiltonian(U, P, arg)
            assert v[4] == parts[0]
            assert abs(v[5] - parts[1]) <= 1e-14 * abs(parts[1]) and abs(v[6] - H) <= 1e-14 * abs(H)
            # against the oracle on the same links and momenta (FP64 storage only)
            if prec == M.F64:
                act = orc.action(arg.gauge.type, arg.beta, arg.gauge.c1)
                h_ref, _ = orc.hamiltonian(dims, U.download(), P.download(), act)
                assert abs(v[6] - h_ref) <= 1e-13 * abs(h_ref)
    finally:
        h.hipFree(ptr)


def test_noalloc_rejects_smeared_actions(ctx, orc):
    h = hip()
    ptr = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(ptr), 64) == 0
    try:
        U, P = case(ctx, orc, (4, 4, 4, 4), M.F64, None, 9)
        arg = M.Arg_chmc(beta=2.6, gauge=M.Gauge(M.IWASAKI, -0.331), n_stout=1)
        with pytest.raises(M.MagError) as e:
            M.get_hamiltonian_noalloc(U, P, arg, ptr.value)
        assert e.value.code == _lib.ERR_UNSUPPORTED
    finally:
        h.hipFree(ptr)
