"""llg_reduce_mag_noalloc (the allocation-free <M>) against llg_reduce through the C ABI: the
same bits, on two calls in a row with the second on a smaller system, and again on the first
system after the array held the second's partial sums.  The FP32 Heun handle is stepped first,
so only the AoS3 copy of its spins is current and the call must rebuild the SoA spins."""
import ctypes as C

import numpy as np
import pytest

import sllg
from sllg import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(g):
    m = np.zeros(g.nsub * 3)
    rc = sllg.lib().llg_reduce_mag_noalloc(g.h, m.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, sllg.lib().llg_last_error()
    return m.reshape(g.nsub, 3)


def test_noalloc_reduce_equals_reduce(gpu):
    big, s_big = W.fcc_3shell(L=16, precision="fp32")  # 16,384 spins, line reader, AoS3 stepping
    small, s_small = W.bct_afm_dmi(Lx=8, Ly=8, Lz=4, precision="fp32", scheme="heun")  # 512 spins, 2 sublattices
    plain, s_plain, _ = W.sc_ferro(L=10, precision="fp64")
    ga, gb, gc = sllg.Context(big), sllg.Context(small), sllg.Context(plain)
    for g, s0, T in ((ga, s_big, 300.0), (gb, s_small, 0.0), (gc, s_plain, 100.0)):
        g.set_spins(s0)
        g.set_temperature(T)
        g.step(0, 30)
    ma = _noalloc(ga)  # first call on the larger system
    assert np.array_equal(ma, ga.reduce(energy=False)[0])
    mb = _noalloc(gb)  # second call, smaller system, two sublattices
    assert mb.shape == (2, 3) and np.array_equal(mb, gb.reduce(energy=False)[0])
    assert np.array_equal(_noalloc(gc), gc.reduce(energy=False)[0])
    ga.step(30, 10)
    assert np.array_equal(_noalloc(ga), ga.reduce(energy=False)[0])
    assert np.array_equal(gb.reduce_mag_noalloc(), mb)  # unchanged spins, same bits
    # the AoS3 stepping and the sync before the sum leave the trajectory unchanged
    ref = sllg.Context(big)
    ref.set_spins(s_big)
    ref.set_temperature(300.0)
    ref.step(0, 40)
    assert np.array_equal(ref.get_spins(), ga.get_spins())
    for g in (ga, gb, gc, ref):
        g.close()
