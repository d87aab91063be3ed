"""PCG of common.wgsl:29-56, pinned by invertibility and the LCG closed form. CPU only."""
import ctypes as C

import numpy as np
import pytest

from . import oracle_ffi as orc

M = 2**32
A, INC, MULT_OUT = 747796405, 1, 277803737


def _inv(a):
    return pow(a, -1, M)


def _unxorshift(x, shift):
    s = x
    for _ in range(32 // shift + 1):
        s = x ^ (s >> shift)
    return s


def _undo_rxs_m(x):
    """Invert ((s >> ((s >> 28) + 4)) ^ s) * 277803737: the top 4 bits of s survive the shift."""
    x = (x * _inv(MULT_OUT)) % M
    return _unxorshift(x, (x >> 28) + 4)


def _state_after(seed_state, n):
    """LCG jump-ahead by the geometric series: a^n s + c (a^n - 1)/(a - 1) mod 2^32."""
    an = pow(A, n, M)
    geo = sum(pow(A, i, M) for i in range(n)) % M
    return (an * seed_state + INC * geo) % M


@pytest.mark.parametrize("seed", [0, 1, 12345, 2**31 + 17, 2**32 - 1])
def test_rng_init_inverts_to_seed(seed):
    s0 = orc.lib().or_rng_init(seed)
    pre = _undo_rxs_m(s0)
    assert ((pre - 2891336453) * _inv(A)) % M == seed


@pytest.mark.parametrize("seed", [7, 99991, 2**32 - 5])
def test_rng_next_is_the_lcg_with_rxs_m_xs_output(seed):
    L = orc.lib()
    s0 = L.or_rng_init(seed)
    st = C.c_uint32(s0)
    outs, states = [], [s0]
    for _ in range(64):
        outs.append(L.or_rng_next(C.byref(st)))
        states.append(st.value)
    for k in (1, 5, 64):
        assert states[k] == _state_after(s0, k)
    for old, o in zip(states[:-1], outs):
        w = o ^ (o >> 22)  # (w >> 22) ^ w with shift 22 >= 16 inverts in one step
        assert _undo_rxs_m(w) == old


def test_rng_f32_range_and_rounding_to_one():
    L = orc.lib()
    # state whose next output is 0xFFFFFFFF: f32(u) rounds up to 2^32, so rng_f32 == 1.0
    for target, expect in ((0xFFFFFFFF, 1.0), (0, 0.0), (2**31, 0.5)):
        w = target ^ (target >> 22)
        st = C.c_uint32(_undo_rxs_m(w))
        assert L.or_rng_f32(C.byref(st)) == expect
    st = C.c_uint32(L.or_rng_init(3))
    v = np.array([L.or_rng_f32(C.byref(st)) for _ in range(20000)])
    assert v.min() >= 0.0 and v.max() <= 1.0 and abs(v.mean() - 0.5) < 0.01
