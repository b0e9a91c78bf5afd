"""Oracle Philox / Box-Muller / Wigner pinned by known answers and closed forms."""
import math

import numpy as np
import pytest

from oracle import gpe_oracle as O


@pytest.mark.parametrize("key,counter,expected", [
    (0, (0, 0, 0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    (0xFFFFFFFFFFFFFFFF, (0xFFFFFFFF,) * 4, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    (0x299F31D0A4093822, (0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_random123_known_answers(key, counter, expected):
    """Random123 KATs, as in tests/unit/test_noise.py:56-64 of the reference."""
    assert tuple(int(v) for v in O.philox4x32(np.asarray(counter, dtype=np.uint32), key)) == expected


def test_counter_layout_and_streams_differ():
    c = O.counter_words(np.arange(3), (5 << 32) + 9, 2)
    assert c.tolist() == [[0, 9, 5, 2], [1, 9, 5, 2], [2, 9, 5, 2]]
    cells = np.arange(1024)
    a = O.philox4x32(O.counter_words(cells, 0), 42)
    assert not np.array_equal(a, O.philox4x32(O.counter_words(cells, 1), 42))
    assert not np.array_equal(a, O.philox4x32(O.counter_words(cells, 0, 2), 42))


def test_uniforms_closed_form_and_range():
    w = np.asarray([[0, 0, 0, 0], [0xFFFFFFFF] * 4, [0x80000000, 0x40, 0x1F, 0x3F]], dtype=np.uint32)
    u1, u2 = O.uniforms(w, "single")
    assert u1[0] == np.float32(2.0 ** -24) and u1[1] == np.float32(1.0)
    d1, d2 = O.uniforms(w, "double")
    assert d1[0] == 2.0 ** -53 and d1[1] == 1.0 and d2[1] == 1.0
    # (w0 >> 5) * 2^26 + (w1 >> 6) + 1, times 2^-53
    assert d1[2] == ((0x80000000 >> 5) * 2 ** 26 + (0x40 >> 6) + 1) * 2.0 ** -53


def test_normal_pair_matches_math_module():
    words = O.philox4x32(O.counter_words(np.arange(64), 3), 99)
    u1, u2 = O.uniforms(words, "double")
    re, im = O.normal_pair(words, "double")
    for i in range(64):
        r = math.sqrt(-2.0 * math.log(u1[i]))
        assert abs(re[i] - r * math.cos(2 * math.pi * u2[i])) < 1e-14
        assert abs(im[i] - r * math.sin(2 * math.pi * u2[i])) < 1e-14


def test_normals_are_standard():
    words = O.philox4x32(O.counter_words(np.arange(200_000), 11), 7)
    for prec in ("single", "double"):
        re, im = O.normal_pair(words, prec)
        for x in (re, im):
            assert abs(np.mean(x)) < 0.01 and abs(np.var(x) - 1) < 0.01


@pytest.mark.parametrize("strength", [1.0, 2.0])
def test_one_step_wigner_variance(strength):
    """<|dpsi|^2> = strength gamma_C dt / (2 dx dy) (reference test_noise.py:111-118)."""
    B, ny, nx = 1, 64, 64
    psi = np.zeros((B, 1, ny, nx), dtype=np.complex128)
    dt, dx, dy, gamma = 0.01, 1.0, 1.0, 0.083
    O.wigner_add(psi, 0, [123456], [gamma], [0.0], [strength], dt, dx * dy, 0, "double")
    expected = strength * gamma * dt / 2.0
    assert abs(np.mean(np.abs(psi) ** 2) - expected) / expected < 0.06
