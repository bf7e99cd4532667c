"""Random-number generators: the reference's RANMAR (src/omc_random.c) and
the per-particle Philox4x32-10 streams the device uses."""
import ctypes as C

import numpy as np


def _ctx(oracle, mode):
    L = oracle.lib()
    return L, L.ora_create(None, None, None, 1, mode)


def test_ranmar_marsaglia_known_answer(oracle):
    # Marsaglia & Zaman's published check: seeds (1802, 9373), after 20000
    # numbers the next six times 2^24 are these integers.
    L, ctx = _ctx(oracle, oracle.RANMAR)
    L.ora_seed_ranmar(ctx, 1802, 9373)
    for _ in range(20000):
        L.ora_random(ctx)
    got = [int(L.ora_random(ctx) * 16777216) for _ in range(6)]
    L.ora_destroy(ctx)
    assert got == [6533892, 14220222, 7275067, 6172232, 8354498, 10633180]


def test_ranmar_reference_seeds_97_33(oracle):
    # first draws of the reference build with the example seeds
    L, ctx = _ctx(oracle, oracle.RANMAR)
    L.ora_seed_ranmar(ctx, 97, 33)
    got = [int(L.ora_random(ctx) * 16777216) for _ in range(6)]
    L.ora_destroy(ctx)
    assert got == [16137026, 11603567, 5109275, 13162860, 12190880, 7217496]


def test_ranmar_invalid_seed_falls_back_to_default(oracle):
    L, ctx = _ctx(oracle, oracle.RANMAR)
    L.ora_seed_ranmar(ctx, 0, 99999)
    a = [L.ora_random(ctx) for _ in range(5)]
    L.ora_seed_ranmar(ctx, 1802, 9373)
    b = [L.ora_random(ctx) for _ in range(5)]
    L.ora_destroy(ctx)
    assert a == b


def _philox(oracle, c, k):
    L = oracle.lib()
    ci = (C.c_uint32 * 4)(*c)
    ki = (C.c_uint32 * 2)(*k)
    o = (C.c_uint32 * 4)()
    L.ora_philox4x32(ci, ki, o)
    return list(o)


def test_philox4x32_10_known_answers(oracle):
    # Random123 (Salmon et al., SC'11) known-answer vectors for philox4x32 R=10
    assert _philox(oracle, [0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _philox(oracle, [0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _philox(oracle, [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_uniforms_are_24_bit_and_uniform(oracle):
    # draws of one stream: k * 2^-24 in [0,1), flat to statistical precision
    L = oracle.lib()
    vals = []
    for blk in range(4000):
        vals += [w >> 8 for w in _philox(oracle, [blk, 0, 0, 0], [123, 456])]
    u = np.array(vals, dtype=np.float64) / 16777216.0
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(np.floor(u * 16777216) == u * 16777216)
    hist, _ = np.histogram(u, bins=16, range=(0, 1))
    chi2 = ((hist - u.size / 16) ** 2 / (u.size / 16)).sum()
    assert chi2 < 40.0          # 15 dof, p ~ 5e-4
    assert abs(u.mean() - 0.5) < 0.005
