"""Philox4x32-10 known-answer vectors (Random123 kat_vectors, Salmon et al.
SC'11) pin the oracle's generator; the product's host copy must agree."""

import numpy as np

import oracle as O

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_kat():
    for ctr, key, want in KAT:
        assert list(O.philox(ctr, key)) == want


def test_draw_distribution():
    # uniforms in [0,1), normals with unit variance: moments over 20000 draws
    us, zs = [], []
    for unit in range(20000):
        u, z = O.draw(7, 0, 3, unit)
        us.append(u)
        zs.append(z)
    us, zs = np.array(us), np.array(zs)
    assert us.min() >= 0 and us.max() < 1
    assert abs(us.mean() - 0.5) < 0.01 and abs(zs.mean()) < 0.03 and abs(zs.std() - 1) < 0.03
    # uniform is the top 24 bits of word 0
    w = O.philox([5, 3, 0, 0], [7, 0])
    u, _ = O.draw(7, 0, 3, 5)
    assert u == (w[0] >> 8) * 2.0 ** -24


def test_product_host_draw_matches_oracle():
    import ctypes
    from gsbn_amd import lib
    L = lib()
    for seed, pop, step, unit in [(0, 0, 0, 0), (1 << 40, 2, 99, 123), (123456789, 0, 5000, 9999)]:
        u = ctypes.c_float()
        z = ctypes.c_float()
        L.gsbn_host_draw(seed, pop, step, unit, ctypes.byref(u), ctypes.byref(z))
        ou, oz = O.draw(seed, pop, step, unit)
        assert np.float32(ou) == np.float32(u.value) and np.float32(oz) == np.float32(z.value)
