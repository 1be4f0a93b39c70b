"""The 16-bit storage forms of Pij / Eij.

Pins, for both forms:
  * the oracle's rounding (double arithmetic, nearbyint) against numpy's
    float16 conversion and torch's bfloat16 conversion on every tie point
    between neighbouring 16-bit values, one float ulp on either side of it,
    the overflow and underflow edges and random values over the full range;
  * the product's host conversion (integer bit manipulation shared with the
    kernels) against the same library conversions, bit for bit;
  * the scale exponent of the binary16 form against its defining bound;
  * the oracle's storage mode: stored values are the rounding of the float
    update, and Wij is computed from the value before rounding.
"""

import ctypes

import numpy as np
import pytest
import torch

import oracle as O
from gsbn_amd import configs

BF16, F16 = 1, 2


def _f16_values():
    """All finite binary16 values, as float32."""
    h = np.arange(0, 0x7C00, dtype=np.uint16)
    v = h.view(np.float16).astype(np.float32)
    return np.concatenate([v, -v])


def _bf16_values():
    h = np.arange(0, 0x7F80, dtype=np.uint32)
    v = (h << 16).view(np.float32)
    return np.concatenate([v, -v])


def _edges(vals):
    """Midpoints between neighbours (exact in float32) and one ulp around them."""
    v = np.unique(vals[vals >= 0].astype(np.float64))
    mid = ((v[:-1] + v[1:]) / 2).astype(np.float32)
    assert np.all(mid.astype(np.float64) == (v[:-1] + v[1:]) / 2)
    up = np.nextafter(mid, np.float32(np.inf))
    dn = np.nextafter(mid, np.float32(0))
    out = np.concatenate([mid, up, dn, v.astype(np.float32)])
    return np.concatenate([out, -out])


def _random_floats(n, seed):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 0x7F800000, size=n, dtype=np.uint32)
    x = bits.view(np.float32)
    sign = rng.integers(0, 2, size=n).astype(bool)
    x[sign] = -x[sign]
    return x


def _np_f16(x, shift=0):
    with np.errstate(over="ignore"):
        y = (x.astype(np.float64) * 2.0 ** shift).astype(np.float32)
        return (y.astype(np.float16).astype(np.float64) / 2.0 ** shift).astype(np.float32)


def _torch_bf16(x):
    return torch.from_numpy(x.copy()).to(torch.bfloat16).to(torch.float32).numpy()


def _same(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_oracle_f16_rounding_against_numpy():
    x = np.concatenate([_edges(_f16_values()), _random_floats(200000, 1),
                        np.array([65504, 65519.99, 65520, 7e4, 1e30, 2.0 ** -25, 2.0 ** -24,
                                  1.5 * 2 ** -25, 2.0 ** -26, 0.0, -0.0], np.float32)])
    got = O.round_trace(F16, x, 0)
    want = _np_f16(x)
    assert _same(got, want)


@pytest.mark.parametrize("shift", [-3, 4, 9, 13])
def test_oracle_f16_scaled_rounding(shift):
    # binary16 of x * 2^shift, decoded and divided by 2^shift again
    x = np.concatenate([_edges(_f16_values()) * np.float32(2.0 ** -shift),
                        _random_floats(100000, 2)]).astype(np.float32)
    got = O.round_trace(F16, x, shift)
    want = _np_f16(x, shift)
    assert _same(got, want)


def test_oracle_bf16_rounding_against_torch():
    x = np.concatenate([_edges(_bf16_values()), _random_floats(200000, 3),
                        np.array([3.3895e38, 3.4e38, np.finfo(np.float32).max, 1e-45, 0.0,
                                  -0.0], np.float32)])
    got = O.round_trace(BF16, x, 0)
    want = _torch_bf16(x)
    assert _same(got, want)


def _product_lib():
    import gsbn_amd
    L = gsbn_amd.lib()
    L.gsbn_host_pack_trace.restype = ctypes.c_uint16
    L.gsbn_host_pack_trace.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float]
    L.gsbn_host_unpack_trace.restype = ctypes.c_float
    L.gsbn_host_unpack_trace.argtypes = [ctypes.c_int, ctypes.c_uint16, ctypes.c_float]
    return L


def test_product_host_conversion_bits():
    L = _product_lib()
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.choice(_edges(_f16_values()), 20000),
                        rng.choice(_edges(_bf16_values()), 20000),
                        _random_floats(5000, 8),
                        np.array([65504, 65520, 1e30, 2.0 ** -25, 0.0, -0.0], np.float32)])
    x = x.astype(np.float32)
    h16 = np.array([L.gsbn_host_pack_trace(F16, float(v), 1.0) for v in x], np.uint16)
    hbf = np.array([L.gsbn_host_pack_trace(BF16, float(v), 1.0) for v in x], np.uint16)
    with np.errstate(over="ignore"):
        assert np.array_equal(h16, x.astype(np.float16).view(np.uint16))
    tb = torch.from_numpy(x.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(hbf, tb)
    # decode: exact, and the round trip equals the oracle's rounding
    d16 = np.array([L.gsbn_host_unpack_trace(F16, int(h), 1.0) for h in h16], np.float32)
    dbf = np.array([L.gsbn_host_unpack_trace(BF16, int(h), 1.0) for h in hbf], np.float32)
    assert _same(d16, O.round_trace(F16, x, 0))
    assert _same(dbf, O.round_trace(BF16, x, 0))
    # scaled binary16: pack(x * 2^k) / 2^k
    k = 9
    xs = x[np.abs(x) < 1e3]
    ds = np.array([L.gsbn_host_unpack_trace(F16, L.gsbn_host_pack_trace(F16, float(v), 2.0 ** k),
                                            2.0 ** -k) for v in xs], np.float32)
    assert _same(ds, O.round_trace(F16, xs, k))


def test_fp16_shift_bound():
    spec = configs.config1()
    q = spec.projs[0]
    k = O.fp16_shift(q, spec.dt)
    # kfti = kftj = 1, kz = 0.1 per step: zmax = 1 / (1 - e^-0.1) = 10.508...
    zmax = 1.0 / (1.0 - np.exp(-0.1))
    assert k == int(np.floor(np.log2(65504.0 / zmax ** 2))) == 9
    assert zmax ** 2 * 2.0 ** k <= 65504 < zmax ** 2 * 2.0 ** (k + 1)
    # a z trace kicked every step reaches the bound and the product of two
    # such traces, scaled, stays finite in binary16
    z = 0.0
    for _ in range(1000):
        z = z * np.float32(np.exp(-0.1)) + 1.0
    assert abs(z - zmax) < 1e-4
    assert np.isfinite(np.float16(np.float32(z * z * 2.0 ** k)))


def _scripted_spec(form, H=3, M=4, steps=120, seed=11):
    rng = np.random.default_rng(seed)
    ext = {s: sorted(rng.choice(H * M, size=rng.integers(0, 4), replace=False).tolist())
           for s in range(1, steps)}
    stim = np.full((1, H), configs.UNCLAMPED, np.uint32)
    mask = np.zeros((1, H), np.float32)
    modes = [configs.ModeSpec(configs.step_time(0, 0.001), configs.step_time(steps, 0.001),
                              time_step=steps, begin_lgidx_id=0, begin_wmask_id=0, prn=1.0,
                              lgidx_step=0, wmask_step=0)]
    pop = configs.PopSpec(hcu=H, mcu=M, snoise=0.0)
    proj = configs.ProjSpec(src=0, dest=0, tauzi=0.01, tauzj=0.01, taue=0.02, taup=1.0,
                            slot_num=float((H - 1) * M))
    spec = configs.NetSpec(dt=0.001, pops=[pop], projs=[proj], modes=modes, stim=stim, mask=mask,
                           conn=[configs.dense_conn(H, M, 1)], name="scripted",
                           trace_precision=form)
    spec.ext_spikes = {0: ext}
    return spec


@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_oracle_storage_mode(form):
    prec = O.TRACE_FORMS[form]
    steps = 120
    plain = O.OracleNet(_scripted_spec("fp32", steps=steps))
    low = O.OracleNet(_scripted_spec(form, steps=steps))
    shift = low.trace_shifts[0] if prec == F16 else 0
    # C*M = 32: the initial Pij = 1/32 is exact in both forms
    assert _same(low.vector("pij_0"), plain.vector("pij_0"))
    first_update = None
    for s in range(steps):
        p0 = low.vector("pij_0").copy()
        plain.step()
        low.step()
        pij = low.vector("pij_0")
        eij = low.vector("eij_0")
        # every stored value is a value of the 16-bit form
        assert _same(O.round_trace(prec, pij, shift), pij)
        assert _same(O.round_trace(prec, eij, shift), eij)
        if first_update is None and not _same(pij, p0):
            first_update = s
            # the first update of every synapse starts from identical state in
            # both runs: the stored value is the rounding of the float result
            # and the weight is computed from the unrounded value
            ch = ~(pij.view(np.uint32) == p0.view(np.uint32))
            assert _same(pij[ch], O.round_trace(prec, plain.vector("pij_0")[ch], shift))
            assert _same(low.vector("wij_0")[ch], plain.vector("wij_0")[ch])
    assert first_update is not None
    # the two runs see the same spikes (external spikes, no noise), so they
    # differ only by storage rounding, which accumulates over the ~30 updates
    # each synapse gets here (measured: bf16 max 1.5e-2 / mean 4.7e-3,
    # fp16 max 1.8e-3 / mean 4.9e-4 relative for Pij)
    for name in ("pij_0", "eij_0"):
        a = low.vector(name).astype(np.float64)
        b = plain.vector(name).astype(np.float64)
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
        mx, mean = (2.0 ** -5, 2.0 ** -7) if prec == BF16 else (2.0 ** -8, 2.0 ** -10)
        assert rel.max() <= mx and rel.mean() <= mean
