"""Error of the storage forms against the exact double-precision replay, on a
small network (CPU, oracle engine; the product is bit-identical to this
engine by the GPU parity tests).  Bounds are from measurement with margin:
20x20, 3000 steps gave Pij max/mean relative error 5.7e-7/6.7e-8 (fp32),
1.9e-2/2.5e-3 (bf16) and 3.0e-3/3.4e-4 (fp16)."""

import pytest

import precision_study as PS

BOUNDS = {  # form: (Pij max, Pij mean, Wij abs max)
    "fp32": (1e-5, 1e-6, 1e-4),
    "bf16": (6e-2, 6e-3, 6e-2),
    "fp16": (1e-2, 1e-3, 1e-2),
}


@pytest.mark.parametrize("form", ["fp32", "bf16", "fp16"])
@pytest.mark.parametrize("point", ["default", "close"])
def test_storage_form_error_against_exact(form, point):
    r = PS.run(form, point, 1500, "oracle", hcu=12, mcu=12, log=lambda m: None)
    mx, mean, wmax = BOUNDS[form]
    assert r["pij"]["max"] <= mx and r["pij"]["mean"] <= mean, r["pij"]
    assert r["wij"]["max"] <= wmax, r["wij"]
    # Pi, bj are float in every form
    assert r["pi"]["max"] <= 1e-5 and r["bj"]["max"] <= 1e-5
