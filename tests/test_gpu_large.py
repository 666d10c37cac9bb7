"""Large permanents on the GPU against committed extended-precision oracle
values (tests/golden/extended_vectors.json, written by
tests/golden/make_extended_vectors.py with the double-double oracle).

The bar here is the engine's: relative error <= 1e-10 against the
extended-precision oracle, at the measured sizes (n = 12, 20, 24, 28, 32 and
the 40-mode / 20-photon boson-sampling amplitude)."""

import numpy as np
import pytest

import piquasso_amd as pa
from conftest import load_json, rel_err

pytestmark = pytest.mark.gpu

CASES = {c["id"]: c for c in load_json("extended_vectors.json")["cases"]}


def _inputs(case):
    m = case["matrix"]
    a = (np.asarray(m["re"]) + 1j * np.asarray(m["im"])).reshape(m["shape"])
    dd = case["dd"]
    return a, np.asarray(case["rows"]), np.asarray(case["cols"]), complex(dd[0] + dd[1], dd[2] + dd[3])


@pytest.mark.parametrize("cid", sorted(CASES))
def test_extended_oracle_values(gpu, cid):
    a, r, c, ref = _inputs(CASES[cid])
    print(f"case {cid}: shape {a.shape}", flush=True)
    got = complex(pa.permanent(a, r, c))
    err = rel_err(got, ref)
    print(f"case {cid}: rel.err {err:.3e}", flush=True)
    assert err <= 1e-10, (cid, got, ref, err)


def test_n20_fp32_products_error_is_stated(gpu):
    a, r, c, ref = _inputs(CASES["large_single_n20"])
    v32, _ = pa.permanent_ex(a, r, c, mode="fp32")
    err = rel_err(v32, ref)
    print(f"n=20 FP32-product variant: rel.err {err:.3e}", flush=True)
    # FP32 row sums and products: ~1e-7 per operation, amplified by the
    # cancellation of the alternating sum; the accumulation itself is dd
    assert err < 1e-2


def test_n32_plain_vs_compensated(gpu):
    a, r, c, ref = _inputs(CASES["large_single_n32"])
    v_dd, _ = pa.permanent_ex(a, r, c, mode="dd")
    v_plain, _ = pa.permanent_ex(a, r, c, mode="plain")
    e_dd, e_plain = rel_err(v_dd, ref), rel_err(v_plain, ref)
    print(f"n=32: dd rel.err {e_dd:.3e}, plain FP64 rel.err {e_plain:.3e}", flush=True)
    assert e_dd <= 1e-10
