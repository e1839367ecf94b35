"""The oracle reproduces the reference build bit for bit on the committed
fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from
oracle/_ref)."""
import numpy as np
import pytest

from helpers import GOLDEN_CASES, load_golden
from oracle.bind import Oracle


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_bitwise_on_fixture(name, oracle_lib):
    cfg, par, a, b, dts = load_golden(name)
    o = Oracle(cfg, par, a)
    for n, dt in enumerate(dts):
        assert o.getmindt() == dt, f"CFL dt differs before step {n}"
        o.timestep(dt)
    s = o.get_state()
    m = b.real_mask()
    np.testing.assert_array_equal(s.prim, b.prim)
    np.testing.assert_array_equal(s.cons[:, m], b.cons[:, m])
    np.testing.assert_array_equal(s.piph, b.piph)
    np.testing.assert_array_equal(s.dphi, b.dphi)
    np.testing.assert_array_equal(s.wiph, b.wiph)
