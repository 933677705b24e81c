"""CPU tests: the SRRF oracle against closed forms (tests/srrf_closed_forms.py).

These pin the radiality magnitude (the squared (1 - d/R) term, its sign, the
divisor, renormalisation and the positivity constraint; kernel.cu:186-253)
and the intensity-weighted temporal modes including ITC orders 3 and 4
(kernel.cu:291-555) independently of the oracle's own formulas.
tests/test_srrf_oracle_mutants.py shows that each check fails on a planted
slip in the oracle.
"""
import pytest

import srrf_closed_forms as cf


@pytest.mark.parametrize("kind", ["ring2", "d6", "d26"])
@pytest.mark.parametrize("renorm,poscon", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_radiality_closed_form(kind, renorm, poscon):
    e64, e32 = cf.check_radiality(kind, renorm, poscon)
    print("%s renorm=%d poscon=%d: FP64 err %.2g, FP32 err %.2g" % (kind, renorm, poscon, e64, e32))


@pytest.mark.parametrize("method", [0, 1, 2])
def test_temporal_weighted_closed_form(method):
    cf.check_temporal_weighted(method)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_trac_weighted_closed_form(order):
    cf.check_temporal_weighted(3, order)


@pytest.mark.parametrize("order", [3, 4])
def test_itc_hand_derived_levels(order):
    cf.check_itc(order)
