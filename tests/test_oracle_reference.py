"""The reference's own test suite (test/runtests.jl) run against the oracle on the CPU.

test/1d_test.jl, 2d_test.jl, 3d_test.jl: hard-source propagation against the analytic
pulse w(z, t_end - z/C0) with the reference's tolerances (1e-2, 5e-2, 1e-1).
test/*_test_losses.jl: Beer-Lambert decay of the accumulated Poynting integral Sa in a
conductive and in a Lorentz half-space (rtol 1e-2), in 1D, 2D and 3D.
"""
import pytest

import cases
from oracle_util import rel_l2, run_oracle


def test_ref_1d_propagation():
    model, Eth, rtol = cases.ref_1d_test()
    h = run_oracle(model)
    assert rel_l2(h.field("Ex"), Eth) <= rtol


def test_ref_2d_propagation():
    model, Eth, rtol = cases.ref_2d_test()
    h = run_oracle(model)
    assert rel_l2(h.field("Ex"), Eth.ravel(order="F")) <= rtol


def test_ref_3d_propagation():
    model, Eth, rtol = cases.ref_3d_test()
    h = run_oracle(model)
    assert rel_l2(h.field("Ex"), Eth.ravel(order="F")) <= rtol


@pytest.mark.parametrize("ndims", [1, 2, 3])
@pytest.mark.parametrize("kind", ["sigma", "lorentz"])
def test_ref_losses(ndims, kind):
    model, ki = cases.ref_losses(ndims, kind)
    h = run_oracle(model)
    Sa = h.output(0).reshape(model.grid.shape, order="F")
    assert cases.beer_lambert_ok(Sa, model.grid.z, ki)


def test_losses_fail_with_wrong_conductivity():
    """The Beer-Lambert check has teeth: a 20% error in the conductivity fails it."""
    model, ki = cases.ref_losses(1, "sigma")
    h = run_oracle(model)
    Sa = h.output(0).reshape(model.grid.shape, order="F")
    assert not cases.beer_lambert_ok(Sa, model.grid.z, ki * 1.2)
