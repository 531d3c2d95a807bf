"""The equilibrated direct solve is the same Newton step as the
reference's direct solve (to rounding): on well-conditioned systems the
iteration counts and converged states coincide with the reference
fixture; on the deep-reverse-bias diode it converges in no more
iterations than the reference."""
import numpy as np

from golden.fixtures import load
from oracle import physics as oph
from current_bar import check
from oracle_build import oracle_sweep


def test_equilibrated_matches_reference_when_well_conditioned():
    spec, d = load("diode_60x40_sweep.npz")
    _, its, cur, _ = oracle_sweep(spec, opts=oph.NewtonOptions(linsolve="direct_equilibrated"))
    assert np.array_equal(its, d["newton_iterations"])
    check(cur, d["currents"], "equilibrated", say=lambda m: None)


def test_equilibrated_never_slower_at_reverse_bias():
    spec, d = load("diode_24x16_defaults_sweep.npz")
    _, its, cur, _ = oracle_sweep(spec, opts=oph.NewtonOptions(linsolve="direct_equilibrated"))
    assert np.all(its <= d["newton_iterations"])
    assert np.array_equal(its, [9, 9, 16, 27, 5, 5, 5, 5, 5])
