"""The terminal-current bar of the parity tests and the bench.

Per bias point and per contact:

    |I - I_ref| <= 1e-6 * max_c |I_ref|  +  floor,
    floor = |sum_c I_ref| + |sum_c I|     (contacts with disjoint node lists)
    floor = 0                             (contacts that share nodes)

1e-6 of the largest current is the stated relative bar.  The floor is
absolute and per device: the contact currents of an exact solution of a
device whose contacts are disjoint sum to zero (the continuity rows are
conservative), so |sum_c I| is how far each side's FP64 state determines
its own currents (both the reference's and this side's).  A reverse- or
low-forward-bias current is a cancellation of fluxes many decades larger;
there the reference's direct solve leaves a sum of a few percent of the
current, and no FP64 state resolves that current further.  Devices whose
contacts share nodes (the tri-gate's body contact overlaps source and
drain) have no such identity and get no floor.

`check` prints, for every point, whether the strict bar (no floor) held
and returns how many points needed the floor.
"""
import numpy as np


def oracle_sweep(spec, opts=None):
    """The oracle's warm-started sweep (equilibrium, then every point):
    Newton counts and currents in contact order."""
    from oracle_build import oracle_device
    dev = oracle_device(spec)
    dev.solve_equilibrium(opts)
    its, cur = [], []
    for bias in spec.sweep:
        dev.solve_bias(bias, opts)
        its.append(dev.last_iterations)
        cur.append([dev.terminal_current(c.name) for c in spec.contacts])
    return np.array(its), np.array(cur)


def contacts_overlap(spec):
    seen = set()
    for c in spec.contacts:
        nodes = set(np.ravel_multi_index(c.idx, spec.shape).tolist())
        if seen & nodes:
            return True
        seen |= nodes
    return False


def floors(cur, ref, overlap):
    cur, ref = np.atleast_2d(cur), np.atleast_2d(ref)
    if overlap:
        return np.zeros(ref.shape[0])
    return np.abs(ref.sum(axis=1)) + np.abs(cur.sum(axis=1))


def check(cur, ref, label="", overlap=False, rel=1e-6, say=print):
    """Assert the bar on every point; return (points, strict, floor_only):
    how many points met 1e-6 of the largest current with no floor, and how
    many needed the floor."""
    cur, ref = np.atleast_2d(cur), np.atleast_2d(ref)
    fl = floors(cur, ref, overlap)
    strict = floor_only = 0
    for p in range(ref.shape[0]):
        big = np.abs(ref[p]).max()
        err = np.abs(cur[p] - ref[p]).max()
        ok_strict = err <= rel * big
        strict += int(ok_strict)
        floor_only += int(not ok_strict and err <= rel * big + fl[p])
        say(f"  {label} point {p}: |dI| {err:.3e} = {err / big:.2e} of max|I|, floor "
            f"{fl[p]:.2e} = {fl[p] / big:.2e} of max|I|: {'1e-6' if ok_strict else 'floor'}")
        assert err <= rel * big + fl[p], (label, p, cur[p], ref[p], fl[p])
    return ref.shape[0], strict, floor_only
