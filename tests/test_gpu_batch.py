"""Batched ensembles (s2t_batch_*): every member is bit-identical to the same instance run as a
plain solver, eager or replayed from hipGraphs, whatever else shares its size class."""

import numpy as np
import pytest

from oracle.spectral import random_field

pytestmark = pytest.mark.gpu


def member_params(i):
    rng = np.random.default_rng(100 + i)
    return dict(beta=float(rng.uniform(0, 20)), drag=0.01, viscosity=float(10 ** rng.uniform(-1, 0.5)),
                hyper_p=4.0, force_amp=float(rng.uniform(0, 4)), force_k=float(rng.choice([8, 12])),
                force_width=2.0, cfl=0.5, dt_min=1e-5, dt_max=0.02, fixed_dt=0.0)


SIZES = [64, 128, 64, 256, 128, 64]


def run_plain(s2t, n, params, seed, steps, k0=None):
    with s2t.Solver(n) as s:
        s.configure(**params)
        s.seed_spectrum(k0=n / 10 if k0 is None else k0, amplitude=6.0, seed=seed)
        s.step(steps)
        return s.read_state(), s.diagnostics()


@pytest.mark.parametrize("grouped", [True, False])
@pytest.mark.parametrize("graphs", [True, False])
def test_batch_members_match_plain(s2t, graphs, grouped):
    steps = 12
    with s2t.Batch(SIZES) as b:
        b.use_graphs(graphs)
        b.use_grouped(grouped)
        for i, n in enumerate(SIZES):
            b.configure(i, **member_params(i))
            b.seed_spectrum(i, k0=n / 10, amplitude=6.0, seed=1000 + i)
        b.step(5)
        mid = b.read_state(2)  # eager work between replays
        b.step(steps - 5)
        got = [b.read_state(i) for i in range(len(SIZES))]
        diags = [b.diagnostics(i) for i in range(len(SIZES))]
    for i, n in enumerate(SIZES):
        want, d = run_plain(s2t, n, member_params(i), 1000 + i, steps)
        assert np.array_equal(got[i], want), "member %d differs" % i
        assert diags[i]["stepIndex"] == steps and diags[i]["time"] == d["time"]
    want_mid, _ = run_plain(s2t, SIZES[2], member_params(2), 1002, 5)
    assert np.array_equal(mid, want_mid)


GROUPED_SIZES = [256, 512, 256, 1024, 128, 512, 256, 2048]


@pytest.mark.parametrize("graphs", [True, False])
def test_grouped_members_match_plain(s2t, graphs):
    """Grouped launches (s2t_batch_use_grouped): the 256, 512 and 1024 classes step together in six
    launches per step, the 128 and 2048 classes on their own; every member is bit-identical to its
    plain solver, eager or replayed, also when the mode is switched between steps."""
    sizes = GROUPED_SIZES
    with s2t.Batch(sizes) as b:
        b.use_graphs(graphs)
        b.use_grouped(True)
        for i, n in enumerate(sizes):
            b.configure(i, **member_params(i))
            b.seed_spectrum(i, k0=n / 10, amplitude=6.0, seed=1000 + i)
        b.step(3)
        b.use_grouped(False)
        b.step(2)
        b.use_grouped(True)
        b.step(3)
        got = [b.read_state(i) for i in range(len(sizes))]
        diags = [b.diagnostics(i) for i in range(len(sizes))]
    for i, n in enumerate(sizes):
        want, d = run_plain(s2t, n, member_params(i), 1000 + i, 8)
        assert np.array_equal(got[i], want), "member %d (%d^2) differs" % (i, n)
        assert diags[i]["stepIndex"] == 8 and diags[i]["time"] == d["time"]


@pytest.mark.parametrize("n", [128, 256])
def test_batch_reseed_member_between_replays(s2t, n):
    """Replacing one member's state re-primes it; the others continue undisturbed (128: a class on
    its own launches; 256: a class of the grouped launches)."""
    sizes = [n, n, n]
    with s2t.Batch(sizes) as b:
        for i in range(3):
            b.configure(i, **member_params(i))
            b.seed_spectrum(i, k0=12, amplitude=6.0, seed=7 + i)
        b.step(4)
        field = random_field(n, k0=10, seed=3, amplitude=4.0)
        b.seed_physical(1, field)
        b.step(6)
        got0 = b.read_state(0)
        got1 = b.read_state(1)
    want0, _ = run_plain(s2t, n, member_params(0), 7, 10, k0=12)
    assert np.array_equal(got0, want0)
    with s2t.Solver(n) as s:
        s.configure(**member_params(1))
        s.seed_physical(field)
        s.step(6)
        assert np.array_equal(got1, s.read_state())


def test_batch_rejects_bad_sizes(s2t):
    with pytest.raises(s2t.S2TError, match="must be 4\\^a or 2 \\* 4\\^a"):
        s2t.Batch([64, 96])


@pytest.mark.parametrize("members, width", [(2, 2), (4, 4)])
def test_2048_classes_of_several_members_match_plain(s2t, members, width):
    """2048^2 classes step by per-class launches at the width measured fastest for their member
    count (two columns per workgroup for two or three members, four with P_hat kept in LDS from
    four): every member, several instances in each launch, is bit-identical to its plain solver
    (which runs the single-instance field-parallel kernel)."""
    steps = 3
    sizes = [2048] * members
    with s2t.Batch(sizes) as b:
        assert b.column_widths() == {2048: width}
        for i in range(members):
            b.configure(i, **member_params(i))
            b.seed_spectrum(i, k0=30.0, amplitude=6.0, seed=2000 + i)
        b.step(steps)
        got = [b.read_state(i) for i in range(members)]
    for i in range(members):
        want, _ = run_plain(s2t, 2048, member_params(i), 2000 + i, steps, k0=30.0)
        assert np.array_equal(got[i], want), "member %d differs" % i
