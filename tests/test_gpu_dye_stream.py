"""Ordering of the stepped dye. In a step with dye the BFECC launch runs on a stream of its own,
beside stages 1 and 2 of its step (it reads the stage-0 velocity and the step's dt); the solver's
stream waits for it at the end of the step. So the next step's stage 0 (which rewrites the
velocity), and every write or read of the dye, come after it: the results below are the same
whether or not the host touches the dye between steps, and the same from run to run."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARAMS = dict(beta=0.0, drag=0.0, viscosity=6.0, hyper_p=4.0, force_amp=0.0, cfl=0.5,
              dt_min=1e-5, dt_max=0.02, dye_fade=1.0)


def texture(m, seed):
    """Random colour in [0, 1], alpha 1, as binary16."""
    rng = np.random.default_rng(seed)
    t = np.ones((m, m, 4), np.float32)
    t[..., :3] = rng.uniform(0.0, 1.0, (m, m, 3))
    return t.astype(np.float16)


def run(s2t, n, scale, steps, touch):
    with s2t.Solver(n, dye_scale=scale) as s:
        s.configure(**PARAMS)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=2)
        s.write_dye(texture(s.dye_size(), 7))
        seen = []
        for _ in range(steps):
            s.step(1, dye=True)
            if touch:
                seen.append(s.read_dye().view(np.uint16).copy())
        return s.read_dye().view(np.uint16), s.read_state(), seen


@pytest.mark.parametrize("n, scale", [(512, 2), (256, 1)])
def test_dye_reads_between_steps_change_nothing(s2t, n, scale):
    dye_a, w_a, seen = run(s2t, n, scale, 6, touch=True)
    dye_b, w_b, _ = run(s2t, n, scale, 6, touch=False)
    assert np.array_equal(dye_a, dye_b) and np.array_equal(w_a, w_b)
    assert np.array_equal(seen[-1], dye_a)
    # every step moved the dye
    assert all(not np.array_equal(seen[i], seen[i + 1]) for i in range(len(seen) - 1))


@pytest.mark.parametrize("n, scale", [(1024, 2), (256, 1)])
def test_stepped_dye_is_deterministic(s2t, n, scale):
    """A dye launch still reading the velocity while the next step's stage 0 rewrote it, or a
    step starting before the dye launch read its dt, would make runs differ."""
    a = run(s2t, n, scale, 8, touch=False)
    b = run(s2t, n, scale, 8, touch=False)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_dye_written_after_a_stepped_dye_is_kept(s2t):
    """A texture written (or painted) right after a step with dye is what the next read returns:
    the step's dye launch, which writes the texture that becomes current, has finished first."""
    with s2t.Solver(512, dye_scale=2) as s:
        s.configure(**PARAMS)
        s.seed_bandlimited(8.0, 16.0, 6.0, seed=3)
        m = s.dye_size()
        s.write_dye(texture(m, 1))
        for k in range(4):
            s.step(1, dye=True)
            t = texture(m, 10 + k)
            s.write_dye(t)
            assert np.array_equal(s.read_dye().view(np.uint16), t.view(np.uint16)), k
        s.step(1, dye=True)
        s.paint_dye("bands")
        painted = s.read_dye()
    with s2t.Solver(512, dye_scale=2) as ref:
        ref.paint_dye("bands")
        assert np.array_equal(painted.view(np.uint16), ref.read_dye().view(np.uint16))
