"""What the trained sketch network gains over plain random-sketch
RSVD-Delta-t on held-out members of the Ginzburg-Landau family (reduced
sizes: 12 training members, 3 held-out): the sigma_1 error after the same
number of power iterations, against a converged reference (plain, q = 4)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rsvd-delta-t_amd"))
import learned  # noqa: E402

pytestmark = pytest.mark.gpu


def test_learned_sketch_gain(pkg):
    rng = np.random.default_rng(77)
    train = learned.family_sample(rng, 12)
    held = learned.family_sample(rng, 3)
    print(f"learned sketch: {len(train)} training members, {len(held)} held-out", flush=True)
    data = learned.training_data(pkg, train)
    net = learned.SketchNet(seed=3)
    hist = net.train(data, epochs=20)
    print(f"alignment loss {hist[0]:.3f} -> {hist[-1]:.3f}", flush=True)
    assert hist[-1] < 0.6 * hist[0]
    rows, summary = learned.evaluate(pkg, net, held, qs=(0, 1))
    for r in rows:
        print(f"mu0={r['mu0']:.3f} Re U={r['Ur']:.3f}: q=0 plain {r['plain'][0][0]:.1e} learned "
              f"{r['learned'][0][0]:.1e}; q=1 plain {r['plain'][1][0]:.1e} learned {r['learned'][1][0]:.1e}", flush=True)
    for q in (0, 1):
        assert summary["learned"]["median_sigma1_err"][q] < 0.1 * summary["plain"]["median_sigma1_err"][q]
    # the learned columns only replace random ones: every run still ends in valid gains
    for r in rows:
        assert all(np.isfinite(r["learned"][q][0]) for q in (0, 1))
