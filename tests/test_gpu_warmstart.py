"""The warm start of the coupling loop on a microstructure outside the
network's training set (128^3, sphere radii 6-16, seed 1003; training used
seeds 0-39): the extrapolation in time (linear, with Galerkin steps along the
second and third differences of phi_P's changes) lowers the phP CG iterations
of the fast solver well below the plain start (the last step's fields, as the
reference), and every solve meets the same tolerance.

The learned correction (the packaged network, trained on 1920 solver pairs of
40 seeded 128^3 packs) is applied as an A-norm-optimal Galerkin step, so it
never makes the start worse in the norm CG minimises (energy gain >= 0 at
every step).  It does not lower the iteration count below the extrapolated
start's: measured on this case, steps 3-8, plain 10944, extrapolated 5105,
learned 5935 phP iterations.  The error the extrapolation leaves is mostly
the previous solves' tolerance-level error, which the network's inputs do
not predict, and a start with lower energy error can still need more CG
iterations to the relative-residual rule.  The test records the numbers and
bounds the cost (learned within 25 % of extrapolated)."""
import numpy as np
import pytest

import elsim_amd as E
from elsim_amd import nn as NN
from elsim_amd._lib import F_CNE, F_PHE, F_PHP

pytestmark = pytest.mark.gpu

N = 128
SEED = 1003
STEPS = 8


def run(ws_mode, net=None):
    rng = np.random.default_rng(10_000 + SEED)
    target = float(rng.uniform(0.40, 0.60))
    sdf, _, _ = E.sphere_pack(N, SEED, 6.0, 16.0, target=target, hardcore=0.85, connect=True)
    s = E.Sbm(sdf, -sdf, E.synthetic_params(ws_mode=ws_mode, cne_rule=1), E.tables())
    s.set(E._lib.F_CNP, E.initial_cnp(N, SEED))
    if net is not None:
        s.set_warm_start(net)
    stats = [s.step(E.MODE_MIXED) for _ in range(STEPS)]
    fields = {f: s.get(f) for f in (F_PHP, F_PHE, F_CNE)}
    s.close()
    return stats, fields


def test_warm_start_saves_iterations(gpu):
    net = NN.WarmStartNet.load(NN.DEFAULT_WEIGHTS)
    plain, fp = run(E.WS_PLAIN)
    extrap, fe = run(E.WS_EXTRAP)
    learned, fl = run(E.WS_LEARNED, net)
    it = {k: sum(st.it_php for st in v[2:]) for k, v in (("plain", plain), ("extrap", extrap), ("learned", learned))}
    print("phP CG iterations, steps 3..%d:" % STEPS, it)
    for st in learned[2:]:
        assert st.ws_used == 2 and st.ws_energy_gain >= 0.0
    for v in (plain, extrap, learned):
        for st in v:
            assert st.status == 0 and st.relres_php <= 1e-8 and st.relres_phe <= 1e-8
    assert it["extrap"] < 0.6 * it["plain"]
    assert it["learned"] < 0.6 * it["plain"]
    assert it["learned"] <= 1.25 * it["extrap"]
    # every run meets relres 1e-8 but from different starts; the coupled
    # fields agree to the level the reference's coupling rule allows
    for f in (F_PHP, F_PHE, F_CNE):
        for other in (fe, fl):
            rel = np.linalg.norm(other[f] - fp[f]) / np.linalg.norm(fp[f])
            assert rel < 2e-5, (f, rel)
    net.close()
