"""The reference's own test configurations on the GPU against the oracle.

test/params_01.in ... params_07.in of the reference (kept as fixtures under
tests/golden/params/) are read by the params reader exactly as the reference
reads them, given a seeded initial state at each configuration's own scales,
and stepped 20 times with the reference's adaptive time step (none of them
fixes dt).  The device and the oracle must agree on q, psi, t and dt to 1e-12,
or to ten times the oracle's own sensitivity where the configuration amplifies
round-off: the oracle is run a second time from the initial state perturbed
by 1e-15 (relative), and that run's distance from the first is the
sensitivity.  Six of the seven configurations keep it below 1e-15; params_06
from this state (a flow that grows 250-fold in the 20 steps) raises a 1e-15
perturbation to ~1e-9, and the device's distance is of that size.  The set
covers what the files exercise:

- params_01: one layer, bounded, Ekman-less, unregistered keys ignored;
- params_02: stochastic ring forcing on a bounded grid;
- params_03: three layers with the double-gyre wind;
- params_04: no-slip walls (bc_fac = 1) with a topographic shelf;
- params_05, params_06: periodic domains with stochastic forcing (params_06
  with the order-6 hyperviscosity);
- params_07: three layers on 512^2.
"""
import os
import sys

import numpy as np
import pytest

import qgw_amd as Q
import oracle as O
from conftest import rel

pytestmark = pytest.mark.gpu

PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "params")
NAMES = ["params_0%d.in" % i for i in range(1, 8)]


def _announce(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", NAMES)
def test_reference_config_fp64_vs_oracle(name):
    p = Q.read_params(os.path.join(PARAMS, name))
    # a seeded k^-3 state: rms velocity a small fraction of the rotation
    # velocity scale f0 Lx (or of Lx per unit time without rotation), eddies
    # at an eighth of the domain
    vel = (p.f0 if p.f0 > 0 else 1.0) * p.Lx
    p.set(ic_kind=1, seed=20261203, ic_urms=2e-4 * vel, ic_lambda0=p.Lx / 8)
    _announce("reference config", name, p.NX, p.NY, p.nl, "bc_fac", p.bc_fac, "k_f", p.k_f)
    modes = Q.eigmode(p)
    q0 = Q.make_ic(p, modes)[1]

    def oracle_run(q_init):
        o = O.Oracle(p)
        o.set_modes(*modes)
        o.set_q(q_init)
        if p.k_f != 0:
            o.set_stochastic(p.k_f, p.sigma_f)
        o.run(20)
        t = o.get_time()
        q = o.field(O.Q).copy()
        o.invert_pv()
        psi = o.field(O.PSI).copy()
        o.close()
        return t, q, psi

    to, q_orc, psi_orc = oracle_run(q0)
    pert = np.random.default_rng(1).standard_normal(q0.shape)
    tp, q_p, psi_p = oracle_run(q0 * (1.0 + 1e-15 * pert))
    m = Q.Model(p, Q.FP64, modes)
    m.set_q(q0)
    for _ in range(20):
        m.timestep()
    tm = m.get_time()
    q_dev = m.q
    m.invert_pv()
    psi_dev = m.psi
    m.close()
    assert tm["it"] == to["it"] == 20

    def rdiff(a, b):
        return abs(a - b) / abs(b)

    errs = dict(t=rdiff(tm["t"], to["t"]), dt=rdiff(tm["dt"], to["dt"]), q=rel(q_dev, q_orc),
                psi=rel(psi_dev, psi_orc))
    sens = dict(t=rdiff(tp["t"], to["t"]), dt=rdiff(tp["dt"], to["dt"]), q=rel(q_p, q_orc),
                psi=rel(psi_p, psi_orc))
    print("[errors]", name, " ".join("%s %.2e (sens %.1e)" % (k, errs[k], sens[k]) for k in errs), flush=True)
    for k in errs:
        assert errs[k] < max(1e-12, 10.0 * sens[k]), (k, errs[k], sens[k])
