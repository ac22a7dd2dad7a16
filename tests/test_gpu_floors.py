"""Ten-step parity of the HIP step with the CPU oracle on the parity grids,
both methods, all fields including the dynamic pressure.

Bar: each field within 1.5 times its rounding floor. The floor
(tests/golden/floors.json, made by tests/golden/make_floors.py on the CPU) is
how far the oracle moves from itself after 10 steps when one-ulp noise is put
on the initial field and on every step's Lamb vector and solves, maximum over
four noise seeds: two correct implementations cannot agree more closely. Where
the floor is below 1e-10 the bar is below the required 1e-10 as well.

Grids:
  re180_64         Re_tau 180, 64x65x64
  re180_128        Re_tau 180, 128x129x128 (box 4pi x 2 x 4pi/3)
  re550_wall_32    the Re_tau 550 wall-normal grid (NY 193, stretch 2.0,
                   nu 1e-4, dkx 1, dkz 2) on a 32x32 horizontal grid
  case180_wall_16  test-cases/Retau-180/case.toml's nu, dt, dkx, dkz, NY 160
                   (even) and stretch 2.0 on a 16x16 horizontal grid
  case550_wall_16  test-cases/Retau-550/case.toml's nu, dt, dkx, dkz, NY 288
                   and stretch 2.4 on a 16x16 horizontal grid
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FLOORS = json.load(open(os.path.join(HERE, "golden", "floors.json")))
NAMES = ("u", "v", "w", "omx", "omy", "omz", "p")
SLACK = 1.5


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def oracle_runs(oracle_mod):
    cache = {}

    def get(name):
        if name not in cache:
            c = FLOORS[name]["case"]
            o = oracle_mod.Oracle(c["nu"], c["dt"], c["dkx"], c["dkz"], c["alpha_mesh"], c["NX"], c["NY"], c["NZ"])
            o.init(2, seed=0, amp=c["amp"])
            o.step(FLOORS[name]["steps"])
            cache[name] = ({n: o.field(n) for n in NAMES}, o.diag())
        return cache[name]
    return get


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("name", ["re180_64", "case180_wall_16", "case550_wall_16", "re550_wall_32", "re180_128"])
def test_ten_steps_within_rounding_floor(owt_mod, oracle_runs, name, method):
    c, floor, steps = FLOORS[name]["case"], FLOORS[name]["floor"], FLOORS[name]["steps"]
    ref, dref = oracle_runs(name)
    p = {k: c[k] for k in ("nu", "dt", "dkx", "dkz", "alpha_mesh", "NX", "NY", "NZ")}
    print(f"[gpu] {name} {c['NX']}x{c['NY']}x{c['NZ']} method {method}", flush=True)
    with owt_mod.Channel(params=dict(p, precision=0, method=method)) as ch:
        ch.init_fields(2, seed=0, amp=c["amp"])
        ch.step(steps)
        err = {n: rel(ch.field(n), ref[n]) for n in NAMES}
        d = ch.diag()
    print(f"[parity] {name} method {method}: " + " ".join(f"{n}={err[n]:.2e}/{err[n] / floor[n]:.2f}" for n in NAMES),
          flush=True)
    for n in NAMES:
        assert err[n] <= SLACK * floor[n], (n, err[n], floor[n])
    assert d.shape == dref.shape
    np.testing.assert_allclose(d, dref, rtol=1e-10, atol=0)
