"""The 32-channel Re_tau 180 ensemble at its real shapes: 16 channels at
64x65x64, 8 at 64x97x64 and 8 at 96x49x96 (box 4pi x 2 x 2pi, stretch 1.5,
nu 1/2800, dt 1e-3), member m from seed m, FP64, batched method.

- Every member after 10 grouped steps equals the same channel stepped alone,
  bit for bit (fields, friction series): the grouped launches change only which
  block runs which tile, and members with equal parameters read one copy of
  operators that each would compute identically.
- The first member of each shape (seeds 0, 16, 24) against the CPU oracle after
  10 steps, every field within 1.5 times its rounding floor
  (tests/golden/floors.json: re180_64, ens180_64x97, ens180_96x49).
- Launches per step depend on the distinct shapes only: the 32-member ensemble
  and a 3-member ensemble of one channel per shape launch the same number of
  kernels, and a 16-member single-shape ensemble launches one per stage.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FLOORS = json.load(open(os.path.join(HERE, "golden", "floors.json")))
STEPS = 10
AMP = 0.15
NAMES = ("u", "v", "w", "omx", "omy", "omz", "p")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def ensemble(owt_mod):
    ps = owt_mod.ENSEMBLES["re180_32"]
    print(f"[gpu] 32-member ensemble, {STEPS} grouped steps", flush=True)
    ens = owt_mod.Ensemble(ps, method=owt_mod.BATCHED, precision=owt_mod.FP64)
    ens.init_fields(2, seeds=range(len(ps)), amp=AMP)
    ens.step(STEPS)
    yield ps, ens
    ens.close()


def test_members_equal_standalone_channels(owt_mod, ensemble):
    ps, ens = ensemble
    assert [m.steps_done() for m in ens.members] == [STEPS] * len(ps)
    for m, p in enumerate(ps):
        with owt_mod.Channel(params=dict(p, precision=0, method=1)) as ch:
            ch.init_fields(2, seed=m, amp=AMP)
            ch.step(STEPS)
            for name in ("u", "v", "w", "omx", "omy", "omz", "Hv", "Hg"):
                assert np.array_equal(ens[m].field(name), ch.field(name)), (m, name)
            assert np.array_equal(ens[m].diag(), ch.diag()), m


@pytest.mark.parametrize("member,case", [(0, "re180_64"), (16, "ens180_64x97"), (24, "ens180_96x49")])
def test_member_of_each_shape_against_oracle(oracle_mod, ensemble, member, case):
    ps, ens = ensemble
    c, floor = FLOORS[case]["case"], FLOORS[case]["floor"]
    p = ps[member]
    assert (p["NX"], p["NY"], p["NZ"]) == (c["NX"], c["NY"], c["NZ"]) and c.get("seed", 0) == member
    o = oracle_mod.Oracle(c["nu"], c["dt"], c["dkx"], c["dkz"], c["alpha_mesh"], c["NX"], c["NY"], c["NZ"])
    o.init(2, seed=member, amp=c["amp"])
    o.step(STEPS)
    err = {n: rel(ens[member].field(n), o.field(n)) for n in NAMES}
    print(f"[parity] ensemble member {member} ({case}): " +
          " ".join(f"{n}={err[n]:.2e}/{err[n] / floor[n]:.2f}" for n in NAMES), flush=True)
    for n in NAMES:
        assert err[n] <= 1.5 * floor[n], (n, err[n], floor[n])
    np.testing.assert_allclose(ens[member].diag(), o.diag(), rtol=1e-10, atol=0)


def test_launches_do_not_grow_with_members(owt_mod, ensemble):
    ps, ens = ensemble
    n32 = ens.launches_per_step()
    with owt_mod.Ensemble([ps[0], ps[16], ps[24]], method=owt_mod.BATCHED, precision=owt_mod.FP64) as e3:
        e3.init_fields(2, seeds=[0, 16, 24], amp=AMP)
        n3 = e3.launches_per_step()
    with owt_mod.Ensemble(ps[:16], method=owt_mod.BATCHED, precision=owt_mod.FP64) as e16:
        e16.init_fields(2, amp=AMP)
        n16 = e16.launches_per_step()
    print(f"[ensemble] launches per step: 32 members {n32}, 3 members {n3}, 16 equal members {n16}", flush=True)
    assert n32 == n3 and 8 <= n32 <= 24
    assert n16 == 8


def test_uneven_ensemble_of_other_kernel_instances(owt_mod):
    """Members that run the grouped kernels' other instances, each equal to the
    same channel stepped alone, bit for bit: mixed precision (FP32 transforms),
    NY = 288 (three row tiles in the matrix-core kernels, the panel LU above 256
    rows), a 480-point transform (radix 5), and an FP64 member of the same shape
    as the mixed one (two lanes of one shape). Also after one member was stepped
    alone (members stepped one by one until the phases agree)."""
    base = dict(owt_mod.ENSEMBLES["re180_32"][0])
    ps = [dict(base, NX=16, NY=33, NZ=16, precision=1), dict(base, NX=16, NY=33, NZ=16, precision=0),
          dict(base, NX=8, NY=288, NZ=8, precision=0), dict(base, NX=320, NY=17, NZ=16, precision=1)]
    print("[gpu] uneven ensemble: mixed, NY 288, 480-point x transform", flush=True)
    with owt_mod.Ensemble(ps, method=owt_mod.BATCHED) as ens:
        ens.init_fields(2, seeds=[3, 4, 5, 6], amp=0.15)
        ens.step(3)
        assert ens.launches_per_step() > 0
        ens[2].step(1)
        ens.step(2)
        for m, (p, seed) in enumerate(zip(ps, (3, 4, 5, 6))):
            with owt_mod.Channel(params=dict(p, method=1)) as ch:
                ch.init_fields(2, seed=seed, amp=0.15)
                ch.step(6 if m == 2 else 5)
                for name in ("u", "v", "w", "omy", "Hv", "Hg", "lmbx"):
                    assert np.array_equal(ens[m].field(name), ch.field(name)), (m, name)
                assert np.array_equal(ens[m].diag(), ch.diag()), m
