"""The mixed-precision step's v solves with the inverses in mirror form
(k_mirror_inverse + k_solve_mirror) against the same step with the full
explicit inverses (OWT_MIRROR_SOLVE=0 at channel creation), and against the
FP64 oracle.

After one step from the same initial field the two mixed channels differ only
in the v solve (everything before it is the same arithmetic on the same data),
so the first new v level compares the two solves directly. The full inverse X
is refined against the LU factors, so x = X b carries the LU substitution's
rounding, which the v operator's conditioning (3e10 at the Re_tau 550 grid)
amplifies for right-hand sides that reach its ill-conditioned directions, as
the FP32 rounding of the Lamb vector does. The mirror form starts from X and
drops the part of its rounding that is odd under the reflection. A CPU model of
both forms at NY 193 (X the inverse of the LU factors, right-hand sides of a
smooth v plus 1e-7 noise) puts each within 2e-9 of a long-double-refined
solution, as dgetrs itself (4e-9), and the two 9e-10 apart. Bar: 1.5 times
the grid's one-step FP64 rounding floor (tests/golden/floors.json,
"floor_1step": how far the oracle moves after one step under one-ulp noise on
its field, Lamb vector and solves), the convention of the parity tests. Printed
beside it: what mixed precision itself changes at the same step (the
full-inverse mixed channel against an FP64 channel).
Over 10 steps the FP32 transform phase turns any difference into FP32-level
rounding flips, so the 10-step check is against the FP64 oracle with the mixed
path's bars: u 1e-6, v and w 2e-5 (the mixed path's distance from FP64 at 64^3
after 20 steps is about 1e-6).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FLOORS = json.load(open(os.path.join(HERE, "golden", "floors.json")))
KEYS = ("nu", "dt", "dkx", "dkz", "alpha_mesh", "NX", "NY", "NZ")
# the floor grids: NY 65 (one row block of the half range), NY 160 (even, no
# middle row), NY 193 (the Re_tau 550 wall-normal grid), NY 256 (a full second
# row block)
GRIDS = {name: ({k: FLOORS[name]["case"][k] for k in KEYS}, FLOORS[name]["case"]["amp"], FLOORS[name]["floor_1step"])
         for name in ("re180_64", "case180_wall_16", "re550_wall_32", "re550_ny256_16")}
SLACK = 1.5


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def mixed_channel(owt, p, mirror):
    old = os.environ.get("OWT_MIRROR_SOLVE")
    os.environ["OWT_MIRROR_SOLVE"] = "1" if mirror else "0"
    try:
        return owt.Channel(params=dict(p, precision=1, method=1))
    finally:
        if old is None:
            del os.environ["OWT_MIRROR_SOLVE"]
        else:
            os.environ["OWT_MIRROR_SOLVE"] = old


@pytest.mark.parametrize("name", list(GRIDS))
def test_mirror_v_solve_equals_full_inverse(owt_mod, name):
    p, amp, floor = GRIDS[name]
    print(f"[gpu] mirror v solve {name} {p['NX']}x{p['NY']}x{p['NZ']}", flush=True)
    a, b = mixed_channel(owt_mod, p, True), mixed_channel(owt_mod, p, False)
    c = owt_mod.Channel(params=dict(p, precision=0, method=1))
    try:
        for ch in (a, b, c):
            ch.init_fields(2, seed=0, amp=amp)
            ch.step(1)
        errs = {n: rel(a.field(n), b.field(n)) for n in ("v", "u", "w")}
        prec = {n: rel(b.field(n), c.field(n)) for n in ("v", "u", "w")}
    finally:
        a.close()
        b.close()
        c.close()
    print(f"[mirror] {name} one step, mirror vs full: " + " ".join(f"{n}={e:.2e}" for n, e in errs.items())
          + "; mixed vs FP64: " + " ".join(f"{n}={e:.2e}" for n, e in prec.items()), flush=True)
    for n, e in errs.items():
        assert e <= SLACK * floor[n], (name, n, e, floor[n])


@pytest.mark.parametrize("name", ["re180_64", "re550_wall_32"])
def test_mirror_mixed_step_against_the_oracle(owt_mod, oracle_mod, name):
    p, amp, _ = GRIDS[name]
    o = oracle_mod.Oracle(p["nu"], p["dt"], p["dkx"], p["dkz"], p["alpha_mesh"], p["NX"], p["NY"], p["NZ"])
    o.init(2, seed=0, amp=amp)
    o.step(10)
    print(f"[gpu] mirror mixed 10 steps {name}", flush=True)
    a = mixed_channel(owt_mod, p, True)
    try:
        a.init_fields(2, seed=0, amp=amp)
        a.step(10)
        errs = {n: rel(a.field(n), o.field(n)) for n in ("u", "v", "w")}
        da, do = a.diag(), o.diag()
    finally:
        a.close()
    print(f"[mirror] {name} 10 steps vs oracle: " + " ".join(f"{n}={e:.2e}" for n, e in errs.items()), flush=True)
    assert errs["u"] <= 1e-6 and errs["v"] <= 2e-5 and errs["w"] <= 2e-5, errs
    assert np.abs(da[:, 0] - do[:, 0]).max() <= 1e-6 * np.abs(do[:, 0]).max()


def test_precision_switch_converts_the_stored_wall_normal_terms(owt_mod):
    """The stored wall-normal term of each history level is DY2 v in FP64 and
    DY2 v - k^2 v in mixed precision (the v right-hand side's combination);
    owt_set_precision converts all three levels. A mixed channel switched to
    FP64 and back steps on like one never switched, within the one-step
    rounding floor (the round trip rounds the stored terms once); an FP64 channel
    switched to mixed steps like an FP64 channel up to the FP32 Lamb vector of
    that one step (a missing or doubled k^2 v term would move v by O(1))."""
    p, amp, floor = GRIDS["re180_64"]
    print("[gpu] precision switch re180_64", flush=True)
    a, b = mixed_channel(owt_mod, p, True), mixed_channel(owt_mod, p, True)
    try:
        for ch in (a, b):
            ch.init_fields(2, seed=0, amp=amp)
            ch.step(3)
        a.set_precision(owt_mod.FP64)
        a.set_precision(owt_mod.MIXED)
        for ch in (a, b):
            ch.step(1)
        errs = {n: rel(a.field(n), b.field(n)) for n in ("v", "u", "w")}
    finally:
        a.close()
        b.close()
    print("[mirror] mixed -> FP64 -> mixed, one step: " + " ".join(f"{n}={e:.2e}" for n, e in errs.items()), flush=True)
    for n, e in errs.items():
        assert e <= SLACK * floor[n], (n, e, floor[n])
    c = owt_mod.Channel(params=dict(p, precision=0, method=1))
    d = owt_mod.Channel(params=dict(p, precision=0, method=1))
    try:
        for ch in (c, d):
            ch.init_fields(2, seed=0, amp=amp)
            ch.step(3)
        c.set_precision(owt_mod.MIXED)
        for ch in (c, d):
            ch.step(1)
        errs = {n: rel(c.field(n), d.field(n)) for n in ("v", "u", "w")}
    finally:
        c.close()
        d.close()
    print("[mirror] FP64 -> mixed, one step against FP64: " + " ".join(f"{n}={e:.2e}" for n, e in errs.items()), flush=True)
    for n, e in errs.items():
        assert e <= 1e-6, (n, e)


def test_mixed_checkpoint_restores_exactly(owt_mod, tmp_path):
    """A mixed-precision checkpoint (stored wall-normal terms in the mixed form)
    restored into a fresh mixed channel and into an FP64 channel (which the
    restore switches to mixed) continues bit for bit."""
    p, amp, _ = GRIDS["re180_64"]
    path = str(tmp_path / "mixed.ckpt")
    print("[gpu] mixed checkpoint re180_64", flush=True)
    a = mixed_channel(owt_mod, p, True)
    try:
        a.init_fields(2, seed=0, amp=amp)
        a.step(3)
        a.checkpoint(path)
        a.step(2)
        ref = {n: a.field(n) for n in ("v", "u", "w", "omy")}
    finally:
        a.close()
    for prec in (owt_mod.MIXED, owt_mod.FP64):
        b = owt_mod.Channel(params=dict(p, precision=prec, method=1))
        try:
            b.restore(path)
            b.step(2)
            for n, r in ref.items():
                assert np.array_equal(b.field(n), r), (prec, n)
        finally:
            b.close()


def test_mixed_step_above_ny256_against_the_oracle(owt_mod, oracle_mod):
    """Above NY 256 the mixed step runs the full-product kernels and the full
    inverses, with the stored wall-normal terms in the mixed form: the Retau-550
    case file's wall-normal grid (NY 288) on 16 x 16, 10 steps against the FP64
    oracle with the mixed path's bars."""
    c = FLOORS["case550_wall_16"]["case"]
    p = {k: c[k] for k in KEYS}
    o = oracle_mod.Oracle(p["nu"], p["dt"], p["dkx"], p["dkz"], p["alpha_mesh"], p["NX"], p["NY"], p["NZ"])
    o.init(2, seed=0, amp=c["amp"])
    o.step(10)
    print("[gpu] mixed 10 steps case550_wall_16 (NY 288)", flush=True)
    with owt_mod.Channel(params=dict(p, precision=1, method=1)) as ch:
        ch.init_fields(2, seed=0, amp=c["amp"])
        ch.step(10)
        errs = {n: rel(ch.field(n), o.field(n)) for n in ("u", "v", "w")}
        d = ch.diag()
    print("[mirror] NY 288 mixed, 10 steps vs oracle: " + " ".join(f"{n}={e:.2e}" for n, e in errs.items()), flush=True)
    assert errs["u"] <= 1e-6 and errs["v"] <= 2e-5 and errs["w"] <= 2e-5, errs
    do = o.diag()
    assert np.abs(d[:, 0] - do[:, 0]).max() <= 1e-6 * np.abs(do[:, 0]).max()
