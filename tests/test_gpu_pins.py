"""GPU twins of the oracle's pins (tests/pins.py, tests/test_oracle.py): the
library's step, through the C ABI, against the same identities, recursions and
eigenproblems, in both methods."""
import numpy as np
import pytest

import pins

pytestmark = pytest.mark.gpu

RE180 = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5)


def chan(owt_mod, method, **kw):
    p = dict(RE180, precision=0, method=method)
    p.update(kw)
    print(f"[gpu] pins {p['NX']}x{p['NY']}x{p['NZ']} method {method}", flush=True)
    return owt_mod.Channel(params=p)


@pytest.mark.parametrize("method", [0, 1])
def test_recovery_identities_on_gpu(owt_mod, method):
    with chan(owt_mod, method, NX=16, NY=33, NZ=16) as ch:
        ch.init_fields(2, seed=3, amp=0.3)
        ch.step(3)
        kx, kz = pins.wavenumbers(16, 16, RE180["dkx"], RE180["dkz"])
        res = pins.recovery_identities(ch, kx, kz, ch.matrix("DY1"), ch.matrix("DY2"))
        print(f"[pin] gpu recovery identities method {method}: {res}", flush=True)
        assert res.max() < 1e-12, res


@pytest.mark.parametrize("method", [0, 1])
def test_squire_mode_on_gpu(owt_mod, method):
    N, dt = 65, 0.01
    with chan(owt_mod, method, NX=4, NY=N, NZ=8, dt=dt) as ch:
        dev = pins.squire_sequence(ch, ch.wave_shape, ch.matrix("DY2"), RE180["dkz"], RE180["nu"], dt, 30)
        print(f"[pin] gpu Squire BDF3 recursion method {method}: {dev:.2e}", flush=True)
        assert dev < 1e-12


@pytest.mark.parametrize("method", [0, 1])
def test_mean_flow_recursion_on_gpu(owt_mod, method):
    with chan(owt_mod, method, NX=4, NY=33, NZ=8, nu=0.02, dt=0.01, dkz=2.0) as ch:
        y = pins.grid_y(33, 1.5)
        dev, fdev = pins.mean_flow_recursion(ch, ch.wave_shape, y, ch.matrix("DY1"), ch.matrix("DY2"), 2.0, 0.02, 0.01, 12)
        print(f"[pin] gpu mean-flow recursion method {method}: {dev:.2e} friction {fdev:.2e}", flush=True)
        assert dev < 1e-12 and fdev < 1e-12


@pytest.mark.parametrize("method", [0, 1])
def test_flow_rate_on_gpu(owt_mod, method):
    with chan(owt_mod, method, NX=16, NY=33, NZ=16) as ch:
        ch.init_fields(2, seed=3, amp=0.3)
        drift = pins.flow_rate_drift(ch, pins.grid_y(33, 1.5), 40)
        print(f"[pin] gpu flow-rate drift method {method}: {drift:.2e}", flush=True)
        assert drift < 1e-7


@pytest.mark.parametrize("method", [0, 1])
def test_oblique_mode_on_gpu(owt_mod, method):
    nu, dt = 1.5e-4, 0.01
    with chan(owt_mod, method, NX=4, NY=65, NZ=4, nu=nu, dt=dt, dkx=1.0, dkz=0.5) as ch:
        lm, lam, ratio = pins.oblique_mode(ch, ch.wave_shape, pins.grid_y(65, 1.5), ch.matrix("DY1"), ch.matrix("DY2"),
                                           ch.matrix("DY4"), 1.0, 0.5, nu, dt, 100, 600)
        print(f"[pin] gpu oblique mode method {method}: measured {lm:.10e} discrete {lam:.10e} ratio {ratio:.8f}",
              flush=True)
        assert abs(ratio - 1) < 1e-5
        assert abs(lm.real - lam.real) < 2e-5 * abs(lam.real) and abs(lm.imag - lam.imag) < 2e-6 * abs(lam.imag)
