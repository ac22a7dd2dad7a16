"""Rounding floors of the 10-step parity tests, measured on the CPU oracle.

Two correct implementations of the step differ by the rounding of every
transform, product and solve. This script measures how far that moves the
fields: the oracle is run from the seeded initial field, and again from the
same field with a one-ulp Gaussian perturbation of u, v, w and with every
step's Lamb vector and solved v, omega_y multiplied by (1 + 2^-52 g), g ~ N(0, 1)
(Oracle.set_noise). The floor of a field is the largest relative L2 difference
over four perturbation seeds. A case's "seed" (0 if absent) is the seed of its
initial field. The GPU parity tests require each field to stay
within 1.5 times its floor (and within 1e-10 where the floor allows).
"floor_1step" is the same measure after one step (the bar of comparisons of a
single solve).

    python tests/golden/make_floors.py [case ...]     (writes tests/golden/floors.json)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

NAMES = ("u", "v", "w", "omx", "omy", "omz", "p")
SEEDS = (1, 2, 3, 4)
STEPS = 10

# The parity cases: reference toml keys, seeded amplitude
CASES = {
    # Re_tau 180, 64x65x64 (box 4pi x 2 x 2pi)
    "re180_64": dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5, NX=64, NY=65, NZ=64, amp=0.15),
    # Re_tau 180, 128x129x128 (box 4pi x 2 x 4pi/3)
    "re180_128": dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.5, alpha_mesh=1.5, NX=128, NY=129, NZ=128, amp=0.15),
    # Re_tau 550 wall-normal grid (NY 193, stretch 2.0, nu 1e-4) on a 32x32 horizontal grid
    "re550_wall_32": dict(nu=1e-4, dt=1e-3, dkx=1.0, dkz=2.0, alpha_mesh=2.0, NX=32, NY=193, NZ=32, amp=0.075),
    # the wall-normal grids of the reference's two case files on 16x16 horizontal grids
    "case180_wall_16": dict(nu=3.571428571428571e-4, dt=2e-3, dkx=0.25, dkz=0.5, alpha_mesh=2.0, NX=16, NY=160, NZ=16,
                            amp=0.15),
    "case550_wall_16": dict(nu=1e-4, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=2.4, NX=16, NY=288, NZ=16, amp=0.075),
    # the Re_tau 550 wall-normal parameters at NY 256 on a 16x16 horizontal grid
    "re550_ny256_16": dict(nu=1e-4, dt=1e-3, dkx=1.0, dkz=2.0, alpha_mesh=2.0, NX=16, NY=256, NZ=16, amp=0.075),
    # the other two member shapes of the 32-channel Re_tau 180 ensemble (4pi x 2 x 2pi box),
    # each from the seed of its first member (16 and 24)
    "ens180_64x97": dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5, NX=64, NY=97, NZ=64, amp=0.15, seed=16),
    "ens180_96x49": dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5, NX=96, NY=49, NZ=96, amp=0.15, seed=24),
}


def make(c):
    return O.Oracle(c["nu"], c["dt"], c["dkx"], c["dkz"], c["alpha_mesh"], c["NX"], c["NY"], c["NZ"])


def run(c, seed=None, steps=STEPS):
    o = make(c)
    o.init(2, seed=c.get("seed", 0), amp=c["amp"])
    if seed is not None:
        rng = np.random.default_rng(seed)
        uvw = [o.field(n) for n in ("u", "v", "w")]
        o.set_velocity(*(f.real * (1 + 2.0 ** -52 * rng.standard_normal(f.shape))
                         + 1j * f.imag * (1 + 2.0 ** -52 * rng.standard_normal(f.shape)) for f in uvw))
        o.set_noise(seed, 2.0 ** -52, O.Oracle.NOISE_LAMB | O.Oracle.NOISE_SOLVE)
    o.step(steps)
    return {n: o.field(n) for n in NAMES}


def floors(c, steps=STEPS):
    ref = run(c, steps=steps)
    out = {n: 0.0 for n in NAMES}
    for s in SEEDS:
        f = run(c, s, steps)
        for n in NAMES:
            out[n] = max(out[n], float(np.linalg.norm(f[n] - ref[n]) / np.linalg.norm(ref[n])))
    return out


def main():
    O.set_threads(os.cpu_count() or 1)
    path = os.path.join(HERE, "floors.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    for name in sys.argv[1:] or list(CASES):
        t = time.time()
        fl, f1 = floors(CASES[name]), floors(CASES[name], 1)
        data[name] = {"steps": STEPS, "seeds": list(SEEDS), "case": CASES[name],
                      "floor": {n: float("%.3e" % v) for n, v in fl.items()},
                      "floor_1step": {n: float("%.3e" % v) for n, v in f1.items()}}
        print(name, f"{time.time() - t:.0f} s", data[name]["floor"], flush=True)
        json.dump(data, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
