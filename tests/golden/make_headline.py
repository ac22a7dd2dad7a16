"""Oracle reference of the Re_tau 550 workload at full size: 384x193x384,
100 FP64 steps from the seeded field, for the GPU test that compares both
precisions with it (tests/test_gpu_headline.py).

The full fields are 228 MB each, so the fixture keeps what the test compares:

- a fixed set of 64 pencils (8 kz indices x 8 kx indices, from the mean and
  the largest scales up to kx index 64 and kz index +-16) of u, v, w, omega_y
  and p after 100 steps, in the library's [kz][kx][y] layout;
- the friction series (the step's diagnostics: fx_fric, wall shear of the
  lower and upper wall) of all 100 steps;
- the rounding floor of the same quantities: the oracle rerun from the same
  field with the one-ulp perturbation of make_floors.py (initial u, v, w and
  every step's Lamb vector and solved v, omega_y), the largest relative L2
  difference over the perturbation seeds on the stored pencils.

The oracle factors each distinct k^2 once (Oracle.set_factor_cache: the same
fields bit for bit as the reference's per-pencil factoring). The script saves
after every run, so a partial run leaves a usable fixture with fewer seeds.

    python tests/golden/make_headline.py [nseeds]   (writes tests/golden/re550_384_100.npz;
                                                     resumes from it when present)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_floors as MF  # noqa: E402

O = MF.O
CASE = dict(nu=1e-4, dt=1e-3, dkx=1.0, dkz=2.0, alpha_mesh=2.0, NX=384, NY=193, NZ=384, amp=0.075)
STEPS = 100
NAMES = ("u", "v", "w", "omy", "p")
KZ = np.array([0, 1, 2, 4, 8, 16, 384 - 1, 384 - 4])
KX = np.array([0, 1, 2, 4, 8, 16, 32, 64])
PATH = os.path.join(HERE, "re550_384_100.npz")


def run(seed=None):
    c = CASE
    o = MF.make(c)
    o.set_factor_cache(True)  # bit-identical to per-pencil factoring, hours faster
    o.init(2, seed=0, amp=c["amp"])
    if seed is not None:
        rng = np.random.default_rng(seed)
        uvw = [o.field(n) for n in ("u", "v", "w")]
        o.set_velocity(*(f.real * (1 + 2.0 ** -52 * rng.standard_normal(f.shape))
                         + 1j * f.imag * (1 + 2.0 ** -52 * rng.standard_normal(f.shape)) for f in uvw))
        del uvw
        o.set_noise(seed, 2.0 ** -52, O.Oracle.NOISE_LAMB | O.Oracle.NOISE_SOLVE)
    t = time.time()
    for s in range(STEPS):
        o.step(1)
        if s % 10 == 9:
            print(f"  seed {seed}: step {s + 1} ({time.time() - t:.0f} s)", flush=True)
    out = {}
    for n in NAMES:
        f = o.field(n)
        out[n] = f[KZ][:, KX].copy()
        del f
    out["diag"] = o.diag()
    return out


def main():
    O.set_threads(os.cpu_count() or 1)
    nseeds = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    if os.path.exists(PATH):
        # resume: keep the reference run and the seeds already measured
        data = dict(np.load(PATH))
        ref = {n: data[n] for n in NAMES + ("diag",)}
    else:
        ref = run()
        data = {n: ref[n] for n in NAMES}
        data.update(diag=ref["diag"], kz=KZ, kx=KX, steps=STEPS, seeds=np.zeros(0, int),
                    case=np.array([CASE[k] for k in ("nu", "dt", "dkx", "dkz", "alpha_mesh", "NX", "NY", "NZ", "amp")]))
        for n in NAMES + ("diag",):
            data["floor_" + n] = 0.0
        np.savez(PATH, **data)
    for seed in range(len(data["seeds"]) + 1, nseeds + 1):
        f = run(seed)
        # the perturbed start runs the initial post-processing once more (one more
        # friction row before the steps): compare the rows of the steps
        f["diag"] = f["diag"][-len(ref["diag"]):]
        for n in NAMES + ("diag",):
            d = float(np.linalg.norm(f[n] - ref[n]) / np.linalg.norm(ref[n]))
            data["floor_" + n] = max(float(data["floor_" + n]), d)
        data["seeds"] = np.arange(1, seed + 1)
        np.savez(PATH, **data)
        print(f"seed {seed}:", {n: "%.2e" % data["floor_" + n] for n in NAMES + ("diag",)}, flush=True)


if __name__ == "__main__":
    main()
