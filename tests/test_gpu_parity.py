"""Parity of the HIP time step (through the C ABI) with the CPU oracle.

Tolerances (relative L2 over a whole field, ||a - b|| / ||b||):
  transforms, init, one Lamb evaluation      1e-12 (roundoff of different FFTs)
  FP64 step after 10 steps, u v w            1e-10 (the required parity bar)
  FP64 step after 10 steps, derived fields   1e-9: omega, Lamb, Hv, Hg, p carry first and
                                             second y-derivatives of the velocity, which
                                             amplify the last-bit differences of fused
                                             multiply-adds (GPU) against separately
                                             rounded products (oracle)
  batched vs plain method                    3x the rounding floor (the batched method
                                             reorders the arithmetic; see test_gpu_floors.py)
  mixed precision vs FP64 after 20 steps     5e-6 on velocity, 2e-5 on p (from the FP32
                                             Lamb vector), 1e-6 on wall shear
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RE180 = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5)
FIELDS = ("u", "v", "w", "omx", "omy", "omz", "lmbx", "lmby", "lmbz", "Hv", "Hg", "p")
VELOCITY = ("u", "v", "w")


def check_fields(got, ref, label):
    errs = {n: rel(got[n], ref[n]) for n in got}
    print(f"[parity] {label}: " + " ".join(f"{n}={e:.2e}" for n, e in errs.items()), flush=True)
    for n, e in errs.items():
        assert e < (1e-10 if n in VELOCITY else 1e-9), (label, n, e)


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def params(NX, NY, NZ, **kw):
    p = dict(RE180, NX=NX, NY=NY, NZ=NZ, precision=0, method=1)
    p.update(kw)
    return p


def gpu(owt_mod, p, **kw):
    print(f"[gpu] channel {p['NX']}x{p['NY']}x{p['NZ']} {kw}", flush=True)
    return owt_mod.Channel(params=p, **kw)


def orc(oracle_mod, p):
    return oracle_mod.Oracle(p["nu"], p["dt"], p["dkx"], p["dkz"], p["alpha_mesh"], p["NX"], p["NY"], p["NZ"])


# padded lengths 24..1152 cover every transform tile shape (<= 96, 192, 576, 1152)
@pytest.mark.parametrize("shape", [(16, 17, 16), (64, 65, 64), (12, 33, 24), (128, 9, 96), (384, 9, 384), (768, 7, 32),
                                   (320, 9, 320), (640, 7, 64), (16, 7, 640), (40, 7, 20)])
def test_transforms_match_oracle(owt_mod, oracle_mod, shape):
    p = params(*shape)
    ch, o = gpu(owt_mod, p), orc(oracle_mod, p)
    rng = np.random.default_rng(1)
    F = rng.standard_normal((3,) + o.wave_shape) + 1j * rng.standard_normal((3,) + o.wave_shape)
    P = ch.f2p(F)
    for q in range(3):
        assert rel(P[q], o.f2p(F[q])) < 1e-13
    back = ch.p2f(P)
    for q in range(3):
        assert rel(back[q], o.p2f(P[q])) < 1e-13


@pytest.mark.parametrize("shape", [(16, 17, 16), (64, 65, 64), (128, 9, 128), (384, 9, 384), (768, 7, 32), (320, 9, 320),
                                   (640, 7, 64)])
def test_lamb_matches_oracle(owt_mod, oracle_mod, shape):
    p = params(*shape)
    ch, o = gpu(owt_mod, p), orc(oracle_mod, p)
    rng = np.random.default_rng(2)
    six = rng.standard_normal((6,) + o.wave_shape) + 1j * rng.standard_normal((6,) + o.wave_shape)
    ref = o.lamb(six)
    got = ch.lamb(six)
    for q in range(3):
        assert rel(got[q], ref[q]) < 1e-13
    ch.set_precision(1)
    got32 = ch.lamb(six)
    for q in range(3):
        assert rel(got32[q], ref[q]) < 2e-6


@pytest.mark.parametrize("mode", [1, 2])
def test_init_fields_match_oracle(owt_mod, oracle_mod, mode):
    p = params(64, 65, 64)
    ch, o = gpu(owt_mod, p), orc(oracle_mod, p)
    ch.init_fields(mode, seed=0, amp=0.15)
    o.init(mode, seed=0, amp=0.15)
    for name in FIELDS:
        assert rel(ch.field(name), o.field(name)) < 1e-12, name
    np.testing.assert_allclose(ch.wall_shear(), o.wall_shear(), rtol=1e-13)


@pytest.fixture(scope="module")
def ten_steps_64(owt_mod, oracle_mod):
    """64x65x64, seeded start, 10 steps: oracle, plain GPU method, batched GPU method."""
    p = params(64, 65, 64)
    o = orc(oracle_mod, p)
    o.init(2, seed=0, amp=0.15)
    o.step(10)
    ref = {n: o.field(n) for n in FIELDS}
    out = {"oracle": ref, "oracle_diag": o.diag()}
    for method in (0, 1):
        ch = gpu(owt_mod, p, method=method)
        ch.init_fields(2, seed=0, amp=0.15)
        ch.step(10)
        out[method] = {n: ch.field(n) for n in FIELDS}
        out[(method, "diag")] = ch.diag()
        out[(method, "steps")] = ch.steps_done()
        ch.close()
    return out


@pytest.mark.parametrize("method", [0, 1])
def test_ten_step_parity_64(ten_steps_64, method):
    got, ref = ten_steps_64[method], ten_steps_64["oracle"]
    assert ten_steps_64[(method, "steps")] == 10
    check_fields(got, ref, f"64x65x64 method {method}")
    d, dref = ten_steps_64[(method, "diag")], ten_steps_64["oracle_diag"]
    assert d.shape == dref.shape == (11, 3)
    np.testing.assert_allclose(d, dref, rtol=1e-10, atol=0)


def test_batched_equals_plain(ten_steps_64):
    """The plain method keeps the reference's arithmetic; the batched method
    reorders it (matrix-core mat-vecs, stored DY2 v, fused substitution). Each
    is within 1.5 times the rounding floor of the oracle (test_gpu_floors.py),
    so they differ from each other by at most three times that floor."""
    import json
    floor = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "floors.json")))["re180_64"]["floor"]
    for name in ("u", "v", "w", "omx", "omy", "omz", "p"):
        e = rel(ten_steps_64[1][name], ten_steps_64[0][name])
        print(f"[parity] batched vs plain {name} {e:.2e} (floor {floor[name]:.2e})", flush=True)
        assert e <= 3 * floor[name], (name, e, floor[name])


def test_reference_initial_wave_parity(owt_mod, oracle_mod):
    """mode 1: laminar + the reference's (2, +-2) wave, 10 steps on a small grid with N = 33"""
    p = params(16, 33, 16, dkx=0.5, dkz=1.0)
    o = orc(oracle_mod, p)
    o.init(1)
    o.step(10)
    for method in (0, 1):
        ch = gpu(owt_mod, p, method=method)
        ch.init_fields(1)
        ch.step(10)
        check_fields({n: ch.field(n) for n in FIELDS}, {n: o.field(n) for n in FIELDS}, f"wave 16x33x16 method {method}")


def test_mixed_precision_close_to_fp64(owt_mod):
    p = params(64, 65, 64)
    a = gpu(owt_mod, p, method=1, precision=0)
    b = gpu(owt_mod, p, method=1, precision=1)
    for ch in (a, b):
        ch.init_fields(2, seed=0, amp=0.15)
        ch.step(20)
    for name in ("u", "v", "w"):
        assert rel(b.field(name), a.field(name)) < 5e-6, name
    assert rel(b.field("p"), a.field("p")) < 2e-5
    da, db = a.diag(), b.diag()
    assert np.abs(db[:, 0] - da[:, 0]).max() <= 1e-6 * np.abs(da[:, 0]).max()


def test_checkpoint_restore_is_exact(owt_mod, tmp_path):
    p = params(32, 33, 32)
    ch = gpu(owt_mod, p)
    ch.init_fields(2, seed=4, amp=0.2)
    ch.step(3)
    path = str(tmp_path / "state.ckpt")
    ch.checkpoint(path)
    ch.step(4)
    ref = {n: ch.field(n) for n in FIELDS}
    ch.restore(path)
    assert ch.steps_done() == 3
    ch.step(4)
    for n in FIELDS:
        assert np.array_equal(ch.field(n), ref[n]), n


def test_laminar_flow_stays_put_on_gpu(owt_mod, oracle_mod):
    p = params(16, 65, 16)
    ch, o = gpu(owt_mod, p), orc(oracle_mod, p)
    y = o.vector("y")
    u = np.zeros(o.wave_shape, dtype=complex)
    u[0, 0] = 1.5 * (1 - y ** 2)
    z = np.zeros_like(u)
    ch.set_velocity(u, z, z)
    ch.step(20)
    assert np.abs(ch.field("u")[0, 0] - u[0, 0]).max() < 1e-6
    assert np.all(np.abs(ch.diag()[:, 0] - 3 * p["nu"]) < 5e-5 * 3 * p["nu"])


def test_step_keeps_hermitian_plane_and_divergence(owt_mod):
    """The kx = 0 plane stays Hermitian in kz and the discrete divergence
    i kx u + DY1 v + i kz w vanishes. The transforms give the +kz and -kz bins
    of a real sequence to rounding only, and the v operator amplifies that
    rounding (near the walls, where v is zero in exact arithmetic, what remains
    is elimination roundoff), so the Hermitian defect is bounded by 1e-10 of the
    field's largest value; a sign or index error would be of order one."""
    p = params(32, 33, 32)
    ch = gpu(owt_mod, p)
    ch.init_fields(2, seed=9, amp=0.3)
    ch.step(5)
    DY1 = ch.matrix("DY1")
    kx = np.arange(ch.NXH) * p["dkx"]
    kz = np.array([k if k < ch.NZH else k - ch.NZ for k in range(ch.NZ)]) * p["dkz"]
    u, v, w = ch.field("u"), ch.field("v"), ch.field("w")
    for name in ("u", "v", "w", "omy", "Hv", "Hg"):
        f = ch.field(name)
        for m in range(1, ch.NZH):
            assert np.abs(f[m, 0] - np.conj(f[ch.NZ - m, 0])).max() < 1e-10 * np.abs(f).max(), name
    for k in range(ch.NZ):
        for i in range(ch.NXH):
            if kx[i] ** 2 + kz[k] ** 2 < 1e-8:
                continue
            d1v = DY1 @ v[k, i]
            div = 1j * kx[i] * u[k, i] + d1v + 1j * kz[k] * w[k, i]
            assert np.abs(div).max() <= 1e-12 * max(np.abs(d1v).max(), 1e-30)


def test_save_snapshot_is_the_reference_format(owt_mod, tmp_path):
    """save_instant_field (src/IO.cpp:198-283): F-<id_step, 8 digits>.H5 with the real
    and imaginary parts of u, v, w as [NZ][NXH][NY] doubles, equal to the fields."""
    import h5read

    if h5read.lib() is None:
        pytest.skip("HDF5 C library not available to read the snapshot back")
    p = params(16, 17, 16)
    ch = gpu(owt_mod, p)
    ch.init_fields(2, seed=0, amp=0.15)
    ch.step(3)
    path = ch.save(str(tmp_path))
    assert os.path.basename(path) == "F-00000003.H5" and os.path.exists(path)
    for name, key in (("u", "velx"), ("v", "vely"), ("w", "velz")):
        f = ch.field(name)
        re, im = h5read.read(path, key + "_real"), h5read.read(path, key + "_imag")
        assert re.shape == im.shape == ch.wave_shape
        assert np.array_equal(re, f.real) and np.array_equal(im, f.imag), name
    ch.close()


def test_ensemble_members_equal_standalone_channels(owt_mod):
    """Members of uneven size stepped together (one captured graph per step) give
    bit for bit the fields of the same channels stepped alone; a member stepped on
    its own first (ring phases differ) is still stepped correctly."""
    shapes = [(16, 17, 16), (16, 33, 16), (24, 25, 24)]
    ps = [params(*s) for s in shapes]
    print(f"[gpu] ensemble {shapes}", flush=True)
    ens = owt_mod.Ensemble(ps)
    ens.init_fields(2, seeds=[5, 6, 7], amp=0.15)
    ens.step(4)
    ens[1].step(1)
    ens.step(2)
    assert [m.steps_done() for m in ens.members] == [6, 7, 6]
    for m, (p, seed, n) in enumerate(zip(ps, (5, 6, 7), (6, 7, 6))):
        ch = gpu(owt_mod, p)
        ch.init_fields(2, seed=seed, amp=0.15)
        ch.step(n)
        for name in ("u", "v", "w", "omy", "Hv"):
            assert np.array_equal(ens[m].field(name), ch.field(name)), (m, name)
        assert np.array_equal(ens[m].diag(), ch.diag()), m
        ch.close()
    assert ens.time_steps(2) > 0
    ens.close()


@pytest.mark.parametrize("shape,precision", [((16, 17, 16), 0), ((128, 17, 128), 0), ((384, 9, 384), 0),
                                             ((768, 7, 32), 0), ((384, 9, 384), 1), ((768, 7, 32), 1)])
def test_every_tuned_launch_shape_gives_identical_results(owt_mod, shape, precision):
    """Each candidate launch shape of each tunable kernel (forced one knob at a time),
    and the autotuner's own choice, step bit for bit like the default shapes; the
    shapes cover the four transform length classes, and the FP32 transforms' extra
    candidates in mixed precision."""
    p = params(*shape, precision=precision)
    print(f"[gpu] tune candidates {shape} precision {precision}", flush=True)
    ch = gpu(owt_mod, p)

    def run(knobs):
        ch.tune = knobs
        ch.init_fields(2, seed=3, amp=0.15)
        ch.step(2)
        return [ch.field(n) for n in ("u", "v", "w", "lmbx")]

    ref = run([0] * 6)
    for k in range(6):
        for cand in (1, 2, 3):
            knobs = [0] * 6
            knobs[k] = cand
            got = run(knobs)
            assert all(np.array_equal(a, b) for a, b in zip(got, ref)), (k, cand)
    ch.init_fields(2, seed=3, amp=0.15)
    rep = ch.autotune(2)
    assert len(rep) == 6 and all(0 <= v["choice"] <= 3 for v in rep.values())
    got = run(ch.tune)
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))
    ch.close()


def test_orr_sommerfeld_eigenmode_on_gpu(owt_mod):
    """Linear dynamics of the HIP step: from the laminar profile plus a 1e-6
    Orr-Sommerfeld eigenmode (Re 1e4, alpha 1, built from the library's own DY
    matrices), the mode grows and turns at the discrete eigenvalue to the scheme's
    O(dt^3), in both methods."""
    import orr_sommerfeld as os_

    N, dt = 129, 0.01
    p = dict(params(4, N, 4, nu=os_.NU, dkx=os_.ALPHA, dt=dt))
    for method in (0, 1):
        ch = gpu(owt_mod, p, method=method)
        DY1, DY2, DY4 = ch.matrix("DY1"), ch.matrix("DY2"), ch.matrix("DY4")
        s = 2.0 * np.arange(N) / (N - 1) - 1.0
        y = np.tanh(p["alpha_mesh"] * s) / np.tanh(p["alpha_mesh"])
        lam, phi = os_.least_stable(*os_.textbook_operator(y, DY1, DY2, DY4))
        phi = phi / phi[np.argmax(np.abs(phi))]
        ch.set_velocity(*os_.eigenmode_fields(ch.wave_shape, y, DY1, phi, 1e-6))
        lm = os_.growth(lambda: ch.step(1), lambda: ch.field("v")[0, 1], phi, dt, 100, 600)
        print(f"[os] method {method}: measured {lm:.10e} discrete {lam:.10e}", flush=True)
        assert abs(lm.real - lam.real) < 2e-5 * abs(lam.real), (lm, lam)
        assert abs(lm.imag - lam.imag) < 1e-6 * abs(lam.imag), (lm, lam)
        ch.close()


def test_statistics_match_physical_space_averages(owt_mod):
    """Plane statistics by Parseval over wave space equal the plane averages of the
    physical fields (the library's dealiased transforms), which fixes the weights
    (2 for kx > 0, 1 on the kx = 0 plane) and the z-Nyquist exclusion; the laminar
    profile gives U = 1.5 (1 - y^2), no stresses and Re_tau = sqrt(3 nu) / nu."""
    p = params(16, 33, 16)
    ch = gpu(owt_mod, p)
    ch.init_fields(2, seed=4, amp=0.3)
    ch.step(3)
    st = ch.statistics()
    P = ch.f2p(np.stack([ch.field(n) for n in ("u", "v", "w")]))
    m = P.mean(axis=(1, 2))
    f = P - m[:, None, None, :]
    ref = {"U": m[0], "V": m[1], "W": m[2], "uu": (f[0] * f[0]).mean(axis=(0, 1)), "vv": (f[1] * f[1]).mean(axis=(0, 1)),
           "ww": (f[2] * f[2]).mean(axis=(0, 1)), "uv": (f[0] * f[1]).mean(axis=(0, 1)),
           "uw": (f[0] * f[2]).mean(axis=(0, 1)), "vw": (f[1] * f[2]).mean(axis=(0, 1))}
    # relative to each family's size: the means to |U|, the stresses to <u'u'>
    for k, r in ref.items():
        scale = np.abs(ref["U"]).max() if k in ("U", "V", "W") else ref["uu"].max()
        assert np.abs(st[k] - r).max() <= 1e-12 * scale, k
    assert st["uu"].max() > 1e-4
    ch.init_fields(1)
    lam = ch.statistics()
    s = 2.0 * np.arange(33) / 32 - 1.0
    y = np.tanh(1.5 * s) / np.tanh(1.5)
    assert np.abs(lam["U"] - 1.5 * (1 - y ** 2)).max() < 1e-12
    fx = ch.wall_shear()[0]
    assert abs(lam["re_tau"] - np.sqrt(fx) / p["nu"]) < 1e-12 * lam["re_tau"]
    assert abs(lam["re_tau"] - np.sqrt(3 * p["nu"]) / p["nu"]) < 1e-3 * lam["re_tau"]
    ch.close()


def test_save_fields_writes_every_dataset_family(owt_mod, tmp_path):
    import h5read

    if h5read.lib() is None:
        pytest.skip("HDF5 C library not available to read the snapshot back")
    ch = gpu(owt_mod, params(16, 17, 16))
    ch.init_fields(2, seed=0, amp=0.15)
    ch.step(2)
    path = ch.save_fields(str(tmp_path), 63)
    for name, key in (("u", "velx"), ("omy", "vory"), ("lmbz", "lmbz")):
        f = ch.field(name)
        assert np.array_equal(h5read.read(path, key + "_real"), f.real), key
        assert np.array_equal(h5read.read(path, key + "_imag"), f.imag), key
        ph = h5read.read(path, key + "_p")
        assert ph.shape == ch.phys_shape
        assert np.array_equal(ph, ch.f2p(f)), key
    ch.close()


@pytest.mark.parametrize("method", [0, 1])
def test_timing_entry_points_keep_or_advance_the_state(owt_mod, method):
    """owt_time_kernel launches a stage on the saved state and restores it: every
    field, the friction series and the step count are unchanged, and the next
    steps equal those of an untimed channel bit for bit. owt_time_steps advances
    the state by nsteps + 3 steps, like owt_step. Both refuse a channel with a
    tangent attached."""
    p = params(16, 17, 16, method=method)
    print(f"[gpu] timing entry points, method {method}", flush=True)
    ch, ref = gpu(owt_mod, p), gpu(owt_mod, p)
    for c in (ch, ref):
        c.init_fields(2, seed=4, amp=0.15)
        c.step(2)
    before = {n: ch.field(n) for n in ("u", "v", "w", "omy", "Hv", "Hg", "lmbx")}
    for name in owt_mod.KERNELS:
        assert ch.time_kernel(name, 2) >= 0
    for n, f in before.items():
        assert np.array_equal(ch.field(n), f), n
    assert ch.steps_done() == 2 and np.array_equal(ch.diag(), ref.diag())
    ch.step(3)
    ref.step(3)
    for n in before:
        assert np.array_equal(ch.field(n), ref.field(n)), n
    assert ch.time_steps(2) > 0
    ref.step(5)
    assert ch.steps_done() == ref.steps_done() == 10
    assert np.array_equal(ch.field("v"), ref.field("v"))
    if method == 1:
        # a tangent starts from an initial field
        ref.init_fields(2, seed=4, amp=0.15)
        z = np.zeros(ref.wave_shape, complex)
        ref.tangent_init(z, z, z)
        with pytest.raises(owt_mod.OwtError):
            ref.time_steps(1)
        with pytest.raises(owt_mod.OwtError):
            ref.time_kernel("solve", 1)
    ch.close()
    ref.close()


def test_autotune_leaves_the_state_unchanged(owt_mod):
    """The autotuner times its candidates on the channel's state between steps (the
    new-level slots it writes hold history level 2 there) and restores it: steps
    after tuning equal the same steps without it, bit for bit."""
    p = params(24, 33, 24)
    print("[gpu] autotune between steps", flush=True)
    a, b = gpu(owt_mod, p), gpu(owt_mod, p)
    for c in (a, b):
        c.init_fields(2, seed=8, amp=0.15)
        c.step(2)
    a.autotune(2)
    for c in (a, b):
        c.step(3)
    for n in ("u", "v", "w", "omy", "Hv", "Hg"):
        assert np.array_equal(a.field(n), b.field(n)), n
    assert np.array_equal(a.diag(), b.diag())
    a.close()
    b.close()
