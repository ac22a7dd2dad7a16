"""Tangent and discrete adjoint of the time step (FP64), through the C ABI.

Nothing in the reference differentiates the step, so the checks are the ones the
mathematics fixes:
  tangent   central finite differences of the forward step (itself checked
            against the oracle): fields after K steps and the objective
  adjoint   the adjoint identity <grad J, d> = dJ[d] against the tangent, which
            holds to roundoff for an exact transpose, and central finite
            differences of J; checkpoint spacing does not change the gradient
The objective is the mean friction force fx_fric (mean wall shear) over steps
first..last, the gradient taken with respect to the initial u, v, w.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RE180 = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5)


def params(NX, NY, NZ):
    return dict(RE180, NX=NX, NY=NY, NZ=NZ, precision=0, method=1)


def directions(shape, seeds):
    out = []
    for s in seeds:
        rng = np.random.default_rng(s)
        out.append([0.1 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) for _ in range(3)])
    return out


def run_J(ch, x, K, first, last):
    ch.set_velocity(*x)
    ch.step(K)
    return ch.diag()[first:last + 1, 0].mean(), [ch.field(n) for n in ("u", "v", "w")]


@pytest.fixture(scope="module")
def base(owt_mod):
    p = params(16, 33, 16)
    print("[gpu] adjoint base 16x33x16", flush=True)
    ch = owt_mod.Channel(params=p)
    ch.init_fields(2, seed=0, amp=0.3)
    x0 = [ch.field(n) for n in ("u", "v", "w")]
    yield ch, x0
    ch.close()


def test_tangent_matches_finite_differences(base):
    ch, x0 = base
    K, first, last, eps = 6, 2, 6, 1e-6
    d = directions(x0[0].shape, [11])[0]
    Jp, Fp = run_J(ch, [a + eps * b for a, b in zip(x0, d)], K, first, last)
    Jm, Fm = run_J(ch, [a - eps * b for a, b in zip(x0, d)], K, first, last)
    ch.set_velocity(*x0)
    t = ch.tangent_init(*d)
    ch.step(K)
    for n, name in enumerate(("u", "v", "w")):
        fd = (Fp[n] - Fm[n]) / (2 * eps)
        err = np.linalg.norm(t.field(name) - fd) / np.linalg.norm(fd)
        print(f"[tangent] {name} vs fd {err:.2e}", flush=True)
        assert err < 1e-6, (name, err)
    dJ_t = t.diag()[first:last + 1, 0].mean()
    dJ_fd = (Jp - Jm) / (2 * eps)
    assert abs(dJ_t - dJ_fd) < 1e-6 * abs(dJ_fd), (dJ_t, dJ_fd)


def test_adjoint_identity_and_finite_differences(base):
    ch, x0 = base
    K, first, last, eps = 6, 2, 6, 1e-6
    ch.set_velocity(*x0)
    J, g = ch.adjoint_run(K, 2, first, last)
    J0, _ = run_J(ch, x0, K, first, last)
    assert abs(J - J0) <= 1e-14 * abs(J0)
    for d in directions(x0[0].shape, [21, 22, 23]):
        dJ_a = sum(np.vdot(gg, dd).real for gg, dd in zip(g, d))
        ch.set_velocity(*x0)
        t = ch.tangent_init(*d)
        ch.step(K)
        dJ_t = t.diag()[first:last + 1, 0].mean()
        Jp, _ = run_J(ch, [a + eps * b for a, b in zip(x0, d)], K, first, last)
        Jm, _ = run_J(ch, [a - eps * b for a, b in zip(x0, d)], K, first, last)
        dJ_fd = (Jp - Jm) / (2 * eps)
        print(f"[adjoint] <g,d>={dJ_a:.15e} tangent={dJ_t:.15e} fd={dJ_fd:.15e}", flush=True)
        assert abs(dJ_a - dJ_t) <= 1e-10 * abs(dJ_t)
        assert abs(dJ_a - dJ_fd) <= 1e-6 * abs(dJ_fd)


def test_adjoint_checkpoint_spacing_does_not_change_gradient(base):
    ch, x0 = base
    ch.set_velocity(*x0)
    J1, g1 = ch.adjoint_run(7, 1, 1, 7)
    ch.set_velocity(*x0)
    J7, g7 = ch.adjoint_run(7, 7, 1, 7)
    ch.set_velocity(*x0)
    J3, g3 = ch.adjoint_run(7, 3, 1, 7)
    assert J1 == J7 == J3
    for a, b, c in zip(g1, g7, g3):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert ch.steps_done() == 7


def test_adjoint_rejects_mixed_precision_and_bad_ranges(owt_mod):
    ch = owt_mod.Channel(params=dict(params(16, 17, 16), precision=1))
    ch.init_fields(2, seed=0, amp=0.15)
    with pytest.raises(owt_mod.OwtError):
        ch.adjoint_run(3, 1, 1, 3)
    ch.set_precision(0)
    ch.init_fields(2, seed=0, amp=0.15)
    with pytest.raises(owt_mod.OwtError):
        ch.adjoint_run(3, 1, 2, 4)
    ch.close()


def test_torch_autograd_function_harness(base):
    """owt_adjoint_run behind torch.autograd.Function: torch's backward returns the
    adjoint gradient (torch's convention for a real loss of complex inputs is
    d/dRe + i d/dIm, the same as the library's), and a directional derivative
    through the Function matches central differences of the Function itself."""
    torch = pytest.importorskip("torch")
    ch, x0 = base
    K, first, last = 6, 2, 6

    class MeanWallShear(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u, v, w):
            ch.set_velocity(u.detach().numpy(), v.detach().numpy(), w.detach().numpy())
            J, g = ch.adjoint_run(K, 3, first, last)
            ctx.save_for_backward(*[torch.from_numpy(a) for a in g])
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gJ):
            return tuple(gJ * gi for gi in ctx.saved_tensors)

    xs = [torch.tensor(a, requires_grad=True) for a in x0]
    J = MeanWallShear.apply(*xs)
    J.backward()
    d = directions(x0[0].shape, [31])[0]
    dJ = sum(float((torch.conj(x.grad) * torch.from_numpy(dd)).real.sum()) for x, dd in zip(xs, d))
    eps = 1e-6
    with torch.no_grad():
        Jp = MeanWallShear.apply(*[torch.from_numpy(a + eps * b) for a, b in zip(x0, d)])
        Jm = MeanWallShear.apply(*[torch.from_numpy(a - eps * b) for a, b in zip(x0, d)])
    fd = float(Jp - Jm) / (2 * eps)
    print(f"[torch] dJ={dJ:.15e} fd={fd:.15e}", flush=True)
    assert abs(dJ - fd) <= 1e-6 * abs(fd)


def _dot(a, b):
    return sum(float(np.vdot(x, y).real) for x, y in zip(a, b))


@pytest.mark.parametrize("shape,K,ckpt", [((16, 33, 16), 6, 2), ((128, 129, 128), 20, 5)])
def test_adjoint_apply_dot_product_identity(owt_mod, shape, K, ckpt):
    """<L d, w> = <d, L^T w> for the map L from the initial velocity to the
    velocity after K steps, with white-noise d (through owt_tangent_init) and
    white-noise w (through owt_adjoint_apply): every mode of every field."""
    print(f"[gpu] adjoint apply {shape} K={K}", flush=True)
    ch = owt_mod.Channel(params=params(*shape))
    ch.init_fields(2, seed=0, amp=0.15)
    x0 = [ch.field(n) for n in ("u", "v", "w")]
    d, w = directions(x0[0].shape, [41, 42])
    t = ch.tangent_init(*d)
    ch.step(K)
    Ld = [t.field(n) for n in ("u", "v", "w")]
    ch.set_velocity(*x0)
    LTw = ch.adjoint_apply(K, *w, ckpt_every=ckpt)
    assert ch.steps_done() == K
    a, b = _dot(Ld, w), _dot(d, LTw)
    scale = sum(float(np.sum(np.abs(x.real * y.real) + np.abs(x.imag * y.imag))) for x, y in zip(Ld, w))
    print(f"[adjoint] <Ld,w>={a:.15e} <d,LTw>={b:.15e} rel {abs(a - b) / abs(a):.2e} "
          f"(terms {abs(a - b) / scale:.2e})", flush=True)
    assert abs(a - b) <= 1e-12 * abs(a)
    ch.close()


def test_adjoint_config_re180_128_K20(owt_mod):
    """The adjoint workload at full size: Re_tau 180, 128x129x128, K = 20 steps,
    checkpoints every 5, J = mean friction force over steps 11..20. The adjoint
    identity against the tangent on two white-noise directions, and central
    finite differences (step 1e-6) on two smooth directions of the state's own
    scale: an O(1) mean-flow profile plus the seeded divergence-free
    perturbation (along white noise J is not linear enough at that step, and
    along divergence-free perturbations alone dJ is of second order)."""
    K, ckpt, first, last, eps = 20, 5, 11, 20, 1e-6
    p = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.5, alpha_mesh=1.5, NX=128, NY=129, NZ=128, precision=0, method=1)
    print("[gpu] adjoint 128x129x128 K=20", flush=True)
    ch = owt_mod.Channel(params=p)
    NY = ch.NY
    s = 2.0 * np.arange(NY) / (NY - 1) - 1.0
    y = np.tanh(p["alpha_mesh"] * s) / np.tanh(p["alpha_mesh"])
    ch.init_fields(2, seed=0, amp=0.15)
    x0 = [ch.field(n) for n in ("u", "v", "w")]
    ch.set_velocity(*x0)
    J, g = ch.adjoint_run(K, ckpt, first, last)

    def J_of(x):
        ch.set_velocity(*x)
        ch.step(K)
        return ch.diag()[first:last + 1, 0].mean()

    assert abs(J - J_of(x0)) <= 1e-14 * abs(J)
    for d in directions(x0[0].shape, [1000, 1001]):
        ch.set_velocity(*x0)
        t = ch.tangent_init(*d)
        ch.step(K)
        dJ_t = float(t.diag()[first:last + 1, 0].mean())
        dJ_a = _dot(g, d)
        print(f"[adjoint 128] identity <g,d>={dJ_a:.15e} tangent={dJ_t:.15e} rel {abs(dJ_a - dJ_t) / abs(dJ_t):.2e}",
              flush=True)
        assert abs(dJ_a - dJ_t) <= 1e-12 * abs(dJ_t)
    for sd in (1000, 1001):
        ch.init_fields(2, seed=sd, amp=0.15)
        d = [ch.field(n) for n in ("u", "v", "w")]
        c = np.random.default_rng(sd).standard_normal(4)
        d[0][0, 0] = (1 - y ** 2) * (c[0] + c[1] * y + c[2] * y ** 2 + c[3] * y ** 3)
        dJ_a = _dot(g, d)
        fd = (J_of([a + eps * b for a, b in zip(x0, d)]) - J_of([a - eps * b for a, b in zip(x0, d)])) / (2 * eps)
        print(f"[adjoint 128] fd <g,d>={dJ_a:.15e} fd={fd:.15e} rel {abs(dJ_a - fd) / abs(fd):.2e}", flush=True)
        assert abs(dJ_a - fd) <= 1e-6 * abs(fd)
    ch.close()
