"""The learned return map: BF16 MFMA forward/backward against the CPU
oracle, GPU training, and its use inside the step with the exact fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAT = dict(type=4, density=2700.0, elastic_modulus=70e9, poisson_ratio=0.33, yield_stress=250e6,
           hardening_c0=15.0, hardening_c1=300e6)


def _random_params(seed, scale=0.3):
    from oracle import mlp_oracle as mo
    rng = np.random.default_rng(seed)
    u = mo.unpack(rng.normal(0, scale, mo.PARAM_COUNT))
    u["W1"][13:] = 0.0
    u["W4"][:, 7:] = 0.0
    u["b4"][7:] = 0.0
    return mo.pack(u).astype(np.float32)


def test_forward_matches_oracle():
    from code_mpm_amd.surrogate import Surrogate
    from oracle import mlp_oracle as mo
    s = Surrogate(seed=1)
    p = _random_params(1)
    s.set_params(p)
    x = np.random.default_rng(2).normal(0, 1, (1000, 13)).astype(np.float32)
    y = s.forward(x)
    yo, _ = mo.forward(p.astype(np.float64), x.astype(np.float64))
    err = np.abs(y - yo)
    assert err.max() < 2e-2 * np.abs(yo).max(), err.max()
    assert np.median(err) < 2e-3 * np.abs(yo).max()


def test_gradient_matches_oracle():
    from code_mpm_amd.surrogate import Surrogate
    from oracle import mlp_oracle as mo
    s = Surrogate(seed=3)
    p = _random_params(3)
    s.set_params(p)
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (333, 13)).astype(np.float32)   # not a multiple of 64
    t = rng.normal(0, 1, (333, 7)).astype(np.float32)
    g, loss = s.gradient(x, t)
    lo, go = mo.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lo) < 1e-3 * lo
    o = 0
    for name, shp in mo.SHAPES:
        k = int(np.prod(shp))
        a, b = g[o:o + k], go[o:o + k]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-12, name
        o += k


def test_training_reduces_loss_deterministically():
    from code_mpm_amd import surrogate as su
    rows = su.synthetic_states(MAT, 60000, seed=5)
    data = su.make_training_set(MAT, rows)
    assert len(data["x"]) > 20000
    losses = []
    for _ in range(2):
        s = su.Surrogate(seed=7)
        s.fit_normalization(data)
        xs, ts = s.standardise(data)
        hist = s.train(xs, ts, steps=1500, batch=2048, lr=2e-3, seed=9)
        losses.append(hist)
    assert np.array_equal(losses[0], losses[1])          # fixed-order gradient reduction
    assert losses[0][-50:].mean() < 0.1 * losses[0][:10].mean()


def _plastic_scene():
    from code_mpm_amd import scenes
    sc = scenes.random_block(seed=21, n=3000, cells=(8, 6, 5), features=False)
    sc.materials = [MAT]
    sc.material = np.zeros(sc.n, np.int32)
    rng = np.random.default_rng(21)
    rows = __import__("code_mpm_amd").surrogate.synthetic_states(MAT, sc.n, seed=22)
    sc.fields["STRESS"] = np.ascontiguousarray(rows[:, :6].T)
    sc.fields["BACKSTRESS"] = rows[:, 12].copy()
    sc.fields["VELOCITY"] = rng.normal(0, 3.0, (3, sc.n))
    return sc


def test_zero_tolerance_falls_back_to_exact_bitwise():
    from code_mpm_amd import Engine
    from code_mpm_amd.surrogate import Surrogate
    sc = _plastic_scene()
    a, b = Engine(), Engine()
    a.load_scene(sc)
    b.load_scene(sc)
    s = Surrogate(seed=0)
    b.set_surrogate(s, material=0, residual_tolerance=0.0)
    a.step(sc.dt, 5)
    b.step(sc.dt, 5)
    h = b.health()
    assert h["surrogate_calls"] > 0 and h["surrogate_fallbacks"] == h["surrogate_calls"]
    for name in ("STRESS", "STRAIN_PLASTIC", "BACKSTRESS", "ENERGY_PLASTIC", "POSITION", "VELOCITY"):
        assert np.array_equal(a.get_particle(name), b.get_particle(name)), name


def test_trained_surrogate_in_step():
    """Trained on the exact map's samples, the network predicts the radial
    scale alpha far better than the elastic predictor (alpha = 0, no
    return).  Inside the step, at the strict tolerance the stresses stay
    within 1e-3 of the exact run; at a loose tolerance predictions are
    accepted and the stress error stays within what that tolerance allows."""
    from code_mpm_amd import Engine
    from code_mpm_amd import surrogate as su
    sc = _plastic_scene()
    cap = Engine()
    cap.load_scene(sc)
    cap.set_sample_capture(200000)
    cap.step(sc.dt, 10)
    rows, produced = cap.captured_samples()
    assert produced > 300
    train_rows = np.concatenate([rows[:, :13], su.synthetic_states(MAT, 100000, seed=3)])
    data = su.make_training_set(MAT, train_rows)
    s = su.Surrogate(seed=11)
    s.fit_normalization(data)
    xs, ts = s.standardise(data)
    s.train(xs, ts, steps=3000, batch=4096, lr=2e-3, seed=12)
    # held-out accuracy of the correction vs the elastic predictor
    test = su.make_training_set(MAT, su.synthetic_states(MAT, 20000, seed=99))
    y = s.forward(((test["x"] - s.in_mean) / s.in_std).astype(np.float32))
    pred = y[:, 0] * s.out_std[0] + s.out_mean[0]
    err_net = np.linalg.norm(pred - test["t"][:, 0])
    err_none = np.linalg.norm(test["t"][:, 0])
    assert err_net < 0.05 * err_none, err_net / err_none
    steps = 5
    exact = Engine()
    exact.load_scene(sc)
    exact.step(sc.dt, steps)
    a = exact.get_particle("STRESS")
    # accepted points sit on the yield surface within tol * sigma_y, and the
    # returned deviator is a scaling of the exact trial deviator, so an
    # accepted point's stress is within about tol * sqrt(2) of the exact one
    # per step
    for tol, bound in ((1e-4, 1e-3), (5e-2, 0.5)):
        sur = Engine()
        sur.load_scene(sc)
        sur.set_surrogate(s, material=0, residual_tolerance=tol)
        sur.step(sc.dt, steps)
        h = sur.health()
        assert h["surrogate_calls"] > 0
        if tol > 1e-3:
            assert h["surrogate_fallbacks"] < h["surrogate_calls"], h
        b = sur.get_particle("STRESS")
        assert np.isfinite(b).all()
        rel = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(a, axis=0), MAT["yield_stress"])
        assert rel.max() < bound, (tol, rel.max())


def test_surrogate_on_the_honeycomb_crush():
    """The learned map on the J2 honeycomb crush (64 x 64 x 8 cells, the
    workload's generator at a test size: 2 x 2 hexagonal cells, platen at
    50 m/s so that the walls yield within the run), trained on return-map samples
    captured from seeded runs of the same generator (seeds 2 and 3), tried
    on seed 1: at tolerance 0 every answer falls back and the run equals the
    exact one bit for bit; at 1e-4 sigma_y the stresses stay within 1e-3 of
    the exact FP64 run (relative, against max(|sigma|, sigma_y))."""
    from code_mpm_amd import Engine, scenes
    from code_mpm_amd import surrogate as su
    kw = dict(cells=(64, 64, 8), hex_cells=(2, 2), platen_speed=50.0)
    rows, info = su.capture_from_runs([2, 3], 300, 400000, scene_kw=kw)
    print("captured", info, flush=True)
    assert rows.shape[0] > 50000 and np.all(rows[:, 13] == 0)
    sc = scenes.honeycomb_crush(seed=1, steps=300, **kw)
    mat = sc.materials[0]
    data = su.make_training_set(mat, rows[:, :13])
    s = su.Surrogate(seed=5)
    s.fit_normalization(data)
    xs, ts = s.standardise(data)
    hist = s.train(xs, ts, steps=3000, batch=4096, lr=2e-3, seed=6)
    assert hist[-50:].mean() < 0.2 * hist[:10].mean()
    steps = 200
    exact = Engine()
    exact.load_scene(sc)
    exact.step(sc.dt, steps)
    a = exact.get_particle("STRESS")
    sy = mat["yield_stress"]
    for tol in (0.0, 1e-4):
        sur = Engine()
        sur.load_scene(sc)
        sur.set_surrogate(s, material=0, residual_tolerance=tol)
        sur.step(sc.dt, steps)
        h = sur.health()
        b = sur.get_particle("STRESS")
        print(f"tol {tol}: {h['surrogate_calls']} network calls, {h['surrogate_fallbacks']} fallbacks", flush=True)
        assert h["surrogate_calls"] > 1000
        if tol == 0.0:
            assert h["surrogate_fallbacks"] == h["surrogate_calls"]
            for name in ("STRESS", "STRAIN_PLASTIC", "BACKSTRESS", "ENERGY_PLASTIC", "POSITION"):
                assert np.array_equal(exact.get_particle(name), sur.get_particle(name)), name
        else:
            rel = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(a, axis=0), sy)
            print(f"tol {tol}: max stress error {rel.max():.3e}, accepted "
                  f"{1 - h['surrogate_fallbacks'] / h['surrogate_calls']:.4f}", flush=True)
            assert rel.max() <= 1e-3, rel.max()
