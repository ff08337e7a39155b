"""The learned return-map surrogate (host side).

``Surrogate`` wraps the C-ABI object: a 13 -> 64 -> 64 -> 64 -> 7 SiLU MLP
whose forward, backward and Adam steps run in hand-written BF16 MFMA kernels
on the GPU.  ``make_training_set`` turns return-map samples (captured from
exact runs) into (input, target) rows.  The von Mises return scales the trial
deviator s_tr = dev(sigma + C : de)
(ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:59-67): the
returned increment is C : de - alpha s_tr with alpha = 1 - K / sqrt(J2), so
the network predicts the scalar alpha from invariants of the trial state
formed in FP64:

  inputs   s_tr / sqrt(J2) (6), sqrt(J2) / K, (sqrt(J2) - K) / s_sigma,
           K / s_sigma, p_tr / s_sigma, R / s_sigma, |de| / s_eps,
           sqrt(J2) / s_sigma
  targets  alpha, dep_bar / s_eps, and five outputs held at zero

with K = K0 + R / sqrt(2), s_sigma = sigma_y and s_eps = sigma_y / E of the
material.  The device kernels form the same inputs (surrogate_features in
mpm_kernels.hip).
"""
import ctypes

import numpy as np

from . import capi
from .capi import check, ptr


def hooke_increment(E, nu, de):
    """Elastic trial increment C : de (engineering shears), rows of 6."""
    C = E / (1.0 + nu) / (1.0 - 2.0 * nu)
    out = np.empty_like(de)
    out[:, 0] = C * ((1.0 - nu) * de[:, 0] + nu * de[:, 1] + nu * de[:, 2])
    out[:, 1] = C * ((1.0 - nu) * de[:, 1] + nu * de[:, 0] + nu * de[:, 2])
    out[:, 2] = C * ((1.0 - nu) * de[:, 2] + nu * de[:, 0] + nu * de[:, 1])
    out[:, 3:] = C * ((0.5 - nu) * de[:, 3:])
    return out


def trial_invariants(material, sig, de, R):
    """Trial deviator s_tr (k, 6), sqrt(J2), K and the trial pressure, as the
    reference's return map forms them (VonMisesHardening.cpp:15-46)."""
    E, nu, sy = material["elastic_modulus"], material["poisson_ratio"], material["yield_stress"]
    st = sig + hooke_increment(E, nu, de)
    p = (st[:, 0] + st[:, 1] + st[:, 2]) / 3.0
    dev = st.copy()
    dev[:, :3] -= p[:, None]
    J2 = ((st[:, 0] - st[:, 1]) ** 2 + (st[:, 1] - st[:, 2]) ** 2 + (st[:, 2] - st[:, 0]) ** 2) / 6.0 \
        + st[:, 3] ** 2 + st[:, 4] ** 2 + st[:, 5] ** 2
    K = np.full(len(R), sy / np.sqrt(3.0))
    if material.get("type", 4) == 4:
        K = K + R / np.sqrt(2.0)
    return dev, np.sqrt(J2), K, p


def features(material, sig, de, R):
    """The network's 13 inputs (before standardisation), rows of 13."""
    sy, E = material["yield_stress"], material["elastic_modulus"]
    s_sig, s_eps = sy, sy / E
    dev, q, K, p = trial_invariants(material, sig, de, R)
    return np.concatenate([dev / q[:, None], (q / K)[:, None], ((q - K) / s_sig)[:, None], (K / s_sig)[:, None],
                           (p / s_sig)[:, None], (R / s_sig)[:, None],
                           (np.sqrt(np.sum(de * de, axis=1)) / s_eps)[:, None], (q / s_sig)[:, None]], axis=1)


def make_training_set(material, rows, device=0, yielding_only=True):
    """rows (k, >=13): sigma6, de6, R.  Targets come from the device's exact
    return map: alpha is the fraction of the trial deviator it removed,
    (C : de - dsigma) : s_tr / (s_tr : s_tr), shears counted twice.  Only
    rows whose trial state yields are kept by default: the engine sends only
    those to the network.  Returns dict with raw features, raw targets,
    scales."""
    sig = np.ascontiguousarray(rows[:, 0:6])
    de = np.ascontiguousarray(rows[:, 6:12])
    R = np.ascontiguousarray(rows[:, 12])
    ds, dp, dR = capi.constitutive_batch(material, R, sig, de, device=device)
    if yielding_only:
        keep = np.any(dp != 0.0, axis=1)
        sig, de, R, ds, dp, dR = sig[keep], de[keep], R[keep], ds[keep], dp[keep], dR[keep]
    E, nu, sy = material["elastic_modulus"], material["poisson_ratio"], material["yield_stress"]
    s_sig, s_eps = sy, sy / E
    corr = hooke_increment(E, nu, de) - ds
    dev, _, _, _ = trial_invariants(material, sig, de, R)
    w = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
    alpha = np.sum(w * corr * dev, axis=1) / np.sum(w * dev * dev, axis=1)
    epbar = np.sqrt(2.0 / 3.0 * np.sum(dp * dp, axis=1))
    x = features(material, sig, de, R)
    t = np.zeros((len(R), 7))
    t[:, 0] = alpha
    t[:, 1] = epbar / s_eps
    return dict(x=x, t=t, stress_scale=s_sig, strain_scale=s_eps, dstress=ds, dplastic=dp, dR=dR)


class Surrogate:
    PARAMS = None

    def __init__(self, device=0, seed=0):
        self._lib = capi.load()
        h = ctypes.c_void_p()
        check(self._lib.mpm_surrogate_create(ctypes.byref(h), int(device), int(seed)), None, "surrogate_create")
        self.handle = h
        self.device = device
        self.nparams = int(self._lib.mpm_surrogate_param_count())
        self.in_mean = np.zeros(13, np.float32)
        self.in_std = np.ones(13, np.float32)
        self.out_mean = np.zeros(7, np.float32)
        self.out_std = np.ones(7, np.float32)
        self.stress_scale = 1.0
        self.strain_scale = 1.0

    def close(self):
        if getattr(self, "handle", None):
            self._lib.mpm_surrogate_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc, what):
        check(rc, None, what)

    def params(self):
        p = np.empty(self.nparams, np.float32)
        self._ck(self._lib.mpm_surrogate_get_params(self.handle, ptr(p), self.nparams), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._ck(self._lib.mpm_surrogate_set_params(self.handle, ptr(p), self.nparams), "set_params")

    def set_normalization(self, stress_scale, strain_scale, in_mean, in_std, in_lo, in_hi, out_mean, out_std):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (in_mean, in_std, in_lo, in_hi, out_mean, out_std)]
        self._ck(self._lib.mpm_surrogate_set_normalization(self.handle, float(stress_scale), float(strain_scale),
                                                           *[ptr(a) for a in arrs]), "set_normalization")
        self.in_mean, self.in_std, self.in_lo, self.in_hi, self.out_mean, self.out_std = arrs
        self.stress_scale, self.strain_scale = stress_scale, strain_scale

    def fit_normalization(self, data, margin=0.05):
        """Standardisation and envelope from a training set (make_training_set output)."""
        x, t = data["x"], data["t"]
        im, isd = x.mean(0), x.std(0) + 1e-12
        om, osd = t.mean(0), t.std(0) + 1e-30
        z = (x - im) / isd
        lo, hi = z.min(0), z.max(0)
        span = np.maximum(hi - lo, 1e-3)   # a constant feature keeps a band that FP32 rounding stays inside
        self.set_normalization(data["stress_scale"], data["strain_scale"], im, isd, lo - margin * span,
                               hi + margin * span, om, osd)

    def standardise(self, data):
        xs = ((data["x"] - self.in_mean) / self.in_std).astype(np.float32)
        ts = ((data["t"] - self.out_mean) / self.out_std).astype(np.float32)
        return np.ascontiguousarray(xs), np.ascontiguousarray(ts)

    def forward(self, xs):
        xs = np.ascontiguousarray(xs, np.float32)
        y = np.empty((xs.shape[0], 7), np.float32)
        self._ck(self._lib.mpm_surrogate_forward(self.handle, ptr(xs), ptr(y), xs.shape[0]), "forward")
        return y

    def gradient(self, xs, ts):
        xs = np.ascontiguousarray(xs, np.float32)
        ts = np.ascontiguousarray(ts, np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = np.zeros(1, np.float32)
        self._ck(self._lib.mpm_surrogate_gradient(self.handle, ptr(xs), ptr(ts), xs.shape[0], ptr(g), ptr(loss)),
                 "gradient")
        return g, float(loss[0])

    def gradient_device(self, x_ptr, t_ptr, n, grad_ptr, loss_ptr=None):
        """mpm_surrogate_gradient_noalloc on rows already in device memory
        (raw device addresses); allocates nothing."""
        self._ck(self._lib.mpm_surrogate_gradient_noalloc(self.handle, x_ptr, t_ptr, int(n), grad_ptr, loss_ptr),
                 "gradient (no allocation)")

    def train(self, xs, ts, steps=2000, batch=4096, lr=1e-3, seed=0):
        """Adam on the GPU; returns the per-step batch losses."""
        rng = np.random.default_rng(seed)
        perm = rng.permutation(xs.shape[0])
        xs = np.ascontiguousarray(xs[perm], np.float32)
        ts = np.ascontiguousarray(ts[perm], np.float32)
        self._ck(self._lib.mpm_surrogate_load_training_set(self.handle, ptr(xs), ptr(ts), xs.shape[0]), "load")
        hist = np.zeros(steps, np.float32)
        self._ck(self._lib.mpm_surrogate_train(self.handle, int(steps), int(batch), float(lr), int(seed), ptr(hist)),
                 "train")
        return hist


def content_order(rows):
    """A permutation of `rows` that depends only on their contents: rows
    sorted by a 64-bit mix of each row's bit patterns.  The capture appends
    rows with one atomic per block, so the same set arrives in a different
    order from run to run; in this order it does not (rows with equal bits
    are interchangeable, and two different rows share a key with
    probability about k^2 / 2^65)."""
    b = np.ascontiguousarray(rows, dtype=np.float64).view(np.uint64)
    h = np.zeros(b.shape[0], dtype=np.uint64)
    mul, sh = np.uint64(0x9E3779B97F4A7C15), np.uint64(29)
    for j in range(b.shape[1]):
        h ^= b[:, j]
        h *= mul
        h ^= h >> sh
    return np.argsort(h, kind="stable")


def capture_from_runs(seeds, steps, capacity_per_run, scene_kw=None, device=0, log=None):
    """Return-map samples the exact map itself produces on seeded honeycomb
    crushes (scenes.honeycomb_crush(seed=s, **scene_kw)): each run steps
    with the exact FP64 map and records every yielding point's inputs
    (sigma6, dstrain6, R, material index) through mpm_set_sample_capture,
    until at least `capacity_per_run` rows are held or `steps` steps are
    done.  The device keeps every row of those steps (its buffer has room
    for the last chunk of steps too), and the run keeps `capacity_per_run`
    of them chosen by content_order, so the training set, and its order,
    is the same in every run.  Returns (rows (k, 14), per-run
    [(seed, steps run, rows kept)])."""
    from . import scenes
    from .engine import Engine
    out, info = [], []
    for seed in seeds:
        sc = scenes.honeycomb_crush(seed=seed, steps=steps, **(scene_kw or {}))
        eng = Engine(device=device, order="pair")
        eng.load_scene(sc)
        eng.set_sample_capture(capacity_per_run + capacity_per_run // 2 + 1)
        done, chunk = 0, max(1, steps // 20)
        while done < steps:
            k = min(chunk, steps - done)
            eng.step(sc.dt, k)
            done += k
            if eng.captured_samples_count() >= capacity_per_run:
                break
        rows, produced = eng.captured_samples()
        eng.close()
        if produced > rows.shape[0] and log:
            log(f"capture seed {seed}: {produced - rows.shape[0]} rows beyond the buffer were dropped")
        rows = rows[content_order(rows)[:capacity_per_run]]
        out.append(rows)
        info.append((int(seed), int(done), int(rows.shape[0])))
        if log:
            log(f"capture seed {seed}: {sc.n} particles, {done} steps, {rows.shape[0]} rows (of {produced})")
    return np.concatenate(out), info


def synthetic_states(material, n, seed=0, strain_scale=None, r_max=None, on_surface=0.8):
    """Return-map inputs spread over the plastic regime: stresses on the
    yield surface (fraction on_surface, as after a plastic step) or just
    inside it (random deviatoric direction, random pressure), strain
    increments of size ~strain_scale, hardening variable in [0, r_max].
    Rows: sigma6, de6, R."""
    rng = np.random.default_rng(seed)
    E, sy = material["elastic_modulus"], material["yield_stress"]
    if strain_scale is None:
        strain_scale = 0.02 * sy / E
    if r_max is None:
        r_max = 0.5 * sy if material.get("type", 4) == 4 else 0.0
    R = rng.uniform(0.0, r_max, n)
    K = sy / np.sqrt(3.0) + R / np.sqrt(2.0)
    d = rng.normal(0, 1, (n, 6))
    d[:, :3] -= d[:, :3].mean(1, keepdims=True)
    j2 = 0.5 * np.sum(d[:, :3] ** 2, 1) + np.sum(d[:, 3:] ** 2, 1)
    radius = np.where(rng.uniform(size=n) < on_surface, 1.0, rng.uniform(0.98, 1.0, n))
    d *= (K * radius / np.sqrt(j2))[:, None]
    p = rng.normal(0, 0.5 * sy, n)
    sig = d.copy()
    sig[:, :3] += p[:, None]
    de = rng.normal(0, strain_scale, (n, 6))
    return np.concatenate([sig, de, R[:, None]], axis=1)
