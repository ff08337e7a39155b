"""The learned return-map surrogate (host side).

``Surrogate`` wraps the C-ABI object: a 13 -> 64 -> 64 -> 64 -> 7 SiLU MLP
whose forward, backward and Adam steps run in hand-written BF16 MFMA kernels
on the GPU.  ``make_training_set`` turns return-map samples (captured from
exact runs) into standardised (input, target) rows:

  inputs   sigma / s_sigma (6), dstrain / s_eps (6), R / s_sigma (1)
  targets  (dsigma_trial - dsigma_exact) / s_sigma (6), dep_bar / s_eps (1)

with s_sigma = sigma_y and s_eps = sigma_y / E of the material; the targets'
stress part is the plastic correction the radial return applies to the
exact elastic trial increment
(ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:59-67).
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


def make_training_set(material, rows, device=0, yielding_only=True):
    """rows (k, >=13): sigma6, de6, R.  Targets come from the device's exact
    return map.  Only rows whose trial state yields are kept by default: the
    engine sends only those to the network.  Returns dict with raw features,
    raw targets, scales."""
    sig = np.ascontiguousarray(rows[:, 0:6])
    de = np.ascontiguousarray(rows[:, 6:12])
    R = np.ascontiguousarray(rows[:, 12])
    ds, dp, dR = capi.constitutive_batch(material, R, sig, de, device=device)
    if yielding_only:
        keep = np.any(dp != 0.0, axis=1)
        sig, de, R, ds, dp, dR = sig[keep], de[keep], R[keep], ds[keep], dp[keep], dR[keep]
    E, nu, sy = material["elastic_modulus"], material["poisson_ratio"], material["yield_stress"]
    s_sig, s_eps = sy, sy / E
    trial = hooke_increment(E, nu, de)
    corr = trial - ds
    epbar = np.sqrt(2.0 / 3.0 * np.sum(dp * dp, axis=1))
    x = np.concatenate([sig / s_sig, de / s_eps, (R / s_sig)[:, None]], axis=1)
    t = np.concatenate([corr / s_sig, (epbar / s_eps)[:, None]], axis=1)
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
        span = hi - lo
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
