"""MLP surrogate of the propagation (the role of PhotonProp.jl's Modelling
module, src/modelling.jl).

The reference fits a weighted Gamma distribution per (distance, angle) cell of
its MC output (fit_photon_dist, :52-79), transforms the inputs to (log distance,
cos angle) (source_to_input, :269-289) and trains a small Flux MLP on the fitted
parameters with an MSE loss (train_mlp, :202-267).  Here the MLP is trained
directly on MC rows with the likelihoods themselves:

  input (8)   log10 distance, cos(emitter axis, target direction),
              cos(target axis +z, direction to the emitter), target radius,
              four medium scalars (absorption length at 450 nm, scattering length
              at 450 nm, HG g, refractive-index offset), each affinely scaled
  output (10) y0 = log expected yield per emitted photon,
              3-component gamma mixture of the residual arrival time
              (logits, log shapes, log rates per 10 ns)
  loss        Poisson NLL of the summed absorption weight of each MC row +
              NLL of up to T sampled residual times per row

Training, inference and the loss run in libphotonprop.so (pp_mlp_*:
hand-written BF16 MFMA kernels, FP32 accumulation).  Queries outside the
training domain, where the predicted yield is below the floor the training MC
could resolve, or where an ensemble of networks trained on bootstrap
resamples disagrees by more than a set spread of the log yield (the
predictive uncertainty), fall back to the plain MC kernel.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .medium import SeawaterConfigMedium
from .propagation import C_VACUUM, make_requests
from .spectrum import InverseSquareSpectrum

N_IN, WIDTH, DEPTH, N_OUT = 8, 256, 4, 10
ACTIVATIONS = {"silu": _abi.PP_MLP_SILU, "relu": _abi.PP_MLP_RELU}
ACTIVATION = "silu"  # hidden activation of the networks below unless one is named

# Training domain and affine scaling of each feature: (lo, hi).
DOMAIN = {
    "log10_distance": (0.0, math.log10(200.0)),
    "cos_axis": (-1.0, 1.0),
    "cos_view": (-1.0, 1.0),
    "radius": (0.2, 0.5),
    "abs_450": (20.0, 40.0),
    "sca_450": (30.0, 70.0),
    "g": (0.85, 0.95),
    "n_offset": (-0.01, 0.01),
}
FEATURES = list(DOMAIN)


def scale_features(raw):
    """raw (n, 8) in physical units -> inputs in [-1, 1]."""
    raw = np.asarray(raw, dtype=np.float64)
    lo = np.array([DOMAIN[k][0] for k in FEATURES])
    hi = np.array([DOMAIN[k][1] for k in FEATURES])
    return (2.0 * (raw - lo) / (hi - lo) - 1.0).astype(np.float32)


def in_domain(raw, slack=1e-6):
    raw = np.asarray(raw, dtype=np.float64)
    lo = np.array([DOMAIN[k][0] for k in FEATURES]) - slack
    hi = np.array([DOMAIN[k][1] for k in FEATURES]) + slack
    return np.all((raw >= lo) & (raw <= hi), axis=1)


def medium_for(abs_450, sca_450, g, n_offset):
    """The seawater-like medium with its four scalars varied."""
    return SeawaterConfigMedium(n_300=1.36 + n_offset, n_600=1.34 + n_offset, abs_peak=abs_450,
                                abs_edge=abs_450 / 3.0, sca_450=sca_450, hg_g=g)


def t_geo(distance, radius, medium):
    """Earliest possible arrival: the straight path at the largest phase speed."""
    n_min = float(min(medium.refractive_index(300.0), medium.refractive_index(600.0)))
    return (np.asarray(distance) - np.asarray(radius)) * n_min / C_VACUUM


def random_rows(rng, n, medium_scalars):
    """n random geometries for one medium: emitter at the origin with a random
    axis, target along a random direction at a log-uniform distance, random
    target axis... the raw feature matrix and matching requests."""
    ld = rng.uniform(*DOMAIN["log10_distance"], n)
    d = 10.0 ** ld
    radius = rng.uniform(*DOMAIN["radius"], n)
    # target direction u (unit), emitter axis a with cos(a, u) uniform
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * math.pi, n)
    u = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], 1)
    ca = rng.uniform(-1, 1, n)
    # any unit vector perpendicular to u
    t = np.cross(u, np.where(np.abs(u[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]]))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    axis = ca[:, None] * u + np.sqrt(1 - ca * ca)[:, None] * t
    cos_view = -u[:, 2]  # target axis +z against the direction target -> emitter (-u)
    raw = np.zeros((n, 8))
    raw[:, 0], raw[:, 1], raw[:, 2], raw[:, 3] = ld, ca, cos_view, radius
    raw[:, 4:] = medium_scalars
    reqs = make_requests(n)
    reqs["center"] = u * d[:, None]
    reqs["axis"] = axis
    reqs["radius"] = radius
    reqs["beta"] = 1.0
    reqs["profile"] = _abi.PP_EMIT_CHERENKOV_CONE
    return raw, reqs


@dataclass
class TrainingSet:
    raw: np.ndarray        # (n, 8) physical features
    x: np.ndarray          # (n, 8) scaled
    log_n: np.ndarray      # (n,) log photons per row
    k: np.ndarray          # (n,) summed absorption weight
    times: np.ndarray      # (n, T) residual times in ns, -1 where absent
    steps: int = 0
    photons: int = 0
    meta: dict = field(default_factory=dict)


def generate_training_set(ctx, n_rows, photons_per_row=200_000, rows_per_medium=4096, n_times=16,
                          seed=0, spectrum=None, allocate=None):
    """MC rows from the batched kernel: each group of rows_per_medium rows shares
    a randomly drawn medium; each row is one request of photons_per_row photons
    with up to n_times hit times kept (a subset independent of the photons'
    histories, in the kernel's flush order, so it can differ between runs while
    the rows' counts and weights do not).  `allocate` (optional) maps a group's raw features to each
    row's photon count instead (see YieldAllocator); the Poisson likelihood
    takes every row's own exposure, so rows may differ."""
    spectrum = spectrum or InverseSquareSpectrum(300.0, 600.0)
    rng = np.random.default_rng(seed)
    raws, lns, ks, ts = [], [], [], []
    steps = 0
    photons = 0
    done = 0
    batch = 0
    while done < n_rows:
        n = min(rows_per_medium, n_rows - done)
        ms = np.array([rng.uniform(*DOMAIN[k]) for k in ("abs_450", "sca_450", "g", "n_offset")])
        medium = medium_for(*ms)
        ctx.set_medium(medium)
        raw, reqs = random_rows(rng, n, ms)
        n_ph = (np.asarray(allocate(raw), dtype=np.uint64) if allocate is not None
                else np.full(n, int(photons_per_row), dtype=np.uint64))
        reqs["n_photons"] = n_ph
        reqs["seed"] = seed * 1_000_003 + batch
        reqs["stream"] = np.arange(n, dtype=np.uint32)
        hits, off, sums = ctx.propagate_batch(reqs, spectrum, per_request_cap=n_times,
                                              hits_cap=int(min(int(n_ph.sum()), 1 << 27)),
                                              allow_overflow=True)
        tg = t_geo(np.linalg.norm(reqs["center"], axis=1), reqs["radius"], medium)
        times = -np.ones((n, n_times), dtype=np.float32)
        cnt = np.diff(off.astype(np.int64))
        row = np.repeat(np.arange(n), cnt)
        slot = np.arange(hits.shape[0]) - np.repeat(off[:-1].astype(np.int64), cnt)
        times[row, slot] = np.maximum(hits["time"].astype(np.float64) - tg[row], 0.0)
        raws.append(raw)
        lns.append(np.log(n_ph.astype(np.float64)).astype(np.float32))
        ks.append(sums["sum_w"].astype(np.float32))
        ts.append(times)
        steps += int(sums["n_steps"].sum())
        photons += int(n_ph.sum())
        done += n
        batch += 1
    raw = np.concatenate(raws)
    return TrainingSet(raw=raw, x=scale_features(raw), log_n=np.concatenate(lns),
                       k=np.concatenate(ks), times=np.concatenate(ts), steps=steps,
                       photons=photons)


class YieldAllocator:
    """Photons per training row from a pilot network's yield estimate ŷ:
    clip(target_hits / ŷ, n_min, n_max), so that rows of low yield carry about
    as many hits as rows of high yield (up to n_max).  The pilot is trained on
    a separate MC set, so a row's own count never decides its exposure."""

    def __init__(self, model, target_hits=20.0, n_min=100_000, n_max=1_000_000):
        self.model = model
        self.target_hits = float(target_hits)
        self.n_min = int(n_min)
        self.n_max = int(n_max)

    def __call__(self, raw):
        ly = self.model.forward(scale_features(raw))[:, 0].astype(np.float64)
        n = self.target_hits * np.exp(-np.clip(ly, -60.0, 0.0))
        return np.clip(np.rint(n), self.n_min, self.n_max).astype(np.uint64)


def pilot_allocator(ctx, target_hits=20.0, n_min=100_000, n_max=1_000_000, rows=131_072,
                    photons_per_row=100_000, epochs=12, seed=0, activation=None):
    """A YieldAllocator whose pilot network is trained on `rows` MC rows of
    photons_per_row photons each (drawn with their own seed).  The caller
    closes allocator.model when done."""
    ts = generate_training_set(ctx, rows, photons_per_row=photons_per_row, n_times=4,
                               seed=seed + 7_777)
    mlp = MLP(ctx, seed=seed + 99, activation=activation)
    train(mlp, ts, epochs=epochs, batch=4096, lr=3e-3, time_weight=0.0, seed=seed + 5)
    return YieldAllocator(mlp, target_hits=target_hits, n_min=n_min, n_max=n_max)


class MLP:
    """The 8 -> 256^4 -> 10 surrogate network on the GPU (pp_mlp_*).
    `activation` is the hidden activation: "relu" (the reference's network,
    src/modelling.jl:212-220) or "silu"."""

    def __init__(self, ctx, seed=0, n_in=N_IN, width=WIDTH, depth=DEPTH, n_out=N_OUT,
                 activation=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n_in, self.n_out = n_in, n_out
        self.activation = activation or ACTIVATION
        if self.activation not in ACTIVATIONS:
            raise ValueError("unknown activation %r (one of %s)" % (activation, list(ACTIVATIONS)))
        self._h = ctypes.c_void_p()
        _abi.check(self.lib.pp_mlp_create(ctx.handle, n_in, width, depth, n_out, int(seed),
                                          ctypes.byref(self._h)), self.lib)
        rc = self.lib.pp_mlp_set_activation(self._h, ACTIVATIONS[self.activation])
        if rc != _abi.PP_OK:
            self.close()
            _abi.check(rc, self.lib)
        n = ctypes.c_uint64()
        _abi.check(self.lib.pp_mlp_n_params(self._h, ctypes.byref(n)), self.lib)
        self.n_params = n.value

    def close(self):
        if self._h:
            self.lib.pp_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        _abi.check(self.lib.pp_mlp_get_params(self._h, _abi.ptr(p), self.n_params), self.lib)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _abi.check(self.lib.pp_mlp_set_params(self._h, _abi.ptr(p), self.n_params), self.lib)

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((x.shape[0], self.n_out), dtype=np.float32)
        _abi.check(self.lib.pp_mlp_forward(self._h, _abi.ptr(x), x.shape[0], _abi.ptr(y)), self.lib)
        return y

    def forward_device(self, d_x, d_y, n):
        """Device pointers (ints, e.g. torch data_ptr()); returns kernel ms."""
        _abi.check(self.lib.pp_mlp_forward_device(self._h, ctypes.c_void_p(d_x), int(n),
                                                  ctypes.c_void_p(d_y)), self.lib)
        ms = ctypes.c_float()
        _abi.check(self.lib.pp_mlp_last_forward_ms(self._h, ctypes.byref(ms)), self.lib)
        return ms.value

    def loss_grad(self, d_x, d_logn, d_k, d_times, n_times, n, time_weight=1.0, want_grad=True):
        g = np.empty(self.n_params, dtype=np.float32) if want_grad else None
        loss = np.zeros(2)
        _abi.check(self.lib.pp_mlp_loss_grad_device(
            self._h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_logn), ctypes.c_void_p(d_k),
            ctypes.c_void_p(d_times), int(n_times), int(n), float(time_weight),
            _abi.ptr(g) if want_grad else None, _abi.ptr(loss)), self.lib)
        return loss, g

    def train_step(self, d_x, d_logn, d_k, d_times, n_times, n, lr, time_weight=1.0):
        loss = np.zeros(2)
        _abi.check(self.lib.pp_mlp_train_step_device(
            self._h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_logn), ctypes.c_void_p(d_k),
            ctypes.c_void_p(d_times), int(n_times), int(n), float(lr), float(time_weight),
            _abi.ptr(loss)), self.lib)
        return loss

    def train_steps(self, d_x, d_logn, d_k, d_times, n_times, batch, lrs, time_weight=1.0):
        """len(lrs) Adam steps over consecutive minibatches of `batch` rows,
        enqueued with no host synchronisation between them; mean loss."""
        lrs = np.ascontiguousarray(lrs, dtype=np.float32)
        loss = np.zeros(2)
        _abi.check(self.lib.pp_mlp_train_steps_device(
            self._h, ctypes.c_void_p(d_x), ctypes.c_void_p(d_logn), ctypes.c_void_p(d_k),
            ctypes.c_void_p(d_times), int(n_times), int(batch), int(lrs.size), _abi.ptr(lrs),
            float(time_weight), _abi.ptr(loss)), self.lib)
        return loss


def device_data(ts, dev="cuda:0"):
    """The training set's arrays as device tensors (torch as plumbing)."""
    import torch

    return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(dev)
                 for a in (ts.x, ts.log_n, ts.k, ts.times))


def train(mlp, ts, epochs=40, batch=4096, lr=3e-3, time_weight=0.5, seed=0, log=None, rows=None,
          data=None):
    """Adam over shuffled minibatches held in device memory (torch tensors as
    plumbing); cosine learning-rate decay.  `rows` (optional) is the multiset of
    training-row indices to draw from (a bootstrap resample); `data` the
    tensors of device_data(ts), shared between networks.  Returns per-epoch
    mean losses."""
    import torch

    dev = "cuda:%d" % 0
    X, LN, K, T = data if data is not None else device_data(ts, dev)
    base = (torch.arange(X.shape[0]) if rows is None
            else torch.from_numpy(np.asarray(rows, dtype=np.int64)))
    n = int(base.shape[0])
    # start the log-yield output at the data's mean yield (the output layer is
    # initialised small, so y0 would otherwise start at log(1) and take
    # thousands of Adam steps to reach yields of 1e-4..1e-8)
    sel = base.numpy()
    p = mlp.get_params()
    p[mlp.n_params - mlp.n_out] = math.log(
        max(float(ts.k[sel].astype(np.float64).sum()), 1e-30) /
        float(np.exp(ts.log_n[sel].astype(np.float64)).sum()))
    mlp.set_params(p)
    g = torch.Generator(device="cpu").manual_seed(seed)
    steps_total = epochs * (n // batch)
    step = 0
    hist = []
    # train_steps takes each minibatch as the next block of rows
    assert X.stride(0) == X.shape[1] and T.stride(0) == T.shape[1]
    for ep in range(epochs):
        perm = base[torch.randperm(n, generator=g)].to(dev)
        # one gather per epoch; minibatches are contiguous row blocks of it
        Xp, LNp, Kp, Tp = X[perm], LN[perm], K[perm], T[perm]
        ptrs = [t.data_ptr() for t in (Xp, LNp, Kp, Tp)]
        # the library's stream does not wait for torch's: finish the gather first
        torch.cuda.current_stream().synchronize()
        nb = n // batch
        steps = step + np.arange(nb)
        lrs = lr * 0.5 * (1 + np.cos(math.pi * steps / max(1, steps_total)))
        # the epoch's minibatches in one call: no host round trip per step
        mean = mlp.train_steps(*ptrs, T.shape[1], batch, lrs, time_weight) if nb else np.zeros(2)
        step += nb
        del Xp, LNp, Kp, Tp
        hist.append(mean.tolist())
        if log:
            log("epoch %d: poisson %.5f time %.5f" % (ep, hist[-1][0], hist[-1][1]))
    return hist


class Ensemble:
    """K surrogate networks trained on bootstrap resamples of one MC training
    set.  The prediction is the members' mean log yield; their spread (sample
    standard deviation of the log yield) is the predictive uncertainty that
    HybridYield gates on."""

    def __init__(self, members):
        self.members = list(members)

    @property
    def k(self):
        return len(self.members)

    def forward_all(self, x):
        """(K, n, 10) outputs of every member."""
        return np.stack([m.forward(x) for m in self.members])

    def predict(self, x):
        """(mean log yield, its spread across members, all outputs)."""
        ys = self.forward_all(x)
        ly = ys[:, :, 0].astype(np.float64)
        std = ly.std(axis=0, ddof=1) if self.k > 1 else np.zeros(ly.shape[1])
        return ly.mean(axis=0), std, ys

    def time_mixture(self, ys):
        """The members' arrival-time mixtures as one mixture of 3K gamma
        components with weights summing to 1: (weights, shapes, rates per ns)."""
        parts = [gamma_mixture(y) for y in ys]
        w = np.concatenate([p[0] for p in parts], axis=1) / len(parts)
        return w, np.concatenate([p[1] for p in parts], axis=1), np.concatenate([p[2] for p in parts], axis=1)

    def close(self):
        for m in self.members:
            m.close()


def train_ensemble(ctx, ts, k=4, epochs=40, batch=4096, lr=3e-3, time_weight=0.5, seed=0,
                   bootstrap=True, log=None, activation=None):
    """Train k networks (different initialisation and order; each on its own
    bootstrap resample of the rows when `bootstrap`) concurrently, one host
    thread and one stream per network.  Returns (Ensemble, per-member loss
    histories)."""
    import threading

    data = device_data(ts)
    n = ts.x.shape[0]
    rng = np.random.default_rng(seed + 17)
    members = [MLP(ctx, seed=seed * 7919 + 101 * i, activation=activation) for i in range(k)]
    rows = [rng.integers(0, n, n) if bootstrap else None for _ in range(k)]
    hists, errors = [None] * k, []

    def run(i):
        try:
            lg = (lambda m: log("member %d %s" % (i, m))) if log else None
            hists[i] = train(members[i], ts, epochs=epochs, batch=batch, lr=lr,
                             time_weight=time_weight, seed=seed * 31 + i, log=lg, rows=rows[i],
                             data=data)
        except Exception as e:  # re-raised on the calling thread
            errors.append(e)

    def train_all(idx):
        threads = [threading.Thread(target=run, args=(i,)) for i in idx]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            for m in members:
                m.close()
            raise errors[0]

    train_all(range(k))
    # A member whose final Poisson loss is not finite or lies well above the
    # others' median has diverged: it is trained again from another seed (at
    # most twice), and dropped if it still does.
    for attempt in (1, 2):
        final = np.array([h[-1][0] for h in hists], dtype=np.float64)
        med = np.median(final[np.isfinite(final)]) if np.isfinite(final).any() else 0.0
        bad = [i for i in range(len(members))
               if not np.isfinite(final[i]) or final[i] > med + 0.1 * abs(med)]
        if not bad:
            break
        if log:
            log("retraining diverged members %s (final losses %s)" % (bad, final[bad]))
        for i in bad:
            members[i].close()
            members[i] = MLP(ctx, seed=seed * 7919 + 101 * i + 100003 * attempt,
                             activation=activation)
        train_all(bad)
    final = np.array([h[-1][0] for h in hists], dtype=np.float64)
    med = np.median(final[np.isfinite(final)]) if np.isfinite(final).any() else 0.0
    keep = [i for i in range(len(members)) if np.isfinite(final[i]) and final[i] <= med + 0.1 * abs(med)]
    for i in range(len(members)):
        if i not in keep:
            members[i].close()
    if len(keep) < min(2, k):
        for i in keep:
            members[i].close()
        raise RuntimeError("surrogate training diverged (final losses %s)" % final)
    return Ensemble([members[i] for i in keep]), [hists[i] for i in keep]


def gamma_mixture(y):
    """(weights, shapes, rates per ns) from network outputs y[:, 1:10]."""
    y = np.asarray(y, dtype=np.float64)
    lg = y[:, 1:4] - y[:, 1:4].max(axis=1, keepdims=True)
    w = np.exp(lg)
    w /= w.sum(axis=1, keepdims=True)
    a = np.exp(np.clip(y[:, 4:7], -4, 6))
    b = np.exp(np.clip(y[:, 7:10], -8, 6)) / 10.0
    return w, a, b


class HybridYield:
    """Yield per emitted photon from the surrogate, with the plain MC as the
    fallback for rows out of the training domain, below the yield floor, or
    (for an Ensemble) whose predicted spread of the log yield exceeds
    max_log_std.  A fallback row runs mc_photons photons when that is given,
    else clip(mc_target_hits / ŷ, mc_min, mc_max) with ŷ the surrogate's own
    estimate, so that the MC answer carries about mc_target_hits hits (a
    relative error near 1/sqrt(mc_target_hits)) wherever mc_max allows."""

    def __init__(self, model, ctx, yield_floor=1e-8, max_log_std=0.1, mc_photons=None,
                 mc_target_hits=400.0, mc_min=10_000_000, mc_max=1_000_000_000, spectrum=None):
        self.model = model
        self.ctx = ctx
        self.floor = yield_floor
        self.max_log_std = max_log_std
        self.mc_photons = mc_photons
        self.mc_target_hits = float(mc_target_hits)
        self.mc_min = int(mc_min)
        self.mc_max = int(mc_max)
        self.spectrum = spectrum or InverseSquareSpectrum(300.0, 600.0)

    def predict(self, raw):
        """Surrogate only: (yield, spread of the log yield, MC-fallback mask)."""
        x = scale_features(raw)
        if isinstance(self.model, Ensemble):
            ly, std, _ = self.model.predict(x)
        else:
            ly = self.model.forward(x)[:, 0].astype(np.float64)
            std = np.zeros_like(ly)
        yld = np.exp(ly)
        use_mc = ~in_domain(raw) | (yld < self.floor) | (std > self.max_log_std)
        return yld, std, use_mc

    def mc_budget(self, yld):
        """Photons per fallback row for surrogate estimates yld."""
        if self.mc_photons is not None:
            return np.full(np.shape(yld), int(self.mc_photons), dtype=np.uint64)
        n = self.mc_target_hits / np.maximum(np.asarray(yld, dtype=np.float64), 1e-30)
        return np.clip(np.rint(n), self.mc_min, self.mc_max).astype(np.uint64)

    def query(self, raw, reqs, medium, seed=None):
        """raw (n, 8) physical features and the matching requests (same medium).
        Returns (yield, used_mc mask, spread of the surrogate's log yield)."""
        yld, std, use_mc = self.predict(raw)
        if use_mc.any():
            self.ctx.set_medium(medium)
            r = reqs[use_mc].copy()
            r["n_photons"] = self.mc_budget(yld[use_mc])
            if seed is not None:
                r["seed"] = seed
            _, _, sums = self.ctx.propagate_batch(r, self.spectrum, hits_cap=0, allow_overflow=True)
            yld[use_mc] = sums["sum_w"] / r["n_photons"].astype(np.float64)
        return yld, use_mc, std


def held_out_pairs(ctx, n_pairs, photons, seed=1234, n_media=4, per_request_cap=0, spectrum=None,
                   allocate=None):
    """High-statistics MC on held-out pairs: n_pairs random geometries spread
    over n_media randomly drawn media, photons each (or allocate(raw) photons
    per pair: see YieldAllocator).  Returns a list per medium of dicts {medium,
    raw, reqs, hits, offsets, sums}; hits are kept up to per_request_cap per
    pair; reqs["n_photons"] holds each pair's budget."""
    spectrum = spectrum or InverseSquareSpectrum(300.0, 600.0)
    rng = np.random.default_rng(seed)
    out = []
    per = n_pairs // n_media
    for g in range(n_media):
        ms = np.array([rng.uniform(*DOMAIN[k]) for k in ("abs_450", "sca_450", "g", "n_offset")])
        medium = medium_for(*ms)
        raw, reqs = random_rows(rng, per, ms)
        ctx.set_medium(medium)
        reqs["n_photons"] = photons if allocate is None else allocate(raw)
        reqs["seed"] = seed * 1009 + g
        reqs["stream"] = np.arange(per, dtype=np.uint32)
        hits, off, sums = ctx.propagate_batch(
            reqs, spectrum, per_request_cap=per_request_cap,
            hits_cap=max(1, per_request_cap * per), allow_overflow=True)
        out.append(dict(medium=medium, raw=raw, reqs=reqs, hits=hits, offsets=off, sums=sums))
    return out


def evaluate_hybrid(hy, pairs, truth_photons=None, mc_seed=4242):
    """Surrogate and hybrid yields of held-out pairs against their MC truth.

    Per pair: truth yield Σw/N (N = truth_photons, or the pair's own budget
    when None) and its relative error bar sqrt(Σw²)/Σw (the hit count is
    Poisson, so Var Σw = E Σw²), with Σw² taken from the pair's stored hit
    weights scaled by n_hits/n_stored (the stored records are a subset
    independent of the photons' histories); a pair without stored hits gets
    1/sqrt(Σw), an upper bound (every weight is <= 1).  A far pair's photons
    arrive strongly absorbed, so its weights are small and 1/sqrt(Σw) can
    overstate the error many times.  A pair is resolved when its error bar is
    <= 5 %.  Also the surrogate yield and spread, the gate's decision and the
    hybrid's answer (the MC fallback runs with its own seed).  Returns
    (arrays, summary)."""
    import time

    cols = {k: [] for k in ("truth", "truth_err", "sumw", "surrogate", "spread", "gated",
                            "hybrid", "in_domain", "below_floor")}
    hybrid_s = mc_kernel_s = 0.0
    mc_photons = all_mc_photons = 0
    for p in pairs:
        sur, std, gate = hy.predict(p["raw"])
        t0 = time.perf_counter()
        hyb, used, _ = hy.query(p["raw"], p["reqs"], p["medium"], seed=mc_seed)
        hybrid_s += time.perf_counter() - t0
        if used.any():
            mc_kernel_s += 1e-3 * hy.ctx.last_kernel_ms()
            mc_photons += int(hy.mc_budget(sur[used]).sum())
        all_mc_photons += int(hy.mc_budget(sur).sum())
        sw = p["sums"]["sum_w"].astype(np.float64)
        n_truth = (p["reqs"]["n_photons"].astype(np.float64) if truth_photons is None
                   else float(truth_photons))
        cols["truth"].append(sw / n_truth)
        err = np.where(sw > 0, 1.0 / np.sqrt(np.maximum(sw, 1e-300)), np.inf)
        stored = p["sums"]["n_stored"].astype(np.float64)
        if p.get("hits") is not None and len(p["hits"]) and stored.sum() > 0:
            cnt = np.diff(p["offsets"].astype(np.int64))
            row = np.repeat(np.arange(len(sw)), cnt)
            w = p["hits"]["weight"].astype(np.float64)
            sw2 = np.bincount(row, weights=w * w, minlength=len(sw))
            scale = p["sums"]["n_hits"].astype(np.float64) / np.maximum(stored, 1.0)
            ok_h = (stored > 0) & (sw > 0)
            err = np.where(ok_h, np.sqrt(sw2 * scale) / np.maximum(sw, 1e-300), err)
        cols["truth_err"].append(err)
        cols["sumw"].append(sw)
        cols["surrogate"].append(sur)
        cols["spread"].append(std)
        cols["gated"].append(used)
        cols["hybrid"].append(hyb)
        cols["in_domain"].append(in_domain(p["raw"]))
        cols["below_floor"].append(sur < hy.floor)
    a = {k: np.concatenate(v) for k, v in cols.items()}
    ok = a["sumw"] > 0
    res = ok & (a["truth_err"] <= 0.05)  # truth known to <= 5 %
    rel_s = np.full(ok.shape, np.inf)
    rel_h = np.full(ok.shape, np.inf)
    rel_s[ok] = np.abs(a["surrogate"][ok] - a["truth"][ok]) / a["truth"][ok]
    rel_h[ok] = np.abs(a["hybrid"][ok] - a["truth"][ok]) / a["truth"][ok]
    a["rel_surrogate"], a["rel_hybrid"] = rel_s, rel_h
    worst = res & (rel_s >= np.percentile(rel_s[res], 90)) if res.any() else res
    keep = res & ~a["gated"]
    s = {
        "pairs": int(ok.size),
        "pairs_without_truth_hits": int((~ok).sum()),
        "pairs_resolved": int(res.sum()),
        # surrogate alone, every pair with a truth hit, with the truth's own error bar
        "surrogate_median_rel_error_all": float(np.median(rel_s[ok])),
        "surrogate_p90_rel_error_all": float(np.percentile(rel_s[ok], 90)),
        "truth_median_rel_error_bar_all": float(np.median(a["truth_err"][ok])),
        "surrogate_within_2sigma_plus_5pct": float(np.mean(
            rel_s[ok] <= 2.0 * a["truth_err"][ok] + 0.05)),
        "surrogate_median_rel_error_resolved": float(np.median(rel_s[res])),
        "surrogate_p90_rel_error_resolved": float(np.percentile(rel_s[res], 90)),
        # the gate
        "max_log_std": hy.max_log_std,
        "fallback_fraction": float(a["gated"].mean()),
        "fallback_out_of_domain": float((~a["in_domain"]).mean()),
        "fallback_below_floor": float((a["in_domain"] & a["below_floor"]).mean()),
        "fallback_spread": float((a["in_domain"] & ~a["below_floor"] & a["gated"]).mean()),
        "kept_resolved_pairs": int(keep.sum()),
        "kept_median_rel_error": float(np.median(rel_s[keep])) if keep.any() else None,
        "kept_p90_rel_error": float(np.percentile(rel_s[keep], 90)) if keep.any() else None,
        "worst_decile_sent_to_mc": float(a["gated"][worst].mean()) if worst.any() else None,
        # the gate picks out the worse answers: median surrogate error of the
        # resolved pairs it sends to MC and of those it keeps
        "gated_resolved_median_rel_error": (float(np.median(rel_s[res & a["gated"]]))
                                            if (res & a["gated"]).any() else None),
        # the hybrid's answers
        "hybrid_median_rel_error_resolved": float(np.median(rel_h[res])),
        "hybrid_median_rel_error_all": float(np.median(rel_h[ok])),
        # cost: wall time of the hybrid's queries (surrogate + MC fallback), the
        # fallback's photons and kernel time, and the photons plain MC would
        # run at the same budget rule for every pair
        "hybrid_s": hybrid_s,
        "hybrid_mc_photons": mc_photons,
        "hybrid_mc_kernel_s": mc_kernel_s,
        "all_mc_photons": all_mc_photons,
    }
    return a, s
