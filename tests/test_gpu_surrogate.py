"""MLP surrogate kernels (BF16 MFMA) against a plain PyTorch FP32 reference of
the same network and loss, plus a short end-to-end train on MC rows."""
import math

import numpy as np
import pytest
import torch

import photonprop as pp
from photonprop import surrogate as sg

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def unpack(params, dims):
    out, o = [], 0
    for l in range(len(dims) - 1):
        ni, no = dims[l], dims[l + 1]
        W = torch.from_numpy(params[o:o + no * ni].reshape(no, ni).copy()); o += no * ni
        b = torch.from_numpy(params[o:o + no].copy()); o += no
        out.append((W, b))
    return out


ACTS = ("silu", "relu")


def ref_forward(layers, x, requires_grad=False, act="silu"):
    """FP32 reference with the kernels' BF16 rounding of weights and of each
    layer's input activations."""
    a = bf(torch.from_numpy(x))
    params = []
    for l, (W, b) in enumerate(layers):
        Wb = bf(W).clone().requires_grad_(requires_grad)
        bb = b.clone().requires_grad_(requires_grad)
        params += [Wb, bb]
        z = a @ Wb.T + bb
        if l + 1 < len(layers):
            s = z * torch.sigmoid(z) if act == "silu" else torch.relu(z)
            a = s + (bf(s) - s).detach()  # BF16 activations, straight-through gradient
        else:
            return z, params


def ref_loss(y, logn, k, times, tw):
    y = y.double()
    M = y.shape[0]
    l = y[:, 0] + logn
    ex = torch.clamp(l - 30.0, min=0.0)
    lp = torch.exp(torch.clamp(l, max=30.0)) * (1 + ex + 0.5 * ex * ex) - k * l
    logpi = torch.log_softmax(y[:, 1:4], dim=1)
    la = torch.clamp(y[:, 4:7], -4, 6)
    lb = torch.clamp(y[:, 7:10], -8, 6)
    a, b = torch.exp(la), torch.exp(lb)
    mask = times >= 0
    ts = torch.where(mask, times, torch.zeros_like(times)) / 10.0 + 1e-3
    lt_ = torch.log(ts)[:, :, None]
    lpc = (logpi[:, None, :] + a[:, None, :] * lb[:, None, :] + (a[:, None, :] - 1) * lt_
           - b[:, None, :] * ts[:, :, None] - torch.lgamma(a)[:, None, :])
    ll = torch.logsumexp(lpc, dim=2)
    nt = mask.sum(1)
    lt = -(ll * mask).sum(1) / torch.clamp(nt, min=1)
    return (lp.sum() + tw * lt.sum()) / M, lp.mean(), lt.mean()


def batch(rng, M, T=8):
    x = rng.uniform(-1, 1, (M, 8)).astype(np.float32)
    logn = np.full(M, math.log(30.0), dtype=np.float32)
    k = rng.poisson(30.0, M).astype(np.float32)
    times = rng.gamma(2.0, 15.0, (M, T)).astype(np.float32)
    times[rng.uniform(size=(M, T)) < 0.3] = -1.0
    return x, logn, k, times


def to_dev(*arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("act", ACTS)
def test_forward_matches_torch_fp32(ctx, act):
    mlp = sg.MLP(ctx, seed=1, activation=act)
    dims = [8, 256, 256, 256, 256, 10]
    layers = unpack(mlp.get_params(), dims)
    rng = np.random.default_rng(0)
    for M in (1, 37, 1000, 4099):
        x = rng.uniform(-1, 1, (M, 8)).astype(np.float32)
        y = mlp.forward(x)
        ref, _ = ref_forward(layers, x, act=act)
        ref = ref.detach().numpy()
        err = np.abs(y - ref).max() / max(np.abs(ref).max(), 1e-6)
        say("%s forward M=%d max rel err %.2e" % (act, M, err))
        assert err < 2e-2
    mlp.close()


@pytest.mark.parametrize("shape", [(5, 48, 2, 3), (8, 256, 4, 16), (20, 112, 3, 1),
                                   (8, 512, 2, 10)])
def test_forward_other_shapes_match_torch_fp32(ctx, shape):
    """Widths that leave half a 32-feature tile (48, 112), an odd input width,
    1 and 16 outputs, and a width past the fused kernel's limit (512, layer by
    layer on the GEMM kernel)."""
    n_in, width, depth, n_out = shape
    mlp = sg.MLP(ctx, seed=3, n_in=n_in, width=width, depth=depth, n_out=n_out, activation="relu")
    dims = [n_in] + [width] * depth + [n_out]
    p = mlp.get_params()
    rng = np.random.default_rng(11)
    p[:] = rng.normal(0, 0.1, p.size).astype(np.float32)  # non-zero biases too
    mlp.set_params(p)
    layers = unpack(mlp.get_params(), dims)
    for M in (3, 300):
        x = rng.uniform(-1, 1, (M, n_in)).astype(np.float32)
        y = mlp.forward(x)
        ref = ref_forward(layers, x, act="relu")[0].detach().numpy()
        err = np.abs(y - ref).max() / max(np.abs(ref).max(), 1e-6)
        say("shape %s M=%d max rel err %.2e" % (shape, M, err))
        assert y.shape == (M, n_out) and err < 2e-2
    mlp.close()


@pytest.mark.parametrize("act", ACTS)
def test_loss_and_gradients_match_torch_autograd(ctx, act):
    mlp = sg.MLP(ctx, seed=2, activation=act)
    dims = [8, 256, 256, 256, 256, 10]
    # make the output layer non-trivial so every mixture term gets gradient
    p = mlp.get_params()
    rng = np.random.default_rng(3)
    p[-(10 * 256 + 10):] = rng.normal(0, 0.05, 10 * 256 + 10).astype(np.float32)
    mlp.set_params(p)
    layers = unpack(mlp.get_params(), dims)
    M, T = 512, 8
    x, logn, k, times = batch(rng, M, T)
    dx, dln, dk, dt = to_dev(x, logn, k, times)
    tw = 0.7
    loss, g = mlp.loss_grad(dx.data_ptr(), dln.data_ptr(), dk.data_ptr(), dt.data_ptr(), T, M, tw)
    y, params = ref_forward(layers, x, requires_grad=True, act=act)
    L, lp, lt = ref_loss(y, torch.from_numpy(logn).double(), torch.from_numpy(k).double(),
                         torch.from_numpy(times).double(), tw)
    L.backward()
    say("loss lib", loss, "ref", float(lp), float(lt))
    assert loss[0] == pytest.approx(float(lp), rel=2e-2, abs=1e-3)
    assert loss[1] == pytest.approx(float(lt), rel=2e-2, abs=1e-3)
    o = 0
    for l, (W, b) in enumerate(layers):
        for ref_g in (params[2 * l].grad, params[2 * l + 1].grad):
            n = ref_g.numel()
            gl = torch.from_numpy(g[o:o + n]).double()
            gr = ref_g.reshape(-1).double()
            o += n
            cos = float(gl @ gr / (gl.norm() * gr.norm() + 1e-30))
            rel = float((gl.norm() - gr.norm()).abs() / (gr.norm() + 1e-30))
            say("%s layer %d grad n=%d cos %.5f |rel norm| %.3e" % (act, l, n, cos, rel))
            assert cos > 0.99 and rel < 0.05
    assert o == mlp.n_params
    mlp.close()


def test_loss_beyond_the_log_mean_cap(ctx):
    """Rows whose log mean exceeds the cap (30) keep a finite, bounded gradient
    equal to the reference's (mu grows linearly there), so an overshooting step
    cannot overflow Adam's second moment."""
    mlp = sg.MLP(ctx, seed=6)
    dims = [8, 256, 256, 256, 256, 10]
    layers = unpack(mlp.get_params(), dims)
    rng = np.random.default_rng(7)
    M, T = 256, 4
    x, _, k, times = batch(rng, M, T)
    logn = np.full(M, 33.0, dtype=np.float32)  # l = y0 + 33 > 30 for every row
    dx, dln, dk, dt = to_dev(x, logn, k, times)
    loss, g = mlp.loss_grad(dx.data_ptr(), dln.data_ptr(), dk.data_ptr(), dt.data_ptr(), T, M, 0.0)
    assert np.all(np.isfinite(g)) and np.isfinite(loss).all()
    y, params = ref_forward(layers, x, requires_grad=True)
    L, lp, _ = ref_loss(y, torch.from_numpy(logn).double(), torch.from_numpy(k).double(),
                        torch.from_numpy(times).double(), 0.0)
    L.backward()
    assert loss[0] == pytest.approx(float(lp), rel=2e-2)
    gl = torch.from_numpy(g[-(10 * 256 + 10):]).double()
    gr = torch.cat([params[-2].grad.reshape(-1), params[-1].grad.reshape(-1)]).double()
    cos = float(gl @ gr / (gl.norm() * gr.norm()))
    say("capped loss lib %.4g ref %.4g, output-layer grad cos %.5f" % (loss[0], float(lp), cos))
    assert cos > 0.99
    mlp.close()


@pytest.mark.parametrize("act", ACTS)
def test_training_fits_a_known_yield_function(ctx, act):
    """Poisson rows with log-yield f(x) = -9 + 2 x0 - x1²: the Poisson NLL must
    approach its floor and the fitted log yield f within 0.1."""
    rng = np.random.default_rng(5)
    M = 65536
    x = rng.uniform(-1, 1, (M, 8)).astype(np.float32)
    f = -9.0 + 2.0 * x[:, 0] - x[:, 1] ** 2
    logn = np.full(M, math.log(1e6), dtype=np.float32)
    k = rng.poisson(np.exp(f + logn)).astype(np.float32)
    times = -np.ones((M, 4), dtype=np.float32)
    ts = sg.TrainingSet(raw=x, x=x, log_n=logn, k=k, times=times)
    mlp = sg.MLP(ctx, seed=4, activation=act)
    hist = sg.train(mlp, ts, epochs=6, batch=2048, lr=3e-3, time_weight=0.0, log=say)
    assert hist[-1][0] < hist[0][0]
    xt = rng.uniform(-0.9, 0.9, (4096, 8)).astype(np.float32)
    ft = -9.0 + 2.0 * xt[:, 0] - xt[:, 1] ** 2
    y = mlp.forward(xt)
    err = np.median(np.abs(y[:, 0] - ft))
    say("%s median |log-yield error| %.4f" % (act, err))
    assert err < 0.1
    mlp.close()


def test_mc_rows_train_and_hybrid_fallback():
    """A small MC training set from the batched kernel, a short training run,
    and the MC fallback for an out-of-domain query."""
    with pp.Context(pp.SeawaterConfigMedium()) as c:
        ts = sg.generate_training_set(c, 16384, photons_per_row=20_000, rows_per_medium=4096,
                                      n_times=8, seed=1)
        assert ts.x.shape == (16384, 8) and np.all(np.abs(ts.x) <= 1.0 + 1e-6)
        assert (ts.times >= 0).any() and ts.k.sum() > 0
        mlp = sg.MLP(c, seed=0)
        hist = sg.train(mlp, ts, epochs=3, batch=2048, lr=2e-3, log=say)
        assert hist[-1][0] < hist[0][0]
        hy = sg.HybridYield(mlp, c, mc_photons=100_000)
        rng = np.random.default_rng(9)
        ms = np.array([30.0, 50.0, 0.9, 0.0])
        raw, reqs = sg.random_rows(rng, 8, ms)
        raw[0, 0] = 3.0  # 1 km: outside the training domain
        reqs["seed"] = 1
        y, used, _ = hy.query(raw, reqs, sg.medium_for(*ms))
        assert used[0] and np.all(np.isfinite(y)) and np.all(y >= 0)
        mlp.close()


def test_ensemble_gate_against_held_out_mc():
    """The surrogate's acceptance on held-out pairs, at a reduced training size.
    Four networks trained on bootstrap resamples of 5e5 MC rows whose photon
    counts come from a pilot network's yield estimate ŷ (clip(20/ŷ, 1e5, 1e6)
    photons per row); 1024 held-out pairs in 4 media against MC of
    clip(1000/ŷ, 1e8, 2e9) photons each.  The gate (out of domain, yield below
    1e-8, or spread of the members' log yields above 0.1) must keep the
    surrogate's own answers at a median yield error <= 5 %, send to MC pairs
    whose median error is larger than that of the pairs it keeps, fall back on
    at most 60 % of the pairs (about a third at the full 2e6 rows), and leave
    the hybrid's answers at a median error <= 5 %."""
    with pp.Context(pp.SeawaterConfigMedium()) as c:
        pilot = sg.pilot_allocator(c, seed=11)
        ts = sg.generate_training_set(c, 500_000, n_times=16, seed=3, allocate=pilot)
        ens, hists = sg.train_ensemble(c, ts, k=4, epochs=40, seed=0)
        del ts
        assert all(h[-1][0] < h[0][0] for h in hists)
        truth = sg.YieldAllocator(pilot.model, target_hits=1000.0, n_min=100_000_000,
                                  n_max=2_000_000_000)
        pairs = sg.held_out_pairs(c, 1024, None, seed=1234, allocate=truth)
        pilot.model.close()
        hy = sg.HybridYield(ens, c)
        a, s = sg.evaluate_hybrid(hy, pairs)
        say(s)
        assert s["pairs_resolved"] >= 500
        assert s["kept_resolved_pairs"] >= 250
        assert s["kept_median_rel_error"] <= 0.05
        assert s["gated_resolved_median_rel_error"] > s["kept_median_rel_error"]
        assert s["fallback_fraction"] <= 0.6
        assert s["hybrid_median_rel_error_resolved"] <= 0.05
        # a wider gate keeps a superset of the pairs
        hy.max_log_std = 0.2
        _, _, wide = hy.predict(np.concatenate([p["raw"] for p in pairs]))
        assert np.all(a["gated"] >= wide)
        ens.close()


def test_training_is_reproducible(ctx):
    """Weight gradients are split-K partials added in a fixed order (no
    atomics), so two trainings from the same seed on the same rows in the same
    order end with bitwise equal parameters."""
    rng = np.random.default_rng(8)
    M = 16384
    x = rng.uniform(-1, 1, (M, 8)).astype(np.float32)
    logn = np.full(M, math.log(1e5), dtype=np.float32)
    k = rng.poisson(np.exp(-6.0 + x[:, 0] + logn)).astype(np.float32)
    times = rng.gamma(2.0, 10.0, (M, 4)).astype(np.float32)
    ts = sg.TrainingSet(raw=x, x=x, log_n=logn, k=k, times=times)
    params = []
    for _ in range(2):
        mlp = sg.MLP(ctx, seed=12)
        sg.train(mlp, ts, epochs=2, batch=4096, lr=3e-3, time_weight=0.5, seed=5)
        params.append(mlp.get_params())
        mlp.close()
    assert np.array_equal(params[0], params[1])


def test_train_steps_equal_single_steps(ctx):
    """pp_mlp_train_steps_device (a run of minibatch steps with no host
    synchronisation between them) leaves the network bit for bit where the same
    steps taken one call at a time leave it, with the same mean loss."""
    rng = np.random.default_rng(13)
    B, S, T = 1024, 5, 4
    x, logn, k, times = batch(rng, B * S, T)
    dx, dln, dk, dt = to_dev(x, logn, k, times)
    lrs = np.float32(3e-3) * np.linspace(1.0, 0.5, S, dtype=np.float32)
    one = sg.MLP(ctx, seed=21)
    many = sg.MLP(ctx, seed=21)
    losses = []
    for i in range(S):
        o = i * B
        losses.append(one.train_step(dx.data_ptr() + 4 * o * 8, dln.data_ptr() + 4 * o,
                                     dk.data_ptr() + 4 * o, dt.data_ptr() + 4 * o * T, T, B,
                                     float(lrs[i]), 0.5))
    mean = many.train_steps(dx.data_ptr(), dln.data_ptr(), dk.data_ptr(), dt.data_ptr(), T, B,
                            lrs, 0.5)
    assert np.array_equal(one.get_params(), many.get_params())
    np.testing.assert_allclose(mean, np.mean(losses, axis=0), rtol=1e-12)
    with pytest.raises(pp.PhotonPropError):
        many.train_steps(dx.data_ptr(), dln.data_ptr(), dk.data_ptr(), dt.data_ptr(), T, B,
                         np.zeros(0, np.float32), 0.5)
    one.close()
    many.close()
