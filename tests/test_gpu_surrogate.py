"""chi_B surrogate: the device MLP kernels against the FP64 oracle, the pilot-only entry point,
and the surrogate-or-pilot decision inside qa_sqa_chi_run.

The MLP products read bf16 copies of the weights with FP32 activations and FP32 accumulation, so
the oracle (oracle/mlp64.py, FP64) is evaluated at the same bf16-rounded weights and FP32 biases;
what remains is FP32 rounding of the accumulations and of tanh/sigmoid, bounded here by 2e-5 on
outputs in (0, 1) and 1e-4 of the largest gradient entry.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from qlatanneal_amd.abi import Context
    return Context(0)


def _bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _as_evaluated(params, dims):
    """One member's parameters as the kernels use them: bf16 weights, FP32 biases."""
    from oracle import mlp64
    p = np.asarray(params, np.float32).copy()
    off = 0
    for din, dout in mlp64.layer_dims(*dims):
        p[off:off + din * dout] = _bf16(p[off:off + din * dout])
        off += din * dout + dout
    return p.astype(np.float64)


@pytest.mark.parametrize("dims,members,batch", [((12, 256, 3, 64), 4, 37), ((5, 70, 2, 9), 3, 130)])
def test_mlp_forward_and_gradient_match_fp64_oracle(ctx, dims, members, batch):
    from qlatanneal_amd import abi
    from oracle import mlp64
    rng = np.random.default_rng(3)
    m = abi.Mlp(ctx, *dims, members=members, seed=11)
    params, norm = m.get_state()
    # non-trivial biases and normalization so that every term is exercised
    params = params + rng.normal(0, 0.05, params.shape).astype(np.float32)
    norm[0] = rng.normal(size=dims[0])
    norm[1] = rng.uniform(0.5, 2.0, dims[0])
    m.set_state(params, norm)
    x = rng.normal(size=(batch, dims[0]))
    t = rng.random((batch, dims[3])).astype(np.float32)
    y = m.forward(x)
    loss, grad = m.loss_grad(x, t)
    for z in range(members):
        pz = _as_evaluated(params[z], dims)
        ref = mlp64.forward(pz, x, *dims, mean=norm[0], scale=norm[1])[-1]
        assert np.max(np.abs(y[z] - ref)) < 2e-5, z
        l_ref, g_ref = mlp64.loss_grad(pz, x, t, *dims, mean=norm[0], scale=norm[1])
        assert abs(loss[z] - l_ref) <= 1e-5 * l_ref, z
        err = np.max(np.abs(grad[z] - g_ref))
        print("member", z, "max grad err", err, "max |grad|", np.max(np.abs(g_ref)), flush=True)
        assert err <= 1e-4 * np.max(np.abs(g_ref)), z


def test_mlp_training_fits_a_smooth_map(ctx):
    from qlatanneal_amd import abi
    rng = np.random.default_rng(4)
    A = rng.normal(0, 0.6, (12, 16))
    x = rng.normal(1.0, 2.0, (512, 12))
    t = (1.0 / (1.0 + np.exp(-((x - 1.0) / 2.0) @ A))).astype(np.float32)
    m = abi.Mlp(ctx, 12, 64, 2, 16, members=2, seed=1)
    assert not m.trained
    hist = m.train(x, t, epochs=120, batch=32, lr=3e-3, seed=2)
    print("loss first/last epoch", hist[0], hist[-1], flush=True)
    assert m.trained
    assert hist[-1] < 0.05 * hist[0]
    _, norm = m.get_state()
    assert np.allclose(norm[0], x.mean(axis=0)) and np.allclose(norm[1], x.std(axis=0))
    y = m.forward(x[:64])
    assert np.mean((y.mean(axis=0) - t[:64]) ** 2) < 2 * hist[-1] + 1e-4
    assert not np.array_equal(y[0], y[1])  # members start from different weights


def test_chi_pilot_equals_run_chi_scan(ctx):
    from qlatanneal_amd import abi, surrogate
    rng = np.random.default_rng(5)
    prob = surrogate.sk_problem(ctx, rng, 64)
    kw = dict(scan_points=8, scan_sweeps=6, scan_burn=2, driver_A0=3.0, seed=77)
    pil = abi.chi_pilot(prob, 8, 2, 4.0, 3.0, 1e-3, **kw)
    run = abi.sqa_chi_run(prob, 8, 2, 4.0, 3.0, 1e-3, num_steps=10, sweeps_per_step=2, **kw)
    assert np.array_equal(pil["scan_chi_B"], run["scan_chi_B"])
    assert np.array_equal(pil["scan_s"], run["scan_s"]) and np.array_equal(pil["scan_gamma"], run["scan_gamma"])
    assert not run["surrogate_used"] and np.isnan(run["surrogate_spread"])


def _small_cfg():
    from qlatanneal_amd.surrogate import PilotConfig
    return PilotConfig(trotter_slices=8, replicas=2, scan_points=16, scan_sweeps=10, scan_burn=4)


def test_surrogate_use_and_fallback(ctx):
    from qlatanneal_amd import abi, surrogate
    cfg = _small_cfg()
    rng = np.random.default_rng(6)
    specs = surrogate.random_specs(24, rng, sk_n=(48, 128), maxcut_n=(48, 160), beta=(2.0, 8.0))
    X, Y, _ = surrogate.make_dataset(ctx, specs, cfg)
    assert X.shape[0] >= 20
    sur = surrogate.ChiSurrogate(ctx, cfg, width=64, depth=3, members=3, seed=3)
    run_kw = dict(num_steps=12, sweeps_per_step=2, scan_points=cfg.scan_points, scan_sweeps=cfg.scan_sweeps,
                  scan_burn=cfg.scan_burn, driver_A0=cfg.driver_A0, seed=9)
    prob = surrogate.build(ctx, specs[0])
    beta = specs[0].beta
    args = (prob, cfg.trotter_slices, cfg.replicas, beta, cfg.gamma_start, cfg.gamma_end)
    with pytest.raises(abi.QaInvalidArgument):  # untrained
        abi.sqa_chi_run(*args, surrogate=sur.mlp, surrogate_max_spread=1e9, **run_kw)
    sur.fit(X, Y, epochs=150, batch=8)
    # trusted: the profile is the ensemble mean of the network's prediction for these features
    r = abi.sqa_chi_run(*args, surrogate=sur.mlp, surrogate_max_spread=1e9, **run_kw)
    assert r["surrogate_used"]
    f = surrogate.features(prob, beta, cfg)
    y = sur.mlp.forward(f[None, :])[:, 0, :].astype(np.float64)
    assert np.array_equal(r["scan_chi_B"], y.mean(axis=0))
    assert r["surrogate_spread"] == pytest.approx(np.max(y.std(axis=0)), rel=1e-12)
    assert r["best_energy"] == prob.energy(r["best_state"])[0]
    # spread above the bound: the pilot runs and the run equals a run without a surrogate
    fb = abi.sqa_chi_run(*args, surrogate=sur.mlp, surrogate_max_spread=0.0, **run_kw)
    plain = abi.sqa_chi_run(*args, **run_kw)
    assert not fb["surrogate_used"] and fb["surrogate_spread"] > 0.0
    for k in ("scan_chi_B", "gamma_schedule", "energy_trace", "best_state"):
        assert np.array_equal(fb[k], plain[k]), k
    # outside the training range (beta far above it): fallback with infinite spread
    ood = abi.sqa_chi_run(prob, cfg.trotter_slices, cfg.replicas, 500.0, cfg.gamma_start, cfg.gamma_end,
                          surrogate=sur.mlp, surrogate_max_spread=1e9, **run_kw)
    assert not ood["surrogate_used"] and np.isinf(ood["surrogate_spread"])
    # save / load round trip keeps predictions bit for bit
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "sur.npz")
        sur.save(path)
        back = surrogate.ChiSurrogate.load(ctx, path)
        assert np.array_equal(back.mlp.forward(X[:5]), sur.mlp.forward(X[:5]))


def test_dataset_threads_equal_sequential_and_chunked_training(ctx, tmp_path):
    """make_dataset on several contexts at once yields the sequential dataset exactly (same scans,
    same order); `generate` chunks feed `train --from-data`."""
    from qlatanneal_amd import surrogate as sg
    specs = sg.random_specs(6, np.random.default_rng(77), sk_n=(64, 160), maxcut_n=(64, 160))
    cfg = sg.PilotConfig()
    X1, Y1, _ = sg.make_dataset(ctx, specs, cfg, threads=1)
    X3, Y3, _ = sg.make_dataset(ctx, specs, cfg, threads=3)
    assert X1.shape == X3.shape and np.array_equal(X1, X3) and np.array_equal(Y1, Y3)
    a, b, out = tmp_path / "a.npz", tmp_path / "b.npz", tmp_path / "m.npz"
    np.savez(a, X=X1[:3], Y=Y1[:3])
    np.savez(b, X=X1[3:], Y=Y1[3:])
    assert sg._main(["train", "--from-data", str(a), str(b), "--out", str(out), "--epochs", "2",
                     "--batch", "4"]) == 0
    model = sg.ChiSurrogate.load(ctx, str(out))
    pred, spread = model.predict(X1)
    assert pred.shape == Y1.shape and np.all(np.isfinite(pred)) and np.all(spread >= 0)
