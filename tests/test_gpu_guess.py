"""The learned CG initial guess on the GPU: MLP kernels against the float64 oracle, and the guess in MD (GPU).

Bounds: the FP32 training forward within 1e-5 relative RMS of the oracle, its gradients within
1e-4 of the largest gradient of each layer, one Adam step within 1e-6; the BF16 matrix-core
inference within 2e-2 relative RMS (8-bit mantissas of inputs, weights and activations).
"""

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _ff(nw=512):
    import pgm_jax_amd as P

    sys_, pos, H = P.models.water_box(nw, 2.483 * (nw / 512) ** (1 / 3), seed=5)
    ff = P.PGMForceField(sys_, H, P.MDSettings(precision="mixed", predictor="learned", pme_grid=(32, 32, 32),
                                               dipole_tol=1e-7, max_iter=300))
    return P, ff, sys_, pos, H


def _to_oracle(flat, sizes):
    params, off = [], 0
    for a, b in zip(sizes[:-1], sizes[1:]):
        W = flat[off:off + a * b].reshape(a, b).astype(np.float64)
        off += a * b
        params.append((W, flat[off:off + b].astype(np.float64)))
        off += b
    return params


SIZES = [16, 128, 128, 128, 3]


def _random_params(ff, seed=1):
    rng = np.random.default_rng(seed)
    flat = ff.guess_params()
    flat = flat + rng.normal(scale=0.05, size=flat.shape).astype(np.float32)  # nonzero biases too
    ff.guess_params(flat)
    return flat


def test_inference_and_training_forward_match_oracle():
    P, ff, sys_, pos, H = _ff()
    ff.guess_config(hidden_layers=3, seed=7)
    print("mlp forward", flush=True)
    flat = _random_params(ff)
    params = _to_oracle(flat, SIZES)
    x = np.random.default_rng(2).normal(size=(1000, 16)).astype(np.float32)
    ref, _ = mlp.forward(params, x.astype(np.float64))
    y32 = ff.guess_forward(x, bf16=False).cpu().numpy()
    ybf = ff.guess_forward(x, bf16=True).cpu().numpy()
    rel = lambda a: float(np.sqrt(np.mean((a - ref) ** 2) / np.mean(ref**2)))  # noqa: E731
    print("fp32", rel(y32), "bf16", rel(ybf), flush=True)
    assert rel(y32) < 1e-5
    assert rel(ybf) < 2e-2


def test_gradients_and_adam_match_oracle():
    P, ff, sys_, pos, H = _ff()
    ff.guess_config(hidden_layers=3, seed=8)
    flat = _random_params(ff, seed=3)
    params = _to_oracle(flat, SIZES)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(777, 16)).astype(np.float32)
    t = rng.normal(size=(777, 3)).astype(np.float32)
    g, loss = ff.guess_grad(x, t)
    lref, gref = mlp.loss_grad(params, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lref) < 1e-5 * lref
    gflat = _to_oracle(g, SIZES)
    for (gW, gb), (rW, rb) in zip(gflat, gref):
        assert np.abs(gW - rW).max() < 1e-4 * np.abs(rW).max()
        assert np.abs(gb - rb).max() < 1e-4 * max(np.abs(rb).max(), 1e-12)


def test_learned_guess_in_md_converges_and_selects():
    """MD with predictor "learned": the ring fills, the linear predictor is fitted at each harvest (every 5 solves),
    the network trains on its errors; the selector picks the better guess; every solve reaches the tolerance."""
    P, ff, sys_, pos, H = _ff()
    ff.guess_config(hidden_layers=3, harvest_every=5, seed=9)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    md = P.RigidWaterMD(ff, pos, vel)
    r0 = md.run(60, 1.0)
    st = ff.guess_status()
    print(st, r0, flush=True)
    assert st["fits"] >= 5 and st["sigma"] > 0 and st["rows"] > 0 and not st["trained"]
    assert abs(sum(st["coefficients"]) - 1.0) < 1e-3  # a consistent extrapolation reproduces constant dipoles
    loss0 = float(np.mean(ff.guess_status()["err_linear"])) / st["sigma"] ** 2
    loss = ff.guess_train(steps=200, batch=2048, lr=1e-3)
    print("loss", loss, "untrained", loss0, flush=True)
    assert np.isfinite(loss)
    r1 = md.run(40, 1.0)
    st = ff.guess_status()
    print(st, r1, flush=True)
    assert st["trained"] and np.isfinite(st["err_net"])
    assert r1["cg_residual_max"] <= 1e-7 and not r1["not_converged"]
    if st["mode"] != "mu4":  # the fitted predictor was chosen: it must beat mu4 on this harvest
        assert st["err_linear"] < 0.8 * st["err_mu4"]


def test_linear_predictor_beats_mu4_in_cg_iterations():
    """The same 512-water trajectory from the same state: mean CG iterations with the learned guess after the fit
    vs mu4 (the selector keeps mu4 when the fit does not help, so never worse by more than one iteration)."""
    import pgm_jax_amd as P

    sys_, pos, H = P.models.water_box(512, 2.483, seed=5)
    vel = P.models.maxwell_boltzmann_rigid(pos, sys_.masses, 298.0, seed=6)
    kw = dict(precision="mixed", pme_grid=(32, 32, 32), dipole_tol=1e-7, max_iter=300)
    ffl = P.PGMForceField(sys_, H, P.MDSettings(predictor="learned", **kw))
    ffl.guess_config(hidden_layers=3, harvest_every=5, seed=9)
    mdl = P.RigidWaterMD(ffl, pos, vel)
    mdl.run(80, 1.0)
    x, v = mdl.pos.clone(), mdl.vel.clone()
    ffm = P.PGMForceField(sys_, H, P.MDSettings(predictor="mu4", **kw))
    mdm = P.RigidWaterMD(ffm, x, v)
    mdm.run(10, 1.0)
    it_m = mdm.run(50, 1.0)["cg_iterations"] / 50
    it_l = mdl.run(50, 1.0)["cg_iterations"] / 50
    print("mu4", it_m, "learned", it_l, ffl.guess_status(), flush=True)
    assert it_l <= it_m + 1.0
