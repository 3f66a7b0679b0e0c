"""The adjoint of the bounces (lightbaker_amd/csrc/adjoint.hip) on the GPU.

Pins, on the 16-material fitting scene (3 bounces):
  * the gather rays' records equal the oracle's (oracle/trace_ref.c orc_gather_records), bit for
    bit;
  * a forward with the scene's own parameters reproduces the PLAIN bake's light atlas bit for bit;
  * the gradient of an L2 loss against a target baked from perturbed albedos (+-0.1), sky
    gradient colours and the sky's L1 coefficients equals the FP64 forward-mode gradient
    (oracle/adjoint_ref.py) within 1e-3 relative, and central finite differences (h = 1e-3) of
    the GPU's own forward within 1e-3 relative, every parameter (3 M + 24) included;
  * the torch.autograd.Function gives the same gradient, and 50 Adam steps fit the target, the
    sky's L1 coefficients among the fitted parameters.
The backward accumulates in FP32 with compensated (two-sum) atomics and reductions, no FP64.
Tolerance: per parameter, |g - g_ref| <= 1e-3 * max(|g_ref|, 1e-2 * max |g_ref|) (parameters
whose gradient is under 1 % of the largest are held to 1e-5 of the largest).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle.adjoint_ref as aref
import oracle.bake_ref as bref
from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession
from lightbaker_amd.scenes import fit_scene

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-3


def session_inputs(scene, sess):
    P = sess.array(N.BAKE_PARAMS, N.PARAMS)
    return {
        "settings": scene.settings, "texels": sess.array(N.BAKE_TEXELS, N.TEXEL),
        "tiles": sess.array(N.BAKE_TILES, np.int32).reshape(-1, 6),
        "aw": int(P["atlas"][0][0]), "ah": int(P["atlas"][0][1]),
        "nodes": sess.array(N.BAKE_NODES, N.NODE), "tris": sess.array(N.BAKE_TRIS, N.TRI),
        "mats": sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4),
        "charts": sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8), "params": P,
        "sun": P["sun_dir"][0][:3].copy(), "irr": np.asarray(scene.sun_irradiance, np.float32),
        "probes": None, "volume_params": None, "sky_sh": scene.sky_sh,
    }


def perturbed(scene, seed=55):
    """The scene's own light parameters, and a target's: albedo +-0.1, the sky's gradient
    colours and L1 coefficients scaled by 0.8..1.2."""
    from lightbaker_amd.adjoint import scene_light_params
    alb, sky, sun, sh = scene_light_params(scene)
    assert np.abs(sh).max() > 0
    rng = np.random.default_rng(seed)
    alb_t = np.clip(alb + rng.uniform(-0.1, 0.1, alb.shape), 0.05, 0.95).astype(np.float32)
    sky_t = (sky * rng.uniform(0.8, 1.2, sky.shape)).astype(np.float32)
    sh_t = (sh * rng.uniform(0.8, 1.2, sh.shape)).astype(np.float32)
    return (alb, sky, sun, sh), (alb_t, sky_t, sun, sh_t)


def check_close(g, ref, what):
    g, ref = np.asarray(g, np.float64), np.asarray(ref, np.float64)
    big = np.abs(ref).max()
    bound = TOL * np.maximum(np.abs(ref), 1e-2 * big)
    bad = np.nonzero(np.abs(g - ref) > bound)[0]
    worst = float((np.abs(g - ref) / np.maximum(np.abs(ref), 1e-2 * big)).max())
    print(f"{what}: worst relative error {worst:.2e} over {len(ref)} parameters", flush=True)
    assert len(bad) == 0, f"{what}: parameters {bad[:8]} differ: {g[bad[:8]]} vs {ref[bad[:8]]}"


@pytest.fixture(scope="module")
def fitted(gpu):
    from lightbaker_amd.adjoint import Adjoint
    scene = fit_scene(5)
    print("adjoint: PLAIN bake of the fitting scene", flush=True)
    sess = BakeSession(scene, gpu, form=N.FORM_PLAIN, graph=False)
    sess.execute()
    adj = Adjoint(sess)
    sess.finish()
    print(f"adjoint: {adj.items} gather items x {adj.rays} rays, atlas {adj.shape}", flush=True)
    inp = session_inputs(scene, sess)
    out = bref.bake(inp)
    S = aref.setup(inp, out)
    yield scene, sess, adj, inp, out, S
    adj.close()
    sess.close()


def test_records_match_oracle(fitted):
    scene, sess, adj, inp, out, S = fitted
    got = adj.records().transpose(1, 0, 2)
    assert got.shape == S["rec"].shape
    bad = np.nonzero((got != S["rec"]).any(axis=2))
    assert len(bad[0]) == 0, f"{len(bad[0])} records differ, first item {bad[0][:3]} ray {bad[1][:3]}"
    kinds = np.bincount((got[:, :, 1] >> 30).reshape(-1), minlength=4)
    print(f"records: chart hits {kinds[0]}, off-chart {kinds[1]}, back {kinds[2]}, sky {kinds[3]}")
    assert kinds[0] > 0 and kinds[1] > 0 and kinds[2] > 0 and kinds[3] > 0


def test_replay_reproduces_the_bake(fitted):
    from lightbaker_amd.adjoint import scene_light_params
    scene, sess, adj, *_ = fitted
    light = adj.forward(*scene_light_params(scene)).cpu().numpy().reshape(-1, 4)
    want = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    bad = np.nonzero((light.view(np.uint32) != want.view(np.uint32)).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} atlas texels differ: {light[bad[:3]]} vs {want[bad[:3]]}"


def loss_and_dlight(light, target):
    w = (target[:, 3:] > 0.5).astype(np.float64)
    d = (light[:, :3].astype(np.float64) - target[:, :3]) * w
    dl = np.zeros((len(light), 4), np.float64)
    dl[:, :3] = 2.0 * d
    return float((d * d).sum()), dl


def test_gradient_matches_fp64(fitted):
    scene, sess, adj, inp, out, S = fitted
    (alb, sky, sun, sh), target_params = perturbed(scene)
    target = adj.forward(*target_params).cpu().numpy().reshape(-1, 4).astype(np.float64)
    light = adj.forward(alb, sky, sun, sh).cpu().numpy().reshape(-1, 4)
    loss, dl = loss_and_dlight(light, target)
    g = adj.backward(torch.tensor(dl, dtype=torch.float32).reshape(adj.shape)).cpu().numpy()
    assert len(g) == 3 * len(alb) + 24
    Lt = aref.forward(S, alb, sky[0], sky[1], sky[2], sun, sh)
    light64 = aref.atlas_light(S, Lt)
    loss64, dl64 = loss_and_dlight(light64, target)
    g64 = aref.gradient(S, Lt, dl64)
    print(f"loss {loss:.6g} (FP64 {loss64:.6g})")
    assert abs(loss - loss64) <= 1e-4 * loss64
    check_close(g, g64, "adjoint vs FP64 forward-mode gradient")


def test_gradient_matches_finite_differences(fitted):
    scene, sess, adj, *_ = fitted
    (alb, sky, sun, sh), target_params = perturbed(scene)
    target = adj.forward(*target_params).cpu().numpy().reshape(-1, 4).astype(np.float64)
    light = adj.forward(alb, sky, sun, sh).cpu().numpy().reshape(-1, 4)
    _, dl = loss_and_dlight(light, target)
    g = adj.backward(torch.tensor(dl, dtype=torch.float32).reshape(adj.shape)).cpu().numpy()
    theta = np.concatenate([alb.reshape(-1), sky.reshape(-1), sh.reshape(-1), sun]).astype(np.float64)
    M = len(alb)
    h = 1e-3

    def loss_at(th):
        a = th[:M * 3].reshape(M, 3).astype(np.float32)
        s = th[M * 3:M * 3 + 9].reshape(3, 3).astype(np.float32)
        c = th[M * 3 + 9:M * 3 + 21].reshape(4, 3).astype(np.float32)
        u = th[M * 3 + 21:].astype(np.float32)
        lt = adj.forward(a, s, u, c).cpu().numpy().reshape(-1, 4)
        return loss_and_dlight(lt, target)[0]

    fd = np.zeros_like(theta)
    for p in range(len(theta)):
        e = np.zeros_like(theta)
        e[p] = h
        fd[p] = (loss_at(theta + e) - loss_at(theta - e)) / (2 * h)
    check_close(g, fd, "adjoint vs central differences (h = 1e-3)")


def test_autograd_function(fitted):
    from lightbaker_amd.adjoint import baked_light, l2_loss
    scene, sess, adj, *_ = fitted
    (alb, sky, sun, sh), target_params = perturbed(scene)
    target = adj.forward(*target_params)
    a = torch.tensor(alb, requires_grad=True)
    s = torch.tensor(sky, requires_grad=True)
    u = torch.tensor(sun, requires_grad=True)
    c = torch.tensor(sh, requires_grad=True)
    loss = l2_loss(baked_light(adj, a, s, u, c), target)
    loss.backward()
    light = adj.forward(alb, sky, sun, sh)
    w = (target[..., 3:] > 0.5).float()
    dl = torch.zeros_like(light)
    dl[..., :3] = 2.0 * (light[..., :3] - target[..., :3]) * w
    g = adj.backward(dl).cpu().numpy()
    M = len(alb)
    got = np.concatenate([a.grad.numpy().reshape(-1), s.grad.numpy().reshape(-1),
                          c.grad.numpy().reshape(-1), u.grad.numpy()])
    assert np.allclose(got, g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    assert a.grad.shape == (M, 3) and s.grad.shape == (3, 3) and u.grad.shape == (3,)
    assert c.grad.shape == (4, 3)


def test_adam_fits_the_target(fitted):
    from lightbaker_amd.adjoint import fit
    scene, sess, adj, *_ = fitted
    (alb, sky, sun, sh), (alb_t, sky_t, _, sh_t) = perturbed(scene)
    target = adj.forward(alb_t, sky_t, sun, sh_t)
    res = fit(adj, target, alb, sky, sun, sh, steps=50, lr=0.02, fit_sky="both")
    first, last = res["loss"][0], res["loss"][-1]
    print(f"Adam: loss {first:.4g} -> {last:.4g}; albedo error "
          f"{np.abs(alb - alb_t).mean():.4f} -> {np.abs(res['albedo'] - alb_t).mean():.4f}; "
          f"L1 sky error {np.abs(sh - sh_t).mean():.4f} -> "
          f"{np.abs(res['sky_sh'] - sh_t).mean():.4f}")
    assert last < 0.1 * first
    assert np.abs(res["sky"] - sky_t).mean() < np.abs(sky - sky_t).mean()


def test_adam_recovers_the_l1_sky(fitted):
    """The gradient sky (a constant and a term linear in y above the horizon) and the L1 term
    overlap, so a fit of both recovers the light, not each parameter. With the gradient sky
    held at the target's, the L1 coefficients are determined: Adam on albedo and L1 sky brings
    the coefficients towards the target's."""
    from lightbaker_amd.adjoint import fit
    scene, sess, adj, *_ = fitted
    (alb, sky, sun, sh), (alb_t, _, _, sh_t) = perturbed(scene)
    target = adj.forward(alb_t, sky, sun, sh_t)
    res = fit(adj, target, alb, sky, sun, sh, steps=60, lr=0.01, fit_sky="sh")
    first, last = res["loss"][0], res["loss"][-1]
    e0, e1 = np.abs(sh - sh_t).mean(), np.abs(res["sky_sh"] - sh_t).mean()
    print(f"Adam (L1 sky): loss {first:.4g} -> {last:.4g}; L1 sky error {e0:.4f} -> {e1:.4f}")
    assert last < 0.1 * first
    assert e1 < 0.5 * e0
