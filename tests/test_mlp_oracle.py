"""Pin the learned-bound network's CPU oracle (oracle/mlp_ref.py) by finite differences."""
import numpy as np

from ebgeometry_amd import workloads as wl
from oracle import mlp_ref


def random_theta(seed, scale=0.2):
    s = wl.Stream(seed, "theta")
    return (s.uniform(mlp_ref.NPARAMS) * 2 - 1) * scale


def test_param_count():
    assert mlp_ref.NPARAMS == 56321


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-5, 65504.0], np.float32)
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125  # tie to even; above half rounds up
    assert abs(r[3] - (-3.0e-5)) < 1e-7 and r[4] == 65536.0


def test_gradient_matches_finite_differences():
    s = wl.Stream(3, "fd")
    pts = (s.uniform(24 * 3).reshape(24, 3) * 2 - 1).astype(np.float32)
    target = s.uniform(24) * 0.5
    theta = random_theta(4)
    lo, scale, diag = np.array([-1.2] * 3), np.array([1 / 1.2] * 3), 3.0
    L, g = mlp_ref.loss_and_grad(theta, pts, target, lo, scale, diag)
    idx = np.r_[0:5, 6520:6525, 6600:6605, 30000:30005, 56190:56196, 56300:56321]
    h = 1e-6
    for i in idx:
        tp = theta.copy()
        tp[i] += h
        tm = theta.copy()
        tm[i] -= h
        fd = (mlp_ref.loss_and_grad(tp, pts, target, lo, scale, diag)[0]
              - mlp_ref.loss_and_grad(tm, pts, target, lo, scale, diag)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-7, (i, fd, g[i])


def test_loss_is_asymmetric():
    theta = np.zeros(mlp_ref.NPARAMS)
    pts = np.zeros((2, 3), np.float32)
    lo, scale = np.array([-1.0] * 3), np.array([1.0] * 3)
    t = np.array([0.1, 0.1])
    lt = np.log(0.1 + mlp_ref.LOSS_C)
    # y = b5 only; an under-prediction costs under_weight times the same over-prediction
    theta[-1] = lt - 0.3
    under, _ = mlp_ref.loss_and_grad(theta, pts, t, lo, scale, 1.0)
    theta[-1] = lt + 0.3
    over, _ = mlp_ref.loss_and_grad(theta, pts, t, lo, scale, 1.0)
    assert np.isclose(over, 0.09) and np.isclose(under, mlp_ref.LOSS_UNDER * 0.09)
    # the prediction inverts the log map
    assert np.allclose(mlp_ref.predict(theta, pts, lo, scale, 2.0), (np.exp(lt + 0.3) - mlp_ref.LOSS_C) * 2.0)


def test_training_schedule_lowers_the_loss():
    """The trainer's schedule (Adam, cosine decay, BF16 rounding points) at the learned-bound test's
    batch size, with its partial last batch (1,400 pairs in batches of 512: 2 full + 376 rows),
    lowers the loss on the unit sphere's exact distances |‖p‖ - 1|."""
    s = wl.Stream(11, "train")
    pts = ((s.uniform(1400 * 3).reshape(1400, 3) * 2 - 1) * 1.5).astype(np.float32)
    target = np.abs(np.linalg.norm(pts.astype(np.float64), axis=1) - 1.0)
    lo, scale, diag = np.array([-1.2 * 1.2] * 3), np.array([1 / (1.2 * 1.2)] * 3), 2.0 * np.sqrt(3.0)
    s2 = wl.Stream(12, "theta-train")
    parts = []
    for shape in mlp_ref.SHAPES:
        fan_in = shape[0] if len(shape) == 2 else 128
        parts.append(((s2.uniform(int(np.prod(shape))) * 2 - 1) * np.sqrt(6.0 / fan_in)).reshape(shape))
    for i in (1, 3, 5, 7):
        parts[i][:] = 0.0
    parts[8] *= 0.01
    parts[9][:] = np.log(0.05)
    theta0 = mlp_ref.pack(parts)
    L0, _ = mlp_ref.loss_and_grad(theta0, pts, target, lo, scale, diag, bf16_mode=True)
    theta, losses = mlp_ref.adam_train(theta0, pts, target, lo, scale, diag, batch=512, epochs=12)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.2 * L0, (L0, losses)
    assert losses[1] < 0.5 * losses[0] and max(losses[6:]) < min(losses[:3]), losses
    # and the trained network's prediction follows |d| (36 Adam steps: median error < 0.6 of
    # |d| + 0.01 diag, measured 0.48; the initial network is uncorrelated)
    pred = mlp_ref.predict(theta, pts, lo, scale, diag, bf16_mode=True)
    rel = np.abs(pred - target) / (target + 0.01 * diag)
    assert np.median(rel) < 0.6, np.median(rel)
    assert np.corrcoef(pred, target)[0, 1] > 0.25
