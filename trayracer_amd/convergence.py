"""Time to reach a relative MSE target, plain accumulation vs learned denoising.

relMSE(x, r) = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2), with r a
high-spp render of the same view. The reference is rendered as two independent halves
(different seeds); their mean is r, and a quarter of the halves' relMSE to each other
estimates r's own error against the converged image (each half has twice r's
variance), which is reported beside every result and subtracted from the measured
relMSE for the `*_true` figures. Wall times are measured on the GPU with the scene
resident: rendering k spp (plus, for the denoised path, one denoiser inference). The
denoiser is trained first, on pairs (low-spp input, 1024-spp target) rendered by the
tracer from the benchmark scenes under seeded random cameras; inputs use 1 to 64 spp so
the network serves every sample count of the progressive render. Training time is
reported separately. The evaluation views are not among the training views.

The 1M-triangle soup is opaque and its emitters lie inside it, so seen from outside it
renders black; its views are lit by the daylight sky (the reference's default sky is on).
"""
import math
import time

import numpy as np

from . import scenes

EPS = 1e-2


def relmse(x, ref):
    x = np.asarray(x, np.float64)[..., :3]
    r = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((x - r) ** 2 / (r * r + EPS)))


def random_camera(rng, kind):
    if kind == "cornell":  # in front of the open side of the box, looking inside
        pos = np.array([rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(-1.6, -0.6)])
        tgt = rng.uniform(0.3, 0.7, 3)
        return dict(position=tuple(pos), target=tuple(tgt), up=(0.0, 1.0, 0.0),
                    fov=math.radians(rng.uniform(30.0, 55.0)))
    if kind == "soup":  # outside the unit cube, looking at a point inside it
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos = 0.5 + d * rng.uniform(0.95, 1.4)
        tgt = rng.uniform(0.3, 0.7, 3)
        fov = math.radians(rng.uniform(35.0, 60.0))
    else:  # inside or near the 100^3 instance box
        pos = rng.uniform(-20.0, 120.0, 3)
        tgt = rng.uniform(25.0, 75.0, 3)
        fov = math.radians(rng.uniform(45.0, 75.0))
    return dict(position=tuple(pos), target=tuple(tgt), up=(0.0, 1.0, 0.0), fov=fov)


INST_CAMERA = dict(position=(50.0, 50.0, -20.0), target=(50.0, 50.0, 50.0), up=(0.0, 1.0, 0.0),
                   fov=math.radians(70.0))


def _render_spp(rend, spp, seed, frame_spp):
    """Accumulates `spp` samples (frames of at most frame_spp) from a reset; returns the
    wall time in seconds."""
    s = rend.settings
    s.seed = seed
    rend.Reset()
    t0 = time.perf_counter()
    left = spp
    while left > 0:
        s.spp = min(frame_spp, left)
        rend.RenderFrame()
        left -= s.spp
    return time.perf_counter() - t0


SKY = dict(enabled=True, draw_sun=True, dir=(0.3, 0.8, -0.5), angular_diameter=32.0, intensity=1000.0,
           turbidity=4.0)
SOUP_CAMERA = dict(position=(0.5, 0.5, -1.2), target=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov=math.radians(50.0))


def make_scene(kind, width=1920, height=1080):
    """The evaluation scenes: "soup" (the 1M-triangle soup under the sky), "inst" (the
    1000-instance scene) and "cornell"."""
    if kind == "soup":
        return dict(scenes.soup(width=width, height=height), sky=SKY)
    if kind == "inst":
        return scenes.instanced(width=width, height=height)
    return scenes.cornell(width=width, height=height)


TRAIN_VIEWS = (("soup", 8), ("inst", 2))
TRAIN_INPUT_SPP = (1, 2, 4, 8, 16, 32, 64)
TRAIN_SCHEDULE = ((6000, 1e-3), (3000, 3e-4), (3000, 1e-4))


def build_training_set(rend, width=1920, height=1080, views=TRAIN_VIEWS, input_spp=TRAIN_INPUT_SPP, target_spp=1024,
                       seed=0, frame_spp=32, log=print):
    """One slot per (view, input spp). Inputs span 1 to 64 spp so that the network learns
    how far to trust its input at every sample count of a progressive render (trained on
    1 to 8 spp only, it over-smoothed cleaner inputs and levelled off at relMSE ~0.027 on
    the soup view); the 1000-instance scene, whose 10,000 log-normal emitters make
    firefly-dominated frames, takes a smaller share of the views."""
    rng = np.random.default_rng(seed)
    plan = []
    for kind, n in views:
        plan += [(kind, tuple(input_spp))] * n
    total = sum(len(levels) for _, levels in plan)
    rend.denoiser_dataset(total, width, height)
    slot, view = 0, 0
    t0 = time.perf_counter()
    current = None
    for kind, levels in plan:
        if kind != current:
            rend.UpdateScene(make_scene(kind, width, height))
            current = kind
        rend.SetCamera(random_camera(rng, kind))
        _render_spp(rend, target_spp, seed=5000 + view, frame_spp=frame_spp)
        first = slot
        for k in levels:
            rend.denoiser_capture(slot, 1)
            slot += 1
        for i, k in enumerate(levels):
            _render_spp(rend, k, seed=1000 + 16 * view + i, frame_spp=frame_spp)
            rend.denoiser_capture(first + i, 0)
        view += 1
    log(f"training set: {slot} pairs from {view} views in {time.perf_counter() - t0:.1f} s")
    return time.perf_counter() - t0


def train(rend, schedule=TRAIN_SCHEDULE, batch=8, crop=128, seed=0):
    """Adam in stages of (steps, learning rate); returns the seconds and the losses."""
    t0 = time.perf_counter()
    losses = [rend.denoiser_train(steps, batch=batch, crop=crop, lr=lr, seed=seed + k)
              for k, (steps, lr) in enumerate(schedule)]
    return time.perf_counter() - t0, np.concatenate(losses)


def time_to_relmse(rend, scene, camera, width=1920, height=1080, target=0.01, ref_spp=8192, max_plain_spp=1024,
                   max_denoise_spp=64, frame_spp=32, log=print):
    """Returns a dict with the plain and denoised times to `target` (seconds) and the
    reference's estimated own error. Plain: sample counts double until the target is
    met, and the crossing is interpolated under relMSE ~ 1/spp. Denoised: sample counts
    double until a denoised frame meets the target, then whole sample counts inside the
    last doubling are bisected; the time is the measured one (render + inference) at
    the smallest count found to meet it."""
    rend.UpdateScene(dict(scene, render=dict(scene["render"], width=width, height=height)))
    rend.SetCamera(camera)
    t_a = _render_spp(rend, ref_spp // 2, seed=999, frame_spp=frame_spp)
    ref_a = rend.color()[..., :3].astype(np.float64)
    t_b = _render_spp(rend, ref_spp // 2, seed=998, frame_spp=frame_spp)
    ref_b = rend.color()[..., :3].astype(np.float64)
    ref = 0.5 * (ref_a + ref_b)
    ref_err = relmse(ref_a, ref_b) / 4.0
    rend.denoise()  # untimed warm-up: first launches load the kernels
    log(f"reference: {ref_spp} spp in {t_a + t_b:.1f} s, estimated own relMSE {ref_err:.5f}")
    plain = []
    k = 1
    while k <= max_plain_spp:
        t = _render_spp(rend, k, seed=17, frame_spp=frame_spp)
        e = relmse(rend.color(), ref)
        plain.append((k, t, e))
        if e - ref_err <= target:
            break
        k *= 2
    den = []

    def denoised_at(k):
        t = _render_spp(rend, k, seed=17, frame_spp=frame_spp)
        t0 = time.perf_counter()
        st = rend.denoise()
        td = time.perf_counter() - t0
        e = relmse(rend.denoised(), ref)
        den.append((k, t + td, e, st["total_ms"]))
        return e - ref_err <= target

    k = 1
    while k <= max_denoise_spp:
        if denoised_at(k):
            break
        k *= 2
    # the doubling brackets the crossing in (k / 2, k]: bisect over whole sample counts
    lo, hi = k // 2, k
    if k <= max_denoise_spp and k > 1:
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if denoised_at(mid):
                hi = mid
            else:
                lo = mid
    passing = [d for d in den if d[2] - ref_err <= target]
    # plain: the sample count where the reference-corrected relMSE crosses the target
    # under the 1/spp law of Monte Carlo error from the last measured count (between the
    # last two doublings when it crossed, beyond the budget otherwise), times the
    # measured seconds per sample at that count.
    kp, tp, ep = plain[-1]
    ep_true = max(ep - ref_err, 1e-12)
    n_star = kp * ep_true / target
    t_plain = tp / kp * n_star
    plain_how = "interpolated (relMSE ~ 1/spp)" if ep_true <= target else "extrapolated (relMSE ~ 1/spp)"
    reached = bool(passing)
    kd, td_, ed, _ = min(passing, key=lambda d: d[0]) if reached else den[-1]
    t_den = td_ if reached else None
    why = None
    if not reached:
        why = (f"the denoised image's relMSE at {kd} spp is {ed:.4f} ({ed - ref_err:.4f} after the reference's own "
               f"{ref_err:.4f}), above the target")
    if ref_err > 0.5 * target:
        why = (why + "; " if why else "") + f"the {ref_spp}-spp reference's own error ({ref_err:.4f}) is comparable " \
            f"to the target, so the figures are uncertain"
    return dict(target=target, plain_time_s=t_plain, plain_spp=n_star, plain_how=plain_how, denoised_time_s=t_den,
                denoised_spp=kd if t_den is not None else None, speedup=(t_plain / t_den) if t_den else None,
                plain_curve=plain, denoised_curve=den, reference_spp=ref_spp, reference_s=t_a + t_b,
                reference_relmse=ref_err, failure=why)
