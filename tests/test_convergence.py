"""Host logic of the time-to-relMSE measurement (trayracer_amd/convergence.py) on a
stand-in renderer whose images have a known error: relMSE(plain) = a / spp and
relMSE(denoised) = b / spp against a reference whose two halves differ by a known
amount. Checks the interpolated plain crossing, the bisected denoised crossing and the
reported failure when the denoised error never meets the target (CPU)."""
import types

import numpy as np
import pytest

from trayracer_amd import convergence as C


class FakeClock:
    """Stand-in for time.perf_counter: the fake renderer advances it by 1 ms per sample
    and 1 ms per denoise, so the measured times (and the speedup) are exact, not the
    host's timing of the stand-in's Python work."""

    def __init__(self):
        self.t = 0.0

    def perf_counter(self):
        return self.t


@pytest.fixture(autouse=True)
def fake_clock(monkeypatch):
    clock = FakeClock()
    monkeypatch.setattr(C, "time", types.SimpleNamespace(perf_counter=clock.perf_counter))
    return clock


class FakeSettings:
    def __init__(self):
        self.spp = 1
        self.seed = 0


class FakeRenderer:
    """Accumulates a frame whose per-pixel error shrinks as 1/sqrt(spp)."""

    def __init__(self, plain_err, den_err, clock, shape=(64, 64)):
        self.settings = FakeSettings()
        self.clock = clock
        self.plain_err, self.den_err = plain_err, den_err
        self.rng = np.random.default_rng(0)
        self.truth = 0.5 + 0.5 * self.rng.random(shape + (3,))
        self.noise = self.rng.standard_normal(shape + (3,))
        self.n = 0

    def UpdateScene(self, scene):
        pass

    def SetCamera(self, cam):
        pass

    def Reset(self):
        self.n = 0

    def RenderFrame(self):
        self.n += self.settings.spp
        self.clock.t += 1e-3 * self.settings.spp

    def _img(self, err):
        # relMSE = mean(err * noise^2 / n * (r^2 + 1e-2) / (r^2 + 1e-2)) ~ err / n for unit-variance noise
        den = self.truth ** 2 + 1e-2
        scale = np.sqrt(err * den / max(self.n, 1))
        if self.settings.seed in (998, 999):  # the reference halves: nearly exact, independent
            scale = scale * 1e-3
            return self.truth + scale * self.rng.standard_normal(self.truth.shape)
        return self.truth + scale * self.noise

    def color(self):
        return self._img(self.plain_err)

    def denoise(self):
        self.clock.t += 1e-3
        return {"total_ms": 1.0}

    def denoised(self):
        return self._img(self.den_err)


def _run(plain_err, den_err, clock, **kw):
    r = FakeRenderer(plain_err, den_err, clock)
    scene = {"render": {"width": 64, "height": 64}}
    return C.time_to_relmse(r, scene, camera=None, width=64, height=64, ref_spp=64, log=lambda m: None, **kw)


def test_plain_crossing_is_interpolated_and_denoised_crossing_bisected(fake_clock):
    res = _run(plain_err=0.73, den_err=0.058, clock=fake_clock)  # plain crosses at ~73 spp, denoised at ~6 spp
    assert abs(res["plain_spp"] - 73.0) < 8.0
    assert res["plain_how"].startswith("interpolated")
    ks = [int(k) for k, *_ in res["denoised_curve"]]
    assert ks[:4] == [1, 2, 4, 8]          # doubling up to the first passing count
    assert set(ks[4:]) <= {5, 6, 7}        # then bisection inside (4, 8]
    assert res["denoised_spp"] in (5, 6, 7)
    passing = [k for k, _, e, _ in res["denoised_curve"] if e - res["reference_relmse"] <= 0.01]
    assert res["denoised_spp"] == min(passing)
    assert res["failure"] is None
    # 1 ms per sample: plain ~73 ms; denoised k samples + 1 ms of inference
    assert abs(res["plain_time_s"] - res["plain_spp"] * 1e-3) < 1e-9
    assert abs(res["denoised_time_s"] - (res["denoised_spp"] + 1) * 1e-3) < 1e-9
    assert res["speedup"] > 5.0


def test_denoised_failure_is_reported(fake_clock):
    res = _run(plain_err=0.73, den_err=5.0, clock=fake_clock)  # needs ~500 spp: beyond max_denoise_spp
    assert res["denoised_time_s"] is None and res["speedup"] is None
    assert "above the target" in res["failure"]
