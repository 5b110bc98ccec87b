"""Python mirror of the reference's Renderer / Scene façade over the C ABI.

Method names follow Tracer::Renderer (Tracer/Renderer/Renderer.h:45-97): UpdateScene,
RenderFrame, Reset, SetCamera, PickRay, Settings / SetSettings (KernelSettings),
Statistics, SampleCount. Device work runs in libtrt.so on the GPU; nothing here
computes pixels.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import TrtError


def _fp(a):
    return a.ctypes.data_as(_lib.c_float_p)


def _u32(a):
    return a.ctypes.data_as(_lib.c_u32_p)


class KernelSettings:
    """Reference KernelSettings (Tracer/Common/CommonStructs.h:145-153) plus the film
    size, Russian-roulette start depth and firefly clamp."""

    def __init__(self, width=256, height=256, spp=1, max_depth=4, rr_depth=3, seed=0, ray_epsilon=1e-3, clamp=0.0,
                 mode=0, mode_flags=0, ao_dist=1.0, zdepth_max=1.0):
        self.width, self.height, self.spp = width, height, spp
        self.max_depth, self.rr_depth, self.seed = max_depth, rr_depth, seed
        self.ray_epsilon, self.clamp = ray_epsilon, clamp
        self.mode, self.mode_flags, self.ao_dist, self.zdepth_max = mode, mode_flags, ao_dist, zdepth_max

    def to_c(self):
        o = _lib.TrtRenderOpts()
        o.width, o.height, o.spp = self.width, self.height, self.spp
        o.max_depth, o.rr_depth, o.seed = self.max_depth, self.rr_depth, self.seed
        o.ray_epsilon, o.clamp = self.ray_epsilon, self.clamp
        o.mode, o.mode_flags, o.ao_dist, o.zdepth_max = self.mode, self.mode_flags, self.ao_dist, self.zdepth_max
        return o

    @classmethod
    def from_dict(cls, d):
        return cls(**d)


def camera_struct(cam):
    c = _lib.TrtCamera()
    c.position[:] = list(cam["position"])
    c.target[:] = list(cam["target"])
    c.up[:] = list(cam["up"])
    c.fov = cam["fov"]
    c.aperture = cam.get("aperture", 0.0)
    c.focal_dist = cam.get("focal_dist", 1.0)
    c.distortion = cam.get("distortion", 0.0)
    c.bokeh_sides = int(cam.get("bokeh_sides", 0))
    c.bokeh_rotation = cam.get("bokeh_rotation", 0.0)
    return c


class Renderer:
    """Owns one trt context on `device`."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        self.ctx = ctypes.c_void_p()
        st = self.lib.trt_create(device, ctypes.byref(self.ctx))
        if st != 0:
            raise TrtError(st, "trt_create failed (no usable HIP device?)")
        self.settings = KernelSettings()
        self.last_stats = None

    def close(self):
        if self.ctx:
            self.lib.trt_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise TrtError(st, self.lib.trt_last_error(self.ctx).decode())

    # -- scene -----------------------------------------------------------------------
    def UpdateScene(self, scene, bins=32, max_leaf=4, threads=0, presplit=6.0):
        """Uploads a scene dict (see trayracer_amd.scenes) and builds its BVH(s).

        presplit: spatial pre-splitting budget in extra references per triangle (0 = off).
        """
        lib, ctx = self.lib, self.ctx
        self._check(lib.trt_clear_scene(ctx))
        mats = np.ascontiguousarray(scene["materials"], np.float32)
        self.n_materials = len(mats)
        cm = (_lib.TrtMaterial * len(mats))()
        ctypes.memmove(cm, mats.ctypes.data, mats.nbytes)
        self._check(lib.trt_set_materials(ctx, cm, len(mats)))
        pos = np.ascontiguousarray(scene["positions"], np.float32)
        nrm = scene.get("normals")
        nrm = None if nrm is None else np.ascontiguousarray(nrm, np.float32)
        mid = np.ascontiguousarray(scene["mat_ids"], np.uint32)
        for off, cnt in scene["meshes"]:
            mesh_id = ctypes.c_uint32()
            self._check(lib.trt_add_mesh(ctx, _fp(pos[off:off + cnt]), None if nrm is None else _fp(nrm[off:off + cnt]),
                                         _u32(mid[off:off + cnt]), cnt, ctypes.byref(mesh_id)))
        inst = scene.get("instances")
        if inst is not None:
            inst = np.ascontiguousarray(inst, np.float32)
            ci = (_lib.TrtInstance * len(inst))()
            ctypes.memmove(ci, inst.ctypes.data, inst.nbytes)
            self._check(lib.trt_set_instances(ctx, ci, len(inst)))
        bo = _lib.TrtBuildOpts(bins, max_leaf, threads, presplit)
        self._check(lib.trt_build(ctx, ctypes.byref(bo)))
        if scene.get("sky") is not None:
            self.SetSky(scene["sky"])
        if "camera" in scene:
            self.SetCamera(scene["camera"])
        if "render" in scene:
            self.SetSettings(KernelSettings.from_dict(scene["render"]))

    def LoadScene(self, path, bins=32, max_leaf=4, threads=0, presplit=6.0):
        """Loads a scene file in the reference's JSON format (OBJ models) and builds it
        (SceneFile::Load, Tracer/FileIO/SceneFile.cpp:683-731). The file's render settings
        (multisample, maxdepth, ray-epsilon) replace the current settings; the film size,
        seed and Russian-roulette depth are kept. Returns the file's info as a dict."""
        info = _lib.TrtSceneFileInfo()
        bo = _lib.TrtBuildOpts(bins, max_leaf, threads, presplit)
        self._check(self.lib.trt_load_scene_file(self.ctx, os.fsencode(path), ctypes.byref(bo), ctypes.byref(info)))
        s = self.settings
        s.spp, s.max_depth, s.ray_epsilon = info.spp, info.max_depth, info.ray_epsilon
        s.ao_dist, s.zdepth_max = info.ao_dist, info.zdepth_max
        return info.as_dict()

    TONEMAP = {"aces": 0, "filmic": 1, "lottes": 2, "reinhard": 3, "reinhard2": 4, "uchimura": 5, "uncharted2": 6,
               "none": 7}

    def tonemap(self, method="aces", exposure=1.0, gamma=2.2, denoised=False):
        """8-bit RGBA display image [H, W, 4] (row 0 = bottom) of the mean colour or the
        denoised frame (Tonemap.frag: exposure, curve, gamma)."""
        m = self.TONEMAP[method] if isinstance(method, str) else int(method)
        o = _lib.TrtTonemapOpts(m, exposure, gamma, 1 if denoised else 0)
        s = self.settings
        out = np.empty((s.height, s.width, 4), np.uint8)
        self._check(self.lib.trt_tonemap(self.ctx, ctypes.byref(o), out.ctypes.data_as(_lib.c_u8_p)))
        return out

    def save_png(self, path, **tonemap_args):
        img = self.tonemap(**tonemap_args)
        self._check(self.lib.trt_write_png(os.fsencode(path), img.shape[1], img.shape[0],
                                           img.ctypes.data_as(_lib.c_u8_p)))

    def save_pfm(self, path, denoised=False):
        img = np.ascontiguousarray(self.denoised() if denoised else self.color(), np.float32)
        self._check(self.lib.trt_write_pfm(os.fsencode(path), img.shape[1], img.shape[0], _fp(img)))

    def SetRenderMode(self, mode, flags=0):
        """Render mode by name or number (Renderer::SetRenderMode; _lib.MODES) and, for
        the material-property mode, the property (name from _lib.MATERIAL_PROPERTIES or id)."""
        m = _lib.MODES[mode] if isinstance(mode, str) else int(mode)
        f = _lib.MATERIAL_PROPERTIES.index(flags) if isinstance(flags, str) else int(flags)
        self.settings.mode, self.settings.mode_flags = m, f
        self.Reset()

    def bvh_info(self):
        info = _lib.TrtBvhInfo()
        self._check(self.lib.trt_bvh_info_get(self.ctx, ctypes.byref(info)))
        return info.as_dict()

    def SetSky(self, sky):
        """sky: dict(enabled, draw_sun, dir, angular_diameter (arc minutes), intensity,
        turbidity) as in the reference's scene files (Tracer/Renderer/Sky.h), and sun_nee
        (the sun disc as a light of next-event estimation; off by default, as in the
        reference)."""
        c = _lib.TrtSky(int(sky.get("enabled", True)), int(sky.get("draw_sun", True)),
                        (ctypes.c_float * 3)(*sky.get("dir", (0.0, 1.0, 0.0))), float(sky.get("angular_diameter", 32.0)),
                        float(sky.get("intensity", 1000.0)), float(sky.get("turbidity", 4.0)),
                        int(sky.get("sun_nee", False)))
        self._check(self.lib.trt_set_sky(self.ctx, ctypes.byref(c)))

    # -- camera / settings -----------------------------------------------------------
    def SetCamera(self, cam):
        c = camera_struct(cam)
        self._check(self.lib.trt_set_camera(self.ctx, ctypes.byref(c)))

    def Settings(self):
        return self.settings

    def SetSettings(self, settings):
        self.settings = settings
        self.Reset()

    # -- rendering -------------------------------------------------------------------
    def RenderFrame(self):
        st = _lib.TrtStats()
        o = self.settings.to_c()
        self._check(self.lib.trt_render(self.ctx, ctypes.byref(o), ctypes.byref(st)))
        self.last_stats = st.as_dict()
        return self.last_stats

    def SetPhaseTiming(self, enabled):
        """Per-phase times in RenderFrame's statistics (on by default). Off, a frame
        records only the events that order its two streams; counts and frame_ms remain."""
        self._check(self.lib.trt_set_phase_timing(self.ctx, 1 if enabled else 0))

    def Reset(self):
        self._check(self.lib.trt_reset(self.ctx))

    def SampleCount(self):
        return int(self.lib.trt_sample_count(self.ctx))

    def frame_id(self):
        """Frames rendered by this context so far (changes with every RenderFrame)."""
        return int(self.lib.trt_frame_id(self.ctx))

    def Statistics(self):
        return self.last_stats

    def measure_traversal(self):
        tc = _lib.TrtTraceCounts()
        o = self.settings.to_c()
        self._check(self.lib.trt_measure_traversal(self.ctx, ctypes.byref(o), ctypes.byref(tc)))
        return tc.as_dict()

    # -- outputs ---------------------------------------------------------------------
    def _pixels(self, fn):
        s = self.settings
        out = np.empty((s.height, s.width, 4), np.float32)
        self._check(fn(self.ctx, _fp(out)))
        return out

    def color(self):
        return self._pixels(self.lib.trt_read_color)

    def accum(self):
        return self._pixels(self.lib.trt_read_accum)

    def write_accum(self, rgba):
        """Restores a progressive render from an accumulator read by accum() (same film
        size as the settings); the next RenderFrame continues it."""
        a = np.ascontiguousarray(rgba, np.float32)
        H, W = a.shape[0], a.shape[1]
        self._check(self.lib.trt_write_accum(self.ctx, W, H, _fp(a)))

    def albedo(self):
        return self._pixels(self.lib.trt_read_albedo)

    def normal(self):
        return self._pixels(self.lib.trt_read_normal)

    def path_radiance(self):
        s = self.settings
        n = s.width * s.height * s.spp
        out = np.empty((n, 3), np.float32)
        self._check(self.lib.trt_read_path_radiance(self.ctx, _fp(out), n))
        return out

    # -- differentiable rendering -----------------------------------------------------
    EVAL, SAMPLING, BOTH = 1, 2, 3

    def set_material_params(self, params, which=3):
        """params [n_materials, 5]: base colour rgb, roughness, metallic."""
        params = np.ascontiguousarray(params, np.float32).reshape(-1, 5)
        self._check(self.lib.trt_set_material_params(self.ctx, _fp(params), params.shape[0], which))

    def backward(self, dl_dimage):
        """Gradient of a loss w.r.t. every material's 5 parameters for the last frame;
        dl_dimage = dLoss/d(frame mean image) [H, W, 3]."""
        s = self.settings
        g = np.ascontiguousarray(dl_dimage, np.float32).reshape(s.height, s.width, 3)
        out = np.empty((self.n_materials, 5), np.float32)
        o = s.to_c()
        self._check(self.lib.trt_backward(self.ctx, ctypes.byref(o), _fp(g), _fp(out)))
        return out

    def frame_image(self):
        """Mean of the last frame's paths per pixel (= color() after a single frame)."""
        s = self.settings
        L = self.path_radiance().reshape(s.spp, s.height, s.width, 3).astype(np.float32)
        acc = np.zeros((s.height, s.width, 3), np.float32)
        for k in range(s.spp):
            acc += L[k]
        return acc * np.float32(1.0 / s.spp)

    # -- denoiser ---------------------------------------------------------------------
    def denoiser_init(self, seed=0):
        self._check(self.lib.trt_denoiser_init(self.ctx, seed))

    def denoiser_params(self):
        n = int(self.lib.trt_denoiser_num_params())
        out = np.empty(n, np.float32)
        self._check(self.lib.trt_denoiser_get_params(self.ctx, _fp(out), n))
        return out

    def denoiser_set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self.lib.trt_denoiser_set_params(self.ctx, _fp(w), w.size))

    def denoiser_save(self, path):
        self._check(self.lib.trt_denoiser_save(self.ctx, str(path).encode()))

    def denoiser_load(self, path):
        self._check(self.lib.trt_denoiser_load(self.ctx, str(path).encode()))

    def denoiser_dataset(self, n, width, height):
        self._check(self.lib.trt_denoiser_dataset(self.ctx, n, width, height))

    def denoiser_capture(self, slot, kind):
        self._check(self.lib.trt_denoiser_capture(self.ctx, slot, kind))

    def denoiser_train(self, steps, batch=8, crop=128, lr=1e-3, seed=0):
        losses = np.empty(steps, np.float32)
        self._check(self.lib.trt_denoiser_train(self.ctx, steps, batch, crop, lr, seed, _fp(losses)))
        return losses

    def denoise(self):
        st = _lib.TrtDenoiseStats()
        self._check(self.lib.trt_denoise(self.ctx, ctypes.byref(st)))
        return st.as_dict()

    def denoised(self):
        return self._pixels(self.lib.trt_read_denoised)

    def SetDenoiser(self, enabled=True, sample_threshold=0):
        """The reference's Denoiser::SetEnabled / SetSampleTreshold."""
        self._check(self.lib.trt_denoiser_set_policy(self.ctx, int(enabled), int(sample_threshold)))

    def DenoiseFrame(self):
        """Renderer::DenoiseFrame: denoises (blended by 1 / (spp + 1)) when the policy
        says so for the current sample count; returns the stats, or None when it did not."""
        st = _lib.TrtDenoiseStats()
        ran = ctypes.c_uint32()
        self._check(self.lib.trt_denoise_frame(self.ctx, ctypes.byref(st), ctypes.byref(ran)))
        return st.as_dict() if ran.value else None

    def denoiser_forward(self, features):
        f = np.ascontiguousarray(features, np.float32)
        N, H, W, C = f.shape
        assert C == 16
        out = np.empty((N, H, W, 32), np.float32)
        self._check(self.lib.trt_denoiser_forward(self.ctx, _fp(f), N, H, W, _fp(out)))
        return out

    def denoiser_apply(self, features):
        """Inference path of denoise() on given features -> radiance [N][H][W][4]."""
        f = np.ascontiguousarray(features, np.float32)
        N, H, W, C = f.shape
        assert C == 16
        out = np.empty((N, H, W, 4), np.float32)
        self._check(self.lib.trt_denoiser_apply(self.ctx, _fp(f), N, H, W, _fp(out)))
        return out

    def denoiser_activation(self, index, shape):
        out = np.empty(shape, np.float32)
        self._check(self.lib.trt_denoiser_activation(self.ctx, index, _fp(out), out.size))
        return out

    def denoiser_backward(self, dout):
        d = np.ascontiguousarray(dout, np.float32)
        N, H, W, _ = d.shape
        g = np.empty(int(self.lib.trt_denoiser_num_params()), np.float32)
        self._check(self.lib.trt_denoiser_backward(self.ctx, _fp(d), N, H, W, _fp(g)))
        return g

    # -- ray queries -----------------------------------------------------------------
    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.empty((n, 4), np.float32)
        inst = np.empty(n, np.uint32)
        self._check(self.lib.trt_trace_closest(self.ctx, _fp(rays), n, _fp(hits), _u32(inst)))
        return hits, inst

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        occ = np.empty(n, np.uint8)
        self._check(self.lib.trt_trace_any(self.ctx, _fp(rays), n, occ.ctypes.data_as(_lib.c_u8_p)))
        return occ

    def PickRay(self, x, y):
        r = _lib.TrtPickResult()
        s = self.settings
        self._check(self.lib.trt_pick(self.ctx, s.width, s.height, x, y, ctypes.byref(r)))
        return dict(origin=tuple(r.origin), direction=tuple(r.direction), t=r.t, instance=r.instance,
                    primitive=r.primitive)
