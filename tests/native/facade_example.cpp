// A C++ program that uses the renderer the way the reference's application does
// (Tracer/Renderer/Renderer.h:34-171 through include/trt.hpp): load a scene file, set the
// settings, render frames progressively, pick a ray, run the denoiser policy, read the
// frame. Built and run by tests/test_facade.py. Prints one line of key=value results;
// exits 2 when no HIP device can be used.
#include <cstdio>
#include <cstring>
#include <numeric>

#include "trt.hpp"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s scene.json [frames]\n", argv[0]);
    return 1;
  }
  const int frames = argc > 2 ? std::atoi(argv[2]) : 4;
  try {
    trt::Renderer r(0);
    trt_build_opts bo{32, 4, 0, 1.0f};
    const trt_scene_file_info info = r.LoadScene(argv[1], &bo);
    trt_render_opts s = r.Settings();
    s.width = 128;
    s.height = 96;
    s.spp = 2;
    s.seed = 5;
    r.SetSettings(s);
    uint64_t rays = 0;
    for (int f = 0; f < frames; ++f) {
      const trt_stats& st = r.RenderFrame();
      rays += st.extend_rays + st.shadow_rays;
    }
    const std::vector<float> c = r.Color();
    double sum[3] = {0, 0, 0};
    for (size_t i = 0; i < c.size(); i += 4)
      for (int k = 0; k < 3; ++k) sum[k] += c[i + k];
    const trt_pick_result pick = r.PickRay(64, 48);
    const uint32_t samples = r.SampleCount();
    // an invalid setting is reported as an exception, not an exit
    bool caught = false;
    try {
      trt_render_opts bad = s;
      bad.max_depth = 0;
      r.SetSettings(bad);
      r.RenderFrame();
    } catch (const trt::Error& e) {
      caught = e.status() == TRT_INVALID_ARG;
    }
    std::printf("ok models=%u instances=%u samples=%u rays=%llu mean=%.9g,%.9g,%.9g pick_prim=%u pick_t=%.9g error_caught=%d\n",
                info.n_models, info.n_instances, samples, static_cast<unsigned long long>(rays),
                sum[0] / (c.size() / 4), sum[1] / (c.size() / 4), sum[2] / (c.size() / 4), pick.primitive, pick.t,
                caught ? 1 : 0);
    return 0;
  } catch (const trt::Error& e) {
    if (e.status() == TRT_DEVICE) {
      std::printf("no device: %s\n", e.what());
      return 2;
    }
    std::fprintf(stderr, "trt error %d: %s\n", int(e.status()), e.what());
    return 1;
  }
}
