// Host-code exercise for the sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer,
// and ThreadSanitizer): the project's host BVH builder and spatial pre-splits, the scene-file
// JSON/OBJ/MTL readers (valid files and malformed ones, whose errors must be reported, not
// crash), the PFM/PNG writers, and the CPU oracle (single- and two-level scenes, closest and
// any hit against brute force, renders with the sky and every strategy, the adjoint, the
// product-tree mode), all on several host threads. Prints "ok" and exits 0, or names the
// first failed check; a sanitizer finding aborts the process with its report.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/trt.h"
#include "bvh_build.h"
#include "scene_file.h"

extern "C" {
void* orc_scene_create(const float* pos, const float* nrm, const uint32_t* mat_ids, uint32_t n_tris,
                       const uint32_t* mesh_off, const uint32_t* mesh_cnt, uint32_t n_meshes, const float* mats,
                       uint32_t n_mats, const float* inst, uint32_t n_inst);
int orc_scene_use_product_bvh(void* p, float presplit, int threads);
void orc_set_sky(void* p, uint32_t enabled, uint32_t draw_sun, const float* dir, float angular_diameter,
                 float intensity, float turbidity, uint32_t sun_nee);
void orc_set_material_params(void* p, const float* params, uint32_t n, uint32_t which);
void orc_scene_destroy(void* p);
void orc_trace_closest(void* p, const float* rays, uint32_t n, float* hits, uint32_t* inst, int mode, int threads);
void orc_trace_any(void* p, const float* rays, uint32_t n, uint8_t* occ, int mode, int threads);
void orc_render(void* p, const float* cam, const uint32_t* uopts, const float* fopts, float* path_rgb,
                float* image_f32, double* image_f64, float* aov, uint64_t* counts, int threads);
void orc_backward(void* p, const float* cam, const uint32_t* uopts, const float* fopts, const float* dldi,
                  double* grad, int threads);
}

namespace {

int g_fail = 0;

void check(bool ok, const char* what) {
  if (!ok && !g_fail) {
    std::printf("FAIL %s\n", what);
    g_fail = 1;
  }
}

void quad(std::vector<float>& pos, std::vector<uint32_t>& mat, const float a[3], const float b[3], const float c[3],
          const float d[3], uint32_t m) {
  const float* t[2][3] = {{a, b, c}, {a, c, d}};
  for (auto& tri : t) {
    for (const float* v : tri) pos.insert(pos.end(), v, v + 3);
    mat.push_back(m);
  }
}

// material record: base rgb, roughness, emission rgb, metallic, specular, specular tint,
// sheen, sheen tint, clearcoat, clearcoat gloss, subsurface, anisotropic, model (bits)
std::vector<float> material(float r, float g, float b, float rough, float em, uint32_t model, float extra) {
  std::vector<float> m(17, 0.f);
  m[0] = r, m[1] = g, m[2] = b, m[3] = rough;
  m[4] = m[5] = m[6] = em;
  m[7] = extra;          // metallic
  m[8] = 0.5f;           // specular
  m[10] = extra;         // sheen
  m[12] = extra;         // clearcoat
  m[13] = 0.7f;          // clearcoat gloss
  m[14] = extra;         // subsurface
  m[15] = 0.5f * extra;  // anisotropic
  std::memcpy(&m[16], &model, 4);
  return m;
}

// A closed unit box (floor, ceiling, three walls), an emitter under the ceiling and a
// cloud of small random triangles inside.
void box_scene(std::vector<float>& pos, std::vector<uint32_t>& mat, uint32_t n_random) {
  const float P[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  quad(pos, mat, P[0], P[1], P[5], P[4], 0);  // floor
  quad(pos, mat, P[3], P[7], P[6], P[2], 0);  // ceiling
  quad(pos, mat, P[4], P[5], P[6], P[7], 1);  // back
  quad(pos, mat, P[0], P[4], P[7], P[3], 2);  // left
  quad(pos, mat, P[1], P[2], P[6], P[5], 3);  // right
  const float L[4][3] = {{0.4f, 0.99f, 0.4f}, {0.6f, 0.99f, 0.4f}, {0.6f, 0.99f, 0.6f}, {0.4f, 0.99f, 0.6f}};
  quad(pos, mat, L[0], L[1], L[2], L[3], 4);
  std::mt19937 rng(3);
  std::uniform_real_distribution<float> u(0.1f, 0.9f), s(-0.05f, 0.05f);
  for (uint32_t i = 0; i < n_random; ++i) {
    const float c[3] = {u(rng), u(rng) * 0.8f, u(rng)};
    for (int v = 0; v < 3; ++v)
      for (int k = 0; k < 3; ++k) pos.push_back(c[k] + s(rng));
    mat.push_back(5 + i % 2);
  }
}

std::vector<float> box_materials() {
  std::vector<float> m;
  const std::vector<float> ms[7] = {material(0.8f, 0.8f, 0.8f, 1.f, 0.f, 1, 0.f),
                                    material(0.7f, 0.7f, 0.7f, 0.5f, 0.f, 0, 0.f),
                                    material(0.8f, 0.1f, 0.1f, 1.f, 0.f, 1, 0.f),
                                    material(0.1f, 0.8f, 0.1f, 1.f, 0.f, 1, 0.f),
                                    material(0.f, 0.f, 0.f, 1.f, 12.f, 1, 0.f),
                                    material(0.5f, 0.6f, 0.7f, 0.3f, 0.f, 0, 0.8f),
                                    material(0.9f, 0.8f, 0.3f, 0.05f, 0.f, 0, 0.2f)};
  for (const auto& x : ms) m.insert(m.end(), x.begin(), x.end());
  return m;
}

void camera(float cam[15]) {
  const float c[15] = {0.5f, 0.5f, -1.5f, 0.5f, 0.5f, 0.5f, 0.f, 1.f, 0.f, 0.8f, 0.02f, 1.8f, 0.1f, 5.f, 0.3f};
  std::memcpy(cam, c, sizeof(c));
}

void exercise_oracle(void* s, uint32_t n_mats, int threads) {
  float cam[15];
  camera(cam);
  const uint32_t W = 24, H = 16, spp = 2, D = 6;
  const float fo[4] = {1e-3f, 0.f, 0.5f, 3.f};
  for (uint32_t strategy = 0; strategy < 3; ++strategy) {
    const uint32_t uo[14] = {W, H, spp, D, 2, 7, 0, 0, 0, 0, 0, 0, 0, strategy};
    std::vector<float> paths(size_t(W) * H * spp * 3), img(size_t(W) * H * 3), aov(size_t(W) * H * 8);
    std::vector<double> img64(size_t(W) * H * 3);
    std::vector<uint64_t> counts(2 * D);
    orc_render(s, cam, uo, fo, paths.data(), img.data(), img64.data(), aov.data(), counts.data(), threads);
    double sum = 0.0;
    for (float v : img) sum += v;
    check(std::isfinite(sum) && sum > 0.0, "render: finite, non-zero image");
    check(counts[0] == uint64_t(W) * H * spp, "render: one primary ray per path");
  }
  // debug modes and a crop
  for (uint32_t mode = 1; mode <= 12; ++mode) {
    const uint32_t uo[14] = {W, H, 1, D, 2, 7, 3, mode, mode == 6 ? 7u : 0u, 4, 2, 8, 6, 0};
    std::vector<float> img(8 * 6 * 3);
    orc_render(s, cam, uo, fo, nullptr, img.data(), nullptr, nullptr, nullptr, threads);
  }
  // adjoint
  {
    const uint32_t uo[14] = {W, H, spp, D, 1, 9, 0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<float> dldi(size_t(W) * H * 3, 1.f / float(W * H));
    std::vector<double> grad(size_t(n_mats) * 5);
    orc_backward(s, cam, uo, fo, dldi.data(), grad.data(), threads);
    double g = 0.0;
    for (double x : grad) g += std::fabs(x);
    check(std::isfinite(g) && g > 0.0, "adjoint: finite, non-zero gradient");
  }
  // closest / any hit: BVH against brute force
  std::mt19937 rng(11);
  std::uniform_real_distribution<float> u(-1.f, 1.f), p(0.f, 1.f);
  const uint32_t n = 2000;
  std::vector<float> rays(size_t(n) * 8);
  for (uint32_t i = 0; i < n; ++i) {
    float* r = &rays[size_t(i) * 8];
    r[0] = p(rng), r[1] = p(rng), r[2] = p(rng), r[3] = 0.f;
    r[4] = u(rng), r[5] = u(rng), r[6] = (i % 7 == 0) ? 0.f : u(rng);
    r[7] = (i % 3 == 0) ? 0.5f : INFINITY;
  }
  std::vector<float> h0(size_t(n) * 4), h1(size_t(n) * 4);
  std::vector<uint32_t> i0(n), i1(n);
  orc_trace_closest(s, rays.data(), n, h0.data(), i0.data(), 0, threads);
  orc_trace_closest(s, rays.data(), n, h1.data(), i1.data(), 1, threads);
  check(std::memcmp(h0.data(), h1.data(), h0.size() * 4) == 0 && i0 == i1, "closest hit: BVH equals brute force");
  std::vector<uint8_t> o0(n), o1(n);
  orc_trace_any(s, rays.data(), n, o0.data(), 0, threads);
  orc_trace_any(s, rays.data(), n, o1.data(), 1, threads);
  check(o0 == o1, "any hit: BVH equals brute force");
}

bool write_file(const std::string& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary);
  f << text;
  return bool(f);
}

void exercise_scene_files(const std::string& dir) {
  using namespace trt;
  // JSON: valid documents and malformed ones (each must fail with a message)
  const char* good[] = {"{}", "[1, -2.5e3, true, false, null, \"a\\u00e9\\n\"]", "{\"a\": {\"b\": [[], {}]}}"};
  for (const char* t : good) {
    JsonValue v;
    std::string err;
    check(json_parse(t, v, err), "json: valid document parses");
  }
  const char* bad[] = {"", "{", "[1, 2", "{\"a\" 1}", "\"unterminated", "[1e]", "\"\\u12\"", "{\"a\": tru}",
                       "[01]", "{} x", "\"\\q\"", "[-]"};
  for (const char* t : bad) {
    JsonValue v;
    std::string err;
    check(!json_parse(t, v, err) && !err.empty(), "json: malformed document is refused with a message");
  }
  std::string deep(5000, '[');
  {
    JsonValue v;
    std::string err;
    json_parse(deep, v, err);  // refused or parsed, but no overflow of the host stack
  }
  // OBJ + MTL + scene
  const std::string obj = dir + "/quad.obj", mtl = dir + "/quad.mtl", scene = dir + "/scene.json";
  check(write_file(mtl, "newmtl lamp\nKd 0.5 0.5 0.5\nKe 4 4 4\nnewmtl wall\nKd 0.8 0.2 0.2\nPr 0.4\nPm 0.1\n"),
        "write mtl");
  check(write_file(obj, "mtllib quad.mtl\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\nusemtl wall\n"
                        "f 1//1 2//1 3//1 4//1\nusemtl lamp\nf -4 -2 -1\n"),
        "write obj");
  check(write_file(scene, "{\"models\": [{\"name\": \"q\", \"path\": \"quad.obj\"}],"
                          " \"instances\": [{\"model\": \"q\", \"name\": \"a\", \"transform\": [{\"translate\": [0, 1, 2]},"
                          " {\"rotate-y\": 30}, {\"scale\": 2}]}, {\"model\": \"q\", \"name\": \"b\"}],"
                          " \"camera\": {\"position\": [0, 0, -3], \"fov\": 60},"
                          " \"renderer\": {\"multisample\": 4, \"maxdepth\": 5}, \"sky\": {\"enabled\": false}}"),
        "write scene");
  SceneFile sf;
  std::string err;
  check(load_scene_file(scene, sf, err), "scene file loads");
  check(sf.instances.size() == 2 && sf.models.size() == 1 && sf.models[0].pos.size() == 9 * 3,
        "scene file: instances and triangles");
  const char* bad_obj[] = {"v 0 0 0\nf 1 2 3\n", "v 0 0 0\nv 1 0 0\nf 1 2\n", "v 0 0\n", "v 0 0 0\nf 0 1 1\n",
                           "v 1 2 3\nvn 0 0 1\nf 1//5 1//1 1//1\n"};
  for (const char* t : bad_obj) {
    check(write_file(obj, t), "write obj");
    FileModel m;
    m.path = obj;
    std::string e;
    check(!load_obj(obj, m, e) && !e.empty(), "obj: malformed model is refused with a message");
  }
  const char* bad_scene[] = {"{\"models\": 3}", "{\"instances\": [{\"model\": \"missing\"}]}",
                             "{\"models\": [{\"name\": \"q\", \"path\": \"nowhere.obj\"}],"
                             " \"instances\": [{\"model\": \"q\"}]}",
                             "{\"instances\": [{\"model\": \"q\", \"transform\": [{\"skew\": 1}]}],"
                             " \"models\": [{\"name\": \"q\", \"path\": \"quad.obj\"}]}"};
  for (const char* t : bad_scene) {
    check(write_file(scene, t), "write scene");
    SceneFile s2;
    std::string e;
    check(!load_scene_file(scene, s2, e) && !e.empty(), "scene: malformed file is refused with a message");
  }
  // image writers
  std::vector<float> rgba(6 * 4 * 4, 0.25f);
  std::vector<uint8_t> rgba8(6 * 4 * 4, 128);
  check(trt_write_pfm((dir + "/x.pfm").c_str(), 6, 4, rgba.data()) == TRT_OK, "pfm");
  check(trt_write_png((dir + "/x.png").c_str(), 6, 4, rgba8.data()) == TRT_OK, "png");
  check(trt_write_png((dir + "/no/such/dir.png").c_str(), 6, 4, rgba8.data()) != TRT_OK, "png: unwritable path");
}

void exercise_bvh(int threads) {
  std::mt19937 rng(5);
  std::uniform_real_distribution<float> u(0.f, 1.f);
  const uint32_t n = 6000;
  std::vector<float> pos(size_t(n) * 9);
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 9; ++k) pos[size_t(i) * 9 + k] = (k < 3 || i % 4) ? u(rng) : pos[size_t(i) * 9 + k % 3];
  for (float beta : {0.f, 1.f, 6.f}) {
    std::vector<float> box;
    std::vector<uint32_t> ref;
    if (beta > 0.f) {
      trt::presplit_triangles(pos.data(), n, beta, box, ref);
    } else {
      box.resize(size_t(n) * 6);
      for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
          const float* p = &pos[size_t(i) * 9];
          box[size_t(i) * 6 + k] = std::fmin(p[k], std::fmin(p[3 + k], p[6 + k]));
          box[size_t(i) * 6 + 3 + k] = std::fmax(p[k], std::fmax(p[3 + k], p[6 + k]));
        }
    }
    const uint32_t nr = uint32_t(box.size() / 6);
    for (int leaf : {1, 4}) {
      const trt::BuildResult b = trt::build_binned_sah(box.data(), nr, 32, leaf, threads, trt::kTravCostTriangles);
      check(b.prim_order.size() == nr, "bvh: every reference once");
      std::vector<trt::DevNode48> packed;
      check(trt::pack_nodes(trt::to_device_layout(b), packed), "bvh: nodes pack into 48 bytes");
    }
  }
  const trt::BuildResult one = trt::build_binned_sah(pos.data(), 1, 32, 4, threads, trt::kTravCostTriangles);
  check(one.nodes.size() == 1, "bvh: one primitive");
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  const int threads = 4;
  exercise_bvh(threads);
  exercise_scene_files(dir);
  // single-level box with a random cloud
  std::vector<float> pos;
  std::vector<uint32_t> mat;
  box_scene(pos, mat, 300);
  const std::vector<float> mats = box_materials();
  const uint32_t n_mats = uint32_t(mats.size() / 17), n_tris = uint32_t(mat.size());
  void* s = orc_scene_create(pos.data(), nullptr, mat.data(), n_tris, nullptr, nullptr, 0, mats.data(), n_mats,
                             nullptr, 0);
  exercise_oracle(s, n_mats, threads);
  const float dir3[3] = {0.3f, 0.8f, -0.5f};
  orc_set_sky(s, 1, 1, dir3, 300.f, 1000.f, 4.f, 1);
  exercise_oracle(s, n_mats, threads);
  std::vector<float> params(size_t(n_mats) * 5, 0.5f);
  orc_set_material_params(s, params.data(), n_mats, 1u);
  exercise_oracle(s, n_mats, threads);
  check(orc_scene_use_product_bvh(s, 2.f, threads) > 0, "product tree mode");
  exercise_oracle(s, n_mats, threads);
  orc_scene_destroy(s);
  // two-level: the box as mesh 0, the cloud as mesh 1, three instances
  const uint32_t box_tris = 12;
  const uint32_t off[2] = {0, box_tris}, cnt[2] = {box_tris, n_tris - box_tris};
  std::vector<float> inst;
  const float T[3][12] = {{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0},
                          {0.5f, 0, 0, 0.2f, 0, 0.5f, 0, 0.1f, 0, 0, 0.5f, 0.3f},
                          {0, 0, 0.4f, 0.5f, 0, 0.4f, 0, 0.2f, -0.4f, 0, 0, 0.6f}};
  const uint32_t meshes[3] = {0, 1, 1}, moff[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i) {
    inst.insert(inst.end(), T[i], T[i] + 12);
    float m, o;
    std::memcpy(&m, &meshes[i], 4);
    std::memcpy(&o, &moff[i], 4);
    inst.push_back(m);
    inst.push_back(o);
  }
  std::vector<float> nrm(size_t(n_tris) * 9);
  for (uint32_t t = 0; t < n_tris; ++t)
    for (int v = 0; v < 3; ++v) {
      nrm[size_t(t) * 9 + v * 3 + 0] = 0.f;
      nrm[size_t(t) * 9 + v * 3 + 1] = 1.f;
      nrm[size_t(t) * 9 + v * 3 + 2] = 0.f;
    }
  void* s2 = orc_scene_create(pos.data(), nrm.data(), mat.data(), n_tris, off, cnt, 2, mats.data(), n_mats,
                              inst.data(), 3);
  exercise_oracle(s2, n_mats, threads);
  check(orc_scene_use_product_bvh(s2, 2.f, threads) == -1, "product tree mode refuses instanced scenes");
  orc_scene_destroy(s2);
  if (g_fail) return 1;
  std::printf("ok\n");
  return 0;
}
