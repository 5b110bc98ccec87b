// Scene files in the reference's JSON format (Tracer/FileIO/SceneFile.cpp:89-376,
// 683-731) with Wavefront OBJ/MTL models (the formats Tracer/FileIO/ModelFile.cpp
// imports through assimp), parsed on the host into plain arrays for the C ABI.
//
// Sections read: "models" (name -> path), "instances" (model, name, transform
// operations: matrix / translate / rotate-x|y|z (degrees) / scale (scalar or xyz)),
// "materials" (per model, per material name: the 13 lower-case property names, e.g.
// "diffuse", "emissive", "roughness"; "...Map" texture keys are reported as
// unsupported), "camera", "renderer" (multisample, maxdepth, ray-epsilon, aodist,
// zdepthmax) and "sky".
#pragma once
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

namespace trt {

// Minimal JSON document model (RFC 8259 subset: no duplicate-key semantics, numbers
// as double).
struct JsonValue {
  enum Type { Null, Bool, Number, String, Array, Object } type = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<JsonValue> arr;
  std::vector<std::pair<std::string, JsonValue>> obj;  // in file order

  const JsonValue* get(const std::string& key) const;  // nullptr when absent / not an object
};

// Parses `text`; on failure returns false and sets `err` (with line number).
bool json_parse(const std::string& text, JsonValue& out, std::string& err);

struct FileMaterial {
  std::string name;
  float diffuse[3] = {1.f, 1.f, 1.f};  // reference defaults (Tracer/Resources/Material.cpp:98-115)
  float emissive[3] = {0.f, 0.f, 0.f};
  float roughness = 1.f, metallic = 0.f, specular = 0.f, specular_tint = 0.f, sheen = 0.f, sheen_tint = 0.f;
  float clearcoat = 0.f, clearcoat_gloss = 0.f, subsurface = 0.f, anisotropic = 0.f;
};

struct FileModel {
  std::string name, path;
  std::vector<float> pos;       // 9 floats per triangle (object space)
  std::vector<float> nrm;       // 9 floats per triangle (vertex normals)
  std::vector<uint32_t> mat;    // model-local material index per triangle
  std::vector<FileMaterial> materials;
  std::vector<std::string> warnings;  // unsupported features met while loading
};

struct FileInstance {
  std::string name;
  uint32_t model = 0;   // index into SceneFile::models
  float rows[12];       // row-major 3x4 object-to-world
};

struct SceneFile {
  std::vector<FileModel> models;  // only the models that instances use, in first-use order
  std::vector<FileInstance> instances;
  // camera (reference defaults, Tracer/FileIO/SceneFile.cpp:98-106)
  float cam_pos[3] = {0.f, 0.f, -1.f}, cam_target[3] = {0.f, 0.f, 0.f}, cam_up[3] = {0.f, 1.f, 0.f};
  float fov_deg = 90.f, aperture = 0.f, distortion = 0.f, focal_dist = 1e5f, bokeh_rotation = 0.f;
  int32_t bokeh_sides = 0;
  bool has_camera = false;
  // renderer (reference defaults: Renderer::DefaultSettings)
  uint32_t spp = 1, max_depth = 16;
  float ray_epsilon = 1e-3f, ao_dist = 1.f, zdepth_max = 1.f;
  // sky (reference defaults, Tracer/Renderer/Sky.h:54-66; enabled unless the file says otherwise)
  bool sky_enabled = true, sky_draw_sun = true;
  float sky_dir[3] = {0.f, 1.f, 0.f};
  float sky_angular_diameter = 32.f, sky_intensity = 1000.f, sky_turbidity = 4.f, sky_bias = 1.f;
  std::vector<std::string> warnings;
};

// Loads `path` (JSON) and the OBJ models its instances use. Relative model paths are
// resolved against the scene file's directory, then its ancestors (the reference's
// scenes name models relative to the application's working directory), then as given.
// Returns false with `err`.
bool load_scene_file(const std::string& path, SceneFile& out, std::string& err);

// Loads one OBJ (+ its MTL libraries) into `m` (m.path must be set).
bool load_obj(const std::string& path, FileModel& m, std::string& err);

// Transform operations of an instance, composed the way the reference does
// (Tracer/FileIO/SceneFile.cpp:150-177 with Tracer/Utility/LinearMath.h:245-299) and read
// back row-major, as the reference's traversal reads its 12 floats
// (Tracer/Resources/Instance.cpp:39-49). Exposed for tests.
bool instance_transform(const JsonValue& ops, float rows[12], std::string& err);

}  // namespace trt
