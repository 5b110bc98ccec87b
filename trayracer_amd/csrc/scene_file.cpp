// Scene-file and OBJ/MTL import (host). See scene_file.h.
#include "scene_file.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace trt {

// ---------------------------------------------------------------------------------
// JSON
// ---------------------------------------------------------------------------------
const JsonValue* JsonValue::get(const std::string& key) const {
  if (type != Object) return nullptr;
  for (const auto& kv : obj)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

namespace {

struct JsonReader {
  const std::string& s;
  size_t i = 0;
  std::string err;

  explicit JsonReader(const std::string& text) : s(text) {}

  int line() const {
    int l = 1;
    for (size_t k = 0; k < i && k < s.size(); ++k) l += s[k] == '\n';
    return l;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = "JSON line " + std::to_string(line()) + ": " + m;
    return false;
  }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if (s.compare(i, n, w) != 0) return false;
    i += n;
    return true;
  }
  static void put_utf8(std::string& o, uint32_t cp) {
    if (cp < 0x80) {
      o += char(cp);
    } else if (cp < 0x800) {
      o += char(0xC0 | (cp >> 6));
      o += char(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      o += char(0xE0 | (cp >> 12));
      o += char(0x80 | ((cp >> 6) & 0x3F));
      o += char(0x80 | (cp & 0x3F));
    } else {
      o += char(0xF0 | (cp >> 18));
      o += char(0x80 | ((cp >> 12) & 0x3F));
      o += char(0x80 | ((cp >> 6) & 0x3F));
      o += char(0x80 | (cp & 0x3F));
    }
  }
  bool hex4(uint32_t& v) {
    if (i + 4 > s.size()) return fail("truncated \\u escape");
    v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = s[i++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
      else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
      else return fail("bad \\u escape");
    }
    return true;
  }
  bool string(std::string& o) {
    if (i >= s.size() || s[i] != '"') return fail("expected string");
    ++i;
    o.clear();
    while (i < s.size()) {
      const char c = s[i++];
      if (c == '"') return true;
      if (static_cast<unsigned char>(c) < 0x20) return fail("control character in string");
      if (c != '\\') {
        o += c;
        continue;
      }
      if (i >= s.size()) break;
      const char e = s[i++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          uint32_t cp;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < s.size() && s[i] == '\\' && s[i + 1] == 'u') {
            i += 2;
            uint32_t lo;
            if (!hex4(lo)) return false;
            if (lo >= 0xDC00 && lo < 0xE000) cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(o, cp);
          break;
        }
        default: return fail("bad escape");
      }
    }
    return fail("unterminated string");
  }
  bool number(double& v) {
    const size_t b = i;
    if (i < s.size() && s[i] == '-') ++i;
    if (i >= s.size() || !(std::isdigit(static_cast<unsigned char>(s[i])))) return fail("bad number");
    if (s[i] == '0') {
      ++i;
    } else {
      while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
    }
    if (i < s.size() && s[i] == '.') {
      ++i;
      if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return fail("bad fraction");
      while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
    }
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
      ++i;
      if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
      if (i >= s.size() || !std::isdigit(static_cast<unsigned char>(s[i]))) return fail("bad exponent");
      while (i < s.size() && std::isdigit(static_cast<unsigned char>(s[i]))) ++i;
    }
    v = std::strtod(s.substr(b, i - b).c_str(), nullptr);
    return true;
  }
  bool value(JsonValue& v, int depth) {
    if (depth > 256) return fail("nesting too deep");
    ws();
    if (i >= s.size()) return fail("unexpected end");
    const char c = s[i];
    if (c == '{') {
      ++i;
      v.type = JsonValue::Object;
      ws();
      if (i < s.size() && s[i] == '}') {
        ++i;
        return true;
      }
      for (;;) {
        ws();
        std::string k;
        if (!string(k)) return false;
        ws();
        if (i >= s.size() || s[i] != ':') return fail("expected ':'");
        ++i;
        JsonValue child;
        if (!value(child, depth + 1)) return false;
        v.obj.emplace_back(std::move(k), std::move(child));
        ws();
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == '}') {
          ++i;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      ++i;
      v.type = JsonValue::Array;
      ws();
      if (i < s.size() && s[i] == ']') {
        ++i;
        return true;
      }
      for (;;) {
        JsonValue child;
        if (!value(child, depth + 1)) return false;
        v.arr.push_back(std::move(child));
        ws();
        if (i < s.size() && s[i] == ',') {
          ++i;
          continue;
        }
        if (i < s.size() && s[i] == ']') {
          ++i;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.type = JsonValue::String;
      return string(v.str);
    }
    if (lit("true")) {
      v.type = JsonValue::Bool;
      v.b = true;
      return true;
    }
    if (lit("false")) {
      v.type = JsonValue::Bool;
      v.b = false;
      return true;
    }
    if (lit("null")) {
      v.type = JsonValue::Null;
      return true;
    }
    v.type = JsonValue::Number;
    return number(v.num);
  }
};

}  // namespace

bool json_parse(const std::string& text, JsonValue& out, std::string& err) {
  JsonReader r(text);
  out = JsonValue();
  if (!r.value(out, 0)) {
    err = r.err;
    return false;
  }
  r.ws();
  if (r.i != text.size()) {
    r.fail("trailing characters");
    err = r.err;
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------------
// typed reads with the reference's leniency (a key of the wrong type is ignored)
// ---------------------------------------------------------------------------------
namespace {

bool read_num(const JsonValue& o, const char* key, double& v) {
  const JsonValue* x = o.get(key);
  if (!x || x->type != JsonValue::Number) return false;
  v = x->num;
  return true;
}
bool read_f(const JsonValue& o, const char* key, float& v) {
  double d;
  if (!read_num(o, key, d)) return false;
  v = float(d);
  return true;
}
bool read_i(const JsonValue& o, const char* key, int& v) {
  double d;
  if (!read_num(o, key, d)) return false;
  v = int(d);
  return true;
}
bool read_b(const JsonValue& o, const char* key, bool& v) {
  const JsonValue* x = o.get(key);
  if (!x || x->type != JsonValue::Bool) return false;
  v = x->b;
  return true;
}
bool read_s(const JsonValue& o, const char* key, std::string& v) {
  const JsonValue* x = o.get(key);
  if (!x || x->type != JsonValue::String) return false;
  v = x->str;
  return true;
}
bool read_arr(const JsonValue& o, const char* key, size_t n, double* v) {
  const JsonValue* x = o.get(key);
  if (!x || x->type != JsonValue::Array || x->arr.size() != n) return false;
  for (size_t k = 0; k < n; ++k)
    if (x->arr[k].type != JsonValue::Number) return false;
  for (size_t k = 0; k < n; ++k) v[k] = x->arr[k].num;
  return true;
}
bool read_f3(const JsonValue& o, const char* key, float* v) {
  double d[3];
  if (!read_arr(o, key, 3, d)) return false;
  for (int k = 0; k < 3; ++k) v[k] = float(d[k]);
  return true;
}

std::string dir_of(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}

bool read_text(const std::string& p, std::string& out) {
  std::ifstream f(p, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// The reference's 3x4 transform record: three axis vectors x, y, z and a translation t
// (Tracer/Utility/LinearMath.h:205-228). Products follow LinearMath.h:245-299:
// (a * b).axis_i[j] = a.axis_i . (b.x[j], b.y[j], b.z[j]), (a * b).t[j] = a.t . (same) + b.t[j].
struct Xf {
  double x[3] = {1, 0, 0}, y[3] = {0, 1, 0}, z[3] = {0, 0, 1}, t[3] = {0, 0, 0};
};

Xf mul(const Xf& a, const Xf& b) {
  Xf r;
  for (int j = 0; j < 3; ++j) {
    const double c[3] = {b.x[j], b.y[j], b.z[j]};
    r.x[j] = a.x[0] * c[0] + a.x[1] * c[1] + a.x[2] * c[2];
    r.y[j] = a.y[0] * c[0] + a.y[1] * c[1] + a.y[2] * c[2];
    r.z[j] = a.z[0] * c[0] + a.z[1] * c[1] + a.z[2] * c[2];
    r.t[j] = a.t[0] * c[0] + a.t[1] * c[1] + a.t[2] * c[2] + b.t[j];
  }
  return r;
}

constexpr double kDegToRad = 3.14159265358979323846 / 180.0;

}  // namespace

bool instance_transform(const JsonValue& ops, float rows[12], std::string& err) {
  Xf m;
  if (ops.type == JsonValue::Array) {
    for (const JsonValue& op : ops.arr) {
      Xf o;
      double v[12], f;
      if (read_arr(op, "matrix", 12, v)) {
        for (int k = 0; k < 3; ++k) {
          o.x[k] = v[k];
          o.y[k] = v[3 + k];
          o.z[k] = v[6 + k];
          o.t[k] = v[9 + k];
        }
      } else if (read_arr(op, "translate", 3, v)) {
        for (int k = 0; k < 3; ++k) o.t[k] = v[k];
      } else if (read_num(op, "rotate-x", f)) {
        const double s = std::sin(f * kDegToRad), c = std::cos(f * kDegToRad);
        o.y[1] = c;
        o.y[2] = s;
        o.z[1] = -s;
        o.z[2] = c;
      } else if (read_num(op, "rotate-y", f)) {
        const double s = std::sin(f * kDegToRad), c = std::cos(f * kDegToRad);
        o.x[0] = c;
        o.x[2] = -s;
        o.z[0] = s;
        o.z[2] = c;
      } else if (read_num(op, "rotate-z", f)) {
        const double s = std::sin(f * kDegToRad), c = std::cos(f * kDegToRad);
        o.x[0] = c;
        o.x[1] = s;
        o.y[0] = -s;
        o.y[1] = c;
      } else if (read_num(op, "scale", f)) {
        o.x[0] = o.y[1] = o.z[2] = f;
      } else if (read_arr(op, "scale", 3, v)) {
        o.x[0] = v[0];
        o.y[1] = v[1];
        o.z[2] = v[2];
      } else {
        continue;  // unknown operation: skipped, as the reference does
      }
      m = mul(m, o);
    }
  }
  // Read back row-major: row i = (axis_i, t[i]), the layout the traced instance uses.
  const double* ax[3] = {m.x, m.y, m.z};
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) rows[r * 4 + k] = float(ax[r][k]);
    rows[r * 4 + 3] = float(m.t[r]);
  }
  for (int k = 0; k < 12; ++k)
    if (!std::isfinite(rows[k])) {
      err = "non-finite instance transform";
      return false;
    }
  return true;
}

// ---------------------------------------------------------------------------------
// OBJ / MTL
// ---------------------------------------------------------------------------------
namespace {

std::vector<std::string> split_ws(const std::string& line) {
  std::vector<std::string> t;
  std::istringstream ss(line);
  std::string w;
  while (ss >> w) t.push_back(w);
  return t;
}

std::string rest_of(const std::string& line, size_t ntok) {
  // text after the first ntok whitespace-separated tokens (file names with spaces)
  size_t i = 0;
  for (size_t k = 0; k < ntok; ++k) {
    while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
    while (i < line.size() && !std::isspace(static_cast<unsigned char>(line[i]))) ++i;
  }
  while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
  size_t e = line.size();
  while (e > i && std::isspace(static_cast<unsigned char>(line[e - 1]))) --e;
  return line.substr(i, e - i);
}

bool parse_float(const std::string& s, float& v) {
  char* end = nullptr;
  v = std::strtof(s.c_str(), &end);
  return end && end != s.c_str() && *end == 0;
}

void load_mtl(const std::string& path, std::vector<FileMaterial>& mats, std::vector<std::string>& warnings) {
  std::ifstream f(path);
  if (!f) {
    warnings.push_back("material library not found: " + path);
    return;
  }
  std::string line;
  FileMaterial* cur = nullptr;
  while (std::getline(f, line)) {
    const std::vector<std::string> t = split_ws(line);
    if (t.empty() || t[0][0] == '#') continue;
    if (t[0] == "newmtl") {
      mats.emplace_back();
      cur = &mats.back();
      cur->name = rest_of(line, 1);
      continue;
    }
    if (!cur) continue;
    float v[3];
    auto f3 = [&](float* dst) {
      if (t.size() >= 4 && parse_float(t[1], v[0]) && parse_float(t[2], v[1]) && parse_float(t[3], v[2]))
        for (int k = 0; k < 3; ++k) dst[k] = v[k];
    };
    if (t[0] == "Kd") f3(cur->diffuse);
    else if (t[0] == "Ke") f3(cur->emissive);
    else if (t[0] == "Pr" && t.size() >= 2) parse_float(t[1], cur->roughness);
    else if (t[0] == "Pm" && t.size() >= 2) parse_float(t[1], cur->metallic);
    else if (t[0].rfind("map_", 0) == 0 || t[0] == "bump" || t[0] == "norm")
      warnings.push_back("texture maps are not supported (" + t[0] + " in " + path + ")");
  }
}

struct PosKey {
  float x, y, z;
  bool operator==(const PosKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct PosHash {
  size_t operator()(const PosKey& k) const {
    uint32_t a, b, c;
    std::memcpy(&a, &k.x, 4);
    std::memcpy(&b, &k.y, 4);
    std::memcpy(&c, &k.z, 4);
    return size_t(a) * 73856093u ^ size_t(b) * 19349663u ^ size_t(c) * 83492791u;
  }
};

}  // namespace

bool load_obj(const std::string& path, FileModel& m, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = "cannot open model " + path;
    return false;
  }
  std::vector<float> v, vn;
  std::vector<std::string> warnings;
  std::vector<FileMaterial> mats;
  std::unordered_map<std::string, uint32_t> mat_index;
  int cur_mat = -1;  // -1: faces before any usemtl use a default material
  bool any_default = false;
  std::vector<int> face_vn;  // per emitted corner: normal index or -1
  std::vector<int> face_mat;
  std::vector<float> corners;  // 9 floats per triangle
  std::string line;
  int lineno = 0;
  const std::string dir = dir_of(path);
  auto resolve = [&](int64_t idx, size_t count, int& out) -> bool {
    if (idx > 0 && size_t(idx) <= count) {
      out = int(idx - 1);
      return true;
    }
    if (idx < 0 && size_t(-idx) <= count) {
      out = int(int64_t(count) + idx);
      return true;
    }
    return false;
  };
  while (std::getline(f, line)) {
    ++lineno;
    const std::vector<std::string> t = split_ws(line);
    if (t.empty() || t[0][0] == '#') continue;
    if (t[0] == "v") {
      float x, y, z;
      if (t.size() < 4 || !parse_float(t[1], x) || !parse_float(t[2], y) || !parse_float(t[3], z)) {
        err = path + ":" + std::to_string(lineno) + ": bad vertex";
        return false;
      }
      v.insert(v.end(), {x, y, z});
    } else if (t[0] == "vn") {
      float x, y, z;
      if (t.size() < 4 || !parse_float(t[1], x) || !parse_float(t[2], y) || !parse_float(t[3], z)) {
        err = path + ":" + std::to_string(lineno) + ": bad normal";
        return false;
      }
      vn.insert(vn.end(), {x, y, z});
    } else if (t[0] == "f") {
      if (t.size() < 4) {
        err = path + ":" + std::to_string(lineno) + ": face with fewer than 3 vertices";
        return false;
      }
      std::vector<int> pi, ni;
      for (size_t k = 1; k < t.size(); ++k) {
        const std::string& w = t[k];
        const size_t s1 = w.find('/');
        int p = -1, n = -1;
        if (!resolve(std::strtoll(w.substr(0, s1).c_str(), nullptr, 10), v.size() / 3, p)) {
          err = path + ":" + std::to_string(lineno) + ": vertex index out of range";
          return false;
        }
        if (s1 != std::string::npos) {
          const size_t s2 = w.find('/', s1 + 1);
          if (s2 != std::string::npos && s2 + 1 < w.size()) {
            if (!resolve(std::strtoll(w.substr(s2 + 1).c_str(), nullptr, 10), vn.size() / 3, n)) {
              err = path + ":" + std::to_string(lineno) + ": normal index out of range";
              return false;
            }
          }
        }
        pi.push_back(p);
        ni.push_back(n);
      }
      if (cur_mat < 0) any_default = true;
      for (size_t k = 1; k + 1 < pi.size(); ++k) {  // fan triangulation
        const int c[3] = {0, int(k), int(k + 1)};
        for (int q = 0; q < 3; ++q) {
          for (int a = 0; a < 3; ++a) corners.push_back(v[size_t(pi[c[q]]) * 3 + a]);
          face_vn.push_back(ni[c[q]]);
        }
        face_mat.push_back(cur_mat);
      }
    } else if (t[0] == "usemtl") {
      const std::string name = rest_of(line, 1);
      auto it = mat_index.find(name);
      if (it == mat_index.end()) {
        it = mat_index.emplace(name, uint32_t(mat_index.size())).first;
      }
      cur_mat = int(it->second);
    } else if (t[0] == "mtllib") {
      const std::string lib = rest_of(line, 1);
      const std::string p1 = dir + lib;
      load_mtl(file_exists(p1) ? p1 : lib, mats, warnings);
    }
  }
  const size_t ntri = face_mat.size();
  if (ntri == 0) {
    err = "model " + path + " has no faces";
    return false;
  }
  // Materials: one per used name (MTL values when defined there, else defaults), plus
  // a default material for faces before any usemtl.
  std::vector<FileMaterial> used(mat_index.size());
  for (const auto& kv : mat_index) {
    FileMaterial fm;
    fm.name = kv.first;
    bool found = false;
    for (const FileMaterial& d : mats)
      if (d.name == kv.first) {
        fm = d;
        found = true;
      }
    if (!found) warnings.push_back("material '" + kv.first + "' not defined in the MTL libraries of " + path);
    used[kv.second] = fm;
  }
  const uint32_t default_id = uint32_t(used.size());
  if (any_default) {
    FileMaterial d;
    d.name = "DefaultMaterial";
    used.push_back(d);
  }
  m.materials = used;
  m.pos = corners;
  m.mat.resize(ntri);
  for (size_t i = 0; i < ntri; ++i) m.mat[i] = face_mat[i] < 0 ? default_id : uint32_t(face_mat[i]);
  // Normals: from the file when all three corners have one; otherwise smooth normals
  // (normalised sum of the unit face normals around each position, as assimp's
  // GenSmoothNormals does after JoinIdenticalVertices; degenerate faces add nothing).
  std::vector<double> fn(ntri * 3, 0.0);
  for (size_t i = 0; i < ntri; ++i) {
    const float* p = &corners[i * 9];
    const double e1[3] = {double(p[3]) - p[0], double(p[4]) - p[1], double(p[5]) - p[2]};
    const double e2[3] = {double(p[6]) - p[0], double(p[7]) - p[1], double(p[8]) - p[2]};
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l > 0)
      for (int k = 0; k < 3; ++k) fn[i * 3 + k] = n[k] / l;
  }
  std::unordered_map<PosKey, size_t, PosHash> pos_id;
  std::vector<double> acc;
  std::vector<size_t> corner_pos(ntri * 3);
  for (size_t c = 0; c < ntri * 3; ++c) {
    const PosKey k{corners[c * 3], corners[c * 3 + 1], corners[c * 3 + 2]};
    auto it = pos_id.find(k);
    if (it == pos_id.end()) {
      it = pos_id.emplace(k, acc.size() / 3).first;
      acc.insert(acc.end(), {0.0, 0.0, 0.0});
    }
    corner_pos[c] = it->second;
    for (int a = 0; a < 3; ++a) acc[it->second * 3 + a] += fn[(c / 3) * 3 + a];
  }
  m.nrm.resize(ntri * 9);
  for (size_t c = 0; c < ntri * 3; ++c) {
    double n[3];
    if (face_vn[c] >= 0) {
      for (int a = 0; a < 3; ++a) n[a] = vn[size_t(face_vn[c]) * 3 + a];
    } else {
      for (int a = 0; a < 3; ++a) n[a] = acc[corner_pos[c] * 3 + a];
    }
    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(l > 0)) {  // fall back to the face normal, then to +z
      for (int a = 0; a < 3; ++a) n[a] = fn[(c / 3) * 3 + a];
      l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (!(l > 0)) {
        n[0] = n[1] = 0.0;
        n[2] = 1.0;
        l = 1.0;
      }
    }
    for (int a = 0; a < 3; ++a) m.nrm[c * 3 + a] = float(n[a] / l);
  }
  m.warnings = warnings;
  return true;
}

// ---------------------------------------------------------------------------------
// scene file
// ---------------------------------------------------------------------------------
namespace {

// Applies a JSON material block to `m`: the reference's property names, lower-cased
// (Tracer/Common/CommonStructs.h:78-96), scalar or RGB as each property takes.
void apply_material(const JsonValue& j, FileMaterial& m, const std::string& where, std::vector<std::string>& warn) {
  read_f3(j, "diffuse", m.diffuse);
  read_f3(j, "emissive", m.emissive);
  struct S {
    const char* key;
    float* dst;
  } scalars[] = {{"roughness", &m.roughness},       {"metallic", &m.metallic},
                 {"specular", &m.specular},         {"speculartint", &m.specular_tint},
                 {"sheen", &m.sheen},               {"sheentint", &m.sheen_tint},
                 {"clearcoat", &m.clearcoat},       {"clearcoatgloss", &m.clearcoat_gloss},
                 {"subsurface", &m.subsurface},     {"anisotropic", &m.anisotropic}};
  for (const S& s : scalars) {
    float v3[3];
    if (!read_f(j, s.key, *s.dst) && read_f3(j, s.key, v3)) *s.dst = v3[0];
  }
  float dv;  // a scalar diffuse / emissive sets all three channels
  if (read_f(j, "diffuse", dv)) m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = dv;
  if (read_f(j, "emissive", dv)) m.emissive[0] = m.emissive[1] = m.emissive[2] = dv;
  for (const auto& kv : j.obj)
    if (kv.first.size() > 3 && kv.first.compare(kv.first.size() - 3, 3, "Map") == 0 &&
        !(kv.second.type == JsonValue::String && kv.second.str.empty()))
      warn.push_back("texture maps are not supported (" + kv.first + " of " + where + ")");
}

}  // namespace

bool load_scene_file(const std::string& path, SceneFile& out, std::string& err) {
  out = SceneFile();
  std::string text;
  if (!read_text(path, text)) {
    err = "cannot open scene file " + path;
    return false;
  }
  JsonValue doc;
  if (!json_parse(text, doc, err)) {
    err = path + ": " + err;
    return false;
  }
  if (doc.type != JsonValue::Object) {
    err = path + ": the document is not a JSON object";
    return false;
  }
  const std::string dir = dir_of(path);

  // models: name -> path
  std::map<std::string, std::string> model_paths;
  if (const JsonValue* ms = doc.get("models"))
    if (ms->type == JsonValue::Array)
      for (const JsonValue& m : ms->arr) {
        std::string name, p;
        if (read_s(m, "name", name) && read_s(m, "path", p)) model_paths[name] = p;
      }

  // instances
  std::map<std::string, uint32_t> model_index;
  if (const JsonValue* is = doc.get("instances"))
    if (is->type == JsonValue::Array)
      for (const JsonValue& ij : is->arr) {
        std::string model;
        if (!read_s(ij, "model", model)) continue;
        auto mp = model_paths.find(model);
        if (mp == model_paths.end()) {
          out.warnings.push_back("model \"" + model + "\" was not declared");
          continue;
        }
        FileInstance inst;
        inst.name = model;
        read_s(ij, "name", inst.name);
        const JsonValue* ops = ij.get("transform");
        if (!instance_transform(ops ? *ops : JsonValue(), inst.rows, err)) {
          err = path + ": instance " + inst.name + ": " + err;
          return false;
        }
        auto mi = model_index.find(model);
        if (mi == model_index.end()) {
          FileModel fm;
          fm.name = model;
          // relative model paths: against the scene file's directory, then its
          // ancestors (the reference's scenes name models relative to the
          // application's working directory, e.g. "../Resources/models/..."), then as given
          const std::string& rel = mp->second;
          fm.path = rel;
          if (!rel.empty() && rel[0] != '/') {
            std::string up;
            for (int k = 0; k < 5; ++k, up += "../")
              if (file_exists(dir + up + rel)) {
                fm.path = dir + up + rel;
                break;
              }
          }
          if (!load_obj(fm.path, fm, err)) {
            err = path + ": " + err;
            return false;
          }
          out.warnings.insert(out.warnings.end(), fm.warnings.begin(), fm.warnings.end());
          mi = model_index.emplace(model, uint32_t(out.models.size())).first;
          out.models.push_back(std::move(fm));
        }
        inst.model = mi->second;
        out.instances.push_back(inst);
      }

  // materials: per model, per material name
  if (const JsonValue* mats = doc.get("materials"))
    if (mats->type == JsonValue::Array)
      for (const JsonValue& mj : mats->arr) {
        std::string model;
        if (!read_s(mj, "model", model)) continue;
        auto mi = model_index.find(model);
        if (mi == model_index.end()) continue;
        FileModel& fm = out.models[mi->second];
        const JsonValue* list = mj.get("materials");
        if (!list || list->type != JsonValue::Array) continue;
        for (const JsonValue& one : list->arr) {
          std::string name;
          if (!read_s(one, "name", name)) continue;
          for (FileMaterial& m : fm.materials)
            if (m.name == name) apply_material(one, m, model + "/" + name, out.warnings);
        }
      }

  // camera (Tracer/FileIO/SceneFile.cpp:89-129)
  if (const JsonValue* cj = doc.get("camera")) {
    out.has_camera = true;
    read_f3(*cj, "position", out.cam_pos);
    read_f3(*cj, "target", out.cam_target);
    read_f3(*cj, "up", out.cam_up);
    read_f(*cj, "aperture", out.aperture);
    read_f(*cj, "distortion", out.distortion);
    read_f(*cj, "focaldist", out.focal_dist);
    read_f(*cj, "fov", out.fov_deg);
    read_f(*cj, "bokehrotation", out.bokeh_rotation);
    int sides;
    if (read_i(*cj, "bokehsides", sides)) out.bokeh_sides = sides;
    out.fov_deg = std::fmin(std::fmax(out.fov_deg, 0.1f), 179.9f);
    if (out.cam_pos[0] == out.cam_target[0] && out.cam_pos[1] == out.cam_target[1] &&
        out.cam_pos[2] == out.cam_target[2])
      out.cam_target[2] = out.cam_pos[2] + 1.f;
  }

  // renderer (Tracer/FileIO/SceneFile.cpp:336-358)
  if (const JsonValue* rj = doc.get("renderer")) {
    int i;
    if (read_i(*rj, "multisample", i) && i > 0) out.spp = uint32_t(i);
    if (read_i(*rj, "maxdepth", i) && i > 0) out.max_depth = uint32_t(i);
    read_f(*rj, "ray-epsilon", out.ray_epsilon);
    read_f(*rj, "aodist", out.ao_dist);
    read_f(*rj, "zdepthmax", out.zdepth_max);
  }

  // sky (Tracer/FileIO/SceneFile.cpp:362-389): a present section resets to defaults first
  if (const JsonValue* sj = doc.get("sky")) {
    SceneFile d;
    out.sky_enabled = d.sky_enabled;
    out.sky_draw_sun = d.sky_draw_sun;
    read_b(*sj, "enabled", out.sky_enabled);
    read_b(*sj, "drawsun", out.sky_draw_sun);
    read_f3(*sj, "dir", out.sky_dir);
    read_f(*sj, "angulardiameter", out.sky_angular_diameter);
    read_f(*sj, "intensity", out.sky_intensity);
    read_f(*sj, "turbidity", out.sky_turbidity);
    read_f(*sj, "bias", out.sky_bias);
  }
  if (out.instances.empty()) {
    err = path + ": no instance refers to a declared, loadable model";
    return false;
  }
  return true;
}

}  // namespace trt
