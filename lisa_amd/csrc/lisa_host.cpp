// liblisa.so host half: the pieces of LiSA's host code that sit on the
// drop-in boundary -- camera frame, sRGB quantisation, PPM writer and the
// .rto / .obj scene reader.  Plain C++17, no device code.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <regex>
#include <sstream>
#include <string>
#include <vector>

#include "lisa.h"
#include "lisa_internal.h"

namespace lisa {
namespace {
thread_local std::string g_error;
}
void set_global_error(const std::string& msg) { g_error = msg; }
lisa_status global_fail(lisa_status code, const std::string& msg) {
  g_error = msg;
  return code;
}
const char* global_error() { return g_error.c_str(); }
}  // namespace lisa

namespace {

using lisa::global_fail;

// ---------------------------------------------------------------------------
// Camera: sutil::Camera::UVWFrame (src/sutil/Camera.cpp:34-45).
// ---------------------------------------------------------------------------
struct H3 {
  float x, y, z;
};
inline H3 sub(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline H3 scale(H3 a, float s) { return H3{a.x * s, a.y * s, a.z * s}; }
inline float hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline H3 hcross(H3 a, H3 b) {
  return H3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline H3 hnormalize(H3 v) {
  const float inv = 1.0f / std::sqrt(hdot(v, v));
  return scale(v, inv);
}

// src/cuda/helpers.h:35-55.
inline float to_srgb(float c) {
  const float p = std::pow(c, 1.0f / 2.4f);
  return c < 0.0031308f ? 12.92f * c : 1.055f * p - 0.055f;
}
inline float clampf(float f, float a, float b) { return std::fmax(a, std::fmin(f, b)); }
inline uint8_t quant8(float x) {
  x = clampf(x, 0.0f, 1.0f);
  const unsigned v = (unsigned)(x * 256.0f);
  return (uint8_t)(v < 255u ? v : 255u);
}

}  // namespace

extern "C" {

lisa_status lisa_camera_uvw(const float eye[3], const float look_at[3], float fov, float aspect,
                            float U[3], float V[3], float W[3]) {
  if (!eye || !look_at || !U || !V || !W)
    return global_fail(LISA_ERR_INVALID_ARG, "lisa_camera_uvw: NULL argument");
  const H3 e{eye[0], eye[1], eye[2]}, l{look_at[0], look_at[1], look_at[2]}, up{0.0f, 1.0f, 0.0f};
  const H3 w = sub(l, e);
  const float wlen = std::sqrt(hdot(w, w));
  H3 u = hnormalize(hcross(w, up));
  H3 v = hnormalize(hcross(u, w));
  const float vlen = wlen * std::tan(0.5f * fov * 3.14159265358979323846f / 180.0f);
  v = scale(v, vlen);
  const float ulen = vlen * aspect;
  u = scale(u, ulen);
  U[0] = u.x, U[1] = u.y, U[2] = u.z;
  V[0] = v.x, V[1] = v.y, V[2] = v.z;
  W[0] = w.x, W[1] = w.y, W[2] = w.z;
  return LISA_OK;
}

lisa_status lisa_srgb_quantize(const float* rgba, int64_t n, uint8_t* rgba8) {
  if (n < 0 || (n > 0 && (!rgba || !rgba8)))
    return global_fail(LISA_ERR_INVALID_ARG, "lisa_srgb_quantize: bad arguments");
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) rgba8[4 * i + c] = quant8(to_srgb(clampf(rgba[4 * i + c], 0.0f, 1.0f)));
    rgba8[4 * i + 3] = 255u;
  }
  return LISA_OK;
}

// sutil::saveImage, UNSIGNED_BYTE4 branch (src/sutil/sutil.cpp:523-554) and
// savePPM (97-117).
lisa_status lisa_save_ppm(const char* path, const uint8_t* rgba8, int32_t width, int32_t height) {
  if (!path) return global_fail(LISA_ERR_INVALID_ARG, "lisa_save_ppm: path is NULL");
  const std::string fn(path);
  if (fn.size() < 5)
    return global_fail(LISA_ERR_INVALID_ARG, "sutil::saveImage(): Failed to determine filename extension");
  const std::string ext = fn.substr(fn.size() - 3);
  if (ext != "PPM" && ext != "ppm")
    return global_fail(LISA_ERR_UNSUPPORTED, "saveImage: only .ppm output is supported: " + fn);
  if (!rgba8 || width < 1 || height < 1)
    return global_fail(LISA_ERR_INVALID_ARG, "savePPM: Image is ill-formed. Not saving");
  std::vector<unsigned char> pix(3 * (size_t)width * height);
  for (int j = height - 1; j >= 0; --j)
    for (int i = 0; i < width; ++i) {
      const size_t dst = 3 * (size_t)width * (height - j - 1) + 3 * (size_t)i;
      const size_t src = 4 * (size_t)width * j + 4 * (size_t)i;
      pix[dst] = rgba8[src];
      pix[dst + 1] = rgba8[src + 1];
      pix[dst + 2] = rgba8[src + 2];
    }
  std::ofstream out(fn, std::ios::out | std::ios::binary);
  if (!out.is_open()) return global_fail(LISA_ERR_IO, "savePPM: Could not open file for " + fn);
  out << "P6\n" << width << " " << height << "\n" << 255 << "\n";
  out.write(reinterpret_cast<const char*>(pix.data()), (std::streamsize)pix.size());
  out.close();
  if (!out) return global_fail(LISA_ERR_IO, "savePPM: write failed for " + fn);
  return LISA_OK;
}

}  // extern "C"

// ===========================================================================
// .rto reader: the behaviour of SceneParser (src/LiSA/src/scene_parser.cc,
// include/scene_parser.hh) and parse_obj (src/LiSA/src/parse_obj.cc).  The
// grammar is defined by the reference's regular expressions, so the same
// expressions are used here (std::regex, ECMAScript) to accept and reject
// exactly the same files; exit(1) becomes LISA_ERR_PARSE with the
// reference's message.
// ===========================================================================
struct lisa_scene_file {
  std::vector<float> vertices, normals;  // 9 per triangle
  std::vector<int32_t> mat_indices;
  std::vector<lisa_material> materials;
  uint32_t width = 0, height = 0, num_samples = 0, num_bounces = 0;
  float eye[3] = {0, 0, 0}, look_at[3] = {0, 0, 0}, fov = 0.0f;
  std::string output_image;
};

namespace {

struct ParseError {
  lisa_status code;
  std::string msg;
};

const std::string kVar = "([a-zA-Z0-9]|_)+";                // scene_parser.hh:43
const std::string kPath = "([a-zA-Z0-9]|_|\\/|.)+";         // scene_parser.hh:44
const std::string kFloat = "\\s*([+-]?([0-9]*[.])?[0-9]+)\\s*";  // scene_parser.cc:71

std::vector<std::string> all_matches(const std::string& s, const std::regex& r) {
  std::vector<std::string> out;
  for (auto it = std::sregex_iterator(s.begin(), s.end(), r); it != std::sregex_iterator(); ++it)
    out.push_back((*it)[0]);
  return out;
}

std::regex vector_rgx(const std::string& key, int n) {  // scene_parser.cc:66-79
  std::string r = key + "\\s*=\\s*\\(";
  for (int i = 0; i < n - 1; ++i) r += kFloat + ",";
  r += kFloat + "\\)";
  return std::regex(r);
}
std::regex float_rgx(const std::string& key) {  // scene_parser.cc:81-83
  return std::regex(key + "\\s*=\\s*([+-]?([0-9]*[.])?[0-9]+)");
}

// std::stof / std::stoi semantics: leading whitespace skipped, longest valid
// prefix converted, failure when nothing converts or the value is out of range.
float to_float(const std::string& s) {
  const char* b = s.c_str();
  char* e = nullptr;
  errno = 0;
  const float v = std::strtof(b, &e);
  if (e == b || errno == ERANGE) throw ParseError{LISA_ERR_PARSE, "stof: invalid number '" + s + "'"};
  return v;
}
int to_int(const std::string& s) {
  const char* b = s.c_str();
  char* e = nullptr;
  errno = 0;
  const long v = std::strtol(b, &e, 10);
  if (e == b || errno == ERANGE || v < INT32_MIN || v > INT32_MAX)
    throw ParseError{LISA_ERR_PARSE, "stoi: invalid integer '" + s + "'"};
  return (int)v;
}

std::string search_param(const std::string& file, const std::string& param, bool is_string) {
  std::smatch m;
  const std::regex r = is_string ? std::regex(param + "\\s*=\\s*(" + kPath + ")")
                                 : std::regex(param + "\\s*=\\s*([0-9]+)");
  if (std::regex_search(file, m, r)) return m[1];
  throw ParseError{LISA_ERR_PARSE, "Param " + param + " not found."};
}

// parse_obj (parse_obj.cc:24-69): `v` / `vn` lines, faces `f a/b/c ...` with the
// normal index required, first three face vertices, 1-based indices, tokens
// separated by single spaces; the soup is de-indexed per face.
void parse_obj(const std::string& path, lisa_scene_file& f, int32_t mat_idx) {
  std::printf("Importing %s...\n", path.c_str());
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) throw ParseError{LISA_ERR_IO, path + " not found."};
  std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  std::vector<float> vt, nt;
  std::vector<std::string> tok, idx;
  auto split = [](const std::string& txt, char ch, std::vector<std::string>& out) {
    out.clear();
    size_t start = 0, pos;
    while ((pos = txt.find(ch, start)) != std::string::npos) {
      out.push_back(txt.substr(start, pos - start));
      start = pos + 1;
    }
    out.push_back(txt.substr(start));
  };
  int nb = 0;
  size_t ls = 0;
  std::string line;
  while (ls < data.size()) {
    size_t le = data.find('\n', ls);
    if (le == std::string::npos) le = data.size();
    line.assign(data, ls, le - ls);
    ls = le + 1;
    split(line, ' ', tok);
    if (tok[0] == "v" || tok[0] == "vn") {
      if (tok.size() < 4) throw ParseError{LISA_ERR_PARSE, "obj: short vertex line in " + path};
      std::vector<float>& dst = tok[0] == "v" ? vt : nt;
      for (int c = 1; c < 4; ++c) dst.push_back(to_float(tok[c]));
    } else if (tok[0] == "f") {
      if (tok.size() < 4) throw ParseError{LISA_ERR_PARSE, "obj: face with fewer than 3 vertices in " + path};
      for (int i = 1; i < 4; ++i) {
        split(tok[i], '/', idx);
        if (idx.size() < 3) throw ParseError{LISA_ERR_PARSE, "obj: face vertex without normal index in " + path};
        const long vi = (long)to_int(idx[0]) - 1, ni = (long)to_int(idx[2]) - 1;
        if (vi < 0 || 3 * (size_t)vi + 2 >= vt.size() || ni < 0 || 3 * (size_t)ni + 2 >= nt.size())
          throw ParseError{LISA_ERR_PARSE, "obj: face index out of range in " + path};
        for (int c = 0; c < 3; ++c) f.vertices.push_back(vt[3 * vi + c]);
        for (int c = 0; c < 3; ++c) f.normals.push_back(nt[3 * ni + c]);
      }
      f.mat_indices.push_back(mat_idx);
      ++nb;
    }
  }
  std::printf("Done. Imported %d triangles.\n", nb);
}

void parse_rto(const char* path, lisa_scene_file& f) {
  std::ifstream in(path);
  if (!in.is_open()) throw ParseError{LISA_ERR_IO, std::string(path) + " not found"};
  std::string file, line;
  while (std::getline(in, line)) file += line + "\n";
  {  // remove /* ... */ comments (scene_parser.cc:59-64)
    const std::regex comments("\\/\\*((.|\n)*?)\\*\\/");
    file = std::regex_replace(file, comments, "");
  }
  f.width = (uint32_t)to_int(search_param(file, "width", false));
  f.height = (uint32_t)to_int(search_param(file, "height", false));
  f.num_samples = (uint32_t)to_int(search_param(file, "num_samples", false));
  f.num_bounces = (uint32_t)to_int(search_param(file, "num_bounces", false));
  f.output_image = search_param(file, "output_image", true);

  // materials (scene_parser.cc:106-163)
  std::map<std::string, int> name_idx;
  {
    const auto mats = all_matches(file, std::regex("material\\s+" + kVar + "+\\s*\\{\n*[^\\}]*"));
    auto err = [](const std::string& e) {
      throw ParseError{LISA_ERR_PARSE, "Error in materials declaration: " + e};
    };
    if (mats.empty()) err("no valid materials declared.");
    for (const std::string& s : mats) {
      std::smatch m;
      std::string name;
      if (std::regex_search(s, m, std::regex("\\s+(" + kVar + ")\\s*\\{"))) name = m[1];
      const bool is_light = std::regex_search(s, m, std::regex("(emit\\s*=\\s*true)"));
      float col[3], alpha;
      if (std::regex_search(s, m, vector_rgx("color", 4))) {
        col[0] = to_float(m[1]);
        col[1] = to_float(m[3]);
        col[2] = to_float(m[5]);
        alpha = to_float(m[7]);
        if (col[0] > 1 || col[1] > 1 || col[2] > 1)
          for (float& c : col) c = (float)((double)c / 255.0);
      } else {
        err("no valid color provided in material " + name + ".");
      }
      lisa_material mt;
      std::memset(&mt, 0, sizeof(mt));
      if (is_light) {  // mk_material_emit (structs.hh:25-31)
        mt.emit = 1;
        for (int c = 0; c < 3; ++c) mt.emission[c] = col[c];
        mt.alpha = 1.0f;
      } else {  // mk_material_diffuse (structs.hh:41-52)
        if (!std::regex_search(s, m, float_rgx("(roughness | n)")))
          err("no valid roughness|refractive index in material " + name + ".");
        const float param = to_float(m[2]);
        if (alpha < 1.0f)
          mt.n = param;
        else
          mt.roughness = param;
        mt.alpha = alpha;
        for (int c = 0; c < 3; ++c) mt.diffuse[c] = col[c];
        mt.emit = 0;
      }
      if (name_idx.find(name) == name_idx.end()) {
        f.materials.push_back(mt);
        name_idx[name] = (int)f.materials.size() - 1;
      }
    }
  }
  // meshes (scene_parser.cc:165-198)
  {
    const auto meshes = all_matches(file, std::regex("mesh\\s*\\{\n*[^\\}]*"));
    auto err = [](const std::string& e) {
      throw ParseError{LISA_ERR_PARSE, "Error in mesh declaration: " + e};
    };
    for (const std::string& s : meshes) {
      std::smatch m;
      int mat_idx = 0;
      if (std::regex_search(s, m, std::regex("material\\s*=\\s*(" + kVar + ")"))) {
        auto it = name_idx.find(m[1]);
        if (it == name_idx.end()) err(std::string(m[1]) + " material not found");
        mat_idx = it->second;
      } else {
        err("no valid material provided");
      }
      if (std::regex_search(s, m, std::regex("obj_file\\s*=\\s*(.+\\.obj)")))
        parse_obj(m[1], f, mat_idx);
      else
        err("no valid obj_file provided");
    }
  }
  // camera (scene_parser.cc:200-245)
  {
    const auto cams = all_matches(file, std::regex("camera\\s*\\{\n*[^\\}]*"));
    auto err = [](const std::string& e) {
      throw ParseError{LISA_ERR_PARSE, "Error in camera declaration: " + e};
    };
    if (cams.size() != 1) err("no valid camera declared.");
    const std::string& s = cams[0];
    std::smatch m;
    if (std::regex_search(s, m, vector_rgx("position", 3))) {
      f.eye[0] = to_float(m[1]);
      f.eye[1] = to_float(m[3]);
      f.eye[2] = to_float(m[5]);
    } else {
      err("no valid position provided.");
    }
    if (std::regex_search(s, m, vector_rgx("look_at", 3))) {
      f.look_at[0] = to_float(m[1]);
      f.look_at[1] = to_float(m[3]);
      f.look_at[2] = to_float(m[5]);
    } else {
      err("no valid look_at provided.");
    }
    if (std::regex_search(s, m, float_rgx("fov")))
      f.fov = to_float(m[1]);
    else
      err("no valid fov provided.");
  }
}

}  // namespace

extern "C" {

lisa_status lisa_scene_file_load(const char* path, lisa_scene_file** out) {
  if (!path || !out) return global_fail(LISA_ERR_INVALID_ARG, "lisa_scene_file_load: NULL argument");
  *out = nullptr;
  lisa_scene_file* f = new lisa_scene_file;
  try {
    parse_rto(path, *f);
  } catch (const ParseError& e) {
    delete f;
    return global_fail(e.code, e.msg);
  } catch (const std::exception& e) {
    delete f;
    return global_fail(LISA_ERR_PARSE, std::string("scene parse failed: ") + e.what());
  }
  *out = f;
  return LISA_OK;
}

void lisa_scene_file_free(lisa_scene_file* f) { delete f; }

lisa_status lisa_scene_file_info(const lisa_scene_file* f, lisa_scene_info* info) {
  if (!f || !info) return global_fail(LISA_ERR_INVALID_ARG, "lisa_scene_file_info: NULL argument");
  info->num_tris = (int32_t)f->mat_indices.size();
  info->num_materials = (int32_t)f->materials.size();
  info->width = f->width;
  info->height = f->height;
  info->num_samples = f->num_samples;
  info->num_bounces = f->num_bounces;
  for (int c = 0; c < 3; ++c) {
    info->eye[c] = f->eye[c];
    info->look_at[c] = f->look_at[c];
  }
  info->fov = f->fov;
  return LISA_OK;
}

const float* lisa_scene_file_vertices(const lisa_scene_file* f) { return f ? f->vertices.data() : nullptr; }
const float* lisa_scene_file_normals(const lisa_scene_file* f) { return f ? f->normals.data() : nullptr; }
const int32_t* lisa_scene_file_mat_indices(const lisa_scene_file* f) {
  return f ? f->mat_indices.data() : nullptr;
}
const lisa_material* lisa_scene_file_materials(const lisa_scene_file* f) {
  return f ? f->materials.data() : nullptr;
}
const char* lisa_scene_file_output_image(const lisa_scene_file* f) {
  return f ? f->output_image.c_str() : nullptr;
}

}  // extern "C"
