// liblisa.so host half: the pieces of LiSA's host code that sit on the
// drop-in boundary -- camera frame, sRGB quantisation, PPM writer and the
// .rto / .obj scene reader.  Plain C++17, no device code.
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
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
#include <thread>
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
// separated by single spaces; the soup is de-indexed per face.  Large files
// are parsed in parallel: chunks split at line ends parse their v / vn / f
// lines independently, then faces are resolved against the concatenated
// vertex lists.  A face may only use vertices defined before it in the file
// (the reference reads vertices_tmp as it grows); an index outside that
// range, a short line or a token that is not a number is an error.
struct ObjChunk {
  const char* begin = nullptr;
  const char* end = nullptr;
  std::vector<float> v, vn;
  std::vector<long> face;       // 6 per face: (vertex, normal) x 3, zero-based
  std::vector<long> v_before;   // v lines of this chunk before each face
  std::vector<long> vn_before;  // vn lines of this chunk before each face
  std::string error;
  long error_line = -1;         // line number inside the chunk
};

// One space-separated token [b, e) as stof / stoi see it.
bool token_float(const char* b, const char* e, float& out) {
  char buf[64];
  std::string big;
  const size_t n = (size_t)(e - b);
  const char* z;
  if (n < sizeof(buf)) {
    std::memcpy(buf, b, n);
    buf[n] = 0;
    z = buf;
  } else {
    big.assign(b, e);
    z = big.c_str();
  }
  char* stop = nullptr;
  errno = 0;
  out = std::strtof(z, &stop);
  return stop != z && errno != ERANGE;
}
bool token_long(const char* b, const char* e, long& out) {
  char buf[32];
  const size_t n = (size_t)(e - b);
  if (n >= sizeof(buf)) return false;
  std::memcpy(buf, b, n);
  buf[n] = 0;
  char* stop = nullptr;
  errno = 0;
  out = std::strtol(buf, &stop, 10);
  return stop != buf && errno != ERANGE && out >= INT32_MIN && out <= INT32_MAX;
}

void parse_chunk(ObjChunk& c) {
  const char* tok_b[8];
  const char* tok_e[8];
  long line_no = 0;
  for (const char* ls = c.begin; ls < c.end; ++line_no) {
    const char* le = static_cast<const char*>(std::memchr(ls, '\n', (size_t)(c.end - ls)));
    if (!le) le = c.end;
    // split on single spaces, keep the first 5 tokens (a trailing `\r` stays
    // in the last token, as in the reference's split)
    int nt = 0;
    const char* t = ls;
    for (;;) {
      const char* sp = static_cast<const char*>(std::memchr(t, ' ', (size_t)(le - t)));
      const char* te = sp ? sp : le;
      if (nt < 8) {
        tok_b[nt] = t;
        tok_e[nt] = te;
      }
      ++nt;
      if (!sp) break;
      t = sp + 1;
    }
    const size_t l0 = (size_t)(tok_e[0] - tok_b[0]);
    auto fail_line = [&](const char* what) {
      if (c.error.empty()) {
        c.error = what;
        c.error_line = line_no;
      }
    };
    if ((l0 == 1 && tok_b[0][0] == 'v') || (l0 == 2 && tok_b[0][0] == 'v' && tok_b[0][1] == 'n')) {
      std::vector<float>& dst = l0 == 1 ? c.v : c.vn;
      if (nt < 4) {
        fail_line("short vertex line");
      } else {
        float x[3];
        bool ok = true;
        for (int i = 0; i < 3; ++i) ok = ok && token_float(tok_b[1 + i], tok_e[1 + i], x[i]);
        if (!ok) fail_line("invalid number");
        dst.insert(dst.end(), x, x + 3);
      }
    } else if (l0 == 1 && tok_b[0][0] == 'f') {
      if (nt < 4) {
        fail_line("face with fewer than 3 vertices");
      } else {
        long idx[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < 3; ++i) {
          // split the token on '/': index 0 = vertex, index 2 = normal
          const char* b = tok_b[1 + i];
          const char* e = tok_e[1 + i];
          const char* s1 = static_cast<const char*>(std::memchr(b, '/', (size_t)(e - b)));
          const char* s2 = s1 ? static_cast<const char*>(std::memchr(s1 + 1, '/', (size_t)(e - s1 - 1))) : nullptr;
          if (!s2) {
            fail_line("face vertex without normal index");
            break;
          }
          const char* s3 = static_cast<const char*>(std::memchr(s2 + 1, '/', (size_t)(e - s2 - 1)));
          long vi, ni;
          if (!token_long(b, s1, vi) || !token_long(s2 + 1, s3 ? s3 : e, ni)) {
            fail_line("invalid face index");
            break;
          }
          idx[2 * i] = vi - 1;
          idx[2 * i + 1] = ni - 1;
        }
        c.face.insert(c.face.end(), idx, idx + 6);
        c.v_before.push_back((long)c.v.size() / 3);
        c.vn_before.push_back((long)c.vn.size() / 3);
      }
    }
    ls = le + 1;
  }
}

int worker_count() {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, std::min(CPU_COUNT(&set), 32));
  return 1;
}

void parse_obj(const std::string& path, lisa_scene_file& f, int32_t mat_idx) {
  std::printf("Importing %s...\n", path.c_str());
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) throw ParseError{LISA_ERR_IO, path + " not found."};
  in.seekg(0, std::ios::end);
  const std::streamoff size = in.tellg();
  in.seekg(0, std::ios::beg);
  std::string data((size_t)std::max<std::streamoff>(size, 0), '\0');
  if (size > 0) in.read(&data[0], size);
  // chunks of >= 4 MiB split at line ends
  const int workers = worker_count();
  const size_t want = std::max<size_t>(1, std::min<size_t>((size_t)workers * 4, data.size() >> 22));
  std::vector<ObjChunk> chunks(want);
  size_t pos = 0;
  for (size_t i = 0; i < want; ++i) {
    size_t stop = i + 1 == want ? data.size() : std::max(pos, data.size() * (i + 1) / want);
    if (stop < data.size()) {
      const size_t nl = data.find('\n', stop);
      stop = nl == std::string::npos ? data.size() : nl + 1;
    }
    chunks[i].begin = data.data() + pos;
    chunks[i].end = data.data() + stop;
    pos = stop;
  }
  {
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    const int nthreads = std::min<int>(workers, (int)chunks.size());
    for (int w = 0; w < nthreads; ++w)
      th.emplace_back([&]() {
        for (size_t i; (i = next.fetch_add(1)) < chunks.size();) parse_chunk(chunks[i]);
      });
    for (auto& t : th) t.join();
  }
  // concatenate vertex lists; offsets per chunk
  std::vector<long> v_off(chunks.size() + 1, 0), vn_off(chunks.size() + 1, 0), f_off(chunks.size() + 1, 0);
  for (size_t i = 0; i < chunks.size(); ++i) {
    if (!chunks[i].error.empty())
      throw ParseError{LISA_ERR_PARSE, "obj: " + chunks[i].error + " in " + path};
    v_off[i + 1] = v_off[i] + (long)chunks[i].v.size() / 3;
    vn_off[i + 1] = vn_off[i] + (long)chunks[i].vn.size() / 3;
    f_off[i + 1] = f_off[i] + (long)chunks[i].face.size() / 6;
  }
  std::vector<float> vt(3 * (size_t)v_off.back()), nt(3 * (size_t)vn_off.back());
  for (size_t i = 0; i < chunks.size(); ++i) {
    std::copy(chunks[i].v.begin(), chunks[i].v.end(), vt.begin() + 3 * v_off[i]);
    std::copy(chunks[i].vn.begin(), chunks[i].vn.end(), nt.begin() + 3 * vn_off[i]);
  }
  const size_t tri0 = f.mat_indices.size();
  const long nb = f_off.back();
  f.vertices.resize(9 * (tri0 + (size_t)nb));
  f.normals.resize(9 * (tri0 + (size_t)nb));
  f.mat_indices.resize(tri0 + (size_t)nb, mat_idx);
  std::vector<char> bad(chunks.size(), 0);
  {
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    const int nthreads = std::min<int>(workers, (int)chunks.size());
    for (int w = 0; w < nthreads; ++w)
      th.emplace_back([&]() {
        for (size_t i; (i = next.fetch_add(1)) < chunks.size();) {
          const ObjChunk& c = chunks[i];
          const size_t nf = c.face.size() / 6;
          for (size_t q = 0; q < nf && !bad[i]; ++q) {
            const long vlim = v_off[i] + c.v_before[q], nlim = vn_off[i] + c.vn_before[q];
            const size_t out = tri0 + (size_t)f_off[i] + q;
            for (int j = 0; j < 3; ++j) {
              const long vi = c.face[6 * q + 2 * j], ni = c.face[6 * q + 2 * j + 1];
              if (vi < 0 || vi >= vlim || ni < 0 || ni >= nlim) {
                bad[i] = 1;
                break;
              }
              std::memcpy(&f.vertices[9 * out + 3 * j], &vt[3 * (size_t)vi], 12);
              std::memcpy(&f.normals[9 * out + 3 * j], &nt[3 * (size_t)ni], 12);
            }
          }
        }
      });
    for (auto& t : th) t.join();
  }
  for (char b : bad)
    if (b) throw ParseError{LISA_ERR_PARSE, "obj: face index out of range in " + path};
  std::printf("Done. Imported %ld triangles.\n", nb);
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
