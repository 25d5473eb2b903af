// Scene I/O without nlohmann::json and nibabel: a small JSON reader for the
// three scene keys and a NIfTI-1 (.nii / .nii.gz) reader via zlib.
#include "scene.h"

#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace mdx {

// ---------------------------------------------------------------- JSON
namespace {

struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  double num = 0;
  bool b = false;
  std::string str;
  std::vector<JVal> arr;
  std::map<std::string, JVal> obj;
};

struct JParser {
  const std::string& s;
  size_t i = 0;
  explicit JParser(const std::string& src) : s(src) {}
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  [[noreturn]] void fail(const std::string& what) {
    throw std::runtime_error("Error parsing JSON file: " + what + " at offset " + std::to_string(i));
  }
  JVal value() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    char c = s[i];
    JVal v;
    if (c == '{') {
      v.kind = JVal::Obj;
      ++i;
      ws();
      if (i < s.size() && s[i] == '}') { ++i; return v; }
      for (;;) {
        ws();
        JVal k = value();
        if (k.kind != JVal::Str) fail("object key is not a string");
        ws();
        if (i >= s.size() || s[i] != ':') fail("expected ':'");
        ++i;
        v.obj[k.str] = value();
        ws();
        if (i < s.size() && s[i] == ',') { ++i; continue; }
        if (i < s.size() && s[i] == '}') { ++i; break; }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind = JVal::Arr;
      ++i;
      ws();
      if (i < s.size() && s[i] == ']') { ++i; return v; }
      for (;;) {
        v.arr.push_back(value());
        ws();
        if (i < s.size() && s[i] == ',') { ++i; continue; }
        if (i < s.size() && s[i] == ']') { ++i; break; }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = JVal::Str;
      ++i;
      while (i < s.size() && s[i] != '"') {
        if (s[i] == '\\' && i + 1 < s.size()) {
          char e = s[i + 1];
          v.str.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
          i += 2;
        } else {
          v.str.push_back(s[i++]);
        }
      }
      if (i >= s.size()) fail("unterminated string");
      ++i;
    } else if (c == 't' && s.compare(i, 4, "true") == 0) {
      v.kind = JVal::Bool; v.b = true; i += 4;
    } else if (c == 'f' && s.compare(i, 5, "false") == 0) {
      v.kind = JVal::Bool; i += 5;
    } else if (c == 'n' && s.compare(i, 4, "null") == 0) {
      i += 4;
    } else {
      char* end = nullptr;
      v.num = std::strtod(s.c_str() + i, &end);
      if (end == s.c_str() + i) fail("unexpected character");
      v.kind = JVal::Num;
      i = end - s.c_str();
    }
    return v;
  }
};

std::array<double, 3> vec3(const JVal& v, const char* key) {
  if (v.kind != JVal::Arr || v.arr.size() != 3) throw std::runtime_error(std::string("scene key '") + key + "' must be a 3-vector");
  std::array<double, 3> o{};
  for (int k = 0; k < 3; ++k) {
    if (v.arr[k].kind != JVal::Num) throw std::runtime_error(std::string("scene key '") + key + "' must hold numbers");
    o[k] = v.arr[k].num;
  }
  return o;
}

}  // namespace

Scene read_scene(const std::string& json_path) {
  namespace fs = std::filesystem;
  if (json_path.find(".json") == std::string::npos) throw std::runtime_error("The file provided is not a JSON file.");
  fs::path abs = fs::absolute(json_path);
  if (!fs::exists(abs) || !fs::is_regular_file(abs)) throw std::runtime_error("JSON file does not exist or is not accessible.");
  std::ifstream f(json_path);
  if (!f.is_open()) throw std::runtime_error("Failed to open JSON file.");
  std::stringstream ss;
  ss << f.rdbuf();
  std::string src = ss.str();
  JParser p(src);
  JVal root = p.value();
  if (root.kind != JVal::Obj) throw std::runtime_error("Error parsing JSON file: top level is not an object");
  Scene sc;
  auto it = root.obj.find("dim_space");
  if (it == root.obj.end()) throw std::runtime_error("scene JSON lacks 'dim_space'");
  sc.dim_space = vec3(it->second, "dim_space");
  it = root.obj.find("background_material_name");
  if (it == root.obj.end() || it->second.kind != JVal::Str) throw std::runtime_error("scene JSON lacks 'background_material_name'");
  sc.background_material_name = it->second.str;
  it = root.obj.find("voxel_grids");
  std::string dir = abs.parent_path().string();
  if (it != root.obj.end()) {
    if (it->second.kind != JVal::Arr) throw std::runtime_error("'voxel_grids' must be an array");
    for (const JVal& g : it->second.arr) {
      auto fp = g.obj.find("file_path");
      auto org = g.obj.find("origin");
      if (fp == g.obj.end() || fp->second.kind != JVal::Str || org == g.obj.end())
        throw std::runtime_error("voxel grid entry needs 'file_path' and 'origin'");
      std::string path = fp->second.str;
      if (path.find(".nii") == std::string::npos) throw std::runtime_error("The file provided is not a NIFTI file.");
      SceneGrid sg;
      sg.path = fs::path(path).is_absolute() ? path : dir + "/" + path;
      sg.origin = vec3(org->second, "origin");
      sc.grids.push_back(sg);
    }
  }
  return sc;
}

// ---------------------------------------------------------------- NIfTI-1
NiftiGrid read_nifti(const std::string& path) {
  gzFile g = gzopen(path.c_str(), "rb");
  if (!g) throw std::runtime_error("cannot open NIfTI file " + path);
  std::vector<unsigned char> buf;
  unsigned char chunk[1 << 16];
  int got;
  while ((got = gzread(g, chunk, sizeof chunk)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  gzclose(g);
  if (buf.size() < 348) throw std::runtime_error("NIfTI file too short: " + path);
  int32_t sizeof_hdr;
  std::memcpy(&sizeof_hdr, buf.data(), 4);
  if (sizeof_hdr != 348) throw std::runtime_error("not a little-endian NIfTI-1 file: " + path);
  int16_t dim[8], datatype, bitpix;
  float pixdim[8], vox_offset, slope, inter;
  std::memcpy(dim, buf.data() + 40, 16);
  std::memcpy(&datatype, buf.data() + 70, 2);
  std::memcpy(&bitpix, buf.data() + 72, 2);
  std::memcpy(pixdim, buf.data() + 76, 32);
  std::memcpy(&vox_offset, buf.data() + 108, 4);
  std::memcpy(&slope, buf.data() + 112, 4);
  std::memcpy(&inter, buf.data() + 116, 4);
  unsigned char units = buf[123];
  if (dim[0] < 3) throw std::runtime_error("NIfTI grid must have 3 dimensions: " + path);
  for (int k = 4; k <= dim[0] && k < 8; ++k)
    if (dim[k] > 1) throw std::runtime_error("NIfTI grid has more than 3 non-trivial dimensions: " + path);
  NiftiGrid out;
  double to_cm;
  switch (units & 0x07) {  // NIFTI_UNITS_METER / MM / MICRON
    case 1: to_cm = 100.0; break;
    case 2: to_cm = 0.1; break;
    case 3: to_cm = 1e-4; break;
    default: throw std::runtime_error("NIfTI grid has no spatial unit: " + path);
  }
  for (int k = 0; k < 3; ++k) {
    out.dims[k] = dim[k + 1];
    out.spacing[k] = (double)std::fabs(pixdim[k + 1]) * to_cm;
    if (out.dims[k] <= 0 || !(out.spacing[k] > 0)) throw std::runtime_error("bad NIfTI dims / spacing: " + path);
  }
  size_t n = (size_t)out.dims[0] * out.dims[1] * out.dims[2];
  size_t off = (size_t)vox_offset;
  if (off < 348) off = 352;
  size_t bpv = (size_t)bitpix / 8;
  if (buf.size() < off + n * bpv) throw std::runtime_error("NIfTI data truncated: " + path);
  const unsigned char* d = buf.data() + off;
  bool scale = std::isfinite(slope) && slope != 0.0f && !(slope == 1.0f && inter == 0.0f);
  out.ids.resize(n);
  for (size_t i = 0; i < n; ++i) {
    double v;
    switch (datatype) {
      case 2: v = d[i]; break;
      case 256: v = (int8_t)d[i]; break;
      case 4: { int16_t x; std::memcpy(&x, d + 2 * i, 2); v = x; break; }
      case 512: { uint16_t x; std::memcpy(&x, d + 2 * i, 2); v = x; break; }
      case 8: { int32_t x; std::memcpy(&x, d + 4 * i, 4); v = x; break; }
      case 768: { uint32_t x; std::memcpy(&x, d + 4 * i, 4); v = x; break; }
      case 16: { float x; std::memcpy(&x, d + 4 * i, 4); v = x; break; }
      case 64: { double x; std::memcpy(&x, d + 8 * i, 8); v = x; break; }
      default: throw std::runtime_error("unsupported NIfTI datatype " + std::to_string(datatype) + ": " + path);
    }
    if (scale) v = v * slope + inter;
    if (!(v >= 0.0 && v < 256.0)) throw std::out_of_range("NIfTI voxel value outside the uint8 material-ID range: " + path);
    out.ids[i] = (uint8_t)v;
  }
  return out;
}

}  // namespace mdx
