// lcr/ply_io.hpp -- PLY reading and writing for the C++ facade.
//
// Replaces pcl::io::loadPLYFile / savePLYFileBinary as the reference's tools call
// them (fine_registration.cpp:94-102,129,135; normal_estimation.cpp:79,123).
// Reads ASCII and binary little-endian PLY (the capture tools write ASCII,
// capture/depth_capture/depth_capture.cpp:281-306; PCL tools write binary), maps
// x/y/z, nx/ny/nz or normal_x/normal_y/normal_z, red/green/blue(/alpha) or a packed
// float rgb, and curvature, and skips every other element (faces, PCL's camera
// element).  Writes binary little-endian PLY with PCL's vertex property names and
// order for PointXYZRGBNormal (x y z, red green blue, normal_x normal_y normal_z,
// curvature) followed by a camera element carrying the sensor origin.
#pragma once
#include <stdint.h>
#include <string.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "point_cloud.hpp"

namespace lcr {
namespace io {
namespace detail {

struct Prop {
  std::string name, type, list_count_type;
  bool is_list = false;
};
struct Element {
  std::string name;
  size_t count = 0;
  std::vector<Prop> props;
};

inline size_t type_size(const std::string& t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32")
    return 4;
  if (t == "double" || t == "float64") return 8;
  return 0;
}

inline double read_binary(const unsigned char* p, const std::string& t) {
  if (t == "char" || t == "int8") return (double)*(const int8_t*)p;
  if (t == "uchar" || t == "uint8") return (double)*p;
  int16_t s16;
  uint16_t u16;
  int32_t s32;
  uint32_t u32;
  float f;
  double d;
  if (t == "short" || t == "int16") return memcpy(&s16, p, 2), (double)s16;
  if (t == "ushort" || t == "uint16") return memcpy(&u16, p, 2), (double)u16;
  if (t == "int" || t == "int32") return memcpy(&s32, p, 4), (double)s32;
  if (t == "uint" || t == "uint32") return memcpy(&u32, p, 4), (double)u32;
  if (t == "float" || t == "float32") return memcpy(&f, p, 4), (double)f;
  memcpy(&d, p, 8);
  return d;
}

struct VertexSink {
  std::vector<std::string> names;
  std::vector<double> values;  // one row per vertex
};

inline bool parse(const std::string& path, VertexSink& out, std::vector<float>* rgb_float) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::string line;
  if (!std::getline(f, line) || line.substr(0, 3) != "ply") return false;
  std::string format;
  std::vector<Element> els;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ss(line);
    std::string kw;
    ss >> kw;
    if (kw == "format") {
      ss >> format;
    } else if (kw == "element") {
      Element e;
      ss >> e.name >> e.count;
      els.push_back(e);
    } else if (kw == "property") {
      if (els.empty()) return false;
      Prop p;
      std::string t;
      ss >> t;
      if (t == "list") {
        p.is_list = true;
        ss >> p.list_count_type >> p.type >> p.name;
      } else {
        p.type = t;
        ss >> p.name;
      }
      if (type_size(p.type) == 0) return false;
      els.back().props.push_back(p);
    } else if (kw == "end_header") {
      break;
    }
  }
  if (format != "ascii" && format != "binary_little_endian") return false;
  const bool ascii = format == "ascii";
  for (const Element& e : els) {
    const bool vertex = e.name == "vertex";
    if (vertex) {
      out.names.clear();
      for (const Prop& p : e.props) out.names.push_back(p.is_list ? std::string() : p.name);
      out.values.assign(e.count * e.props.size(), 0.0);
    }
    int rgb_idx = -1;
    if (vertex && rgb_float)
      for (size_t k = 0; k < e.props.size(); ++k)
        if (e.props[k].name == "rgb" && e.props[k].type == "float") rgb_idx = (int)k;
    if (vertex && rgb_idx >= 0) rgb_float->assign(e.count, 0.f);
    for (size_t i = 0; i < e.count; ++i) {
      for (size_t k = 0; k < e.props.size(); ++k) {
        const Prop& p = e.props[k];
        if (ascii) {
          if (p.is_list) {
            double n;
            if (!(f >> n)) return false;
            for (int m = 0; m < (int)n; ++m) {
              double v;
              if (!(f >> v)) return false;
            }
          } else {
            double v;
            if (!(f >> v)) return false;
            if (vertex) out.values[i * e.props.size() + k] = v;
          }
        } else {
          unsigned char buf[8];
          if (p.is_list) {
            const size_t cs = type_size(p.list_count_type);
            if (!f.read((char*)buf, cs)) return false;
            const size_t n = (size_t)read_binary(buf, p.list_count_type);
            f.ignore((std::streamsize)(n * type_size(p.type)));
          } else {
            const size_t s = type_size(p.type);
            if (!f.read((char*)buf, s)) return false;
            if (vertex) {
              out.values[i * e.props.size() + k] = read_binary(buf, p.type);
              if ((int)k == rgb_idx) memcpy(&(*rgb_float)[i], buf, 4);
            }
          }
        }
      }
    }
    if (vertex && !f && !f.eof()) return false;
  }
  return true;
}

inline int index_of(const VertexSink& s, std::initializer_list<const char*> names) {
  for (const char* n : names)
    for (size_t k = 0; k < s.names.size(); ++k)
      if (s.names[k] == n) return (int)k;
  return -1;
}

template <class PointT>
inline void set_xyz(PointT& p, double x, double y, double z) {
  p.x = (float)x;
  p.y = (float)y;
  p.z = (float)z;
}

}  // namespace detail

// Fills xyz, and normals / colour / curvature when the point type has them.
template <class PointT>
int loadPLYFile(const std::string& path, PointCloud<PointT>& cloud) {
  detail::VertexSink s;
  std::vector<float> rgbf;
  if (!detail::parse(path, s, &rgbf)) return -1;
  const int ix = detail::index_of(s, {"x"}), iy = detail::index_of(s, {"y"}),
            iz = detail::index_of(s, {"z"});
  if (ix < 0 || iy < 0 || iz < 0) return -1;
  const int inx = detail::index_of(s, {"normal_x", "nx"}), iny = detail::index_of(s, {"normal_y", "ny"}),
            inz = detail::index_of(s, {"normal_z", "nz"});
  const int ir = detail::index_of(s, {"red", "r"}), ig = detail::index_of(s, {"green", "g"}),
            ib = detail::index_of(s, {"blue", "b"}), ia = detail::index_of(s, {"alpha"});
  const int irgb = detail::index_of(s, {"rgb"});
  const int ic = detail::index_of(s, {"curvature"});
  const size_t np = s.names.size();
  const size_t n = np ? s.values.size() / np : 0;
  cloud.resize(n);
  cloud.is_dense = true;
  for (size_t i = 0; i < n; ++i) {
    const double* v = &s.values[i * np];
    PointT& p = cloud.points[i];
    detail::set_xyz(p, v[ix], v[iy], v[iz]);
    if (!(std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z))) cloud.is_dense = false;
    if constexpr (std::is_same<PointT, PointXYZRGBNormal>::value) {
      if (inx >= 0 && iny >= 0 && inz >= 0) {
        p.normal_x = (float)v[inx];
        p.normal_y = (float)v[iny];
        p.normal_z = (float)v[inz];
      }
      if (ic >= 0) p.curvature = (float)v[ic];
    }
    if constexpr (std::is_same<PointT, PointXYZRGBNormal>::value ||
                  std::is_same<PointT, PointXYZRGB>::value) {
      if (ir >= 0 && ig >= 0 && ib >= 0) {
        p.r = (uint8_t)v[ir];
        p.g = (uint8_t)v[ig];
        p.b = (uint8_t)v[ib];
        p.a = ia >= 0 ? (uint8_t)v[ia] : 255;
      } else if (irgb >= 0 && !rgbf.empty()) {
        memcpy(&p.rgba, &rgbf[i], 4);
      }
    }
  }
  return 0;
}

// Binary little-endian PLY in PCL's property layout for PointXYZRGBNormal.
inline int savePLYFileBinary(const std::string& path, const PointCloud<PointXYZRGBNormal>& cloud) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return -1;
  std::ostringstream h;
  h << "ply\nformat binary_little_endian 1.0\ncomment PCL generated\n"
    << "element vertex " << cloud.size() << "\n"
    << "property float x\nproperty float y\nproperty float z\n"
    << "property uchar red\nproperty uchar green\nproperty uchar blue\n"
    << "property float normal_x\nproperty float normal_y\nproperty float normal_z\n"
    << "property float curvature\n"
    << "element camera 1\n"
    << "property float view_px\nproperty float view_py\nproperty float view_pz\n"
    << "property float x_axisx\nproperty float x_axisy\nproperty float x_axisz\n"
    << "property float y_axisx\nproperty float y_axisy\nproperty float y_axisz\n"
    << "property float z_axisx\nproperty float z_axisy\nproperty float z_axisz\n"
    << "property float focal\nproperty float scalex\nproperty float scaley\n"
    << "property float centerx\nproperty float centery\n"
    << "property int viewportx\nproperty int viewporty\n"
    << "property float k1\nproperty float k2\n"
    << "end_header\n";
  const std::string hs = h.str();
  bool ok = fwrite(hs.data(), 1, hs.size(), f) == hs.size();
  std::vector<unsigned char> rec(3 * 4 + 3 + 4 * 4);
  for (const PointXYZRGBNormal& p : cloud.points) {
    unsigned char* o = rec.data();
    memcpy(o, &p.x, 4);
    memcpy(o + 4, &p.y, 4);
    memcpy(o + 8, &p.z, 4);
    o[12] = p.r;
    o[13] = p.g;
    o[14] = p.b;
    memcpy(o + 15, &p.normal_x, 4);
    memcpy(o + 19, &p.normal_y, 4);
    memcpy(o + 23, &p.normal_z, 4);
    memcpy(o + 27, &p.curvature, 4);
    ok = ok && fwrite(rec.data(), 1, rec.size(), f) == rec.size();
  }
  const float cam_f[17] = {cloud.sensor_origin[0], cloud.sensor_origin[1], cloud.sensor_origin[2],
                           1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0,
                           (float)cloud.width, (float)cloud.height};
  const int32_t vp[2] = {(int32_t)cloud.width, (int32_t)cloud.height};
  const float k[2] = {0.f, 0.f};
  ok = ok && fwrite(cam_f, 4, 17, f) == 17 && fwrite(vp, 4, 2, f) == 2 && fwrite(k, 4, 2, f) == 2;
  ok = (fclose(f) == 0) && ok;
  return ok ? 0 : -1;
}

}  // namespace io
}  // namespace lcr
