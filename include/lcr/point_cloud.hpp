// lcr/point_cloud.hpp -- point types and clouds for the C++ facade of liblcr.
//
// Mirrors the parts of pcl::PointXYZRGBNormal / pcl::PointCloud that the
// reference's fine-registration and normal-estimation tools use
// (pcl_tools/fine_registration.cpp:20-21,89-91,131-134;
//  pcl_tools/normal_estimation.cpp:20-22,78,121-122): the 48-byte point record with
// xyz, normal and rgb/curvature blocks, a shared-pointer cloud type with size(),
// points, width/height and concatenation.  The layout is kept so a caller can
// pass the points array straight to the C ABI with a 48-byte stride.
#pragma once
#include <stdint.h>

#include <cmath>
#include <iomanip>
#include <limits>
#include <memory>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

namespace lcr {

struct PointXYZ {
  float x = 0.f, y = 0.f, z = 0.f, pad_ = 1.f;
};

struct PointXYZRGB {
  float x = 0.f, y = 0.f, z = 0.f, pad_ = 1.f;
  union {
    struct {
      uint8_t b, g, r, a;
    };
    uint32_t rgba = 0xff000000u;
  };
  float pad2_[3] = {0.f, 0.f, 0.f};
};

struct Normal {
  float normal_x = 0.f, normal_y = 0.f, normal_z = 0.f, pad_ = 0.f;
  float curvature = 0.f, pad2_[3] = {0.f, 0.f, 0.f};
};

// 48 bytes: [x y z 1] [nx ny nz 0] [rgba curvature 0 0], like PCL's.
struct PointXYZRGBNormal {
  float x = 0.f, y = 0.f, z = 0.f, pad_ = 1.f;
  float normal_x = 0.f, normal_y = 0.f, normal_z = 0.f, pad2_ = 0.f;
  union {
    struct {
      uint8_t b, g, r, a;
    };
    uint32_t rgba = 0xff000000u;
  };
  float curvature = 0.f;
  float pad3_[2] = {0.f, 0.f};
};
static_assert(sizeof(PointXYZRGBNormal) == 48, "PCL-compatible 48-byte record");

template <class PointT>
struct PointCloud {
  using Ptr = std::shared_ptr<PointCloud<PointT>>;
  using ConstPtr = std::shared_ptr<const PointCloud<PointT>>;
  std::vector<PointT> points;
  uint32_t width = 0, height = 1;
  bool is_dense = true;
  float sensor_origin[3] = {0.f, 0.f, 0.f};

  size_t size() const { return points.size(); }
  bool empty() const { return points.empty(); }
  void resize(size_t n) {
    points.resize(n);
    width = (uint32_t)n;
    height = 1;
  }
  void push_back(const PointT& p) {
    points.push_back(p);
    width = (uint32_t)points.size();
    height = 1;
  }
  PointCloud& operator+=(const PointCloud& o) {
    points.insert(points.end(), o.points.begin(), o.points.end());
    width = (uint32_t)points.size();
    height = 1;
    is_dense = is_dense && o.is_dense;
    return *this;
  }
  PointCloud operator+(const PointCloud& o) const {
    PointCloud r = *this;
    r += o;
    return r;
  }
};

// copyPointCloud between clouds of the same point type (fine_registration.cpp:133)
template <class PointT>
void copyPointCloud(const PointCloud<PointT>& in, PointCloud<PointT>& out) {
  out = in;
}

// concatenateFields(PointXYZRGB, Normal) -> PointXYZRGBNormal (normal_estimation.cpp:122)
inline void concatenateFields(const PointCloud<PointXYZRGB>& a, const PointCloud<Normal>& b,
                              PointCloud<PointXYZRGBNormal>& out) {
  out.resize(a.size());
  for (size_t i = 0; i < a.size(); ++i) {
    PointXYZRGBNormal& o = out.points[i];
    o.x = a.points[i].x;
    o.y = a.points[i].y;
    o.z = a.points[i].z;
    o.rgba = a.points[i].rgba;
    o.normal_x = b.points[i].normal_x;
    o.normal_y = b.points[i].normal_y;
    o.normal_z = b.points[i].normal_z;
    o.curvature = b.points[i].curvature;
  }
  out.is_dense = a.is_dense && b.is_dense;
}

// Row-major 4x4 with Eigen's accessors and stream format (the matrix printed at
// fine_registration.cpp:127).
template <class Scalar>
struct Matrix4 {
  Scalar m[16];
  Matrix4() { setIdentity(); }
  static Matrix4 Identity() { return Matrix4(); }
  void setIdentity() {
    for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? Scalar(1) : Scalar(0);
  }
  Scalar& operator()(int r, int c) { return m[4 * r + c]; }
  Scalar operator()(int r, int c) const { return m[4 * r + c]; }
  Scalar coeff(int r, int c) const { return m[4 * r + c]; }
  const Scalar* data() const { return m; }
};
using Matrix4f = Matrix4<float>;
using Matrix4d = Matrix4<double>;

// Eigen's default IOFormat: every coefficient printed at the stream's precision,
// right-aligned to the widest one, separated by one space, one row per line.
// The cloud summary the reference's filter tools print with std::cout << *cloud
// (cloud_downsampling.cpp:69-78): header, point count, width, height, density,
// sensor origin and orientation.
template <class PointT>
std::ostream& operator<<(std::ostream& os, const PointCloud<PointT>& c) {
  os << "header: seq: 0 stamp: 0 frame_id: \n";
  os << "points[]: " << c.points.size() << "\n";
  os << "width: " << c.width << "\n";
  os << "height: " << c.height << "\n";
  os << "is_dense: " << (c.is_dense ? 1 : 0) << "\n";
  os << "sensor origin (xyz): [" << c.sensor_origin[0] << ", " << c.sensor_origin[1] << ", "
     << c.sensor_origin[2] << "] / orientation (xyzw): [0, 0, 0, 1]\n";
  return os;
}

template <class Scalar>
std::ostream& operator<<(std::ostream& os, const Matrix4<Scalar>& M) {
  std::streamsize width = 0;
  for (int i = 0; i < 16; ++i) {
    std::ostringstream s;
    s.copyfmt(os);
    s << M.m[i];
    width = std::max<std::streamsize>(width, (std::streamsize)s.str().size());
  }
  for (int r = 0; r < 4; ++r) {
    for (int c = 0; c < 4; ++c) {
      if (c) os << ' ';
      os.width(width);
      os << M(r, c);
    }
    if (r < 3) os << '\n';
  }
  return os;
}

}  // namespace lcr
