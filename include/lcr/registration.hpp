// lcr/registration.hpp -- C++ facade over the C ABI with pcl::Registration's
// method names, so the ICP block of the reference's fine-registration tool
// (pcl_tools/fine_registration.cpp:105-127) reads unchanged against it:
//
//   lcr::IterativeClosestPoint<PointT, PointT> icp;
//   icp.setInputSource(cloud_src);            icp.setInputTarget(cloud_tgt);
//   icp.setMaxCorrespondenceDistance(d);      icp.setMaximumIterations(n);
//   icp.setTransformationEpsilon(e);          icp.setEuclideanFitnessEpsilon(f);
//   icp.align(*registered);
//   Matrix4f T = icp.getFinalTransformation();
//   icp.hasConverged();  icp.getFitnessScore();
//
// Like PCL the object keeps the shared pointers it is given and reads the clouds
// at align() time.  Like PCL it does not throw: failures leave hasConverged()
// false; lastError() describes them.
#pragma once
#include <float.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../lcr.h"
#include "point_cloud.hpp"

namespace lcr {

template <class PointSource, class PointTarget, class Scalar = float>
class IterativeClosestPoint {
 public:
  using PointCloudSource = PointCloud<PointSource>;
  using PointCloudTarget = PointCloud<PointTarget>;
  using PointCloudSourceConstPtr = std::shared_ptr<const PointCloudSource>;
  using PointCloudTargetConstPtr = std::shared_ptr<const PointCloudTarget>;
  using Matrix4 = lcr::Matrix4<Scalar>;

  IterativeClosestPoint() {
    lcr_icp_params_default(&params_);
    if (lcr_icp_create(&params_, &h_) != LCR_OK) {
      h_ = nullptr;
      error_ = lcr_last_error();
    }
  }
  ~IterativeClosestPoint() { lcr_icp_destroy(h_); }
  IterativeClosestPoint(const IterativeClosestPoint&) = delete;
  IterativeClosestPoint& operator=(const IterativeClosestPoint&) = delete;

  void setInputSource(const PointCloudSourceConstPtr& cloud) { src_ = cloud; }
  void setInputTarget(const PointCloudTargetConstPtr& cloud) { tgt_ = cloud; }
  PointCloudSourceConstPtr getInputSource() const { return src_; }
  PointCloudTargetConstPtr getInputTarget() const { return tgt_; }

  void setMaxCorrespondenceDistance(double d) { params_.max_corr_dist = d; }
  double getMaxCorrespondenceDistance() const { return params_.max_corr_dist; }
  void setMaximumIterations(int n) { params_.max_iterations = n; }
  int getMaximumIterations() const { return params_.max_iterations; }
  void setTransformationEpsilon(double e) { params_.transformation_epsilon = e; }
  double getTransformationEpsilon() const { return params_.transformation_epsilon; }
  void setEuclideanFitnessEpsilon(double e) { params_.euclidean_fitness_epsilon = e; }
  double getEuclideanFitnessEpsilon() const { return params_.euclidean_fitness_epsilon; }
  // IterativeClosestPointWithNormals' estimator: TransformationEstimationPointToPlaneLLS
  void setUsePointToPlane(bool on) { params_.point_to_plane = on ? 1 : 0; }

  void align(PointCloudSource& output) { align(output, Matrix4::Identity()); }

  void align(PointCloudSource& output, const Matrix4& guess) {
    converged_ = false;
    final_.setIdentity();
    if (!h_ || !src_ || !tgt_) {
      error_ = !h_ ? error_ : "source or target not set";
      return;
    }
    output = *src_;
    if (!call(lcr_icp_set_params(h_, &params_))) return;
    const float* sn = has_normals<PointSource>() ? &src_->points.data()->normal_x : nullptr;
    const float* tn = has_normals<PointTarget>() ? &tgt_->points.data()->normal_x : nullptr;
    if (src_->empty()) sn = nullptr;
    if (tgt_->empty()) tn = nullptr;
    if (!call(lcr_icp_set_source(h_, src_->empty() ? nullptr : &src_->points.data()->x, sn,
                                 (int64_t)src_->size(), sizeof(PointSource))))
      return;
    if (!call(lcr_icp_set_target(h_, tgt_->empty() ? nullptr : &tgt_->points.data()->x, tn,
                                 (int64_t)tgt_->size(), sizeof(PointTarget))))
      return;
    double g[16];
    for (int i = 0; i < 16; ++i) g[i] = (double)guess.m[i];
    std::vector<float> xyz(3 * src_->size()), nrm(sn ? 3 * src_->size() : 0);
    if (!call(lcr_icp_align(h_, g, xyz.empty() ? nullptr : xyz.data(),
                            nrm.empty() ? nullptr : nrm.data())))
      return;
    for (size_t i = 0; i < output.size(); ++i) {
      output.points[i].x = xyz[3 * i];
      output.points[i].y = xyz[3 * i + 1];
      output.points[i].z = xyz[3 * i + 2];
      if constexpr (has_normals<PointSource>()) {
        if (sn) {
          output.points[i].normal_x = nrm[3 * i];
          output.points[i].normal_y = nrm[3 * i + 1];
          output.points[i].normal_z = nrm[3 * i + 2];
        }
      }
    }
    double T[16];
    lcr_icp_final_transform(h_, T);
    for (int i = 0; i < 16; ++i) final_.m[i] = (Scalar)T[i];
    converged_ = lcr_icp_has_converged(h_) == 1;
  }

  Matrix4 getFinalTransformation() const { return final_; }
  // the FP64 transform before the cast to Scalar
  Matrix4d getFinalTransformationDouble() const {
    Matrix4d T;
    if (h_) lcr_icp_final_transform(h_, T.m);
    return T;
  }
  bool hasConverged() const { return converged_; }
  double getFitnessScore(double max_range = DBL_MAX) {
    double f = DBL_MAX;
    if (h_ && lcr_icp_fitness(h_, max_range, &f) != LCR_OK) error_ = lcr_last_error();
    return f;
  }
  int nrIterations() const { return h_ ? lcr_icp_iterations(h_) : 0; }
  int getConvergenceState() const { return h_ ? lcr_icp_convergence_state(h_) : 0; }
  const std::string& lastError() const { return error_; }

 private:
  template <class P>
  static constexpr bool has_normals() {
    return std::is_same<P, PointXYZRGBNormal>::value;
  }
  bool call(int rc) {
    if (rc == LCR_OK) return true;
    error_ = lcr_last_error();
    return false;
  }

  lcr_icp* h_ = nullptr;
  lcr_icp_params params_{};
  PointCloudSourceConstPtr src_;
  PointCloudTargetConstPtr tgt_;
  Matrix4 final_;
  bool converged_ = false;
  std::string error_;
};

// pcl::NormalEstimation<PointInT, Normal> (normal_estimation.cpp:84-108)
template <class PointInT, class PointOutT>
class NormalEstimation {
 public:
  void setInputCloud(const std::shared_ptr<const PointCloud<PointInT>>& c) { cloud_ = c; }
  void setKSearch(int k) { k_ = k; }
  void setViewPoint(float x, float y, float z) {
    vp_[0] = x;
    vp_[1] = y;
    vp_[2] = z;
    sensor_origin_ = false;
  }
  void useSensorOriginAsViewPoint() { sensor_origin_ = true; }
  // Returns false (and leaves NaN normals) when the GPU call fails.
  bool compute(PointCloud<PointOutT>& out) {
    out.resize(cloud_ ? cloud_->size() : 0);
    if (!cloud_ || cloud_->empty()) return true;
    const size_t n = cloud_->size();
    float vp[3] = {vp_[0], vp_[1], vp_[2]};
    if (sensor_origin_)
      for (int a = 0; a < 3; ++a) vp[a] = cloud_->sensor_origin[a];
    std::vector<float> nrm(3 * n), curv(n);
    const int rc = lcr_normals(&cloud_->points.data()->x, (int64_t)n, sizeof(PointInT), k_, vp, 0,
                               nrm.data(), curv.data());
    if (rc != LCR_OK) {
      error_ = lcr_last_error();
      return false;
    }
    for (size_t i = 0; i < n; ++i) {
      out.points[i].normal_x = nrm[3 * i];
      out.points[i].normal_y = nrm[3 * i + 1];
      out.points[i].normal_z = nrm[3 * i + 2];
      out.points[i].curvature = curv[i];
    }
    return true;
  }
  const std::string& lastError() const { return error_; }

 private:
  std::shared_ptr<const PointCloud<PointInT>> cloud_;
  int k_ = 0;
  float vp_[3] = {0.f, 0.f, 0.f};
  bool sensor_origin_ = true;  // PCL's default viewpoint: the cloud's sensor origin
  std::string error_;
};

// compute3DCentroid over finite points (normal_estimation.cpp:101)
template <class PointT>
void compute3DCentroid(const PointCloud<PointT>& c, float out[3]) {
  double s[3] = {0, 0, 0};
  size_t n = 0;
  for (const PointT& p : c.points) {
    if (!(std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z))) continue;
    s[0] += p.x;
    s[1] += p.y;
    s[2] += p.z;
    ++n;
  }
  for (int a = 0; a < 3; ++a) out[a] = n ? (float)(s[a] / n) : 0.f;
}

}  // namespace lcr
