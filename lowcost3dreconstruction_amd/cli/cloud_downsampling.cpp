// cloud_downsampling -- VoxelGrid downsampling on the GPU.
//
// Flag-compatible with the reference tool (pcl_tools/cloud_downsampling.cpp:35-38):
// -i/--input, -o/--output (required), -s/--leaf_size (1.0).  Prints the cloud
// summaries before and after filtering and writes a binary PLY, like the
// reference; exit status 255 (-1) on error.  The filter is lcr_voxel_grid: one
// centroid per occupied voxel (xyz, normal, curvature, colour).
#include <iostream>
#include <string>
#include <vector>

#include "lcr.h"
#include "lcr/options.hpp"
#include "lcr/ply_io.hpp"

using PointT = lcr::PointXYZRGBNormal;
using PointC = lcr::PointCloud<PointT>;

int main(int argc, char* argv[]) {
  try {
    lcr::Options desc("Options");
    desc.add("help,h", "Print help message", false)
        .add("input,i", "Input file (.ply)")
        .add("output,o", "Output file (.ply)")
        .add("leaf_size,s", "Leaf size for pcl::VoxelGrid filter in meters", true, "1", true);
    desc.parse(argc, argv);
    if (desc.count("help")) {
      std::cout << "Reduce the number of points of a point cloud, using a voxelized grid approach."
                << std::endl
                << std::endl
                << desc.help() << std::endl;
      return 0;
    }
    if (!(desc.count("input") && desc.count("output")))
      throw std::string("Correct mode of use: " + std::string(argv[0]) +
                        " -i input.ply -o output.ply [opts]");
    const float leaf = desc.as<float>("leaf_size");
    PointC cloud;
    if (lcr::io::loadPLYFile<PointT>(desc.str("input"), cloud) == -1)
      throw std::string("Couldn't load input point cloud: " + desc.str("input"));
    std::cout << "Cloud before filtering: " << std::endl << cloud << std::endl;
    const size_t n = cloud.size();
    std::vector<float> xyz(3 * n), nrm(3 * n), curv(n);
    std::vector<uint8_t> rgb(3 * n);
    for (size_t i = 0; i < n; ++i) {
      const PointT& p = cloud.points[i];
      xyz[3 * i] = p.x;
      xyz[3 * i + 1] = p.y;
      xyz[3 * i + 2] = p.z;
      nrm[3 * i] = p.normal_x;
      nrm[3 * i + 1] = p.normal_y;
      nrm[3 * i + 2] = p.normal_z;
      curv[i] = p.curvature;
      rgb[3 * i] = p.r;
      rgb[3 * i + 1] = p.g;
      rgb[3 * i + 2] = p.b;
    }
    std::vector<float> ox(3 * n + 3), on(3 * n + 3), oc(n + 1);
    std::vector<uint8_t> orgb(3 * n + 3);
    int64_t m = 0;
    if (lcr_voxel_grid(xyz.data(), nrm.data(), curv.data(), rgb.data(), (int64_t)n, leaf, ox.data(),
                       on.data(), oc.data(), orgb.data(), &m) != LCR_OK)
      throw std::string("liblcr: ") + lcr_last_error();
    PointC out;
    out.resize((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
      PointT& q = out.points[(size_t)i];
      q.x = ox[3 * i];
      q.y = ox[3 * i + 1];
      q.z = ox[3 * i + 2];
      q.normal_x = on[3 * i];
      q.normal_y = on[3 * i + 1];
      q.normal_z = on[3 * i + 2];
      q.curvature = oc[i];
      q.r = orgb[3 * i];
      q.g = orgb[3 * i + 1];
      q.b = orgb[3 * i + 2];
    }
    std::cout << "Cloud after filtering: " << std::endl << out << std::endl;
    if (lcr::io::savePLYFileBinary(desc.str("output"), out) != 0)
      throw std::string("Couldn't write output file");
    return 0;
  } catch (const lcr::options_error& e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
  } catch (const std::string& msg) {
    std::cerr << msg << std::endl;
  }
  return -1;
}
