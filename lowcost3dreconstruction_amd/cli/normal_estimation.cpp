// normal_estimation -- k-nearest-neighbour PCA normals on the GPU.
//
// Flag-compatible with the reference tool (pcl_tools/normal_estimation.cpp:42-48):
//   -i/--input, -o/--output (required), -n/--neighbors (50),
//   -r/--reverse_normals, -c/--centroid (viewpoint = centroid), -z/--origin
//   (viewpoint = sensor origin); -c with -z is an error.
// Normals are flipped towards the viewpoint, then all negated when
// (centroid && !reverse) || (!centroid && reverse) (normal_estimation.cpp:112-118).
// Input normals are discarded; output is binary PLY with x y z rgb normals curvature.
#include <iostream>
#include <stdexcept>
#include <string>

#include "lcr/options.hpp"
#include "lcr/ply_io.hpp"
#include "lcr/registration.hpp"

using PointT = lcr::PointXYZRGB;
using PointC = lcr::PointCloud<PointT>;
using PointN = lcr::PointCloud<lcr::Normal>;

int main(int argc, char* argv[]) {
  try {
    lcr::Options desc("Options");
    desc.add("help,h", "Print help message", false)
        .add("input,i", "Input file (.ply)")
        .add("output,o", "Output file (.ply)")
        .add("neighbors,n", "N. of neighbors to analyze for each point", true, "50", true)
        .add("reverse_normals,r", "Reverse normals' direction", false)
        .add("centroid,c", "Use the centroid as defined view point", false)
        .add("origin,z", "Use the origin as defined view point", false);
    desc.parse(argc, argv);
    if (desc.count("help")) {
      std::cout << "Estimate a set of normals for all the points in the input dataset." << std::endl
                << std::endl;
      std::cout << desc.help() << std::endl;
      return 0;
    }
    if (!(desc.count("input") && desc.count("output")))
      throw std::logic_error("Correct mode of use: " + std::string(argv[0]) +
                             " -i input.ply -o output.ply [opts]");
    const std::string in = desc.str("input"), out = desc.str("output");
    const long neighbors = desc.as<long>("neighbors");
    if (neighbors < 0) throw lcr::options_error("the argument ('" + desc.str("neighbors") +
                                                "') for option '--neighbors' is invalid");
    const bool reverse = desc.count("reverse_normals") > 0;
    const bool centroid = desc.count("centroid") > 0;
    const bool origin = desc.count("origin") > 0;
    if (origin && centroid)
      throw std::runtime_error(
          "It is not possible to use the centroid and origin as viewpoint at the same time");

    PointC::Ptr cloud(new PointC);
    if (lcr::io::loadPLYFile<PointT>(in, *cloud) == -1)
      throw std::runtime_error("Couldn't load input point cloud: " + in);

    lcr::NormalEstimation<PointT, lcr::Normal> ne;
    ne.setInputCloud(cloud);
    ne.setKSearch((int)neighbors);
    PointN normals;
    if (centroid) {
      float c[3];
      lcr::compute3DCentroid(*cloud, c);
      ne.setViewPoint(c[0], c[1], c[2]);
    } else if (origin) {
      ne.useSensorOriginAsViewPoint();
    }
    if (!ne.compute(normals)) throw std::runtime_error("normal estimation failed: " + ne.lastError());

    if ((centroid && !reverse) || (!centroid && reverse)) {
      for (auto& n : normals.points) {
        n.normal_x *= -1;
        n.normal_y *= -1;
        n.normal_z *= -1;
      }
    }
    lcr::PointCloud<lcr::PointXYZRGBNormal> cloud_normals;
    lcr::concatenateFields(*cloud, normals, cloud_normals);
    if (lcr::io::savePLYFileBinary(out, cloud_normals) != 0)
      throw std::runtime_error("Couldn't write output file: " + out);
    return 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
  } catch (...) {
    std::cerr << "An unknown error has occurred." << std::endl;
  }
  return -1;
}
