// outlier_removal -- statistical outlier removal on the GPU.
//
// Flag-compatible with the reference tool (pcl_tools/outlier_removal.cpp:41-46):
// -i/--input, -o/--output (required), -f/--outliers_file, -n/--neighbors (50),
// -d/--dev_mult (1.0).  Prints the cloud summaries, writes the kept points as a
// binary PLY and, with -f, the removed ones to <output>_outliers.ply; exit status
// 255 (-1) on error.  The filter is lcr_outlier_removal.
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
        .add("outliers_file,f", "Saves the outliers in a ply file", false)
        .add("neighbors,n", "Neighbors to analyze for each point", true, "50", true)
        .add("dev_mult,d", "Standard deviation multiplier", true, "1", true);
    desc.parse(argc, argv);
    if (desc.count("help")) {
      std::cout << "Remove noisy measurements from a point cloud dataset using statistical "
                   "analysis techniques."
                << std::endl
                << std::endl
                << desc.help() << std::endl;
      return 0;
    }
    if (!(desc.count("input") && desc.count("output")))
      throw std::string("Correct mode of use: " + std::string(argv[0]) +
                        " -i input.ply -o output.ply [opts]");
    const int k = desc.as<int>("neighbors");
    const double mult = desc.as<double>("dev_mult");
    PointC cloud;
    if (lcr::io::loadPLYFile<PointT>(desc.str("input"), cloud) == -1)
      throw std::string("Couldn't load input point cloud: " + desc.str("input"));
    std::cout << "Cloud before filtering: " << std::endl << cloud << std::endl;
    const size_t n = cloud.size();
    std::vector<float> xyz(3 * n + 3);
    for (size_t i = 0; i < n; ++i) {
      xyz[3 * i] = cloud.points[i].x;
      xyz[3 * i + 1] = cloud.points[i].y;
      xyz[3 * i + 2] = cloud.points[i].z;
    }
    std::vector<uint8_t> keep(n + 1);
    if (lcr_outlier_removal(xyz.data(), (int64_t)n, k, mult, keep.data(), nullptr, nullptr,
                            nullptr) != LCR_OK)
      throw std::string("liblcr: ") + lcr_last_error();
    PointC kept, removed;
    for (size_t i = 0; i < n; ++i) (keep[i] ? kept : removed).points.push_back(cloud.points[i]);
    kept.width = (uint32_t)kept.points.size();
    removed.width = (uint32_t)removed.points.size();
    std::cout << "Cloud after filtering: " << std::endl << kept << std::endl;
    std::string out = desc.str("output");
    if (lcr::io::savePLYFileBinary(out, kept) != 0) throw std::string("Couldn't write output file");
    if (desc.count("outliers_file")) {
      const size_t pos = out.rfind(".ply");
      if (pos != std::string::npos) out.erase(pos, 4);
      if (lcr::io::savePLYFileBinary(out + "_outliers.ply", removed) != 0)
        throw std::string("Couldn't write outliers file");
    }
    return 0;
  } catch (const lcr::options_error& e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
  } catch (const std::string& msg) {
    std::cerr << msg << std::endl;
  }
  return -1;
}
