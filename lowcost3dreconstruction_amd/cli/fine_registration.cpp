// fine_registration -- ICP of one cloud onto another on the GPU.
//
// Flag-compatible with the reference tool (pcl_tools/fine_registration.cpp:44-54):
//   -i/--input, -t/--target, -o/--output (required), -a/--accumulated,
//   --distance_threshold (0.1), --max_iterations (50),
//   --transformation_epsilon (1e-9), --euclidean_fitness_epsilon (0.001)
// with the same stdout ("Loaded N data points from F", "Has converged:", "Score:",
// the 4x4 matrix), the same binary PLY outputs and exit status 255 (-1) on error.
// Additions: --matrix_out FILE writes the final 4x4 as the whitespace-separated
// row-major text that the transform tool reads (pcl_tools/transform.cpp:68-82), and
// --point_to_plane switches to the point-to-plane estimator (uses target normals).
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>

#include "lcr/options.hpp"
#include "lcr/ply_io.hpp"
#include "lcr/registration.hpp"

using PointT = lcr::PointXYZRGBNormal;
using PointC = lcr::PointCloud<PointT>;

int main(int argc, char* argv[]) {
  try {
    lcr::Options desc("Options");
    desc.add("help,h", "Print help message", false)
        .add("input,i", "Input cloud file (.ply)")
        .add("target,t", "Input target file (.ply)")
        .add("output,o", "Output file (.ply)")
        .add("accumulated,a", "Save the accumulated point cloud")
        .add("distance_threshold",
             "The maximum distance threshold between two correspondent points", true, "0.1", true)
        .add("max_iterations",
             "The maximum number of iterations the internal optimization should run for", true,
             "50", true)
        .add("transformation_epsilon",
             "Maximum allowable difference between two consecutive transformations to be "
             "considered as having converged",
             true, "1e-09", true)
        .add("euclidean_fitness_epsilon",
             "Maximum allowed Euclidean error between two consecutive steps in the ICP loop, "
             "before the algorithm is considered to have converged",
             true, "0.001", true)
        .add("matrix_out", "Write the final 4x4 transformation (row-major text) to this file")
        .add("point_to_plane", "Minimise point-to-plane distances (uses the target normals)",
             false);
    desc.parse(argc, argv);
    if (desc.count("help")) {
      std::cout << desc.help() << std::endl;
      return 0;
    }
    if (!(desc.count("input") && desc.count("target") && desc.count("output")))
      throw std::string("Correct mode of use: " + std::string(argv[0]) +
                        " -i input.ply -t target.ply -o output.ply [opts]");
    const std::string src_file = desc.str("input");
    const std::string tgt_file = desc.str("target");
    const std::string out_file = desc.str("output");
    const int max_iterations = desc.as<int>("max_iterations");
    if (max_iterations <= 0) throw std::string("max_iterations needs to be greater than zero.");
    const double distance_threshold = desc.as<double>("distance_threshold");
    const double transformation_epsilon = desc.as<double>("transformation_epsilon");
    const double euclidean_fitness_epsilon = desc.as<double>("euclidean_fitness_epsilon");

    PointC::Ptr cloud_src(new PointC), cloud_tgt(new PointC), registered(new PointC);
    if (lcr::io::loadPLYFile<PointT>(src_file, *cloud_src) == -1)
      throw std::string("Couldn't load input cloud file");
    std::cout << "Loaded " << cloud_src->size() << " data points from " << src_file << std::endl;
    if (lcr::io::loadPLYFile<PointT>(tgt_file, *cloud_tgt) == -1)
      throw std::string("Couldn't load input target file");
    std::cout << "Loaded " << cloud_tgt->size() << " data points from " << tgt_file << std::endl;

    lcr::IterativeClosestPoint<PointT, PointT> icp;
    icp.setInputSource(cloud_src);
    icp.setInputTarget(cloud_tgt);
    icp.setMaxCorrespondenceDistance(distance_threshold);
    icp.setMaximumIterations(max_iterations);
    icp.setTransformationEpsilon(transformation_epsilon);
    icp.setEuclideanFitnessEpsilon(euclidean_fitness_epsilon);
    icp.setUsePointToPlane(desc.count("point_to_plane") > 0);
    icp.align(*registered);
    if (!icp.lastError().empty()) std::cerr << "liblcr: " << icp.lastError() << std::endl;

    const lcr::Matrix4f transformation = icp.getFinalTransformation();
    std::cout << "Has converged: " << (icp.hasConverged() ? "True" : "False") << std::endl
              << "Score: " << icp.getFitnessScore() << std::endl;
    std::cout << transformation << std::endl;

    if (desc.count("matrix_out")) {
      std::ofstream m(desc.str("matrix_out"));
      const lcr::Matrix4d T = icp.getFinalTransformationDouble();
      m << std::setprecision(17);
      for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) m << (c ? " " : "") << T(r, c);
        m << "\n";
      }
      if (!m) throw std::string("Couldn't write matrix file");
    }
    if (lcr::io::savePLYFileBinary(out_file, *registered) != 0)
      throw std::string("Couldn't write output file");
    if (desc.count("accumulated")) {
      PointC accumulated;
      lcr::copyPointCloud(*registered, accumulated);
      accumulated += *cloud_tgt;
      if (lcr::io::savePLYFileBinary(desc.str("accumulated"), accumulated) != 0)
        throw std::string("Couldn't write accumulated file");
    }
    return 0;
  } catch (const lcr::options_error& e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
  } catch (const std::string& msg) {
    std::cerr << msg << std::endl;
  }
  return -1;
}
