// align_cycle -- fine registration of a whole turntable capture cycle on the GPU.
//
// Replaces the pair loop of scripts/alignment.sh:105-113 and the top/bottom step
// of alignment.sh:117-118 with one batched run: the captures are read as the
// script names them (DIR/0.ply .. DIR/N-1.ply, plus DIR/top.ply and
// DIR/bottom.ply with --top_bottom), every capture i >= 1 is aligned onto capture
// i-1 (and, with --closing, capture 0 onto N-1), the top and bottom views onto
// capture 0, all in one lcr_cycle run.  Poses are chained as the script's
// in-place rewrites do (alignment.sh:109-110 applies each pair's transform to all
// later captures): capture j ends in the frame of capture 0 by
// T_1 T_2 ... T_j, with T_i the transform of capture i onto i-1.  With --closing
// the closing pair's loop residual is spread evenly over the cycle's edges
// (lcr_chain_poses).  The top and bottom pairs start from centroid_align's
// transform (pcl_tools/centroid_align.cpp:104-130) with the script's roll, pitch,
// yaw and elevation (alignment.sh:117-118: top -90, 0, 180, +1; bottom 90, 0, 0,
// -1; --top_seed / --bottom_seed change them).
//
// Output: OUT/<name>.ply (binary PLY, points and normals moved by the chained
// pose), OUT/<name>.txt (the 4x4 in the row-major text the transform tool reads,
// pcl_tools/transform.cpp:68-82) and, per pair, the fine_registration lines
// ("Has converged:", "Score:").  Flags and defaults of the ICP follow
// fine_registration (pcl_tools/fine_registration.cpp:44-54).  --yaw_step seeds
// each pair with the rig's turntable rotation (rotate_align.cpp:224-235) about the
// vertical axis through --centre; otherwise pairs start from the identity (the
// captures are coarsely aligned already, as after alignment.sh:104-113).
#include <math.h>

#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "lcr.h"
#include "lcr/options.hpp"
#include "lcr/ply_io.hpp"

using PointT = lcr::PointXYZRGBNormal;
using PointC = lcr::PointCloud<PointT>;

namespace {

void mat_mul(const double* A, const double* B, double* C) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double v = 0.0;
      for (int k = 0; k < 4; ++k) v += A[4 * r + k] * B[4 * k + c];
      C[4 * r + c] = v;
    }
}

void identity(double* T) {
  for (int i = 0; i < 16; ++i) T[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

// rotation by -yaw about the vertical (y) axis through c
void turntable_prior(double yaw_deg, const double* c, double* T) {
  const double a = -yaw_deg * M_PI / 180.0, ca = cos(a), sa = sin(a);
  const double R[9] = {ca, 0, sa, 0, 1, 0, -sa, 0, ca};
  identity(T);
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) T[4 * r + k] = R[3 * r + k];
    T[4 * r + 3] = c[r] - (R[3 * r] * c[0] + R[3 * r + 1] * c[1] + R[3 * r + 2] * c[2]);
  }
}

void check(int rc) {
  if (rc != LCR_OK) throw std::string("liblcr: ") + lcr_last_error();
}

// "roll,pitch,yaw,elevation"
void parse_seed(const std::string& text, const std::string& flag, double* v) {
  std::string c = text;
  for (char& ch : c)
    if (ch == ',') ch = ' ';
  std::istringstream is(c);
  if (!(is >> v[0] >> v[1] >> v[2] >> v[3]))
    throw std::string(flag + " must be roll,pitch,yaw,elevation");
}

}  // namespace

int main(int argc, char* argv[]) {
  try {
    lcr::Options desc("Options");
    desc.add("help,h", "Print help message", false)
        .add("dir,d", "Directory holding 0.ply .. N-1.ply (and top.ply, bottom.ply)")
        .add("num_captures,n", "Number of turntable captures")
        .add("output_dir,o", "Directory for the registered clouds and matrices")
        .add("top_bottom", "Also align top.ply and bottom.ply onto capture 0", false)
        .add("closing", "Also align capture 0 onto capture N-1", false)
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
        .add("point_to_plane", "Minimise point-to-plane distances (uses the target normals)",
             false)
        .add("normals_k", "Neighbours for normals computed on the device when a target has none",
             true, "30", true)
        .add("yaw_step", "Turntable step in degrees used as the initial pose (0 = identity)",
             true, "0", true)
        .add("centre", "Point on the turntable axis, x,y,z", true, "0,0,0", true)
        .add("top_seed", "centroid_align roll,pitch,yaw,elevation of top.ply onto 0.ply", true,
             "-90,0,180,1", true)
        .add("bottom_seed", "centroid_align roll,pitch,yaw,elevation of bottom.ply onto 0.ply",
             true, "90,0,0,-1", true)
        .add("precision", "Correspondence search: fp32, fp16 or bf16", true, "fp32", true);
    desc.parse(argc, argv);
    if (desc.count("help")) {
      std::cout << desc.help() << std::endl;
      return 0;
    }
    if (!(desc.count("dir") && desc.count("num_captures") && desc.count("output_dir")))
      throw std::string("Correct mode of use: " + std::string(argv[0]) +
                        " -d captures_dir -n num_captures -o output_dir [opts]");
    const std::string dir = desc.str("dir"), out_dir = desc.str("output_dir");
    const int ncap = desc.as<int>("num_captures");
    if (ncap < 2) throw std::string("num_captures needs to be at least 2.");
    const int max_iterations = desc.as<int>("max_iterations");
    if (max_iterations <= 0) throw std::string("max_iterations needs to be greater than zero.");
    const bool tb = desc.count("top_bottom") > 0, closing = desc.count("closing") > 0;
    const std::string prec = desc.str("precision");
    if (prec != "fp32" && prec != "fp16" && prec != "bf16")
      throw std::string("precision must be fp32, fp16 or bf16");
    double centre[3] = {0, 0, 0};
    {
      std::string c = desc.str("centre");
      for (char& ch : c)
        if (ch == ',') ch = ' ';
      std::istringstream is(c);
      if (!(is >> centre[0] >> centre[1] >> centre[2]))
        throw std::string("centre must be x,y,z");
    }

    std::vector<std::string> names;
    for (int i = 0; i < ncap; ++i) names.push_back(std::to_string(i));
    if (tb) {
      names.push_back("top");
      names.push_back("bottom");
    }
    std::vector<PointC> clouds(names.size());
    std::vector<std::vector<float>> xyz(names.size()), nrm(names.size());
    bool all_nrm = true;
    for (size_t c = 0; c < names.size(); ++c) {
      const std::string f = dir + "/" + names[c] + ".ply";
      if (lcr::io::loadPLYFile<PointT>(f, clouds[c]) == -1)
        throw std::string("Couldn't load input cloud file " + f);
      std::cout << "Loaded " << clouds[c].size() << " data points from " << f << std::endl;
      const size_t n = clouds[c].size();
      xyz[c].resize(3 * n);
      nrm[c].resize(3 * n);
      bool finite_nrm = false;
      for (size_t i = 0; i < n; ++i) {
        const PointT& p = clouds[c].points[i];
        xyz[c][3 * i] = p.x;
        xyz[c][3 * i + 1] = p.y;
        xyz[c][3 * i + 2] = p.z;
        nrm[c][3 * i] = p.normal_x;
        nrm[c][3 * i + 1] = p.normal_y;
        nrm[c][3 * i + 2] = p.normal_z;
        finite_nrm = finite_nrm || (p.normal_x != 0.f || p.normal_y != 0.f || p.normal_z != 0.f);
      }
      all_nrm = all_nrm && finite_nrm;
    }

    // pairs: i -> i-1, optional 0 -> N-1, top/bottom -> 0
    std::vector<int32_t> src, tgt;
    for (int i = 1; i < ncap; ++i) {
      src.push_back(i);
      tgt.push_back(i - 1);
    }
    if (closing) {
      src.push_back(0);
      tgt.push_back(ncap - 1);
    }
    if (tb) {
      src.push_back(ncap);
      tgt.push_back(0);
      src.push_back(ncap + 1);
      tgt.push_back(0);
    }
    const int npairs = (int)src.size();
    std::vector<double> guesses(16 * (size_t)npairs);
    const double yaw = desc.as<double>("yaw_step");
    double seeds[2][4];
    parse_seed(desc.str("top_seed"), "top_seed", seeds[0]);
    parse_seed(desc.str("bottom_seed"), "bottom_seed", seeds[1]);
    for (int p = 0; p < npairs; ++p) {
      double* g = &guesses[16 * (size_t)p];
      if (src[p] >= ncap) {
        // top / bottom onto capture 0: centroid_align's transform
        const double* v = seeds[src[p] - ncap];
        check(lcr_centroid_align(xyz[src[p]].data(), (int64_t)clouds[src[p]].size(),
                                 xyz[0].data(), (int64_t)clouds[0].size(), v[0], v[1], v[2], v[3],
                                 g));
      } else if (yaw != 0.0) {
        turntable_prior(yaw, centre, g);
      } else {
        identity(g);
      }
    }

    std::vector<const float*> xp(names.size()), np_(names.size());
    std::vector<int64_t> counts(names.size());
    for (size_t c = 0; c < names.size(); ++c) {
      xp[c] = xyz[c].data();
      np_[c] = nrm[c].data();
      counts[c] = (int64_t)clouds[c].size();
    }
    lcr_cycle_desc d;
    memset(&d, 0, sizeof d);
    lcr_icp_params_default(&d.params);
    d.params.max_corr_dist = desc.as<double>("distance_threshold");
    d.params.max_iterations = max_iterations;
    d.params.transformation_epsilon = desc.as<double>("transformation_epsilon");
    d.params.euclidean_fitness_epsilon = desc.as<double>("euclidean_fitness_epsilon");
    d.params.point_to_plane = desc.count("point_to_plane") ? 1 : 0;
    d.params.precision_mode = prec == "fp16" ? LCR_PREC_FP16 : (prec == "bf16" ? LCR_PREC_BF16 : 0);
    d.n_clouds = (int32_t)names.size();
    d.xyz = xp.data();
    d.nrm = (all_nrm && d.params.point_to_plane) ? np_.data() : nullptr;
    d.n_points = counts.data();
    d.n_pairs = npairs;
    d.pair_src = src.data();
    d.pair_tgt = tgt.data();
    d.n_hyp = 1;
    d.guesses = guesses.data();
    d.normals_k = desc.as<int>("normals_k");

    lcr_cycle* cyc = nullptr;
    check(lcr_cycle_create(&d, &cyc));
    std::vector<double> T(16 * (size_t)npairs), fit(npairs);
    std::vector<int32_t> iters(npairs), conv(npairs), state(npairs);
    lcr_cycle_result r;
    memset(&r, 0, sizeof r);
    r.T = T.data();
    r.iterations = iters.data();
    r.converged = conv.data();
    r.state = state.data();
    r.fitness = fit.data();
    const int rc = lcr_cycle_run(cyc);
    const int rc2 = rc == LCR_OK ? lcr_cycle_results(cyc, &r) : rc;
    if (rc2 != LCR_OK) {
      const std::string msg = lcr_last_error();
      lcr_cycle_destroy(cyc);
      throw std::string("liblcr: ") + msg;
    }
    lcr_cycle_destroy(cyc);

    // chained poses into the frame of capture 0, the closing pair's residual spread
    // over the loop
    std::vector<double> pose(16 * names.size());
    check(lcr_chain_poses(ncap, T.data(), closing ? &T[16 * (size_t)(ncap - 1)] : nullptr,
                          pose.data()));
    if (tb) {
      const int base = npairs - 2;
      mat_mul(&pose[0], &T[16 * base], &pose[16 * ncap]);
      mat_mul(&pose[0], &T[16 * (base + 1)], &pose[16 * (ncap + 1)]);
    }
    for (int p = 0; p < npairs; ++p) {
      std::cout << "Pair " << names[src[p]] << " -> " << names[tgt[p]]
                << ": Has converged: " << (conv[p] ? "True" : "False") << ", Score: " << fit[p]
                << ", Iterations: " << iters[p] << std::endl;
    }

    for (size_t c = 0; c < names.size(); ++c) {
      const size_t n = clouds[c].size();
      std::vector<float> ox(3 * n), on(3 * n);
      check(lcr_transform_points(xyz[c].data(), nrm[c].data(), (int64_t)n, 12, &pose[16 * c],
                                 ox.data(), on.data()));
      PointC out = clouds[c];
      for (size_t i = 0; i < n; ++i) {
        PointT& q = out.points[i];
        q.x = ox[3 * i];
        q.y = ox[3 * i + 1];
        q.z = ox[3 * i + 2];
        q.normal_x = on[3 * i];
        q.normal_y = on[3 * i + 1];
        q.normal_z = on[3 * i + 2];
      }
      const std::string fo = out_dir + "/" + names[c] + ".ply";
      if (lcr::io::savePLYFileBinary(fo, out) != 0) throw std::string("Couldn't write " + fo);
      std::ofstream m(out_dir + "/" + names[c] + ".txt");
      m << std::setprecision(17);
      for (int rr = 0; rr < 4; ++rr) {
        for (int cc = 0; cc < 4; ++cc) m << (cc ? " " : "") << pose[16 * c + 4 * rr + cc];
        m << "\n";
      }
      if (!m) throw std::string("Couldn't write matrix for " + names[c]);
    }
    return 0;
  } catch (const lcr::options_error& e) {
    std::cerr << "ERROR: " << e.what() << std::endl;
  } catch (const std::string& msg) {
    std::cerr << msg << std::endl;
  }
  return -1;
}
