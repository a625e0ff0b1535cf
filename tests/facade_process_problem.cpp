// facade_process_problem.cpp -- ProcessProblem's call sequence (main.cpp:257-631)
// written against hpmhost::HPM with the reference's method names, for the
// photometric pass and the prior pass at distance 0, with hpmhost::MapF in
// place of cv::Mat.  tests/test_gpu_facade.py builds nothing: build.py
// compiles this file into tests/_build/facade_process_problem, and the test
// compares the files it writes with those of the library's own ProcessProblem.
//
//   facade_process_problem <dense_folder> <seed>
//
// For every problem of pair.txt: the photometric pass (seed + i); then
// ConfidenceEvaluation; then the prior pass (seed + 1000 + i).
#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "hpm_host.hpp"

using namespace hpmhost;

namespace {

std::string ResultDir(const std::string& dense_folder, int id) {
  std::stringstream result_path;
  result_path << dense_folder << "/HPM_MVS_plusplus" << "/2333_" << std::setw(8) << std::setfill('0') << id;
  return result_path.str();
}

MapU8 ReferenceImage(const std::string& dense_folder, int id) {
  std::stringstream p;
  p << dense_folder << "/images/" << std::setw(8) << std::setfill('0') << id << ".pgm";
  MapU8 img;
  if (!ReadNetpbm(p.str(), img)) throw Error("cannot read " + p.str());
  return img;
}

void ProcessProblem(const std::string& dense_folder, const std::vector<Problem>& problems, const int idx,
                    bool prior_consistency, uint64_t seed) {
  const Problem problem = problems[idx];
  const std::string result_folder = ResultDir(dense_folder, problem.ref_image_id);
  HPM hpm;
  hpm.SetSeed(seed);
  hpm.InuputInitialization(dense_folder, problems, idx);
  hpm.CudaSpaceInitialization(dense_folder, problem);

  const int width = hpm.GetReferenceImageWidth();
  const int height = hpm.GetReferenceImageHeight();
  MapF depths(width, height), normals(width, height, 3), costs(width, height), texture(width, height);

  hpm.SetMandConsistencyParams(false);
  const MapU8 image = ReferenceImage(dense_folder, problem.ref_image_id);
  hpm.CudaCannyInitialization(CannyEdges(image, 50, 150));  // cv::Canny(Image_scaled_color, Canny_edge, 50, 150)
  if (prior_consistency) hpm.CudaConfidenceInitialization(dense_folder, problems, idx);
  if (!prior_consistency) hpm.TextureInformationInitialization();

  if (!prior_consistency) {
    hpm.RunPatchMatch();
    for (int col = 0; col < width; ++col) {
      for (int row = 0; row < height; ++row) {
        const int center = row * width + col;
        const Float4 plane_hypothesis = hpm.GetPlaneHypothesis(center);
        depths.at(col, row) = plane_hypothesis.w;
        normals.at(col, row, 0) = plane_hypothesis.x;
        normals.at(col, row, 1) = plane_hypothesis.y;
        normals.at(col, row, 2) = plane_hypothesis.z;
        costs.at(col, row) = hpm.GetCost(center);
        texture.at(col, row) = hpm.GetTexture(center);
      }
    }
  } else {
    hpm.SetPlanarPriorParams();
    MapF confidences, textures;
    if (!ReadDmb(result_folder + "/depths.dmb", depths) || !ReadDmb(result_folder + "/normals.dmb", normals) ||
        !ReadDmb(result_folder + "/confidence.dmb", confidences) || !ReadDmb(result_folder + "/costs.dmb", costs) ||
        !ReadDmb(result_folder + "/texture0.dmb", textures))
      throw Error("cannot read the photometric pass's maps in " + result_folder);
    const Rect imageRC{0, 0, width, height};
    std::vector<Point2i> support2DPoints;
    hpm.GetSupportPoints_Classify_Check(support2DPoints, costs, confidences, textures, 1);
    const auto triangles = hpm.DelaunayTriangulation(imageRC, support2DPoints);

    MapF mask_tri(width, height);
    std::vector<Float4> planeParams_tri;
    uint32_t idx_tri = 0;
    for (const auto& triangle : triangles) {
      if (imageRC.contains(triangle.pt1) && imageRC.contains(triangle.pt2) && imageRC.contains(triangle.pt3)) {
        float L01 = std::sqrt(std::pow(triangle.pt1.x - triangle.pt2.x, 2) + std::pow(triangle.pt1.y - triangle.pt2.y, 2));
        float L02 = std::sqrt(std::pow(triangle.pt1.x - triangle.pt3.x, 2) + std::pow(triangle.pt1.y - triangle.pt3.y, 2));
        float L12 = std::sqrt(std::pow(triangle.pt2.x - triangle.pt3.x, 2) + std::pow(triangle.pt2.y - triangle.pt3.y, 2));
        float max_edge_length = std::max(L01, std::max(L02, L12));
        float step = 1.0 / max_edge_length;
        for (float p = 0; p < 1.0; p += step) {
          for (float q = 0; q < 1.0 - p; q += step) {
            int x = p * triangle.pt1.x + q * triangle.pt2.x + (1.0 - p - q) * triangle.pt3.x;
            int y = p * triangle.pt1.y + q * triangle.pt2.y + (1.0 - p - q) * triangle.pt3.y;
            mask_tri.at(x, y) = idx_tri + 1.0;
          }
        }
        planeParams_tri.push_back(hpm.GetPriorPlaneParams(triangle, depths));
        idx_tri++;
      }
    }
    for (int i = 0; i < width; ++i) {
      for (int j = 0; j < height; ++j) {
        if (mask_tri.at(i, j) > 0) {
          const float d = hpm.GetDepthFromPlaneParam(planeParams_tri[(size_t)mask_tri.at(i, j) - 1], i, j);
          if (!(d <= hpm.GetMaxDepth() && d >= hpm.GetMinDepth())) mask_tri.at(i, j) = 0;
        }
      }
    }
    hpm.CudaPlanarPriorInitialization(planeParams_tri, mask_tri);
    hpm.CudaHypothesesReload(depths, costs, normals);
    hpm.RunPatchMatch();
    for (int col = 0; col < width; ++col) {
      for (int row = 0; row < height; ++row) {
        const int center = row * width + col;
        const Float4 plane_hypothesis = hpm.GetPlaneHypothesis(center);
        depths.at(col, row) = plane_hypothesis.w;
        normals.at(col, row, 0) = plane_hypothesis.x;
        normals.at(col, row, 1) = plane_hypothesis.y;
        normals.at(col, row, 2) = plane_hypothesis.z;
        costs.at(col, row) = hpm.GetCost(center);
      }
    }
    hpm.CudaPlanarPriorRelease();
  }
  if (!WriteDmb(result_folder + "/depths.dmb", depths) || !WriteDmb(result_folder + "/normals.dmb", normals) ||
      !WriteDmb(result_folder + "/costs.dmb", costs))
    throw Error("cannot write the maps in " + result_folder);
  if (!prior_consistency && !WriteDmb(result_folder + "/texture0.dmb", texture))
    throw Error("cannot write " + result_folder + "/texture0.dmb");
  hpm.CudaSpaceRelease(false);
  hpm.ReleaseProblemHostMemory();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <dense_folder> <seed>\n", argv[0]);
    return 2;
  }
  try {
    const std::string folder = argv[1];
    const uint64_t seed = std::strtoull(argv[2], nullptr, 10);
    const std::vector<Problem> problems = ReadPairFile(folder + "/pair.txt");
    mkdir((folder + "/HPM_MVS_plusplus").c_str(), 0755);
    for (size_t i = 0; i < problems.size(); ++i) {
      mkdir(ResultDir(folder, problems[i].ref_image_id).c_str(), 0755);  // main.cpp:255
      ProcessProblem(folder, problems, (int)i, false, seed + i);
    }
    ConfidenceEvaluation(folder, problems, false);
    for (size_t i = 0; i < problems.size(); ++i) ProcessProblem(folder, problems, (int)i, true, seed + 1000 + i);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "facade_process_problem: %s\n", e.what());
    return 1;
  }
  std::printf("facade_process_problem: done\n");
  return 0;
}
