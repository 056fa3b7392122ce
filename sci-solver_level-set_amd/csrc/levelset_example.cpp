// Command-line driver with the options and presets of the reference's
// examples: Example1 on tetgen tet meshes (src/examples/example1.cu:4-142)
// and, built with -DLS_EXAMPLE_TRI, Example2 on PLY triangle meshes
// (src/examples/example2.cu:3-145). Runs LevelSet::solveLevelSet and writes
// result<step>.vtk into the working directory.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "LevelSet.h"

#ifdef LS_EXAMPLE_TRI
static const bool kTri = true;
static const char* kName = "Example2";
#else
static const bool kTri = false;
static const char* kName = "Example1";
#endif

static void usage() {
  std::cout << "Usage: ./" << kName << " [OPTIONS]\n"
            << "   -h                 Print this help message.\n"
            << "   -v                 Print verbose runtime information.\n"
            << "   -i FILENAME        Use this input mesh (" << (kTri ? "PLY" : "node/ele") << ").\n"
            << "   -n NSTEPS          # of steps to take of TIMESTEP amount.\n"
            << "   -t TIMESTEP        Duration of a timestep.\n"
            << "   -s INSIDE_NITER    # of inside iterations.\n"
            << "   -c                 Bound each timestep by the narrowband's CFL condition.\n"
            << "   -d NSIDE           # of sides for Square partition type.\n"
            << "   -p PARTITION_TYPE  1 for Square, otherwise is it METIS.\n"
            << "   -b NUM_BLOCKS      # of blocks for Square partition type.\n"
            << "   -m METIS_SIZE      The size for METIS partiation type.\n"
            << "   -w BANDWIDTH       The Bandwidth for the algorithm.\n"
            << "   -y EXAMPLE_TYPE    Example type: 'center', 'revolve', 'x', 'curvature'\n"
            << "   -o ISOVALUE        The isovalue for curvature." << std::endl;
}

static bool ends_with(const std::string& s, const std::string& t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

int main(int argc, char* argv[]) {
  LevelSet data(kTri, kTri ? "../src/test/test_data/sphere_266verts.ply" : "../src/test/test_data/sphere334", false);
  std::string type = "x";
  float isovalue = 0.f;
  for (int i = 1; i < argc; ++i) {
    const bool has = i + 1 < argc;
    if (!strcmp(argv[i], "-v")) {
      data.verbose_ = true;
    } else if (!strcmp(argv[i], "-h")) {
      usage();
      return 0;
    } else if (!has) {
      break;
    } else if (!strcmp(argv[i], "-i")) {
      data.filename_ = argv[++i];
      if (!kTri) {
        if (ends_with(data.filename_, ".node")) data.filename_.resize(data.filename_.size() - 5);
        if (ends_with(data.filename_, ".ele")) data.filename_.resize(data.filename_.size() - 4);
      }
    } else if (!strcmp(argv[i], "-n")) {
      data.numSteps_ = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-t")) {
      data.timeStep_ = atof(argv[++i]);
    } else if (!strcmp(argv[i], "-s")) {
      data.insideIterations_ = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-c")) {
      data.cflTimeStep_ = true;
    } else if (!strcmp(argv[i], "-d")) {
      data.sideLengths_ = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-p")) {
      data.partitionType_ = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-o")) {
      isovalue = (float)atof(argv[++i]);
    } else if (!strcmp(argv[i], "-m")) {
      data.metisSize_ = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-b")) {
      data.blockSize_ = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "-w")) {
      data.bandwidth_ = atof(argv[++i]);
    } else if (!strcmp(argv[i], "-y")) {
      type = argv[++i];
    }
  }
  if (type == "center" || type == "revolve" || type == "curvature") {
    data.initializeMesh();
    const std::vector<point>& V = kTri ? data.triMesh_->vertices : data.tetMesh_->vertices;
    point center(0, 0, 0);
    for (const point& p : V) center = center + p;
    center = center / (float)V.size();
    float rmax = 0.f;
    for (const point& p : V) rmax = std::max(rmax, len(p - center));
    // Initial values (example1.cu:96-119, example2.cu:98-119).
    std::vector<float> vals;
    vals.reserve(V.size());
    for (const point& q : V) {
      const point p = q - center;
      float theta = 0.f;
      if (type == "revolve") {
        theta = std::acos(std::abs(p[0]) / std::sqrt(p[0] * p[0] + p[1] * p[1]));
        if (p[1] < 0.f) theta = -theta;
      }
      if (type == "revolve")
        vals.push_back(10.f * theta);
      else if (type == "center")
        vals.push_back((float)(len(p) - rmax / 2.));
      else
        vals.push_back(std::max(std::abs(p[0]), std::max(std::abs(p[1]), std::abs(p[2]))) - isovalue);
    }
    // Per-element velocity (example1.cu:121-137, example2.cu:121-140).
    std::vector<point> adv;
    if (kTri) {
      for (const auto& f : data.triMesh_->faces) {
        const point p = (V[f[0]] + V[f[1]] + V[f[2]]) / 3.f;
        if (type == "center") {
          const point pt = p - center;
          const float mag = len(pt) / (rmax / 10.f);
          adv.push_back(pt / mag);
        } else if (type == "revolve") {
          adv.push_back(cross(p, point(0, 0, 1)) * len(p) / 100.f);
        } else {
          adv.push_back(point(0.f, 0.f, 0.f));
        }
      }
    } else {
      for (const auto& t : data.tetMesh_->tets) {
        const point p = (V[t[0]] + V[t[1]] + V[t[2]] + V[t[3]]) / 4.f - center;
        const float mag = len(p) / (rmax / 20.f);
        if (type == "revolve") {
          point p2 = p;
          p2[2] = 0.f;
          const point p3 = cross(p2, point(0, 0, 1));
          adv.push_back(p3 * len(p2) / (100.f * len(p3)));
        } else if (type == "center") {
          adv.push_back(p / mag / mag);
        } else {
          adv.push_back(point(0.f, 0.f, 0.f));
        }
      }
    }
    data.initializeVertices(vals);
    data.initializeAdvection(adv);
  }
  data.solveLevelSet();
  data.writeVTK();
  return 0;
}
