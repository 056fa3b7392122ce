// Drop-in C++ LevelSet class (src/LevelSet.h:19-67 of the reference), running
// the MI355X hot path through the C-ABI in include/levelset_hip.h.
//
// Same constructor, methods and public fields. The mesh classes the reference
// takes from SCI-Solver_Eikonal (TriMesh, TetMesh, point) are replaced by the
// minimal types below, which expose the members the reference's examples use
// (vertices, faces / tets, vertT, normals as per-element velocity).
#pragma once

#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/levelset_hip.h"

struct point {
  float v[3];
  point(float x = 0.f, float y = 0.f, float z = 0.f) : v{x, y, z} {}
  float& operator[](int i) { return v[i]; }
  const float& operator[](int i) const { return v[i]; }
  point operator+(const point& o) const { return point(v[0] + o[0], v[1] + o[1], v[2] + o[2]); }
  point operator-(const point& o) const { return point(v[0] - o[0], v[1] - o[1], v[2] - o[2]); }
  point operator*(float s) const { return point(v[0] * s, v[1] * s, v[2] * s); }
  point operator/(float s) const { return point(v[0] / s, v[1] / s, v[2] / s); }
};
inline float len(const point& p) { return std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }
inline point cross(const point& a, const point& b) {
  return point(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}

// Triangle surface mesh read from an ASCII PLY file.
struct TriMesh {
  std::vector<point> vertices;
  std::vector<std::array<int, 3>> faces;
  std::vector<float> vertT;    // level-set values per vertex
  std::vector<point> normals;  // velocity per face
  static TriMesh* read(const char* filename, bool verbose = false);
};

// Tetrahedral mesh read from tetgen .node/.ele files.
struct TetMesh {
  std::vector<point> vertices;
  std::vector<std::array<int, 4>> tets;
  std::vector<float> vertT;    // level-set values per vertex
  std::vector<point> normals;  // velocity per tet
  static TetMesh* read(const char* basename, bool verbose = false);  // basename.node + basename.ele
  void reorient();  // positive orientation of every tet
};

// Solver handle owned by LevelSet (one C-ABI context on device 0).
class meshFIM {
 public:
  meshFIM() = default;
  ~meshFIM();
  ls_ctx* ctx = nullptr;
};
using meshFIM2d = meshFIM;
using meshFIM3d = meshFIM;

/** The class that represents all of the available options for LevelSet */
class LevelSet {
 public:
  LevelSet(bool isTriMesh, std::string fname = "../src/test/test_data/sphere_266verts.ply",
           bool verbose = false);
  virtual ~LevelSet();

  // helper functions
  std::vector<float>& getResultAtIteration(size_t i);
  size_t numIterations();
  void writeVTK();
  void writeFLD();
  // main functions
  void initializeVertices(std::vector<float> values);
  void initializeAdvection(std::vector<point> values);
  void initializeMesh();
  /** Creates the mesh, partitions the mesh, and runs the algorithm. */
  void solveLevelSet();
  /** RMS error of every stored iteration against `solution`, as a log plot. */
  void printErrorGraph(std::vector<float> solution);
  // data
  bool verbose_;
  std::string filename_;
  int partitionType_;
  int numSteps_;
  double timeStep_;
  int insideIterations_;
  int blockSize_;
  int sideLengths_;
  double bandwidth_;
  int metisSize_;
  bool userSetInitial_;
  bool userSetAdvection_;
  TriMesh* triMesh_;
  TetMesh* tetMesh_;
  std::vector<std::vector<float>> time_values_;
  meshFIM2d* FIMPtr2d_;
  meshFIM3d* FIMPtr3d_;
  bool isTriMesh_;
  // Opt-in (no reference counterpart; default off): advect each timestep
  // with min(timeStep_, the CFL bound of the new narrowband), the bound the
  // reference computes per step and discards (meshFIM2d::compute_deltaT,
  // src/core2d/cuda/meshFIM2d.cu:609-674); see ls_compute_dt.
  bool cflTimeStep_;
};
