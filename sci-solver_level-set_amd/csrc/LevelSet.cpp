// Drop-in LevelSet class on the MI355X C-ABI.
//
// Behaviour follows src/LevelSet.cu of the reference: the same defaults
// (:4-22), the same misuse messages followed by exit(0) (:52-108, :111-144),
// the same default initial values and velocities (:176-231), one result array
// per step including step 0, shorter when the narrowband empties
// (src/core3d/cuda/meshFIM3d.cu:634-677). The mesh readers replace the
// tetgen and TriMesh loaders of the external SCI-Solver_Eikonal module; the
// writers produce the reference's VTK and SCIRun FLD text formats
// (src/core3d/cuda/meshFIM3d.cu:20-81, src/core2d/cuda/meshFIM2d.cu:18-91).
#include "LevelSet.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>

namespace {

// Next non-empty, non-comment line of a tetgen file.
bool next_data_line(std::ifstream& f, std::istringstream& ss) {
  std::string line;
  while (std::getline(f, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    if (line.find_first_not_of(" \t\r\n") == std::string::npos) continue;
    ss.clear();
    ss.str(line);
    return true;
  }
  return false;
}

// Signed volume x 6 of a tet, evaluated in double.
double orient6(const point& a, const point& b, const point& c, const point& d) {
  const double ab[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
  const double ac[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
  const double ad[3] = {(double)d[0] - a[0], (double)d[1] - a[1], (double)d[2] - a[2]};
  return (ab[1] * ac[2] - ab[2] * ac[1]) * ad[0] + (ab[2] * ac[0] - ab[0] * ac[2]) * ad[1] +
         (ab[0] * ac[1] - ab[1] * ac[0]) * ad[2];
}

struct PlyProp {
  std::string name, type, count_type;  // count_type non-empty for list properties
};
struct PlyElement {
  std::string name;
  int64_t count = 0;
  std::vector<PlyProp> props;
};

int ply_type_size(const std::string& t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32")
    return 4;
  if (t == "double" || t == "float64") return 8;
  return 0;
}

double ply_binary_value(const unsigned char* p, const std::string& t) {
  if (t == "char" || t == "int8") return (double)*(const int8_t*)p;
  if (t == "uchar" || t == "uint8") return (double)*p;
  int16_t s; uint16_t us; int32_t i; uint32_t ui; float f; double d;
  if (t == "short" || t == "int16") { std::memcpy(&s, p, 2); return s; }
  if (t == "ushort" || t == "uint16") { std::memcpy(&us, p, 2); return us; }
  if (t == "int" || t == "int32") { std::memcpy(&i, p, 4); return i; }
  if (t == "uint" || t == "uint32") { std::memcpy(&ui, p, 4); return ui; }
  if (t == "float" || t == "float32") { std::memcpy(&f, p, 4); return f; }
  std::memcpy(&d, p, 8);
  return d;
}

}  // namespace

// tetgen .node/.ele reader (the role of tetgenio::load_tetmesh in
// src/LevelSet.cu:126-133). Point indices may start at 0 or 1; the first
// index in the .node file decides, as in tetgen. Extra corners of quadratic
// tets, attributes and boundary markers are ignored.
TetMesh* TetMesh::read(const char* basename, bool verbose) {
  std::ifstream fn(std::string(basename) + ".node"), fe(std::string(basename) + ".ele");
  if (!fn || !fe) return nullptr;
  std::istringstream ss;
  int64_t nv = 0;
  int dim = 0;
  if (!next_data_line(fn, ss) || !(ss >> nv >> dim) || nv <= 0 || dim != 3) return nullptr;
  TetMesh* m = new TetMesh();
  m->vertices.resize(nv);
  int64_t first = 0;
  for (int64_t i = 0; i < nv; ++i) {
    int64_t idx;
    double x, y, z;
    if (!next_data_line(fn, ss) || !(ss >> idx >> x >> y >> z)) { delete m; return nullptr; }
    if (i == 0) first = idx;
    if (idx - first != i) { delete m; return nullptr; }
    m->vertices[i] = point((float)x, (float)y, (float)z);
  }
  int64_t nt = 0;
  int per = 0;
  if (!next_data_line(fe, ss) || !(ss >> nt >> per) || nt < 0 || per < 4) { delete m; return nullptr; }
  m->tets.resize(nt);
  for (int64_t t = 0; t < nt; ++t) {
    int64_t idx, v[4];
    if (!next_data_line(fe, ss) || !(ss >> idx >> v[0] >> v[1] >> v[2] >> v[3])) { delete m; return nullptr; }
    for (int c = 0; c < 4; ++c) {
      const int64_t k = v[c] - first;
      if (k < 0 || k >= nv) { delete m; return nullptr; }
      m->tets[t][c] = (int)k;
    }
  }
  if (verbose) printf("Read %lld vertices and %lld tets from %s.node/.ele\n", (long long)nv, (long long)nt, basename);
  return m;
}

// Positive orientation for every tet (the role of TetMesh::reorient,
// src/LevelSet.cu:139): a tet with negative signed volume has its corners 1
// and 2 exchanged.
void TetMesh::reorient() {
  for (auto& t : tets)
    if (orient6(vertices[t[0]], vertices[t[1]], vertices[t[2]], vertices[t[3]]) < 0.0) std::swap(t[1], t[2]);
}

// PLY reader (the role of TriMesh::read, src/LevelSet.cu:114): ASCII and
// binary little-endian; vertex x, y, z and a face index list. Polygons with
// more than three corners are split into a triangle fan.
TriMesh* TriMesh::read(const char* filename, bool verbose) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) return nullptr;
  std::string line, format;
  std::getline(f, line);
  if (line.compare(0, 3, "ply") != 0) return nullptr;
  std::vector<PlyElement> elems;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ss(line);
    std::string kw;
    ss >> kw;
    if (kw == "format") {
      ss >> format;
    } else if (kw == "element") {
      PlyElement e;
      ss >> e.name >> e.count;
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) return nullptr;
      PlyProp p;
      ss >> p.type;
      if (p.type == "list") ss >> p.count_type >> p.type;
      ss >> p.name;
      elems.back().props.push_back(p);
    } else if (kw == "end_header") {
      break;
    }
  }
  const bool ascii = format == "ascii";
  if (!ascii && format != "binary_little_endian") return nullptr;
  TriMesh* m = new TriMesh();
  auto fail = [&]() -> TriMesh* { delete m; return nullptr; };
  for (const PlyElement& e : elems) {
    int ix = -1, iy = -1, iz = -1, il = -1;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const auto& p = e.props[k];
      if (p.name == "x") ix = (int)k;
      if (p.name == "y") iy = (int)k;
      if (p.name == "z") iz = (int)k;
      if (!p.count_type.empty() && (p.name == "vertex_indices" || p.name == "vertex_index")) il = (int)k;
    }
    for (int64_t r = 0; r < e.count; ++r) {
      std::vector<double> vals(e.props.size(), 0.0);
      std::vector<int64_t> list;
      if (ascii) {
        if (!std::getline(f, line)) return fail();
        std::istringstream ss(line);
        for (size_t k = 0; k < e.props.size(); ++k) {
          if (e.props[k].count_type.empty()) {
            if (!(ss >> vals[k])) return fail();
          } else {
            int64_t n;
            if (!(ss >> n) || n < 0) return fail();
            for (int64_t q = 0; q < n; ++q) {
              double v;
              if (!(ss >> v)) return fail();
              if ((int)k == il) list.push_back((int64_t)v);
            }
          }
        }
      } else {
        for (size_t k = 0; k < e.props.size(); ++k) {
          const auto& p = e.props[k];
          unsigned char buf[8];
          if (p.count_type.empty()) {
            const int sz = ply_type_size(p.type);
            if (!sz || !f.read((char*)buf, sz)) return fail();
            vals[k] = ply_binary_value(buf, p.type);
          } else {
            const int csz = ply_type_size(p.count_type), sz = ply_type_size(p.type);
            if (!csz || !sz || !f.read((char*)buf, csz)) return fail();
            const int64_t n = (int64_t)ply_binary_value(buf, p.count_type);
            for (int64_t q = 0; q < n; ++q) {
              if (!f.read((char*)buf, sz)) return fail();
              if ((int)k == il) list.push_back((int64_t)ply_binary_value(buf, p.type));
            }
          }
        }
      }
      if (e.name == "vertex") {
        if (ix < 0 || iy < 0) return fail();
        m->vertices.emplace_back((float)vals[ix], (float)vals[iy], iz >= 0 ? (float)vals[iz] : 0.f);
      } else if (e.name == "face") {
        for (size_t q = 2; q < list.size(); ++q) m->faces.push_back({(int)list[0], (int)list[q - 1], (int)list[q]});
      }
    }
  }
  // Index base: a file whose faces never name vertex 0 but do name vertex
  // nv is 1-based (src/test/test_data/SquareMesh_size16.ply is: its 450 faces
  // name vertices 1..256 of a 256-vertex grid), and its indices are shifted
  // down with a warning. Faces that still name a missing vertex are dropped.
  const int nv = (int)m->vertices.size();
  int lo = std::numeric_limits<int>::max(), hi = -1;
  for (const auto& t : m->faces)
    for (int c = 0; c < 3; ++c) lo = std::min(lo, t[c]), hi = std::max(hi, t[c]);
  if (!m->faces.empty() && lo == 1 && hi == nv) {
    fprintf(stderr, "%s: face indices run 1..%d; reading them as 1-based\n", filename, nv);
    for (auto& t : m->faces)
      for (int c = 0; c < 3; ++c) --t[c];
  }
  const size_t before = m->faces.size();
  m->faces.erase(std::remove_if(m->faces.begin(), m->faces.end(),
                                [nv](const std::array<int, 3>& t) {
                                  return t[0] < 0 || t[0] >= nv || t[1] < 0 || t[1] >= nv || t[2] < 0 || t[2] >= nv;
                                }),
                 m->faces.end());
  if (m->faces.size() != before)
    fprintf(stderr, "%s: skipped %zu faces with out-of-range vertex indices\n", filename, before - m->faces.size());
  if (verbose) printf("Read %zu vertices and %zu faces from %s\n", m->vertices.size(), m->faces.size(), filename);
  return m;
}

meshFIM::~meshFIM() {
  if (ctx) ls_destroy(ctx);
}

LevelSet::LevelSet(bool isTriMesh, std::string fname, bool verbose)
    : verbose_(verbose),
      filename_(fname),
      partitionType_(0),
      numSteps_(10),
      timeStep_(1.),
      insideIterations_(1),
      blockSize_(16),
      sideLengths_(1),
      bandwidth_(16.),
      metisSize_(16),
      userSetInitial_(false),
      userSetAdvection_(false),
      triMesh_(nullptr),
      tetMesh_(nullptr),
      FIMPtr2d_(nullptr),
      FIMPtr3d_(nullptr),
      isTriMesh_(isTriMesh),
      cflTimeStep_(false) {}

LevelSet::~LevelSet() {
  delete tetMesh_;
  delete triMesh_;
  delete FIMPtr2d_;
  delete FIMPtr3d_;
}

std::vector<float>& LevelSet::getResultAtIteration(size_t i) { return time_values_.at(i); }

size_t LevelSet::numIterations() { return time_values_.size(); }

// One ASCII VTK unstructured grid per stored step: result<j>.vtk.
void LevelSet::writeVTK() {
  const bool tri = isTriMesh_;
  if ((tri && !triMesh_) || (!tri && !tetMesh_)) {
    std::cerr << "You must initialize the mesh first!" << std::endl;
    exit(0);
  }
  const std::vector<point>& V = tri ? triMesh_->vertices : tetMesh_->vertices;
  const size_t nv = V.size(), nt = tri ? triMesh_->faces.size() : tetMesh_->tets.size();
  const int nc = tri ? 3 : 4;
  for (size_t j = 0; j < time_values_.size(); ++j) {
    const std::string name = "result" + std::to_string(j) + ".vtk";
    FILE* f = fopen(name.c_str(), "w+");
    if (!f) {
      std::cerr << "Cannot write " << name << std::endl;
      return;
    }
    fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    fprintf(f, "POINTS %zu float\n", nv);
    for (const point& p : V) fprintf(f, "%.12f %.12f %.12f\n", p[0], p[1], p[2]);
    fprintf(f, "CELLS %zu %zu\n", nt, nt * (nc + 1));
    for (size_t i = 0; i < nt; ++i) {
      if (tri)
        fprintf(f, "3 %d %d %d\n", triMesh_->faces[i][0], triMesh_->faces[i][1], triMesh_->faces[i][2]);
      else
        fprintf(f, "4 %d %d %d %d\n", tetMesh_->tets[i][0], tetMesh_->tets[i][1], tetMesh_->tets[i][2],
                tetMesh_->tets[i][3]);
    }
    fprintf(f, "CELL_TYPES %zu\n", nt);
    for (size_t i = 0; i < nt; ++i) fprintf(f, tri ? "5\n" : "10\n");
    fprintf(f, "POINT_DATA %zu\nSCALARS traveltime float 1\nLOOKUP_TABLE default\n", nv);
    for (float v : time_values_[j]) fprintf(f, tri ? "%.12f\n" : "%.12f\n ", v);
    fclose(f);
  }
}

// SCIRun text field result.fld (mesh only, constant zero cell data); triangle
// meshes also write the final vertex values to result.txt.
void LevelSet::writeFLD() {
  const bool tri = isTriMesh_;
  if ((tri && !triMesh_) || (!tri && !tetMesh_)) {
    std::cerr << "You must initialize the mesh first!" << std::endl;
    exit(0);
  }
  const std::vector<point>& V = tri ? triMesh_->vertices : tetMesh_->vertices;
  const size_t nv = V.size(), nt = tri ? triMesh_->faces.size() : tetMesh_->tets.size();
  FILE* f = fopen("result.fld", "w+");
  if (!f) {
    std::cerr << "Cannot write result.fld" << std::endl;
    return;
  }
  if (tri)
    fprintf(f,
            "SCI\nASC\n2\n{@1 {GenericField<TriSurfMesh<TriLinearLgn<Point>>,ConstantBasis<float>,vector<float>> 3 "
            "{Field 3 {PropertyManager 2 1 \"name\" @2 {Property<string> 2 0 \"TorsoVolumeMesh\" }\n}\n}\n"
            "{@3 {TriSurfMesh<TriLinearLgn<Point>> 4 {Mesh 2 {PropertyManager 2 0 }\n}\n");
  else
    fprintf(f,
            "SCI\nASC\n2\n{@1 {GenericField<TetVolMesh<TetLinearLgn<Point>>,ConstantBasis<float>,vector<float>> 3 "
            "{Field 3 {PropertyManager 2 0 }\n}\n{@2 {TetVolMesh<TetLinearLgn<Point>> 4 {Mesh 2 {PropertyManager 2 0 }\n}\n");
  fprintf(f, "{STLVector 2 %zu ", nv);
  for (const point& p : V) fprintf(f, "{%.12f %.12f %.12f}", p[0], p[1], p[2]);
  fprintf(f, "}\n{STLIndexVector 1 %zu 8 ", nt * (tri ? 3 : 4));
  for (size_t i = 0; i < nt; ++i) {
    if (tri)
      fprintf(f, "%d %d %d ", triMesh_->faces[i][0], triMesh_->faces[i][1], triMesh_->faces[i][2]);
    else
      fprintf(f, "%d %d %d %d ", tetMesh_->tets[i][0], tetMesh_->tets[i][1], tetMesh_->tets[i][2], tetMesh_->tets[i][3]);
  }
  fprintf(f, "}\n");
  fprintf(f, tri ? "{TriLinearLgn<Point>  1 }\n}\n}{ConstantBasis<float>  1 }\n"
                 : "{TetLinearLgn<Point>  1 }\n}\n}{ConstantBasis<float>  1 }\n");
  fprintf(f, "{STLVector 2 %zu ", nt);
  for (size_t i = 0; i < nt; ++i) fprintf(f, " 0");
  fprintf(f, "}\n}\n}");
  fclose(f);
  if (tri) {
    FILE* t = fopen("result.txt", "w+");
    if (t) {
      for (float v : triMesh_->vertT) fprintf(t, "%.12f\n", v);
      fclose(t);
    }
  }
}

void LevelSet::initializeVertices(std::vector<float> values) {
  if (!triMesh_ && !tetMesh_) {
    std::cerr << "You must initialize the mesh first!" << std::endl;
    exit(0);
  }
  const size_t nv = triMesh_ ? triMesh_->vertices.size() : tetMesh_->vertices.size();
  if (values.size() != nv) {
    std::cerr << "Initialize values size does not match number of vertices!" << std::endl;
    exit(0);
  }
  (triMesh_ ? triMesh_->vertT : tetMesh_->vertT) = values;
  userSetInitial_ = true;
}

void LevelSet::initializeAdvection(std::vector<point> values) {
  if (!triMesh_ && !tetMesh_) {
    std::cerr << "You must initialize the mesh first!" << std::endl;
    exit(0);
  }
  const size_t ne = triMesh_ ? triMesh_->faces.size() : tetMesh_->tets.size();
  if (values.size() != ne) {
    std::cerr << "Initialize values size does not match number of elements!" << std::endl;
    exit(0);
  }
  (triMesh_ ? triMesh_->normals : tetMesh_->normals) = values;
  userSetAdvection_ = true;
}

void LevelSet::initializeMesh() {
  if (isTriMesh_) {
    if (!triMesh_) {
      triMesh_ = TriMesh::read(filename_.c_str(), verbose_);
      if (!triMesh_) {
        std::cerr << "File open failed!!" << std::endl;
        exit(0);
      }
    }
  } else if (!tetMesh_) {
    tetMesh_ = TetMesh::read(filename_.c_str(), verbose_);
    if (!tetMesh_) {
      std::cerr << "File open failed!!" << std::endl;
      exit(0);
    }
    tetMesh_->reorient();
  }
}

namespace {

// A failed C-ABI call ends the process with status 1 (the reference aborts
// on CUDA errors through cudaSafeCall).
[[noreturn]] void die(ls_ctx* ctx, const char* what) {
  std::cerr << "LevelSet: " << what << " failed: " << (ctx ? ls_last_error(ctx) : "no context") << std::endl;
  exit(1);
}

// The per-step loop of meshFIM3d::GenerateData / meshFIM2d::GenerateData
// (src/core3d/cuda/meshFIM3d.cu:572-690, src/core2d/cuda/meshFIM2d.cu:676-863)
// on one C-ABI context. `vertT` is updated after every step, as the
// reference updates m_meshPtr->vertT.
std::vector<std::vector<float>> generate(meshFIM* fim, int dim, const std::vector<point>& V, const int* cells,
                                         size_t ne, std::vector<float>& vertT, const std::vector<point>& vel,
                                         const LevelSet& o) {
  const int64_t nv = (int64_t)V.size();
  if (ls_create(&fim->ctx, 0) != LS_OK) die(fim->ctx, "ls_create");
  ls_ctx* c = fim->ctx;
  std::vector<double> xyz(3 * nv);
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = V[i][a];
  std::vector<int32_t> cl(cells, cells + ne * (dim + 1));
  if (ls_set_mesh(c, dim, nv, xyz.data(), (int64_t)ne, cl.data()) != LS_OK) die(c, "ls_set_mesh");
  int rc;
  if (o.partitionType_ == 1) {
    const int n = o.sideLengths_, b = o.blockSize_;
    rc = dim == 3 ? ls_partition_grid(c, n, n, n, b, b, b) : ls_partition_grid(c, n, n, 1, b, b, 1);
  } else {
    // Patch counts of Partition_METIS: ceil(nn / size) in 3D
    // (src/core3d/cuda/meshFIM3d.cu:361), nn / size with a minimum of 2 in 2D
    // (src/core2d/cuda/meshFIM2d.cu:380-382); never more patches than vertices.
    const int64_t m = std::max(1, o.metisSize_);
    int64_t np = dim == 3 ? (nv + m - 1) / m : std::max<int64_t>(2, nv / m);
    np = std::max<int64_t>(1, std::min<int64_t>(np, nv));
    rc = ls_partition_rcb(c, (int)np);
  }
  if (rc != LS_OK) die(c, "partitioning");
  if (vertT.size() != (size_t)nv) vertT.assign(nv, 0.f);
  if (ls_set_phi(c, vertT.data(), LS_F32) != LS_OK) die(c, "ls_set_phi");
  std::vector<double> v(3 * ne, 0.0);
  for (size_t e = 0; e < ne && e < vel.size(); ++e)
    for (int a = 0; a < 3; ++a) v[3 * e + a] = vel[e][a];
  if (ls_set_velocity(c, v.data(), LS_VEL_ELEMENT) != LS_OK) die(c, "ls_set_velocity");

  ls_step_params p{};
  // 3D GenerateData takes the timestep and bandwidth as float
  // (src/core3d/include/meshFIM3d.h:32-36); 2D takes double.
  p.dt = dim == 3 ? (double)(float)o.timeStep_ : o.timeStep_;
  p.bandwidth = dim == 3 ? (double)(float)o.bandwidth_ : o.bandwidth_;
  p.tol = 1e-5;
  p.lambda_eps = 1e-5;
  p.niter = dim == 3 ? 5 : 10;
  p.inside_niter = o.insideIterations_;
  p.mode = LS_MODE_ADVECT;
  p.max_fim_iters = 0;
  p.cfl = o.cflTimeStep_ ? 1 : 0;

  std::vector<std::vector<float>> ans;
  ans.push_back(vertT);
  for (int s = 0; s < o.numSteps_; ++s) {
    ls_stats st{};
    if (ls_step(c, &p, &st) != LS_OK) die(c, "ls_step");
    if (st.nb_count == 0) {
      std::cout << "NOTE: Ending at timestep " << s << " due to zero narrow band." << std::endl;
      break;
    }
    if (ls_get_phi(c, vertT.data(), LS_F32) != LS_OK) die(c, "ls_get_phi");
    ans.push_back(vertT);
    if (o.verbose_)
      printf("step %d: %d narrowband patches, %d FIM iterations, dt %g, %.3f ms\n", s, st.nb_count,
             st.fim_iters, st.dt_used, st.ms_total);
  }
  return ans;
}

}  // namespace

void LevelSet::solveLevelSet() {
  if (verbose_) std::cout << "Narrowband width is " << bandwidth_ << std::endl;
  const auto t0 = std::chrono::steady_clock::now();
  if ((isTriMesh_ && !triMesh_) || (!isTriMesh_ && !tetMesh_)) initializeMesh();
  const std::vector<point>& V = isTriMesh_ ? triMesh_->vertices : tetMesh_->vertices;
  float mn = std::numeric_limits<float>::max();
  float mx = 1.f - mn;
  for (const point& p : V) {
    mn = std::min(mn, p[0]);
    mx = std::max(mx, p[0]);
  }
  const double midX = (mn + mx) / 2.;
  const double rangeX = mx - mn;
  if (isTriMesh_) {
    if (!userSetAdvection_) triMesh_->normals.assign(triMesh_->faces.size(), point(0.f, 0.f, 0.f));
    if (!userSetInitial_) {
      triMesh_->vertT.resize(V.size());
      for (size_t i = 0; i < V.size(); ++i) triMesh_->vertT[i] = (float)(V[i][0] - midX);
    }
    delete FIMPtr2d_;
    FIMPtr2d_ = new meshFIM2d();
    time_values_ = generate(FIMPtr2d_, 2, V, triMesh_->faces.empty() ? nullptr : triMesh_->faces[0].data(),
                            triMesh_->faces.size(), triMesh_->vertT, triMesh_->normals, *this);
  } else {
    if (!userSetAdvection_) tetMesh_->normals.assign(tetMesh_->tets.size(), point((float)(rangeX / 100.), 0.f, 0.f));
    if (!userSetInitial_) {
      tetMesh_->vertT.resize(V.size());
      for (size_t i = 0; i < V.size(); ++i) tetMesh_->vertT[i] = (float)(V[i][0] - midX);
    }
    delete FIMPtr3d_;
    FIMPtr3d_ = new meshFIM3d();
    time_values_ = generate(FIMPtr3d_, 3, V, tetMesh_->tets.empty() ? nullptr : tetMesh_->tets[0].data(),
                            tetMesh_->tets.size(), tetMesh_->vertT, tetMesh_->normals, *this);
  }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (verbose_) std::cout << "Computing time : " << ms << " ms" << std::endl;
}

// Root-mean-square error of every stored step against `solution`, printed as
// a log-scale text plot (src/LevelSet.cu:245-292).
void LevelSet::printErrorGraph(std::vector<float> solution) {
  const size_t n = numIterations();
  if (n == 0 || solution.empty()) return;
  std::vector<float> rms(n);
  for (size_t i = 0; i < n; ++i) {
    const std::vector<float>& r = getResultAtIteration(i);
    float sum = 0.f;
    for (size_t j = 0; j < solution.size() && j < r.size(); ++j) {
      const float e = std::abs(solution[j] - r[j]);
      sum += e * e;
    }
    rms[i] = std::sqrt(sum / (float)solution.size());
  }
  int max_log = -10, min_log = 10;
  while (std::pow(10.f, (float)max_log) < rms[0]) max_log++;
  while (std::pow(10.f, (float)min_log) > rms[n - 1]) min_log--;
  printf("\n\nlog(Err)|\n");
  bool tick = true;
  for (int i = max_log; i >= min_log; --i) {
    if (tick)
      printf("   10^%2d|", i);
    else
      printf("        |");
    for (size_t j = 0; j < n; ++j)
      putchar(rms[j] > std::pow(10.f, (float)i) && rms[j] < std::pow(10.f, (float)(i + 1)) ? '*' : ' ');
    putchar('\n');
    tick = !tick;
  }
  printf("--------|------------------------------------------  Converged to: %.4f\n", rms[n - 1]);
  printf("        |1   5    10   15   20   25   30   35\n");
  printf("                   Iteration\n");
}
