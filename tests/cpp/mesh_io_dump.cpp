// Prints a mesh as the C++ LevelSet readers load it (after the tet
// reorientation LevelSet::initializeMesh applies), for comparison with the
// Python readers. Usage: mesh_io_dump tet|tri <path>
#include <cstdio>
#include <cstring>

#include "LevelSet.h"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  if (!strcmp(argv[1], "tet")) {
    TetMesh* m = TetMesh::read(argv[2]);
    if (!m) return 1;
    m->reorient();
    printf("%zu %zu\n", m->vertices.size(), m->tets.size());
    for (const point& p : m->vertices) printf("%.9g %.9g %.9g\n", p[0], p[1], p[2]);
    for (const auto& t : m->tets) printf("%d %d %d %d\n", t[0], t[1], t[2], t[3]);
    delete m;
  } else {
    TriMesh* m = TriMesh::read(argv[2]);
    if (!m) return 1;
    printf("%zu %zu\n", m->vertices.size(), m->faces.size());
    for (const point& p : m->vertices) printf("%.9g %.9g %.9g\n", p[0], p[1], p[2]);
    for (const auto& t : m->faces) printf("%d %d %d\n", t[0], t[1], t[2]);
    delete m;
  }
  return 0;
}
