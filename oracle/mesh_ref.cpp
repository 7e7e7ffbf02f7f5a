// mesh_ref.cpp -- C entry points over the reference's own mesh code, built
// with it into oracle/_ref/libmeshref.so (recipe: oracle/Makefile, target
// ref).  TEST INFRASTRUCTURE ONLY: the parity checker of the engine's OBJ
// reader, exact predicates and point-in-mesh test.
//
// Compiled together with src/examples/mesh_query0.1/{mesh_query,
// bounding_box_tree,predicates}.cpp and src/examples/cyTriMesh.h of the
// reference (mesh_query 0.1, R. Bridson, public domain; predicates by
// J. R. Shewchuk, public domain; cyCodeBase TriMesh, C. Yuksel).
#include <string.h>

#include "cyTriMesh.h"
#include "mesh_query.h"
#include "predicates.h"

extern "C" {

// examples/scene.cpp:31-33: cyTriMesh::LoadFromFileObj + ComputeBoundingBox.
// With null outputs only the counts are returned.
int mref_load_obj(const char* path, int* nv, int* nf, float* v, int* f, float* box6) {
  cyTriMesh m;
  if (!m.LoadFromFileObj(path)) return -1;
  *nv = (int)m.NV();
  *nf = (int)m.NF();
  if (!v) return 0;
  for (int i = 0; i < *nv; ++i)
    for (int a = 0; a < 3; ++a) v[3 * i + a] = m.V(i)[a];
  for (int i = 0; i < *nf; ++i)
    for (int a = 0; a < 3; ++a) f[3 * i + a] = (int)m.F(i).v[a];
  m.ComputeBoundingBox();
  for (int a = 0; a < 3; ++a) {
    box6[a] = m.GetBoundMin()[a];
    box6[3 + a] = m.GetBoundMax()[a];
  }
  return 0;
}

// mesh_query's point_inside_mesh for n points
void mref_inside(int nv, const double* xyz, int nt, const int* tri, int n, const double* pts,
                 int* out) {
  MeshObject* m = construct_mesh_object(nv, xyz, nt, tri);
  for (int i = 0; i < n; ++i) out[i] = point_inside_mesh(pts + 3 * i, m) ? 1 : 0;
  destroy_mesh_object(m);
}

// signs of Shewchuk's orient2d / orient3d
void mref_orient(int n, const double* a, const double* b, const double* c, const double* d,
                 int* o2, int* o3) {
  exactinit();
  for (int i = 0; i < n; ++i) {
    const double v2 = orient2d(a + 3 * i, b + 3 * i, c + 3 * i);
    const double v3 = orient3d(a + 3 * i, b + 3 * i, c + 3 * i, d + 3 * i);
    o2[i] = v2 > 0 ? 1 : (v2 < 0 ? -1 : 0);
    o3[i] = v3 > 0 ? 1 : (v3 < 0 ? -1 : 0);
  }
}

}  // extern "C"
