// mesh_fill.h -- OBJ triangle meshes and an exact point-in-mesh test for
// mesh-filled scenes (examples/scene.cpp:18-136, 530-598 of the reference,
// which uses the mesh_query library's point_inside_mesh).
#ifndef PD_HOST_MESH_FILL_H
#define PD_HOST_MESH_FILL_H

#include <string>
#include <vector>

struct ObjMesh {
  std::vector<float> v;  // 3 per vertex, as parsed ("%f")
  std::vector<int> f;    // 3 per triangle, 0-based (polygons fanned from their first vertex)
};

// Reads "v x y z" and "f a b c ..." lines (a face token's vertex index is
// the integer before its first '/'; 1-based, negative = relative to the
// end).  Other lines are ignored.  Returns false with `err` set on failure.
bool load_obj(const char* path, ObjMesh& mesh, std::string& err);

// Parity of the crossings of the ray from p along +z with the triangles of a
// closed mesh, decided with exact orientation predicates; a crossing through
// an edge or a vertex is counted for exactly one of the triangles sharing it
// (half-open rules on the xy projection).  Points outside the closed
// bounding box of the mesh are outside.  Exact for every double input.
class MeshInside {
 public:
  MeshInside(const std::vector<double>& xyz, const std::vector<int>& tri);
  bool inside(const double p[3]) const;

 private:
  bool crosses(const double p[3], int t) const;
  int cell_of(double v, int axis) const;

  const std::vector<double>& x_;
  const std::vector<int>& tri_;
  double lo_[3], hi_[3];
  bool empty_ = true;
  // triangles binned by the xy cells their xy bounding boxes overlap
  int nx_ = 1, ny_ = 1;
  double cell_[2] = {1.0, 1.0};
  std::vector<int> start_;
  std::vector<int> items_;
  std::vector<double> zmax_;  // per triangle
};

#endif
