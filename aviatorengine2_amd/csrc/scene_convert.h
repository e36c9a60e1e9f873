// scene_convert.h - host conversion of the reference ABI arrays into the
// traversal layouts of avrt_internal.h (no HIP dependency).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/avrt.h"
#include "avrt_internal.h"

namespace avrt {

// Appends the tree rooted at nodes[base] (children relative to base, as
// MeshLibrary::Add stores a BLAS and BuildBVH a TLAS) behind a super-root to
// `out`; *super receives the super-root's index. Leaves of more than 127
// primitives become small subtrees over the same box. `slot_class` (kClassLG
// / kClassOther per leaf slot, `nslots` of them) fills the child masks; null
// marks every child as holding both classes. `src`, when given, receives for
// every child slot of every Node4 in `out` the index in `nodes` of the node
// whose box fills it (~0u for empty slots), which is what a refit needs.
bool convert_tree(const AvrtBVHNode *nodes, uint32_t base, uint32_t avail, std::vector<Node4> &out, uint32_t *super,
                  std::string *err, const uint8_t *slot_class = nullptr, uint32_t nslots = 0,
                  std::vector<uint32_t> *src = nullptr);

// IsRasterProxy or a mesh: what the hybrid renderer rasterizes
// (deferred.comp.hlsl:1011-1017, renderer.h IsRasterProxy).
bool is_rasterized(const AvrtObject &o);

// kClassLG for DIFFUSE_LIGHT and DIELECTRIC objects, else kClassOther; plus
// kClassRaster or kClassNonRaster.
uint8_t object_class(const AvrtObject &o);

// Triangle record of HitMeshTriangle's derived quantities.
TriPre make_tripre(const float p0[3], const float p1[3], const float p2[3]);

// Per-slot ObjGeom records. blas_root maps a BLAS base to its super-root.
// Returns AVRT_OK or an error code with *err set.
int build_geom_records(const AvrtObject *objects, uint32_t n, const AvrtMeshInstance *inst, uint32_t ninst,
                       const std::unordered_map<uint32_t, uint32_t> &blas_root, std::vector<ObjGeom> &geom,
                       bool *has_volume, bool *has_tri_or_mesh, std::string *err);

} // namespace avrt
