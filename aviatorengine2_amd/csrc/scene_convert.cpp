// scene_convert.cpp - host conversion of the reference ABI arrays into the
// traversal layouts (Node2 trees, TriPre records, ObjGeom records).
//
// Every derived float is computed with the same operations, in the same
// order, as the reference computes it per ray (HitQuad / HitTriangle /
// HitMeshTriangle, ray_trace.comp.hybrid.hlsl:425-598), so precomputing it
// here changes no result. Built with -ffp-contract=off.
#include "scene_convert.h"

#include <cmath>
#include <cstring>

namespace avrt {

static inline void hcross(const float a[3], const float b[3], float o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline float hdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

namespace {

struct Converter {
    const AvrtBVHNode *nodes;
    uint32_t base;      // index of the tree's root in `nodes`
    uint32_t avail;     // nodes available from base
    std::vector<Node2> *out;
    std::string err;

    static void set_child(Node2 &n, int c, const float lo[3], const float hi[3], uint32_t ref)
    {
        n.bx[c] = lo[0];
        n.bx[c + 2] = hi[0];
        n.by[c] = lo[1];
        n.by[c + 2] = hi[1];
        n.bz[c] = lo[2];
        n.bz[c + 2] = hi[2];
        n.ref[c] = ref;
    }
    static void set_empty(Node2 &n, int c)
    {
        const float far_pt[3] = {1e30f, 1e30f, 1e30f};
        set_child(n, c, far_pt, far_pt, kRefInvalid);
    }

    // Reference for a leaf [first, first+count) whose box is lo/hi. Leaves
    // over 127 primitives become small subtrees of the same box.
    bool leaf_ref(uint32_t first, uint32_t count, const float lo[3], const float hi[3], uint32_t *ref)
    {
        if (count == 0) {
            err = "empty leaf";
            return false;
        }
        if (first + count > kMaxLeafFirst) {
            err = "leaf slot beyond 2^24";
            return false;
        }
        if (count <= kMaxLeafCount) {
            *ref = make_leaf(first, count);
            return true;
        }
        const uint32_t me = static_cast<uint32_t>(out->size());
        out->push_back(Node2{});
        const uint32_t half = count / 2;
        uint32_t r0, r1;
        if (!leaf_ref(first, half, lo, hi, &r0)) return false;
        if (!leaf_ref(first + half, count - half, lo, hi, &r1)) return false;
        set_child((*out)[me], 0, lo, hi, r0);
        set_child((*out)[me], 1, lo, hi, r1);
        *ref = me;
        return true;
    }

    bool ref_of(uint32_t local, int depth, uint32_t *ref)
    {
        if (local >= avail || depth > 64) {
            err = "BVH child index out of range";
            return false;
        }
        const AvrtBVHNode &n = nodes[base + local];
        if (n.count > 0) {
            const float lo[3] = {n.min_x, n.min_y, n.min_z};
            const float hi[3] = {n.max_x, n.max_y, n.max_z};
            if (n.left < 0) {
                err = "negative leaf slot";
                return false;
            }
            return leaf_ref(static_cast<uint32_t>(n.left), static_cast<uint32_t>(n.count), lo, hi, ref);
        }
        if (n.left < 0 || n.right < 0) {
            err = "negative child index";
            return false;
        }
        const uint32_t me = static_cast<uint32_t>(out->size());
        out->push_back(Node2{});
        uint32_t r[2];
        const uint32_t kids[2] = {static_cast<uint32_t>(n.left), static_cast<uint32_t>(n.right)};
        for (int c = 0; c < 2; c++) {
            if (!ref_of(kids[c], depth + 1, &r[c])) return false;
        }
        for (int c = 0; c < 2; c++) {
            const AvrtBVHNode &ch = nodes[base + kids[c]];
            const float lo[3] = {ch.min_x, ch.min_y, ch.min_z};
            const float hi[3] = {ch.max_x, ch.max_y, ch.max_z};
            set_child((*out)[me], c, lo, hi, r[c]);
        }
        *ref = me;
        return true;
    }

    // Appends the tree rooted at nodes[base] behind a super-root; returns the
    // super-root's index.
    bool convert(uint32_t *super)
    {
        const uint32_t me = static_cast<uint32_t>(out->size());
        out->push_back(Node2{});
        uint32_t r;
        if (!ref_of(0, 0, &r)) return false;
        const AvrtBVHNode &root = nodes[base];
        const float lo[3] = {root.min_x, root.min_y, root.min_z};
        const float hi[3] = {root.max_x, root.max_y, root.max_z};
        set_child((*out)[me], 0, lo, hi, r);
        set_empty((*out)[me], 1);
        *super = me;
        return true;
    }
};

} // namespace

bool convert_tree(const AvrtBVHNode *nodes, uint32_t base, uint32_t avail, std::vector<Node2> &out, uint32_t *super,
                  std::string *err)
{
    Converter cv{nodes, base, avail, &out, {}};
    if (!cv.convert(super)) {
        *err = cv.err;
        return false;
    }
    return true;
}

TriPre make_tripre(const float p0[3], const float p1[3], const float p2[3])
{
    TriPre t{};
    const float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    float n[3];
    hcross(e1, e2, n);
    const float nl = hdot(n, n);
    for (int k = 0; k < 3; k++) {
        t.p0[k] = p0[k];
        t.e1[k] = e1[k];
        t.e2[k] = e2[k];
        t.w[k] = n[k] / nl;
    }
    t.d0 = hdot(n, p0);
    t.valid = (nl < 1e-12f) ? 0u : 1u;
    return t;
}

int build_geom_records(const AvrtObject *objects, uint32_t n, const AvrtMeshInstance *inst, uint32_t ninst,
                       const std::unordered_map<uint32_t, uint32_t> &blas_root, std::vector<ObjGeom> &geom,
                       bool *has_volume, bool *has_tri_or_mesh, std::string *err)
{
    geom.assign(n, ObjGeom{});
    *has_volume = false;
    *has_tri_or_mesh = false;
    for (uint32_t i = 0; i < n; i++) {
        const AvrtObject &o = objects[i];
        ObjGeom &g = geom[i];
        if (o.colorType == AVRT_ISOTROPIC) *has_volume = true;
        uint32_t flags = o.shapeType & 0xFFu;
        switch (o.shapeType) {
        case AVRT_SHAPE_SPHERE:
            g.a[0] = o.x; g.a[1] = o.y; g.a[2] = o.z; g.a[3] = o.radius;
            g.b[0] = o.x2 - o.x; g.b[1] = o.y2 - o.y; g.b[2] = o.z2 - o.z;
            flags |= kGeomValid;
            break;
        case AVRT_SHAPE_QUAD:
        case AVRT_SHAPE_TRIANGLE: {
            *has_tri_or_mesh |= (o.shapeType == AVRT_SHAPE_TRIANGLE);
            const float u[3] = {o.u_x, o.u_y, o.u_z}, v[3] = {o.v_x, o.v_y, o.v_z}, P[3] = {o.x, o.y, o.z};
            float nn[3];
            hcross(u, v, nn);
            const float nl = hdot(nn, nn);
            g.a[0] = o.x; g.a[1] = o.y; g.a[2] = o.z; g.a[3] = hdot(nn, P);
            for (int k = 0; k < 3; k++) {
                g.b[k] = u[k];
                g.c[k] = v[k];
                const float w = nn[k] / nl;
                std::memcpy(&g.d[k], &w, 4);
            }
            if (!(nl < 1e-8f)) flags |= kGeomValid;
            break;
        }
        case AVRT_SHAPE_MESH: {
            *has_tri_or_mesh = true;
            if (o.instanceIndex >= ninst) {
                *err = "mesh object instanceIndex out of range";
                return AVRT_E_INVALID_ARG;
            }
            const AvrtMeshInstance &m = inst[o.instanceIndex];
            g.a[0] = m.px; g.a[1] = m.py; g.a[2] = m.pz; g.a[3] = (m.scale == 0.0f) ? 1.0f : m.scale;
            g.b[0] = m.qx; g.b[1] = m.qy; g.b[2] = m.qz; g.b[3] = m.qw;
            if (m.triangleCount > 0) {
                auto it = blas_root.find(m.blasBase);
                if (it == blas_root.end()) {
                    *err = "mesh BLAS not prepared";
                    return AVRT_E_STATE;
                }
                const uint32_t words[4] = {it->second, m.indexBase / 3, m.vertexBase, m.indexBase};
                std::memcpy(g.c, words, sizeof(words));
                flags |= kGeomValid;
            }
            break;
        }
        default: {
            const float s = std::sin(o.rotation), cs = std::cos(o.rotation);
            g.a[0] = o.x; g.a[1] = o.y; g.a[2] = o.z; g.a[3] = s;
            g.b[0] = o.half_x; g.b[1] = o.half_y; g.b[2] = o.half_z; g.b[3] = cs;
            flags |= kGeomValid;
            break;
        }
        }
        g.d[3] = flags;
    }
    return AVRT_OK;
}


} // namespace avrt
