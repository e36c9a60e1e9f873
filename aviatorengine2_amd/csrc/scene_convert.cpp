// scene_convert.cpp - host conversion of the reference ABI arrays into the
// traversal layouts (Node4 trees, TriPre records, ObjGeom records).
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
    std::vector<Node4> *out;
    const uint8_t *slot_class; // null: every subtree holds both classes
    uint32_t nslots;
    std::vector<uint32_t> *src; // optional: source node (base-relative + base) of each child box
    std::string err;

    void note_src(uint32_t me, int c, uint32_t node)
    {
        if (src == nullptr) return;
        if (src->size() < 4 * static_cast<size_t>(me + 1)) src->resize(4 * static_cast<size_t>(me + 1), ~0u);
        (*src)[4 * static_cast<size_t>(me) + c] = node;
    }

    static void set_child(Node4 &n, int c, const float lo[3], const float hi[3], uint32_t ref)
    {
        n.lox[c] = lo[0];
        n.hix[c] = hi[0];
        n.loy[c] = lo[1];
        n.hiy[c] = hi[1];
        n.loz[c] = lo[2];
        n.hiz[c] = hi[2];
        n.ref[c] = ref;
    }
    static Node4 empty_node()
    {
        Node4 n{};
        const float far_pt[3] = {1e30f, 1e30f, 1e30f};
        for (int c = 0; c < 4; c++) set_child(n, c, far_pt, far_pt, kRefInvalid);
        return n;
    }
    uint32_t class_of_range(uint32_t first, uint32_t count)
    {
        if (slot_class == nullptr) return kClassAll;
        uint32_t cls = 0;
        for (uint32_t i = first; i < first + count && i < nslots; i++) cls |= slot_class[i];
        return cls;
    }
    bool box_of(uint32_t local, float lo[3], float hi[3])
    {
        if (local >= avail) {
            err = "BVH child index out of range";
            return false;
        }
        const AvrtBVHNode &n = nodes[base + local];
        lo[0] = n.min_x;
        lo[1] = n.min_y;
        lo[2] = n.min_z;
        hi[0] = n.max_x;
        hi[1] = n.max_y;
        hi[2] = n.max_z;
        return true;
    }

    // Reference for a leaf [first, first+count) whose box is lo/hi. Leaves
    // over 127 primitives become small subtrees over the same box.
    bool leaf_ref(uint32_t first, uint32_t count, const float lo[3], const float hi[3], uint32_t *ref, uint32_t *cls,
                  uint32_t node)
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
            *cls = class_of_range(first, count);
            return true;
        }
        const uint32_t me = static_cast<uint32_t>(out->size());
        out->push_back(empty_node());
        const uint32_t half = count / 2;
        uint32_t r0, r1, c0, c1;
        if (!leaf_ref(first, half, lo, hi, &r0, &c0, node)) return false;
        if (!leaf_ref(first + half, count - half, lo, hi, &r1, &c1, node)) return false;
        set_child((*out)[me], 0, lo, hi, r0);
        set_child((*out)[me], 1, lo, hi, r1);
        note_src(me, 0, node);
        note_src(me, 1, node);
        (*out)[me].pad[0] = child_mask_bits(0, c0) | child_mask_bits(1, c1);
        *ref = me;
        *cls = c0 | c1;
        return true;
    }

    // Node4 of binary interior node `local`: its children, each interior
    // child replaced by its own two children, left to right.
    bool ref_of(uint32_t local, int depth, uint32_t *ref, uint32_t *cls)
    {
        if (local >= avail || depth > 64) {
            err = "BVH child index out of range";
            return false;
        }
        const AvrtBVHNode &n = nodes[base + local];
        if (n.count > 0) {
            float lo[3], hi[3];
            box_of(local, lo, hi);
            if (n.left < 0) {
                err = "negative leaf slot";
                return false;
            }
            return leaf_ref(static_cast<uint32_t>(n.left), static_cast<uint32_t>(n.count), lo, hi, ref, cls, base + local);
        }
        if (n.left < 0 || n.right < 0 || static_cast<uint32_t>(n.left) >= avail || static_cast<uint32_t>(n.right) >= avail) {
            err = "bad child index";
            return false;
        }
        uint32_t kids[4];
        int nk = 0;
        for (int32_t c : {n.left, n.right}) {
            const AvrtBVHNode &ch = nodes[base + c];
            if (ch.count == 0) {
                if (ch.left < 0 || ch.right < 0) {
                    err = "bad child index";
                    return false;
                }
                kids[nk++] = static_cast<uint32_t>(ch.left);
                kids[nk++] = static_cast<uint32_t>(ch.right);
            } else {
                kids[nk++] = static_cast<uint32_t>(c);
            }
        }
        const uint32_t me = static_cast<uint32_t>(out->size());
        out->push_back(empty_node());
        uint32_t mask = 0, all = 0;
        for (int c = 0; c < nk; c++) {
            uint32_t r, cc;
            if (!ref_of(kids[c], depth + 2, &r, &cc)) return false;
            float lo[3], hi[3];
            if (!box_of(kids[c], lo, hi)) return false;
            set_child((*out)[me], c, lo, hi, r);
            note_src(me, c, base + kids[c]);
            mask |= child_mask_bits(c, cc);
            all |= cc;
        }
        (*out)[me].pad[0] = mask;
        *ref = me;
        *cls = all;
        return true;
    }

    // Appends the tree rooted at nodes[base] behind a super-root; returns the
    // super-root's index.
    bool convert(uint32_t *super)
    {
        const uint32_t me = static_cast<uint32_t>(out->size());
        out->push_back(empty_node());
        uint32_t r, cls;
        if (!ref_of(0, 0, &r, &cls)) return false;
        float lo[3], hi[3];
        box_of(0, lo, hi);
        set_child((*out)[me], 0, lo, hi, r);
        note_src(me, 0, base);
        (*out)[me].pad[0] = child_mask_bits(0, cls);
        *super = me;
        return true;
    }
};

} // namespace

bool is_rasterized(const AvrtObject &o)
{
    if (o.shapeType == AVRT_SHAPE_MESH) return true;
    return o.shapeType == AVRT_SHAPE_BOX && o.colorType != AVRT_ISOTROPIC && o.textureID == AVRT_NO_TEXTURE &&
           o.rotation == 0.0f;
}

uint8_t object_class(const AvrtObject &o)
{
    const uint32_t lg = (o.colorType == AVRT_DIFFUSE_LIGHT || o.colorType == AVRT_DIELECTRIC) ? kClassLG : kClassOther;
    const uint32_t ras = is_rasterized(o) ? kClassRaster : kClassNonRaster;
    return static_cast<uint8_t>(lg | ras);
}

bool convert_tree(const AvrtBVHNode *nodes, uint32_t base, uint32_t avail, std::vector<Node4> &out, uint32_t *super,
                  std::string *err, const uint8_t *slot_class, uint32_t nslots, std::vector<uint32_t> *src)
{
    if (src) src->resize(4 * out.size(), ~0u);
    Converter cv{nodes, base, avail, &out, slot_class, nslots, src, {}};
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
            g.a[0] = m.px; g.a[1] = m.py; g.a[2] = m.pz; g.a[3] = 1.0f / ((m.scale == 0.0f) ? 1.0f : m.scale);
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
        if (object_class(o) & kClassLG) flags |= kGeomLG;
        if (is_rasterized(o)) flags |= kGeomRaster;
        g.d[3] = flags;
    }
    return AVRT_OK;
}


} // namespace avrt
