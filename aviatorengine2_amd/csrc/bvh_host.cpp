// bvh_host.cpp - host-side acceleration-structure builders of libavrt.
//
// avrt_build_bvh reproduces BuildBVH (scene/bvh.h:275-295) bit for bit:
//   * node box = surround of the primitive boxes, inflated by 1e-4 on any axis
//     thinner than that (MakeNode, bvh.h:58-89);
//   * leaves of <= 4 primitives or at depth 20 (bvh.h:150);
//   * 12-bin SAH per axis over centroid bounds, bins priced by a left sweep
//     and a right sweep, strict '<' so the first axis and the highest plane
//     of equal cost wins (bvh.h:177-235);
//   * the partition is the two-ended swap loop of libstdc++'s std::partition
//     for bidirectional iterators (which the reference calls, bvh.h:241), so
//     the leaf order matches a libstdc++ build of the reference;
//   * median fallback when no plane separates (bvh.h:237-257);
//   * nodes numbered in DFS preorder, a leaf's `left` is its first slot.
// The recursion of the reference is replaced by an explicit preorder work
// list; numbering and slot order come out identical.
//
// Compiled with -ffp-contract=off: the SAH costs must be rounded exactly like
// the reference's separate multiplies and adds.
#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/avrt.h"

namespace {

struct Box {
    float lo[3];
    float hi[3];
};

inline Box empty_box()
{
    return Box{{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}};
}

inline void grow(Box &a, const Box &b)
{
    for (int k = 0; k < 3; k++) {
        a.lo[k] = std::min(a.lo[k], b.lo[k]);
        a.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
}

inline void grow_point(Box &a, float x, float y, float z)
{
    a.lo[0] = std::min(a.lo[0], x);
    a.lo[1] = std::min(a.lo[1], y);
    a.lo[2] = std::min(a.lo[2], z);
    a.hi[0] = std::max(a.hi[0], x);
    a.hi[1] = std::max(a.hi[1], y);
    a.hi[2] = std::max(a.hi[2], z);
}

inline float half_surface(const Box &b)
{
    const float dx = b.hi[0] - b.lo[0];
    const float dy = b.hi[1] - b.lo[1];
    const float dz = b.hi[2] - b.lo[2];
    return dx * dy + dy * dz + dz * dx;
}

inline float mid_of(const Box &b, int axis) { return (b.lo[axis] + b.hi[axis]) * 0.5f; }

inline int bin_of(float c, float cmin, float scale)
{
    const int b = static_cast<int>((c - cmin) * scale);
    return b < 11 ? b : 11;
}

AvrtBVHNode finish_node(const Box &b, int left, int right, int count)
{
    const float thin = 0.0001f;
    AvrtBVHNode n;
    std::memset(&n, 0, sizeof(n));
    float lo[3] = {b.lo[0], b.lo[1], b.lo[2]};
    float hi[3] = {b.hi[0], b.hi[1], b.hi[2]};
    for (int k = 0; k < 3; k++) {
        if (hi[k] - lo[k] < thin) {
            lo[k] -= thin * 0.5f;
            hi[k] += thin * 0.5f;
        }
    }
    n.min_x = lo[0];
    n.min_y = lo[1];
    n.min_z = lo[2];
    n.max_x = hi[0];
    n.max_y = hi[1];
    n.max_z = hi[2];
    n.left = left;
    n.right = right;
    n.count = count;
    return n;
}

struct Task {
    int start, end, depth;
    int parent; // node whose child slot this task fills, -1 for the root
    int side;   // 0 = left, 1 = right
};

class Builder {
public:
    Builder(const Box *boxes, uint32_t n) : boxes_(boxes), n_(n) {}

    void run(AvrtBVHNode *out_nodes, uint32_t *out_count, uint32_t *out_order)
    {
        std::vector<uint32_t> perm(n_);
        for (uint32_t i = 0; i < n_; i++) perm[i] = i;
        std::vector<Task> todo;
        todo.push_back(Task{0, static_cast<int>(n_), 0, -1, 0});
        uint32_t nodes = 0;
        uint32_t slots = 0;
        while (!todo.empty()) {
            const Task t = todo.back();
            todo.pop_back();
            const int me = static_cast<int>(nodes++);
            if (t.parent >= 0) {
                if (t.side == 0) out_nodes[t.parent].left = me;
                else out_nodes[t.parent].right = me;
            }
            Box box = empty_box();
            float cmin[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
            float cmax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int i = t.start; i < t.end; i++) {
                const Box &b = boxes_[perm[i]];
                grow(box, b);
                for (int a = 0; a < 3; a++) {
                    const float c = mid_of(b, a);
                    cmin[a] = std::min(cmin[a], c);
                    cmax[a] = std::max(cmax[a], c);
                }
            }
            const int count = t.end - t.start;
            if (count <= 4 || t.depth >= 20) {
                // Preorder visits leaves left to right, so slots are handed out
                // in the same order the reference appends to outOrder.
                out_nodes[me] = finish_node(box, static_cast<int>(slots), -1, count);
                for (int i = t.start; i < t.end; i++) out_order[slots++] = perm[i];
                continue;
            }
            int axis_best = -1;
            int plane_best = 0;
            float cost_best = FLT_MAX;
            for (int axis = 0; axis < 3; axis++) {
                const float extent = cmax[axis] - cmin[axis];
                if (extent <= 0.0f) continue;
                const float scale = 12.0f / extent;
                Box bins[12];
                int counts[12];
                for (int b = 0; b < 12; b++) {
                    bins[b] = empty_box();
                    counts[b] = 0;
                }
                for (int i = t.start; i < t.end; i++) {
                    const Box &p = boxes_[perm[i]];
                    const int b = bin_of(mid_of(p, axis), cmin[axis], scale);
                    grow(bins[b], p);
                    counts[b]++;
                }
                float below_cost[11];
                int below_n[11];
                Box sweep = empty_box();
                int sweep_n = 0;
                for (int s = 0; s < 11; s++) {
                    grow(sweep, bins[s]);
                    sweep_n += counts[s];
                    below_n[s] = sweep_n;
                    below_cost[s] = sweep_n > 0 ? half_surface(sweep) * sweep_n : 0.0f;
                }
                sweep = empty_box();
                sweep_n = 0;
                for (int s = 11; s > 0; s--) {
                    grow(sweep, bins[s]);
                    sweep_n += counts[s];
                    if (below_n[s - 1] == 0 || sweep_n == 0) continue;
                    const float cost = below_cost[s - 1] + half_surface(sweep) * sweep_n;
                    if (cost < cost_best) {
                        cost_best = cost;
                        axis_best = axis;
                        plane_best = s;
                    }
                }
            }
            int mid = t.start + count / 2;
            if (axis_best >= 0) {
                const float scale = 12.0f / (cmax[axis_best] - cmin[axis_best]);
                mid = two_ended_partition(perm, t.start, t.end, axis_best, cmin[axis_best], scale, plane_best);
            }
            if (mid <= t.start || mid >= t.end) mid = t.start + count / 2;
            out_nodes[me] = finish_node(box, -1, -1, 0);
            // Right pushed first so the left subtree is numbered next.
            todo.push_back(Task{mid, t.end, t.depth + 1, me, 1});
            todo.push_back(Task{t.start, mid, t.depth + 1, me, 0});
        }
        *out_count = nodes;
    }

private:
    // Scan from the left for an element that belongs right, from the right for
    // one that belongs left, swap, continue until the scans meet.
    int two_ended_partition(std::vector<uint32_t> &perm, int lo, int hi, int axis, float cmin, float scale,
                            int plane) const
    {
        auto goes_left = [&](uint32_t prim) { return bin_of(mid_of(boxes_[prim], axis), cmin, scale) < plane; };
        int i = lo;
        int j = hi;
        for (;;) {
            while (i != j && goes_left(perm[i])) i++;
            if (i == j) return i;
            j--;
            while (i != j && !goes_left(perm[j])) j--;
            if (i == j) return i;
            std::swap(perm[i], perm[j]);
            i++;
        }
    }

    const Box *boxes_;
    uint32_t n_;
};

inline Box to_box(const AvrtAABB &a)
{
    return Box{{a.min_x, a.min_y, a.min_z}, {a.max_x, a.max_y, a.max_z}};
}

} // namespace

extern "C" int avrt_build_bvh(const AvrtAABB *bounds, uint32_t n, AvrtBVHNode *out_nodes, uint32_t *out_node_count,
                              uint32_t *out_order)
{
    if (out_node_count == nullptr) return AVRT_E_INVALID_ARG;
    *out_node_count = 0;
    if (n == 0) return AVRT_OK;
    if (bounds == nullptr || out_nodes == nullptr || out_order == nullptr) return AVRT_E_INVALID_ARG;
    std::vector<Box> boxes(n);
    for (uint32_t i = 0; i < n; i++) boxes[i] = to_box(bounds[i]);
    Builder b(boxes.data(), n);
    b.run(out_nodes, out_node_count, out_order);
    return AVRT_OK;
}

extern "C" int avrt_build_blas(const float *positions, uint32_t vertex_count, const uint32_t *tri_indices,
                               uint32_t tri_count, AvrtBVHNode *out_nodes, uint32_t *out_node_count,
                               uint32_t *out_indices, AvrtAABB *out_local_bounds)
{
    if (out_node_count == nullptr || out_local_bounds == nullptr) return AVRT_E_INVALID_ARG;
    if (vertex_count > 0 && positions == nullptr) return AVRT_E_INVALID_ARG;
    if (tri_count > 0 && (tri_indices == nullptr || out_nodes == nullptr || out_indices == nullptr))
        return AVRT_E_INVALID_ARG;
    // Vertex bounds, SurroundPoint per vertex (mesh_library.h:104-112).
    Box local = empty_box();
    for (uint32_t i = 0; i < vertex_count; i++)
        grow_point(local, positions[3 * i], positions[3 * i + 1], positions[3 * i + 2]);
    *out_local_bounds = AvrtAABB{local.lo[0], local.lo[1], local.lo[2], local.hi[0], local.hi[1], local.hi[2]};
    *out_node_count = 0;
    if (tri_count == 0) return AVRT_OK;
    // Triangle boxes: Surround2Points(a, b) then SurroundPoint(c)
    // (mesh_library.h:124-132).
    std::vector<Box> tb(tri_count);
    for (uint32_t t = 0; t < tri_count; t++) {
        const uint32_t ia = tri_indices[3 * t], ib = tri_indices[3 * t + 1], ic = tri_indices[3 * t + 2];
        if (ia >= vertex_count || ib >= vertex_count || ic >= vertex_count) return AVRT_E_INVALID_ARG;
        const float *a = positions + 3 * ia;
        const float *b = positions + 3 * ib;
        const float *c = positions + 3 * ic;
        Box e;
        for (int k = 0; k < 3; k++) {
            e.lo[k] = std::min(a[k], b[k]);
            e.hi[k] = std::max(a[k], b[k]);
        }
        grow_point(e, c[0], c[1], c[2]);
        tb[t] = e;
    }
    std::vector<uint32_t> order(tri_count);
    Builder builder(tb.data(), tri_count);
    builder.run(out_nodes, out_node_count, order.data());
    // The index buffer follows the BLAS slot order (mesh_library.h:147-153).
    for (uint32_t s = 0; s < tri_count; s++) {
        const uint32_t t = order[s];
        out_indices[3 * s + 0] = tri_indices[3 * t + 0];
        out_indices[3 * s + 1] = tri_indices[3 * t + 1];
        out_indices[3 * s + 2] = tri_indices[3 * t + 2];
    }
    return AVRT_OK;
}

// v rotated by unit quaternion q: v + w*t + cross(u, t), t = 2 cross(u, v)
// (math/quat.hpp:212-217).
static inline void quat_apply(float qx, float qy, float qz, float qw, const float v[3], float out[3])
{
    const float tx = 2.0f * (qy * v[2] - qz * v[1]);
    const float ty = 2.0f * (qz * v[0] - qx * v[2]);
    const float tz = 2.0f * (qx * v[1] - qy * v[0]);
    out[0] = v[0] + qw * tx + (qy * tz - qz * ty);
    out[1] = v[1] + qw * ty + (qz * tx - qx * tz);
    out[2] = v[2] + qw * tz + (qx * ty - qy * tx);
}

extern "C" int avrt_instance_bounds(const AvrtMeshInstance *instances, uint32_t n, const AvrtAABB *local,
                                    AvrtAABB *out)
{
    if (n == 0) return AVRT_OK;
    if (instances == nullptr || local == nullptr || out == nullptr) return AVRT_E_INVALID_ARG;
    for (uint32_t i = 0; i < n; i++) {
        const AvrtAABB &b = local[i];
        Box w = empty_box();
        if (b.min_x > b.max_x) {
            out[i] = AvrtAABB{w.lo[0], w.lo[1], w.lo[2], w.hi[0], w.hi[1], w.hi[2]};
            continue;
        }
        const AvrtMeshInstance &q = instances[i];
        const float s = q.scale == 0.0f ? 1.0f : q.scale;
        for (int c = 0; c < 8; c++) {
            float p[3] = {(c & 1) ? b.max_x : b.min_x, (c & 2) ? b.max_y : b.min_y, (c & 4) ? b.max_z : b.min_z};
            if (s != 1.0f) {
                p[0] = p[0] * s;
                p[1] = p[1] * s;
                p[2] = p[2] * s;
            }
            float r[3];
            quat_apply(q.qx, q.qy, q.qz, q.qw, p, r);
            grow_point(w, q.px + r[0], q.py + r[1], q.pz + r[2]);
        }
        out[i] = AvrtAABB{w.lo[0], w.lo[1], w.lo[2], w.hi[0], w.hi[1], w.hi[2]};
    }
    return AVRT_OK;
}
