/*
 * pt_host.cpp -- host side of libpentrack_gpu.so: world-model builders (BVH,
 * field tables, tricubic pre-interpolation), device memory of a particle batch,
 * and the C ABI declared in include/pentrack_gpu.h.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <climits>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pentrack_gpu.h"
#include "pt_decay.h"
#include "pt_device.h"
#include "pt_formula.h"
#include "pt_kernels.h"

namespace ptf {
bool compile(const char *formula, std::vector<uint32_t> &code, std::vector<double> &cst, std::string &err);
bool compile_x(const char *formula, std::vector<uint32_t> &code, std::vector<double> &cst, std::string &err);
}

namespace {

thread_local std::string g_last_error;
std::atomic<int> g_quit{0};

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define PT_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) return fail(PT_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

/* device allocation list owned by a handle */
struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **out, size_t count) {
        void *p = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail(PT_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        ptrs.push_back(p);
        *out = static_cast<T *>(p);
        return PT_OK;
    }
    template <class T>
    int upload(T **out, const T *src, size_t count) {
        int r = alloc(out, count);
        if (r) return r;
        if (count) {
            hipError_t e = hipMemcpy(*out, src, count * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(PT_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        }
        return PT_OK;
    }
};

} // namespace

/* ======================================================================== */
/* geometry                                                                  */
/* ======================================================================== */

struct pt_geometry {
    DevMem mem;
    ptd::Geom g{};
    int n_nodes = 0, depth = 0;
    int n_solid = 0;
};

namespace {

struct BuildPrim {
    float lo[3], hi[3];
    float c[3];
    int idx;
};

/* median split on the longest axis of the centroid box (the AABB_traits split of
   the reference's CGAL tree, AABB_traits.h:296-328, applied to centroids) */
int build_bvh(std::vector<BuildPrim> &prims, int begin, int end, int leaf_size, std::vector<ptd::BvhNode> &nodes,
              std::vector<int> &order, int depth, int &max_depth) {
    int me = (int)nodes.size();
    nodes.emplace_back();
    max_depth = std::max(max_depth, depth);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = begin; i < end; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], prims[i].lo[a]);
            hi[a] = std::max(hi[a], prims[i].hi[a]);
            clo[a] = std::min(clo[a], prims[i].c[a]);
            chi[a] = std::max(chi[a], prims[i].c[a]);
        }
    ptd::BvhNode nd;
    for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[a]; nd.hi[a] = hi[a]; }
    if (end - begin <= leaf_size) {
        nd.a = -((int)order.size() + 1);
        nd.b = end - begin;
        for (int i = begin; i < end; ++i) order.push_back(prims[i].idx);
        nodes[me] = nd;
        return me;
    }
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > ext) { ext = chi[a] - clo[a]; axis = a; }
    int mid = begin + (end - begin) / 2;
    std::nth_element(prims.begin() + begin, prims.begin() + mid, prims.begin() + end,
                     [axis](const BuildPrim &p, const BuildPrim &q) {
                         if (p.c[axis] != q.c[axis]) return p.c[axis] < q.c[axis];
                         return p.idx < q.idx;
                     });
    int left = build_bvh(prims, begin, mid, leaf_size, nodes, order, depth + 1, max_depth);
    int right = build_bvh(prims, mid, end, leaf_size, nodes, order, depth + 1, max_depth);
    nd.a = left;
    nd.b = right;
    nodes[me] = nd;
    return me;
}

/* squared distance from point p to triangle (a, b, c): closest point by the
   Voronoi regions of the vertices, edges and face */
double point_tri_dist2(const double *p, const double *a, const double *b, const double *c) {
    auto sub = [](const double *u, const double *v, double *w) { w[0] = u[0] - v[0]; w[1] = u[1] - v[1]; w[2] = u[2] - v[2]; };
    auto dot = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double ab[3], ac[3], ap[3], q[3];
    sub(b, a, ab); sub(c, a, ac); sub(p, a, ap);
    auto at = [&](double s, double t) {
        for (int k = 0; k < 3; ++k) q[k] = a[k] + s * ab[k] + t * ac[k];
        double d[3];
        sub(p, q, d);
        return dot(d, d);
    };
    double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return at(0, 0);
    double bp[3];
    sub(p, b, bp);
    double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return at(1, 0);
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return at(d1 / (d1 - d3), 0);
    double cp[3];
    sub(p, c, cp);
    double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return at(0, 1);
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return at(0, d2 / (d2 - d6));
    double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return at(1 - w, w);
    }
    double den = 1.0 / (va + vb + vc);
    if (!std::isfinite(den)) {   /* degenerate triangle: nearest of its edges */
        double best = INFINITY;
        const double *V[3] = {a, b, c};
        for (int e = 0; e < 3; ++e) {
            const double *u = V[e], *v = V[(e + 1) % 3];
            double uv[3], up[3];
            sub(v, u, uv); sub(p, u, up);
            double l2 = dot(uv, uv), t = l2 > 0 ? std::min(1.0, std::max(0.0, dot(up, uv) / l2)) : 0.0;
            double d[3] = {up[0] - t * uv[0], up[1] - t * uv[1], up[2] - t * uv[2]};
            best = std::min(best, dot(d, d));
        }
        return best;
    }
    return at(vb * den, vc * den);
}

/* distance from p to the nearest triangle: branch and bound over the BVH */
double nearest_tri_dist(const double *p, const std::vector<ptd::BvhNode> &nodes, const std::vector<double> &tv) {
    double best2 = INFINITY;
    int stack[ptd::BVH_STACK];
    int sp = 0;
    stack[sp++] = 0;
    auto box_d2 = [&](const ptd::BvhNode &n) {
        double d2 = 0;
        for (int a = 0; a < 3; ++a) {
            double d = std::max(std::max((double)n.lo[a] - p[a], p[a] - (double)n.hi[a]), 0.0);
            d2 += d * d;
        }
        return d2;
    };
    while (sp > 0) {
        const ptd::BvhNode &n = nodes[stack[--sp]];
        if (box_d2(n) >= best2) continue;
        if (n.a < 0) {
            int first = -n.a - 1;
            for (int k = first; k < first + n.b; ++k) {
                const double *T = &tv[(size_t)9 * k];
                best2 = std::min(best2, point_tri_dist2(p, T, T + 3, T + 6));
            }
        } else {
            double dl = box_d2(nodes[n.a]), dr = box_d2(nodes[n.b]);
            /* nearer child on top */
            if (dl < dr) { if (dr < best2) stack[sp++] = n.b; if (dl < best2) stack[sp++] = n.a; }
            else { if (dl < best2) stack[sp++] = n.a; if (dr < best2) stack[sp++] = n.b; }
        }
    }
    return std::sqrt(best2);
}

/*
 * Clearance grid over the mesh's bounding box: for each cell a lower bound of the
 * distance from any point of the cell to any triangle (distance from the cell
 * centre to the nearest triangle minus half the cell diagonal, rounded down to
 * float).  A segment that starts in a cell and is shorter than its bound crosses
 * no triangle, so the tracker skips the BVH query for it.
 */
void build_clearance(const std::vector<ptd::BvhNode> &nodes, const std::vector<double> &tv, const double lo[3],
                     const double hi[3], int cells_long, std::vector<float> &clear, int gn[3], double ginv_h[3],
                     double *hs_out = nullptr) {
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    double h = ext / cells_long;
    double hs[3];
    for (int a = 0; a < 3; ++a) {
        gn[a] = std::max(1, (int)std::ceil((hi[a] - lo[a]) / h));
        hs[a] = (hi[a] - lo[a]) / gn[a];
        if (!(hs[a] > 0)) { hs[a] = h; }
        ginv_h[a] = 1.0 / hs[a];
    }
    if (hs_out)
        for (int a = 0; a < 3; ++a) hs_out[a] = hs[a];
    double half_diag = 0.5 * std::sqrt(hs[0] * hs[0] + hs[1] * hs[1] + hs[2] * hs[2]);
    size_t ncell = (size_t)gn[0] * gn[1] * gn[2];
    clear.assign(ncell, 0.f);
    int threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    auto work = [&](int t) {
        for (size_t cidx = (size_t)t; cidx < ncell; cidx += (size_t)threads) {
            int iz = (int)(cidx % gn[2]);
            int iy = (int)((cidx / gn[2]) % gn[1]);
            int ix = (int)(cidx / ((size_t)gn[2] * gn[1]));
            double p[3] = {lo[0] + (ix + 0.5) * hs[0], lo[1] + (iy + 0.5) * hs[1], lo[2] + (iz + 0.5) * hs[2]};
            double b = nearest_tri_dist(p, nodes, tv) - half_diag;
            b = b > 0 ? b * (1 - 1e-6) : 0.0;
            float f = (float)b;
            if ((double)f > b) f = std::nextafter(f, 0.f);
            clear[cidx] = f;
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work, t);
    for (auto &t : th) t.join();
}

/* Candidate lists of the clearance cells (Geom::cl_rng, cl_tri): for every cell, the
   triangles in leaf order whose boxes, padded as the BVH traversals pad them, meet the
   cell widened by pad on every side (and by a small slack beyond, so that the device's
   own rounding of the cell bounds stays inside).  A cell with more than CELL_LIST_MAX
   triangles keeps no list (count -1). */
template <class Fn>
void visit_cells(const ptd::Tri &t, const double glo[3], const double hs[3], const int gn[3], double pad, Fn &&fn) {
    double blo[3], bhi[3];
    int ilo[3], ihi[3];
    for (int a = 0; a < 3; ++a) {
        const double v0 = t.v[a], v1 = t.v[3 + a], v2 = t.v[6 + a];
        const double tlo = std::min(v0, std::min(v1, v2)), thi = std::max(v0, std::max(v1, v2));
        const double tp = 1e-9 + 1e-7 * (std::fabs(tlo) + std::fabs(thi));
        blo[a] = tlo - tp;
        bhi[a] = thi + tp;
        ilo[a] = std::max(0, (int)std::floor((blo[a] - pad - glo[a]) / hs[a]) - 1);
        ihi[a] = std::min(gn[a] - 1, (int)std::floor((bhi[a] + pad - glo[a]) / hs[a]) + 1);
    }
    auto meets = [&](int a, int i) {
        const double slack = 1e-9 + 1e-6 * hs[a];
        const double clo = glo[a] + i * hs[a] - pad - slack, chi = glo[a] + (i + 1) * hs[a] + pad + slack;
        return blo[a] <= chi && bhi[a] >= clo;
    };
    for (int i = ilo[0]; i <= ihi[0]; ++i) {
        if (!meets(0, i)) continue;
        for (int j = ilo[1]; j <= ihi[1]; ++j) {
            if (!meets(1, j)) continue;
            for (int l = ilo[2]; l <= ihi[2]; ++l)
                if (meets(2, l)) fn(((size_t)i * gn[1] + j) * gn[2] + l);
        }
    }
}

void build_cell_lists(const std::vector<ptd::Tri> &tris, const double glo[3], const double hs[3], const int gn[3],
                      double pad, std::vector<int2> &rng, std::vector<ptd::Tri> &out) {
    const size_t ncell = (size_t)gn[0] * gn[1] * gn[2];
    std::vector<int> cnt(ncell, 0);
    auto visit = [&](const ptd::Tri &t, auto &&fn) { visit_cells(t, glo, hs, gn, pad, fn); };
    for (const ptd::Tri &t : tris) visit(t, [&](size_t c) { ++cnt[c]; });
    rng.assign(ncell, int2{0, -1});
    size_t total = 0;
    for (size_t c = 0; c < ncell; ++c)
        if (cnt[c] <= ptd::CELL_LIST_MAX) { rng[c] = int2{(int)total, cnt[c]}; total += (size_t)cnt[c]; }
    out.resize(total);
    std::vector<int> pos(ncell, 0);
    for (size_t c = 0; c < ncell; ++c) pos[c] = rng[c].x;
    for (const ptd::Tri &t : tris)
        visit(t, [&](size_t c) { if (rng[c].y >= 0) out[(size_t)pos[c]++] = t; });
}

/* Per cell the deepest BVH node whose subtree holds every triangle whose padded box meets
   the cell widened by pad (Geom::cl_root; -2 where there is none): the triangles of a subtree
   are a contiguous run of the leaf order, so the node is found by descending from the root
   while one child's run contains the cell's lowest and highest leaf-order index. */
void build_cell_roots(const std::vector<ptd::BvhNode> &nodes, const std::vector<ptd::Tri> &tris, const double glo[3],
                      const double hs[3], const int gn[3], double pad, std::vector<int> &root) {
    const size_t ncell = (size_t)gn[0] * gn[1] * gn[2];
    std::vector<int> kmin(ncell, INT_MAX), kmax(ncell, -1);
    for (int k = 0; k < (int)tris.size(); ++k)
        visit_cells(tris[k], glo, hs, gn, pad, [&](size_t c) {
            kmin[c] = std::min(kmin[c], k);
            kmax[c] = std::max(kmax[c], k);
        });
    /* leaf-order runs of the subtrees (host nodes: a = left child, b = right child) */
    std::vector<int> lo(nodes.size()), hi(nodes.size());
    for (int n = (int)nodes.size() - 1; n >= 0; --n) {
        if (nodes[n].a < 0) { lo[n] = -nodes[n].a - 1; hi[n] = lo[n] + nodes[n].b - 1; }
        else { lo[n] = lo[nodes[n].a]; hi[n] = hi[nodes[n].b]; }
    }
    root.assign(ncell, -2);
    for (size_t c = 0; c < ncell; ++c) {
        if (kmax[c] < 0) continue;
        int n = 0;
        while (nodes[n].a >= 0) {
            const int l = nodes[n].a, r = nodes[n].b;
            if (lo[l] <= kmin[c] && kmax[c] <= hi[l]) n = l;
            else if (lo[r] <= kmin[c] && kmax[c] <= hi[r]) n = r;
            else break;
        }
        root[c] = n;
    }
}

} // namespace

extern "C" int pt_geometry_create(int n_tri, const float *tri_vertices, const int32_t *tri_solid, int n_solid,
                                  const uint32_t *solid_id, const int32_t *solid_mat, const int32_t *ign_off,
                                  const double *ign, int default_mat, int n_mat, const double *mats, int leaf_size,
                                  pt_geometry **out) {
    if (!out || n_tri < 0 || n_solid < 0 || n_solid > 64 || n_mat <= 0 || !mats) return fail(PT_EINVAL, "pt_geometry_create: bad sizes");
    if (n_tri > 0 && (!tri_vertices || !tri_solid)) return fail(PT_EINVAL, "pt_geometry_create: null triangle arrays");
    if (n_solid > 0 && (!solid_id || !solid_mat || !ign_off)) return fail(PT_EINVAL, "pt_geometry_create: null solid arrays");
    if (default_mat < 0 || default_mat >= n_mat) return fail(PT_EINVAL, "pt_geometry_create: default material out of range");
    if (leaf_size < 1 || leaf_size > 32) return fail(PT_EINVAL, "pt_geometry_create: leaf_size must be 1..32");
    for (int s = 0; s < n_solid; ++s) {
        if (solid_id[s] <= 1) return fail(PT_EINVAL, "pt_geometry_create: solid IDs must be > 1");
        if (solid_mat[s] < 0 || solid_mat[s] >= n_mat) return fail(PT_EINVAL, "pt_geometry_create: solid material out of range");
        for (int t = 0; t < s; ++t)
            if (solid_id[t] == solid_id[s]) return fail(PT_EINVAL, "pt_geometry_create: duplicate solid ID");
        if (ign_off[s + 1] < ign_off[s]) return fail(PT_EINVAL, "pt_geometry_create: bad ignore offsets");
    }
    for (int i = 0; i < n_tri; ++i)
        if (tri_solid[i] < 0 || tri_solid[i] >= n_solid) return fail(PT_EINVAL, "pt_geometry_create: triangle solid index out of range");

    std::unique_ptr<pt_geometry> G(new pt_geometry());
    /* vertex snap |v| < 1e-8 -> 0 (src/trianglemesh.cpp:42) */
    std::vector<float> v(tri_vertices, tri_vertices + (size_t)9 * n_tri);
    for (float &c : v)
        if (std::fabs(c) < 1e-8) c = 0.f;
    std::vector<BuildPrim> prims(n_tri);
    std::vector<double> bbox((size_t)6 * std::max(n_solid, 1));
    for (int s = 0; s < n_solid; ++s) {
        for (int a = 0; a < 3; ++a) { bbox[6 * s + a] = INFINITY; bbox[6 * s + 3 + a] = -INFINITY; }
    }
    std::vector<double> nrm((size_t)3 * n_tri);
    for (int i = 0; i < n_tri; ++i) {
        const float *T = &v[(size_t)9 * i];
        BuildPrim &p = prims[i];
        p.idx = i;
        for (int a = 0; a < 3; ++a) {
            p.lo[a] = std::min(T[a], std::min(T[3 + a], T[6 + a]));
            p.hi[a] = std::max(T[a], std::max(T[3 + a], T[6 + a]));
            p.c[a] = (p.lo[a] + p.hi[a]) * 0.5f;
            double *b = &bbox[6 * tri_solid[i]];
            b[a] = std::min(b[a], (double)p.lo[a]);
            b[3 + a] = std::max(b[3 + a], (double)p.hi[a]);
        }
        /* CGAL compute_face_normal: 0.5*cross(v1-v0, v2-v0), normalised */
        double v0[3] = {T[0], T[1], T[2]}, v1[3] = {T[3], T[4], T[5]}, v2[3] = {T[6], T[7], T[8]};
        double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        n[0] = n[0] * 0.5; n[1] = n[1] * 0.5; n[2] = n[2] * 0.5;
        double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len != 0) { n[0] = n[0] / len; n[1] = n[1] / len; n[2] = n[2] / len; }
        nrm[3 * i] = n[0]; nrm[3 * i + 1] = n[1]; nrm[3 * i + 2] = n[2];
    }
    std::vector<ptd::BvhNode> nodes;
    std::vector<int> order;
    int max_depth = 0;
    if (n_tri > 0) build_bvh(prims, 0, n_tri, leaf_size, nodes, order, 0, max_depth);
    if (max_depth + 2 >= ptd::BVH_STACK) return fail(PT_EINVAL, "pt_geometry_create: BVH too deep");
    std::vector<ptd::Tri> tris(n_tri);
    std::vector<double> nrm_ord((size_t)3 * n_tri);
    for (int k = 0; k < n_tri; ++k) {
        int i = order[k];
        std::memcpy(tris[k].v, &v[(size_t)9 * i], 9 * sizeof(float));
        tris[k].solid = tri_solid[i];
        tris[k].orig = i;
        tris[k].pad = 0;
        for (int a = 0; a < 3; ++a) nrm_ord[3 * k + a] = nrm[3 * i + a];
    }
    ptd::Geom &g = G->g;
    g.n_tri = n_tri;
    g.n_node = (int)nodes.size();
    g.n_solid = n_solid;
    g.default_mat = default_mat;
    int r;
    ptd::Tri *dtri; double *dnrm; ptd::BvhNode *dnode; uint32_t *dsid; int *dsmat, *dign_off; double *dign, *dmats, *dbbox;
    std::vector<int> ioff(n_solid + 1, 0);
    for (int s = 0; s <= n_solid; ++s) ioff[s] = n_solid ? ign_off[s] - ign_off[0] : 0;
    int n_ign = n_solid ? ign_off[n_solid] - ign_off[0] : 0;
    std::vector<int> smat(solid_mat, solid_mat + n_solid);
    if ((r = G->mem.upload(&dtri, tris.data(), tris.size()))) return r;
    if ((r = G->mem.upload(&dnrm, nrm_ord.data(), nrm_ord.size()))) return r;
    {
        /* device copy of the nodes with escape indices in place of left children: the end of
           a node's subtree in preorder, a leaf's is the next node */
        std::vector<ptd::BvhNode> dev(nodes);
        std::vector<int> end(nodes.size());
        for (int k = (int)nodes.size() - 1; k >= 0; --k) {
            if (nodes[k].a < 0) { end[k] = k + 1; continue; }
            if (nodes[k].a != k + 1 || nodes[k].b <= k || nodes[k].b >= (int)nodes.size())
                return fail(PT_EINVAL, "pt_geometry_create: BVH nodes not in preorder");
            end[k] = end[nodes[k].b];
            dev[k].a = end[k];
        }
        if ((r = G->mem.upload(&dnode, dev.data(), dev.size()))) return r;
    }
    if ((r = G->mem.upload(&dsid, solid_id, (size_t)n_solid))) return r;
    if ((r = G->mem.upload(&dsmat, smat.data(), smat.size()))) return r;
    if ((r = G->mem.upload(&dign_off, ioff.data(), ioff.size()))) return r;
    if ((r = G->mem.upload(&dign, n_ign ? ign + 2 * ign_off[0] : nullptr, (size_t)2 * n_ign))) return r;
    if ((r = G->mem.upload(&dmats, mats, (size_t)9 * n_mat))) return r;
    if ((r = G->mem.upload(&dbbox, bbox.data(), bbox.size()))) return r;
    g.tri = dtri; g.nrm = dnrm; g.node = dnode; g.solid_id = dsid; g.solid_mat = dsmat;
    g.ign_off = dign_off; g.ign = dign; g.mats = dmats; g.mesh_bbox = dbbox;
    g.clear = nullptr;
    g.cl_rng = nullptr; g.cl_tri = nullptr; g.cl_pad = 0.0;
    g.cl_root = nullptr; g.rt_pad = 0.0;
    if (n_tri > 0) {
        std::vector<double> tv((size_t)9 * n_tri);
        double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < n_tri; ++k)
            for (int c = 0; c < 9; ++c) {
                tv[(size_t)9 * k + c] = tris[k].v[c];
                glo[c % 3] = std::min(glo[c % 3], (double)tris[k].v[c]);
                ghi[c % 3] = std::max(ghi[c % 3], (double)tris[k].v[c]);
            }
        std::vector<float> clear;
        double hs[3];
        /* 128 cells on the longest axis: 2.36e8 RK steps/s on the benchmark against 2.30e8
           with 96 and no gain from 160 */
        build_clearance(nodes, tv, glo, ghi, 128, clear, g.gn, g.ginv_h, hs);
        for (int a = 0; a < 3; ++a) g.glo[a] = glo[a];
        float *dclear;
        if ((r = G->mem.upload(&dclear, clear.data(), clear.size()))) return r;
        g.clear = dclear;
        /* candidate lists of the cells widened by two cells on every side */
        std::vector<int2> rng;
        std::vector<ptd::Tri> ltri;
        const double pad = 2.0 * std::max(hs[0], std::max(hs[1], hs[2]));
        build_cell_lists(tris, glo, hs, g.gn, pad, rng, ltri);
        if (ltri.size() < (size_t)INT_MAX) {
            int2 *drng; ptd::Tri *dltri;
            if ((r = G->mem.upload(&drng, rng.data(), rng.size()))) return r;
            if ((r = G->mem.upload(&dltri, ltri.data(), ltri.size()))) return r;
            g.cl_rng = drng; g.cl_tri = dltri; g.cl_pad = pad;
        }
        /* BVH subtree roots of the cells widened by four cells on every side */
        std::vector<int> roots;
        const double rpad = 2.0 * pad;
        build_cell_roots(nodes, tris, glo, hs, g.gn, rpad, roots);
        int *droot;
        if ((r = G->mem.upload(&droot, roots.data(), roots.size()))) return r;
        g.cl_root = droot; g.rt_pad = rpad;
    }
    G->n_nodes = (int)nodes.size();
    G->depth = max_depth;
    G->n_solid = n_solid;
    *out = G.release();
    return PT_OK;
}

extern "C" void pt_geometry_destroy(pt_geometry *g) { delete g; }

/* host-only: the clearance grid pt_geometry_create builds (for tests and diagnostics);
   with out == NULL only gn, lo and inv_h are returned */
extern "C" int pt_clearance_grid(int n_tri, const float *tri_vertices, int cells_long, float *out, int32_t *gn,
                                 double *lo, double *inv_h) {
    if (n_tri <= 0 || !tri_vertices || cells_long < 1 || cells_long > 1024 || !gn || !lo || !inv_h)
        return fail(PT_EINVAL, "pt_clearance_grid: bad arguments");
    std::vector<BuildPrim> prims(n_tri);
    double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < n_tri; ++i) {
        const float *T = tri_vertices + (size_t)9 * i;
        BuildPrim &p = prims[i];
        p.idx = i;
        for (int a = 0; a < 3; ++a) {
            p.lo[a] = std::min(T[a], std::min(T[3 + a], T[6 + a]));
            p.hi[a] = std::max(T[a], std::max(T[3 + a], T[6 + a]));
            p.c[a] = (p.lo[a] + p.hi[a]) * 0.5f;
            glo[a] = std::min(glo[a], (double)p.lo[a]);
            ghi[a] = std::max(ghi[a], (double)p.hi[a]);
        }
    }
    std::vector<ptd::BvhNode> nodes;
    std::vector<int> order;
    int depth = 0;
    build_bvh(prims, 0, n_tri, 4, nodes, order, 0, depth);
    if (depth + 2 >= ptd::BVH_STACK) return fail(PT_EINVAL, "pt_clearance_grid: BVH too deep");
    std::vector<double> tv((size_t)9 * n_tri);
    for (int k = 0; k < n_tri; ++k)
        for (int c = 0; c < 9; ++c) tv[(size_t)9 * k + c] = tri_vertices[(size_t)9 * order[k] + c];
    std::vector<float> clear;
    int g[3];
    double ih[3];
    if (!out) {
        double ext = std::max(ghi[0] - glo[0], std::max(ghi[1] - glo[1], ghi[2] - glo[2]));
        for (int a = 0; a < 3; ++a) {
            g[a] = std::max(1, (int)std::ceil((ghi[a] - glo[a]) / (ext / cells_long)));
            gn[a] = g[a]; lo[a] = glo[a];
        }
        return PT_OK;
    }
    build_clearance(nodes, tv, glo, ghi, cells_long, clear, g, ih);
    std::memcpy(out, clear.data(), clear.size() * sizeof(float));
    for (int a = 0; a < 3; ++a) { gn[a] = g[a]; lo[a] = glo[a]; inv_h[a] = ih[a]; }
    return PT_OK;
}

extern "C" int pt_geometry_info(const pt_geometry *g, int *n_nodes, int *depth) {
    if (!g) return fail(PT_EINVAL, "null geometry");
    if (n_nodes) *n_nodes = g->n_nodes;
    if (depth) *depth = g->depth;
    return PT_OK;
}

/* ======================================================================== */
/* fields                                                                    */
/* ======================================================================== */

struct pt_fields {
    DevMem mem;
    std::vector<ptd::Field> host;
    ptd::Field *dev = nullptr;
    int n = 0;
    bool has_fp32 = false;  /* mixed-precision tables built */
    long long mix_blocks = 0, mix_fp64 = 0;   /* blocks of the mixed tables, of which kept in FP64 */
    bool general = false;   /* a scale formula, analytic field or conductor is present */
    bool b_reuse = true;    /* no 2D map tabulates B (value-only B equals B with derivatives) */
};

namespace {
int tricubic_on_device(int nx, int ny, int nz, const double *const gh[3], int ncomp, const double *values_h,
                       int device_order, double *coef_dev);
}

extern "C" int pt_fields_create(int n, const pt_field_desc *descs, int with_fp32, pt_fields **out) {
    if (!out || n < 0 || (n > 0 && !descs)) return fail(PT_EINVAL, "pt_fields_create: bad arguments");
    if (n > ptd::MAX_FIELDS)
        return fail(PT_EINVAL, "pt_fields_create: at most " + std::to_string(ptd::MAX_FIELDS) + " field containers");
    std::unique_ptr<pt_fields> F(new pt_fields());
    F->n = n;
    F->has_fp32 = with_fp32 != 0;
    F->host.resize(std::max(n, 1));
    for (int k = 0; k < n; ++k) {
        const pt_field_desc &d = descs[k];
        ptd::Field f{};
        f.type = d.type;
        f.bscale = d.bscale; f.escale = d.escale; f.bw = d.bw;
        for (int a = 0; a < 6; ++a) f.box[a] = d.box[a];
        for (int a = 0; a < 3; ++a) f.e[a] = d.e[a];
        for (int a = 0; a < 8; ++a) f.p[a] = d.p[a];
        {
            /* TFieldBoundaryBox constructor check (src/field.cpp:116-125) */
            const double *b = d.box;
            if (d.bw < 0 || d.bw > (b[0] - b[1]) / 2 || d.bw > (b[2] - b[3]) / 2 || d.bw > (b[4] - b[5]) / 2)
                return fail(PT_EINVAL, "pt_fields_create: You defined a field with an invalid boundary. Check field parameters.");
        }
        /* TFieldScaler formulas (src/field.cpp:34-45): compiled to device programs */
        const char *forms[2] = {d.bscale_formula, d.escale_formula};
        for (int w = 0; w < 2; ++w) {
            if (!forms[w]) continue;
            std::vector<uint32_t> code;
            std::vector<double> cst;
            std::string err;
            if (!ptf::compile(forms[w], code, cst, err))
                return fail(PT_EINVAL, "pt_fields_create: field " + std::to_string(k) + ": " + err);
            if (code.size() == 1 && (code[0] & 0xffu) == PF_CONST) {   /* a constant formula */
                (w == 0 ? f.bscale : f.escale) = cst[0];
                continue;
            }
            uint32_t *dcode;
            double *dcst;
            int r = F->mem.upload(&dcode, code.data(), code.size());
            if (r) return r;
            r = F->mem.upload(&dcst, cst.data(), cst.size());
            if (r) return r;
            if (w == 0) { f.bcode = dcode; f.bcst = dcst; f.blen = (int)code.size(); }
            else { f.ecode = dcode; f.ecst = dcst; f.elen = (int)code.size(); }
        }
        if (d.type == PT_FIELD_CONDUCTOR) {
            /* TConductorField constructor (src/conductor.cpp:9-19) and the prefactor
               mu0 I / (4 pi) in long double (src/conductor.cpp:35, src/globals.cpp:13,19) */
            if (d.p[1] == d.p[4] && d.p[2] == d.p[5] && d.p[3] == d.p[6])
                return fail(PT_EINVAL, "pt_fields_create: You have defined a conductor with length zero! Aborting...");
            const long double pi = 3.1415926535897932384626L, mu0 = 4 * pi * 1e-7L;
            f.p[7] = (double)(mu0 * (long double)d.p[0] / (4 * pi));
        }
        if (d.type == PT_FIELD_TRICUBIC) {
            if (d.nx < 2 || d.ny < 2 || d.nz < 2 || !d.xg || !d.yg || !d.zg || (!d.coef && !d.values) ||
                (!d.has_b && !d.has_v))
                return fail(PT_EINVAL, "pt_fields_create: bad tricubic descriptor");
            const double *gs[3] = {d.xg, d.yg, d.zg};
            int ns[3] = {d.nx, d.ny, d.nz};
            for (int a = 0; a < 3; ++a) {
                for (int i = 1; i < ns[a]; ++i)
                    if (!(gs[a][i] > gs[a][i - 1])) return fail(PT_EINVAL, "pt_fields_create: grid not strictly increasing");
                double *dg;
                int r = F->mem.upload(&dg, gs[a], (size_t)ns[a]);
                if (r) return r;
                f.g[a] = dg;
                f.x0[a] = gs[a][0];
                f.inv_h[a] = (double)(ns[a] - 1) / (gs[a][ns[a] - 1] - gs[a][0]);
            }
            f.nx = d.nx; f.ny = d.ny; f.nz = d.nz;
            f.has_b = d.has_b ? 1 : 0;
            f.has_v = d.has_v ? 1 : 0;
            f.ncomp = 3 * f.has_b + f.has_v;
            size_t ncoef = (size_t)(d.nx - 1) * (d.ny - 1) * (d.nz - 1) * f.ncomp * 64;
            double *dc;
            int r;
            if (!d.coef) {
                /* samples only: pre-interpolate on the device, straight into the tracker's order */
                if ((r = F->mem.alloc(&dc, ncoef))) return r;
                if ((r = tricubic_on_device(d.nx, d.ny, d.nz, gs, f.ncomp, d.values, 1, dc))) return r;
            } else {
                /* libtricubic order a[i + 4j + 16k] (coefficient of x^i y^j z^k) becomes
                   a[16i + 4j + k]: the 16 coefficients of one power of x are contiguous */
                std::vector<double> cdev(ncoef);
                for (size_t b = 0; b < ncoef; b += 64)
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j)
                            for (int k = 0; k < 4; ++k) cdev[b + 16 * i + 4 * j + k] = d.coef[b + i + 4 * j + 16 * k];
                if ((r = F->mem.upload(&dc, cdev.data(), ncoef))) return r;
            }
            f.coef = dc;
            if (with_fp32) {
                /* the mixed-precision table, built on the device from the FP64 table */
                if (!(d.mixed_tol >= 0)) return fail(PT_EINVAL, "pt_fields_create: negative mixed_tol");
                const double tol = d.mixed_tol > 0 ? d.mixed_tol : PT_MIXED_TOL_DEFAULT;
                const size_t nblk = ncoef / 64;
                unsigned char *dm;
                int *dcount;
                const int zero = 0;
                if ((r = F->mem.alloc(&dm, nblk / f.ncomp * ptd::mix_cell_bytes(f.ncomp))) ||
                    (r = F->mem.upload(&dcount, &zero, 1)))
                    return r;
                PT_HIP(pt_launch_mixed_table(f.g[0], f.g[1], f.g[2], d.nx, d.ny, d.nz, f.ncomp, dc, tol, dm, dcount, 0));
                int nfp64 = 0;
                PT_HIP(hipMemcpy(&nfp64, dcount, sizeof(int), hipMemcpyDeviceToHost));
                f.mix = dm;
                F->mix_blocks += (long long)nblk;
                F->mix_fp64 += nfp64;
            }
        } else if (d.type == PT_FIELD_TAB2D) {
            /* TabField (src/field_2d.cpp): r grid in xg[nx], z grid in zg[nz] */
            if (d.nx < 2 || d.nz < 2 || !d.xg || !d.zg || !d.coef || (d.has_b & ~7) || (d.has_v & ~15) ||
                (!d.has_b && !d.has_v))
                return fail(PT_EINVAL, "pt_fields_create: bad 2D table descriptor");
            if ((d.has_v & 1) && (d.has_v & 14))
                return fail(PT_EINVAL, "pt_fields_create: 2D table with both V and E columns (the potential is ignored "
                                       "when E is tabulated: leave it out)");
            const double *gs[2] = {d.xg, d.zg};
            int ns[2] = {d.nx, d.nz};
            for (int a = 0; a < 2; ++a) {
                for (int i = 1; i < ns[a]; ++i)
                    if (!(gs[a][i] > gs[a][i - 1])) return fail(PT_EINVAL, "pt_fields_create: grid not strictly increasing");
                double *dg;
                int r = F->mem.upload(&dg, gs[a], (size_t)ns[a]);
                if (r) return r;
                f.g[2 * a] = dg;
                f.x0[2 * a] = gs[a][0];
                f.inv_h[2 * a] = (double)(ns[a] - 1) / (gs[a][ns[a] - 1] - gs[a][0]);
            }
            f.nx = d.nx; f.ny = 1; f.nz = d.nz;
            /* TabField::CheckTab (src/field_2d.cpp:199-206): the map's range from the first
               coordinate and the first spacing */
            f.p[0] = d.xg[0];
            f.p[1] = d.xg[0] + (d.xg[1] - d.xg[0]) * (d.nx - 1);
            f.p[2] = d.zg[0];
            f.p[3] = d.zg[0] + (d.zg[1] - d.zg[0]) * (d.nz - 1);
            f.has_b = d.has_b;
            f.has_v = d.has_v;
            f.ncomp = __builtin_popcount((unsigned)d.has_b) + __builtin_popcount((unsigned)d.has_v);
            size_t ncoef = (size_t)(d.nx - 1) * (d.nz - 1) * f.ncomp * 16;
            double *dc;
            int r = F->mem.upload(&dc, d.coef, ncoef);
            if (r) return r;
            f.coef = dc;
        } else if (d.type < PT_FIELD_UNIFORM_E || d.type > PT_FIELD_B0_XY) {
            return fail(PT_EINVAL, "pt_fields_create: unknown field type");
        }
        F->general = F->general || f.blen > 0 || f.elen > 0 || f.type >= PT_FIELD_CONDUCTOR;
        F->b_reuse = F->b_reuse && !(f.type == PT_FIELD_TAB2D && f.has_b);
        F->host[k] = f;
    }
    int r = F->mem.upload(&F->dev, F->host.data(), F->host.size());
    if (r) return r;
    *out = F.release();
    return PT_OK;
}

extern "C" void pt_fields_destroy(pt_fields *f) { delete f; }

/* blocks (cell x component) of the mixed-precision tables and those kept in FP64 */
extern "C" int pt_fields_mixed_info(const pt_fields *f, int64_t *blocks, int64_t *fp64_blocks) {
    if (!f) return fail(PT_EINVAL, "pt_fields_mixed_info: bad arguments");
    if (blocks) *blocks = f->mix_blocks;
    if (fp64_blocks) *fp64_blocks = f->mix_fp64;
    return PT_OK;
}

/* ======================================================================== */
/* tricubic pre-interpolation (TabField3::PreInterpol, src/field_3d.cpp:336-414) */
/* ======================================================================== */

namespace {

/* spline1dgriddiffcubic with parabolically terminated ends (boundltype = boundrtype = 0),
   ALGLIB interpolation.cpp:26709-26882 with the tridiagonal solver at :26970-27010 */
void spline_grid_diff(const double *x, const double *y, int n, double *d, std::vector<double> &buf) {
    if (n == 2) {
        d[0] = (y[1] - y[0]) / (x[1] - x[0]);
        d[1] = d[0];
        return;
    }
    buf.resize((size_t)4 * n);
    double *a1 = buf.data(), *a2 = a1 + n, *a3 = a2 + n, *b = a3 + n;
    a1[0] = 0; a2[0] = 1; a3[0] = 1;
    b[0] = 2 * (y[1] - y[0]) / (x[1] - x[0]);
    for (int i = 1; i <= n - 2; i++) {
        a1[i] = x[i + 1] - x[i];
        a2[i] = 2 * (x[i + 1] - x[i - 1]);
        a3[i] = x[i] - x[i - 1];
        b[i] = 3 * (y[i] - y[i - 1]) / (x[i] - x[i - 1]) * (x[i + 1] - x[i]) + 3 * (y[i + 1] - y[i]) / (x[i + 1] - x[i]) * (x[i] - x[i - 1]);
    }
    a1[n - 1] = 1; a2[n - 1] = 1; a3[n - 1] = 0;
    b[n - 1] = 2 * (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
    /* Thomas algorithm on (sub a1, diag a2, super a3, rhs b) */
    for (int k = 1; k <= n - 1; k++) {
        double t = a1[k] / a2[k - 1];
        a2[k] = a2[k] - t * a3[k - 1];
        b[k] = b[k] - t * b[k - 1];
    }
    d[n - 1] = b[n - 1] / a2[n - 1];
    for (int k = n - 2; k >= 0; k--) d[k] = (b[k] - a3[k] * d[k + 1]) / a2[k];
}

/* d/d(axis) of a [nx][ny][nz] table, one spline per grid line (TabField3::CalcDerivs) */
void calc_derivs(const std::vector<double> *xyz, int nx, int ny, int nz, const double *tab, int axis, double *out) {
    int n[3] = {nx, ny, nz};
    int len = n[axis];
    int d1 = (axis + 1) % 3, d2 = (axis + 2) % 3;
    std::vector<double> yv(len), dv(len), buf;
    size_t stride[3] = {(size_t)ny * nz, (size_t)nz, 1};
    for (int i1 = 0; i1 < n[d1]; ++i1)
        for (int i2 = 0; i2 < n[d2]; ++i2) {
            size_t base = (size_t)i1 * stride[d1] + (size_t)i2 * stride[d2];
            for (int k = 0; k < len; ++k) yv[k] = tab[base + (size_t)k * stride[axis]];
            spline_grid_diff(xyz[axis].data(), yv.data(), len, dv.data(), buf);
            for (int k = 0; k < len; ++k) out[base + (size_t)k * stride[axis]] = dv[k];
        }
}

/*
 * The 64x64 matrix mapping (f, fx, fy, fz, fxy, fxz, fyz, fxyz at the 8 cell
 * corners) to the coefficients a[i+4j+16k] of sum a x^i y^j z^k on the unit cube
 * (Lekien & Marsden, Int. J. Numer. Meth. Engng 63 (2005) 455).  It is the inverse
 * of the constraint matrix, computed here by exact Gauss-Jordan elimination over
 * the rationals (the inverse is an integer matrix).
 */
struct TricubicMatrix {
    int A[64][64];
    TricubicMatrix() {
        /* constraint matrix M[8*d + c][i + 4j + 16k] */
        static const int der[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        auto mono_d = [](int p, int m, int x) -> long long {
            /* d^m/dx^m x^p at x in {0,1} */
            if (p < m) return 0;
            long long c = 1;
            for (int q = 0; q < m; ++q) c *= (p - q);
            if (p - m == 0) return c;
            return x ? c : 0;
        };
        /* rational Gauss-Jordan with numerators/denominators in __int128 */
        typedef __int128 I;
        static I num[64][128], den[64][128];
        auto gcd = [](I a, I b) { if (a < 0) a = -a; if (b < 0) b = -b; while (b) { I t = a % b; a = b; b = t; } return a; };
        auto norm = [&](I &n, I &d) { if (d < 0) { n = -n; d = -d; } I g = gcd(n, d); if (g > 1) { n /= g; d /= g; } if (n == 0) d = 1; };
        for (int d = 0; d < 8; ++d)
            for (int c = 0; c < 8; ++c) {
                int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
                int row = 8 * d + c;
                for (int k = 0; k < 4; ++k)
                    for (int j = 0; j < 4; ++j)
                        for (int i = 0; i < 4; ++i) {
                            long long v = mono_d(i, der[d][0], cx) * mono_d(j, der[d][1], cy) * mono_d(k, der[d][2], cz);
                            num[row][i + 4 * j + 16 * k] = v;
                            den[row][i + 4 * j + 16 * k] = 1;
                        }
                for (int q = 0; q < 64; ++q) { num[row][64 + q] = (q == row); den[row][64 + q] = 1; }
            }
        for (int col = 0; col < 64; ++col) {
            int piv = col;
            while (num[piv][col] == 0) ++piv;
            if (piv != col)
                for (int q = 0; q < 128; ++q) { std::swap(num[piv][q], num[col][q]); std::swap(den[piv][q], den[col][q]); }
            I pn = num[col][col], pd = den[col][col];
            for (int q = 0; q < 128; ++q) { num[col][q] *= pd; den[col][q] *= pn; norm(num[col][q], den[col][q]); }
            for (int r = 0; r < 64; ++r) {
                if (r == col || num[r][col] == 0) continue;
                I fn = num[r][col], fd = den[r][col];
                for (int q = 0; q < 128; ++q) {
                    if (num[col][q] == 0) continue;
                    /* row_r -= f * row_col */
                    I n2 = fn * num[col][q], d2 = fd * den[col][q];
                    I nn = num[r][q] * d2 - n2 * den[r][q], dd = den[r][q] * d2;
                    norm(nn, dd);
                    num[r][q] = nn; den[r][q] = dd;
                }
            }
        }
        for (int r = 0; r < 64; ++r)
            for (int q = 0; q < 64; ++q) A[r][q] = (int)(num[r][64 + q] / den[r][64 + q]);
    }
};

const TricubicMatrix &tricubic_matrix() {
    static TricubicMatrix m;
    return m;
}

} // namespace

extern "C" int pt_tricubic_precompute(int nx, int ny, int nz, const double *xg, const double *yg, const double *zg,
                                      int ncomp, const double *values, double *coef, int threads) {
    if (nx < 2 || ny < 2 || nz < 2 || ncomp < 1 || !xg || !yg || !zg || !values || !coef)
        return fail(PT_EINVAL, "pt_tricubic_precompute: bad arguments");
    std::vector<double> xyz[3] = {std::vector<double>(xg, xg + nx), std::vector<double>(yg, yg + ny), std::vector<double>(zg, zg + nz)};
    for (int a = 0; a < 3; ++a)
        for (size_t i = 1; i < xyz[a].size(); ++i)
            if (!(xyz[a][i] > xyz[a][i - 1])) return fail(PT_EINVAL, "pt_tricubic_precompute: grid not strictly increasing");
    const TricubicMatrix &M = tricubic_matrix();
    size_t np = (size_t)nx * ny * nz;
    size_t ncell = (size_t)(nx - 1) * (ny - 1) * (nz - 1);
    if (threads < 1) threads = 1;
    for (int comp = 0; comp < ncomp; ++comp) {
        const double *tab = values + comp * np;
        std::vector<double> dx(np), dy(np), dz(np), dxy(np), dxz(np), dyz(np), dxyz(np);
        calc_derivs(xyz, nx, ny, nz, tab, 0, dx.data());
        calc_derivs(xyz, nx, ny, nz, tab, 1, dy.data());
        calc_derivs(xyz, nx, ny, nz, tab, 2, dz.data());
        calc_derivs(xyz, nx, ny, nz, dx.data(), 1, dxy.data());
        calc_derivs(xyz, nx, ny, nz, dx.data(), 2, dxz.data());
        calc_derivs(xyz, nx, ny, nz, dy.data(), 2, dyz.data());
        calc_derivs(xyz, nx, ny, nz, dxy.data(), 2, dxyz.data());
        auto work = [&](int t) {
            for (size_t cell = (size_t)t; cell < ncell; cell += (size_t)threads) {
                int iz = (int)(cell % (nz - 1));
                int iy = (int)((cell / (nz - 1)) % (ny - 1));
                int ix = (int)(cell / ((size_t)(nz - 1) * (ny - 1)));
                double cellx = xyz[0][ix + 1] - xyz[0][ix];
                double celly = xyz[1][iy + 1] - xyz[1][iy];
                double cellz = xyz[2][iz + 1] - xyz[2][iz];
                double xv[64];
                for (int c = 0; c < 8; ++c) {
                    size_t p = ((size_t)(ix + (c & 1)) * ny + (iy + ((c >> 1) & 1))) * nz + (iz + ((c >> 2) & 1));
                    xv[c] = tab[p];
                    xv[8 + c] = dx[p] * cellx;
                    xv[16 + c] = dy[p] * celly;
                    xv[24 + c] = dz[p] * cellz;
                    xv[32 + c] = dxy[p] * cellx * celly;
                    xv[40 + c] = dxz[p] * cellx * cellz;
                    xv[48 + c] = dyz[p] * celly * cellz;
                    xv[56 + c] = dxyz[p] * cellx * celly * cellz;
                }
                double *a = coef + (cell * ncomp + comp) * 64;
                /* tricubic_get_coeff_stacked: a[i] = sum_j A[i][j] x[j] in order */
                for (int i = 0; i < 64; ++i) {
                    double s = 0.0;
                    for (int j = 0; j < 64; ++j) s += (double)M.A[i][j] * xv[j];
                    a[i] = s;
                }
            }
        };
        if (threads == 1) work(0);
        else {
            std::vector<std::thread> th;
            for (int t = 0; t < threads; ++t) th.emplace_back(work, t);
            for (auto &t : th) t.join();
        }
    }
    return PT_OK;
}

namespace {

/* TabField3::PreInterpol on the device for samples values_h [ncomp][nx][ny][nz] (host) into
   coef_dev [cells][ncomp][64]; synchronous */
int tricubic_on_device(int nx, int ny, int nz, const double *const gh[3], int ncomp, const double *values_h,
                       int device_order, double *coef_dev) {
    DevMem tmp;
    const size_t np = (size_t)nx * ny * nz;
    const int n[3] = {nx, ny, nz};
    double *g[3], *der[8], *scr, *A;
    double **der_dev;
    int r;
    for (int a = 0; a < 3; ++a)
        if ((r = tmp.upload(&g[a], gh[a], (size_t)n[a]))) return r;
    for (int k = 0; k < 8; ++k)
        if ((r = tmp.alloc(&der[k], np))) return r;
    if ((r = tmp.alloc(&scr, 2 * np)) || (r = tmp.upload(&der_dev, der, 8))) return r;
    {
        const TricubicMatrix &M = tricubic_matrix();
        std::vector<double> Ad(64 * 64);
        for (int i = 0; i < 64; ++i)
            for (int j = 0; j < 64; ++j) Ad[64 * i + j] = (double)M.A[i][j];
        if ((r = tmp.upload(&A, Ad.data(), Ad.size()))) return r;
    }
    for (int comp = 0; comp < ncomp; ++comp) {
        PT_HIP(hipMemcpy(der[0], values_h + comp * np, np * sizeof(double), hipMemcpyHostToDevice));
        PT_HIP(pt_launch_tricubic_derivs(g, nx, ny, nz, der, scr, 0));
        PT_HIP(pt_launch_tricubic_coef(g[0], g[1], g[2], nx, ny, nz, der_dev, A, ncomp, comp, device_order, coef_dev,
                                       0));
        PT_HIP(hipStreamSynchronize(0));
    }
    return PT_OK;
}

int check_grid(const double *const g[3], const int n[3], const char *who) {
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 2 || !g[a]) return fail(PT_EINVAL, std::string(who) + ": bad grid");
        for (int i = 1; i < n[a]; ++i)
            if (!(g[a][i] > g[a][i - 1])) return fail(PT_EINVAL, std::string(who) + ": grid not strictly increasing");
    }
    return PT_OK;
}

} // namespace

/* TabField3::PreInterpol on the GPU, host arrays in and out (libtricubic coefficient order,
   as pt_tricubic_precompute); synchronous on the null stream */
extern "C" int pt_tricubic_precompute_gpu(int nx, int ny, int nz, const double *xg, const double *yg, const double *zg,
                                          int ncomp, const double *values, double *coef) {
    const double *gh[3] = {xg, yg, zg};
    const int n[3] = {nx, ny, nz};
    if (ncomp < 1 || !values || !coef) return fail(PT_EINVAL, "pt_tricubic_precompute_gpu: bad arguments");
    int r = check_grid(gh, n, "pt_tricubic_precompute_gpu");
    if (r) return r;
    const size_t ncoef = (size_t)(nx - 1) * (ny - 1) * (nz - 1) * ncomp * 64;
    DevMem mem;
    double *dc;
    if ((r = mem.alloc(&dc, ncoef))) return r;
    if ((r = tricubic_on_device(nx, ny, nz, gh, ncomp, values, 0, dc))) return r;
    PT_HIP(hipMemcpy(coef, dc, ncoef * sizeof(double), hipMemcpyDeviceToHost));
    return PT_OK;
}

/* exported for tests: the spline derivatives of one grid line (ALGLIB spline1dgriddiffcubic) */
extern "C" int pt_spline_grid_diff(int n, const double *x, const double *y, double *d) {
    if (n < 2 || !x || !y || !d) return fail(PT_EINVAL, "pt_spline_grid_diff: bad arguments");
    std::vector<double> buf;
    spline_grid_diff(x, y, n, d, buf);
    return PT_OK;
}

/* ======================================================================== */
/* bicubic tables of axisymmetric maps (TabField, src/field_2d.cpp:235-285)  */
/* ======================================================================== */

namespace {

/* Derivatives at the nodes of the cubic spline through (x, y): ALGLIB spline1dbuildcubic
   with parabolic ends (node derivatives d as spline1dgriddiffcubic), its Hermite pieces
   (spline1dbuildhermite, interpolation.cpp:23642-23700) and spline1ddiff at each node
   (:23914-23980).  The node x[j], j > 0, lies at the right end of piece j-1, so the
   value is that piece's derivative at its length, not d[j] itself. */
void spline_node_diff(const double *x, const double *y, int n, double *ds, std::vector<double> &d,
                      std::vector<double> &buf) {
    d.resize(n);
    spline_grid_diff(x, y, n, d.data(), buf);
    for (int j = 0; j < n; ++j) {
        int l = j > 0 ? j - 1 : 0;
        double delta = x[l + 1] - x[l];
        double delta2 = delta * delta;
        double delta3 = delta * delta2;
        double c1 = d[l];
        double c2 = (3 * (y[l + 1] - y[l]) - 2 * d[l] * delta - d[l + 1] * delta) / delta2;
        double c3 = (2 * (y[l] - y[l + 1]) + d[l] * delta + d[l + 1] * delta) / delta3;
        double s = x[j] - x[l];
        ds[j] = c1 + 2 * s * c2 + 3 * (s * s) * c3;
    }
}

} // namespace

/*
 * ALGLIB spline2dbuildbicubicv (interpolation.cpp:44925-45075) and its derivative tables
 * (spline2d_bicubiccalcderivatives, :46791-46886): df/dr along every z row, df/dz along
 * every r column, d2f/drdz along every z row of df/dz.  Output per cell (iz, ir) and
 * component: the four tables (f, df/dr, df/dz, d2f/drdz) at the corners (iz,ir),
 * (iz,ir+1), (iz+1,ir), (iz+1,ir+1) -- the 16 numbers spline2dcalc/spline2ddiff read.
 */
extern "C" int pt_bicubic_precompute(int m, int n, const double *rg, const double *zg, int ncomp,
                                     const double *values, double *coef) {
    if (m < 2 || n < 2 || ncomp < 1 || !rg || !zg || !values || !coef)
        return fail(PT_EINVAL, "pt_bicubic_precompute: bad arguments");
    for (int i = 1; i < m; ++i)
        if (!(rg[i] > rg[i - 1])) return fail(PT_EINVAL, "pt_bicubic_precompute: r grid not strictly increasing");
    for (int i = 1; i < n; ++i)
        if (!(zg[i] > zg[i - 1])) return fail(PT_EINVAL, "pt_bicubic_precompute: z grid not strictly increasing");
    const size_t np = (size_t)m * n;
    std::vector<double> dr(np), dz(np), drz(np), line(std::max(m, n)), out(std::max(m, n)), d, buf;
    for (int comp = 0; comp < ncomp; ++comp) {
        const double *a = values + comp * np;        /* a[iz * m + ir] */
        for (int iz = 0; iz < n; ++iz) {
            spline_node_diff(rg, a + (size_t)iz * m, m, out.data(), d, buf);
            for (int ir = 0; ir < m; ++ir) dr[(size_t)iz * m + ir] = out[ir];
        }
        for (int ir = 0; ir < m; ++ir) {
            for (int iz = 0; iz < n; ++iz) line[iz] = a[(size_t)iz * m + ir];
            spline_node_diff(zg, line.data(), n, out.data(), d, buf);
            for (int iz = 0; iz < n; ++iz) dz[(size_t)iz * m + ir] = out[iz];
        }
        for (int iz = 0; iz < n; ++iz) {
            spline_node_diff(rg, dz.data() + (size_t)iz * m, m, out.data(), d, buf);
            for (int ir = 0; ir < m; ++ir) drz[(size_t)iz * m + ir] = out[ir];
        }
        const double *tab[4] = {a, dr.data(), dz.data(), drz.data()};
        for (int iz = 0; iz < n - 1; ++iz)
            for (int ir = 0; ir < m - 1; ++ir) {
                double *c = coef + (((size_t)iz * (m - 1) + ir) * ncomp + comp) * 16;
                const size_t s1 = (size_t)iz * m + ir, s2 = s1 + 1, s3 = s1 + m, s4 = s3 + 1;
                for (int k = 0; k < 4; ++k) {
                    c[4 * k + 0] = tab[k][s1];
                    c[4 * k + 1] = tab[k][s2];
                    c[4 * k + 2] = tab[k][s3];
                    c[4 * k + 3] = tab[k][s4];
                }
            }
    }
    return PT_OK;
}

/* field-scaling formula f(t) evaluated on the host by the program the device runs */
extern "C" int pt_formula_eval(const char *formula, int n, const double *t, double *out) {
    if (!formula || n < 0 || (n > 0 && (!t || !out))) return fail(PT_EINVAL, "pt_formula_eval: bad arguments");
    std::vector<uint32_t> code;
    std::vector<double> cst;
    std::string err;
    if (!ptf::compile(formula, code, cst, err)) return fail(PT_EINVAL, "pt_formula_eval: " + err);
    for (int i = 0; i < n; ++i) out[i] = ptf_run(code.data(), (int)code.size(), cst.data(), t[i]);
    return PT_OK;
}

/* ======================================================================== */
/* neutron decay products (TNeutron::Decay, src/neutron.cpp:285-362)          */
/* ======================================================================== */

namespace {

/* ProtonBetaSpectrum (src/mc.cpp:84-92) as the reference evaluates it on x86-64: long
   double constants (src/globals.cpp:13-21), double intermediates, pow(x, 2) as x*x
   (what the compilers make of it) and the integer quotient 4/3 == 1 */
double proton_beta_spectrum(double E) {
    const long double ele_e = 1.602176487E-19L, c_0 = 299792458.L;
    const long double m_n = 1.674927211E-27L / ele_e, m_p = 1.672621637E-27L / ele_e, m_e = 9.10938215e-31L / ele_e;
    double DeltaM = m_n - m_p;
    double Xi = m_e / DeltaM;
    double Sigma = 1 - 2 * E * m_n / (DeltaM * DeltaM) / (c_0 * c_0);
    double q = (Sigma - Xi * Xi) / Sigma;
    double g1 = (q * q) * std::sqrt(1 - Sigma) * (4 * (1 + Xi * Xi / Sigma) - 4 / 3 * ((Sigma - Xi * Xi) / Sigma * (1 - Sigma)));
    double g2 = (q * q) * std::sqrt(1 - Sigma) * (4 * (1 + Xi * Xi / Sigma - 2 * Sigma) - 4 / 3 * ((Sigma - Xi * Xi) / Sigma * (1 - Sigma)));
    double littlea = -0.1017;
    return 0.5 * (g1 + littlea * g2);
}

/* adaptive_linear_interpolation (include/mc.h:244-270): bisect until the trapezoid
   integral changes by less than the tolerance and the step is small enough */
template <class F>
void adaptive_interpolation(const F &f, double next_abscissa, double next_value, std::vector<double> &abscissas,
                            std::vector<double> &values, double tolerance, double max_diff) {
    double h = next_abscissa - abscissas.back();
    double I0 = h * (values.back() + next_value) / 2.;
    double ia = (abscissas.back() + next_abscissa) / 2.;
    double iv = f(ia);
    double I1 = (ia - abscissas.back()) * (values.back() + iv) / 2.;
    double I2 = (next_abscissa - ia) * (iv + next_value) / 2.;
    if ((h > max_diff) || ((std::fabs(I1 + I2 - I0) > tolerance) && (h > max_diff / 100.))) {
        adaptive_interpolation(f, ia, iv, abscissas, values, tolerance / 2, max_diff);
        adaptive_interpolation(f, next_abscissa, next_value, abscissas, values, tolerance / 2, max_diff);
    } else {
        abscissas.push_back(ia);
        values.push_back(iv);
        abscissas.push_back(next_abscissa);
        values.push_back(next_value);
    }
}

struct BetaTable { std::vector<double> x, raw, den, cp, m; };

/* interpolate_distribution(f, lo, hi) (include/mc.h:281-296) and the parameter
   std::piecewise_linear_distribution<double> builds from it (libstdc++
   param_type::_M_initialize: trapezoid partial sums, slopes, all divided by the total,
   last cumulative probability set to 1) */
template <class F>
int distribution_table(const F &f, double lo, double hi, BetaTable &t) {
    if (lo > hi) return fail(PT_EINVAL, "Invalid range used to parse distribution");
    t.x.assign(1, lo);
    t.den.assign(1, f(lo));
    if (lo == hi) {
        t.x.push_back(std::nextafter(hi, std::numeric_limits<double>::max()));
        t.den.push_back(f(t.x.back()));
    } else {
        adaptive_interpolation(f, hi, f(hi), t.x, t.den, std::numeric_limits<double>::epsilon() * 1000., (hi - lo) / 10.);
    }
    t.raw = t.den;
    double sum = 0.0;
    t.cp.clear();
    t.m.clear();
    for (size_t k = 0; k + 1 < t.x.size(); ++k) {
        const double delta = t.x[k + 1] - t.x[k];
        sum += 0.5 * (t.den[k + 1] + t.den[k]) * delta;
        t.cp.push_back(sum);
        t.m.push_back((t.den[k + 1] - t.den[k]) / delta);
    }
    if (!(sum > 0)) return fail(PT_EINVAL, "distribution with zero or negative integral");
    for (auto &v : t.den) v = v / sum;
    for (auto &v : t.cp) v = v / sum;
    for (auto &v : t.m) v = v / sum;
    t.cp.back() = 1.0;
    return PT_OK;
}

/* interpolate_distribution(ProtonBetaSpectrum, 0., 750.) (include/mc.h:337) */
const BetaTable &proton_beta_table() {
    static const BetaTable T = [] {
        BetaTable t;
        distribution_table(proton_beta_spectrum, 0., 750., t);
        return t;
    }();
    return T;
}

pt_beta_table host_view(const BetaTable &T) {
    return pt_beta_table{T.x.data(), T.den.data(), T.cp.data(), T.m.data(), (int)T.cp.size()};
}

}  // namespace

double ptf_proton_beta_spectrum(double E) { return proton_beta_spectrum(E); }

/* the sampled proton spectrum (for tests): n_int intervals */
extern "C" int pt_proton_beta_table(int cap, double *x, double *raw, double *den, double *cp, double *m, int *n_int) {
    const BetaTable &T = proton_beta_table();
    if (!n_int) return fail(PT_EINVAL, "pt_proton_beta_table: bad arguments");
    *n_int = (int)T.cp.size();
    if (!x) return PT_OK;
    if (cap < (int)T.x.size() || !raw || !den || !cp || !m) return fail(PT_EINVAL, "pt_proton_beta_table: buffer too small");
    std::copy(T.x.begin(), T.x.end(), x);
    std::copy(T.raw.begin(), T.raw.end(), raw);
    std::copy(T.den.begin(), T.den.end(), den);
    std::copy(T.cp.begin(), T.cp.end(), cp);
    std::copy(T.m.begin(), T.m.end(), m);
    return PT_OK;
}

/* parse_distribution (include/mc.h:310-335): the piecewise linear distribution of a
   formula of x on [lo, hi] (the source's spectra and angular distributions) */
extern "C" int pt_distribution_table(const char *formula, double lo, double hi, int cap, double *x, double *raw,
                                     double *den, double *cp, double *m, int *n_int) {
    if (!formula || !n_int) return fail(PT_EINVAL, "pt_distribution_table: bad arguments");
    std::vector<uint32_t> code;
    std::vector<double> cst;
    std::string err;
    if (!ptf::compile_x(formula, code, cst, err)) return fail(PT_EINVAL, "pt_distribution_table: " + err);
    auto f = [&](double v) { return ptf_run(code.data(), (int)code.size(), cst.data(), v); };
    BetaTable t;
    int r = distribution_table(f, lo, hi, t);
    if (r) return r;
    *n_int = (int)t.cp.size();
    if (!x) return PT_OK;
    if (cap < (int)t.x.size() || !raw || !den || !cp || !m) return fail(PT_EINVAL, "pt_distribution_table: buffer too small");
    std::copy(t.x.begin(), t.x.end(), x);
    std::copy(t.raw.begin(), t.raw.end(), raw);
    std::copy(t.den.begin(), t.den.end(), den);
    std::copy(t.cp.begin(), t.cp.end(), cp);
    std::copy(t.m.begin(), t.m.end(), m);
    return PT_OK;
}

/* the proton energies the device samples for variates p[i] in [0, 1) (for tests) */
extern "C" int pt_beta_sample(int n, const double *p, double *out) {
    if (n < 0 || (n > 0 && (!p || !out))) return fail(PT_EINVAL, "pt_beta_sample: bad arguments");
    const pt_beta_table T = host_view(proton_beta_table());
    for (int i = 0; i < n; ++i) out[i] = ptk_beta_sample(&T, p[i]);
    return PT_OK;
}

/* the decay kinematics the device runs, on the host (for tests): stream key and
   position, neutron velocity -> Ep phip thetap Ee phie thetae, polarisations, new position */
extern "C" int pt_neutron_decay_host(uint64_t key, uint64_t ctr, const double *v, double *out6, int32_t *pol2,
                                     uint64_t *ctr_out) {
    if (!v || !out6 || !pol2) return fail(PT_EINVAL, "pt_neutron_decay_host: bad arguments");
    const pt_beta_table T = host_view(proton_beta_table());
    pt_decay_out o;
    ptk_neutron_decay(&T, key, &ctr, v, &o);
    out6[0] = o.Ep; out6[1] = o.phip; out6[2] = o.thetap; out6[3] = o.Ee; out6[4] = o.phie; out6[5] = o.thetae;
    pol2[0] = o.polp; pol2[1] = o.pole;
    if (ctr_out) *ctr_out = ctr;
    return PT_OK;
}

/* exported for tests: the 64x64 integer matrix */
extern "C" int pt_seg_tri(const double *a, const double *b, const double *c, const double *p, const double *q, double *pt) {
    return ptd::seg_tri(a, b, c, p, q, pt);
}

extern "C" int pt_tricubic_matrix(int32_t *out) {
    const TricubicMatrix &M = tricubic_matrix();
    for (int i = 0; i < 64; ++i)
        for (int j = 0; j < 64; ++j) out[64 * i + j] = M.A[i][j];
    return PT_OK;
}

/* ======================================================================== */
/* particle batches                                                          */
/* ======================================================================== */

struct pt_batch {
    DevMem mem;
    /* captured launch sequences of pt_batch_run, keyed by the launch configuration */
    struct GraphEntry { std::vector<int> key; hipGraph_t graph; hipGraphExec_t exec; };
    std::vector<GraphEntry> graphs;
    /* optional per-kernel timing: event pairs around step / collision launches */
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[2];
    double ms[2] = {0, 0};
    long long launches[2] = {0, 0};
    ~pt_batch() {
        for (auto &g : graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
        for (auto &v : ev)
            for (auto &e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    }
    ptd::Params P{};
    int n = 0;
    int hit_cap = 0;
    int *d_quit = nullptr;
    int *d_count = nullptr;
    ptd::Params *d_params = nullptr;   /* device copy of P, uploaded before each launch sequence */
    const pt_geometry *geom = nullptr;
    const pt_fields *fields = nullptr;
    ptd::Species *d_species = nullptr;
    std::vector<ptd::Species> h_species;
    /* decay products: the proton spectrum on the device and the record buffer (lazily) */
    pt_beta_table d_beta{};
    pt_secondary *d_sec = nullptr;
    int *d_sec_count = nullptr;
};

extern "C" int pt_batch_create(const pt_geometry *g, const pt_fields *f, int n_species, const pt_species *species,
                               double tmax_global, int n, const int32_t *species_idx, const double *t0,
                               const double *y0, const double *spin0, uint64_t seed, const uint64_t *pid,
                               int hit_cap, pt_batch **out) {
    if (!g || !f || !out || n < 0 || n_species <= 0 || !species || hit_cap < 0) return fail(PT_EINVAL, "pt_batch_create: bad arguments");
    if (n_species > ptd::MAX_SPECIES)
        return fail(PT_EINVAL, "pt_batch_create: at most " + std::to_string(ptd::MAX_SPECIES) + " species");
    if (n > 0 && (!species_idx || !t0 || !y0 || !spin0 || !pid)) return fail(PT_EINVAL, "pt_batch_create: null particle arrays");
    for (int i = 0; i < n; ++i)
        if (species_idx[i] < 0 || species_idx[i] >= n_species) return fail(PT_EINVAL, "pt_batch_create: species index out of range");
    for (int s = 0; s < n_species; ++s)
        if (species[s].kind < 0 || species[s].kind > 2) return fail(PT_EINVAL, "pt_batch_create: unknown species kind");
    std::unique_ptr<pt_batch> B(new pt_batch());
    B->n = n;
    B->hit_cap = hit_cap;
    B->geom = g;
    B->fields = f;
    std::vector<ptd::Species> sp(n_species);
    for (int s = 0; s < n_species; ++s) {
        const pt_species &q = species[s];
        sp[s].kind = q.kind; sp[s].q = q.q; sp[s].m = q.m; sp[s].mu = q.mu; sp[s].gamma = q.gamma;
        sp[s].tau = q.tau; sp[s].tmax = q.tmax; sp[s].lmax = q.lmax; sp[s].Bmax = q.Bmax; sp[s].flipspin = q.flipspin;
        if (q.n_spintimes < 0 || q.n_spintimes > 8 || (q.n_spintimes & 1))
            return fail(PT_EINVAL, "pt_batch_create: spintimes needs pairs of times, at most four windows");
        sp[s].n_spintimes = q.n_spintimes;
        sp[s].interpolatefields = q.interpolatefields;
        for (int k = 0; k < 8; ++k) sp[s].spintimes[k] = k < q.n_spintimes ? q.spintimes[k] : 0.0;
    }
    int r;
    if ((r = B->mem.upload(&B->d_species, sp.data(), sp.size()))) return r;
    B->h_species = sp;
    ptd::State &S = B->P.st;
    S.n = n;
    size_t N = (size_t)std::max(n, 1);
    std::vector<double> y(9 * N), sp5(5 * N), ones(N, 1.0), zeros(9 * N, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 9; ++k) y[(size_t)k * n + i] = y0[9 * (size_t)i + k];
        for (int k = 0; k < 5; ++k) sp5[(size_t)k * n + i] = spin0[5 * (size_t)i + k];
    }
    std::vector<unsigned long long> key(N), zero64(N, 0ull);
    for (int i = 0; i < n; ++i) key[i] = ptm_stream_key(seed, pid[i]);
    std::vector<int> izero(N, 0), nsol(N, 1), spidx(species_idx, species_idx + n);
    std::vector<long long> lzero(N, 0);
    if ((r = B->mem.upload(&S.t, t0, (size_t)n))) return r;
    if ((r = B->mem.upload(&S.y, y.data(), 9 * (size_t)n))) return r;
    if ((r = B->mem.upload(&S.xc, y.data(), 9 * (size_t)n))) return r;
    if ((r = B->mem.upload(&S.dxc, zeros.data(), 9 * (size_t)n))) return r;
    if ((r = B->mem.upload(&S.ts, t0, (size_t)n))) return r;
    if ((r = B->mem.upload(&S.dt, zeros.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.spin, sp5.data(), 5 * (size_t)n))) return r;
    if ((r = B->mem.upload(&S.Hmax, zeros.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.noflip, ones.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.tau, zeros.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.stop, izero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.flags, izero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.Nhit, izero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.Nstep, izero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.Nspinflip, izero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.nsolid, nsol.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.nhitlog, izero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.species, spidx.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.n_acc, lzero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.n_rej, lzero.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.ctr, zero64.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.key, key.data(), (size_t)n))) return r;
    if ((r = B->mem.upload(&S.solids, zero64.data(), (size_t)n))) return r;
    if ((r = B->mem.alloc(&S.hits, (size_t)n * std::max(hit_cap, 0)))) return r;
    S.hit_cap = hit_cap;
    if ((r = B->mem.alloc(&S.Bc, 3 * N))) return r;
    if ((r = B->mem.alloc(&S.Vc, N))) return r;
    if ((r = B->mem.alloc(&S.pend, 20 * N))) return r;
    if ((r = B->mem.alloc(&S.dense, (size_t)ptd::DENSE * N))) return r;
    if ((r = B->mem.upload(&S.n_field, lzero.data(), N))) return r;
    if ((r = B->mem.upload(&S.budget, izero.data(), N))) return r;
    if ((r = B->mem.alloc(&S.queue, N))) return r;
    if ((r = B->mem.upload(&S.qcount, izero.data(), 1))) return r;
    if ((r = B->mem.alloc(&S.work, N))) return r;
    const std::vector<int> wzero(ptd::WCOUNT_LEN, 0);
    if ((r = B->mem.upload(&S.wcount, wzero.data(), wzero.size()))) return r;
    if ((r = B->mem.alloc(&S.pkey, N))) return r;
    if ((r = B->mem.alloc(&S.bhist, (size_t)ptd::NBUCKET + 1))) return r;
    if ((r = B->mem.alloc(&S.qkey, N))) return r;
    if ((r = B->mem.alloc(&S.qsorted, N))) return r;
    if ((r = B->mem.alloc(&S.qhist, (size_t)ptd::QKEYS + 1))) return r;
    int zero = 0;
    if ((r = B->mem.upload(&B->d_quit, &zero, 1))) return r;
    if ((r = B->mem.upload(&B->d_count, &zero, 1))) return r;
    if ((r = B->mem.alloc(&B->d_params, 1))) return r;
    ptd::Params &P = B->P;
    P.self = B->d_params;
    P.geom = g->g;
    P.n_fields = f->n;
    P.b_reuse = f->b_reuse ? 1 : 0;
    for (int k = 0; k < f->n; ++k) P.fields[k] = f->host[k];
    for (int k = 0; k < n_species; ++k) P.species[k] = B->h_species[k];
    P.tmax_global = tmax_global;
    P.quit = B->d_quit;
    *out = B.release();
    return PT_OK;
}

extern "C" void pt_batch_destroy(pt_batch *b) { delete b; }

extern "C" int pt_batch_run(pt_batch *b, void *stream, const pt_launch_cfg *cfg) {
    if (!b || !cfg) return fail(PT_EINVAL, "pt_batch_run: bad arguments");
    if (cfg->variant != PT_VARIANT_FP64 && cfg->variant != PT_VARIANT_MIXED) return fail(PT_EINVAL, "pt_batch_run: unknown variant");
    if (cfg->variant == PT_VARIANT_MIXED && !b->fields->has_fp32) {
        bool any_map = false;
        for (int k = 0; k < b->fields->n; ++k) any_map |= b->fields->host[k].type == PT_FIELD_TRICUBIC;
        if (any_map) return fail(PT_EINVAL, "pt_batch_run: mixed precision needs fields created with FP32 tables");
    }
    if (cfg->spatial_order < 0 || cfg->spatial_order > 2) return fail(PT_EINVAL, "pt_batch_run: spatial_order must be 0, 1 or 2");
    if (cfg->collision_order < 0 || cfg->collision_order > 1) return fail(PT_EINVAL, "pt_batch_run: collision_order must be 0 or 1");
    int block = cfg->block > 0 ? cfg->block : 256;
    if (block % 64 != 0 || block > 256) return fail(PT_EINVAL, "pt_batch_run: block must be a multiple of 64 and <= 256");
    {
        /* the step kernels hold their dopri5 stage vectors in PT_STEP_LDS_BYTES of LDS per block */
        int dev = 0, lds = 0;
        PT_HIP(hipGetDevice(&dev));
        PT_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (lds < PT_STEP_LDS_BYTES)
            return fail(PT_EDEVICE, "pt_batch_run: the step kernels need " + std::to_string(PT_STEP_LDS_BYTES) +
                                        " bytes of LDS per block, the device allows " + std::to_string(lds));
    }
    /* mixed-precision tables that keep every block in FP64 give exactly the FP64 kernels'
       arithmetic (a kept block is evaluated as the FP64 kernels evaluate it): such a batch
       runs the FP64 kernels, with their cooperative gathers */
    const bool mixed = cfg->variant == PT_VARIANT_MIXED && b->fields->mix_fp64 < b->fields->mix_blocks;
    const int mode = (mixed ? 1 : 0) | (b->fields->general ? 2 : 0);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int q = g_quit.load();
    PT_HIP(hipMemcpyAsync(b->d_quit, &q, sizeof(int), hipMemcpyHostToDevice, s));
    /* the kernels read the parameters from their device copy (also under graph replay:
       the captured launches take its address, this copy stays outside the graph) */
    PT_HIP(hipMemcpyAsync(b->d_params, &b->P, sizeof(ptd::Params), hipMemcpyHostToDevice, s));
    int steps = cfg->steps_per_launch > 0 ? cfg->steps_per_launch : INT_MAX;
    int rounds = cfg->rounds > 0 ? cfg->rounds : 4;
    int coll_grid = std::max(1, std::min((b->n + 255) / 256, 2048));
    int pgrid = cfg->grid;
    if (cfg->persistent && pgrid <= 0) {
        int dev = 0, cus = 0;
        PT_HIP(hipGetDevice(&dev));
        PT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        pgrid = 2 * std::max(cus, 1);   /* two 256-thread blocks per CU (2 waves per SIMD) */
    }
    auto launch_all = [&](bool timed) -> int {
        PT_HIP(pt_launch_init(b->P, mode, steps, s));
        for (int r = 0; r < rounds; ++r) {
            for (int k = 0; k < 2; ++k) {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (timed) {
                    PT_HIP(hipEventCreate(&e0));
                    PT_HIP(hipEventCreate(&e1));
                    PT_HIP(hipEventRecord(e0, s));
                }
                if (k == 1) PT_HIP(pt_launch_collision(b->P, mode, coll_grid, cfg->collision_order, s));
                else if (cfg->persistent) PT_HIP(pt_launch_step_persistent(b->P, mode, pgrid, cfg->spatial_order, s));
                else PT_HIP(pt_launch_step(b->P, mode, block, s));
                if (timed) {
                    PT_HIP(hipEventRecord(e1, s));
                    b->ev[k].emplace_back(e0, e1);
                }
            }
        }
        return PT_OK;
    };
    if (!cfg->graph) return launch_all(b->profiling);
    if (s == nullptr) return fail(PT_EINVAL, "pt_batch_run: graph mode needs a non-default stream");
    std::vector<int> key = {cfg->variant, block, steps, rounds, cfg->persistent, pgrid, cfg->spatial_order,
                            cfg->collision_order};
    pt_batch::GraphEntry *ge = nullptr;
    for (auto &g : b->graphs)
        if (g.key == key) ge = &g;
    if (!ge) {
        pt_batch::GraphEntry g{key, nullptr, nullptr};
        PT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int r = launch_all(false);
        hipError_t ec = hipStreamEndCapture(s, &g.graph);
        if (r) { if (g.graph) (void)hipGraphDestroy(g.graph); return r; }
        if (ec != hipSuccess) return fail(PT_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        PT_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
        b->graphs.push_back(g);
        ge = &b->graphs.back();
    }
    PT_HIP(hipGraphLaunch(ge->exec, s));
    return PT_OK;
}

namespace {
__global__ void count_active_kernel(const int *stop, const int *flags, int n, int *count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < n && (stop[i] == 0 || !(flags[i] & 1));
    unsigned long long ballot = __ballot(active);
    if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(count, __popcll(ballot));
}
} // namespace

extern "C" int pt_batch_active(pt_batch *b, void *stream, int *active) {
    if (!b || !active) return fail(PT_EINVAL, "pt_batch_active: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int zero = 0;
    PT_HIP(hipMemcpyAsync(b->d_count, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    if (b->n > 0) {
        int grid = (b->n + 255) / 256;
        hipLaunchKernelGGL(count_active_kernel, dim3(grid), dim3(256), 0, s, b->P.st.stop, b->P.st.flags, b->n, b->d_count);
        PT_HIP(hipGetLastError());
    }
    PT_HIP(hipMemcpyAsync(active, b->d_count, sizeof(int), hipMemcpyDeviceToHost, s));
    PT_HIP(hipStreamSynchronize(s));
    return PT_OK;
}

/* TParticle::DoDecay -> TNeutron::Decay (src/particle.cpp:219-223, src/neutron.cpp:285-362) for
   every neutron of the batch that stopped with ID_DECAYED: the proton and the electron,
   ordered by parent index (proton first).  Idempotent: the products are recomputed from
   the final state and the neutron's stream each call. */
extern "C" int pt_batch_secondaries(pt_batch *b, void *stream, int cap, pt_secondary *out, int *count) {
    if (!b || !count || cap < 0 || (cap > 0 && !out)) return fail(PT_EINVAL, "pt_batch_secondaries: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    *count = 0;
    if (b->n == 0) return PT_OK;
    int r;
    if (!b->d_sec) {
        const BetaTable &T = proton_beta_table();
        double *x, *den, *cp, *m;
        if ((r = b->mem.upload(&x, T.x.data(), T.x.size()))) return r;
        if ((r = b->mem.upload(&den, T.den.data(), T.den.size()))) return r;
        if ((r = b->mem.upload(&cp, T.cp.data(), T.cp.size()))) return r;
        if ((r = b->mem.upload(&m, T.m.data(), T.m.size()))) return r;
        b->d_beta = pt_beta_table{x, den, cp, m, (int)T.cp.size()};
        if ((r = b->mem.alloc(&b->d_sec, 2 * (size_t)b->n))) return r;
        if ((r = b->mem.alloc(&b->d_sec_count, 1))) return r;
    }
    PT_HIP(hipMemcpyAsync(b->d_params, &b->P, sizeof(ptd::Params), hipMemcpyHostToDevice, s));
    PT_HIP(pt_launch_decay(b->P, b->d_beta, b->d_sec, b->d_sec_count, s));
    int n2 = 0;
    PT_HIP(hipMemcpyAsync(&n2, b->d_sec_count, sizeof(int), hipMemcpyDeviceToHost, s));
    PT_HIP(hipStreamSynchronize(s));
    std::vector<pt_secondary> rec(n2);
    if (n2) PT_HIP(hipMemcpy(rec.data(), b->d_sec, n2 * sizeof(pt_secondary), hipMemcpyDeviceToHost));
    std::stable_sort(rec.begin(), rec.end(), [](const pt_secondary &a, const pt_secondary &c) {
        return a.parent != c.parent ? a.parent < c.parent : a.kind < c.kind;
    });
    *count = n2;
    if (out) std::copy(rec.begin(), rec.begin() + std::min(n2, cap), out);
    if (out && cap < n2) return fail(PT_EINVAL, "pt_batch_secondaries: capacity below the number of products");
    return PT_OK;
}

/* TFieldManager::BField / EField at n points and times (host arrays), FP64 tables */
namespace {
int fields_eval(const pt_fields *f, void *stream, int n, const double *pts, const double *t, int mixed, int coop,
                double *B, double *dB, double *V, double *E) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevMem mem;
    double *dp, *dt, *dB_, *ddB, *dV, *dE;
    int r;
    if ((r = mem.upload(&dp, pts, 3 * (size_t)n)) || (r = mem.upload(&dt, t, (size_t)n)) || (r = mem.alloc(&dB_, 3 * (size_t)n)) ||
        (dB && (r = mem.alloc(&ddB, 9 * (size_t)n))) || (r = mem.alloc(&dV, (size_t)n)) || (r = mem.alloc(&dE, 3 * (size_t)n)))
        return r;
    if (!dB) ddB = nullptr;
    ptd::Params P{};
    P.n_fields = f->n;
    for (int k = 0; k < f->n; ++k) P.fields[k] = f->host[k];
    ptd::Params *dpar;
    if ((r = mem.alloc(&dpar, 1))) return r;
    P.self = dpar;
    PT_HIP(hipMemcpy(dpar, &P, sizeof(ptd::Params), hipMemcpyHostToDevice));
    PT_HIP(pt_launch_fields_eval(P, (mixed ? 1 : 0) | (f->general ? 2 : 0), coop, n, dp, dt, dB_, ddB, dV, dE, s));
    PT_HIP(hipStreamSynchronize(s));
    PT_HIP(hipMemcpy(B, dB_, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (dB) PT_HIP(hipMemcpy(dB, ddB, 9 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(V, dV, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(E, dE, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    return PT_OK;
}
}  // namespace

extern "C" int pt_fields_eval(const pt_fields *f, void *stream, int n, const double *pts, const double *t, double *B,
                              double *dB, double *V, double *E) {
    if (!f || n < 0 || (n > 0 && (!pts || !t || !B || !V || !E))) return fail(PT_EINVAL, "pt_fields_eval: bad arguments");
    if (n == 0) return PT_OK;
    return fields_eval(f, stream, n, pts, t, 0, 0, B, dB, V, E);
}

extern "C" int pt_fields_eval_tables(const pt_fields *f, void *stream, int n, const double *pts, const double *t,
                                     int variant, int cooperative, double *B, double *dB, double *V, double *E) {
    if (!f || n < 0 || (n > 0 && (!pts || !t || !B || !V || !E)) ||
        (variant != PT_VARIANT_FP64 && variant != PT_VARIANT_MIXED))
        return fail(PT_EINVAL, "pt_fields_eval_tables: bad arguments");
    if (variant == PT_VARIANT_MIXED && !f->has_fp32)
        return fail(PT_EINVAL, "pt_fields_eval_tables: the fields were created without mixed-precision tables");
    if (n == 0) return PT_OK;
    return fields_eval(f, stream, n, pts, t, variant == PT_VARIANT_MIXED, cooperative ? 1 : 0, B, dB, V, E);
}

/* TGeometry::GetSolid (src/geometry.cpp:151-175) at n points and times: solid IDs (1 = default);
   -1 where more solids nest than the tracker supports */
extern "C" int pt_geometry_solid_at(const pt_geometry *g, void *stream, int n, const double *pts, const double *t,
                                    int32_t *solid_id) {
    if (!g || n < 0 || (n > 0 && (!pts || !t || !solid_id))) return fail(PT_EINVAL, "pt_geometry_solid_at: bad arguments");
    if (n == 0) return PT_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevMem mem;
    double *dp, *dt;
    int *dout;
    int r;
    if ((r = mem.upload(&dp, pts, 3 * (size_t)n)) || (r = mem.upload(&dt, t, (size_t)n)) || (r = mem.alloc(&dout, (size_t)n)))
        return r;
    PT_HIP(pt_launch_solid_at(g->g, n, dp, dt, dout, s));
    PT_HIP(hipStreamSynchronize(s));
    PT_HIP(hipMemcpy(solid_id, dout, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return PT_OK;
}

/* intersection counts of n segments p[i] -> q[i] with the mesh (segment_hits_kernel's modes) */
extern "C" int pt_geometry_segment_hits(const pt_geometry *g, void *stream, int n, const double *p, const double *q,
                                        int mode, int32_t *out) {
    if (!g || n < 0 || mode < 0 || mode > 2 || (n > 0 && (!p || !q || !out)))
        return fail(PT_EINVAL, "pt_geometry_segment_hits: bad arguments");
    if (n == 0) return PT_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevMem mem;
    double *dp, *dq;
    int *dout;
    int r;
    if ((r = mem.upload(&dp, p, 3 * (size_t)n)) || (r = mem.upload(&dq, q, 3 * (size_t)n)) || (r = mem.alloc(&dout, (size_t)n)))
        return r;
    PT_HIP(pt_launch_segment_hits(g->g, n, dp, dq, mode, dout, s));
    PT_HIP(hipStreamSynchronize(s));
    PT_HIP(hipMemcpy(out, dout, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return PT_OK;
}

extern "C" int pt_batch_download(pt_batch *b, void *stream, pt_particle_out *out, pt_hit *hits) {
    if (!b || !out) return fail(PT_EINVAL, "pt_batch_download: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    PT_HIP(hipStreamSynchronize(s));
    int n = b->n;
    if (n == 0) return PT_OK;
    const ptd::State &S = b->P.st;
    std::vector<double> t(n), y(9 * (size_t)n), spin(5 * (size_t)n), Hmax(n), noflip(n), tau(n);
    std::vector<int> stop(n), Nhit(n), Nstep(n), Nsf(n), nsolid(n), nhit(n);
    std::vector<long long> acc(n), rej(n);
    std::vector<unsigned long long> solids(n);
    PT_HIP(hipMemcpy(t.data(), S.t, n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(y.data(), S.y, 9 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(spin.data(), S.spin, 5 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(Hmax.data(), S.Hmax, n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(noflip.data(), S.noflip, n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(tau.data(), S.tau, n * sizeof(double), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(stop.data(), S.stop, n * sizeof(int), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(Nhit.data(), S.Nhit, n * sizeof(int), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(Nstep.data(), S.Nstep, n * sizeof(int), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(Nsf.data(), S.Nspinflip, n * sizeof(int), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(nsolid.data(), S.nsolid, n * sizeof(int), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(nhit.data(), S.nhitlog, n * sizeof(int), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(acc.data(), S.n_acc, n * sizeof(long long), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(rej.data(), S.n_rej, n * sizeof(long long), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(solids.data(), S.solids, n * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    /* solid IDs for the end solid */
    const pt_geometry *G = b->geom;
    std::vector<uint32_t> sid(G->n_solid);
    if (G->n_solid) PT_HIP(hipMemcpy(sid.data(), G->g.solid_id, G->n_solid * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        pt_particle_out &o = out[i];
        o.t = t[i];
        for (int k = 0; k < 9; ++k) o.y[k] = y[(size_t)k * n + i];
        for (int k = 0; k < 5; ++k) o.spin[k] = spin[(size_t)k * n + i];
        o.stopID = stop[i]; o.Nhit = Nhit[i]; o.Nstep = Nstep[i]; o.Nspinflip = Nsf[i];
        o.n_acc = acc[i]; o.n_rej = rej[i];
        o.Hmax = Hmax[i]; o.noflipprob = noflip[i]; o.tau = tau[i];
        o.n_hits = nhit[i];
        /* TTracker::GetCurrentsolid on the final stack */
        unsigned long long sk = solids[i];
        int best = 0;
        auto id_of = [&](int e) -> uint32_t { int sidx = ptd::sk_sidx(sk, e); return sidx < 0 ? 1u : sid[sidx]; };
        for (int e = 1; e < nsolid[i]; ++e) {
            int ai = ptd::sk_ign(sk, best), bi = ptd::sk_ign(sk, e);
            if (ai || (!bi && id_of(best) < id_of(e))) best = e;
        }
        o.solidend = (int32_t)id_of(best);
    }
    if (hits && b->hit_cap > 0) {
        std::vector<ptd::Hit> h((size_t)n * b->hit_cap);
        PT_HIP(hipMemcpy(h.data(), S.hits, h.size() * sizeof(ptd::Hit), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < h.size(); ++k) {
            pt_hit &o = hits[k];
            o.t = h[k].t; o.tri = h[k].tri; o.leaving = h[k].leaving; o.entering = h[k].entering;
            for (int a = 0; a < 3; ++a) {
                o.normal[a] = h[k].normal[a]; o.v1[a] = h[k].v1[a]; o.v2[a] = h[k].v2[a]; o.pos[a] = h[k].pos[a];
            }
            o.pol1 = h[k].pol1; o.pol2 = h[k].pol2;
        }
    }
    return PT_OK;
}

/* TLogger::PrintSnapshot (src/logger.cpp:128-140): snapshot times in the caller's order (a
   non-decreasing list is de-duplicated and searched by bisection, which finds the same
   time as find_if) and the records kept per particle; n_times = 0 switches snapshots off */
extern "C" int pt_batch_set_snapshots(pt_batch *b, int n_times, const double *times, int cap) {
    if (!b || n_times < 0 || (n_times > 0 && (!times || cap <= 0))) return fail(PT_EINVAL, "pt_batch_set_snapshots: bad arguments");
    ptd::State &S = b->P.st;
    std::vector<double> t(times, times + n_times);
    const bool sorted = std::is_sorted(t.begin(), t.end());
    if (sorted) t.erase(std::unique(t.begin(), t.end()), t.end());
    S.snap_sorted = sorted ? 1 : 0;
    S.snap_n = 0;
    if (!t.empty()) {
        double *dt;
        ptd::Snap *ds;
        int *dn;
        int r;
        if ((r = b->mem.upload(&dt, t.data(), t.size())) || (r = b->mem.alloc(&ds, (size_t)b->n * cap))) return r;
        std::vector<int> zero((size_t)std::max(b->n, 1), 0);
        if ((r = b->mem.upload(&dn, zero.data(), zero.size()))) return r;
        S.snap_t = dt; S.snaps = ds; S.nsnap = dn; S.snap_cap = cap;
        S.snap_n = (int)t.size();
    }
    /* launch sequences captured before hold the old state by value */
    for (auto &g : b->graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
    b->graphs.clear();
    return PT_OK;
}

/* per-particle loss per bounce (n values; NULL switches it off): replaces the materials'
   LossPerBounce in TNeutron::OnHit's absorption probability (src/neutron.cpp:241) */
extern "C" int pt_batch_set_loss(pt_batch *b, const double *loss) {
    if (!b) return fail(PT_EINVAL, "pt_batch_set_loss: bad arguments");
    ptd::State &S = b->P.st;
    S.loss = nullptr;
    if (loss) {
        for (int i = 0; i < b->n; ++i)
            if (!(loss[i] >= 0.0 && loss[i] <= 1.0))
                return fail(PT_EINVAL, "pt_batch_set_loss: loss per bounce " + std::to_string(loss[i]) +
                                           " of particle " + std::to_string(i) + " outside [0, 1]");
        double *d;
        int r = b->mem.upload(&d, loss, (size_t)std::max(b->n, 1));
        if (r) return r;
        S.loss = d;
    }
    /* launch sequences captured before hold the old state by value */
    for (auto &g : b->graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
    b->graphs.clear();
    return PT_OK;
}

/* TLogger::PrintTrack: records of step ends where the trajectory length crossed a multiple of
   interval; interval = 0 switches the track log off */
extern "C" int pt_batch_set_track(pt_batch *b, double interval, int cap) {
    if (!b || interval < 0 || (interval > 0 && cap <= 0)) return fail(PT_EINVAL, "pt_batch_set_track: bad arguments");
    ptd::State &S = b->P.st;
    S.trk_interval = 0;
    if (interval > 0) {
        ptd::TrackRec *d;
        int *dn;
        int r;
        std::vector<int> zero((size_t)std::max(b->n, 1), 0);
        if ((r = b->mem.alloc(&d, (size_t)b->n * cap)) || (r = b->mem.upload(&dn, zero.data(), zero.size()))) return r;
        S.trk = d; S.ntrk = dn; S.trk_cap = cap; S.trk_interval = interval;
    }
    for (auto &g : b->graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
    b->graphs.clear();
    return PT_OK;
}

extern "C" int pt_batch_track(pt_batch *b, void *stream, pt_track_point *out, int32_t *count) {
    if (!b || !out || !count) return fail(PT_EINVAL, "pt_batch_track: bad arguments");
    const ptd::State &S = b->P.st;
    if (!(S.trk_interval > 0)) return fail(PT_EINVAL, "pt_batch_track: no track log set");
    PT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    const size_t n = (size_t)b->n, cap = (size_t)S.trk_cap;
    std::vector<ptd::TrackRec> h(n * cap);
    PT_HIP(hipMemcpy(h.data(), S.trk, h.size() * sizeof(ptd::TrackRec), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(count, S.ntrk, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < h.size(); ++k) {
        out[k].t = h[k].t;
        for (int q = 0; q < 9; ++q) out[k].y[q] = h[k].y[q];
        out[k].solid = h[k].solid;
        out[k].pad = 0;
    }
    return PT_OK;
}

/* TLogger::PrintSpin: a row after each BMT spin step that crossed a multiple of interval and
   at the start of every spin integration; interval = 0 switches the spin log off */
extern "C" int pt_batch_set_spinlog(pt_batch *b, double interval, int cap) {
    if (!b || interval < 0 || (interval > 0 && cap <= 0)) return fail(PT_EINVAL, "pt_batch_set_spinlog: bad arguments");
    ptd::State &S = b->P.st;
    S.slog_interval = 0;
    if (interval > 0) {
        ptd::SpinRec *d;
        int *dn;
        int r;
        std::vector<int> zero((size_t)std::max(b->n, 1), 0);
        if ((r = b->mem.alloc(&d, (size_t)b->n * cap)) || (r = b->mem.upload(&dn, zero.data(), zero.size()))) return r;
        S.slog = d; S.nslog = dn; S.slog_cap = cap; S.slog_interval = interval;
    }
    for (auto &g : b->graphs) { (void)hipGraphExecDestroy(g.exec); (void)hipGraphDestroy(g.graph); }
    b->graphs.clear();
    return PT_OK;
}

extern "C" int pt_batch_spinlog(pt_batch *b, void *stream, pt_spin_point *out, int32_t *count) {
    if (!b || !out || !count) return fail(PT_EINVAL, "pt_batch_spinlog: bad arguments");
    const ptd::State &S = b->P.st;
    if (!(S.slog_interval > 0)) return fail(PT_EINVAL, "pt_batch_spinlog: no spin log set");
    PT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    const size_t n = (size_t)b->n, cap = (size_t)S.slog_cap;
    static_assert(sizeof(ptd::SpinRec) == sizeof(pt_spin_point), "spin-log record layout");
    PT_HIP(hipMemcpy(out, S.slog, n * cap * sizeof(ptd::SpinRec), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(count, S.nslog, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    return PT_OK;
}

/* the snapshot records: out [n][cap] (cap of pt_batch_set_snapshots), count [n] (may exceed cap) */
extern "C" int pt_batch_snapshots(pt_batch *b, void *stream, pt_snapshot *out, int32_t *count) {
    if (!b || !out || !count) return fail(PT_EINVAL, "pt_batch_snapshots: bad arguments");
    const ptd::State &S = b->P.st;
    if (S.snap_n == 0) return fail(PT_EINVAL, "pt_batch_snapshots: no snapshot times set");
    PT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    const size_t n = (size_t)b->n, cap = (size_t)S.snap_cap;
    std::vector<ptd::Snap> h(n * cap);
    PT_HIP(hipMemcpy(h.data(), S.snaps, h.size() * sizeof(ptd::Snap), hipMemcpyDeviceToHost));
    PT_HIP(hipMemcpy(count, S.nsnap, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < h.size(); ++k) {
        pt_snapshot &o = out[k];
        o.t = h[k].t;
        for (int q = 0; q < 9; ++q) o.y[q] = h[k].y[q];
        for (int q = 0; q < 5; ++q) o.spin[q] = h[k].spin[q];
        o.Hmax = h[k].Hmax; o.noflipprob = h[k].noflip;
        o.stopID = h[k].stop; o.Nhit = h[k].Nhit; o.Nstep = h[k].Nstep; o.Nspinflip = h[k].Nspinflip;
    }
    return PT_OK;
}

extern "C" int pt_batch_step_counts(pt_batch *b, void *stream, int64_t *accepted, int64_t *rejected,
                                    int64_t *field_evals) {
    if (!b) return fail(PT_EINVAL, "pt_batch_step_counts: bad arguments");
    PT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    int n = b->n;
    std::vector<long long> acc(std::max(n, 1)), rej(std::max(n, 1)), nf(std::max(n, 1));
    if (n) {
        PT_HIP(hipMemcpy(acc.data(), b->P.st.n_acc, n * sizeof(long long), hipMemcpyDeviceToHost));
        PT_HIP(hipMemcpy(rej.data(), b->P.st.n_rej, n * sizeof(long long), hipMemcpyDeviceToHost));
        PT_HIP(hipMemcpy(nf.data(), b->P.st.n_field, n * sizeof(long long), hipMemcpyDeviceToHost));
    }
    long long sa = 0, sr = 0, sf = 0;
    for (int i = 0; i < n; ++i) { sa += acc[i]; sr += rej[i]; sf += nf[i]; }
    if (accepted) *accepted = sa;
    if (rejected) *rejected = sr;
    if (field_evals) *field_evals = sf;
    return PT_OK;
}

extern "C" int pt_batch_set_profiling(pt_batch *b, int on) {
    if (!b) return fail(PT_EINVAL, "pt_batch_set_profiling: bad arguments");
    b->profiling = on != 0;
    return PT_OK;
}

extern "C" int pt_batch_kernel_times(pt_batch *b, void *stream, double *ms, int64_t *launches) {
    if (!b || !ms || !launches) return fail(PT_EINVAL, "pt_batch_kernel_times: bad arguments");
    PT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    for (int k = 0; k < 2; ++k) {
        for (auto &e : b->ev[k]) {
            float t = 0;
            PT_HIP(hipEventElapsedTime(&t, e.first, e.second));
            b->ms[k] += t;
            b->launches[k] += 1;
            PT_HIP(hipEventDestroy(e.first));
            PT_HIP(hipEventDestroy(e.second));
        }
        b->ev[k].clear();
        ms[k] = b->ms[k];
        launches[k] = b->launches[k];
    }
    return PT_OK;
}

extern "C" int pt_track_batch(void *stream, const pt_geometry *g, const pt_fields *f, int n_species,
                              const pt_species *species, double tmax_global, int n, const int32_t *species_idx,
                              const double *t0, const double *y0, const double *spin0, uint64_t seed,
                              const uint64_t *pid, pt_particle_out *out, pt_hit *hits, int hit_cap,
                              const pt_launch_cfg *cfg) {
    pt_batch *b = nullptr;
    int r = pt_batch_create(g, f, n_species, species, tmax_global, n, species_idx, t0, y0, spin0, seed, pid, hit_cap, &b);
    if (r) return r;
    std::unique_ptr<pt_batch> guard(b);
    pt_launch_cfg c = cfg ? *cfg : pt_launch_cfg{PT_VARIANT_FP64, 256, 0, 0, 0, 0, 0};
    for (;;) {
        r = pt_batch_run(b, stream, &c);
        if (r) return r;
        int active = 0;
        r = pt_batch_active(b, stream, &active);
        if (r) return r;
        if (active == 0 || g_quit.load()) break;
    }
    return pt_batch_download(b, stream, out, hits);
}

extern "C" void pt_request_quit(int on) { g_quit.store(on ? 1 : 0); }

extern "C" const char *pt_last_error(void) { return g_last_error.c_str(); }

/* for the library's other translation units: record msg as this thread's last error */
int pt_fail(int code, const char *msg) { return fail(code, msg); }
