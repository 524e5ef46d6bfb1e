// Device-side ray queries: two-level BVH2 traversal with an LDS stack, the FP32
// watertight triangle test and the FP64 sphere test.
//
// Contraction is off in every function here (#pragma clang fp contract(off)):
// the Woop edge functions must round each product on its own so shared edges
// get exactly negated values (reference geometry/triangle.cpp:115-122, built
// with -ffp-contract=off, src/CMakeLists.txt:118-124), and the instance
// transform must match the FP32 restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"

namespace ab {

constexpr uint32_t kMissId = 0xffffffffu;

struct RayF {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};

struct HitF {
    float t;
    uint32_t object;
    uint32_t prim;
    float b1, b2;
};

__device__ __forceinline__ bool betterHit(float t, uint32_t obj, uint32_t prim, const HitF& h)
{
    // (t, object, prim) lexicographic order, any hit before none; written with
    // bitwise operators so it compiles to compares without branches (a NaN t
    // compares false both ways and never wins over a hit).
    return (h.object == kMissId) | (t < h.t) | ((t == h.t) & ((obj < h.object) | ((obj == h.object) & (prim < h.prim))));
}

// Sphere test in FP64 (reference geometry/sphere.cpp:92-135): Haines'
// discriminant a (r^2 - |l|^2) and the stable roots q/a, c/q.
__device__ __forceinline__ bool sphereHit(const double* s, const RayF& r, double& t_out)
{
#pragma clang fp contract(off)
    const double ox = double(r.ox) - s[0], oy = double(r.oy) - s[1], oz = double(r.oz) - s[2];
    const double dx = r.dx, dy = r.dy, dz = r.dz;
    const double a = dx * dx + dy * dy + dz * dz;
    if (!(a >= 2.2250738585072014e-308 && a <= 1.7976931348623157e308)) return false;
    const double h = dx * ox + dy * oy + dz * oz;
    const double rad = s[3];
    const double c = (ox * ox + oy * oy + oz * oz) - rad * rad;
    const double ha = h / a;
    const double lx = ox - dx * ha, ly = oy - dy * ha, lz = oz - dz * ha;
    const double disc = a * (rad * rad - (lx * lx + ly * ly + lz * lz));
    if (disc < 0) return false;
    const double q = -(h + copysign(sqrt(disc), h));
    double tn = 0, tf = 0;
    if (q != 0) {
        tn = q / a;
        tf = c / q;
        if (tn > tf) {
            const double tmp = tn;
            tn = tf;
            tf = tmp;
        }
    }
    const double tmin = r.tmin, tmax = r.tmax;
    double t = tn;
    if (!(tmin < t && t < tmax)) {
        t = tf;
        if (!(tmin < t && t < tmax)) return false;
    }
    // The reference forms p = o + d t first, then p - C (sphere.cpp:120-122).
    const double qx = (double(r.ox) + dx * t) - s[0], qy = (double(r.oy) + dy * t) - s[1],
                 qz = (double(r.oz) + dz * t) - s[2];
    if (!(qx * qx + qy * qy + qz * qz > 0)) return false;
    t_out = t;
    return true;
}

// The ray in the sheared space of Woop et al. 2013, set up once per (object
// space) ray: kz, the axis of the direction's largest component, the axes
// (kx, ky) = (kz + 1, kz + 2) mod 3, and the shear (sx, sy, sz). The triangle
// test reads the copy of the triangles stored in that order (DevScene::tris),
// so it selects no coordinates per lane. The reference's test
// (geometry/triangle.cpp:95-131) also swaps kx and ky when the direction's
// largest component is negative; the swap only negates the edge functions u,
// v, w and their sum det, and negation is exact in floating point, so the
// hit's t = (u az + v bz + w cz) / det and barycentrics v / det, w / det are
// the same bits without it.
struct ShearRay {
    float ox, oy, oz; // origin in the order (kx, ky, kz)
    float sx, sy, sz;
    uint32_t kz;
    bool valid;       // |d|^2 is a normal number
    uint32_t tri0;    // index of the first triangle of the copy in this order (set with useCopy)
    __device__ __forceinline__ void useCopy(const DevScene& sc) { tri0 = kz * sc.tri_count; }
};

// Triangle i of the ray's copy; the byte offset fits 32 bits (kMaxTriangles), so
// where the scene's arrays are wave-uniform the loads take a scalar base.
__device__ __forceinline__ const float4* triOf(const DevScene& sc, const ShearRay& sh, uint32_t i)
{
    return reinterpret_cast<const float4*>(reinterpret_cast<const char*>(sc.tris) +
                                           (sh.tri0 + i) * uint32_t(sizeof(DevTri)));
}

__device__ __forceinline__ float pick(float x, float y, float z, int k)
{
    return k == 0 ? x : (k == 1 ? y : z);
}

__device__ __forceinline__ ShearRay makeShear(float ox, float oy, float oz, float dx, float dy, float dz)
{
#pragma clang fp contract(off)
    ShearRay s;
    const float l2 = (dx * dx + dy * dy) + dz * dz;
    s.valid = l2 >= 1.17549435e-38f && l2 <= 3.40282347e38f;
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    int kz;
    if (ax > ay) kz = ax > az ? 0 : 2;
    else kz = ay > az ? 1 : 2;
    const int kx = kz == 2 ? 0 : kz + 1;
    const int ky = kx == 2 ? 0 : kx + 1;
    const float dzv = pick(dx, dy, dz, kz);
    s.kz = uint32_t(kz);
    s.sx = pick(dx, dy, dz, kx) / dzv;
    s.sy = pick(dx, dy, dz, ky) / dzv;
    s.sz = 1.0f / dzv;
    s.ox = pick(ox, oy, oz, kx);
    s.oy = pick(ox, oy, oz, ky);
    s.oz = pick(ox, oy, oz, kz);
    return s;
}

// Watertight triangle test (reference geometry/triangle.cpp:132-176) in FP32,
// with the edge functions recomputed in FP64 when one is exactly zero, on a
// triangle from the copy of the ray's kz (coordinates already in the order
// (kx, ky, kz)). The barycentrics are v / det and w / det; callers divide only
// for a hit they keep.
__device__ __forceinline__ bool triHit(const ShearRay& s, const float4& v0, const float4& v1, const float4& v2,
                                       float tmin, float tmax, float& t_out, float& v_out, float& w_out,
                                       float& det_out)
{
#pragma clang fp contract(off)
    const float az_ = v0.z - s.oz;
    const float bz_ = v1.z - s.oz;
    const float cz_ = v2.z - s.oz;
    const float ax = (v0.x - s.ox) - s.sx * az_;
    const float ay = (v0.y - s.oy) - s.sy * az_;
    const float bx = (v1.x - s.ox) - s.sx * bz_;
    const float by = (v1.y - s.oy) - s.sy * bz_;
    const float cx = (v2.x - s.ox) - s.sx * cz_;
    const float cy = (v2.y - s.oy) - s.sy * cz_;
    float u = cx * by - cy * bx;
    float v = ax * cy - ay * cx;
    float w = bx * ay - by * ax;
    if (u == 0.0f || v == 0.0f || w == 0.0f) {
        u = float(double(cx) * double(by) - double(cy) * double(bx));
        v = float(double(ax) * double(cy) - double(ay) * double(cx));
        w = float(double(bx) * double(ay) - double(by) * double(ax));
    }
    const bool inside = (u >= 0.0f && v >= 0.0f && w >= 0.0f) || (u <= 0.0f && v <= 0.0f && w <= 0.0f);
    if (!inside) return false;
    const float det = (u + v) + w;
    if (det == 0.0f) return false;
    const float az = s.sz * az_, bz = s.sz * bz_, cz = s.sz * cz_;
    const float t = ((u * az + v * bz) + w * cz) / det;
    if (!(tmin < t && t < tmax)) return false;
    t_out = t;
    v_out = v;
    w_out = w;
    det_out = det;
    return true;
}

// Slab test of a node's two children (Ize 2013: exits widened by 1 + 2^-20, more
// than the 1 + 2 gamma(3) rounding bound). A zero direction component is
// replaced by +-1e-20 so no 0 * inf NaN can arise; the test stays conservative.
// The inverse direction comes from the hardware reciprocal (v_rcp_f32, within
// 1 ulp) instead of an IEEE division: that adds at most 1.5 ulp to each slab
// distance, well inside the 16-ulp (2^-20) widening, so the box test stays
// conservative and hits do not change (the test only decides which nodes are
// visited). Rays start, enter and leave instances with it: coherent trefoil set
// 11 459-11 588 -> 11 715-11 725 Mrays/s, full-feature scene 1.59 -> 1.56 s.
struct BoxRay {
    float ox, oy, oz, ix, iy, iz;
};

__device__ __forceinline__ BoxRay makeBoxRay(float ox, float oy, float oz, float dx, float dy, float dz)
{
    BoxRay b;
    b.ox = ox;
    b.oy = oy;
    b.oz = oz;
    b.ix = __builtin_amdgcn_rcpf(fabsf(dx) > 1e-20f ? dx : copysignf(1e-20f, dx));
    b.iy = __builtin_amdgcn_rcpf(fabsf(dy) > 1e-20f ? dy : copysignf(1e-20f, dy));
    b.iz = __builtin_amdgcn_rcpf(fabsf(dz) > 1e-20f ? dz : copysignf(1e-20f, dz));
    return b;
}

// f * (1 + 2^-20) for f >= 0 and f * (1 - 2^-20) for f < 0, i.e. f + |f| 2^-20
// rounded once: one fused multiply-add gives the same value for every finite f
// (and +inf for +inf; -inf, which no finite box yields, would give NaN and then
// pass the box, never lose it).
__device__ __forceinline__ float widenFar(float f)
{
    return fmaf(fabsf(f), 9.5367431640625e-07f, f);
}

// One axis: the near and the (unwidened) far distance of the slab.
__device__ __forceinline__ void slab(float lo, float hi, float o, float inv, float& tn, float& tf)
{
#pragma clang fp contract(off)
    const float t0 = (lo - o) * inv, t1 = (hi - o) * inv;
    tn = fminf(t0, t1);
    tf = fmaxf(t0, t1);
}

// widenFar and its rounding are non-decreasing, so widening the smallest far
// distance equals taking the smallest widened one: one widening per box.
__device__ __forceinline__ void testChildren(const DevNode& n, const BoxRay& r, float tlo, float thi, bool& h0,
                                             bool& h1, float& n0, float& n1)
{
    float nx0, fx0, ny0, fy0, nz0, fz0, nx1, fx1, ny1, fy1, nz1, fz1;
    slab(n.a[0], n.a[1], r.ox, r.ix, nx0, fx0);
    slab(n.a[2], n.a[3], r.oy, r.iy, ny0, fy0);
    slab(n.b[0], n.b[1], r.oz, r.iz, nz0, fz0);
    slab(n.b[2], n.b[3], r.ox, r.ix, nx1, fx1);
    slab(n.c[0], n.c[1], r.oy, r.iy, ny1, fy1);
    slab(n.c[2], n.c[3], r.oz, r.iz, nz1, fz1);
    const float tn0 = fmaxf(fmaxf(tlo, nx0), fmaxf(ny0, nz0));
    const float tn1 = fmaxf(fmaxf(tlo, nx1), fmaxf(ny1, nz1));
    const float tf0 = fminf(thi, widenFar(fminf(fminf(fx0, fy0), fz0)));
    const float tf1 = fminf(thi, widenFar(fminf(fminf(fx1, fy1), fz1)));
    h0 = tn0 <= tf0;
    h1 = tn1 <= tf1 && n.child[1] != kNoChild;
    n0 = tn0;
    n1 = tn1;
}

// Node i of a scene (i >= 0, fewer than kMaxNodes): its byte offset fits 32
// bits, so where `nodes` is wave-uniform the loads take a scalar base and a
// 32-bit vector offset instead of a 64-bit address per lane.
__device__ __forceinline__ DevNode loadNode(const DevNode* nodes, int32_t i)
{
    const float4* p = reinterpret_cast<const float4*>(reinterpret_cast<const char*>(nodes) +
                                                      uint32_t(i) * uint32_t(sizeof(DevNode)));
    const float4 a = p[0], b = p[1], c = p[2];
    const int4 m = reinterpret_cast<const int4*>(p)[3];
    DevNode n;
    n.a[0] = a.x; n.a[1] = a.y; n.a[2] = a.z; n.a[3] = a.w;
    n.b[0] = b.x; n.b[1] = b.y; n.b[2] = b.z; n.b[3] = b.w;
    n.c[0] = c.x; n.c[1] = c.y; n.c[2] = c.z; n.c[3] = c.w;
    n.child[0] = m.x;
    n.child[1] = m.y;
    return n;
}

// Upper end of the t range used to cull boxes: a little above the best hit so a
// triangle whose rounded t ties or beats it is never culled by box rounding.
__device__ __forceinline__ float cullLimit(float t)
{
    return t * (1.0f + 3.814697265625e-06f);
}

// LDS stack: entry k of this thread lives at stack[k * BLOCK].
template <int BLOCK>
struct LdsStack {
    int32_t* base;
    __device__ __forceinline__ void put(int sp, int32_t v) { base[sp * BLOCK] = v; }
    __device__ __forceinline__ int32_t get(int sp) const { return base[sp * BLOCK]; }
};

// BLAS traversal of one instance with the ray already in object space.
// ANY: stop at the first hit. Returns true when ANY found a hit.
template <int BLOCK, bool ANY>
__device__ __forceinline__ bool traverseBlas(const DevScene& sc, int32_t root, uint32_t object, const RayF& r,
                                             HitF& best, LdsStack<BLOCK> stk, int sp0)
{
    ShearRay sh = makeShear(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz);
    if (!sh.valid) return false;
    sh.useCopy(sc);
    const BoxRay br = makeBoxRay(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz);
    int sp = sp0;
    int32_t ref = root;
    while (true) {
        if (ref >= 0) {
            const DevNode n = loadNode(sc.nodes, ref);
            bool h0, h1;
            float t0, t1;
            const float lim = cullLimit(ANY ? r.tmax : best.t);
            testChildren(n, br, r.tmin, lim, h0, h1, t0, t1);
            if (h0 && h1) {
                const bool swap = t1 < t0;
                stk.put(sp++, swap ? n.child[0] : n.child[1]);
                ref = swap ? n.child[1] : n.child[0];
                continue;
            }
            if (h0 || h1) {
                ref = h0 ? n.child[0] : n.child[1];
                continue;
            }
        } else {
            const uint32_t enc = uint32_t(~ref);
            const uint32_t first = enc >> kLeafCountBits, count = enc & ((1u << kLeafCountBits) - 1);
            for (uint32_t i = 0; i < count; ++i) {
                const float4* tp = triOf(sc, sh, first + i);
                const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
                float t, v, w, det;
                if (triHit(sh, v0, v1, v2, r.tmin, r.tmax, t, v, w, det)) {
                    if (ANY) return true;
                    const uint32_t prim = __float_as_uint(v0.w);
                    if (betterHit(t, object, prim, best)) {
                        best.t = t;
                        best.object = object;
                        best.prim = prim;
                        best.b1 = v / det;
                        best.b2 = w / det;
                    }
                }
            }
        }
        if (sp == sp0) return false;
        ref = stk.get(--sp);
    }
}

// Object-space ray of an instance (reference MeshInstance::intersect,
// geometry/mesh_instance.cpp:32-36, Transform::apply core/transform.cpp:191-210).
__device__ __forceinline__ RayF toObject(const float* m, const RayF& r)
{
#pragma clang fp contract(off)
    RayF o;
    o.ox = ((m[0] * r.ox + m[1] * r.oy) + m[2] * r.oz) + m[3];
    o.oy = ((m[4] * r.ox + m[5] * r.oy) + m[6] * r.oz) + m[7];
    o.oz = ((m[8] * r.ox + m[9] * r.oy) + m[10] * r.oz) + m[11];
    o.dx = (m[0] * r.dx + m[1] * r.dy) + m[2] * r.dz;
    o.dy = (m[4] * r.dx + m[5] * r.dy) + m[6] * r.dz;
    o.dz = (m[8] * r.dx + m[9] * r.dy) + m[10] * r.dz;
    o.tmin = r.tmin;
    o.tmax = r.tmax;
    return o;
}

// Tests one scene object. Returns true when ANY found a hit.
template <int BLOCK, bool ANY>
__device__ __forceinline__ bool testObject(const DevScene& sc, uint32_t obj, const RayF& r, HitF& best,
                                           LdsStack<BLOCK> stk, int sp)
{
    const DevObject& o = sc.objects[obj];
    if (o.kind == kObjSphere) {
        double t;
        if (sphereHit(sc.spheres + 4 * size_t(obj), r, t)) {
            if (ANY) return true;
            const float tf = float(t);
            if (betterHit(tf, obj, kMissId, best)) {
                best.t = tf;
                best.object = obj;
                best.prim = kMissId;
                best.b1 = 0.0f;
                best.b2 = 0.0f;
            }
        }
        return false;
    }
    float m[12];
    const float4* mp = reinterpret_cast<const float4*>(o.to_object);
    const float4 m0 = mp[0], m1 = mp[1], m2 = mp[2];
    m[0] = m0.x; m[1] = m0.y; m[2] = m0.z; m[3] = m0.w;
    m[4] = m1.x; m[5] = m1.y; m[6] = m1.z; m[7] = m1.w;
    m[8] = m2.x; m[9] = m2.y; m[10] = m2.z; m[11] = m2.w;
    const RayF ro = toObject(m, r);
    return traverseBlas<BLOCK, ANY>(sc, o.blas_root, obj, ro, best, stk, sp);
}

// Full two-level query. ANY: returns true at the first hit in (tmin, tmax);
// otherwise fills `best` with the lexicographic minimum of (t, object, prim).
template <int BLOCK, bool ANY>
__device__ __forceinline__ bool traceRay(const DevScene& sc, const RayF& r, HitF& best, LdsStack<BLOCK> stk)
{
    best.t = r.tmax;
    best.object = kMissId;
    best.prim = kMissId;
    best.b1 = best.b2 = 0.0f;
    if (sc.tlas_root == kNoChild) return false;
    const BoxRay br = makeBoxRay(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz);
    int sp = 0;
    int32_t ref = sc.tlas_root;
    while (true) {
        if (ref >= 0) {
            const DevNode n = loadNode(sc.nodes, ref);
            bool h0, h1;
            float t0, t1;
            const float lim = cullLimit(ANY ? r.tmax : best.t);
            testChildren(n, br, r.tmin, lim, h0, h1, t0, t1);
            if (h0 && h1) {
                const bool swap = t1 < t0;
                stk.put(sp++, swap ? n.child[0] : n.child[1]);
                ref = swap ? n.child[1] : n.child[0];
                continue;
            }
            if (h0 || h1) {
                ref = h0 ? n.child[0] : n.child[1];
                continue;
            }
        } else {
            const uint32_t enc = uint32_t(~ref);
            const uint32_t first = enc >> kLeafCountBits, count = enc & ((1u << kLeafCountBits) - 1);
            for (uint32_t i = 0; i < count; ++i) {
                const uint32_t obj = sc.leafobj[first + i];
                if (testObject<BLOCK, ANY>(sc, obj, r, best, stk, sp)) return true;
            }
        }
        if (sp == 0) return false;
        ref = stk.get(--sp);
    }
}

// The two-level query as a resumable state machine: step() does a bounded amount
// of work, so a lane whose ray has finished can take a new ray while the other
// lanes of its wave continue (see the kernels that refill lanes).
//
// A lane's triangle tests overlap its node visits. A BLAS leaf the traversal
// reaches moves into a leaf slot and the lane goes on to its next stack entry;
// each step then tests one triangle of the slot's leaf AND visits the lane's
// next node. A wave whose lanes sit at different points of their traversals
// runs both kinds of code in most steps anyway; this way a lane busy with a leaf
// keeps walking nodes in the same steps, so a ray takes about max(node visits,
// triangle tests) steps rather than their sum (trefoil sets: 3 % more rays per
// second than one node visit or one triangle per step). A leaf that arrives
// while the slot is busy waits in `ref`, and an instance is left only once its
// triangles are done (the triangle test uses the instance's ray). Pushes happen
// where traceRay pushes, and a pop only ever replaces a leaf that has moved to
// the slot, so the stack never holds more than traceRay's does.
//
// The culling limit lags by the triangles still in the slot, which can add node
// visits but never loses a hit: the answer is the lexicographic minimum of
// (t, object, prim) however the nodes are visited, so it is bit for bit
// traceRay's. A TLAS leaf is taken one object at a time; the leaf's other
// objects wait as a shorter leaf reference, either on the stack (one entry more
// than traceRay's) or, with LEAF_IN_REG, in a register (pending) taken up before
// the next TLAS stack entry (no extra entry, one more live register and check).
// SPHERES = false compiles the FP64 sphere test out, for scenes whose objects are
// all mesh instances.
// Compare-exchange of two (key, child) pairs into ascending key order.
__device__ __forceinline__ void orderPair(float& ka, int32_t& ca, float& kb, int32_t& cb)
{
    const bool swap = kb < ka;
    const float k = swap ? kb : ka;
    const int32_t c = swap ? cb : ca;
    kb = swap ? ka : kb;
    cb = swap ? ca : cb;
    ka = k;
    ca = c;
}

template <int BLOCK, bool ANY, bool SPHERES = true, bool LEAF_IN_REG = false, bool WIDE = false>
struct RayQuery {
    RayF r;          // world-space ray
    BoxRay box;      // box-test form of the ray of the current level (world or instance space)
    ShearRay sh;     // triangle-test form of the instance-space ray
    HitF best;       // b1, b2 hold Woop's v, w until result() divides them by bdet
    float bdet;      // the best triangle's det (1 for a sphere or no hit)
    float lim;       // box-culling limit: cullLimit of best.t, updated with it
    int32_t ref;     // next node or leaf to visit; kNoChild when the lane must pop
    int32_t leaf;    // BLAS leaf whose triangles are being tested, or kNoChild
    int32_t pending; // the current TLAS leaf's objects not yet taken (LEAF_IN_REG), or kNoChild
    int sp, sp_blas; // stack depth, and where the current instance's entries start
    uint32_t obj;    // instance being traversed
    bool in_blas, done;

    __device__ __forceinline__ void begin(const DevScene& sc, const RayF& ray)
    {
        r = ray;
        best.t = r.tmax;
        best.object = kMissId;
        best.prim = kMissId;
        best.b1 = best.b2 = 0.0f;
        bdet = 1.0f;
        lim = cullLimit(r.tmax);
        box = makeBoxRay(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz);
        ref = WIDE ? sc.wide_root : sc.tlas_root;
        leaf = kNoChild;
        pending = kNoChild;
        sp = sp_blas = 0;
        in_blas = false;
        // An empty range (tmax <= tmin, or a NaN bound) admits no hit.
        done = ref == kNoChild || !(r.tmax > r.tmin);
    }

    // Starts from a hit found over other objects of the scene (a closest-hit
    // query continued over the rest): the answer is then the lexicographic
    // minimum over both sets, and boxes beyond the hit are culled from the start.
    __device__ __forceinline__ void seed(const HitF& h)
    {
        if (h.object == kMissId) return;
        best = h;
        bdet = 1.0f;
        lim = cullLimit(h.t);
    }

    // The best hit with its barycentrics. Each accepted triangle keeps v, w and
    // det; the two divisions run once here instead of at every acceptance, with
    // the same operands, so the values are those an immediate division gives.
    __device__ __forceinline__ HitF result() const
    {
        HitF h = best;
        if (!ANY) {
            h.b1 = best.b1 / bdet;
            h.b2 = best.b2 / bdet;
        }
        return h;
    }

    __device__ __forceinline__ bool found() const { return ANY && best.object != kMissId; }

    // Moves a BLAS leaf in `ref` into a free slot, and takes the next entry when
    // `ref` is empty: the rest of the TLAS leaf (LEAF_IN_REG) or the stack top,
    // leaving the instance first when its entries are used up and its last
    // triangles tested.
    __device__ __forceinline__ void absorb()
    {
        if (in_blas && ref < 0 && ref != kNoChild && leaf == kNoChild) {
            leaf = ref;
            ref = kNoChild;
        }
    }

    __device__ __forceinline__ void refill(LdsStack<BLOCK> stk)
    {
        if (ref != kNoChild) return;
        if (in_blas && sp == sp_blas) {
            if (leaf != kNoChild) return;
            in_blas = false;
            box = makeBoxRay(r.ox, r.oy, r.oz, r.dx, r.dy, r.dz);
        }
        if (LEAF_IN_REG && !in_blas && pending != kNoChild) {
            ref = pending;
            pending = kNoChild;
            return;
        }
        if (sp == 0) {
            // Not in an instance here, so the leaf slot is empty too.
            done = true;
            return;
        }
        ref = stk.get(--sp);
    }

    __device__ __forceinline__ void step(const DevScene& sc, LdsStack<BLOCK> stk)
    {
        const bool has_leaf = leaf != kNoChild;
        const bool has_node = ref >= 0;
        if (has_leaf) {
            const uint32_t enc = uint32_t(~leaf);
            const uint32_t first = enc >> kLeafCountBits, count = enc & ((1u << kLeafCountBits) - 1);
            const float4* tp = triOf(sc, sh, first);
            const float4 v0 = tp[0], v1 = tp[1], v2 = tp[2];
            float t, v, w, det;
            if (triHit(sh, v0, v1, v2, r.tmin, r.tmax, t, v, w, det)) {
                const uint32_t prim = __float_as_uint(v0.w);
                if (ANY || betterHit(t, obj, prim, best)) {
                    best.t = t;
                    best.object = obj;
                    best.prim = prim;
                    best.b1 = ANY ? 0.0f : v;
                    best.b2 = ANY ? 0.0f : w;
                    bdet = ANY ? 1.0f : det;
                    if (!ANY) lim = cullLimit(t);
                }
            }
            leaf = count > 1 ? encodeLeaf(first + 1, count - 1) : kNoChild;
            if (found()) {
                done = true;
                return;
            }
        }
        if (has_node && WIDE) {
            // Four children: the entered ones in ascending entry distance, the
            // nearest next and the others pushed farthest first.
            const DevNode n0 = loadNode(sc.nodes, ref), n1 = loadNode(sc.nodes, ref + 1);
            bool h0, h1, h2, h3;
            float k0, k1, k2, k3;
            testChildren(n0, box, r.tmin, lim, h0, h1, k0, k1);
            testChildren(n1, box, r.tmin, lim, h2, h3, k2, k3);
            constexpr float kInf = __builtin_huge_valf();
            int32_t c0 = h0 ? n0.child[0] : kNoChild, c1 = h1 ? n0.child[1] : kNoChild;
            int32_t c2 = h2 ? n1.child[0] : kNoChild, c3 = h3 ? n1.child[1] : kNoChild;
            k0 = h0 ? k0 : kInf;
            k1 = h1 ? k1 : kInf;
            k2 = c2 != kNoChild ? k2 : kInf;
            k3 = h3 ? k3 : kInf;
            orderPair(k0, c0, k1, c1);
            orderPair(k2, c2, k3, c3);
            orderPair(k0, c0, k2, c2);
            orderPair(k1, c1, k3, c3);
            orderPair(k1, c1, k2, c2);
            if (c3 != kNoChild) stk.put(sp++, c3);
            if (c2 != kNoChild) stk.put(sp++, c2);
            if (c1 != kNoChild) stk.put(sp++, c1);
            ref = c0;
        } else if (has_node) {
            const DevNode n = loadNode(sc.nodes, ref);
            bool h0, h1;
            float t0, t1;
            testChildren(n, box, r.tmin, lim, h0, h1, t0, t1);
            if (h0 && h1) {
                const bool swap = t1 < t0;
                stk.put(sp++, swap ? n.child[0] : n.child[1]);
                ref = swap ? n.child[1] : n.child[0];
            } else if (h0 || h1) {
                ref = h0 ? n.child[0] : n.child[1];
            } else {
                ref = kNoChild;
            }
        } else if (!in_blas && ref != kNoChild) {
            takeObject(sc, stk);
            if (done) return;
        }
        absorb();
        refill(stk);
        absorb();
    }

    // A TLAS leaf in `ref`: its first object is tested (sphere) or entered
    // (instance); the others wait as a shorter leaf reference.
    __device__ __forceinline__ void takeObject(const DevScene& sc, LdsStack<BLOCK> stk)
    {
        const uint32_t enc = uint32_t(~ref);
        const uint32_t first = enc >> kLeafCountBits, count = enc & ((1u << kLeafCountBits) - 1);
        ref = kNoChild;
        if (count == 0) return;
        const uint32_t o = sc.leafobj[first];
        if (count > 1) {
            if (LEAF_IN_REG) pending = encodeLeaf(first + 1, count - 1);
            else stk.put(sp++, encodeLeaf(first + 1, count - 1));
        }
        const DevObject& d = sc.objects[o];
        if (SPHERES && d.kind == kObjSphere) {
            double t;
            if (sphereHit(sc.spheres + 4 * size_t(o), r, t)) {
                const float tf = float(t);
                if (ANY || betterHit(tf, o, kMissId, best)) {
                    best.t = tf;
                    best.object = o;
                    best.prim = kMissId;
                    best.b1 = 0.0f;
                    best.b2 = 0.0f;
                    bdet = 1.0f;
                    if (!ANY) lim = cullLimit(tf);
                }
            }
            if (found()) done = true;
            return;
        }
        float m[12];
        const float4* mp = reinterpret_cast<const float4*>(d.to_object);
        const float4 m0 = mp[0], m1 = mp[1], m2 = mp[2];
        m[0] = m0.x; m[1] = m0.y; m[2] = m0.z; m[3] = m0.w;
        m[4] = m1.x; m[5] = m1.y; m[6] = m1.z; m[7] = m1.w;
        m[8] = m2.x; m[9] = m2.y; m[10] = m2.z; m[11] = m2.w;
        const RayF ro = toObject(m, r);
        sh = makeShear(ro.ox, ro.oy, ro.oz, ro.dx, ro.dy, ro.dz);
        if (!sh.valid) return;
        sh.useCopy(sc);
        box = makeBoxRay(ro.ox, ro.oy, ro.oz, ro.dx, ro.dy, ro.dz);
        obj = o;
        in_blas = true;
        sp_blas = sp;
        ref = WIDE ? d.wide_root : d.blas_root;
    }
};

} // namespace ab
