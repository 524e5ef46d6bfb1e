// Oracle scene, estimator and C API (liboracle.so). Test infrastructure only:
// called from tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker; never linked into or called by the product library.
//
// Scene:        reference scene/scene.cpp:20-182 (TLAS over objects, nearest hit,
//               occlusion), geometry/mesh_geometry.cpp, geometry/mesh_instance.cpp.
// Estimator:    reference estimator/path_estimator.cpp:47-185, surface_ray.hpp:18-31.
// Renderer:     reference render/render_image.cpp:27-77 (per-pixel stream,
//               stratified jitter, row-major cells).
// Camera:       reference camera/perspective_camera.cpp:74-110.
//
// Closest hits use the lexicographic minimum of (t, object, prim); the reference
// keeps the first of equal-t hits in its traversal order (accel/bvh.hpp:218-222),
// which depends on its std::partition order and cannot be pinned without running
// it. With that rule the answer is independent of BVH layout, so the oracle's own
// BVH, brute force and the GPU must all agree.

#include "../include/albedo_gpu.h"
#include "oracle_shade.hpp"

#include <atomic>
#include <memory>
#include <thread>

namespace oracle {

struct Mesh {
    std::vector<std::array<V3f, 3>> tri_f;    // kept triangles, FP32 device values
    std::vector<std::array<V3d, 3>> tri_d;    // the same values in FP64
    std::vector<V3d> face_normal;             // FP64 unit front normal
    std::vector<std::array<V3d, 3>> corner_n; // smooth only
    std::vector<std::array<double, 6>> uv;    // corner uv
    Bvh blas;
    Box box;
};

struct Object {
    uint32_t kind, material;
    int32_t light;
    uint32_t mesh;
    V3d center;
    double radius;
    Affine<double> to_world, to_object;
    Affine<float> to_object_f;
    Box world_box;
};

struct Hit {
    bool hit{false};
    double t{std::numeric_limits<double>::infinity()};
    uint32_t object{AB_MISS}, prim{AB_NO_PRIM};
    double b1{0}, b2{0};
};

inline bool better(double t, uint32_t obj, uint32_t prim, const Hit& best)
{
    if (!best.hit) return true;
    if (t != best.t) return t < best.t;
    if (obj != best.object) return obj < best.object;
    return prim < best.prim;
}

struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Object> objects;
    std::vector<ab_material_desc> materials;
    std::vector<ab_light_desc> lights;
    int sky_light{-1};
    Bvh tlas;

    explicit Scene(const ab_scene_desc& d)
    {
        for (uint32_t mi = 0; mi < d.num_meshes; ++mi) {
            const ab_mesh_desc& md = d.meshes[mi];
            Mesh m;
            auto P = [&](uint32_t i) {
                return V3d{double(md.positions[3 * i]), double(md.positions[3 * i + 1]),
                           double(md.positions[3 * i + 2])};
            };
            auto Pf = [&](uint32_t i) {
                return V3f{md.positions[3 * i], md.positions[3 * i + 1], md.positions[3 * i + 2]};
            };
            // Area-weighted vertex normals, reference triangle_mesh.cpp:92-125.
            std::vector<V3d> sums(md.smooth ? md.num_vertices : 0);
            if (md.smooth) {
                for (uint32_t f = 0; f < md.num_faces; ++f) {
                    const uint32_t* ix = md.indices + 3 * f;
                    const V3d p0 = P(ix[0]), p1 = P(ix[1]), p2 = P(ix[2]);
                    if (!formsTriangle(p0, p1, p2)) continue;
                    const V3d an = cross(p1 - p0, p2 - p0);
                    for (int c = 0; c < 3; ++c) sums[ix[c]] = sums[ix[c]] + an;
                }
            }
            for (uint32_t f = 0; f < md.num_faces; ++f) {
                const uint32_t* ix = md.indices + 3 * f;
                const V3d p0 = P(ix[0]), p1 = P(ix[1]), p2 = P(ix[2]);
                if (!formsTriangle(p0, p1, p2)) continue;
                m.tri_f.push_back({Pf(ix[0]), Pf(ix[1]), Pf(ix[2])});
                m.tri_d.push_back({p0, p1, p2});
                const V3d fn = normalized(cross(p1 - p0, p2 - p0));
                m.face_normal.push_back(fn);
                if (md.smooth) {
                    std::array<V3d, 3> cn;
                    for (int c = 0; c < 3; ++c) {
                        const V3d s = sums[ix[c]];
                        cn[c] = std::isnormal(lengthSquared(s)) ? normalized(s) : fn;
                    }
                    m.corner_n.push_back(cn);
                }
                std::array<double, 6> uv = {0, 0, 1, 0, 1, 1};
                if (md.uv)
                    for (int c = 0; c < 6; ++c) uv[c] = md.uv[6 * f + c];
                m.uv.push_back(uv);
            }
            std::vector<Box> boxes(m.tri_d.size());
            for (size_t i = 0; i < boxes.size(); ++i)
                for (int c = 0; c < 3; ++c) boxes[i].grow(m.tri_d[i][c]);
            for (const Box& b : boxes) m.box.grow(b);
            m.blas.build(boxes);
            meshes.push_back(std::move(m));
        }
        std::vector<Box> wb;
        for (uint32_t oi = 0; oi < d.num_objects; ++oi) {
            const ab_object_desc& od = d.objects[oi];
            Object o{};
            o.kind = od.kind;
            o.material = od.material;
            o.light = od.light;
            o.mesh = od.mesh;
            o.center = {od.center[0], od.center[1], od.center[2]};
            o.radius = od.radius;
            for (int i = 0; i < 12; ++i) {
                o.to_world.m[i] = od.to_world[i];
                o.to_object.m[i] = od.to_object[i];
                o.to_object_f.m[i] = float(od.to_object[i]);
            }
            if (o.kind == AB_OBJ_SPHERE) {
                const V3d r{o.radius, o.radius, o.radius};
                o.world_box.lo = o.center - r;
                o.world_box.hi = o.center + r;
            } else {
                o.world_box = transformBox(o.to_world, meshes[o.mesh].box);
            }
            wb.push_back(o.world_box);
            objects.push_back(o);
        }
        tlas.build(wb);
        materials.assign(d.materials, d.materials + d.num_materials);
        lights.assign(d.lights, d.lights + d.num_lights);
        for (uint32_t i = 0; i < d.num_lights; ++i)
            if (d.lights[i].kind == AB_LIGHT_SKY) sky_light = int(i);
    }

    // ------------------------------------------------------------ queries

    static void consRange(double tmin, const Hit& best, double tmax, double& lo, double& hi)
    {
        lo = tmin - 1e-4 * (1 + std::abs(tmin));
        const double top = best.hit ? best.t : tmax;
        hi = std::isinf(top) ? top : top + std::abs(top) * (1.0 / 4096) + 1e-30;
    }

    // Closest hit. fp32 = the GPU's FP32 operation order; otherwise FP64.
    Hit closest(const ab_ray& r, bool fp32, bool brute) const
    {
        Hit best;
        const V3d od{r.org[0], r.org[1], r.org[2]}, dd{r.dir[0], r.dir[1], r.dir[2]};
        const V3f of{r.org[0], r.org[1], r.org[2]}, df{r.dir[0], r.dir[1], r.dir[2]};
        const double tmin = r.tmin, tmax = r.tmax;
        auto testObject = [&](uint32_t oi) {
            const Object& o = objects[oi];
            if (o.kind == AB_OBJ_SPHERE) {
                double t;
                if (sphereIntersect(o.center, o.radius, od, dd, tmin, tmax, t)) {
                    const double tt = fp32 ? double(float(t)) : t;
                    if (better(tt, oi, AB_NO_PRIM, best)) best = {true, tt, oi, AB_NO_PRIM, 0, 0};
                }
                return;
            }
            const Mesh& m = meshes[o.mesh];
            V3d lo_d, ld_d;
            V3f lo_f, ld_f;
            if (fp32) {
                lo_f = o.to_object_f.point(of);
                ld_f = o.to_object_f.vector(df);
                lo_d = cast<double>(lo_f);
                ld_d = cast<double>(ld_f);
            } else {
                lo_d = o.to_object.point(od);
                ld_d = o.to_object.vector(dd);
            }
            auto testTri = [&](uint32_t ti) {
                if (fp32) {
                    float t, b1, b2;
                    V3f v[3] = {m.tri_f[ti][0], m.tri_f[ti][1], m.tri_f[ti][2]};
                    if (triangleIntersect<float>(v, lo_f, ld_f, r.tmin, r.tmax, t, b1, b2) &&
                        better(t, oi, ti, best))
                        best = {true, t, oi, ti, b1, b2};
                } else {
                    double t, b1, b2;
                    V3d v[3] = {m.tri_d[ti][0], m.tri_d[ti][1], m.tri_d[ti][2]};
                    if (triangleIntersect<double>(v, lo_d, ld_d, tmin, tmax, t, b1, b2) &&
                        better(t, oi, ti, best))
                        best = {true, t, oi, ti, b1, b2};
                }
                return false;
            };
            if (brute) {
                for (uint32_t ti = 0; ti < m.tri_d.size(); ++ti) testTri(ti);
            } else {
                m.blas.traverse(lo_d, ld_d, [&](double& lo, double& hi) { consRange(tmin, best, tmax, lo, hi); },
                                testTri);
            }
        };
        if (brute) {
            for (uint32_t oi = 0; oi < objects.size(); ++oi) testObject(oi);
        } else {
            tlas.traverse(od, dd, [&](double& lo, double& hi) { consRange(tmin, best, tmax, lo, hi); },
                          [&](uint32_t oi) {
                              testObject(oi);
                              return false;
                          });
        }
        return best;
    }

    bool anyHit(const ab_ray& r, bool fp32, bool brute) const
    {
        Hit h = closest(r, fp32, brute);
        return h.hit;
    }

    // ------------------------------------------------------------ shading

    struct HitRecord {
        double t;
        V3d point, normal, shading;
        bool front;
        double u, v;
        uint32_t object;
    };

    std::optional<HitRecord> intersect(V3d o, V3d d) const
    {
        ab_ray r{};
        // The FP64 path tracer needs FP64 rays: query through a double ray.
        Hit best;
        const double tmin = 0, tmax = std::numeric_limits<double>::infinity();
        tlas.traverse(o, d, [&](double& lo, double& hi) { consRange(tmin, best, tmax, lo, hi); },
                      [&](uint32_t oi) {
                          const Object& ob = objects[oi];
                          if (ob.kind == AB_OBJ_SPHERE) {
                              double t;
                              if (sphereIntersect(ob.center, ob.radius, o, d, tmin, tmax, t) &&
                                  better(t, oi, AB_NO_PRIM, best))
                                  best = {true, t, oi, AB_NO_PRIM, 0, 0};
                              return false;
                          }
                          const Mesh& m = meshes[ob.mesh];
                          const V3d lo_ = ob.to_object.point(o), ld = ob.to_object.vector(d);
                          m.blas.traverse(lo_, ld,
                                          [&](double& lo, double& hi) { consRange(tmin, best, tmax, lo, hi); },
                                          [&](uint32_t ti) {
                                              double t, b1, b2;
                                              V3d v[3] = {m.tri_d[ti][0], m.tri_d[ti][1], m.tri_d[ti][2]};
                                              if (triangleIntersect<double>(v, lo_, ld, tmin, tmax, t, b1, b2) &&
                                                  better(t, oi, ti, best))
                                                  best = {true, t, oi, ti, b1, b2};
                                              return false;
                                          });
                          return false;
                      });
        (void)r;
        if (!best.hit) return std::nullopt;
        const Object& ob = objects[best.object];
        HitRecord h{};
        h.t = best.t;
        h.object = best.object;
        V3d n, sn;
        if (ob.kind == AB_OBJ_SPHERE) {
            h.point = o + d * best.t;
            n = normalized(h.point - ob.center);
            sn = n;
            // reference geometry/sphere.cpp:137-147
            h.u = (std::atan2(-n.z, n.x) + kPi) / (2 * kPi);
            h.v = std::acos(std::clamp(-n.y, -1.0, 1.0)) / kPi;
        } else {
            const Mesh& m = meshes[ob.mesh];
            const auto& tv = m.tri_d[best.prim];
            const double b1 = best.b1, b2 = best.b2, b0 = 1 - b1 - b2;
            // reference triangle.cpp:162-175 (b0 = u / det there; 1 - b1 - b2 here)
            const V3d p_obj = tv[0] + (tv[1] - tv[0]) * b1 + (tv[2] - tv[0]) * b2;
            const V3d fn = m.face_normal[best.prim];
            V3d sn_obj = fn;
            bool smooth = !m.corner_n.empty();
            if (smooth) {
                // reference mesh_geometry.cpp:53-64 with barycentricCoordinatesOf (triangle.cpp:183-196)
                const V3d e1 = tv[1] - tv[0], e2 = tv[2] - tv[0], off = p_obj - tv[0];
                const V3d an = cross(e1, e2);
                const double a2 = lengthSquared(an);
                const double c1 = dot(cross(off, e2), an) / a2, c2 = dot(cross(e1, off), an) / a2;
                const double c0 = 1 - c1 - c2;
                const auto& cn = m.corner_n[best.prim];
                const V3d bl = cn[0] * c0 + cn[1] * c1 + cn[2] * c2;
                sn_obj = (lengthSquared(bl) >= 1e-12) ? normalized(bl) : fn;
            }
            const auto& uv = m.uv[best.prim];
            h.u = b0 * uv[0] + b1 * uv[2] + b2 * uv[4];
            h.v = b0 * uv[1] + b1 * uv[3] + b2 * uv[5];
            // reference mesh_instance.cpp:32-60
            h.point = ob.to_world.point(p_obj);
            n = normalized(ob.to_object.transposeVector(fn));
            if (smooth) {
                const V3d s = ob.to_object.transposeVector(sn_obj);
                sn = std::isnormal(lengthSquared(s)) ? normalized(s) : n;
            } else {
                sn = n;
            }
        }
        h.front = dot(d, n) < 0;
        h.normal = h.front ? n : -n;
        h.shading = h.front ? sn : -sn;
        return h;
    }

    bool occluded(V3d o, V3d d, double tmax) const
    {
        bool found = false;
        const double tmin = 0;
        Hit none;
        tlas.traverse(o, d, [&](double& lo, double& hi) { consRange(tmin, none, tmax, lo, hi); },
                      [&](uint32_t oi) {
                          const Object& ob = objects[oi];
                          if (ob.kind == AB_OBJ_SPHERE) {
                              double t;
                              found = sphereIntersect(ob.center, ob.radius, o, d, tmin, tmax, t);
                              return found;
                          }
                          const Mesh& m = meshes[ob.mesh];
                          const V3d lo_ = ob.to_object.point(o), ld = ob.to_object.vector(d);
                          m.blas.traverse(lo_, ld,
                                          [&](double& lo, double& hi) { consRange(tmin, none, tmax, lo, hi); },
                                          [&](uint32_t ti) {
                                              double t, b1, b2;
                                              V3d v[3] = {m.tri_d[ti][0], m.tri_d[ti][1], m.tri_d[ti][2]};
                                              found = triangleIntersect<double>(v, lo_, ld, tmin, tmax, t, b1, b2);
                                              return found;
                                          });
                          return found;
                      });
        return found;
    }

    Rgb texture(const ab_texture_desc& t, const HitRecord& h) const
    {
        const Rgb a{t.a[0], t.a[1], t.a[2]}, b{t.b[0], t.b[1], t.b[2]};
        switch (t.kind) {
        case AB_TEX_CHECKER: {
            // reference texture/checker_texture.cpp:8-17
            const double f = t.frequency;
            const double s = std::floor(h.u * f) + std::floor(h.v * f);
            return std::abs(std::fmod(s, 2.0)) == 1 ? b : a;
        }
        case AB_TEX_PERLIN: {
            const double w = patternWeight(h.point, t.frequency, t.veins);
            return a * (1 - w) + b * w;
        }
        default:
            return a;
        }
    }

    Scattering scatteringAt(const HitRecord& h) const
    {
        const ab_material_desc* m = &materials[objects[h.object].material];
        if (m->kind == AB_MAT_MIX) {
            const double w = patternWeight(h.point, m->scale, m->veins);
            m = &materials[w < m->threshold ? m->first : m->second];
        }
        Lobe lobe;
        switch (m->kind) {
        case AB_MAT_CONDUCTOR:
            lobe.kind = LobeKind::Conductor;
            lobe.eta = {m->eta[0], m->eta[1], m->eta[2]};
            lobe.k = {m->k[0], m->k[1], m->k[2]};
            lobe.alpha = double(m->roughness) * double(m->roughness);
            break;
        case AB_MAT_DIELECTRIC:
            lobe.kind = LobeKind::Dielectric;
            lobe.eta_d = h.front ? double(m->ior) : 1 / double(m->ior);
            lobe.alpha = double(m->roughness) * double(m->roughness);
            break;
        default:
            lobe.kind = LobeKind::Lambert;
            lobe.albedo = texture(m->albedo, h);
            break;
        }
        return Scattering{lobe, Onb::from(h.shading), h.normal};
    }

    Rgb skyRadiance(V3d dir) const
    {
        if (sky_light < 0) return {};
        const ab_light_desc& l = lights[sky_light];
        const double ht = 0.5 * (dir.y + 1);
        return Rgb{l.nadir[0], l.nadir[1], l.nadir[2]} * (1 - ht) + Rgb{l.zenith[0], l.zenith[1], l.zenith[2]} * ht;
    }

    std::optional<LightSample> sampleLight(const ab_light_desc& l, V3d p, double ux, double uy) const
    {
        if (l.kind == AB_LIGHT_POINT) {
            const V3d to = V3d{l.position[0], l.position[1], l.position[2]} - p;
            const double d2 = lengthSquared(to);
            if (d2 == 0) return std::nullopt;
            const double dd = std::sqrt(d2);
            return LightSample{to / dd, Rgb{l.intensity[0], l.intensity[1], l.intensity[2]} / d2, dd, 1, true};
        }
        if (l.kind == AB_LIGHT_SKY) {
            const V3d dir = uniformSphere(ux, uy);
            return LightSample{dir, skyRadiance(dir), std::numeric_limits<double>::infinity(),
                               uniformSpherePdf(), false};
        }
        const Object& ob = objects[l.object];
        const auto cone = coneTowards(ob.center, ob.radius, p);
        if (!cone) return std::nullopt;
        const double omc = ux * cone->one_minus_cos_max;
        const double ct = 1 - omc;
        const double s2 = omc * (2 - omc);
        const double st = std::sqrt(s2);
        const double phi = 2 * kPi * uy;
        const Onb basis = Onb::from(cone->axis);
        const V3d dir = basis.toWorld({st * std::cos(phi), st * std::sin(phi), ct});
        const double r2 = ob.radius * ob.radius;
        const double root = std::sqrt(std::max(0.0, r2 - cone->distance * cone->distance * s2));
        return LightSample{dir, Rgb{l.radiance[0], l.radiance[1], l.radiance[2]},
                           cone->tangent2 / (cone->distance * ct + root),
                           1 / (2 * kPi * cone->one_minus_cos_max), false};
    }

    double lightPdf(const ab_light_desc& l, V3d p, V3d dir) const
    {
        if (l.kind == AB_LIGHT_POINT) return 0;
        if (l.kind == AB_LIGHT_SKY) return uniformSpherePdf();
        const Object& ob = objects[l.object];
        const auto cone = coneTowards(ob.center, ob.radius, p);
        if (!cone) return 0;
        if (dot(dir, cone->axis) < cone->cos_max) return 0;
        return 1 / (2 * kPi * cone->one_minus_cos_max);
    }

    Rgb emitted(const HitRecord& h) const
    {
        const int li = objects[h.object].light;
        if (li < 0 || !h.front) return {};
        const ab_light_desc& l = lights[li];
        if (l.kind != AB_LIGHT_AREA) return {};
        return {l.radiance[0], l.radiance[1], l.radiance[2]};
    }
};

constexpr double kSurfaceOffset = 1e-7; // reference estimator/surface_ray.hpp:18
constexpr double kShadowMargin = 1e-6;  // reference estimator/path_estimator.cpp:24

inline double powerHeuristic(double chosen, double other)
{
    const double c2 = chosen * chosen;
    const double tot = c2 + other * other;
    return tot > 0 ? c2 / tot : 0;
}

struct RayCounts {
    uint64_t camera{0}, extension{0}, shadow{0};
};

// reference estimator/path_estimator.cpp:153-185
inline Rgb radiance(const Scene& s, V3d ro, V3d rd, Generator& g, int max_depth, int rr_start,
                    RayCounts& rc)
{
    Rgb thr{1, 1, 1}, L{};
    V3d arr_origin = ro;
    double arr_pdf = 0;
    bool arr_spec = true;
    const double nl = double(s.lights.size());
    for (int depth = 0;; ++depth) {
        if (depth == 0) ++rc.camera;
        else ++rc.extension;
        const auto hit = s.intersect(ro, rd);
        if (!hit) {
            Rgb e = s.skyRadiance(rd);
            if (!isBlack(e) && !arr_spec) {
                const double lp = (1 / nl) * uniformSpherePdf();
                e = e * powerHeuristic(arr_pdf, lp);
            }
            return L + thr * e;
        }
        {
            Rgb e = s.emitted(*hit);
            if (!isBlack(e) && !arr_spec) {
                const int li = s.objects[hit->object].light;
                const double lp = (1 / nl) * s.lightPdf(s.lights[li], arr_origin, rd);
                e = e * powerHeuristic(arr_pdf, lp);
            }
            L = L + thr * e;
        }
        if (depth >= max_depth) return L;
        const Scattering sc = s.scatteringAt(*hit);
        const V3d wo = -rd;
        // Next-event estimation, reference path_estimator.cpp:73-102.
        {
            const double ul = g.next1D();
            double ux, uy;
            g.next2D(ux, uy);
            if (!s.lights.empty()) {
                const size_t idx = std::min(size_t(ul * nl), s.lights.size() - 1);
                const auto ls = s.sampleLight(s.lights[idx], hit->point, ux, uy);
                if (ls && !isBlack(ls->radiance)) {
                    const Rgb f = sc.evaluate(wo, ls->direction);
                    double smax;
                    if (std::isinf(ls->distance)) smax = ls->distance;
                    else smax = ls->distance - std::max(ls->distance * kShadowMargin, 2 * kSurfaceOffset);
                    bool blocked = false;
                    if (!isBlack(f) && smax > 0) {
                        const double side = dot(ls->direction, hit->normal) >= 0 ? 1.0 : -1.0;
                        ++rc.shadow;
                        blocked = s.occluded(hit->point + hit->normal * (side * kSurfaceOffset), ls->direction, smax);
                    }
                    if (!isBlack(f) && !blocked) {
                        const double lpdf = (1 / nl) * ls->pdf;
                        const double w = ls->delta ? 1.0 : powerHeuristic(lpdf, sc.pdf(wo, ls->direction));
                        const double ct = std::abs(dot(ls->direction, hit->shading));
                        L = L + thr * (f * ls->radiance * (ct * w / lpdf));
                    }
                }
            }
        }
        const double uc = g.next1D();
        double ux, uy;
        g.next2D(ux, uy);
        const auto smp = sc.sample(wo, uc, ux, uy);
        if (!smp) return L;
        thr = thr * smp->weight;
        if (isBlack(thr)) return L;
        if (depth >= rr_start) {
            const double surv = std::min(0.95, std::max({thr.r, thr.g, thr.b}));
            if (g.next1D() >= surv) return L;
            thr = thr / surv;
        }
        arr_origin = hit->point;
        arr_pdf = smp->pdf;
        arr_spec = smp->specular;
        const double side = dot(smp->direction, hit->normal) >= 0 ? 1.0 : -1.0;
        ro = hit->point + hit->normal * (side * kSurfaceOffset);
        rd = smp->direction;
    }
}

// reference estimator/scatter_only_estimator.cpp:10-48
inline Rgb scatterOnly(const Scene& s, V3d ro, V3d rd, Generator& g, int max_count, RayCounts& rc)
{
    Rgb thr{1, 1, 1}, L{};
    for (int count = 0;; ++count) {
        if (count == 0) ++rc.camera;
        else ++rc.extension;
        const auto hit = s.intersect(ro, rd);
        if (!hit) return L + thr * s.skyRadiance(rd);
        L = L + thr * s.emitted(*hit);
        if (count >= max_count) return L;
        const double uc = g.next1D();
        double ux, uy;
        g.next2D(ux, uy);
        const auto smp = s.scatteringAt(*hit).sample(-rd, uc, ux, uy);
        if (!smp) return L;
        thr = thr * smp->weight;
        if (isBlack(thr)) return L;
        const double side = dot(smp->direction, hit->normal) >= 0 ? 1.0 : -1.0;
        ro = hit->point + hit->normal * (side * kSurfaceOffset);
        rd = smp->direction;
    }
}

// reference estimator/normals_estimator.cpp:15-40
inline Rgb normalsView(const Scene& s, V3d ro, V3d rd, RayCounts& rc)
{
    ++rc.camera;
    const auto hit = s.intersect(ro, rd);
    if (hit) return {0.5 * (hit->shading.x + 1), 0.5 * (hit->shading.y + 1), 0.5 * (hit->shading.z + 1)};
    const double h = 0.5 * (normalized(rd).y + 1);
    return Rgb{1, 1, 1} * (1 - h) + Rgb{0.5, 0.7, 1.0} * h;
}

// The estimator an ab_render_params selects (reference radiance_estimator.hpp:36-38).
inline Rgb estimate(const Scene& s, const ab_render_params& p, V3d ro, V3d rd, Generator& g, RayCounts& rc)
{
    if (p.estimator == AB_ESTIMATOR_NORMALS) return normalsView(s, ro, rd, rc);
    if (p.estimator == AB_ESTIMATOR_SCATTER_ONLY) return scatterOnly(s, ro, rd, g, int(p.max_depth), rc);
    return radiance(s, ro, rd, g, int(p.max_depth), int(p.roulette_start), rc);
}

struct Camera {
    V3d origin, forward, half_w, half_h;
    explicit Camera(const ab_camera& c)
    {
        const V3d from{c.look_from[0], c.look_from[1], c.look_from[2]};
        const V3d at{c.look_at[0], c.look_at[1], c.look_at[2]};
        const V3d up{c.up[0], c.up[1], c.up[2]};
        const V3d w = normalized(from - at);
        const V3d u = normalized(cross(normalized(up), w));
        const V3d v = cross(w, u);
        const double hh = std::tan(c.vfov_degrees * kPi / 360);
        const double hw = c.aspect * hh;
        origin = from;
        forward = -w;
        half_w = u * hw;
        half_h = v * hh;
    }
    V3d direction(double fx, double fy) const
    {
        return normalized(forward + half_w * (2 * fx - 1) + half_h * (1 - 2 * fy));
    }
};

} // namespace oracle

using namespace oracle;

namespace {

template <class F>
static void parallelFor(uint64_t n, int threads, F f)
{
    if (threads <= 1 || n < 1024) {
        for (uint64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const uint64_t b = next.fetch_add(256);
                if (b >= n) return;
                const uint64_t e = std::min(n, b + 256);
                for (uint64_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto& th : pool) th.join();
}

static Lobe makeLobe(int kind, const double* p)
{
    Lobe l;
    if (kind == 0) {
        l.kind = LobeKind::Lambert;
        l.albedo = {p[0], p[1], p[2]};
    } else if (kind == 1) {
        l.kind = LobeKind::Conductor;
        l.eta = {p[0], p[1], p[2]};
        l.k = {p[3], p[4], p[5]};
        l.alpha = p[6];
    } else {
        l.kind = LobeKind::Dielectric;
        l.eta_d = p[0];
        l.alpha = p[1];
    }
    return l;
}
} // namespace

extern "C" {

struct or_scene {
    std::unique_ptr<Scene> s;
};

or_scene* or_scene_create(const ab_scene_desc* d)
{
    try {
        return new or_scene{std::make_unique<Scene>(*d)};
    } catch (...) {
        return nullptr;
    }
}
void or_scene_destroy(or_scene* s) { delete s; }
uint32_t or_scene_tlas_depth(const or_scene* s) { return s->s->tlas.max_depth; }

// mode 0 = FP64 reference arithmetic, 1 = FP32 mirror of the GPU.
void or_trace_closest(const or_scene* s, const ab_ray* rays, uint64_t n, int mode, int brute,
                      int threads, double* t, uint32_t* object, uint32_t* prim, double* b1, double* b2)
{
    parallelFor(n, threads, [&](uint64_t i) {
        const Hit h = s->s->closest(rays[i], mode == 1, brute != 0);
        t[i] = h.t;
        object[i] = h.object;
        prim[i] = h.prim;
        if (b1) b1[i] = h.b1;
        if (b2) b2[i] = h.b2;
    });
}

// FP64 closest hits of n rays, counting the ordered traversal's work: box tests
// (TLAS and BLAS nodes) and primitive visits (instances, spheres, triangles).
void or_trace_counts(const or_scene* s, const ab_ray* rays, uint64_t n, int threads, uint64_t* nodes,
                     uint64_t* prims)
{
    std::vector<uint64_t> cn(n), cp(n);
    parallelFor(n, threads, [&](uint64_t i) {
        tl_counts = TraversalCounts{};
        (void)s->s->closest(rays[i], false, false);
        cn[i] = tl_counts.nodes;
        cp[i] = tl_counts.prims;
    });
    uint64_t a = 0, b = 0;
    for (uint64_t i = 0; i < n; ++i) {
        a += cn[i];
        b += cp[i];
    }
    *nodes = a;
    *prims = b;
}

void or_trace_any(const or_scene* s, const ab_ray* rays, uint64_t n, int mode, int brute, int threads,
                  uint8_t* occluded)
{
    parallelFor(n, threads, [&](uint64_t i) { occluded[i] = s->s->anyHit(rays[i], mode == 1, brute != 0); });
}

// Renders with the FP64 estimator. sample_mode 0: the reference's draw order (one
// stream per pixel, consumed sample after sample); 1: the GPU's (each sample at
// its own offset in the pixel's stream, FP32-exact draws). rgb_mean: w*h*3.
// Rows [row0, row1) only, so tests can render a band.
void or_render(const or_scene* s, const ab_camera* cam, const ab_render_params* p, int sample_mode,
               int threads, uint32_t row0, uint32_t row1, double* rgb_mean, uint64_t* counts)
{
    const Camera camera(*cam);
    const uint32_t W = p->width, H = p->height, n = p->samples_per_axis;
    row1 = std::min(row1, H);
    const uint64_t npx = uint64_t(row1 - row0) * W;
    std::vector<RayCounts> per(npx);
    parallelFor(npx, threads, [&](uint64_t i) {
        const uint32_t x = uint32_t(i % W), y = row0 + uint32_t(i / W);
        Generator g = pixelGenerator(p->seed, x, y, W);
        Rgb sum{};
        RayCounts& rc = per[i];
        for (uint32_t sy = 0; sy < n; ++sy)
            for (uint32_t sx = 0; sx < n; ++sx) {
                if (sample_mode == 1) g = sampleGenerator(p->seed, x, y, W, sy * n + sx);
                double jx, jy;
                g.next2D(jx, jy);
                const double fx = (x + (sx + jx) / n) / double(W);
                const double fy = (y + (sy + jy) / n) / double(H);
                const V3d d = camera.direction(fx, fy);
                sum = sum + estimate(*s->s, *p, camera.origin, d, g, rc);
            }
        const Rgb mean = sum / double(n * n);
        const uint64_t o = (uint64_t(y) * W + x) * 3;
        rgb_mean[o] = mean.r;
        rgb_mean[o + 1] = mean.g;
        rgb_mean[o + 2] = mean.b;
    });
    if (counts) {
        counts[0] = counts[1] = counts[2] = 0;
        for (const auto& r : per) {
            counts[0] += r.camera;
            counts[1] += r.extension;
            counts[2] += r.shadow;
        }
    }
}

// One radiance estimate along a given ray, drawing from the reference stream of
// pixel (0,0) of a 1-pixel image under `seed` (for analytic tests), averaged
// over `count` estimates from that one generator.
void or_radiance(const or_scene* s, const double* o, const double* d, uint64_t seed, int max_depth,
                 int rr_start, uint64_t count, double* rgb)
{
    Generator g = pixelGenerator(seed, 0, 0, 1);
    RayCounts rc;
    Rgb sum{};
    for (uint64_t i = 0; i < count; ++i)
        sum = sum + radiance(*s->s, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, g, max_depth, rr_start, rc);
    rgb[0] = sum.r / double(count);
    rgb[1] = sum.g / double(count);
    rgb[2] = sum.b / double(count);
}

// ------------------------------------------------------------ unit helpers

void or_pcg32(uint64_t seed, uint64_t stream, uint64_t advance, uint32_t n, uint32_t* out)
{
    Pcg32 g(seed, stream);
    g.advance(advance);
    for (uint32_t i = 0; i < n; ++i) out[i] = g.next();
}
uint64_t or_mixbits(uint64_t v) { return mixBits(v); }
double or_perlin(double x, double y, double z) { return perlinNoise({x, y, z}); }
double or_turbulence(double x, double y, double z) { return turbulence({x, y, z}); }
double or_pattern(double x, double y, double z, double scale, double veins)
{
    return patternWeight({x, y, z}, scale, veins);
}
double or_fresnel_dielectric(double c, double eta) { return fresnelDielectric(c, eta); }
void or_fresnel_conductor(double c, const double* eta, const double* k, double* out)
{
    const Rgb r = fresnelConductor(c, {eta[0], eta[1], eta[2]}, {k[0], k[1], k[2]});
    out[0] = r.r;
    out[1] = r.g;
    out[2] = r.b;
}
int or_sphere_t(const double* c, double r, const double* o, const double* d, double tmin, double tmax,
                double* t)
{
    return sphereIntersect({c[0], c[1], c[2]}, r, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, tmin, tmax, *t);
}
int or_triangle_f(const float* v, const float* o, const float* d, float tmin, float tmax, float* out)
{
    const V3f vv[3] = {{v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]}};
    return triangleIntersect<float>(vv, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, tmin, tmax, out[0], out[1], out[2]);
}
int or_triangle_d(const double* v, const double* o, const double* d, double tmin, double tmax, double* out)
{
    const V3d vv[3] = {{v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]}};
    return triangleIntersect<double>(vv, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, tmin, tmax, out[0], out[1], out[2]);
}
// Lobe helpers: kind 0 Lambert(albedo=p[0..2]), 1 conductor(eta p[0..2], k p[3..5], alpha p[6]),
// 2 dielectric(eta p[0], alpha p[1]). Local-frame directions.
void or_lobe_eval(int kind, const double* p, const double* wo, const double* wi, double* f, double* pdf)
{
    const Lobe l = makeLobe(kind, p);
    const Rgb r = l.evaluate({wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
    f[0] = r.r;
    f[1] = r.g;
    f[2] = r.b;
    *pdf = l.pdf({wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]});
}
int or_lobe_sample(int kind, const double* p, const double* wo, double uc, double ux, double uy, double* wi,
                   double* weight, double* pdf, int* specular)
{
    const Lobe l = makeLobe(kind, p);
    const auto s = l.sample({wo[0], wo[1], wo[2]}, uc, ux, uy);
    if (!s) return 0;
    wi[0] = s->direction.x;
    wi[1] = s->direction.y;
    wi[2] = s->direction.z;
    weight[0] = s->weight.r;
    weight[1] = s->weight.g;
    weight[2] = s->weight.b;
    *pdf = s->pdf;
    *specular = s->specular;
    return 1;
}
double or_ggx_density(double alpha, const double* m) { return Ggx{alpha}.density({m[0], m[1], m[2]}); }
double or_ggx_masking(double alpha, const double* w) { return Ggx{alpha}.masking({w[0], w[1], w[2]}); }

// Array forms for the distribution tests, which compare Monte Carlo means of
// samples with midpoint-rule quadrature of the densities over 10^5-10^6 points
// (reference tests/scattering/ggx_distribution_test.cpp:154-188,
// conductor_lobe_test.cpp:121-161, dielectric_lobe_test.cpp:199-237).
// u: n x 2 (or n x 3 with the choice draw first); vectors n x 3.
void or_ggx_vn_sample_many(double alpha, const double* wo, uint64_t n, const double* u, double* m)
{
    const Ggx g{alpha};
    const V3d w{wo[0], wo[1], wo[2]};
    for (uint64_t i = 0; i < n; ++i) {
        const V3d s = g.sampleVisibleNormal(w, u[2 * i], u[2 * i + 1]);
        m[3 * i] = s.x;
        m[3 * i + 1] = s.y;
        m[3 * i + 2] = s.z;
    }
}
void or_ggx_vn_pdf_many(double alpha, const double* wo, uint64_t n, const double* m, double* pdf)
{
    const Ggx g{alpha};
    const V3d w{wo[0], wo[1], wo[2]};
    for (uint64_t i = 0; i < n; ++i) pdf[i] = g.visibleNormalPdf(w, {m[3 * i], m[3 * i + 1], m[3 * i + 2]});
}
void or_lobe_eval_many(int kind, const double* p, const double* wo, uint64_t n, const double* wi, double* f,
                       double* pdf)
{
    const Lobe l = makeLobe(kind, p);
    const V3d w{wo[0], wo[1], wo[2]};
    for (uint64_t i = 0; i < n; ++i) {
        const V3d d{wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]};
        const Rgb r = l.evaluate(w, d);
        f[3 * i] = r.r;
        f[3 * i + 1] = r.g;
        f[3 * i + 2] = r.b;
        pdf[i] = l.pdf(w, d);
    }
}
// Returns the number of samples kept; kept[i] = 0 where sample() gave nothing.
uint64_t or_lobe_sample_many(int kind, const double* p, const double* wo, uint64_t n, const double* u, double* wi,
                             double* weight, double* pdf, uint8_t* kept)
{
    const Lobe l = makeLobe(kind, p);
    const V3d w{wo[0], wo[1], wo[2]};
    uint64_t k = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const auto s = l.sample(w, u[3 * i], u[3 * i + 1], u[3 * i + 2]);
        kept[i] = s ? 1 : 0;
        if (!s) {
            wi[3 * i] = wi[3 * i + 1] = wi[3 * i + 2] = 0;
            weight[3 * i] = weight[3 * i + 1] = weight[3 * i + 2] = 0;
            pdf[i] = 0;
            continue;
        }
        ++k;
        wi[3 * i] = s->direction.x;
        wi[3 * i + 1] = s->direction.y;
        wi[3 * i + 2] = s->direction.z;
        weight[3 * i] = s->weight.r;
        weight[3 * i + 1] = s->weight.g;
        weight[3 * i + 2] = s->weight.b;
        pdf[i] = s->pdf;
    }
    return k;
}

// or_radiance with the estimator chosen as in ab_render_params (AB_ESTIMATOR_*;
// for scatter-only, max_depth is the scatter limit), the mean of `count`
// estimates drawn one after another from the stream of pixel (0, 0) of a
// one-pixel image (the reference's test generator, path_estimator_test.cpp:31-34).
void or_radiance_est(const or_scene* s, const double* o, const double* d, uint64_t seed, uint32_t estimator,
                     int max_depth, int rr_start, uint64_t count, double* rgb)
{
    Generator g = pixelGenerator(seed, 0, 0, 1);
    ab_render_params p{};
    p.max_depth = uint32_t(max_depth);
    p.roulette_start = uint32_t(rr_start);
    p.estimator = estimator;
    RayCounts rc;
    Rgb sum{};
    for (uint64_t i = 0; i < count; ++i) sum = sum + estimate(*s->s, p, {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, g, rc);
    rgb[0] = sum.r / double(count);
    rgb[1] = sum.g / double(count);
    rgb[2] = sum.b / double(count);
}

} // extern "C"
