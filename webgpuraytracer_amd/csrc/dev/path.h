// path.h — one bounce of shader.wgsl `main` (2319-2480) after its closest hit
// is known, shared by the megakernel and the wavefront shade kernel.
//
// Next-event estimation hands each shadow query to a sink: the megakernel
// traces it at once, the wavefront tracer queues it for the shadow kernel.
// Either way the contribution is added to the path radiance in the
// reference's order (light NEE, then env NEE, then the emission of a surface
// the height field blocks the sampled direction at, before the next bounce),
// so both backends produce the same bits.
#pragma once
#include "shade.h"

namespace pt {

struct PathState {
    v3 ro, rd, tp, rad;
    float last_pdf, last_wsum;
    Pcg rng;
};

__device__ __forceinline__ v3 camera_dir(const pt_scene_params& P, float su, float sv) {
    const v3 r00 = ld3(P.ray00), r10 = ld3(P.ray10), r01 = ld3(P.ray01), r11 = ld3(P.ray11);
    return normalize(lerp(lerp(r00, r10, su), lerp(r01, r11, su), 1.0f - sv));
}

// Seeding, warm-up draw, jittered camera ray (shader.wgsl:2280-2287).
__device__ __forceinline__ void path_begin(const pt_scene_params& PP, uint32_t px, uint32_t py, uint32_t sample_number,
                                           uint32_t seed, PathState& ps) {
    const uint32_t idx = py * PP.width + px;
    ps.rng.s = (idx * 1973u + sample_number * 9277u + seed * 26699u) | 1u;
    ps.rng.next();
    const float jx = ps.rng.next();
    const float jy = ps.rng.next();
    ps.rd = camera_dir(PP, (float(px) + jx) / float(PP.width), (float(py) + jy) / float(PP.height));
    ps.ro = ld3(PP.eye);
    ps.tp = splat(1.0f);
    ps.rad = splat(0.0f);
    ps.last_pdf = 1.0f;
    ps.last_wsum = 0.0f;
}

// Everything after trace_scene in one bounce.  Returns false when the path
// ends (miss, bright emitter, failed BSDF sample, Russian roulette).
template <uint32_t F, bool CN, class Sink>
__device__ bool shade_bounce(const DevScene& S, Flags<F> fl, const Params& P, int bounce, const HitRec& h,
                             const SurfHit& sh, PathState& ps, Media& md, Counters& c, Sink& sink) {
    // Beer-Lambert attenuation and medium emission (shader.wgsl:2325-2342)
    if (md.n > 0) {
        const Medium cm = md.top();
        if (cm.ior > 1.0f || len(cm.sigma) > 0.0f) {
            const float dist = sh.m_idx == -1 ? 1000.0f : h.t;
            const v3 att = mk(expf(-cm.sigma.x * dist), expf(-cm.sigma.y * dist), expf(-cm.sigma.z * dist));
            const float ex = cm.sigma.x > 1e-4f ? cm.emission.x * (1.0f - att.x) / cm.sigma.x : cm.emission.x * dist;
            const float ey = cm.sigma.y > 1e-4f ? cm.emission.y * (1.0f - att.y) / cm.sigma.y : cm.emission.y * dist;
            const float ez = cm.sigma.z > 1e-4f ? cm.emission.z * (1.0f - att.z) / cm.sigma.z : cm.emission.z * dist;
            ps.rad = ps.rad + ps.tp * firefly(mk(ex, ey, ez), bounce);
            ps.tp = ps.tp * att;
        }
    }
    if (sh.m_idx == -1) {  // miss (shader.wgsl:2344-2358)
        if (fl.has(PT_HAS_SKYBOX)) {
            const v3 sky = sample_sky(S, ps.rd);
            float w = 1.0f;
            if (bounce > 0 && fl.has(PT_MIS_SKYBOX)) w = mis_weight(ps.last_pdf, sky_pdf(S, P, ps.rd));
            ps.rad = ps.rad + ps.tp * firefly(sky * w, bounce);
        } else {
            ps.rad = ps.rad + ps.tp * firefly(mk(0.02f, 0.03f, 0.05f), bounce);
        }
        return false;
    }
    if (CN) c.shade++;
    const pt_material& m = S.mats[sh.m_idx];
    Surface sc = base_surface(m, sh.n);
    // texture coordinates, moved by parallax mapping (shader.wgsl:2361-2373)
    v2 fuv{0.0f, 0.0f};
    float fdepth = 0.0f;
    bool pom = false;
    if (fl.has(kTexturesBit)) {
        fuv = v2{sh.uv.x * m.uv_scale[0], sh.uv.y * m.uv_scale[1]};
        pom = fl.has(PT_HAS_HEIGHTMAPS) && m.height_idx >= 0;
        if (pom) {
            const v3 vts = normalize(to_tangent(sh.tg, sh.bt, sh.n, -ps.rd));
            if (vts.z > 0.0f) {
                const PomRes pr = pom_march(S, m, fuv, vts);
                fuv = pr.uv;
                fdepth = pr.h;
            }
        }
        apply_textures(S, m, fuv, sh.tg, sh.bt, sc);
    }
    // implicit light hit (shader.wgsl:2377-2395)
    if (len(sc.emittance) > 0.0f) {
        float w = 1.0f;
        if (fl.has(PT_HAS_LIGHTS) && bounce > 0 && sh.o_idx >= 0) {
            const pt_object& ob = S.objs[sh.o_idx];
            if (ob.light_idx >= 0) {
                const pt_light& L = S.lights[ob.light_idx];
                float raw = 0.0f;
                if (fl.has(PT_HAS_SPHERES) && ob.object_type == PT_OBJ_SPHERE) raw = sphere_light_pdf(L, ob, ps.ro, ps.rd);
                else if (fl.has(PT_HAS_CUBES) && ob.object_type == PT_OBJ_CUBE) raw = box_light_pdf(L, ps.ro, ps.rd, h.t, sh.n);
                const v3 tl = ld3(ob.world_position) - ps.ro;
                const float d2 = mx(dot(tl, tl), 0.001f);
                const float imp = L.power / d2;
                const float sel = ps.last_wsum > 0.0f ? imp / ps.last_wsum : 0.0f;
                w = mis_weight(ps.last_pdf, raw * sel);
            }
        }
        ps.rad = ps.rad + ps.tp * firefly(sc.emittance * w, bounce);
        if (len(sc.emittance) > 1.0f) return false;
    }
    const v3 hp = ps.ro + ps.rd * h.t;
    const v3 V = -ps.rd;
    const v3 no = dot(sc.normal, V) > 0.0f ? sc.normal : -sc.normal;
    // evaluate_nee (shader.wgsl:2226-2271) up to the shadow ray: inc is the
    // unoccluded incoming radiance; need_ray says whether a shadow ray decides it
    auto nee_eval = [&](const LightSample& ls, int ignore, int k) {
        v3 inc = splat(0.0f);
        bool need = false;
        v3 o = hp;
        if (ls.pdf > 0.0f) {
            const float nl = dot(no, ls.dir);
            if (!(nl <= 0.0f && m.transmission <= 0.0f)) {
                const float bp = pdf_bsdf(V, ls.dir, m, sc, md);
                const float w = mis_weight(ls.pdf, bp);
                const v3 f = eval_bsdf(V, ls.dir, m, sc, md);
                const v3 cw = w * f / mx(ls.pdf, 1e-6f);
                // a light the height map hides gets no shadow ray (shader.wgsl:2254-2260)
                if (!(hmax(ls.color * cw) < 1e-5f) &&
                    !(pom && pom_shadow(S, m, fuv, fdepth, normalize(to_tangent(sh.tg, sh.bt, sh.n, ls.dir))).hit)) {
                    o = hp + (nl < 0.0f ? -no * 0.001f : no * 0.001f);
                    inc = ls.color * cw;
                    need = true;
                }
            }
        }
        sink(k, need, o, ls.dir, ls.dist, ignore, inc, ps.tp, bounce, ps.rad);
    };
    // light NEE with one-pass reservoir selection (shader.wgsl:2402-2446)
    if (fl.has(PT_HAS_LIGHTS)) {
        int sel = -1;
        float wsum = 0.0f, sw = 0.0f;
        for (uint32_t i = 0; i < S.n_lights; ++i) {
            const pt_light& L = S.lights[i];
            if (L.obj_idx < 0) continue;
            const v3 tl = ld3(S.objs[L.obj_idx].world_position) - hp;
            const float d2 = mx(dot(tl, tl), 0.001f);
            const float w = L.power / d2;
            wsum += w;
            if (ps.rng.next() < (w / wsum)) { sel = int(i); sw = w; }
        }
        if (sel >= 0 && wsum > 0.0f) {
            const pt_light& L = S.lights[sel];
            const pt_object& lo = S.objs[L.obj_idx];
            LightSample ls;
            ls.dir = splat(0.0f); ls.color = splat(0.0f); ls.dist = 0.0f; ls.pdf = 0.0f;
            if (fl.has(PT_HAS_SPHERES) && lo.object_type == PT_OBJ_SPHERE) {
                const float x0 = ps.rng.next();
                const float x1 = ps.rng.next();
                ls = sphere_light_sample(S, L, lo, hp, x0, x1);
            } else if (fl.has(PT_HAS_CUBES) && lo.object_type == PT_OBJ_CUBE) {
                const float x0 = ps.rng.next();
                const float x1 = ps.rng.next();
                ls = box_light_sample(S, L, lo, hp, x0, x1, ps.rng);
            }
            if (ls.pdf > 0.0f) {
                ls.pdf *= sw / wsum;
                nee_eval(ls, L.obj_idx, 0);
            }
        }
        ps.last_wsum = wsum;
    }
    // env NEE (shader.wgsl:2449-2453)
    if (fl.has(PT_HAS_SKYBOX) && fl.has(PT_MIS_SKYBOX)) {
        const float ux = ps.rng.next();
        const float uy = ps.rng.next();
        nee_eval(sample_env(S, P, ux, uy), -1, 1);
    }
    if (!sample_bsdf(ps.ro, ps.rd, ps.tp, m, sc, hp, md, ps.last_pdf, ps.rng)) return false;
    // a sampled direction the height map blocks: the surface where it is
    // blocked adds its emission and is sampled again (shader.wgsl:2458-2472)
    if (pom) {
        const PomRes sr = pom_shadow(S, m, fuv, fdepth, normalize(to_tangent(sh.tg, sh.bt, sh.n, ps.rd)));
        if (sr.hit) {
            Surface s2 = base_surface(m, sh.n);
            apply_textures(S, m, sr.uv, sh.tg, sh.bt, s2);
            sink.emit(ps.tp * firefly(s2.emittance, bounce), ps.rad);  // after the NEE contributions
            if (len(s2.emittance) > 1.0f) return false;
            if (!sample_bsdf(ps.ro, ps.rd, ps.tp, m, s2, hp, md, ps.last_pdf, ps.rng)) return false;
        }
    }
    // Russian roulette from the third bounce (shader.wgsl:2474-2479)
    if (bounce < 2) return true;
    const float surv = clampf(hmax(ps.tp), 0.05f, 0.95f);
    if (ps.rng.next() > surv) return false;
    ps.tp = ps.tp / surv;
    return true;
}

}  // namespace pt
