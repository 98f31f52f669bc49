// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle_core.h).
//
// FP64 evaluation of the built-in fragment programs; see oracle_shade.h for
// the contract.  The sampler follows the CUDA spec in
// include/ifrit/softgraphics/engine/math/ShaderBuiltinCuda.cuh with three
// documented corrections: texel centres at (i + 0.5) / size instead of
// uv * (size - 1) (SBC:112-113), the footprint of the x and y derivatives
// measured as |d(uv)/dx| and |d(uv)/dy| (SBC:236-241, 248-250 mix the axes),
// and anisotropic taps placed symmetrically at (i + 0.5) / N - 0.5 instead of
// i / (N + 1) - 0.5 (SBC:259-260), which shifts even a single tap by half a
// footprint.
#include "oracle_shade.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

// Planted mistake (0 = none), set only by the closed-form tests through
// oracle_set_plant to show that each of them fails on a plausible error:
//   1 Phong reflection vector instead of the half vector,
//   2 point-light attenuation 1 / (1 + k d) instead of 1 / (1 + k d^2),
//   3 view vector flipped (P - camera),
//   4 resolve divides by the covered samples instead of S,
//   5 resolve ignores the sample counts (each distinct triangle once).
int g_plant = 0;

struct D3 {
    double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 normalize(D3 a) {
    const double l = std::sqrt(dot(a, a));
    return l > 0 ? a * (1.0 / l) : a;
}
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Layered GGX (the build's config-5 reference material): Trowbridge-Reitz
// distribution with alpha = roughness, Smith-Schlick geometry with
// k = alpha / 2, Schlick Fresnel with F0 = 0.04, Lambert diffuse weighted by
// (1 - F); roughness = 0.05 + 0.95 * texture red.
const double kPi = 3.14159265358979323846;
D3 ggx_light(D3 n, D3 V, D3 L, D3 alb, double alpha, double kd, double ks, bool* flag) {
    const double ndl = dot(n, L), ndv = dot(n, V);
    if (std::fabs(ndl) < 1e-4 || std::fabs(ndv) < 1e-4) *flag = true;
    if (!(ndl > 0 && ndv > 0)) return {0, 0, 0};
    const D3 H = normalize(L + V);
    const double ndh = std::max(dot(n, H), 0.0), vdh = std::max(dot(V, H), 0.0);
    const double a2 = alpha * alpha;
    const double den = ndh * ndh * (a2 - 1.0) + 1.0;
    const double D = a2 / (kPi * den * den);
    const double k = alpha * 0.5;
    const double G = (ndv / (ndv * (1.0 - k) + k)) * (ndl / (ndl * (1.0 - k) + k));
    const double F = 0.04 + 0.96 * std::pow(1.0 - vdh, 5.0);
    const double spec = D * G * F / (4.0 * ndl * ndv);
    const double diff = kd * (1.0 - F) / kPi;
    return D3{alb.x * diff + ks * spec, alb.y * diff + ks * spec, alb.z * diff + ks * spec} * ndl;
}

// Addressing, filter within a level and border colour of a lookup
// (IfritSamplerT, Structures.h:52-61); the default is repeat + bilinear.
struct Mode {
    int u = 0, v = 0;  // IfritSamplerAddressMode
    int nearest = 0;
    double border[4] = {0, 0, 0, 1};
};

struct Tex {
    const uint8_t* base;
    int w, h, levels;
    std::vector<size_t> off;  // byte offset of each level's first layer
    int layers = 1;
};

// Texel index i of a level n texels wide, or -1 for the border (the Vulkan
// definitions: mirrored repeat reflects every other period, mirror clamp to
// edge reflects once about 0 and then clamps).
long address(long i, long n, int mode) {
    switch (mode) {
        case 1: {
            const long m = ((i % (2 * n)) + 2 * n) % (2 * n);
            return m >= n ? 2 * n - 1 - m : m;
        }
        case 2: return std::min(std::max(i, 0l), n - 1);
        case 3: return (i < 0 || i >= n) ? -1 : i;
        case 4: return std::min(i < 0 ? -1 - i : i, n - 1);
        default: return ((i % n) + n) % n;
    }
}

void texel(const Tex& t, int l, long x, long y, double* c, const Mode& m = Mode(), int layer = 0) {
    const long w = std::max(1, t.w >> l), h = std::max(1, t.h >> l);
    x = address(x, w, m.u);
    y = address(y, h, m.v);
    if (x < 0 || y < 0) {
        for (int k = 0; k < 4; k++) c[k] = m.border[k];
        return;
    }
    const uint8_t* p = t.base + t.off[l] + 4 * ((size_t)w * h * layer + (size_t)y * w + x);
    for (int k = 0; k < 4; k++) c[k] = p[k] / 255.0;
}

bool near_integer(double r) { return std::fabs(r - std::nearbyint(r)) < 1e-4 * std::max(1.0, std::fabs(r)); }

// Filters one level.  Returns true when the nearest filter's texel choice is
// too close to call (the coordinate lies within FP32 rounding of a texel
// edge); bilinear filtering is continuous in (u, v) and never is.
bool bilinear(const Tex& t, int l, double u, double v, double* c, const Mode& m = Mode(), int layer = 0) {
    const double w = std::max(1, t.w >> l), h = std::max(1, t.h >> l);
    if (m.nearest) {
        texel(t, l, (long)std::floor(u * w), (long)std::floor(v * h), c, m, layer);
        return near_integer(u * w) || near_integer(v * h);
    }
    const double s = u * w - 0.5, r = v * h - 0.5;
    const double fx0 = std::floor(s), fy0 = std::floor(r);
    const double fx = s - fx0, fy = r - fy0;
    const long x0 = (long)fx0, y0 = (long)fy0;
    double a[4], b[4], cc[4], d[4];
    texel(t, l, x0, y0, a, m, layer);
    texel(t, l, x0 + 1, y0, b, m, layer);
    texel(t, l, x0, y0 + 1, cc, m, layer);
    texel(t, l, x0 + 1, y0 + 1, d, m, layer);
    for (int k = 0; k < 4; k++) {
        const double top = a[k] + (b[k] - a[k]) * fx;
        const double bot = cc[k] + (d[k] - cc[k]) * fx;
        c[k] = top + (bot - top) * fy;
    }
    return false;
}

bool trilinear(const Tex& t, double u, double v, double lod, double* c, const Mode& m = Mode(), int layer = 0) {
    lod = std::min(std::max(lod, 0.0), (double)(t.levels - 1));
    const int l0 = (int)std::floor(lod);
    const int l1 = std::min(l0 + 1, t.levels - 1);
    const double f = lod - l0;
    double a[4], b[4];
    bool flag = bilinear(t, l0, u, v, a, m, layer);
    if (f > 0) flag |= bilinear(t, l1, u, v, b, m, layer);
    else bilinear(t, l1, u, v, b, m, layer);
    for (int k = 0; k < 4; k++) c[k] = a[k] + (b[k] - a[k]) * f;
    return flag;
}

// Cube-map face and (u, v) from a direction's major axis (the Vulkan face
// table that textureCubeCoordConversion, ShaderBuiltinCuda.cuh:184-208,
// follows); true when the choice is too close to call.
bool cube_face(const double* d, int* face, double* u, double* v) {
    const double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    double ma, sc, tc;
    if (az >= ax && az >= ay) {
        *face = d[2] < 0 ? 5 : 4;
        ma = az;
        sc = d[2] < 0 ? -d[0] : d[0];
        tc = -d[1];
    } else if (ay >= ax) {
        *face = d[1] < 0 ? 3 : 2;
        ma = ay;
        sc = d[0];
        tc = d[1] < 0 ? -d[2] : d[2];
    } else {
        *face = d[0] < 0 ? 1 : 0;
        ma = ax;
        sc = d[0] < 0 ? d[2] : -d[2];
        tc = -d[1];
    }
    *u = sc * (0.5 / ma) + 0.5;
    *v = tc * (0.5 / ma) + 0.5;
    const double top = std::max(ax, std::max(ay, az));
    const double second = ax + ay + az - top - std::min(ax, std::min(ay, az));
    return top - second <= 1e-6 * top;
}

// Returns true when a discontinuous decision is too close to call.
bool sample_tex(const Tex& t, int sampler, int maxAniso, double u, double v, double dudx, double dvdx, double dudy,
                double dvdy, double* c, const Mode& m = Mode()) {
    if (sampler == ORS_BILINEAR) return bilinear(t, 0, u, v, c, m);
    const double px = std::sqrt((dudx * t.w) * (dudx * t.w) + (dvdx * t.h) * (dvdx * t.h));
    const double py = std::sqrt((dudy * t.w) * (dudy * t.w) + (dvdy * t.h) * (dvdy * t.h));
    const double pmax = std::max(px, py), pmin = std::min(px, py);
    if (sampler == ORS_TRILINEAR || pmax <= 0) {
        const double lod = pmax > 0 ? std::log2(pmax) : 0.0;
        return trilinear(t, u, v, lod, c, m);
    }
    bool flag = std::fabs(px - py) <= 1e-4 * pmax;
    double ratio = pmin > 0 ? pmax / pmin : (double)maxAniso;
    int n;
    if (ratio >= maxAniso) {
        n = maxAniso;
        if (std::fabs(ratio - maxAniso) < 1e-4 * maxAniso) flag = true;
    } else {
        n = (int)std::ceil(ratio);
        if (near_integer(ratio)) flag = true;
    }
    n = std::max(1, n);
    const double lod = std::log2(pmax / n);
    const double au = px >= py ? dudx : dudy;
    const double av = px >= py ? dvdx : dvdy;
    double acc[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        const double o = (i + 0.5) / n - 0.5;
        double s[4];
        flag |= trilinear(t, u + au * o, v + av * o, lod, s, m);
        for (int k = 0; k < 4; k++) acc[k] += s[k];
    }
    for (int k = 0; k < 4; k++) c[k] = acc[k] / n;
    return flag;
}

struct Shader {
    const OracleShadeParams* p;
    Mode mode;
    const std::vector<OracleTri>* tris;
    const int32_t* idx;
    const float* vtx;
    const std::vector<Tex>* tex;

    // Perspective-correct vertex weights of triangle t at (x, y).  Returns a
    // bound on how far the FP32 evaluation of the three planes (whose terms
    // may cancel: a sliver seen far from its vertices) can move a weight:
    // each plane value is off by up to 2^-23 of the sum of its terms'
    // magnitudes.  Above 0.01 a pixel is too close to call.
    double weights(const OracleTri& t, double x, double y, double* w) const {
        const double b1 = (double)t.f1[0] * x + (double)t.f1[1] * y + (double)t.f1[2];
        const double b2 = (double)t.f2[0] * x + (double)t.f2[1] * y + (double)t.f2[2];
        const double b3 = (double)t.f3[0] * x + (double)t.f3[1] * y + (double)t.f3[2];
        const double s = b1 + b2 + b3;
        const double B[3] = {b1 / s, b2 / s, b3 / s};
        w[0] = B[0] * t.bx[0] + B[1] * t.bx[1] + B[2] * t.bx[2];
        w[1] = B[0] * t.by[0] + B[1] * t.by[1] + B[2] * t.by[2];
        w[2] = 1.0 - w[0] - w[1];
        const double u = std::ldexp(1.0, -23);
        const double e[3] = {
            u * (std::fabs(t.f1[0] * x) + std::fabs(t.f1[1] * y) + std::fabs((double)t.f1[2])),
            u * (std::fabs(t.f2[0] * x) + std::fabs(t.f2[1] * y) + std::fabs((double)t.f2[2])),
            u * (std::fabs(t.f3[0] * x) + std::fabs(t.f3[1] * y) + std::fabs((double)t.f3[2]))};
        const double es = e[0] + e[1] + e[2], as = std::fabs(s);
        const double bb[3] = {b1, b2, b3};
        double worst = 0.0;
        for (int k = 0; k < 3; k++) worst = std::max(worst, e[k] / as + std::fabs(bb[k]) * es / (as * as));
        return worst;
    }
    // vertex k of triangle t in the order setup took them (exchanged triangles: 2, 1, 0)
    const float* vert(const OracleTri& t, int k) const {
        const int kk = t.swapped ? 2 - k : k;
        return vtx + (size_t)idx[3 * t.prim + kk] * p->stride_floats;
    }
    void uv_at(const OracleTri& t, double x, double y, double* uv) const {
        double w[3];
        weights(t, x, y, w);
        uv[0] = uv[1] = 0;
        for (int k = 0; k < 3; k++) {
            const float* v = vert(t, k);
            uv[0] += w[k] * v[p->off_uv];
            uv[1] += w[k] * v[p->off_uv + 1];
        }
    }

    // Colour of triangle ti shaded at pixel (px, py); o = 0 (1 spp, pixel
    // corner) or 0.5 (MSAA, pixel centre).
    bool shade(uint32_t ti, int px, int py, double o, double* rgba) const {
        const OracleTri& t = (*tris)[ti];
        if (p->program == ORP_FLAT) {
            const uint32_t h = oracle_flat_hash((uint32_t)t.prim);
            rgba[0] = (h & 0xff) / 255.0;
            rgba[1] = ((h >> 8) & 0xff) / 255.0;
            rgba[2] = ((h >> 16) & 0xff) / 255.0;
            rgba[3] = 1.0;
            return false;
        }
        if (p->program == ORP_NORMAL_LAMBERT) {
            // the shader of the reference CPU timing harness (oracle/ref/ref_harness.cpp LambertFS)
            double w[3];
            const bool ill = weights(t, px + o, py + o, w) > 0.01;
            D3 N{0, 0, 0};
            for (int k = 0; k < 3; k++) {
                const float* v = vert(t, k);
                N = N + D3{v[p->off_normal], v[p->off_normal + 1], v[p->off_normal + 2]} * w[k];
            }
            const double l = std::sqrt(dot(N, N));
            const double nl = dot(N, D3{p->light_pos[0][0], p->light_pos[0][1], p->light_pos[0][2]});
            double d = l > 0 ? nl / l : 0.0;
            const bool flag = std::fabs(d) < 1e-4 || ill;
            d = std::max(d, 0.0);
            const double sp = std::pow(d, (double)p->shininess);
            for (int k = 0; k < 3; k++) rgba[k] = p->ambient + p->light_color[0][k] * d + sp;
            rgba[3] = 1.0;
            return flag;
        }
        if (p->program == ORP_CUBE_MAP) {
            double w[3];
            const bool ill = weights(t, px + o, py + o, w) > 0.01;
            double N[3] = {0, 0, 0};
            for (int k = 0; k < 3; k++) {
                const float* v = vert(t, k);
                for (int c = 0; c < 3; c++) N[c] += w[k] * v[p->off_normal + c];
            }
            const int mat = p->tris_per_material > 0 ? t.prim / p->tris_per_material : 0;
            if (tex->empty()) {
                rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0;
                return false;
            }
            const Tex& tx = (*tex)[std::min<int>(mat, (int)tex->size() - 1)];
            int face;
            double u, v;
            const bool flag = cube_face(N, &face, &u, &v) || ill;
            return trilinear(tx, u, v, (double)p->lod, rgba, mode, face) || flag;
        }
        if (p->program != ORP_BLINN_PHONG && p->program != ORP_GGX_LAYERED) {
            rgba[0] = rgba[1] = rgba[2] = rgba[3] = 0;
            return false;
        }
        const double x = px + o, y = py + o;
        double w[3];
        const bool ill = weights(t, x, y, w) > 0.01;
        D3 P{0, 0, 0}, N{0, 0, 0};
        double uv[2] = {0, 0};
        for (int k = 0; k < 3; k++) {
            const float* v = vert(t, k);
            P = P + D3{v[p->off_world], v[p->off_world + 1], v[p->off_world + 2]} * w[k];
            N = N + D3{v[p->off_normal], v[p->off_normal + 1], v[p->off_normal + 2]} * w[k];
            uv[0] += w[k] * v[p->off_uv];
            uv[1] += w[k] * v[p->off_uv + 1];
        }
        // Fine quad derivatives: the same triangle evaluated at the quad's
        // neighbouring shading positions (SBC:36-76).
        const int qx = px & ~1, qy = py & ~1;
        double a[2], b[2], c[2], d[2];
        uv_at(t, qx + o, y, a);
        uv_at(t, qx + 1 + o, y, b);
        uv_at(t, x, qy + o, c);
        uv_at(t, x, qy + 1 + o, d);
        if (p->program == ORP_GGX_LAYERED) {
            D3 T{0, 0, 0};
            for (int k = 0; k < 3; k++) {
                const float* v = vert(t, k);
                T = T + D3{v[p->off_tangent], v[p->off_tangent + 1], v[p->off_tangent + 2]} * w[k];
            }
            double alb[4] = {1, 1, 1, 1}, nm[4] = {0.5, 0.5, 1, 1}, rg[4] = {0, 0, 0, 0};
            bool flag = ill;
            if (tex->size() >= 3) {
                const double du = b[0] - a[0], dv = b[1] - a[1], eu = d[0] - c[0], ev = d[1] - c[1];
                flag |= sample_tex((*tex)[0], p->sampler, p->max_aniso, uv[0], uv[1], du, dv, eu, ev, alb);
                flag |= sample_tex((*tex)[1], p->sampler, p->max_aniso, uv[0], uv[1], du, dv, eu, ev, nm);
                flag |= sample_tex((*tex)[2], p->sampler, p->max_aniso, uv[0], uv[1], du, dv, eu, ev, rg);
            }
            N = normalize(N);
            T = normalize(T - N * dot(N, T));
            const D3 B = cross(N, T);
            const D3 nts = normalize(D3{2 * nm[0] - 1, 2 * nm[1] - 1, 2 * nm[2] - 1});
            const D3 n = normalize(T * nts.x + B * nts.y + N * nts.z);
            const double alpha = 0.05 + 0.95 * rg[0];
            const D3 V = normalize(D3{p->camera[0], p->camera[1], p->camera[2]} - P);
            const D3 A{alb[0], alb[1], alb[2]};
            D3 col = A * p->ambient;
            for (int l = 0; l < p->num_lights; l++) {
                const float* lp = p->light_pos[l];
                const float* lc = p->light_color[l];
                D3 L;
                double att = 1.0;
                if (lp[3] == 0.0f) {
                    L = normalize(D3{lp[0], lp[1], lp[2]});
                } else {
                    const D3 Lv = D3{lp[0], lp[1], lp[2]} - P;
                    const double d2 = dot(Lv, Lv);
                    L = Lv * (1.0 / std::sqrt(d2));
                    att = 1.0 / (1.0 + lc[3] * d2);
                }
                const D3 c = ggx_light(n, V, L, A, alpha, p->kd, p->ks, &flag);
                col = col + D3{c.x * lc[0], c.y * lc[1], c.z * lc[2]} * att;
            }
            rgba[0] = col.x;
            rgba[1] = col.y;
            rgba[2] = col.z;
            rgba[3] = alb[3];
            return flag;
        }
        const int mat = p->tris_per_material > 0 ? t.prim / p->tris_per_material : 0;
        double alb[4] = {1, 1, 1, 1};
        bool flag = ill;
        if (!tex->empty()) {
            const Tex& tx = (*tex)[std::min<int>(mat, (int)tex->size() - 1)];
            flag |= sample_tex(tx, p->sampler, p->max_aniso, uv[0], uv[1], b[0] - a[0], b[1] - a[1], d[0] - c[0],
                              d[1] - c[1], alb, mode);
        }
        N = normalize(N);
        const D3 cam{p->camera[0], p->camera[1], p->camera[2]};
        const D3 V = normalize(g_plant == 3 ? P - cam : cam - P);
        D3 col = D3{alb[0], alb[1], alb[2]} * p->ambient;
        for (int l = 0; l < p->num_lights; l++) {
            const float* lp = p->light_pos[l];
            const float* lc = p->light_color[l];
            D3 L;
            double att = 1.0;
            if (lp[3] == 0.0f) {
                L = normalize(D3{lp[0], lp[1], lp[2]});
            } else {
                const D3 Lv = D3{lp[0], lp[1], lp[2]} - P;
                const double d2 = dot(Lv, Lv);
                L = Lv * (1.0 / std::sqrt(d2));
                att = 1.0 / (1.0 + lc[3] * (g_plant == 2 ? std::sqrt(d2) : d2));
            }
            const double ndl = dot(N, L);
            if (std::fabs(ndl) < 1e-4) flag = true;
            if (ndl > 0) {
                const D3 Hh = g_plant == 1 ? N * (2.0 * ndl) - L : normalize(L + V);
                const double ndh = std::max(dot(g_plant == 1 ? V : N, Hh), 0.0);
                const double spec = p->ks * std::pow(ndh, (double)p->shininess);
                col.x += att * lc[0] * (p->kd * ndl * alb[0] + spec);
                col.y += att * lc[1] * (p->kd * ndl * alb[1] + spec);
                col.z += att * lc[2] * (p->kd * ndl * alb[2] + spec);
            }
        }
        rgba[0] = col.x;
        rgba[1] = col.y;
        rgba[2] = col.z;
        rgba[3] = alb[3];
        return flag;
    }
};

inline uint8_t to_u8(double c) {
    c = std::min(std::max(c, 0.0), 1.0);
    return (uint8_t)std::floor(c * 255.0 + 0.5);
}

}  // namespace

extern "C" {

void oracle_set_plant(int plant) { g_plant = plant; }

// One light of the layered-GGX material (ggx_light above), for the closed-form tests.
void oracle_ggx_light(const double* n, const double* v, const double* l, const double* alb, double alpha, double kd,
                      double ks, double* out) {
    bool flag = false;
    const D3 c = ggx_light(D3{n[0], n[1], n[2]}, D3{v[0], v[1], v[2]}, D3{l[0], l[1], l[2]},
                           D3{alb[0], alb[1], alb[2]}, alpha, kd, ks, &flag);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

// One texture lookup (sample_tex above) of an RGBA8 texture with its packed
// mip chain, for the closed-form sampler tests; returns 1 when a discrete
// decision (tap count) was too close to call.
int oracle_sample_tex(const uint8_t* texels, int width, int height, int levels, int sampler, int maxAniso, double u,
                      double v, double dudx, double dvdx, double dudy, double dvdy, double* out) {
    Tex t;
    t.base = texels;
    t.w = width;
    t.h = height;
    t.levels = levels;
    size_t off = 0;
    for (int l = 0; l < levels; l++) {
        t.off.push_back(off);
        off += (size_t)std::max(1, width >> l) * std::max(1, height >> l) * 4;
    }
    return sample_tex(t, sampler, maxAniso, u, v, dudx, dvdx, dudy, dvdy, out) ? 1 : 0;
}

int oracle_sample_tex_mode(const uint8_t* texels, int width, int height, int levels, int sampler, int maxAniso,
                           int addr_u, int addr_v, int nearest, const float* border, double u, double v, double dudx,
                           double dvdx, double dudy, double dvdy, double* out) {
    Tex t;
    t.base = texels;
    t.w = width;
    t.h = height;
    t.levels = levels;
    size_t off = 0;
    for (int l = 0; l < levels; l++) {
        t.off.push_back(off);
        off += (size_t)std::max(1, width >> l) * std::max(1, height >> l) * 4;
    }
    Mode m;
    m.u = addr_u;
    m.v = addr_v;
    m.nearest = nearest;
    for (int k = 0; k < 4; k++) m.border[k] = border[k];
    return sample_tex(t, sampler, maxAniso, u, v, dudx, dvdx, dudy, dvdy, out, m) ? 1 : 0;
}

int oracle_cube_lookup(const uint8_t* texels, int size, int levels, int addr_u, int addr_v, int nearest,
                       const float* border, const double* d, double lod, double* out) {
    Tex t;
    t.base = texels;
    t.w = t.h = size;
    t.levels = levels;
    t.layers = 6;
    size_t off = 0;
    for (int l = 0; l < levels; l++) {
        t.off.push_back(off);
        off += (size_t)std::max(1, size >> l) * std::max(1, size >> l) * 4 * 6;
    }
    Mode m;
    m.u = addr_u;
    m.v = addr_v;
    m.nearest = nearest;
    for (int k = 0; k < 4; k++) m.border[k] = border[k];
    int face;
    double u, v;
    const bool flag = cube_face(d, &face, &u, &v);
    return (trilinear(t, u, v, lod, out, m, face) || flag) ? 1 : 0;
}

uint32_t oracle_flat_hash(uint32_t prim) {
    uint32_t h = prim * 0x9E3779B1u;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    h ^= h >> 13;
    return h;
}

void oracle_build_mips(uint8_t* texels, int width, int height, int levels) {
    size_t off = 0;
    for (int l = 1; l < levels; l++) {
        const int pw = std::max(1, width >> (l - 1)), ph = std::max(1, height >> (l - 1));
        const int w = std::max(1, width >> l), h = std::max(1, height >> l);
        const uint8_t* src = texels + off;
        uint8_t* dst = texels + off + (size_t)pw * ph * 4;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++)
                for (int k = 0; k < 4; k++) {
                    const int x0 = std::min(2 * x, pw - 1), x1 = std::min(2 * x + 1, pw - 1);
                    const int y0 = std::min(2 * y, ph - 1), y1 = std::min(2 * y + 1, ph - 1);
                    const int s = src[4 * ((size_t)y0 * pw + x0) + k] + src[4 * ((size_t)y0 * pw + x1) + k] +
                                  src[4 * ((size_t)y1 * pw + x0) + k] + src[4 * ((size_t)y1 * pw + x1) + k];
                    dst[4 * ((size_t)y * w + x) + k] = (uint8_t)((s + 2) >> 2);
                }
        off += (size_t)pw * ph * 4;
    }
}

int oracle_shade(const OracleFrame* f, const OracleShadeParams* p, int W, int H, int S, const uint32_t* tri_ids,
                 const int32_t* indices, const float* vertices, const OracleTexture* textures, int ntex,
                 uint8_t* rgba8_out, uint8_t* flag_out, int num_threads) {
    std::vector<OracleTri> tris((size_t)oracle_num_tris(f));
    if (!tris.empty()) oracle_get_tris(f, tris.data());
    std::vector<Tex> tex((size_t)std::max(0, ntex));
    for (int i = 0; i < ntex; i++) {
        tex[i].base = textures[i].texels;
        tex[i].w = textures[i].width;
        tex[i].h = textures[i].height;
        tex[i].levels = textures[i].levels;
        tex[i].layers = std::max(1, textures[i].layers);
        size_t off = 0;
        for (int l = 0; l < textures[i].levels; l++) {
            tex[i].off.push_back(off);
            off += (size_t)std::max(1, textures[i].width >> l) * std::max(1, textures[i].height >> l) * 4 *
                   tex[i].layers;
        }
    }
    Mode mode;
    mode.u = p->addr_u;
    mode.v = p->addr_v;
    mode.nearest = p->nearest;
    for (int k = 0; k < 4; k++) mode.border[k] = p->border[k];
    Shader sh{p, mode, &tris, indices, vertices, &tex};
    const double o = S == 1 ? 0.0 : 0.5;
    std::atomic<int> next{0}, flagged{0};
    auto worker = [&]() {
        for (;;) {
            const int row = next.fetch_add(1);
            if (row >= H) break;
            for (int x = 0; x < W; x++) {
                const size_t pix = (size_t)row * W + x;
                const uint32_t* ids = tri_ids + pix * S;
                double sum[4] = {0, 0, 0, 0};
                bool flag = false;
                int covered = 0;
                for (int s = 0; s < S; s++) {
                    if (ids[s] == 0xffffffffu) continue;
                    int first = s;
                    for (int q = 0; q < s; q++)
                        if (ids[q] == ids[s]) { first = q; break; }
                    if (first != s) continue;  // shade each distinct triangle once
                    int count = 0;
                    for (int q = s; q < S; q++) count += ids[q] == ids[s];
                    covered += count;
                    if (g_plant == 5) count = 1;
                    double c[4];
                    flag |= sh.shade(ids[s], x, row, o, c);
                    for (int k = 0; k < 4; k++) sum[k] += std::min(std::max(c[k], 0.0), 1.0) * count;
                }
                if (g_plant != 4)
                    for (int k = 0; k < 4; k++) sum[k] += (double)p->clear_color[k] * (S - covered);
                const int div = g_plant == 4 ? std::max(1, covered) : S;
                for (int k = 0; k < 4; k++) rgba8_out[4 * pix + k] = to_u8(sum[k] / div);
                if (flag_out) flag_out[pix] = flag;
                if (flag) flagged++;
            }
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < std::max(1, num_threads); i++) th.emplace_back(worker);
    worker();
    for (auto& t : th) t.join();
    return flagged.load();
}

}  // extern "C"
