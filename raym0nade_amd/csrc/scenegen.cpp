// Seeded procedural scenes and ray batches for the benchmark workloads.
// All randomness comes from Philox-4x32-10 keyed by (seed, tag), so every
// array is a pure function of (config, seed, options).
//
//   config 0  Cornell box: 5 wall quads, a 0.25x0.25 light quad at y=0.999
//             (radiance 15), two closed boxes rotated U(0,45 deg) about y;
//             36 triangles; albedo U(0.2,0.8)^3 per mesh; 512x512, fov 40.
//   config 1  triangle soup: centroids U[0,1]^3, vertices c + N(0, 0.01)^3.
//   config 2  64 geodesic spheres (frequency 90 -> 162000 triangles each,
//             radius 0.4, 8x8 grid, spacing 1) displaced by 3-octave value
//             noise (amplitude 0.05) on a ground quad; 50/30/20 % Lambertian /
//             GGX metal (roughness U(0.05,0.8)) / dielectric (IOR 1.5);
//             procedural fp32 sky; 1920x1080, fov 40.
//   config 4  config 0 with a 256x256 fp32 albedo texture per wall (smooth
//             value noise in [0.1,0.9]) and roughness U(0.1,0.6) per wall.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <unordered_map>
#include <vector>
#include "rm_internal.h"
#include "rm_vec.h"

namespace rm {
namespace {

constexpr double PI_D = 3.14159265358979323846;

struct Rng {
    uint32_t seed, tag;
    // k-th uniform in [0,1) of item i
    double u(uint32_t i, uint32_t k) const {
        U4 r = philox4x32_10(U4{i, k >> 2, 0u, 0u}, seed, tag);
        uint32_t w = (k & 3) == 0 ? r.x : ((k & 3) == 1 ? r.y : ((k & 3) == 2 ? r.z : r.w));
        return (double)(w >> 8) * (1.0 / 16777216.0);
    }
    // standard normal (Box-Muller on two uniforms)
    double n(uint32_t i, uint32_t k) const {
        double a = u(i, 2 * k), b = u(i, 2 * k + 1);
        a = a < 1e-12 ? 1e-12 : a;
        return sqrt(-2.0 * log(a)) * cos(2.0 * PI_D * b);
    }
};

struct Gen {
    std::vector<float> verts, uvs, normals;
    std::vector<int32_t> material, mesh;
    std::vector<rm_material> mats;
    std::vector<rm_texture> texs;
    std::vector<std::vector<float>> texdata;
    std::vector<float> sky;
    int32_t sky_w = 0, sky_h = 0;
    rm_render_args args{};
    int32_t next_mesh = 0;

    void tri(const double a[3], const double b[3], const double c[3], const double na[3], const double nb[3],
             const double nc[3], const float ua[2], const float ub[2], const float uc[2], int32_t mat, int32_t m) {
        const double *P[3] = {a, b, c}, *N[3] = {na, nb, nc};
        const float *U[3] = {ua, ub, uc};
        for (int k = 0; k < 3; ++k) {
            for (int j = 0; j < 3; ++j) verts.push_back((float)P[k][j]);
            for (int j = 0; j < 3; ++j) normals.push_back((float)N[k][j]);
            uvs.push_back(U[k][0]);
            uvs.push_back(U[k][1]);
        }
        material.push_back(mat);
        mesh.push_back(m);
    }
    // quad p0 p1 p2 p3 (counter-clockwise) as two triangles with corner uvs
    void quad(const double p[4][3], int32_t mat, int32_t m) {
        double e1[3], e2[3], n[3];
        for (int j = 0; j < 3; ++j) { e1[j] = p[1][j] - p[0][j]; e2[j] = p[2][j] - p[0][j]; }
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (double &x : n) x /= l;
        const float uv[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
        tri(p[0], p[1], p[2], n, n, n, uv[0], uv[1], uv[2], mat, m);
        tri(p[0], p[2], p[3], n, n, n, uv[0], uv[2], uv[3], mat, m);
    }
    int32_t add_mat(rm_material m) {
        mats.push_back(m);
        return (int32_t)mats.size() - 1;
    }
};

rm_material lambert(float r, float g, float b) {
    rm_material m{};
    m.base_color[0] = r; m.base_color[1] = g; m.base_color[2] = b;
    m.roughness = 0.8f;  // Material default (src/material.cpp:109-111)
    m.metallic = 0.0f;
    m.opacity = 1.0f;
    m.ior = 1.0f;
    m.albedo_tex = m.normal_tex = m.emissive_tex = m.specular_tex = -1;
    return m;
}

// Camera basis for RenderArgs: direction, right, up (up points down the image).
void camera(rm_render_args &a, const double pos[3], const double target[3], int w, int h, double fov_deg) {
    double d[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (double &x : d) x /= l;
    double wu[3] = {0, 1, 0};
    // right = d x world_up, image-up = d x right (points to world-down)
    double r[3] = {d[1] * wu[2] - d[2] * wu[1], d[2] * wu[0] - d[0] * wu[2], d[0] * wu[1] - d[1] * wu[0]};
    double rl = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double &x : r) x /= rl;
    double u[3] = {d[1] * r[2] - d[2] * r[1], d[2] * r[0] - d[0] * r[2], d[0] * r[1] - d[1] * r[0]};
    for (int j = 0; j < 3; ++j) {
        a.position[j] = (float)pos[j];
        a.direction[j] = (float)d[j];
        a.right[j] = (float)r[j];
        a.up[j] = (float)u[j];
    }
    a.width = w;
    a.height = h;
    a.accuracy = (float)(2.0 * tan(fov_deg * PI_D / 360.0) / w);
    a.focus = 1.0f;
    a.coc = 0.0f;
    a.exposure = 1.0f;
    a.p_direct = 0.7f;
    a.threads = 0;
    a.seed = 1;
}

void cornell(Gen &G, uint32_t seed, bool textured, int tex_size) {
    Rng R{seed, TAG_SCENE};
    uint32_t item = 0;
    auto albedo = [&](uint32_t i) {
        return lambert((float)(0.2 + 0.6 * R.u(i, 0)), (float)(0.2 + 0.6 * R.u(i, 1)), (float)(0.2 + 0.6 * R.u(i, 2)));
    };
    // walls: floor, ceiling, back, left, right; open side at z = 0
    const double W[5][4][3] = {
        {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}},
        {{0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}},
        {{0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}},
        {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},
        {{1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}},
    };
    for (int w = 0; w < 5; ++w) {
        rm_material m = albedo(item);
        if (textured) {
            m.roughness = (float)(0.1 + 0.5 * R.u(item, 3));
            // smooth value noise texture in [0.1, 0.9]
            int S = tex_size;
            std::vector<float> t((size_t)S * S * 3);
            const int L = 8;  // lattice cells per side
            std::vector<double> lat((size_t)L * L * 3);
            for (int k = 0; k < L * L * 3; ++k) lat[k] = R.u(1000 + item * 64 + k / 4, k & 3);
            for (int y = 0; y < S; ++y)
                for (int x = 0; x < S; ++x) {
                    double fx = (x + 0.5) / S * L, fy = (y + 0.5) / S * L;
                    int x0 = (int)floor(fx), y0 = (int)floor(fy);
                    double tx = fx - x0, ty = fy - y0;
                    tx = tx * tx * (3 - 2 * tx);
                    ty = ty * ty * (3 - 2 * ty);
                    for (int c = 0; c < 3; ++c) {
                        auto at = [&](int i, int j) { return lat[(((j % L) * L) + (i % L)) * 3 + c]; };
                        double v = (at(x0, y0) * (1 - tx) + at(x0 + 1, y0) * tx) * (1 - ty) +
                                   (at(x0, y0 + 1) * (1 - tx) + at(x0 + 1, y0 + 1) * tx) * ty;
                        t[((size_t)y * S + x) * 3 + c] = (float)(0.1 + 0.8 * v);
                    }
                }
            G.texdata.push_back(std::move(t));
            m.albedo_tex = (int32_t)G.texdata.size() - 1;
        }
        int32_t mi = G.add_mat(m);
        G.quad(W[w], mi, G.next_mesh++);
        ++item;
    }
    // light quad at y = 0.999, 0.25 x 0.25, facing down
    {
        rm_material m = albedo(item++);
        m.emission[0] = m.emission[1] = m.emission[2] = 15.0f;
        int32_t mi = G.add_mat(m);
        const double y = 0.999, a = 0.375, b = 0.625;
        const double Q[4][3] = {{a, y, a}, {a, y, b}, {b, y, b}, {b, y, a}};
        G.quad(Q, mi, G.next_mesh++);
    }
    // two closed boxes rotated about y by U(0, 45 deg)
    const double cx[2] = {0.33, 0.66}, cz[2] = {0.37, 0.63}, hy[2] = {0.3, 0.6};
    for (int bx = 0; bx < 2; ++bx) {
        double ang = R.u(item, 3) * 45.0 * PI_D / 180.0;
        rm_material m = albedo(item++);
        int32_t mi = G.add_mat(m);
        int32_t mesh = G.next_mesh++;
        const double s = 0.15;
        double c[8][3];
        for (int k = 0; k < 8; ++k) {
            double lx = (k & 1) ? s : -s, ly = (k & 2) ? hy[bx] : 0.0, lz = (k & 4) ? s : -s;
            c[k][0] = cx[bx] + lx * cos(ang) - lz * sin(ang);
            c[k][1] = ly;
            c[k][2] = cz[bx] + lx * sin(ang) + lz * cos(ang);
        }
        const int F[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
        for (int f = 0; f < 6; ++f) {
            double Q[4][3];
            for (int k = 0; k < 4; ++k) memcpy(Q[k], c[F[f][k]], sizeof(Q[k]));
            G.quad(Q, mi, mesh);
        }
    }
    for (size_t t = 0; t < G.texdata.size(); ++t) {
        rm_texture x{};
        x.width = x.height = tex_size;
        x.rgb = G.texdata[t].data();
        G.texs.push_back(x);
    }
    const double pos[3] = {0.5, 0.5, -1.4}, tgt[3] = {0.5, 0.5, 0.5};
    camera(G.args, pos, tgt, 512, 512, 40.0);
    G.args.spp = textured ? 128 : 64;
    G.args.max_depth = 8;
    G.args.rr_min_depth = textured ? 1 : 3;
}

void soup(Gen &G, uint32_t seed, int64_t n) {
    Rng R{seed, TAG_SCENE};
    G.mats.push_back(lambert(0.5f, 0.5f, 0.5f));
    G.verts.resize(n * 9);
    G.material.assign(n, 0);
    G.mesh.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) {
        double c[3] = {R.u((uint32_t)i, 0), R.u((uint32_t)i, 1), R.u((uint32_t)i, 2)};
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) G.verts[i * 9 + k * 3 + j] = (float)(c[j] + 0.01 * R.n((uint32_t)i, 2 + k * 3 + j));
    }
    const double pos[3] = {0.5, 0.5, -1.5}, tgt[3] = {0.5, 0.5, 0.5};
    camera(G.args, pos, tgt, 2828, 2828, 40.0);
}

// ---------------------------------------------------------------- spheres
struct Icosa {
    std::vector<double> v;          // unit vertices [nv][3]
    std::vector<int32_t> f;         // [nf][3]
};

Icosa geodesic(int freq) {
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    double V[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    Icosa I;
    struct Key {
        uint64_t a, b, c;
        bool operator==(const Key &o) const { return a == o.a && b == o.b && c == o.c; }
    };
    struct H {
        size_t operator()(const Key &k) const { return k.a * 0x9E3779B97F4A7C15ull ^ (k.b + 0x632BE59BD9B4E019ull) * 31 ^ k.c; }
    };
    std::unordered_map<Key, int32_t, H> index;
    auto vid = [&](double p[3]) {
        double l = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        double q[3] = {p[0] / l + 0.0, p[1] / l + 0.0, p[2] / l + 0.0};  // +0.0 folds -0 into +0
        Key k;
        memcpy(&k.a, &q[0], 8);
        memcpy(&k.b, &q[1], 8);
        memcpy(&k.c, &q[2], 8);
        auto it = index.find(k);
        if (it != index.end()) return it->second;
        int32_t id = (int32_t)(I.v.size() / 3);
        I.v.insert(I.v.end(), q, q + 3);
        index.emplace(k, id);
        return id;
    };
    for (int fi = 0; fi < 20; ++fi) {
        const double *A = V[F[fi][0]], *B = V[F[fi][1]], *C = V[F[fi][2]];
        std::vector<int32_t> grid((freq + 1) * (freq + 2) / 2);
        auto g = [&](int i, int j) -> int32_t & { return grid[i * (freq + 1) - i * (i - 1) / 2 + j]; };
        for (int i = 0; i <= freq; ++i)
            for (int j = 0; j + i <= freq; ++j) {
                int k = freq - i - j;
                // edge points are sums of two terms, exact under reordering, so
                // neighbouring faces produce bit-identical shared vertices
                double p[3];
                for (int d = 0; d < 3; ++d) p[d] = (A[d] * k + B[d] * i) + C[d] * j;
                if (j == 0) for (int d = 0; d < 3; ++d) p[d] = A[d] * k + B[d] * i;
                else if (i == 0) for (int d = 0; d < 3; ++d) p[d] = A[d] * k + C[d] * j;
                else if (k == 0) for (int d = 0; d < 3; ++d) p[d] = B[d] * i + C[d] * j;
                g(i, j) = vid(p);
            }
        for (int i = 0; i < freq; ++i)
            for (int j = 0; j + i < freq; ++j) {
                I.f.insert(I.f.end(), {g(i, j), g(i + 1, j), g(i, j + 1)});
                if (i + j + 1 < freq) I.f.insert(I.f.end(), {g(i + 1, j), g(i + 1, j + 1), g(i, j + 1)});
            }
    }
    return I;
}

double hash01(uint32_t seed, int32_t x, int32_t y, int32_t z, uint32_t oct) {
    U4 r = philox4x32_10(U4{(uint32_t)x, (uint32_t)y, (uint32_t)z, oct}, seed, TAG_SCENE + 7);
    return (double)(r.x >> 8) * (1.0 / 16777216.0);
}

double value_noise3(uint32_t seed, const double p[3]) {
    double sum = 0.0, amp = 1.0, norm = 0.0, f = 4.0;
    for (uint32_t o = 0; o < 3; ++o) {
        double x = p[0] * f, y = p[1] * f, z = p[2] * f;
        int32_t x0 = (int32_t)floor(x), y0 = (int32_t)floor(y), z0 = (int32_t)floor(z);
        double tx = x - x0, ty = y - y0, tz = z - z0;
        tx = tx * tx * (3 - 2 * tx);
        ty = ty * ty * (3 - 2 * ty);
        tz = tz * tz * (3 - 2 * tz);
        double c[2][2][2];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int d = 0; d < 2; ++d) c[a][b][d] = hash01(seed, x0 + a, y0 + b, z0 + d, o);
        double xy0 = (c[0][0][0] * (1 - tx) + c[1][0][0] * tx) * (1 - ty) + (c[0][1][0] * (1 - tx) + c[1][1][0] * tx) * ty;
        double xy1 = (c[0][0][1] * (1 - tx) + c[1][0][1] * tx) * (1 - ty) + (c[0][1][1] * (1 - tx) + c[1][1][1] * tx) * ty;
        sum += amp * (2.0 * (xy0 * (1 - tz) + xy1 * tz) - 1.0);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

void sky_env(Gen &G, int w, int h) {
    G.sky_w = w;
    G.sky_h = h;
    G.sky.resize((size_t)w * h * 3);
    double s[3] = {0.3, 0.8, 0.5};
    double sl = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double &x : s) x /= sl;
    for (int v = 0; v < h; ++v)
        for (int u = 0; u < w; ++u) {
            // the pixel -> direction map of sampleSkyBox (src/sampling.cpp:454-457)
            double phi = PI_D * (v + 0.5) / h, th = 2.0 * PI_D * (u + 0.5) / w;
            double d[3] = {-sin(phi) * sin(th), cos(phi), sin(phi) * cos(th)};
            double c[3];
            if (d[1] >= 0) {
                double k = sqrt(d[1]);
                const double hor[3] = {0.9, 0.85, 0.8}, zen[3] = {0.25, 0.45, 1.1};
                for (int j = 0; j < 3; ++j) c[j] = hor[j] * (1 - k) + zen[j] * k;
            } else {
                const double gnd[3] = {0.08, 0.07, 0.06};
                for (int j = 0; j < 3; ++j) c[j] = gnd[j];
            }
            double cs = d[0] * s[0] + d[1] * s[1] + d[2] * s[2];
            double sun = exp((cs - 1.0) * 3000.0);
            const double sc[3] = {400.0, 380.0, 350.0};
            for (int j = 0; j < 3; ++j) G.sky[((size_t)v * w + u) * 3 + j] = (float)(c[j] + sc[j] * sun);
        }
}

void spheres(Gen &G, uint32_t seed, int grid, int freq, int sky_w, int sky_h) {
    Rng R{seed, TAG_SCENE};
    Icosa I = geodesic(freq);
    const int nv = (int)(I.v.size() / 3), nf = (int)(I.f.size() / 3);
    const double radius = 0.4, amp = 0.05;
    // ground quad
    {
        int32_t mi = G.add_mat(lambert(0.5f, 0.5f, 0.5f));
        const double Q[4][3] = {{-50, 0, -50}, {-50, 0, 50}, {50, 0, 50}, {50, 0, -50}};
        G.quad(Q, mi, G.next_mesh++);
    }
    std::vector<double> P((size_t)nv * 3), N((size_t)nv * 3);
    for (int s = 0; s < grid * grid; ++s) {
        uint32_t item = 100 + s;
        double kind = R.u(item, 0);
        rm_material m = lambert((float)(0.2 + 0.6 * R.u(item, 1)), (float)(0.2 + 0.6 * R.u(item, 2)),
                                (float)(0.2 + 0.6 * R.u(item, 3)));
        if (kind >= 0.5 && kind < 0.8) {  // GGX metal
            m.metallic = 1.0f;
            m.roughness = (float)(0.05 + 0.75 * R.u(item, 4));
        } else if (kind >= 0.8) {         // dielectric
            m.opacity = 0.0f;
            m.ior = 1.5f;
            m.roughness = 5e-3f;  // the reference's glass roughness (src/material.cpp:313)
            m.transmitting_color[0] = m.transmitting_color[1] = m.transmitting_color[2] = 1.0f;
        }
        int32_t mi = G.add_mat(m);
        int32_t mesh = G.next_mesh++;
        double cx = (s % grid) - (grid - 1) / 2.0, cz = (s / grid) - (grid - 1) / 2.0, cy = radius + amp;
        uint32_t nseed = seed * 977u + (uint32_t)s;
        for (int i = 0; i < nv; ++i) {
            const double *q = &I.v[(size_t)i * 3];
            double r = radius + amp * value_noise3(nseed, q);
            P[i * 3] = cx + r * q[0];
            P[i * 3 + 1] = cy + r * q[1];
            P[i * 3 + 2] = cz + r * q[2];
        }
        std::fill(N.begin(), N.end(), 0.0);
        for (int f = 0; f < nf; ++f) {
            const int32_t *t = &I.f[(size_t)f * 3];
            const double *a = &P[t[0] * 3], *b = &P[t[1] * 3], *c = &P[t[2] * 3];
            double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            for (int k = 0; k < 3; ++k)
                for (int j = 0; j < 3; ++j) N[t[k] * 3 + j] += n[j];
        }
        for (int i = 0; i < nv; ++i) {
            double *n = &N[i * 3];
            double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            for (int j = 0; j < 3; ++j) n[j] /= l;
        }
        const float uv0[2] = {0, 0};
        for (int f = 0; f < nf; ++f) {
            const int32_t *t = &I.f[(size_t)f * 3];
            G.tri(&P[t[0] * 3], &P[t[1] * 3], &P[t[2] * 3], &N[t[0] * 3], &N[t[1] * 3], &N[t[2] * 3], uv0, uv0, uv0, mi,
                  mesh);
        }
    }
    if (sky_w > 0 && sky_h > 0) sky_env(G, sky_w, sky_h);
    const double pos[3] = {0.0, 3.2, -7.5}, tgt[3] = {0.0, 0.2, 0.3};
    camera(G.args, pos, tgt, 1920, 1080, 40.0);
    G.args.spp = 256;
    G.args.max_depth = 16;
    G.args.rr_min_depth = 1;
}

}  // namespace
}  // namespace rm

using namespace rm;

struct rm_generated {
    Gen g;
    rm_scene_desc desc{};
};

extern "C" int32_t rm_generate_scene(int32_t config, uint32_t seed, const rm_gen_options *opt, rm_generated **out) {
    if (!out) return fail(RM_ERR_INVALID, "rm_generate_scene: out is NULL");
    *out = nullptr;
    rm_gen_options o{};
    if (opt) o = *opt;
    rm_generated *G = new (std::nothrow) rm_generated();
    if (!G) return fail(RM_ERR_OOM, "rm_generate_scene: host allocation");
    switch (config) {
        case 0: cornell(G->g, seed, false, 0); break;
        case 1: soup(G->g, seed, o.n_tris > 0 ? o.n_tris : 1000000); break;
        case 2:
            spheres(G->g, seed, o.grid > 0 ? o.grid : 8, o.freq > 0 ? o.freq : 90, o.sky_width >= 0 ? (o.sky_width ? o.sky_width : 2048) : 0,
                    o.sky_height > 0 ? o.sky_height : 1024);
            break;
        case 4: cornell(G->g, seed, true, o.tex_size > 0 ? o.tex_size : 256); break;
        default:
            delete G;
            return fail(RM_ERR_INVALID, "rm_generate_scene: unknown config %d", config);
    }
    Gen &g = G->g;
    rm_scene_desc &d = G->desc;
    d.n_tris = (int64_t)g.material.size();
    d.verts = g.verts.data();
    d.uvs = g.uvs.empty() ? nullptr : g.uvs.data();
    d.normals = g.normals.empty() ? nullptr : g.normals.data();
    d.material = g.material.data();
    d.mesh = g.mesh.data();
    d.n_materials = (int32_t)g.mats.size();
    d.materials = g.mats.data();
    d.n_textures = (int32_t)g.texs.size();
    d.textures = g.texs.empty() ? nullptr : g.texs.data();
    d.sky_width = g.sky_w;
    d.sky_height = g.sky_h;
    d.sky_rgb = g.sky.empty() ? nullptr : g.sky.data();
    g.args.seed = seed;
    *out = G;
    return RM_OK;
}

extern "C" int32_t rm_generated_desc(const rm_generated *G, rm_scene_desc *desc, rm_render_args *args) {
    if (!G) return fail(RM_ERR_INVALID, "rm_generated_desc: NULL");
    if (desc) *desc = G->desc;
    if (args) *args = G->g.args;
    return RM_OK;
}

extern "C" int32_t rm_generated_destroy(rm_generated *G) {
    delete G;
    return RM_OK;
}

extern "C" int32_t rm_generate_rays(int32_t kind, uint32_t seed, int64_t n, const rm_render_args *cam, float *rays) {
    if (!rays || n < 0) return fail(RM_ERR_INVALID, "rm_generate_rays: bad arguments");
    if (kind == 0) {
        // pinhole primary rays exactly as renderPixel forms them (src/render.cpp:466-476),
        // row-major over cam->width x cam->height (n must equal width*height)
        if (!cam || (int64_t)cam->width * cam->height != n) return fail(RM_ERR_INVALID, "rm_generate_rays: coherent rays need a camera with width*height == n");
        V3 view{cam->direction[0], cam->direction[1], cam->direction[2]};
        V3 right{cam->right[0], cam->right[1], cam->right[2]};
        V3 up{cam->up[0], cam->up[1], cam->up[2]};
        for (int64_t i = 0; i < n; ++i) {
            int x = (int)(i % cam->width), y = (int)(i / cam->width);
            float rx = float(x) - float(cam->width) / 2.0f, ry = float(y) - float(cam->height) / 2.0f;
            V3 d = normalize(view + cam->accuracy * (rx * right + ry * up));
            float *p = rays + i * 6;
            p[0] = cam->position[0]; p[1] = cam->position[1]; p[2] = cam->position[2];
            p[3] = d.x; p[4] = d.y; p[5] = d.z;
        }
        return RM_OK;
    }
    if (kind == 1) {
        // origins U[0,1]^3, directions uniform on the sphere (z ~ U[-1,1], phi ~ U[0,2pi))
        Rng R{seed, TAG_RAYS};
        for (int64_t i = 0; i < n; ++i) {
            double z = 2.0 * R.u((uint32_t)i, 3) - 1.0, ph = 2.0 * PI_D * R.u((uint32_t)i, 4);
            double s = sqrt(std::max(0.0, 1.0 - z * z));
            float *p = rays + i * 6;
            p[0] = (float)R.u((uint32_t)i, 0);
            p[1] = (float)R.u((uint32_t)i, 1);
            p[2] = (float)R.u((uint32_t)i, 2);
            p[3] = (float)(s * cos(ph));
            p[4] = (float)(s * sin(ph));
            p[5] = (float)z;
        }
        return RM_OK;
    }
    return fail(RM_ERR_INVALID, "rm_generate_rays: unknown kind %d", kind);
}
