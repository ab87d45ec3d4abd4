/*
 * scenes.cpp - deterministic synthetic scene generators (libnxb_scenes.so).
 *
 * These produce the five workloads the pipeline is measured and checked on.
 * The reference ships no scenes (its benchmark draws random triangles with
 * std::mt19937, Test/src/Main.cpp:31-64); these generators replace it.
 *
 * Randomness: PCG32 XSH-RR (O'Neill 2014): 64-bit LCG state, output
 * rotr32(((s >> 18) ^ s) >> 27, s >> 59). Floats are (u >> 8) * 2^-24 in
 * [0, 1). Work is split into fixed chunks of 65536 triangles; chunk c draws
 * from stream (seed, c), so the output does not depend on the thread count.
 */
#include "../../include/nxb.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t seed, uint64_t stream) {
        state = 0;
        inc = (stream << 1) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    double uniformd() { return (double)(next() >> 8) * (1.0 / 16777216.0); }
    float range(float a, float b) { return a + (b - a) * uniform(); }
    /* Box-Muller on two uniforms; the first uniform is kept away from 0. */
    double normal() {
        double u1 = ((double)(next() >> 8) + 0.5) * (1.0 / 16777216.0);
        double u2 = uniformd();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

constexpr uint32_t kChunk = 65536;

template <typename F>
void parallel_chunks(uint64_t n, F&& f) {
    uint64_t chunks = (n + kChunk - 1) / kChunk;
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    unsigned nt = (unsigned)std::min<uint64_t>(hw, chunks);
    if (nt <= 1) {
        for (uint64_t c = 0; c < chunks; ++c) f(c, c * kChunk, std::min<uint64_t>(n, (c + 1) * kChunk));
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            for (uint64_t c = t; c < chunks; c += nt)
                f(c, c * kChunk, std::min<uint64_t>(n, (c + 1) * kChunk));
        });
    }
    for (auto& x : th) x.join();
}

struct V3 { double x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 norm(V3 a) { double l = std::sqrt(dot(a, a)); return mul(a, 1.0 / l); }
inline nxb_float3 f3(V3 a) { return {(float)a.x, (float)a.y, (float)a.z}; }

/* Lattice hash for value noise (PCG output permutation of a mixed integer). */
inline double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
    uint64_t h = seed ^ 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)x * 0xD6E8FEB86659FD93ull;
    h ^= (uint64_t)y * 0xA0761D6478BD642Full;
    h ^= (uint64_t)z * 0xE7037ED1A0B428DBull;
    h = h * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t xs = (uint32_t)(((h >> 18) ^ h) >> 27);
    uint32_t rot = (uint32_t)(h >> 59);
    uint32_t v = (xs >> rot) | (xs << ((32 - rot) & 31));
    return (double)v * (2.0 / 4294967295.0) - 1.0;
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

double value_noise(V3 p, uint64_t seed) {
    double fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = smooth(p.x - fx), ty = smooth(p.y - fy), tz = smooth(p.z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty;
    double y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

/* 4 octaves, base frequency 4, frequency x2 and amplitude x1/2 per octave,
 * normalized to [-1, 1]. */
double fbm4(V3 p, uint64_t seed) {
    double amp = 1.0, freq = 4.0, sum = 0.0, norm_ = 0.0;
    for (int o = 0; o < 4; ++o) {
        sum += amp * value_noise(mul(p, freq), seed + (uint64_t)o * 7919u);
        norm_ += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    return sum / norm_;
}

const V3 kIco[12] = {
    {-1, 1.618033988749895, 0}, {1, 1.618033988749895, 0}, {-1, -1.618033988749895, 0},
    {1, -1.618033988749895, 0}, {0, -1, 1.618033988749895}, {0, 1, 1.618033988749895},
    {0, -1, -1.618033988749895}, {0, 1, -1.618033988749895}, {1.618033988749895, 0, -1},
    {1.618033988749895, 0, 1}, {-1.618033988749895, 0, -1}, {-1.618033988749895, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                          {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                          {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                          {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};

/* Random rotation from a uniform unit quaternion (Shoemake). */
void random_rotation(Pcg32& g, double R[3][3]) {
    double u1 = g.uniformd(), u2 = g.uniformd(), u3 = g.uniformd();
    double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    double qx = a * std::sin(6.283185307179586 * u2), qy = a * std::cos(6.283185307179586 * u2);
    double qz = b * std::sin(6.283185307179586 * u3), qw = b * std::cos(6.283185307179586 * u3);
    R[0][0] = 1 - 2 * (qy * qy + qz * qz); R[0][1] = 2 * (qx * qy - qz * qw); R[0][2] = 2 * (qx * qz + qy * qw);
    R[1][0] = 2 * (qx * qy + qz * qw); R[1][1] = 1 - 2 * (qx * qx + qz * qz); R[1][2] = 2 * (qy * qz - qx * qw);
    R[2][0] = 2 * (qx * qz - qy * qw); R[2][1] = 2 * (qy * qz + qx * qw); R[2][2] = 1 - 2 * (qx * qx + qy * qy);
}

}  // namespace

extern "C" {

/* Uniform small triangles: centroid U[0,1)^3, vertices centroid + U[-h,h]^3
 * with h = 0.005 (300,000-triangle workload). */
void nxbs_uniform(uint32_t n, uint64_t seed, float half_extent, nxb_triangle* out) {
    parallel_chunks(n, [&](uint64_t c, uint64_t b, uint64_t e) {
        Pcg32 g(seed, c);
        for (uint64_t i = b; i < e; ++i) {
            float cx = g.uniform(), cy = g.uniform(), cz = g.uniform();
            nxb_float3 v[3];
            for (int k = 0; k < 3; ++k) {
                v[k].x = cx + g.range(-half_extent, half_extent);
                v[k].y = cy + g.range(-half_extent, half_extent);
                v[k].z = cz + g.range(-half_extent, half_extent);
            }
            out[i] = nxb_triangle{v[0], v[1], v[2]};
        }
    });
}

/* Geodesic icosphere of radius 1 at frequency f (20*f^2 triangles, face-major,
 * row-major within a face), vertices displaced radially by 0.02 * fbm4;
 * truncated to the first n triangles. Returns the number written. */
uint32_t nxbs_icosphere(uint32_t n, uint32_t f, uint64_t seed, float amplitude, nxb_triangle* out) {
    uint64_t total = 20ull * f * f;
    if (n > total) n = (uint32_t)total;
    uint64_t per_face = (uint64_t)f * f;
    auto vert = [&](const V3& A, const V3& B, const V3& C, uint32_t i, uint32_t j) {
        V3 p = add(A, add(mul(sub(B, A), (double)i / f), mul(sub(C, A), (double)j / f)));
        p = norm(p);
        double r = 1.0 + amplitude * fbm4(p, seed);
        return f3(mul(p, r));
    };
    std::vector<uint32_t> faces((n + per_face - 1) / per_face);
    for (uint32_t k = 0; k < faces.size(); ++k) faces[k] = k;
    std::vector<std::thread> th;
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            for (size_t fi = t; fi < faces.size(); fi += nt) {
                V3 A = norm(kIco[kIcoF[fi][0]]), B = norm(kIco[kIcoF[fi][1]]), C = norm(kIco[kIcoF[fi][2]]);
                uint64_t o = fi * per_face;
                for (uint32_t i = 0; i < f; ++i) {
                    for (uint32_t j = 0; j + i < f; ++j) {
                        if (o < n) out[o] = nxb_triangle{vert(A, B, C, i, j), vert(A, B, C, i + 1, j), vert(A, B, C, i, j + 1)};
                        ++o;
                        if (j + i + 1 < f) {
                            if (o < n) out[o] = nxb_triangle{vert(A, B, C, i + 1, j), vert(A, B, C, i + 1, j + 1), vert(A, B, C, i, j + 1)};
                            ++o;
                        }
                    }
                }
            }
        });
    }
    for (auto& x : th) x.join();
    return n;
}

/* Hair-like slivers: segment p0 -> p0 + L d inside [0,1]^3, d uniform on the
 * sphere, L ~ U(0.02, 0.08); third vertex p0 + width * w with w a unit
 * vector orthogonal to d at a random angle. */
void nxbs_slivers(uint32_t n, uint64_t seed, float width, nxb_triangle* out) {
    parallel_chunks(n, [&](uint64_t c, uint64_t b, uint64_t e) {
        Pcg32 g(seed, c);
        for (uint64_t i = b; i < e; ++i) {
            double z = 2.0 * g.uniformd() - 1.0, phi = 6.283185307179586 * g.uniformd();
            double s = std::sqrt(std::max(0.0, 1.0 - z * z));
            V3 d = {s * std::cos(phi), s * std::sin(phi), z};
            double L = 0.02 + 0.06 * g.uniformd();
            V3 p0, p1;
            for (int tries = 0;; ++tries) {
                p0 = {g.uniformd(), g.uniformd(), g.uniformd()};
                p1 = add(p0, mul(d, L));
                if ((p1.x >= 0 && p1.x <= 1 && p1.y >= 0 && p1.y <= 1 && p1.z >= 0 && p1.z <= 1) || tries > 64) break;
            }
            V3 h = std::fabs(d.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
            V3 e1 = norm(cross(d, h)), e2 = cross(d, e1);
            double th = 6.283185307179586 * g.uniformd();
            V3 w = add(mul(e1, std::cos(th)), mul(e2, std::sin(th)));
            out[i] = nxb_triangle{f3(p0), f3(p1), f3(add(p0, mul(w, width)))};
        }
    });
}

/* Triangle counts of the batched workload: round(exp(U(ln lo, ln hi))). */
uint64_t nxbs_batch_counts(uint32_t meshes, uint64_t seed, uint32_t lo, uint32_t hi, uint32_t* counts) {
    Pcg32 g(seed, 0xBA7C4ull);
    uint64_t total = 0;
    double a = std::log((double)lo), b = std::log((double)hi);
    for (uint32_t m = 0; m < meshes; ++m) {
        double v = std::exp(a + (b - a) * g.uniformd());
        uint32_t c = (uint32_t)std::llround(v);
        counts[m] = std::max(lo, std::min(hi, c));
        total += counts[m];
    }
    return total;
}

/* One mesh of the batched workload: a sphere, box or torus (by mesh index
 * and seed) tessellated to at least `count` triangles and truncated to
 * exactly `count`, under a random rotation, log-uniform scale in [0.1, 10]
 * and translation in [-100, 100]^3. Returns the shape id (0/1/2). */
uint32_t nxbs_batch_mesh(uint32_t mesh, uint32_t count, uint64_t seed, nxb_triangle* out) {
    Pcg32 g(seed, 0x10000ull + mesh);
    uint32_t shape = g.next() % 3;
    double R[3][3];
    random_rotation(g, R);
    double scale = std::exp(std::log(0.1) + (std::log(10.0) - std::log(0.1)) * g.uniformd());
    V3 t = {-100 + 200 * g.uniformd(), -100 + 200 * g.uniformd(), -100 + 200 * g.uniformd()};
    auto xf = [&](V3 p) {
        V3 q = {R[0][0] * p.x + R[0][1] * p.y + R[0][2] * p.z, R[1][0] * p.x + R[1][1] * p.y + R[1][2] * p.z,
                R[2][0] * p.x + R[2][1] * p.y + R[2][2] * p.z};
        return f3(add(mul(q, scale), t));
    };
    uint32_t o = 0;
    auto emit = [&](V3 a, V3 b, V3 c) { if (o < count) out[o++] = nxb_triangle{xf(a), xf(b), xf(c)}; };
    const double PI = 3.141592653589793;
    if (shape == 0) {  /* UV sphere: slices x stacks quads, 2 triangles each */
        uint32_t sl = (uint32_t)std::ceil(std::sqrt(count / 2.0)) + 1;
        uint32_t st = (count + 2 * sl - 1) / (2 * sl) + 1;
        auto P = [&](uint32_t i, uint32_t j) {
            double th = PI * j / st, ph = 2 * PI * i / sl;
            return V3{std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
        };
        for (uint32_t j = 0; j < st && o < count; ++j)
            for (uint32_t i = 0; i < sl && o < count; ++i) {
                emit(P(i, j), P(i + 1, j), P(i + 1, j + 1));
                emit(P(i, j), P(i + 1, j + 1), P(i, j + 1));
            }
    } else if (shape == 1) {  /* box: 6 faces of g x g quads */
        uint32_t gq = (uint32_t)std::ceil(std::sqrt(count / 12.0));
        if (gq < 1) gq = 1;
        for (int face = 0; face < 6 && o < count; ++face) {
            int ax = face / 2;
            double sgn = (face & 1) ? 1.0 : -1.0;
            auto P = [&](uint32_t i, uint32_t j) {
                double u = -1 + 2.0 * i / gq, v = -1 + 2.0 * j / gq;
                V3 p;
                if (ax == 0) p = {sgn, u, v};
                else if (ax == 1) p = {u, sgn, v};
                else p = {u, v, sgn};
                return p;
            };
            for (uint32_t j = 0; j < gq && o < count; ++j)
                for (uint32_t i = 0; i < gq && o < count; ++i) {
                    emit(P(i, j), P(i + 1, j), P(i + 1, j + 1));
                    emit(P(i, j), P(i + 1, j + 1), P(i, j + 1));
                }
        }
    } else {  /* torus R=1, r=0.3 */
        uint32_t nu = (uint32_t)std::ceil(std::sqrt(count * 2.0)) + 1;
        uint32_t nv = (count + 2 * nu - 1) / (2 * nu) + 1;
        auto P = [&](uint32_t i, uint32_t j) {
            double u = 2 * PI * i / nu, v = 2 * PI * j / nv;
            return V3{(1 + 0.3 * std::cos(v)) * std::cos(u), (1 + 0.3 * std::cos(v)) * std::sin(u), 0.3 * std::sin(v)};
        };
        for (uint32_t j = 0; j < nv && o < count; ++j)
            for (uint32_t i = 0; i < nu && o < count; ++i) {
                emit(P(i, j), P(i + 1, j), P(i + 1, j + 1));
                emit(P(i, j), P(i + 1, j + 1), P(i, j + 1));
            }
    }
    return shape;
}

/* Gaussian blobs over a ground plane. The plane is a side x side grid of
 * quads of size `cell` at z = 0 (2 triangles per quad, placed last); the
 * remaining n - 2*side^2 triangles come from `blobs` Gaussian blobs with
 * sigma log-uniform in [0.001, 0.05], centres uniform over the plane at
 * heights U(0.05, 2), triangle size 0.2 sigma. Returns the plane count. */
uint32_t nxbs_blobs_plane(uint32_t n, uint32_t blobs, uint32_t side, float cell, uint64_t seed, nxb_triangle* out) {
    uint64_t plane = 2ull * side * side;
    if (plane > n) plane = n;
    uint64_t nb = n - plane;
    float extent = cell * side;
    std::vector<V3> mu(blobs);
    std::vector<double> sigma(blobs);
    Pcg32 g(seed, 0xB10Bull);
    for (uint32_t k = 0; k < blobs; ++k) {
        mu[k] = {extent * g.uniformd(), extent * g.uniformd(), 0.05 + 1.95 * g.uniformd()};
        sigma[k] = std::exp(std::log(0.001) + (std::log(0.05) - std::log(0.001)) * g.uniformd());
    }
    parallel_chunks(nb, [&](uint64_t c, uint64_t b, uint64_t e) {
        Pcg32 r(seed, c);
        for (uint64_t i = b; i < e; ++i) {
            uint32_t k = (uint32_t)((i * blobs) / (nb ? nb : 1));
            double s = sigma[k];
            V3 cen = {mu[k].x + s * r.normal(), mu[k].y + s * r.normal(), mu[k].z + s * r.normal()};
            double h = 0.2 * s;
            nxb_float3 v[3];
            for (int q = 0; q < 3; ++q)
                v[q] = f3(V3{cen.x + h * (r.uniformd() - 0.5), cen.y + h * (r.uniformd() - 0.5), cen.z + h * (r.uniformd() - 0.5)});
            out[i] = nxb_triangle{v[0], v[1], v[2]};
        }
    });
    parallel_chunks(plane, [&](uint64_t, uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t) {
            uint64_t q = t / 2;
            uint32_t i = (uint32_t)(q % side), j = (uint32_t)(q / side);
            float x0 = cell * i, y0 = cell * j, x1 = cell * (i + 1), y1 = cell * (j + 1);
            nxb_triangle tri;
            if ((t & 1) == 0) tri = nxb_triangle{{x0, y0, 0.0f}, {x1, y0, 0.0f}, {x1, y1, 0.0f}};
            else tri = nxb_triangle{{x0, y0, 0.0f}, {x1, y1, 0.0f}, {x0, y1, 0.0f}};
            out[nb + t] = tri;
        }
    });
    return (uint32_t)plane;
}

}  // extern "C"
