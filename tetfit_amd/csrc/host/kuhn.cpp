// kuhn.cpp - synthetic Kuhn (Freudenthal) tetrahedral meshes of the unit cube
// written in the netgen neutral format that ReadUnstructuredMesh reads
// (src/DGA/Discretization.hpp:5868-6130: points, "mat n1 n2 n3 n4" tets,
// "bid n1 n2 n3" boundary triangles, 1-based node ids).
#include <cmath>
#include <cstdio>
#include <vector>

#include "build.hpp"

namespace tfh {

namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};
}  // namespace

int32_t write_kuhn_mesh(const char* path, int32_t n, double jitter, uint64_t seed, int32_t block,
                        double* min_vol_ratio) {
    if (n <= 0 || block <= 0) throw Error("kuhn mesh: n and block must be positive");
    const int32_t n1 = n + 1;
    const double h = 1.0 / n;
    auto nid = [&](int i, int j, int k) { return (int64_t)i + (int64_t)n1 * ((int64_t)j + (int64_t)n1 * k); };
    const int64_t NP = (int64_t)n1 * n1 * n1;
    std::vector<double> X(NP * 3), X0(NP * 3);
    SplitMix64 rng{seed};
    for (int k = 0; k <= n; ++k)
        for (int j = 0; j <= n; ++j)
            for (int i = 0; i <= n; ++i) {
                int64_t p = nid(i, j, k);
                double c[3] = {i * h, j * h, k * h};
                if (i == n) c[0] = 1.0;
                if (j == n) c[1] = 1.0;
                if (k == n) c[2] = 1.0;
                X0[p * 3] = c[0]; X0[p * 3 + 1] = c[1]; X0[p * 3 + 2] = c[2];
                bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
                for (int a = 0; a < 3; ++a) {
                    double d = interior ? (2 * rng.uniform() - 1) * jitter * h : 0.0;
                    X[p * 3 + a] = c[a] + d;
                }
            }
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int nb = (n + block - 1) / block;
    std::vector<int64_t> tets;
    std::vector<int32_t> mats;
    tets.reserve((size_t)n * n * n * 24);
    double minr = 1e300;
    auto vol = [&](const std::vector<double>& P, const int64_t v[4]) {
        double a[3], b[3], c[3];
        for (int q = 0; q < 3; ++q) {
            a[q] = P[v[1] * 3 + q] - P[v[0] * 3 + q];
            b[q] = P[v[2] * 3 + q] - P[v[0] * 3 + q];
            c[q] = P[v[3] * 3 + q] - P[v[0] * 3 + q];
        }
        return (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
    };
    int32_t max_label = 0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int64_t lin = (int64_t)(i / block) + nb * ((int64_t)(j / block) + (int64_t)nb * (k / block));
                int32_t label = (int32_t)lin + 1;
                if (label >= 7) label += 1;  // label 7 would be dropped by the reader (D:6091)
                if (label > max_label) max_label = label;
                for (auto& pm : perms) {
                    int c[3] = {i, j, k};
                    int64_t v[4];
                    v[0] = nid(c[0], c[1], c[2]);
                    c[pm[0]]++;
                    v[1] = nid(c[0], c[1], c[2]);
                    c[pm[1]]++;
                    v[2] = nid(c[0], c[1], c[2]);
                    c[pm[2]]++;
                    v[3] = nid(c[0], c[1], c[2]);
                    double v0 = vol(X0, v), v1 = vol(X, v);
                    if (!(v0 * v1 > 0)) return -1;
                    double r = std::fabs(v1) / (h * h * h / 6.0);
                    if (r < minr) minr = r;
                    for (int q = 0; q < 4; ++q) tets.push_back(v[q]);
                    mats.push_back(label);
                }
            }
    if (min_vol_ratio) *min_vol_ratio = minr;
    // boundary triangles: faces of the tets with all three nodes on one wall
    std::vector<int64_t> tris;
    std::vector<int32_t> tlab;
    const int fidx[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    for (size_t t = 0; t < mats.size(); ++t) {
        for (auto& f : fidx) {
            int64_t v[3] = {tets[t * 4 + f[0]], tets[t * 4 + f[1]], tets[t * 4 + f[2]]};
            int ijk[3][3];
            for (int q = 0; q < 3; ++q) {
                ijk[q][0] = (int)(v[q] % n1);
                ijk[q][1] = (int)((v[q] / n1) % n1);
                ijk[q][2] = (int)(v[q] / ((int64_t)n1 * n1));
            }
            int label = 0;
            for (int a = 0; a < 3 && !label; ++a) {
                for (int side = 0; side < 2 && !label; ++side) {
                    int want = side ? n : 0;
                    if (ijk[0][a] == want && ijk[1][a] == want && ijk[2][a] == want) {
                        // z-min=1, x-min=2, y-max=3, x-max=4, y-min=5, z-max=6
                        static const int lab[3][2] = {{2, 4}, {5, 3}, {1, 6}};
                        label = lab[a][side];
                    }
                }
            }
            if (label) {
                for (int q = 0; q < 3; ++q) tris.push_back(v[q]);
                tlab.push_back(label);
            }
        }
    }
    FILE* fp = std::fopen(path, "w");
    if (!fp) throw Error(std::string("cannot write ") + path);
    std::fprintf(fp, "%lld\n", (long long)NP);
    for (int64_t p = 0; p < NP; ++p)
        std::fprintf(fp, "  %.17g  %.17g  %.17g\n", X[p * 3], X[p * 3 + 1], X[p * 3 + 2]);
    std::fprintf(fp, "%lld\n", (long long)mats.size());
    for (size_t t = 0; t < mats.size(); ++t)
        std::fprintf(fp, "%d %lld %lld %lld %lld\n", mats[t], (long long)tets[t * 4] + 1, (long long)tets[t * 4 + 1] + 1,
                     (long long)tets[t * 4 + 2] + 1, (long long)tets[t * 4 + 3] + 1);
    std::fprintf(fp, "%lld\n", (long long)tlab.size());
    for (size_t t = 0; t < tlab.size(); ++t)
        std::fprintf(fp, "%d %lld %lld %lld\n", tlab[t], (long long)tris[t * 3] + 1, (long long)tris[t * 3 + 1] + 1,
                     (long long)tris[t * 3 + 2] + 1);
    std::fclose(fp);
    return max_label;
}

}  // namespace tfh
