// matrix_io.hpp -- the reference's binary volume format (Glioma/Matrix.h:
// 71-241, 593-665): int32 header {1234, dim, nx, ny[, nz]}, int32 type id
// (0 double, 1 float, 2 int), then nx*ny[*nz] values, x fastest.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace gq {
namespace io {

struct Volume {
    int dim = 0;
    int n[3] = {0, 0, 1};
    std::vector<double> v;  // x fastest
    size_t size() const { return (size_t)n[0] * n[1] * n[2]; }
    double operator()(size_t i, size_t j, size_t k = 0) const { return v[i + (size_t)n[0] * (j + (size_t)n[1] * k)]; }
};

inline Volume read_matrix(const std::string &path, int dim)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    Volume m;
    int32_t h[2];
    if (std::fread(h, 4, 2, f) != 2 || h[0] != 1234) {
        std::fclose(f);
        throw std::runtime_error(path + ": magic number did not match");
    }
    if (h[1] != dim) {
        std::fclose(f);
        throw std::runtime_error(path + ": dimensions of matrix do not match");
    }
    m.dim = dim;
    int32_t sz[3] = {1, 1, 1};
    int32_t tid = -1;
    if (std::fread(sz, 4, dim, f) != (size_t)dim || std::fread(&tid, 4, 1, f) != 1) {
        std::fclose(f);
        throw std::runtime_error(path + ": truncated header");
    }
    for (int a = 0; a < 3; a++) m.n[a] = a < dim ? sz[a] : 1;
    if (m.n[0] <= 0 || m.n[1] <= 0 || m.n[2] <= 0) {
        std::fclose(f);
        throw std::runtime_error(path + ": bad size");
    }
    const size_t n = m.size();
    m.v.resize(n);
    bool ok = true;
    if (tid == 0) ok = std::fread(m.v.data(), 8, n, f) == n;
    else if (tid == 1) {
        std::vector<float> t(n);
        ok = std::fread(t.data(), 4, n, f) == n;
        for (size_t i = 0; i < n; i++) m.v[i] = t[i];
    } else if (tid == 2) {
        std::vector<int32_t> t(n);
        ok = std::fread(t.data(), 4, n, f) == n;
        for (size_t i = 0; i < n; i++) m.v[i] = t[i];
    } else ok = false;
    std::fclose(f);
    if (!ok) throw std::runtime_error(path + ": bad type id or truncated data");
    return m;
}

inline void write_matrix3d(const std::string &path, const double *v, int nx, int ny, int nz)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    const int32_t h[6] = {1234, 3, nx, ny, nz, 0};
    const size_t n = (size_t)nx * ny * nz;
    const bool ok = std::fwrite(h, 4, 6, f) == 6 && std::fwrite(v, 8, n, f) == n;
    std::fclose(f);
    if (!ok) throw std::runtime_error("short write " + path);
}

}  // namespace io
}  // namespace gq
