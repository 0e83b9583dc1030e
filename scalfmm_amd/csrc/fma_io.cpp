// fma_io.cpp -- the reference's particle file formats and its M2L operator
// cache file.
//
//   .fma / .bfma   Src/Files/FFmaGenericLoader.hpp:34-850 (FFmaGenericLoader,
//                  FFmaGenericWriter): a header (element size, values per
//                  particle, particle count, half box width, box centre), then
//                  the values of every particle; ".bfma" in the name selects
//                  binary, else ".fma" selects text (whitespace-separated, the
//                  writer prints 10 significant digits).
//   operator cache Src/Kernels/Chebyshev/FChebSymM2LHandler.hpp:471-567: for
//                  every stored operator in increasing index, int index, int
//                  rank, U then V (rank * p^3 FReal each, column-major), in a
//                  file named sym2l_<d|f>_o<ORDER>_e<epsilon>.bin.
// The reference exits the process on a bad file; these return HFMM_ERR_IO.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hfmm_internal.hpp"

namespace {

enum class FmaKind { kText, kBinary, kNone };

FmaKind kind_of(const char* path) {
    const std::string s(path);
    if (s.find(".bfma") != std::string::npos) return FmaKind::kBinary;
    if (s.find(".fma") != std::string::npos) return FmaKind::kText;
    return FmaKind::kNone;
}

}  // namespace

extern "C" {

int hfmm_fma_read(const char* path, int64_t* n, int32_t* nvals, double* box_width, double* center, double* data) {
    if (!path || !n || !nvals) return HFMM_ERR_ARG;
    const FmaKind k = kind_of(path);
    if (k == FmaKind::kNone) return HFMM_ERR_ARG;
    uint32_t type[2] = {0, 0};
    int64_t count = 0;
    double half = 0., ctr[3] = {0., 0., 0.};
    if (k == FmaKind::kBinary) {
        std::ifstream f(path, std::ios::in | std::ios::binary);
        if (!f.is_open()) return HFMM_ERR_IO;
        f.read(reinterpret_cast<char*>(type), sizeof(type));
        if (!f || type[0] != sizeof(double) || type[1] < 4) return HFMM_ERR_IO;
        f.read(reinterpret_cast<char*>(&count), sizeof(count));
        f.read(reinterpret_cast<char*>(&half), sizeof(half));
        f.read(reinterpret_cast<char*>(ctr), sizeof(ctr));
        if (!f || count < 0) return HFMM_ERR_IO;
        if (data) {
            f.read(reinterpret_cast<char*>(data), static_cast<std::streamsize>(sizeof(double) * count * type[1]));
            if (!f) return HFMM_ERR_IO;
        }
    } else {
        std::ifstream f(path, std::ios::in);
        if (!f.is_open()) return HFMM_ERR_IO;
        f >> type[0] >> type[1] >> count >> half >> ctr[0] >> ctr[1] >> ctr[2];
        if (!f || type[1] < 4 || count < 0) return HFMM_ERR_IO;
        if (data) {
            const int64_t m = count * static_cast<int64_t>(type[1]);
            for (int64_t i = 0; i < m; ++i) f >> data[i];
            if (f.fail()) return HFMM_ERR_IO;
        }
    }
    *n = count;
    *nvals = static_cast<int32_t>(type[1]);
    if (box_width) *box_width = 2. * half;  // the header holds the half width
    if (center)
        for (int d = 0; d < 3; ++d) center[d] = ctr[d];
    return HFMM_OK;
}

int hfmm_fma_write(const char* path, int64_t n, int32_t nvals, double box_width, const double* center,
                   const double* data) {
    if (!path || n < 0 || nvals < 4 || !center || (n > 0 && !data)) return HFMM_ERR_ARG;
    const FmaKind k = kind_of(path);
    if (k == FmaKind::kNone) return HFMM_ERR_ARG;
    const uint32_t type[2] = {static_cast<uint32_t>(sizeof(double)), static_cast<uint32_t>(nvals)};
    const double half = box_width * 0.5;
    if (k == FmaKind::kBinary) {
        std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
        if (!f.is_open()) return HFMM_ERR_IO;
        f.write(reinterpret_cast<const char*>(type), sizeof(type));
        f.write(reinterpret_cast<const char*>(&n), sizeof(n));
        f.write(reinterpret_cast<const char*>(&half), sizeof(half));
        f.write(reinterpret_cast<const char*>(center), sizeof(double) * 3);
        f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(sizeof(double) * n * nvals));
        if (!f) return HFMM_ERR_IO;
    } else {
        std::ofstream f(path, std::ios::out | std::ios::trunc);
        if (!f.is_open()) return HFMM_ERR_IO;
        f.precision(10);
        f << type[0] << "   " << type[1] << '\n';
        f << n << "   " << half << "   " << center[0] << "  " << center[1] << " " << center[2] << '\n';
        int64_t e = 0;
        for (int64_t i = 0; i < n; ++i) {
            for (int32_t j = 0; j < nvals; ++j, ++e) f << data[e] << "    ";
            f << '\n';
        }
        if (!f) return HFMM_ERR_IO;
    }
    return HFMM_OK;
}

int hfmm_operator_file_name(int32_t order, double epsilon, char* buf, int32_t size) {
    if (!buf || size <= 0 || order < 5 || order > 8) return HFMM_ERR_ARG;
    const double eps = epsilon > 0 ? epsilon : [&] {
        double e = 1.;
        for (int i = 0; i < order; ++i) e /= 10.;
        return e;
    }();
    std::ostringstream s;
    s << "sym2l_d_o" << order << "_e" << eps << ".bin";
    const std::string name = s.str();
    if (static_cast<int32_t>(name.size()) + 1 > size) return HFMM_ERR_ARG;
    std::memcpy(buf, name.c_str(), name.size() + 1);
    return HFMM_OK;
}

}  // extern "C"

namespace hfmm {

int save_operators(Context* c, const char* path) {
    const ChebOperators& op = c->ops;
    std::ofstream f(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f.is_open()) return fail(c, HFMM_ERR_IO, std::string("cannot open ") + path);
    for (int idx = 0; idx < 343; ++idx) {
        const int r = op.rank[idx];
        if (!r) continue;
        f.write(reinterpret_cast<const char*>(&idx), sizeof(int));
        f.write(reinterpret_cast<const char*>(&r), sizeof(int));
        f.write(reinterpret_cast<const char*>(op.K[idx].data()), sizeof(double) * 2 * r * op.nn);
    }
    if (!f) return fail(c, HFMM_ERR_IO, std::string("write failed: ") + path);
    return HFMM_OK;
}

int read_operators(Context* c, const char* path, std::array<int, 343>& rank,
                   std::array<std::vector<double>, 343>& K) {
    const int nn = c->ops.nn;
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f.is_open()) return fail(c, HFMM_ERR_IO, std::string("cannot open ") + path);
    rank.fill(0);
    for (auto& k : K) k.clear();
    int count = 0;
    for (;;) {
        int idx = 0, r = 0;
        if (!f.read(reinterpret_cast<char*>(&idx), sizeof(int))) break;
        if (!f.read(reinterpret_cast<char*>(&r), sizeof(int)) || idx < 0 || idx >= 343 || r <= 0 || r > nn ||
            rank[idx])
            return fail(c, HFMM_ERR_IO, "operator file: bad record");
        K[idx].resize(static_cast<size_t>(2) * r * nn);
        if (!f.read(reinterpret_cast<char*>(K[idx].data()), static_cast<std::streamsize>(sizeof(double) * K[idx].size())))
            return fail(c, HFMM_ERR_IO, "operator file: truncated");
        rank[idx] = r;
        ++count;
    }
    // the 16 canonical positions (i in {2,3}, 0 <= j <= i, 0 <= k <= j) and no others
    int expect = 0;
    for (int i = 2; i <= 3; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k <= j; ++k) {
                if (!rank[(i + 3) * 49 + (j + 3) * 7 + (k + 3)])
                    return fail(c, HFMM_ERR_IO, "operator file: a canonical operator is missing");
                ++expect;
            }
    if (count != expect) return fail(c, HFMM_ERR_IO, "operator file: operators at non-canonical positions");
    return HFMM_OK;
}

}  // namespace hfmm
