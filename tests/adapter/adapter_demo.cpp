// A C++ caller of the engine through include/scalfmm_hip_algorithm.hpp, with
// tree types of its own that lay out their blocks as the reference's do:
//   GroupTree      particle blocks (GroupTree/Core/FGroupOfParticles.hpp:19-200:
//                  64-byte header, 24-byte leaf headers, x y z q arrays from the
//                  next 64-byte boundary, 8-particle packs per leaf) and cell
//                  blocks per level (GroupTree/Core/FGroupOfCells.hpp:125-168:
//                  24-byte header, Morton indexes, 24-byte cells; FChebCellPOD
//                  buffers of 2 p^3 doubles per cell), built as the FGroupTree
//                  particle constructor builds them (GroupTree/Core/FGroupTree.hpp:
//                  178-356)
//   Octree         leaves in Morton order holding indexed containers (the shape
//                  of FOctree + FP2PParticleContainerIndexed that hip_execute uses)
// Usage: adapter_demo <particles.bfma> <output.bin>.  Output (doubles unless
// noted): int64 n, int64 cells of level 3, then the group tree's results
// (pot, fx, fy, fz, input order), the same after the split executions
// execute(P2M | M2M), execute(M2L | L2L | L2P | P2P), the classic octree's
// results, the level-3 cell keys (int64), their multipoles and locals
// (p^3 values per cell, block order).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "scalfmm_hip_algorithm.hpp"

namespace {

constexpr int kHeight = 5, kOrder = 6, kClassicOrder = 7, kPerBlock = 250;

struct Point {
    double v[3];
    double getX() const { return v[0]; }
    double getY() const { return v[1]; }
    double getZ() const { return v[2]; }
};

int64_t morton(const int64_t c[3], int levels) {
    int64_t k = 0;
    for (int b = 0; b < levels; ++b)
        k |= (((c[0] >> b) & 1) << (3 * b + 2)) | (((c[1] >> b) & 1) << (3 * b + 1)) | (((c[2] >> b) & 1) << (3 * b));
    return k;
}

// leaf coordinate of FCoordinateComputer (truncation, the upper face in the last leaf)
int64_t leaf_key(const double p[3], const Point& center, double width, int height) {
    const int64_t leaves = int64_t(1) << (height - 1);
    const double leafw = width / static_cast<double>(leaves);
    int64_t c[3];
    for (int d = 0; d < 3; ++d) {
        const double rel = p[d] - (center.v[d] - width / 2);
        c[d] = rel == width ? leaves - 1 : static_cast<int64_t>(rel / leafw);
    }
    return morton(c, height - 1);
}

class ParticleGroup {
    std::vector<unsigned char> raw_;
    std::vector<double> attributes_;

public:
    int64_t nalloc = 0;
    std::vector<int64_t> leaf_offset;  // offset of each leaf's first particle in the arrays

    ParticleGroup(const std::vector<int64_t>& keys, const std::vector<int64_t>& counts) {
        const int64_t nleaves = static_cast<int64_t>(keys.size());
        const int64_t np = std::accumulate(counts.begin(), counts.end(), int64_t(0));
        auto round_up = [](int64_t m) { return m + (8 - m % 8); };  // RoundToUpperParticles
        nalloc = round_up(np + 7 * nleaves);
        const int64_t off_pos = (64 + 24 * nleaves + 63) / 64 * 64;
        raw_.assign(static_cast<size_t>(off_pos + nalloc * 4 * 8), 0);
        int64_t hdr[8] = {keys.front(), keys.back() + 1, nleaves, nalloc, off_pos, 8 * nalloc, 8 * nalloc, np};
        std::memcpy(raw_.data(), hdr, 16);
        const int32_t nl32[2] = {static_cast<int32_t>(nleaves), 0};
        std::memcpy(raw_.data() + 16, nl32, 8);
        std::memcpy(raw_.data() + 24, hdr + 3, 40);
        int64_t off = 0;
        for (int64_t i = 0; i < nleaves; ++i) {
            const int64_t leaf[3] = {keys[i], counts[i], off};
            std::memcpy(raw_.data() + 64 + 24 * i, leaf, 24);
            leaf_offset.push_back(off);
            off = round_up(off + counts[i]);
        }
        attributes_.assign(static_cast<size_t>(4 * nalloc), 0.0);
    }
    double* array(int k) { return reinterpret_cast<double*>(raw_.data() + header(4)) + k * nalloc; }
    int64_t header(int i) const {
        int64_t v;
        std::memcpy(&v, raw_.data() + 24 + 8 * (i - 3), 8);
        return v;
    }
    const unsigned char* getRawBuffer() const { return raw_.data(); }
    unsigned char* getRawAttributesBuffer() { return reinterpret_cast<unsigned char*>(attributes_.data()); }
    const double* attribute(int a) const { return attributes_.data() + a * nalloc; }
};

class CellGroup {
    std::vector<unsigned char> raw_;
    std::vector<double> multipoles_, locals_;

public:
    std::vector<int64_t> keys;
    CellGroup(const std::vector<int64_t>& k, int level, int order) : keys(k) {
        const int64_t n = static_cast<int64_t>(k.size());
        raw_.assign(static_cast<size_t>(24 + 8 * n + 24 * n), 0);
        const int64_t se[2] = {k.front(), k.back() + 1};
        std::memcpy(raw_.data(), se, 16);
        const int32_t nn = static_cast<int32_t>(n);
        std::memcpy(raw_.data() + 16, &nn, 4);
        std::memcpy(raw_.data() + 24, k.data(), 8 * n);
        for (int64_t i = 0; i < n; ++i) {  // FBasicCellPOD: key, coordinates from the key, level 0
            int32_t pod[4] = {0, 0, 0, 0};
            for (int b = 0; b < level; ++b)
                for (int d = 0; d < 3; ++d) pod[d] |= static_cast<int32_t>(((k[i] >> (3 * b + 2 - d)) & 1) << b);
            std::memcpy(raw_.data() + 24 + 8 * n + 24 * i, &k[i], 8);
            std::memcpy(raw_.data() + 24 + 8 * n + 24 * i + 8, pod, 16);
        }
        const size_t vec = 2 * static_cast<size_t>(order) * order * order;
        multipoles_.assign(n * vec, 0.0);
        locals_.assign(n * vec, 0.0);
    }
    const unsigned char* getRawBuffer() const { return raw_.data(); }
    double* getRawMultipoleBuffer() { return multipoles_.data(); }
    double* getRawLocalBuffer() { return locals_.data(); }
};

class GroupTree {
    int height_;
    double width_;
    Point center_;
    std::vector<ParticleGroup> particle_groups_;
    std::vector<std::vector<CellGroup>> cell_groups_;  // per level

public:
    std::vector<int64_t> original;  // per block particle (block order) -> input index
    std::vector<std::pair<int, int64_t>> where;  // input index -> (block, array position)

    GroupTree(int height, double width, Point center, int per_block, int order, const std::vector<double>& x,
              const std::vector<double>& y, const std::vector<double>& z, const std::vector<double>& q)
        : height_(height), width_(width), center_(center), cell_groups_(height) {
        const int64_t n = static_cast<int64_t>(x.size());
        std::vector<int64_t> key(n);
        for (int64_t i = 0; i < n; ++i) {
            const double p[3] = {x[i], y[i], z[i]};
            key[i] = leaf_key(p, center, width, height);
        }
        original.resize(n);
        std::iota(original.begin(), original.end(), 0);
        std::stable_sort(original.begin(), original.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
        std::vector<int64_t> lkeys, lstart;
        for (int64_t s = 0; s < n; ++s)
            if (s == 0 || key[original[s]] != key[original[s - 1]]) {
                lkeys.push_back(key[original[s]]);
                lstart.push_back(s);
            }
        lstart.push_back(n);
        where.resize(n);
        const int64_t nl = static_cast<int64_t>(lkeys.size());
        for (int64_t b0 = 0; b0 < nl; b0 += per_block) {
            const int64_t b1 = std::min(nl, b0 + per_block);
            std::vector<int64_t> k(lkeys.begin() + b0, lkeys.begin() + b1), cnt;
            for (int64_t l = b0; l < b1; ++l) cnt.push_back(lstart[l + 1] - lstart[l]);
            ParticleGroup g(k, cnt);
            for (int64_t l = b0; l < b1; ++l)
                for (int64_t s = lstart[l]; s < lstart[l + 1]; ++s) {
                    const int64_t e = g.leaf_offset[l - b0] + (s - lstart[l]);
                    const int64_t i = original[s];
                    g.array(0)[e] = x[i];
                    g.array(1)[e] = y[i];
                    g.array(2)[e] = z[i];
                    g.array(3)[e] = q[i];
                    where[i] = {static_cast<int>(particle_groups_.size()), e};
                }
            particle_groups_.push_back(std::move(g));
        }
        std::vector<int64_t> level_keys = lkeys;
        for (int level = height - 1; level >= 1; --level) {
            for (size_t b0 = 0; b0 < level_keys.size(); b0 += per_block) {
                const size_t b1 = std::min(level_keys.size(), b0 + per_block);
                cell_groups_[level].emplace_back(
                    std::vector<int64_t>(level_keys.begin() + b0, level_keys.begin() + b1), level, order);
            }
            std::vector<int64_t> parents;
            for (int64_t k : level_keys)
                if (parents.empty() || parents.back() != (k >> 3)) parents.push_back(k >> 3);
            level_keys.swap(parents);
        }
    }
    int getHeight() const { return height_; }
    double getBoxWidth() const { return width_; }
    Point getBoxCenter() const { return center_; }
    int getNbParticleGroup() const { return static_cast<int>(particle_groups_.size()); }
    ParticleGroup* getParticleGroup(int i) { return &particle_groups_[i]; }
    int getNbCellGroupAtLevel(int level) const { return static_cast<int>(cell_groups_[level].size()); }
    CellGroup* getCellGroup(int level, int i) { return &cell_groups_[level][i]; }

    void results(std::vector<double>& out) const {  // pot, fx, fy, fz in input order
        const int64_t n = static_cast<int64_t>(where.size());
        out.assign(4 * n, 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < 4; ++a) out[a * n + i] = particle_groups_[where[i].first].attribute(a)[where[i].second];
    }
};

struct Container {
    std::vector<double> pos[3], phys, pot, fx, fy, fz;
    std::vector<int64_t> index;
    int64_t getNbParticles() const { return static_cast<int64_t>(phys.size()); }
    const double* const* getPositions() const {
        static thread_local const double* p[3];
        for (int d = 0; d < 3; ++d) p[d] = pos[d].data();
        return p;
    }
    const double* getPhysicalValues() const { return phys.data(); }
    double* getPotentials() { return pot.data(); }
    double* getForcesX() { return fx.data(); }
    double* getForcesY() { return fy.data(); }
    double* getForcesZ() { return fz.data(); }
};

struct Leaf {
    Container c;
    Container* getTargets() { return &c; }
};

class Octree {
    int height_;
    double width_;
    Point center_;
    std::vector<std::pair<int64_t, Leaf>> leaves_;  // Morton order

public:
    Octree(int height, double width, Point center) : height_(height), width_(width), center_(center) {}
    void insert(const double p[3], int64_t index, double q) {  // FOctree::insert with an indexed container
        const int64_t k = leaf_key(p, center_, width_, height_);
        auto it = std::lower_bound(leaves_.begin(), leaves_.end(), k,
                                   [](const std::pair<int64_t, Leaf>& a, int64_t key) { return a.first < key; });
        if (it == leaves_.end() || it->first != k) it = leaves_.insert(it, {k, Leaf{}});
        Container& c = it->second.c;
        for (int d = 0; d < 3; ++d) c.pos[d].push_back(p[d]);
        c.phys.push_back(q);
        c.index.push_back(index);
        for (auto* v : {&c.pot, &c.fx, &c.fy, &c.fz}) v->push_back(0.0);
    }
    int getHeight() const { return height_; }
    template <class F>
    void forEachLeaf(F fn) {
        for (auto& l : leaves_) fn(&l.second);
    }
    void results(std::vector<double>& out, int64_t n) {
        out.assign(4 * n, 0.0);
        forEachLeaf([&](Leaf* leaf) {
            const Container& c = leaf->c;
            for (size_t i = 0; i < c.index.size(); ++i) {
                out[c.index[i]] = c.pot[i];
                out[n + c.index[i]] = c.fx[i];
                out[2 * n + c.index[i]] = c.fy[i];
                out[3 * n + c.index[i]] = c.fz[i];
            }
        });
    }
};

template <int ORDER>
using Kernel = scalfmm_hip::FHipChebSymKernel<double, void, Container, void, ORDER>;

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s particles.bfma output.bin\n", argv[0]);
        return 2;
    }
    try {
        int64_t n = 0;
        int32_t nv = 0;
        double width = 0, c[3];
        if (hfmm_fma_read(argv[1], &n, &nv, &width, c, nullptr) != HFMM_OK) throw std::runtime_error("read header");
        std::vector<double> data(static_cast<size_t>(n) * nv);
        if (hfmm_fma_read(argv[1], &n, &nv, &width, c, data.data()) != HFMM_OK) throw std::runtime_error("read data");
        std::vector<double> x(n), y(n), z(n), q(n);
        for (int64_t i = 0; i < n; ++i) {
            x[i] = data[i * nv];
            y[i] = data[i * nv + 1];
            z[i] = data[i * nv + 2];
            q[i] = data[i * nv + 3];
        }
        const Point center{{c[0], c[1], c[2]}};
        using namespace scalfmm_hip;

        // group tree: one execution
        GroupTree tree(kHeight, width, center, kPerBlock, kOrder, x, y, z, q);
        Kernel<kOrder> kernel(kHeight, width, center, static_cast<const void*>(nullptr));
        FGroupHipAlgorithm<GroupTree, Kernel<kOrder>> algo(&tree, &kernel);
        algo.execute();
        std::vector<double> full, split, classic;
        tree.results(full);

        // group tree: split executions on another tree
        GroupTree tree2(kHeight, width, center, kPerBlock, kOrder, x, y, z, q);
        Kernel<kOrder> kernel2(kHeight, width, center, static_cast<const void*>(nullptr));
        FGroupHipAlgorithm<GroupTree, Kernel<kOrder>> algo2(&tree2, &kernel2);
        algo2.execute(OpP2M | OpM2M);
        algo2.execute(OpM2L | OpL2L | OpL2P | OpP2P);
        tree2.results(split);

        // classic octree
        Octree oct(kHeight, width, center);
        for (int64_t i = 0; i < n; ++i) {
            const double p[3] = {x[i], y[i], z[i]};
            oct.insert(p, i, q[i]);
        }
        Kernel<kClassicOrder> ckernel(kHeight, width, center, static_cast<const void*>(nullptr));
        hip_execute(oct, ckernel);
        oct.results(classic, n);

        // level-3 cells of the first tree, block order
        std::vector<int64_t> keys;
        std::vector<double> M, L;
        const int nn = kOrder * kOrder * kOrder;
        for (int b = 0; b < tree.getNbCellGroupAtLevel(3); ++b) {
            CellGroup* g = tree.getCellGroup(3, b);
            for (size_t i = 0; i < g->keys.size(); ++i) {
                keys.push_back(g->keys[i]);
                M.insert(M.end(), g->getRawMultipoleBuffer() + 2 * nn * i, g->getRawMultipoleBuffer() + 2 * nn * i + nn);
                L.insert(L.end(), g->getRawLocalBuffer() + 2 * nn * i, g->getRawLocalBuffer() + 2 * nn * i + nn);
            }
        }
        FILE* f = std::fopen(argv[2], "wb");
        if (!f) throw std::runtime_error("cannot write the output");
        const int64_t head[2] = {n, static_cast<int64_t>(keys.size())};
        std::fwrite(head, sizeof(int64_t), 2, f);
        for (const auto* v : {&full, &split, &classic}) std::fwrite(v->data(), sizeof(double), v->size(), f);
        std::fwrite(keys.data(), sizeof(int64_t), keys.size(), f);
        std::fwrite(M.data(), sizeof(double), M.size(), f);
        std::fwrite(L.data(), sizeof(double), L.size(), f);
        std::fclose(f);
        std::printf("adapter_demo: n %lld, level-3 cells %lld\n", static_cast<long long>(n),
                    static_cast<long long>(keys.size()));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "adapter_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}
