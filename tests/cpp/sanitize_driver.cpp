// sanitize_driver.cpp -- runs the library's host-side parsers and mesh
// routines on mutated inputs under AddressSanitizer and UBSan (built with g++
// -fsanitize=address,undefined by tests/test_sanitize.py; no GPU code).
//
//   sanitize_driver h5   <file> <n> <seed> <tmpdir>   byte flips and cuts of an HDF5 file
//   sanitize_driver gmsh <file> <n> <seed> <tmpdir>   byte flips, line cuts and digit edits of a .msh
//   sanitize_driver pair <n> <seed>                   face building and both periodic pairings
//
// Every call must end in success or an error return; the sanitizers abort the
// process on any out-of-bounds access, use after free, leak or undefined
// behaviour.  Prints one line of counts.
#include "zhijian_amd.h"
#include "zj_h5.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

namespace {

std::vector<char> slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void spit(const std::string& path, const std::vector<char>& b) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    f.write(b.data(), (std::streamsize)b.size());
}

int read_h5(const std::string& path) {
    zj::h5::Reader r;
    std::string err;
    if (!r.open(path, err)) return 0;
    for (const auto& a : r.attributeNames()) {
        double d; int64_t i; std::string s;
        (void)r.attribute(a, d, err); (void)r.attribute(a, i, err); (void)r.attribute(a, s, err);
    }
    int ok = 1;
    for (const auto& n : r.datasetNames()) {
        zj::h5::DatasetInfo info;
        if (!r.dataset(n, info, err) || info.count > (1u << 24)) continue;
        std::vector<double> d(info.count);
        std::vector<int32_t> i(info.count);
        ok += r.read(n, d.data(), info.count, err);
        (void)r.read(n, i.data(), info.count, err);
    }
    return ok;
}

int read_msh(const std::string& path) {
    zj_gmsh_handle g = nullptr;
    if (zj_gmsh_read(path.c_str(), &g) != ZJ_OK) return 0;
    int64_t nn = 0, np = 0;
    int ne = 0, mx = 0, ng = 0;
    zj_gmsh_counts(g, &nn, &ne, &mx, &ng);
    std::vector<double> xy((size_t)nn * 2 + 1);
    zj_gmsh_nodes(g, xy.data());
    std::vector<int> kind(ne + 1), order(ne + 1), cnt(ne + 1);
    std::vector<int64_t> nodes((size_t)ne * mx + 1);
    zj_gmsh_elements(g, kind.data(), order.data(), cnt.data(), nodes.data(), mx);
    zj_gmsh_periodic(g, &np, nullptr, nullptr);
    std::vector<int64_t> s((size_t)np + 1), m((size_t)np + 1);
    zj_gmsh_periodic(g, &np, s.data(), m.data());
    for (int i = 0; i < ng; ++i) {
        int tag, type;
        char name[64];
        zj_gmsh_group(g, i, &tag, &type, name, (int)sizeof(name));
    }
    int tag;
    if (ne > 0 && cnt[0] >= 2) zj_gmsh_edge_tag(g, nodes[0], nodes[1], &tag);
    zj_gmsh_free(g);
    return 1;
}

int pair_case(int n, std::mt19937_64& rng) {
    // n x n quads, lattice links in both directions plus a few random ones
    const int nx = n, ny = n + (int)(rng() % 3);
    std::vector<double> xy;
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) { xy.push_back(i + 0.3 * j); xy.push_back(j); }
    std::vector<int64_t> en;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const int64_t a = j * (nx + 1) + i;
            en.insert(en.end(), {a, a + 1, a + nx + 2, a + nx + 1});
        }
    const int E = nx * ny;
    std::vector<int> le(4 * E), ll(4 * E), re(4 * E), rl(4 * E);
    int nf = 0;
    if (zj_build_faces(E, 4, 4, en.data(), &nf, le.data(), ll.data(), re.data(), rl.data())) return 0;
    std::vector<int64_t> sl, ms;
    for (int j = 0; j <= ny; ++j) { sl.push_back(j * (nx + 1) + nx); ms.push_back(j * (nx + 1)); }
    for (int i = 0; i <= nx; ++i) { sl.push_back(ny * (nx + 1) + i); ms.push_back(i); }
    const int extra = (int)(rng() % 4);
    const int64_t N = (int64_t)xy.size() / 2;
    for (int k = 0; k < extra; ++k) { sl.push_back((int64_t)(rng() % (N + 2)) - 1); ms.push_back((int64_t)(rng() % N)); }
    std::vector<int> a = le, b = ll, c = re, d = rl;
    int nf2 = nf;
    const int rc1 = zj_pair_periodic_nodes(&nf2, a.data(), b.data(), c.data(), d.data(), 4, 4, E, en.data(), xy.data(), N,
                                           (int64_t)sl.size(), sl.data(), ms.data(), 1e-9);
    int nf3 = nf;
    const int rc2 = zj_pair_periodic(&nf3, le.data(), ll.data(), re.data(), rl.data(), 4, 4, en.data(), xy.data(),
                                     (double)nx, 0.0, 1e-9);
    return (rc1 == ZJ_OK) + (rc2 == ZJ_OK);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "pair" && argc >= 4) {
        std::mt19937_64 rng(std::strtoull(argv[3], nullptr, 10));
        int ok = 0;
        for (int k = 0; k < std::atoi(argv[2]); ++k) ok += pair_case(1 + k % 7, rng);
        std::printf("{\"mode\": \"pair\", \"ok\": %d}\n", ok);
        return 0;
    }
    if (argc < 6) return 2;
    const std::string src = argv[2], tmp = std::string(argv[5]) + "/mut";
    const int n = std::atoi(argv[3]);
    std::mt19937_64 rng(std::strtoull(argv[4], nullptr, 10));
    const std::vector<char> base = slurp(src);
    if (base.empty()) return 3;
    int opened = 0;
    auto run = [&](const std::vector<char>& b) {
        spit(tmp, b);
        opened += mode == "h5" ? read_h5(tmp) : read_msh(tmp);
    };
    run(base);
    const int first = opened;
    for (int k = 0; k < n; ++k) {
        std::vector<char> b = base;
        const int kind = (int)(rng() % 4);
        if (kind == 0) {                       // flip 1-4 bytes
            for (int f = 1 + (int)(rng() % 4); f > 0; --f) b[rng() % b.size()] ^= (char)(1 + rng() % 255);
        } else if (kind == 1) {                // cut the tail
            b.resize(rng() % b.size());
        } else if (kind == 2 && mode == "gmsh") {   // replace a digit by another digit or a '-'
            for (int f = 1 + (int)(rng() % 3); f > 0; --f) {
                const size_t at = rng() % b.size();
                if (b[at] >= '0' && b[at] <= '9') b[at] = (rng() % 5 == 0) ? '-' : (char)('0' + rng() % 10);
            }
        } else {                               // drop a span of bytes (whole lines, mostly)
            const size_t at = rng() % b.size(), len = 1 + rng() % 64;
            b.erase(b.begin() + (long)at, b.begin() + (long)std::min(b.size(), at + len));
        }
        run(b);
    }
    std::printf("{\"mode\": \"%s\", \"base_ok\": %d, \"opened\": %d, \"runs\": %d}\n", mode.c_str(), first, opened,
                n + 1);
    return 0;
}
