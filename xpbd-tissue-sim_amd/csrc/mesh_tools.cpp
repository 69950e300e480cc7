#include "mesh_tools.h"

#include <algorithm>
#include <array>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace xpbd {

void grid_mesh(int nx, int ny, int nz, double hx, double hy, double hz,
               const double origin[3], std::vector<double>& x, std::vector<int32_t>& tets)
{
    const int vx = nx + 1, vy = ny + 1, vz = nz + 1;
    x.resize(3ull * vx * vy * vz);
    for (int k = 0; k < vz; k++)
        for (int j = 0; j < vy; j++)
            for (int i = 0; i < vx; i++) {
                const size_t id = i + (size_t)vx * (j + (size_t)vy * k);
                x[3 * id + 0] = origin[0] + i * hx;
                x[3 * id + 1] = origin[1] + j * hy;
                x[3 * id + 2] = origin[2] + k * hz;
            }
    static const int perm[6][3] = { {0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0} };
    tets.resize(24ull * nx * ny * nz);
    size_t t = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                for (int p = 0; p < 6; p++) {
                    int c[3] = { i, j, k };
                    auto vid = [&](const int* cc) {
                        return (int32_t)(cc[0] + vx * (cc[1] + vy * cc[2]));
                    };
                    tets[t++] = vid(c);
                    c[perm[p][0]]++;
                    tets[t++] = vid(c);
                    c[perm[p][1]]++;
                    tets[t++] = vid(c);
                    c[perm[p][2]]++;
                    tets[t++] = vid(c);
                }
}

namespace {
struct TriKey {
    int32_t a, b, c;
    bool operator==(const TriKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct TriHash {
    size_t operator()(const TriKey& k) const
    {
        uint64_t h = (uint64_t)(uint32_t)k.a * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.b + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.c * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
}  // namespace

int surface_faces(int nv, int nt, const int32_t* tets, std::vector<int32_t>& faces)
{
    static const int lf[4][3] = { {1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1} };
    std::unordered_map<TriKey, std::pair<int64_t, int>, TriHash> count;
    count.reserve((size_t)nt * 3);
    for (int e = 0; e < nt; e++)
        for (int f = 0; f < 4; f++) {
            std::array<int32_t, 3> v = { tets[4 * e + lf[f][0]], tets[4 * e + lf[f][1]], tets[4 * e + lf[f][2]] };
            for (int32_t q : v)
                if (q < 0 || q >= nv) return -1;
            std::sort(v.begin(), v.end());
            auto& slot = count[TriKey{ v[0], v[1], v[2] }];
            if (slot.second == 0) slot.first = 4ll * e + f;
            slot.second++;
        }
    std::vector<int64_t> owners;
    for (auto& kv : count) {
        if (kv.second.second > 2) return -2;
        if (kv.second.second == 1) owners.push_back(kv.second.first);
    }
    std::sort(owners.begin(), owners.end());
    faces.resize(3 * owners.size());
    for (size_t i = 0; i < owners.size(); i++) {
        const int64_t e = owners[i] / 4;
        const int f = (int)(owners[i] % 4);
        for (int j = 0; j < 3; j++) faces[3 * i + j] = tets[4 * e + lf[f][j]];
    }
    return 0;
}

namespace {
constexpr int kWords = 4;  // up to 256 colours
#ifndef XPBD_COLOR_PASSES
#define XPBD_COLOR_PASSES 24
#endif
constexpr int kPasses = XPBD_COLOR_PASSES;
using Mask = std::array<uint64_t, kWords>;

inline bool has(const Mask& m, int c) { return (m[c / 64] >> (c % 64)) & 1ull; }

// First-fit greedy in the given visiting order.  Returns the colour count.
int first_fit(int nv, int n, int k, const int32_t* items, const std::vector<int32_t>& order,
              std::vector<int32_t>& color)
{
    std::vector<Mask> used((size_t)nv);
    for (auto& u : used) u.fill(0);
    color.assign(n, -1);
    int ncol = 0;
    for (int e : order) {
        Mask m{};
        for (int j = 0; j < k; j++) {
            const int32_t v = items[(size_t)k * e + j];
            if (v < 0) continue;
            for (int w = 0; w < kWords; w++) m[w] |= used[v][w];
        }
        int c = -1;
        for (int w = 0; w < kWords && c < 0; w++)
            if (~m[w]) c = 64 * w + __builtin_ctzll(~m[w]);
        if (c < 0) return -2;
        color[e] = c;
        for (int j = 0; j < k; j++) {
            const int32_t v = items[(size_t)k * e + j];
            if (v >= 0) used[v][c / 64] |= 1ull << (c % 64);
        }
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}
}  // namespace

int greedy_color(int nv, int n_items, int verts_per_item, const int32_t* items,
                 std::vector<int32_t>& color)
{
    const int k = verts_per_item;
    for (size_t i = 0; i < (size_t)k * n_items; i++)
        if (items[i] < -1 || items[i] >= nv) return -1;
    std::vector<int32_t> order(n_items);
    for (int e = 0; e < n_items; e++) order[e] = e;
    int ncol = first_fit(nv, n_items, k, items, order, color);
    if (ncol < 0 || n_items == 0) return ncol;

    // Iterated greedy: re-colour class by class.  Visiting every class
    // contiguously can never need more colours than before, and alternating
    // the class order (reverse, largest first, fixed shuffle) removes colours.
    uint64_t rs = 0x9E3779B97F4A7C15ull;
    std::vector<int32_t> trial;
    for (int pass = 0; pass < kPasses; pass++) {
        std::vector<int64_t> cnt(ncol, 0);
        for (int e = 0; e < n_items; e++) cnt[color[e]]++;
        std::vector<int32_t> cls(ncol);
        for (int c = 0; c < ncol; c++) cls[c] = c;
        if (pass % 3 == 0) std::reverse(cls.begin(), cls.end());
        else if (pass % 3 == 1) std::stable_sort(cls.begin(), cls.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
        else
            for (int c = ncol - 1; c > 0; c--) {
                rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
                std::swap(cls[c], cls[rs % (uint64_t)(c + 1)]);
            }
        std::vector<int32_t> rank(ncol);
        for (int c = 0; c < ncol; c++) rank[cls[c]] = c;
        std::vector<int64_t> off(ncol + 1, 0);
        for (int e = 0; e < n_items; e++) off[rank[color[e]] + 1]++;
        for (int c = 0; c < ncol; c++) off[c + 1] += off[c];
        for (int e = 0; e < n_items; e++) order[off[rank[color[e]]]++] = e;
        const int nc = first_fit(nv, n_items, k, items, order, trial);
        if (nc > 0 && nc <= ncol) {
            ncol = nc;
            color.swap(trial);
        }
    }

    // Balance: move items from crowded classes to the least-populated class
    // their vertices allow (never adds a colour).
    std::vector<Mask> used((size_t)nv);
    for (auto& u : used) u.fill(0);
    std::vector<int64_t> cnt(ncol, 0);
    for (int e = 0; e < n_items; e++) {
        cnt[color[e]]++;
        for (int j = 0; j < k; j++) {
            const int32_t v = items[(size_t)k * e + j];
            if (v >= 0) used[v][color[e] / 64] |= 1ull << (color[e] % 64);
        }
    }
    for (int sweep = 0; sweep < 4; sweep++)
        for (int e = 0; e < n_items; e++) {
            const int c0 = color[e];
            Mask m{};
            for (int j = 0; j < k; j++) {
                const int32_t v = items[(size_t)k * e + j];
                if (v >= 0)
                    for (int w = 0; w < kWords; w++) m[w] |= used[v][w];
            }
            int best = c0;
            for (int c = 0; c < ncol; c++)
                if (c != c0 && !has(m, c) && cnt[c] < cnt[best] - 1) best = c;
            if (best == c0) continue;
            for (int j = 0; j < k; j++) {
                const int32_t v = items[(size_t)k * e + j];
                if (v < 0) continue;
                used[v][c0 / 64] &= ~(1ull << (c0 % 64));
                used[v][best / 64] |= 1ull << (best % 64);
            }
            cnt[c0]--;
            cnt[best]++;
            color[e] = best;
        }
    return ncol;
}

int build_clusters(int nt, const int32_t* tets, std::vector<int32_t>& start, std::vector<int32_t>& verts,
                   std::vector<uint32_t>& tloc)
{
    start.assign(1, 0);
    verts.clear();
    tloc.assign(nt, 0u);
    int32_t cv[kClusterVerts];
    int nvc = 0, ntc = 0;
    auto flush = [&]() {
        for (int j = 0; j < kClusterVerts; j++) verts.push_back(j < nvc ? cv[j] : -1);
        start.push_back(start.back() + ntc);
        nvc = ntc = 0;
    };
    for (int e = 0; e < nt; e++) {
        const int32_t* t = tets + 4 * (size_t)e;
        // vertices of t not yet in the cluster
        int add = 0;
        for (int j = 0; j < 4; j++) {
            bool found = false;
            for (int q = 0; q < nvc; q++) found = found || cv[q] == t[j];
            for (int q = 0; q < j; q++) found = found || t[q] == t[j];
            add += !found;
        }
        if (ntc > 0 && (nvc + add > kClusterVerts || ntc == kClusterTets)) flush();
        uint32_t loc = 0;
        for (int j = 0; j < 4; j++) {
            int q = 0;
            while (q < nvc && cv[q] != t[j]) q++;
            if (q == nvc) cv[nvc++] = t[j];
            loc |= (uint32_t)q << (4 * j);
        }
        tloc[e] = loc;
        ntc++;
    }
    if (ntc > 0) flush();
    return (int)start.size() - 1;
}

namespace {
uint64_t spread3(uint64_t v)  // 21 bits -> every third bit
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
}  // namespace

int build_patches(int nv, int nt, const int32_t* tets, const double* x, const int32_t* tet_instance, int patch,
                  Patches& P)
{
    P.patch = patch;
    // ---- order: (instance, Morton code of the centroid)
    double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
    for (int v = 0; v < nv; v++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], x[3 * (size_t)v + k]);
            hi[k] = std::max(hi[k], x[3 * (size_t)v + k]);
        }
    std::vector<std::pair<uint64_t, int32_t>> key(nt);
    for (int e = 0; e < nt; e++) {
        uint64_t m = 0;
        for (int k = 0; k < 3; k++) {
            double c = 0;
            for (int j = 0; j < 4; j++) c += x[3 * (size_t)tets[4 * (size_t)e + j] + k];
            c *= 0.25;
            const double span = hi[k] - lo[k];
            const uint64_t q = span > 0 ? (uint64_t)std::min(1048575.0, (c - lo[k]) / span * 1048575.0) : 0;
            m |= spread3(q) << k;
        }
        // with instances: 20 bits of instance id above the top 44 bits of the code
        const uint64_t inst = tet_instance ? (uint64_t)tet_instance[e] : 0;
        key[e] = { tet_instance ? ((inst << 44) | (m >> 16)) : m, e };
    }
    std::stable_sort(key.begin(), key.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    P.order.resize(nt);
    for (int i = 0; i < nt; i++) P.order[i] = key[i].second;
    P.npatch = (nt + patch - 1) / patch;
    // ---- per-patch local vertices
    std::vector<int32_t> owner_cnt(nv, 0);  // number of patches touching each vertex
    P.pvoff.assign(1, 0);
    P.pverts.clear();
    P.tloc.assign(2 * (size_t)nt, 0u);
    P.maxl = 0;
    std::vector<int32_t> loc;
    for (int p = 0; p < P.npatch; p++) {
        const int b = p * patch, e = std::min(nt, b + patch);
        loc.clear();
        for (int i = b; i < e; i++)
            for (int j = 0; j < 4; j++) loc.push_back(tets[4 * (size_t)P.order[i] + j]);
        std::sort(loc.begin(), loc.end());
        loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
        if (loc.size() > 65535) return -1;
        for (int i = b; i < e; i++)
            for (int j = 0; j < 4; j++) {
                const int32_t g = tets[4 * (size_t)P.order[i] + j];
                const uint32_t l = (uint32_t)(std::lower_bound(loc.begin(), loc.end(), g) - loc.begin());
                P.tloc[2 * (size_t)i + j / 2] |= l << (16 * (j % 2));
            }
        for (int32_t g : loc) owner_cnt[g]++;
        P.pverts.insert(P.pverts.end(), loc.begin(), loc.end());
        P.pvoff.push_back((int32_t)P.pverts.size());
        P.maxl = std::max(P.maxl, (int)loc.size());
    }
    // ---- slots for shared vertices (touched by more than one patch)
    P.pslot.assign(P.pverts.size(), -1);
    P.nslots = 0;
    std::vector<std::vector<int32_t>> vslots;  // only for shared vertices, indexed via map below
    std::vector<int32_t> shared_id(nv, -1);
    P.shared.clear();
    for (int p = 0; p < P.npatch; p++)
        for (int j = P.pvoff[p]; j < P.pvoff[p + 1]; j++) {
            const int32_t g = P.pverts[j];
            if (owner_cnt[g] < 2) continue;
            if (shared_id[g] < 0) {
                shared_id[g] = (int32_t)P.shared.size();
                P.shared.push_back(g);
                vslots.emplace_back();
            }
            P.pslot[j] = P.nslots;
            vslots[shared_id[g]].push_back(P.nslots++);
        }
    // renumber the slots so that those of one shared vertex are contiguous (in
    // patch order): the shared pass then reads partial[soff[j] .. soff[j+1])
    P.soff.assign(1, 0);
    std::vector<int32_t> renum(P.nslots, -1);
    int32_t next = 0;
    for (auto& v : vslots) {
        for (int32_t old : v) renum[old] = next++;
        P.soff.push_back(next);
    }
    for (int32_t& s : P.pslot)
        if (s >= 0) s = renum[s];
    P.pshare.assign(2 * P.pverts.size(), 0);
    for (size_t e = 0; e < P.pverts.size(); e++) {
        const int32_t j = shared_id[P.pverts[e]];
        P.pshare[2 * e] = j;
        if (j >= 0) {
            const int32_t cnt = P.soff[j + 1] - P.soff[j];
            if (cnt > 15 || P.soff[j] >= (1 << 28)) P.fusable = false;  // does not fit the packing
            P.pshare[2 * e + 1] = P.soff[j] | ((cnt & 15) << 28);
        }
    }
    // ---- per-patch local incidence lists (local vertex -> tet corners)
    P.inc_off.clear();
    P.inc.assign((size_t)P.npatch * 4 * patch, 0);
    for (int p = 0; p < P.npatch; p++) {
        const int b = p * patch, e = std::min(nt, b + patch);
        const int nl = P.pvoff[p + 1] - P.pvoff[p];
        std::vector<int32_t> cnt(nl + 1, 0);
        for (int i = b; i < e; i++)
            for (int j = 0; j < 4; j++) cnt[((P.tloc[2 * (size_t)i + j / 2] >> (16 * (j % 2))) & 0xffff) + 1]++;
        for (int l = 0; l < nl; l++) cnt[l + 1] += cnt[l];
        P.inc_off.insert(P.inc_off.end(), cnt.begin(), cnt.end());
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        for (int i = b; i < e; i++)
            for (int j = 0; j < 4; j++) {
                const int l = (P.tloc[2 * (size_t)i + j / 2] >> (16 * (j % 2))) & 0xffff;
                P.inc[(size_t)p * 4 * patch + fill[l]++] = (uint16_t)((i - b) * 4 + j);
            }
    }
    return P.maxl;
}

int read_msh41(const std::string& path, std::vector<double>& x, std::vector<int32_t>& tets,
               std::vector<int32_t>& tris, std::string& err)
{
    std::ifstream in(path);
    if (!in) { err = "cannot open " + path; return -1; }
    std::string line;
    std::unordered_map<long long, int32_t> tag2id;
    x.clear(); tets.clear(); tris.clear();
    bool have_format = false, have_nodes = false, have_elements = false;
    while (std::getline(in, line)) {
        if (line.rfind("$MeshFormat", 0) == 0) {
            std::getline(in, line);
            std::istringstream ss(line);
            double ver = 0; int ftype = 1, dsize = 0;
            ss >> ver >> ftype >> dsize;
            if (ver < 4.0 || ver >= 5.0 || ftype != 0) { err = "only GMSH 4.x ASCII is supported"; return -2; }
            have_format = true;
        } else if (line.rfind("$Nodes", 0) == 0) {
            have_nodes = true;
            long long nblocks = -1, nnodes = -1, mintag, maxtag;
            in >> nblocks >> nnodes >> mintag >> maxtag;
            if (!in || nblocks < 0 || nnodes < 0) { err = "malformed $Nodes header"; return -6; }
            // the counts are the file's claim: reserve no more than a modest amount up front
            x.reserve(3 * (size_t)std::min<long long>(nnodes, 1 << 20));
            long long left = nnodes;  // a block may not hold more nodes than the header declares
            for (long long b = 0; b < nblocks; b++) {
                int edim = 0, etag = 0, param = 0; long long n = -1;
                in >> edim >> etag >> param >> n;
                if (!in || n < 0 || n > left || edim < 0 || edim > 3) { err = "malformed $Nodes block header"; return -6; }
                left -= n;
                std::vector<long long> tags(n);
                for (long long i = 0; i < n; i++) in >> tags[i];
                if (!in) { err = "malformed $Nodes block"; return -6; }
                for (long long i = 0; i < n; i++) {
                    double p[3];
                    in >> p[0] >> p[1] >> p[2];
                    if (!in) { err = "malformed $Nodes block"; return -6; }
                    if (param) {  // skip parametric coordinates
                        for (int d = 0; d < edim; d++) { double u; in >> u; }
                    }
                    tag2id[tags[i]] = (int32_t)(x.size() / 3);
                    x.insert(x.end(), p, p + 3);
                }
                if (!in) { err = "malformed $Nodes block"; return -6; }
            }
            if (left != 0) { err = "$Nodes blocks do not add up to the header's node count"; return -6; }
        } else if (line.rfind("$Elements", 0) == 0) {
            have_elements = true;
            long long nblocks = -1, nel = -1, mintag, maxtag;
            in >> nblocks >> nel >> mintag >> maxtag;
            if (!in || nblocks < 0 || nel < 0) { err = "malformed $Elements header"; return -6; }
            long long left = nel;  // a block may not hold more elements than the header declares
            for (long long b = 0; b < nblocks; b++) {
                int edim = 0, etag = 0, etype = 0; long long n = -1;
                in >> edim >> etag >> etype >> n;
                if (!in || n < 0 || n > left) { err = "malformed $Elements block header"; return -6; }
                left -= n;
                const int nn = etype == 4 ? 4 : etype == 2 ? 3 : etype == 1 ? 2 : etype == 15 ? 1 :
                               etype == 11 ? 10 : etype == 9 ? 6 : etype == 5 ? 8 : -1;
                if (nn < 0) { err = "unsupported element type " + std::to_string(etype); return -3; }
                for (long long i = 0; i < n; i++) {
                    long long etagv; in >> etagv;
                    long long nt[10];
                    for (int j = 0; j < nn; j++) in >> nt[j];
                    if (!in) { err = "malformed $Elements block"; return -6; }
                    if (etype == 4 || etype == 2) {
                        auto& dst = etype == 4 ? tets : tris;
                        for (int j = 0; j < nn; j++) {
                            auto it = tag2id.find(nt[j]);
                            if (it == tag2id.end()) { err = "element references unknown node"; return -4; }
                            dst.push_back(it->second);
                        }
                    }
                }
                if (!in) { err = "malformed $Elements block"; return -6; }
            }
            if (left != 0) { err = "$Elements blocks do not add up to the header's element count"; return -6; }
        }
    }
    if (!have_format) { err = "missing $MeshFormat"; return -5; }
    // a file without these sections is not a GMSH 4.1 mesh (the gmsh loader of
    // MeshUtils.cpp:92-218 reads both), however its body is laid out
    if (!have_nodes) { err = "missing $Nodes section"; return -7; }
    if (!have_elements) { err = "missing $Elements section"; return -7; }
    return 0;
}

// One integer per line until the first line that does not start with one
// (FileUtils::readVectorFromFile<int>, include/utils/FileUtils.hpp:12-30).
int read_int_lines(const std::string& path, std::vector<int32_t>& out, std::string& err)
{
    std::ifstream in(path);
    if (!in) { err = "cannot open " + path; return -1; }
    out.clear();
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        long long v;
        if (!(ss >> v)) break;
        if (v < INT32_MIN || v > INT32_MAX) { err = "value out of range in " + path; return -2; }
        out.push_back((int32_t)v);
    }
    return 0;
}

}  // namespace xpbd
