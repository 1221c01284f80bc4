// rs_host.cpp -- host-side model preparation: structured Kuhn-tet meshes, seeded fields,
// node books (DOF numbering), CSR patterns of the free DOFs and node-disjoint element
// colourings.  No device code; every entry point here works without a GPU.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>

#include "rs_internal.hpp"

namespace rs {

void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)> &fn)
{
    unsigned hw = std::thread::hardware_concurrency();
    unsigned nt = hw == 0 ? 1 : std::min(hw, 32u);
    if (n < 4096 || nt == 1) {
        fn(0, n);
        return;
    }
    nt = (unsigned)std::min<uint64_t>(nt, n / 1024);
    std::vector<std::thread> th;
    uint64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        uint64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------------------- fields
// SplitMix64 finaliser (Steele, Lea, Flood 2014) over a per-index counter: value i of a
// stream depends only on (seed, i), so any subset can be regenerated anywhere.
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double u01(uint64_t seed, uint64_t i)
{
    uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull * (i + 1));
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
}

}  // namespace rs

using namespace rs;

extern "C" int rs_field_uniform(uint64_t seed, uint64_t n, double *out)
{
    if (!out && n) return fail(RS_ERR_INVALID_ARG, "rs_field_uniform: null output");
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) out[i] = u01(seed, i);
    });
    return RS_OK;
}

extern "C" int rs_field_lognormal(uint64_t seed, uint64_t n, double mu, double sigma, double *out)
{
    if (!out && n) return fail(RS_ERR_INVALID_ARG, "rs_field_lognormal: null output");
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            // Box-Muller on two counter draws; 1-u keeps the log argument in (0, 1].
            double u1 = 1.0 - u01(seed, 2 * i), u2 = u01(seed, 2 * i + 1);
            double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
            out[i] = std::exp(mu + sigma * g);
        }
    });
    return RS_OK;
}

// ---------------------------------------------------------------------------- meshes
extern "C" int rs_mesh_kuhn_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly,
                                double lz, double *xyz, uint32_t *tets)
{
    if (!nx || !ny || !nz || !xyz || !tets) return fail(RS_ERR_INVALID_ARG, "rs_mesh_kuhn_box: bad arguments");
    uint64_t px = nx + 1ull, py = ny + 1ull, pz = nz + 1ull;
    if (px * py * pz > 0xFFFFFFF0ull) return fail(RS_ERR_INVALID_ARG, "rs_mesh_kuhn_box: too many nodes");
    parallel_for(px * py * pz, [&](uint64_t b, uint64_t e) {
        for (uint64_t id = b; id < e; id++) {
            uint64_t i = id % px, j = (id / px) % py, k = id / (px * py);
            xyz[3 * id + 0] = lx * (double)i / nx;
            xyz[3 * id + 1] = ly * (double)j / ny;
            xyz[3 * id + 2] = lz * (double)k / nz;
        }
    });
    // The 6 tets of a cell follow the monotone lattice paths from corner 0 to corner 7:
    // {0, e_a, e_a+e_b, 7} for each ordering (a, b, c) of the axes.  Odd orderings are
    // negatively oriented and get their last two vertices swapped.
    static const int perm[6][3] = { { 0, 1, 2 }, { 0, 2, 1 }, { 1, 0, 2 }, { 1, 2, 0 }, { 2, 0, 1 }, { 2, 1, 0 } };
    static const int odd[6] = { 0, 1, 1, 0, 0, 1 };
    uint64_t ncell = (uint64_t)nx * ny * nz;
    parallel_for(ncell, [&](uint64_t b, uint64_t e) {
        for (uint64_t c = b; c < e; c++) {
            uint64_t ci = c % nx, cj = (c / nx) % ny, ck = c / ((uint64_t)nx * ny);
            auto corner = [&](int bits) -> uint32_t {
                return (uint32_t)((ci + (bits & 1)) + px * ((cj + ((bits >> 1) & 1)) + py * (ck + ((bits >> 2) & 1))));
            };
            for (int t = 0; t < 6; t++) {
                int a = 1 << perm[t][0], ab = a | (1 << perm[t][1]);
                uint32_t v[4] = { corner(0), corner(a), corner(ab), corner(7) };
                if (odd[t]) std::swap(v[2], v[3]);
                std::memcpy(tets + 4 * (6 * c + t), v, sizeof(v));
            }
        }
    });
    return RS_OK;
}

extern "C" int rs_mesh_plane_faces(uint32_t n_nodes, const double *xyz, uint32_t n_tets,
                                   const uint32_t *tets, int32_t axis, double value, double tol,
                                   uint32_t *tris, uint64_t cap, uint64_t *count)
{
    if (!xyz || !tets || !count || axis < 0 || axis > 2)
        return fail(RS_ERR_INVALID_ARG, "rs_mesh_plane_faces: bad arguments");
    static const int faces[4][3] = { { 1, 2, 3 }, { 0, 2, 3 }, { 0, 1, 3 }, { 0, 1, 2 } };
    uint64_t cnt = 0;
    for (uint64_t e = 0; e < n_tets; e++) {
        const uint32_t *nd = tets + 4 * e;
        for (int f = 0; f < 4; f++) {
            bool on = true;
            for (int v = 0; v < 3 && on; v++) {
                uint32_t id = nd[faces[f][v]];
                if (id >= n_nodes) return fail(RS_ERR_INVALID_ARG, "rs_mesh_plane_faces: node id out of range");
                on = std::fabs(xyz[3 * (uint64_t)id + axis] - value) <= tol;
            }
            if (!on) continue;
            if (tris) {
                if (cnt >= cap) return fail(RS_ERR_INVALID_ARG, "rs_mesh_plane_faces: capacity exceeded");
                for (int v = 0; v < 3; v++) tris[3 * cnt + v] = nd[faces[f][v]];
            }
            cnt++;
        }
    }
    *count = cnt;
    return RS_OK;
}

// ---------------------------------------------------------------------------- node book
// RSolverGeneric::rebuildNodeBook (src/rsolvergeneric.cpp:565-582): enabled positions are
// numbered 0..n-1 in increasing position order; prescribed positions and positions of nodes
// that belong to no computable element (:542-560, rsolverstress.cpp:1631-1651) get eod.
extern "C" int rs_node_book(uint32_t n_nodes, uint32_t n_elems, uint32_t nodes_per_elem,
                            const uint32_t *conn, uint32_t dofs_per_node, const uint8_t *fixed,
                            uint32_t *book, uint32_t *n_enabled)
{
    if (!conn || !book || !n_enabled || dofs_per_node == 0 || nodes_per_elem == 0)
        return fail(RS_ERR_INVALID_ARG, "rs_node_book: bad arguments");
    std::vector<uint8_t> used(n_nodes, 0);
    for (uint64_t i = 0; i < (uint64_t)n_elems * nodes_per_elem; i++) {
        if (conn[i] >= n_nodes) return fail(RS_ERR_INVALID_ARG, "rs_node_book: node id out of range");
        used[conn[i]] = 1;
    }
    uint32_t next = 0;
    for (uint64_t p = 0; p < (uint64_t)n_nodes * dofs_per_node; p++) {
        bool disabled = (fixed && fixed[p]) || !used[p / dofs_per_node];
        book[p] = disabled ? RS_EOD : next++;
    }
    *n_enabled = next;
    return RS_OK;
}

namespace rs {

// ---------------------------------------------------------------------------- pattern
// Node -> element adjacency, then per node the sorted set of nodes sharing an element.
// Rows of one node list the enabled DOFs of its neighbour nodes; because the node book is
// monotone in (node, component), the columns come out sorted.
int build_pattern(uint32_t n_nodes, uint32_t n_elems, uint32_t npe, const uint32_t *conn,
                  uint32_t dpn, const uint32_t *book, uint32_t n_dof, Pattern &pat)
{
    std::vector<uint32_t> adj_ptr(n_nodes + 1ull, 0);
    for (uint64_t i = 0; i < (uint64_t)n_elems * npe; i++) adj_ptr[conn[i] + 1]++;
    for (uint32_t i = 0; i < n_nodes; i++) adj_ptr[i + 1] += adj_ptr[i];
    std::vector<uint32_t> adj(adj_ptr[n_nodes]);
    {
        std::vector<uint32_t> fill(adj_ptr.begin(), adj_ptr.end() - 1);
        for (uint64_t e = 0; e < n_elems; e++)
            for (uint32_t k = 0; k < npe; k++) adj[fill[conn[e * npe + k]]++] = (uint32_t)e;
    }
    auto row_free = [&](uint32_t node) {
        for (uint32_t c = 0; c < dpn; c++)
            if (book[(uint64_t)node * dpn + c] != RS_EOD) return true;
        return false;
    };
    // neighbour lists (free neighbours only) per node
    std::vector<uint32_t> ncount(n_nodes, 0);
    std::vector<std::vector<uint32_t>> nbrs(n_nodes);
    parallel_for(n_nodes, [&](uint64_t b, uint64_t e) {
        std::vector<uint32_t> tmp;
        for (uint64_t a = b; a < e; a++) {
            if (!row_free((uint32_t)a)) continue;
            tmp.clear();
            for (uint32_t j = adj_ptr[a]; j < adj_ptr[a + 1]; j++) {
                const uint32_t *nd = conn + (uint64_t)adj[j] * npe;
                for (uint32_t k = 0; k < npe; k++)
                    if (row_free(nd[k])) tmp.push_back(nd[k]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            nbrs[a] = tmp;
        }
    });
    pat.n_dof = n_dof;
    pat.rowptr.assign(n_dof + 1ull, 0);
    // row lengths
    for (uint64_t a = 0; a < n_nodes; a++) {
        if (nbrs[a].empty()) continue;
        uint32_t ncol = 0;
        for (uint32_t nb : nbrs[a])
            for (uint32_t c = 0; c < dpn; c++) ncol += book[(uint64_t)nb * dpn + c] != RS_EOD;
        for (uint32_t c = 0; c < dpn; c++) {
            uint32_t r = book[a * dpn + c];
            if (r != RS_EOD) pat.rowptr[r + 1ull] = ncol;
        }
    }
    for (uint32_t r = 0; r < n_dof; r++) {
        uint64_t s = (uint64_t)pat.rowptr[r] + pat.rowptr[r + 1];
        if (s > 0xFFFFFFFFull) return fail(RS_ERR_UNSUPPORTED, "pattern: nnz exceeds 2^32 (uint32 row pointers)");
        pat.rowptr[r + 1] = (uint32_t)s;
    }
    pat.col.resize(pat.rowptr[n_dof]);
    parallel_for(n_nodes, [&](uint64_t b, uint64_t e) {
        for (uint64_t a = b; a < e; a++) {
            for (uint32_t c = 0; c < dpn; c++) {
                uint32_t r = book[a * dpn + c];
                if (r == RS_EOD) continue;
                uint32_t w = pat.rowptr[r];
                for (uint32_t nb : nbrs[a])
                    for (uint32_t cc = 0; cc < dpn; cc++) {
                        uint32_t d = book[(uint64_t)nb * dpn + cc];
                        if (d != RS_EOD) pat.col[w++] = d;
                    }
            }
        }
    });
    return RS_OK;
}

// ---------------------------------------------------------------------------- SpMV tiles
std::vector<uint32_t> build_tiles(const uint32_t *rowptr, uint32_t n, uint32_t cap, uint32_t max_rows)
{
    std::vector<uint32_t> tiles;
    tiles.reserve(n / std::max(1u, max_rows) + 2);
    tiles.push_back(0);
    uint32_t r = 0;
    while (r < n) {
        uint32_t start = r;
        while (r < n && r - start < max_rows && rowptr[r + 1] - rowptr[start] <= cap) r++;
        if (r == start) r++;  // one row longer than cap
        tiles.push_back(r);
    }
    return tiles;
}

// ---------------------------------------------------------------------------- colouring
// First-fit greedy colouring in element order: an element takes the lowest colour that
// none of its nodes has seen yet.  Elements of one colour share no node, so one launch per
// colour scatters without write conflicts, and the colour order fixes the summation order.
int colour_elements(uint32_t n_nodes, uint32_t n_elems, uint32_t npe, const uint32_t *conn,
                    std::vector<uint16_t> &colour, uint32_t &n_colours)
{
    constexpr int kWords = 4;  // up to 256 colours
    std::vector<uint64_t> seen((uint64_t)n_nodes * kWords, 0);
    colour.assign(n_elems, 0);
    n_colours = 0;
    for (uint64_t e = 0; e < n_elems; e++) {
        const uint32_t *nd = conn + e * npe;
        int c = -1;
        for (int w = 0; w < kWords && c < 0; w++) {
            uint64_t busy = 0;
            for (uint32_t k = 0; k < npe; k++) busy |= seen[(uint64_t)nd[k] * kWords + w];
            if (~busy) c = 64 * w + __builtin_ctzll(~busy);
        }
        if (c < 0) return fail(RS_ERR_UNSUPPORTED, "colouring: more than 256 colours needed");
        for (uint32_t k = 0; k < npe; k++) seen[(uint64_t)nd[k] * kWords + c / 64] |= 1ull << (c % 64);
        colour[e] = (uint16_t)c;
        n_colours = std::max<uint32_t>(n_colours, (uint32_t)c + 1);
    }
    return RS_OK;
}

}  // namespace rs
