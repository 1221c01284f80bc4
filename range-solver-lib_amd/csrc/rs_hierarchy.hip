// rs_hierarchy.hip -- the learned preconditioner's aggregation hierarchy built on the GPU (no
// counterpart in the reference): nested coordinate bins, and per level the members of every
// aggregate, the next level's pattern and the gather lists that sum the Galerkin values.
//
// Every list comes out in the order the host builder of the same lists uses (aggregate ids in
// first-appearance order of the current level's nodes; members by row; next-level entries by
// column; gathered positions ascending), so both builders give bit-identical hierarchies.
// The sorts are rocPRIM's stable LSD radix sorts; the scans are rocPRIM scans of integers.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <memory>

#include "rs_device.hpp"
#include "rs_internal.hpp"

namespace rs {

namespace {

constexpr uint32_t kHierBlock = 256;

uint32_t hier_grid(uint64_t n)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kHierBlock - 1) / kHierBlock, 65536));
}

#define HIER_LOOP(i, n) for (uint64_t i = (uint64_t)blockIdx.x * kHierBlock + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * kHierBlock)

__global__ void k_fill_u32(uint32_t *a, uint64_t n, uint32_t v)
{
    HIER_LOOP(i, n) a[i] = v;
}

__global__ void k_iota_u32(uint32_t *a, uint64_t n)
{
    HIER_LOOP(i, n) a[i] = (uint32_t)i;
}

// node_of[d]: the node of DOF d; comp_of[d]: its component (position within the node).
// bad: 1 = a DOF number >= n_dof, 2 = a DOF numbered twice, 4 = a DOF without a node.
__global__ void k_node_of(const uint32_t *__restrict__ book, uint64_t npos, uint32_t dpn, uint32_t n_dof,
                          uint32_t *__restrict__ node_of, uint32_t *__restrict__ comp_of, uint32_t *__restrict__ bad)
{
    HIER_LOOP(q, npos)
    {
        const uint32_t d = book[q];
        if (d == RS_EOD) continue;
        if (d >= n_dof) {
            atomicOr(bad, 1u);
            continue;
        }
        if (atomicExch(&node_of[d], (uint32_t)(q / dpn)) != RS_EOD) atomicOr(bad, 2u);
        comp_of[d] = (uint32_t)(q % dpn);
    }
}

__global__ void k_check_assigned(const uint32_t *__restrict__ node_of, uint32_t n, uint32_t *__restrict__ bad)
{
    HIER_LOOP(d, n)
    if (node_of[d] == RS_EOD) atomicOr(bad, 4u);
}

// Bounding box of all nodes: per workgroup (lo, hi) x 3 into part[6 * gridDim.x], then one
// workgroup folds them (min and max do not depend on the order).
__global__ __launch_bounds__(kHierBlock) void k_bbox(const double *__restrict__ xyz, uint64_t n, double *__restrict__ part)
{
    double v[6] = { 1e300, 1e300, 1e300, -1e300, -1e300, -1e300 };
    HIER_LOOP(i, n)
    for (int a = 0; a < 3; a++) {
        v[a] = fmin(v[a], xyz[3 * i + a]);
        v[3 + a] = fmax(v[3 + a], xyz[3 * i + a]);
    }
    __shared__ double s[6][kHierBlock];
    for (int a = 0; a < 6; a++) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (uint32_t h = kHierBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int a = 0; a < 6; a++)
                s[a][threadIdx.x] = a < 3 ? fmin(s[a][threadIdx.x], s[a][threadIdx.x + h])
                                          : fmax(s[a][threadIdx.x], s[a][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kHierBlock) void k_bbox_fin(const double *__restrict__ part, uint32_t np, double *__restrict__ box)
{
    __shared__ double s[6][kHierBlock];
    for (int a = 0; a < 6; a++) {
        double v = a < 3 ? 1e300 : -1e300;
        for (uint32_t k = threadIdx.x; k < np; k += kHierBlock) v = a < 3 ? fmin(v, part[(uint64_t)a * np + k]) : fmax(v, part[(uint64_t)a * np + k]);
        s[a][threadIdx.x] = v;
    }
    __syncthreads();
    for (uint32_t h = kHierBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int a = 0; a < 6; a++)
                s[a][threadIdx.x] = a < 3 ? fmin(s[a][threadIdx.x], s[a][threadIdx.x + h])
                                          : fmax(s[a][threadIdx.x], s[a][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = s[threadIdx.x][0];
}

struct BinArgs {
    const double *xyz;
    const uint32_t *node_of, *rep;
    const uint32_t *comp_of;    // per component: the component of each fine DOF; nullptr: per node
    const double *box;          // lo[3], hi[3]
    uint32_t cur, nb0, nb1, nb2, dpn;
};

// Bin of each node J of the current level (the coordinates of a fine node it contains), and the
// first J of every bin.  Per component, bin b holds one aggregate per component c (key dpn b + c),
// so an aggregate never mixes components.  The expressions are those of the host builder, so the
// bins agree.
__global__ void k_bin(BinArgs a, uint32_t *__restrict__ bin, uint32_t *__restrict__ first)
{
    const double lo[3] = { a.box[0], a.box[1], a.box[2] }, hi[3] = { a.box[3], a.box[4], a.box[5] };
    const uint32_t nb[3] = { a.nb0, a.nb1, a.nb2 };
    HIER_LOOP(J, a.cur)
    {
        const double *p = a.xyz + 3ull * a.node_of[a.rep[J]];
        uint64_t c[3];
        for (int d = 0; d < 3; d++) {
            const double t = hi[d] > lo[d] ? (p[d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
            const uint64_t q = (uint64_t)(t * nb[d]);
            c[d] = q < nb[d] - 1ull ? q : nb[d] - 1ull;
        }
        uint32_t b = (uint32_t)(c[0] + nb[0] * (c[1] + (uint64_t)nb[1] * c[2]));
        if (a.comp_of) b = b * a.dpn + a.comp_of[a.rep[J]];
        bin[J] = b;
        atomicMin(&first[b], (uint32_t)J);
    }
}

__global__ void k_is_first(const uint32_t *__restrict__ bin, const uint32_t *__restrict__ first, uint32_t cur,
                           uint32_t *__restrict__ flag)
{
    HIER_LOOP(J, cur) flag[J] = first[bin[J]] == J ? 1u : 0u;
}

// Aggregate of J: the rank of its bin's first node among all first nodes (first-appearance
// order); each first node carries its fine representative to the next level.
__global__ void k_assign(const uint32_t *__restrict__ bin, const uint32_t *__restrict__ first,
                         const uint32_t *__restrict__ rank, const uint32_t *__restrict__ flag,
                         const uint32_t *__restrict__ rep, uint32_t cur, uint32_t *__restrict__ agg,
                         uint32_t *__restrict__ next_rep)
{
    HIER_LOOP(J, cur)
    {
        agg[J] = rank[first[bin[J]]];
        if (flag[J]) next_rep[rank[J]] = rep[J];
    }
}

// Key of every entry p of the level matrix: (aggregate of its row) * nn + aggregate of its column.
__global__ void k_entry_keys(const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ col, uint32_t n,
                             const uint32_t *__restrict__ agg, uint32_t nn, uint64_t *__restrict__ key,
                             uint32_t *__restrict__ pos)
{
    HIER_LOOP(i, n)
    {
        const uint64_t I = agg[i];
        for (uint32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
            key[p] = I * nn + agg[col[p]];
            pos[p] = p;
        }
    }
}

template <typename K>
__global__ void k_run_flags(const K *__restrict__ key, uint64_t m, uint32_t *__restrict__ flag)
{
    HIER_LOOP(t, m) flag[t] = (t == 0 || key[t] != key[t - 1]) ? 1u : 0u;
}

// Runs of equal keys (one run per next-level entry): start of each run, its column, and the
// number of runs per next-level row.
__global__ void k_runs(const uint64_t *__restrict__ key, const uint32_t *__restrict__ flag,
                       const uint32_t *__restrict__ run, uint64_t m, uint32_t nn, uint32_t *__restrict__ gptr,
                       uint32_t *__restrict__ ccol, uint32_t *__restrict__ row_runs)
{
    HIER_LOOP(t, m)
    {
        if (!flag[t]) continue;
        const uint32_t s = run[t];
        gptr[s] = (uint32_t)t;
        ccol[s] = (uint32_t)(key[t] % nn);
        atomicAdd(&row_runs[key[t] / nn], 1u);
    }
}

__global__ void k_count_keys(const uint32_t *__restrict__ key, uint64_t m, uint32_t *__restrict__ count)
{
    HIER_LOOP(t, m) atomicAdd(&count[key[t]], 1u);
}

__global__ void k_set_u32(uint32_t *a, uint32_t v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *a = v;
}

__global__ void k_compose(uint32_t *__restrict__ total, uint64_t n, const uint32_t *__restrict__ map)
{
    HIER_LOOP(i, n) total[i] = map[total[i]];
}

uint32_t key_bits(uint64_t max_key)
{
    uint32_t b = 1;
    while (b < 64 && (max_key >> b) != 0) b++;
    return b;
}

// Workspace of the library routines, grown on demand.
// rocPRIM temporary storage: its own buffer, or a HierWork's when one is given.
struct Temp {
    DBuf<uint8_t> own;
    DBuf<uint8_t> *ext = nullptr;
    DBuf<uint8_t> &buf() { return ext ? *ext : own; }
    int fit(size_t bytes)
    {
        DBuf<uint8_t> &b = buf();
        return bytes <= b.n && b.p ? RS_OK : b.alloc(std::max<size_t>(bytes, 256));
    }
};

// At least count elements (a HierWork buffer keeps a larger earlier allocation).
template <typename T>
int grow(DBuf<T> &b, size_t count)
{
    return b.p && b.n >= count ? RS_OK : b.alloc(count);
}

template <typename K>
int sort_pairs(Temp &tmp, const K *kin, K *kout, const uint32_t *vin, uint32_t *vout, uint64_t m, uint32_t bits,
               hipStream_t st)
{
    size_t bytes = 0;
    RS_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, m, 0u, bits, st));
    int rc = tmp.fit(bytes);
    if (rc) return rc;
    bytes = tmp.buf().n;
    RS_HIP(rocprim::radix_sort_pairs(tmp.buf().p, bytes, kin, kout, vin, vout, m, 0u, bits, st));
    return RS_OK;
}

int scan_u32(Temp &tmp, const uint32_t *in, uint32_t *out, uint64_t m, hipStream_t st)
{
    size_t bytes = 0;
    RS_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)m, rocprim::plus<uint32_t>(), st));
    int rc = tmp.fit(bytes);
    if (rc) return rc;
    bytes = tmp.buf().n;
    RS_HIP(rocprim::exclusive_scan(tmp.buf().p, bytes, in, out, 0u, (size_t)m, rocprim::plus<uint32_t>(), st));
    return RS_OK;
}

// ptr[0..nn] = exclusive scan of count[0..nn) with ptr[nn] = the total.
int counts_to_ptr(Temp &tmp, DBuf<uint32_t> &count, uint32_t nn, DBuf<uint32_t> &ptr, hipStream_t st)
{
    int rc;
    if ((rc = ptr.alloc(nn + 1ull))) return rc;
    // count has nn + 1 slots, the last one 0, so the scan's last output is the total
    return scan_u32(tmp, count.p, ptr.p, nn + 1ull, st);
}

}  // namespace

int hier_members(const uint32_t *d_map, uint32_t n, uint32_t nn, DBuf<uint32_t> &ptr, DBuf<uint32_t> &rows,
                 HierWork *work)
{
    if (nn == 0) return fail(RS_ERR_INVALID_ARG, "hierarchy: empty level");
    hipStream_t st = runtime().stream;
    Temp tmp;
    HierWork local;
    HierWork &w = work ? *work : local;
    if (work) tmp.ext = &work->tmp;
    DBuf<uint32_t> &kout = w.kout32, &vin = w.vin, &count = w.count;
    int rc;
    if ((rc = rows.alloc(std::max(n, 1u))) || (rc = grow(kout, std::max(n, 1u))) || (rc = grow(vin, std::max(n, 1u))) ||
        (rc = grow(count, nn + 1ull)))
        return rc;
    k_fill_u32<<<hier_grid(nn + 1ull), kHierBlock, 0, st>>>(count.p, nn + 1ull, 0u);
    if (n) {
        k_iota_u32<<<hier_grid(n), kHierBlock, 0, st>>>(vin.p, n);
        if ((rc = sort_pairs<uint32_t>(tmp, d_map, kout.p, vin.p, rows.p, n, key_bits(nn - 1ull), st))) return rc;
        k_count_keys<<<hier_grid(n), kHierBlock, 0, st>>>(d_map, n, count.p);
    }
    RS_HIP(hipGetLastError());
    if ((rc = counts_to_ptr(tmp, count, nn, ptr, st))) return rc;
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

int hier_level_lists(const uint32_t *d_rowptr, const uint32_t *d_col, uint32_t n, uint64_t nnz, const uint32_t *d_map,
                     uint32_t nn, DBuf<uint32_t> &cptr, DBuf<uint32_t> &ccol, DBuf<uint32_t> &gptr,
                     DBuf<uint32_t> &gsrc, uint32_t &cnnz, HierWork *work)
{
    if (nn == 0) return fail(RS_ERR_INVALID_ARG, "hierarchy: empty level");
    if (nnz >= 0xFFFFFFFFull) return fail(RS_ERR_UNSUPPORTED, "hierarchy: level matrix has 2^32 entries or more");
    hipStream_t st = runtime().stream;
    Temp tmp;
    HierWork local;
    HierWork &w = work ? *work : local;
    if (work) tmp.ext = &work->tmp;
    int rc;
    const uint64_t m = nnz;
    DBuf<uint64_t> &kin = w.kin, &kout = w.kout;
    DBuf<uint32_t> &pin = w.pin, &flag = w.flag, &run = w.run, &row_runs = w.count;
    const size_t m1 = std::max<uint64_t>(m, 1);
    if ((rc = grow(kin, m1)) || (rc = grow(kout, m1)) || (rc = grow(pin, m1)) || (rc = gsrc.alloc(m1)) ||
        (rc = grow(flag, m1)) || (rc = grow(run, m1)) || (rc = grow(row_runs, nn + 1ull)))
        return rc;
    k_fill_u32<<<hier_grid(nn + 1ull), kHierBlock, 0, st>>>(row_runs.p, nn + 1ull, 0u);
    uint32_t nruns = 0;
    if (m) {
        k_entry_keys<<<hier_grid(n), kHierBlock, 0, st>>>(d_rowptr, d_col, n, d_map, nn, kin.p, pin.p);
        RS_HIP(hipGetLastError());
        if ((rc = sort_pairs<uint64_t>(tmp, kin.p, kout.p, pin.p, gsrc.p, m, key_bits((uint64_t)nn * nn - 1), st)))
            return rc;
        k_run_flags<uint64_t><<<hier_grid(m), kHierBlock, 0, st>>>(kout.p, m, flag.p);
        RS_HIP(hipGetLastError());
        if ((rc = scan_u32(tmp, flag.p, run.p, m, st))) return rc;
        uint32_t last[2] = { 0, 0 };
        RS_HIP(hipMemcpyAsync(&last[0], run.p + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        RS_HIP(hipMemcpyAsync(&last[1], flag.p + (m - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        nruns = last[0] + last[1];
    }
    if ((rc = ccol.alloc(std::max(nruns, 1u))) || (rc = gptr.alloc(nruns + 1ull))) return rc;
    if (m) k_runs<<<hier_grid(m), kHierBlock, 0, st>>>(kout.p, flag.p, run.p, m, nn, gptr.p, ccol.p, row_runs.p);
    k_set_u32<<<1, 64, 0, st>>>(gptr.p + nruns, (uint32_t)m);
    RS_HIP(hipGetLastError());
    if ((rc = counts_to_ptr(tmp, row_runs, nn, cptr, st))) return rc;
    RS_HIP(hipStreamSynchronize(st));
    cnnz = nruns;
    return RS_OK;
}

int hier_compose(uint32_t *d_total, uint64_t n, const uint32_t *d_map)
{
    hipStream_t st = runtime().stream;
    if (n) k_compose<<<hier_grid(n), kHierBlock, 0, st>>>(d_total, n, d_map);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

int hier_identity(uint32_t *d_a, uint64_t n)
{
    hipStream_t st = runtime().stream;
    if (n) k_iota_u32<<<hier_grid(n), kHierBlock, 0, st>>>(d_a, n);
    RS_HIP(hipGetLastError());
    return RS_OK;
}

int hier_bins(uint32_t n_nodes, const double *d_xyz, uint32_t dpn, const uint32_t *d_book, uint32_t n_dof,
              const uint32_t bins[3], uint32_t max_coarse, uint32_t max_levels, bool by_component,
              std::vector<std::unique_ptr<DBuf<uint32_t>>> &maps, std::vector<uint32_t> &n_agg)
{
    maps.clear();
    n_agg.clear();
    if (!d_xyz || !d_book || !n_nodes || !dpn || !n_dof || !bins[0] || !bins[1] || !bins[2] || !max_coarse || !max_levels)
        return fail(RS_ERR_INVALID_ARG, "hierarchy by bins: bad arguments");
    hipStream_t st = runtime().stream;
    Temp tmp;
    int rc;
    DBuf<uint32_t> node_of, comp_of, bad, rep, next_rep, bin, first, flag, rank;
    DBuf<double> part, box;
    const uint64_t npos = (uint64_t)n_nodes * dpn;
    const uint32_t nbb = hier_grid(n_nodes);
    if ((rc = node_of.alloc(n_dof)) || (rc = comp_of.alloc(n_dof)) || (rc = bad.alloc(1)) || (rc = rep.alloc(n_dof)) ||
        (rc = next_rep.alloc(n_dof)) || (rc = bin.alloc(n_dof)) || (rc = flag.alloc(n_dof)) || (rc = rank.alloc(n_dof)) ||
        (rc = part.alloc(6ull * nbb)) || (rc = box.alloc(6)))
        return rc;
    k_fill_u32<<<hier_grid(n_dof), kHierBlock, 0, st>>>(node_of.p, n_dof, RS_EOD);
    k_fill_u32<<<1, kHierBlock, 0, st>>>(bad.p, 1, 0u);
    k_node_of<<<hier_grid(npos), kHierBlock, 0, st>>>(d_book, npos, dpn, n_dof, node_of.p, comp_of.p, bad.p);
    k_check_assigned<<<hier_grid(n_dof), kHierBlock, 0, st>>>(node_of.p, n_dof, bad.p);
    k_bbox<<<nbb, kHierBlock, 0, st>>>(d_xyz, n_nodes, part.p);
    k_bbox_fin<<<1, kHierBlock, 0, st>>>(part.p, nbb, box.p);
    k_iota_u32<<<hier_grid(n_dof), kHierBlock, 0, st>>>(rep.p, n_dof);
    RS_HIP(hipGetLastError());
    uint32_t hbad = 0;
    RS_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    RS_HIP(hipStreamSynchronize(st));
    if (hbad) return fail(RS_ERR_INVALID_ARG, "hierarchy by bins: book is not a numbering 0..n-1 of the matrix rows");
    uint32_t cur = n_dof, levels = 0;
    for (uint32_t k = 0; cur > max_coarse || levels == 0; k++) {
        if (levels == max_levels) return fail(RS_ERR_UNSUPPORTED, "hierarchy by bins: coarsest level still too large");
        const uint32_t sh = std::min(k, 31u);
        const uint32_t nb[3] = { std::max(1u, bins[0] >> sh), std::max(1u, bins[1] >> sh), std::max(1u, bins[2] >> sh) };
        const uint64_t nbins = (uint64_t)nb[0] * nb[1] * nb[2] * (by_component ? dpn : 1u);
        if (nbins >= 0xFFFFFFFFull) return fail(RS_ERR_UNSUPPORTED, "hierarchy by bins: 2^32 bins or more");
        if ((rc = first.alloc(nbins))) return rc;
        k_fill_u32<<<hier_grid(nbins), kHierBlock, 0, st>>>(first.p, nbins, RS_EOD);
        BinArgs a{ d_xyz, node_of.p, rep.p, by_component ? comp_of.p : nullptr, box.p, cur, nb[0], nb[1], nb[2], dpn };
        k_bin<<<hier_grid(cur), kHierBlock, 0, st>>>(a, bin.p, first.p);
        k_is_first<<<hier_grid(cur), kHierBlock, 0, st>>>(bin.p, first.p, cur, flag.p);
        RS_HIP(hipGetLastError());
        if ((rc = scan_u32(tmp, flag.p, rank.p, cur, st))) return rc;
        uint32_t last[2] = { 0, 0 };
        RS_HIP(hipMemcpyAsync(&last[0], rank.p + (cur - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        RS_HIP(hipMemcpyAsync(&last[1], flag.p + (cur - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        RS_HIP(hipStreamSynchronize(st));
        const uint32_t count = last[0] + last[1];
        if (count == cur) {
            if (nb[0] == 1 && nb[1] == 1 && nb[2] == 1)
                return fail(RS_ERR_UNSUPPORTED, "hierarchy by bins: cannot coarsen further");
            continue;   // the bins are still as fine as the nodes: halve again
        }
        std::unique_ptr<DBuf<uint32_t>> agg(new DBuf<uint32_t>());
        if ((rc = agg->alloc(cur))) return rc;
        k_assign<<<hier_grid(cur), kHierBlock, 0, st>>>(bin.p, first.p, rank.p, flag.p, rep.p, cur, agg->p, next_rep.p);
        RS_HIP(hipGetLastError());
        maps.push_back(std::move(agg));
        n_agg.push_back(count);
        levels++;
        cur = count;
        std::swap(rep.p, next_rep.p);
        std::swap(rep.n, next_rep.n);
    }
    RS_HIP(hipStreamSynchronize(st));
    return RS_OK;
}

}  // namespace rs
