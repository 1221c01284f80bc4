// rs_internal.hpp -- shared internal declarations of the range-solver MI355X library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "range_solver_amd.h"

namespace rs {

// ----------------------------------------------------------------- errors / runtime
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define RS_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return ::rs::hip_fail(_e, #call);             \
    } while (0)

struct Runtime {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    int n_cu = 0;
};
Runtime &runtime();
int require_runtime();
// Looks up, on `device` (the current device), the address of the allocation-free solve's
// workspace array; called by rs_init once per device.
int scratch_setup(int device);

// RConstants::eps lives in range-base-lib (absent); value chosen here.
constexpr double kEps = 1.0e-14;

// ----------------------------------------------------------------- device arrays
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count);
    void release();
    ~DBuf() { release(); }
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

template <typename T>
int DBuf<T>::alloc(size_t count)
{
    if (count == n && p) return RS_OK;
    release();
    if (count == 0) return RS_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    n = count;
    return RS_OK;
}

template <typename T>
void DBuf<T>::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
}

// Two timing events, destroyed on every return path of the function that owns them.
struct EventPair {
    hipEvent_t e[2] = { nullptr, nullptr };
    hipError_t create()
    {
        hipError_t r = hipEventCreate(&e[0]);
        return r == hipSuccess ? hipEventCreate(&e[1]) : r;
    }
    ~EventPair()
    {
        for (hipEvent_t &x : e)
            if (x) (void)hipEventDestroy(x);
    }
    EventPair() = default;
    EventPair(const EventPair &) = delete;
    EventPair &operator=(const EventPair &) = delete;
};

// ----------------------------------------------------------------- host preparation
// Parallel loop over [0, n) on host threads.
void parallel_for(uint64_t n, const std::function<void(uint64_t, uint64_t)> &fn);

struct Pattern {
    uint32_t n_dof = 0;
    std::vector<uint32_t> book;      // n_nodes*dpn -> dof or RS_EOD
    std::vector<uint32_t> rowptr;    // n_dof+1
    std::vector<uint32_t> col;       // nnz, sorted per row
};

// CSR pattern of the free DOFs of a mesh whose elements couple all their nodes.
int build_pattern(uint32_t n_nodes, uint32_t n_elems, uint32_t npe, const uint32_t *conn,
                  uint32_t dpn, const uint32_t *book, uint32_t n_dof, Pattern &pat);

// SpMV tiles: consecutive rows grouped so that a tile has <= max_rows rows and <= cap
// nonzeros (a single longer row forms its own tile).  Returns ntiles+1 start rows.
std::vector<uint32_t> build_tiles(const uint32_t *rowptr, uint32_t n, uint32_t cap, uint32_t max_rows);

// Greedy node-disjoint colouring; returns colour per element and the number of colours.
int colour_elements(uint32_t n_nodes, uint32_t n_elems, uint32_t npe, const uint32_t *conn,
                    std::vector<uint16_t> &colour, uint32_t &n_colours);

// ----------------------------------------------------------------- device matrix
struct Matrix {
    uint32_t n = 0;
    uint64_t nnz = 0;
    double frob = 0.0;
    DBuf<uint32_t> rowptr;
    DBuf<uint32_t> col;
    DBuf<double> val;
    DBuf<float> val32;          // FP32 copy for the mixed-precision inner solve
    bool val32_valid = false;
    bool borrowed = false;      // owned by an assembly
    uint64_t version = 0;       // bumped on every value update
    DBuf<uint32_t> tiles[2];    // SpMV tile starts for the two tile capacities
    uint32_t ntiles[2] = { 0, 0 };
    // SELL-64 copy: slice offsets, column-major columns (padding RS_EOD) and values
    DBuf<uint32_t> sell_ptr, sell_col;
    DBuf<int16_t> sell_off;     // int16 column offsets from each slice's first row
    bool sell_off_ok = false;   // every column fits such an offset
    DBuf<double> sell_val;
    DBuf<float> sell_val32;
    uint32_t nslices = 0;
    bool sell_valid = false, sell32_valid = false;
    int plan[2] = { -1, -1 };   // autotuned SpMV shape for FP64 / FP32
    double plan_ms[2][11] = {}; // tuning times per shape
};

// Every change of the FP64 values goes through here: derived copies become stale.
inline void matrix_values_changed(Matrix &A)
{
    A.val32_valid = false;
    A.sell_valid = false;
    A.sell32_valid = false;
    A.version++;
}

int matrix_frobenius(Matrix &A);   // device reduction, sets A.frob
int matrix_refresh_fp32(Matrix &A);
int matrix_set_layout(Matrix &A, const uint32_t *host_rowptr);
int matrix_tune(Matrix &A, bool fp32);
// A matrix with the given device pattern (taken over: rowptr and col are left empty) and zero
// values.
int matrix_create_device(uint32_t n, DBuf<uint32_t> &rowptr, DBuf<uint32_t> &col, uint64_t nnz, rs_matrix **out);

// ----------------------------------------------------------------- aggregation hierarchy (GPU)
// Nested coordinate bins over the DOFs of a mesh (device coordinates and DOF book), halved per
// level until a level has <= max_coarse nodes: maps[k] sends the nodes of level k to level k + 1.
// by_component: one aggregate per bin and DOF component (the DOFs of a node go to dpn different
// aggregates, every coarse node keeps the component of its members); else the DOFs of a node share
// their bin's aggregate.
int hier_bins(uint32_t n_nodes, const double *d_xyz, uint32_t dpn, const uint32_t *d_book, uint32_t n_dof,
              const uint32_t bins[3], uint32_t max_coarse, uint32_t max_levels, bool by_component,
              std::vector<std::unique_ptr<DBuf<uint32_t>>> &maps, std::vector<uint32_t> &n_agg);
// Scratch of the two builders below, kept across the levels of one hierarchy build so that the
// finest level's sort buffers (hundreds of MB at 128^3) are allocated once and reused by the
// coarser levels instead of a hipMalloc / hipFree pair per buffer and level.  Grow-only; the
// builders pass every size explicitly.  Null: each call allocates its own.
struct HierWork {
    DBuf<uint64_t> kin, kout;
    DBuf<uint32_t> pin, flag, run, count, vin, kout32;
    DBuf<uint8_t> tmp;
};
// Members of each of the nn aggregates of d_map (n entries), rows ascending: ptr (nn + 1), rows (n).
int hier_members(const uint32_t *d_map, uint32_t n, uint32_t nn, DBuf<uint32_t> &ptr, DBuf<uint32_t> &rows,
                 HierWork *work = nullptr);
// Next-level pattern (cptr, ccol) of a level matrix under d_map, and the gather lists: the entries
// of the level matrix summed into each next-level entry, by position (gptr: cnnz + 1, gsrc: nnz).
int hier_level_lists(const uint32_t *d_rowptr, const uint32_t *d_col, uint32_t n, uint64_t nnz, const uint32_t *d_map,
                     uint32_t nn, DBuf<uint32_t> &cptr, DBuf<uint32_t> &ccol, DBuf<uint32_t> &gptr,
                     DBuf<uint32_t> &gsrc, uint32_t &cnnz, HierWork *work = nullptr);
int hier_compose(uint32_t *d_total, uint64_t n, const uint32_t *d_map);   // total[i] = map[total[i]]
int hier_identity(uint32_t *d_a, uint64_t n);

}  // namespace rs

// C-ABI handle types are the internal structs.
struct rs_matrix : rs::Matrix {};
