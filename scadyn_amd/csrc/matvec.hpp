// Matvec kernel launchers.
#pragma once

#include "internal.hpp"

namespace scadyn {

template <typename C>
struct MatvecArgs {
    // input multi-vector in the matvec precision: element (row, b) at
    // xin[in_off[b] + row*B + b]
    const C *xin = nullptr;
    const long long *in_off = nullptr;
    int B = 0;
    int N = 0;
    int Nx = 0, Ny = 0, Nz = 0, ncx = 0, ncy = 0, ncz = 0;
    const int *cube_start = nullptr;
    const int *tet_cube = nullptr;
    const C *S4 = nullptr;
    const int64_t *row_start = nullptr;
    const int *ent_col = nullptr;
    const uint32_t *ent_blk = nullptr;
    const cf *blocks = nullptr;
    // non-empty cubes; set only when the near-zone rows are in cube order
    // (NearPairs::cube_rows), which the per-cube gather relies on
    const int *ne_cubes = nullptr;
    int n_ne = 0;
    double k2 = 0;
    const int *slot_active = nullptr;  // optional per-slot mask
    int gather_kernel = 0;             // complex64: 1 = near zone on the matrix cores
    C *grid_in = nullptr;              // scatter output [Nz][Ny][Nx][4][B]
    const C *grid_out = nullptr;       // convolved grid, same layout
    cd *y = nullptr;                   // output [3N][B]
};

template <typename C>
void launch_scatter(const MatvecArgs<C> &a, hipStream_t s);
template <typename C>
void launch_gather_near(const MatvecArgs<C> &a, hipStream_t s);

}  // namespace scadyn
