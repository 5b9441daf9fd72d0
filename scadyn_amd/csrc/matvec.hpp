// Matvec kernel launchers.
#pragma once

#include "internal.hpp"

namespace scadyn {

template <typename C>
struct MatvecArgs {
    // input multi-vector: element (row, b) at xin[in_off[b] + row*B + b]
    const cd *xin = nullptr;
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
    double k2 = 0;
    const int *slot_active = nullptr;  // optional per-slot mask
    C *grid_in = nullptr;              // scatter output [Nz][Ny][Nx][4][B]
    const C *grid_out = nullptr;       // convolved grid, same layout
    cd *y = nullptr;                   // output [3N][B]
};

template <typename C>
void launch_scatter(const MatvecArgs<C> &a, hipStream_t s);
template <typename C>
void launch_gather_near(const MatvecArgs<C> &a, hipStream_t s);

}  // namespace scadyn
