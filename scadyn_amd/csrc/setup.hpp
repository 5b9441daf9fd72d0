// Setup-kernel launchers (FP64).
#pragma once

#include "internal.hpp"

namespace scadyn {

struct GridDesc {
    double min[3];
    double delta, box_delta;
    int Nx, Ny, Nz;
    int ncx, ncy, ncz;
};

void upload_tables(const double *cube98_scaled, hipStream_t s);
const double *host_tet24_p();
const double *host_tet24_w();
const double *host_tri7_p();
const double *host_tri7_w();

void launch_projection(const double *tcoord, const double *vol, const int *tet_cube, const GridDesc &gd,
                       double k, const cd *WW, cd *S4, int N, hipStream_t s);
void launch_green_grid(cd *A, const GridDesc &gd, int Lx, int Ly, int Lz, double k, hipStream_t s);
void launch_fg_octant(const cd *F, cd *fg, int Lx, int Ly, int Lz, hipStream_t s);
void launch_pg_table(cd *Gtab, const GridDesc &gd, double k, hipStream_t s);
void launch_near_blocks(const int *PT, const int *PB, const uint8_t *POFF, long long P, const double *tcoord,
                        const double *vol, const cd *param, const cd *S4, const cd *Gtab, double k, cf *blocks,
                        hipStream_t s);
void launch_vswf(const double *tcoord, const double *vol, int N, int Nmax, double k, cd *mat, int ncol,
                 hipStream_t s);

}  // namespace scadyn
