// Internal declarations of the MI355X JVIE T-matrix library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/scadyn_amd.h"

namespace scadyn {

// ---------------------------------------------------------------- errors
struct Error {
    int code;
    std::string msg;
};
void set_error(int code, const std::string &msg);
[[noreturn]] void fail(int code, const std::string &msg);

#define HIP_CHECK(expr)                                                              \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            ::scadyn::fail(SCADYN_ERR_DEVICE, std::string(#expr) + ": " +            \
                                                  hipGetErrorString(e_));            \
    } while (0)

// ---------------------------------------------------------------- complex
typedef double2 cd;
typedef float2 cf;
// complex with IEEE binary16 parts: the storage of a reduced-precision Krylov
// basis (never computed in; converted through float on every load and store)
struct __attribute__((aligned(4))) ch {
    _Float16 x, y;
};

// Workgroup b of a G-workgroup launch -> work item, a bijection of [0, G).
// Workgroups are dealt to the 8 XCDs round-robin (b % 8), each XCD with its own
// L2; XCD x is given the x-th contiguous range of items, the first G % 8 ranges
// one longer, so that neighbouring items (which share cache lines) run on the
// same XCD.
__host__ __device__ inline long long xcd_remap(long long b, long long G) {
    const long long per = (G + 7) / 8, tall = G % 8 == 0 ? 8 : G % 8;
    const long long x = b % 8, j = b / 8;
    return x < tall ? x * per + j : tall * per + (x - tall) * (per - 1) + j;
}

__host__ __device__ inline cd cmk(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline cd cadd(cd a, cd b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cd csub(cd a, cd b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cd cmul(cd a, cd b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline cd cmulc(cd a, cd b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline cd cscale(cd a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline cd cconj(cd a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double cabs2(cd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline cd cdiv(cd a, cd b) {
    // plain formula a * conj(b) / |b|^2, as compilers emit for complex(8) division
    double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// acc += a * b
__host__ __device__ inline void cfma(cd &acc, cd a, cd b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}

// single-precision complex (the mixed-precision matvec)
__host__ __device__ inline cf cmkf(float r, float i) { return make_float2(r, i); }
__host__ __device__ inline cf cadd(cf a, cf b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cf csub(cf a, cf b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cf cmul(cf a, cf b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline cf cscale(cf a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline void cfma(cf &acc, cf a, cf b) {
    acc.x = fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x));
    acc.y = fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y));
}

// precision traits: real type and conversions
template <typename C>
struct Prec;
template <>
struct Prec<cd> {
    typedef double real;
    __host__ __device__ static cd from(cd a) { return a; }
    __host__ __device__ static cd from(cf a) { return make_double2(a.x, a.y); }
    __host__ __device__ static cd from(ch a) { return make_double2((double)(float)a.x, (double)(float)a.y); }
    __host__ __device__ static cd zero() { return make_double2(0.0, 0.0); }
};
template <>
struct Prec<cf> {
    typedef float real;
    __host__ __device__ static cf from(cd a) { return make_float2((float)a.x, (float)a.y); }
    __host__ __device__ static cf from(cf a) { return a; }
    __host__ __device__ static cf from(ch a) { return make_float2((float)a.x, (float)a.y); }
    __host__ __device__ static cf zero() { return make_float2(0.0f, 0.0f); }
};
template <>
struct Prec<ch> {
    typedef float real;  // arithmetic on a stored element runs in float or double
    // double -> float -> binary16, both rounding to nearest even: the same
    // result on the host and the device.  Contraction is off here so that the
    // two conversions are not merged into one double -> binary16 conversion,
    // which gfx950 has no instruction for (it becomes some 35 integer
    // instructions per value instead of two conversions).
    __host__ __device__ static ch from(cd a) {
#pragma clang fp contract(off)
        const float fx = (float)a.x, fy = (float)a.y;
        return {(_Float16)fx, (_Float16)fy};
    }
    __host__ __device__ static ch from(cf a) { return {(_Float16)a.x, (_Float16)a.y}; }
    __host__ __device__ static ch zero() { return {(_Float16)0.0f, (_Float16)0.0f}; }
};

// ---------------------------------------------------------------- geometry (host)
struct HostGeometry {
    int N = 0, Nnode = 0;
    int Nx = 0, Ny = 0, Nz = 0;             // grid nodes per axis (src/geometry.f90:90-98)
    int ncx = 0, ncy = 0, ncz = 0;          // cubes per axis (src/geometry.f90:211-217)
    int N_cubes = 0, N_tet_cube = 0;
    double delta = 0, box_delta = 0;
    double min_coord[3] = {0, 0, 0};
    std::vector<int> perm;        // internal tet i -> reference tet perm[i] (0-based)
    std::vector<int> inv;         // reference tet -> internal
    std::vector<int> tet_cube;    // per internal tet
    std::vector<int> cube_start;  // CSR over cubes, internal tets sorted by cube
    std::vector<double> tcoord;   // internal tets: [N][4][3]
    std::vector<double> vol;      // internal tets
    std::vector<cd> param;        // internal tets
};

// Builds the grid/cube structures; throws on orientation and grid errors.
void build_geometry(const scadyn_mesh &mesh, const scadyn_opts &opts, HostGeometry &g);

// Near-zone pair list in the reference's storage order: for block p,
// (T[p], B[p]) are internal tet ids with reference ids perm[T] <= perm[B],
// off[p] in [0,27) is the cube offset (dx+1) + 3(dy+1) + 9(dz+1) from T's cube
// to B's cube.
struct NearPairs {
    std::vector<int> T, B;
    std::vector<uint8_t> off;
    // row lists over internal tets: entries (column tet, block id | TRANSPOSE)
    std::vector<int64_t> row_start;
    std::vector<int> ent_col;
    std::vector<uint32_t> ent_blk;
    // true when every row lists exactly the tets of its cube's near cubes, in
    // near-cube order (dz, dy, dx ascending) and ascending tet id within a
    // cube: then the entries of all tets of one cube line up by position
    bool cube_rows = false;
};
constexpr uint32_t BLK_TRANSPOSE = 0x80000000u;
void build_near_pairs(const HostGeometry &g, NearPairs &np);

// SVD pseudo-inverse (one-sided Jacobi), singular values <= thr dropped.
// A is m x n column-major; returns n x m column-major.
std::vector<cd> pinv_jacobi(const std::vector<cd> &A, int m, int n, double thr);

// Smallest 2,3,5,7-smooth length >= n, and its radix list (largest first).
int smooth_length(int n);
std::vector<int> factor_radices(int L);

}  // namespace scadyn
