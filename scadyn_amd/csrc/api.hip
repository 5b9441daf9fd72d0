// Operator lifecycle, batched solve driver and the C ABI (include/scadyn_amd.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "fft.hpp"
#include "internal.hpp"
#include "krylov.hpp"
#include "matvec.hpp"
#include "setup.hpp"

namespace scadyn {

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;
void set_error(int, const std::string &msg) { g_err = msg; }
void fail(int code, const std::string &msg) { throw Error{code, msg}; }

// ---------------------------------------------------------------- device memory
struct DevMem {
    std::vector<void *> ptrs;
    size_t bytes = 0;
    template <class T>
    T *alloc(size_t n) {
        if (n == 0) n = 1;
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess)
            fail(SCADYN_ERR_MEMORY, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                                        hipGetErrorString(e));
        ptrs.push_back(p);
        bytes += n * sizeof(T);
        return (T *)p;
    }
    template <class T>
    T *upload(const std::vector<T> &v, hipStream_t s) {
        T *p = alloc<T>(v.size());
        // synchronous: the host vector may be a temporary
        if (!v.empty()) {
            HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
            HIP_CHECK(hipStreamSynchronize(s));
        }
        return p;
    }
    void release(void *p) {
        auto it = std::find(ptrs.begin(), ptrs.end(), p);
        if (it != ptrs.end()) {
            (void)hipFree(p);
            ptrs.erase(it);
        }
    }
    void clear() {
        for (void *p : ptrs) (void)hipFree(p);
        ptrs.clear();
        bytes = 0;
    }
    ~DevMem() { clear(); }
};

// ---------------------------------------------------------------- profiling
struct Profiler {
    bool on = false;
    struct Pending {
        int cls;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pend;
    std::vector<hipEvent_t> pool;
    // per kernel class, FP64 launches in [0, COUNT), complex64 ones in [COUNT, 2 COUNT)
    double ms[2 * SCADYN_KCLASS_COUNT] = {0}, cnt[2 * SCADYN_KCLASS_COUNT] = {0}, byt[2 * SCADYN_KCLASS_COUNT] = {0};
    hipEvent_t ev() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        return e;
    }
    void begin(int cls, double bytes, hipStream_t s, Pending &p) {
        p.cls = cls;
        p.bytes = bytes;
        p.a = ev();
        p.b = ev();
        HIP_CHECK(hipEventRecord(p.a, s));
    }
    void end(Pending &p, hipStream_t s) {
        HIP_CHECK(hipEventRecord(p.b, s));
        pend.push_back(p);
    }
    void collect() {
        for (auto &p : pend) {
            HIP_CHECK(hipEventSynchronize(p.b));
            float t = 0;
            HIP_CHECK(hipEventElapsedTime(&t, p.a, p.b));
            ms[p.cls] += t;
            cnt[p.cls] += 1;
            byt[p.cls] += p.bytes;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pend.clear();
    }
    ~Profiler() {
        for (auto &p : pend) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------- workspace
// Grids, FFT plans and input offsets of one batched matvec of B slots.
struct MatvecSpace {
    int B = 0;
    DevMem mem;
    void *grid0 = nullptr, *grid1 = nullptr;  // complex128 or complex64 grids (one matvec at a time)
    long long *in_off = nullptr;              // slot b's input at xin + in_off[b] + row*B + b
    int *active = nullptr;
    LinePass pass[2][5];  // [fp64, fp32][x fwd, y fwd, z conv, y inv, x inv]
};

struct Workspace {
    int B = 0, m = 0;
    long long n = 0;
    DevMem mem;
    MatvecSpace mv;  // the batch's matvec
    // mixed precision: FP64 matvec of the restarting slots only, packed into
    // a narrow batch of at most rs_cap slots (x in rx, A x in ry)
    std::unique_ptr<MatvecSpace> rs;
    int rs_cap = 0;
    cd *rx = nullptr, *ry = nullptr;
    int *rslot = nullptr;
    // Krylov basis: complex128 V [(m+2)][n][B] with x in block m+1 (X points
    // there), or for the mixed-precision solve (ortho 0, mgs_block > 1) a
    // complex64 basis Vf or a binary16 one Vh [(m+1)][n][B] and x in its own
    // complex128 X
    cd *V = nullptr, *Rhs = nullptr, *W = nullptr, *X = nullptr;
    cf *Vf = nullptr, *Wf = nullptr;  // complex64 basis and orthogonalised w (mixed precision)
    ch *Vh = nullptr;                 // binary16 basis (mixed precision; Xs holds the matvec input)
    // mixed precision with a complex128 or binary16 basis: the batch's inputs
    // converted to complex64 (Xs [n][B], read with zero offsets zoff)
    cf *Xs = nullptr;
    long long *zoff = nullptr;
    int *ract = nullptr;  // mixed precision: restarting slots of a full-width FP64 matvec
    int *ierr = nullptr;  // set by the device when a matvec input offset was out of range
    SlotState *st = nullptr;
    cd *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr, *y = nullptr;
    double *nrm2 = nullptr, *nav2 = nullptr, *pnrm = nullptr, *pnav = nullptr;
    cd *pdot = nullptr;
    int *col = nullptr, *ret = nullptr, *flag = nullptr;
    cd *cc = nullptr, *hprev = nullptr;
    int nwg = 0;
};

}  // namespace scadyn

using namespace scadyn;

struct scadyn_operator {
    int device = 0;
    hipStream_t s = nullptr;
    scadyn_opts opts;
    double k = 0;
    HostGeometry g;
    NearPairs np;
    GridDesc gd;
    int L[3] = {0, 0, 0}, Hh[3] = {0, 0, 0};
    DevMem mem;
    double *d_tcoord = nullptr, *d_vol = nullptr;
    cd *d_param = nullptr, *d_S4 = nullptr, *d_fg = nullptr, *d_tw = nullptr;
    cf *d_S4f = nullptr, *d_fgf = nullptr, *d_twf = nullptr;  // complex64 copies (mixed precision)
    int *d_tet_cube = nullptr, *d_cube_start = nullptr, *d_perm = nullptr;
    int *d_ne_cubes = nullptr;  // non-empty cubes (only with np.cube_rows)
    int n_ne_cubes = 0;
    cf *d_blocks = nullptr;
    int64_t *d_row_start = nullptr;
    int *d_ent_col = nullptr;
    uint32_t *d_ent_blk = nullptr;
    long long nblocks = 0;
    std::unique_ptr<Workspace> ws;
    Profiler prof;
    int32_t last_outer_rounds = 0;
    // times the reference's re-orthogonalisation test (src/gmres_module.f90:251)
    // fired, summed over the systems finished on this operator
    long long reorth_fired = 0;
    ~scadyn_operator() {
        ws.reset();
        mem.clear();
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace scadyn {

static const double PI_H = 3.141592653589793;

static std::vector<cd> twiddle_table(int L) {
    std::vector<cd> t(L);
    for (int m = 0; m < L; ++m) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)L;
        t[m] = cmk((double)cosl(a), (double)sinl(a));
    }
    return t;
}

static void fft3d_full(scadyn_operator *op, cd *A, cd *tmp) {
    // forward unnormalised 3D DFT of a [Lz][Ly][Lx] grid (FFTW_FORWARD sign)
    const int Lx = op->L[0], Ly = op->L[1], Lz = op->L[2];
    const size_t budget = 64 * 1024;
    LinePass p;
    const bool reg = op->opts.fft_kernel == 0;
    make_line_pass(p, Lx, Lx, Lx, 1, (long long)Ly * Lz, budget, sizeof(cd), reg);
    launch_line_fft<cd>(p, A, tmp, op->d_tw, -1, op->s);
    make_line_pass(p, Ly, Ly, Ly, Lx, Lz, budget, sizeof(cd), reg);
    launch_line_fft<cd>(p, tmp, A, op->d_tw + Lx, -1, op->s);
    make_line_pass(p, Lz, Lz, Lz, (long long)Lx * Ly, 1, budget, sizeof(cd), reg);
    launch_line_fft<cd>(p, A, tmp, op->d_tw + Lx + Ly, -1, op->s);
    HIP_CHECK(hipMemcpyAsync(A, tmp, sizeof(cd) * (size_t)Lx * Ly * Lz, hipMemcpyDeviceToDevice, op->s));
}

// basis vectors per MGS pass when opts.mgs_block = 0: 4 moves 2.5 vectors of
// HBM traffic per coefficient against 4 for one per pass, at 98 VGPRs (8 needs
// 234 and runs slower at 2 waves per SIMD)
constexpr int MGS_BLOCK_AUTO = 4;

// The mixed-precision solve keeps its Krylov basis in complex64 (arithmetic
// stays FP64) when the blocked MGS runs: the basis is what its sweeps and the
// complex64 matvec read.
static bool f32_basis_for(const scadyn_opts &o) {
    const int bs = o.mgs_block == 0 ? MGS_BLOCK_AUTO : o.mgs_block;
    return o.precision == SCADYN_PREC_MIXED && o.ortho == 0 && bs > 1;
}
static bool f32_basis(const scadyn_operator *op) { return f32_basis_for(op->opts); }
// Of those, the ones whose basis is stored in binary16.  A unit basis vector
// of n rows rounds with a relative error below 2^-11 per element; elements
// under binary16's normal range (2^-14) add at most sqrt(n) 2^-25 to the
// vector's norm error, which stays below 2^-12 while n <= 2^26.  The rounding
// limits what one restart cycle can reduce the residual by to about 2^-10, so
// such a cycle ends at that reduction at the latest (F16_CYCLE_TOL): a cycle
// asked for more would run on without gaining it.  The automatic choice takes
// binary16 from 2^20 rows, where the basis sweeps dominate the memory traffic;
// smaller systems, which often converge in a few short cycles, keep complex64.
constexpr long long F16_BASIS_MAX_ROWS = 1LL << 26;
constexpr long long F16_BASIS_AUTO_MIN_ROWS = 1LL << 20;
constexpr double F16_CYCLE_TOL = 1.0 / 1024.0;
static bool f16_basis_for(const scadyn_opts &o, long long n) {
    if (!f32_basis_for(o)) return false;
    if (o.basis == SCADYN_BASIS_F16) return true;
    return o.basis == SCADYN_BASIS_AUTO && n >= F16_BASIS_AUTO_MIN_ROWS && n <= F16_BASIS_MAX_ROWS;
}
static bool f16_basis(const scadyn_operator *op) { return f16_basis_for(op->opts, 3LL * op->g.N); }

// Krylov vectors of one slot of n rows with restart length m: the binary16
// basis Vh with x, W, Rhs (complex128), the complex64 w copy and matvec input;
// the complex64 basis Vf with x, W, Rhs and the w copy; or the complex128
// basis with x, W, Rhs and one vector of slack
static size_t slot_vec_bytes(const scadyn_opts &o, size_t n, int m) {
    if (f16_basis_for(o, (long long)n)) return (size_t)(m + 1) * n * sizeof(ch) + n * (3 * sizeof(cd) + 2 * sizeof(cf));
    if (f32_basis_for(o)) return (size_t)(m + 1) * n * sizeof(cf) + n * (3 * sizeof(cd) + sizeof(cf));
    return (size_t)(m + 4) * n * sizeof(cd);
}

// xs[row*B + b] = (complex64) v[in_off[b] + row*B + b] for active slots
__global__ void k_stage_f32(const cd *__restrict__ v, const long long *__restrict__ in_off,
                            const int *__restrict__ active, long long n, int B, cf *__restrict__ xs) {
    const long long tot = n * B;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e % B);
        if (!active[b]) continue;
        const cd a = v[in_off[b] + e];
        xs[e] = make_float2((float)a.x, (float)a.y);
    }
}
static void launch_stage_f32(const cd *v, const long long *in_off, const int *active, long long n, int B, cf *xs,
                             hipStream_t s) {
    const long long tot = n * B;
    const unsigned g = (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16);
    hipLaunchKernelGGL(k_stage_f32, dim3(g), dim3(256), 0, s, v, in_off, active, n, B, xs);
    HIP_CHECK(hipGetLastError());
}

__global__ void k_to_f32(const cd *__restrict__ a, cf *__restrict__ b, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        b[e] = make_float2((float)a[e].x, (float)a[e].y);
}

// complex64 copies of S*, FFT(G) and the twiddles for the mixed-precision matvec
static void make_fp32_copies(scadyn_operator *op) {
    if (op->d_S4f) return;
    DevMem &M = op->mem;
    const size_t nS = (size_t)op->g.N * 108, nF = (size_t)op->Hh[0] * op->Hh[1] * op->Hh[2];
    op->d_S4f = M.alloc<cf>(nS);
    op->d_fgf = M.alloc<cf>(nF);
    hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, op->s, op->d_S4, op->d_S4f, nS);
    hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, op->s, op->d_fg, op->d_fgf, nF);
    HIP_CHECK(hipGetLastError());
    std::vector<cf> tw;
    for (int c = 0; c < 3; ++c) {
        auto t = twiddle_table(op->L[c]);
        for (auto &v : t) tw.push_back(make_float2((float)v.x, (float)v.y));
    }
    op->d_twf = M.upload(tw, op->s);
    HIP_CHECK(hipStreamSynchronize(op->s));
}

// Morton code of cube coordinates (bits of x, y, z interleaved, 21 per axis)
static uint64_t morton3(int x, int y, int z) {
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffULL;
        v = (v | v << 16) & 0x1f0000ff0000ffULL;
        v = (v | v << 8) & 0x100f00f00f00f00fULL;
        v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
        v = (v | v << 2) & 0x1249249249249249ULL;
        return v;
    };
    return spread((uint64_t)x) | spread((uint64_t)y) << 1 | spread((uint64_t)z) << 2;
}

static void create_operator(scadyn_operator *op, const scadyn_mesh &mesh, const scadyn_opts &opts, double k) {
    if (!(k > 0.0)) fail(SCADYN_ERR_ARG, "wavenumber must be positive");
    op->opts = opts;
    op->k = k;
    op->device = opts.device;
    // host-side validation and indexing first: mesh errors surface without a device
    HostGeometry &g = op->g;
    build_geometry(mesh, opts, g);
    build_near_pairs(g, op->np);
    HIP_CHECK(hipSetDevice(op->device));
    HIP_CHECK(hipStreamCreateWithFlags(&op->s, hipStreamNonBlocking));
    hipStream_t s = op->s;
    op->nblocks = (long long)op->np.T.size();
    GridDesc &gd = op->gd;
    for (int c = 0; c < 3; ++c) gd.min[c] = g.min_coord[c];
    gd.delta = g.delta;
    gd.box_delta = g.box_delta;
    gd.Nx = g.Nx;
    gd.Ny = g.Ny;
    gd.Nz = g.Nz;
    gd.ncx = g.ncx;
    gd.ncy = g.ncy;
    gd.ncz = g.ncz;
    const int Nax[3] = {g.Nx, g.Ny, g.Nz};
    for (int c = 0; c < 3; ++c) {
        op->L[c] = opts.fft_pad == 1 ? 2 * Nax[c] : smooth_length(2 * Nax[c] - 1);
        op->Hh[c] = op->L[c] / 2 + 1;
    }
    DevMem &M = op->mem;
    const int N = g.N;
    op->d_tcoord = M.upload(g.tcoord, s);
    op->d_vol = M.upload(g.vol, s);
    op->d_param = M.upload(g.param, s);
    op->d_tet_cube = M.upload(g.tet_cube, s);
    op->d_cube_start = M.upload(g.cube_start, s);
    op->d_perm = M.upload(g.perm, s);
    if (op->np.cube_rows) {
        // non-empty cubes in Morton (Z-curve) order of their coordinates: the
        // gather's waves in flight at once then cover a compact block of
        // cubes rather than a slab of whole x rows (measured 2-3% faster at
        // the same HBM traffic)
        std::vector<std::pair<uint64_t, int>> key;
        for (int c = 0; c < g.N_cubes; ++c)
            if (g.cube_start[c + 1] > g.cube_start[c]) {
                const int cx = c % g.ncx, cy = (c / g.ncx) % g.ncy, cz = c / (g.ncx * g.ncy);
                key.push_back({morton3(cx, cy, cz), c});
            }
        std::sort(key.begin(), key.end());
        std::vector<int> ne;
        for (const auto &kc : key) ne.push_back(kc.second);
        op->n_ne_cubes = (int)ne.size();
        if (!ne.empty()) op->d_ne_cubes = M.upload(ne, s);
    }
    {
        std::vector<cd> tw;
        for (int c = 0; c < 3; ++c) {
            auto t = twiddle_table(op->L[c]);
            tw.insert(tw.end(), t.begin(), t.end());
        }
        op->d_tw = M.upload(tw, s);
    }
    // ---- W (98 x 27) from the first tet's cube and its pseudo-inverse
    //      (src/projection.f90:71-101)
    std::vector<double> P0;
    {
        const double gv[5] = {-1.0, -0.5, 0.0, 0.5, 1.0};
        for (int z = 0; z < 5; ++z)
            for (int y = 0; y < 5; ++y)
                for (int x = 0; x < 5; ++x)
                    if (std::max(std::fabs(gv[x]), std::max(std::fabs(gv[y]), std::fabs(gv[z]))) == 1.0) {
                        P0.push_back(gv[x]);
                        P0.push_back(gv[y]);
                        P0.push_back(gv[z]);
                    }
    }
    std::vector<double> Psph(98 * 3);
    for (int i = 0; i < 98 * 3; ++i) Psph[i] = 2.0 * g.box_delta * P0[i];
    {
        const int c0 = g.tet_cube[g.inv[0]];
        const int cx = c0 % g.ncx, cy = (c0 / g.ncx) % g.ncy, cz = c0 / (g.ncx * g.ncy);
        const int base[3] = {2 * cx, 2 * cy, 2 * cz};
        double node1[3], cp[3];
        for (int c = 0; c < 3; ++c) {
            node1[c] = g.min_coord[c] + g.delta * base[c];
            cp[c] = node1[c] + g.box_delta / 2.0;
        }
        std::vector<cd> W(98 * 27);
        for (int u = 0; u < 27; ++u) {
            const int o[3] = {u % 3, (u / 3) % 3, u / 9};
            double pn[3];
            for (int c = 0; c < 3; ++c) pn[c] = g.min_coord[c] + g.delta * (base[c] + o[c]);
            for (int m = 0; m < 98; ++m) {
                double d2 = 0;
                for (int c = 0; c < 3; ++c) {
                    double a = (Psph[3 * m + c] + cp[c]) - pn[c];
                    d2 += a * a;
                }
                double R = std::sqrt(d2), den = 4.0 * PI_H * R;
                W[(size_t)u * 98 + m] = cmk(std::cos(k * R) / den, std::sin(k * R) / den);
            }
        }
        std::vector<cd> WW = pinv_jacobi(W, 98, 27, 1e-6);
        cd *d_WW = M.upload(WW, s);
        upload_tables(Psph.data(), s);
        op->d_S4 = M.alloc<cd>((size_t)N * 108);
        launch_projection(op->d_tcoord, op->d_vol, op->d_tet_cube, gd, k, d_WW, op->d_S4, N, s);
        HIP_CHECK(hipStreamSynchronize(s));
        M.release(d_WW);
    }
    // ---- near zone
    {
        cd *d_G = M.alloc<cd>(27 * 27 * 27);
        launch_pg_table(d_G, gd, k, s);
        int *d_PT = M.upload(op->np.T, s);
        int *d_PB = M.upload(op->np.B, s);
        uint8_t *d_PO = M.upload(op->np.off, s);
        op->d_blocks = M.alloc<cf>((size_t)op->nblocks * 9);
        launch_near_blocks(d_PT, d_PB, d_PO, op->nblocks, op->d_tcoord, op->d_vol, op->d_param, op->d_S4, d_G, k,
                           op->d_blocks, s);
        op->d_row_start = M.upload(op->np.row_start, s);
        op->d_ent_col = M.upload(op->np.ent_col, s);
        op->d_ent_blk = M.upload(op->np.ent_blk, s);
        HIP_CHECK(hipStreamSynchronize(s));
        M.release(d_PT);
        M.release(d_PB);
        M.release(d_PO);
        M.release(d_G);
        std::vector<int>().swap(op->np.ent_col);
        std::vector<uint32_t>().swap(op->np.ent_blk);
        std::vector<uint8_t>().swap(op->np.off);
    }
    // ---- FFT(G) (src/build_G.f90)
    {
        const size_t tot = (size_t)op->L[0] * op->L[1] * op->L[2];
        cd *A = M.alloc<cd>(tot), *T = M.alloc<cd>(tot);
        launch_green_grid(A, gd, op->L[0], op->L[1], op->L[2], k, s);
        fft3d_full(op, A, T);
        op->d_fg = M.alloc<cd>((size_t)op->Hh[0] * op->Hh[1] * op->Hh[2]);
        launch_fg_octant(A, op->d_fg, op->L[0], op->L[1], op->L[2], s);
        HIP_CHECK(hipStreamSynchronize(s));
        M.release(A);
        M.release(T);
    }
    if (opts.precision == SCADYN_PREC_MIXED) make_fp32_copies(op);
}

// ---------------------------------------------------------------- workspace / matvec
static size_t slot_bytes(const scadyn_operator *op, int m) {
    const size_t n = 3 * (size_t)op->g.N;
    const size_t grid = (size_t)op->L[0] * op->L[1] * op->g.Nz * 4;
    return slot_vec_bytes(op->opts, n, m) + 2 * grid * sizeof(cd);  // two grids (complex128 size: both precisions)
}

static void make_matvec_space(scadyn_operator *op, int B, MatvecSpace &ms) {
    ms.B = B;
    const int Lx = op->L[0], Ly = op->L[1], Lz = op->L[2];
    const int Nx = op->g.Nx, Ny = op->g.Ny, Nz = op->g.Nz;
    const long long nb = 4LL * B;
    const size_t gsz = (size_t)Lx * Ly * Nz * nb;
    ms.grid0 = ms.mem.alloc<cd>(gsz);
    ms.grid1 = ms.mem.alloc<cd>(gsz);
    ms.in_off = ms.mem.alloc<long long>(B);
    ms.active = ms.mem.alloc<int>(B);
    const size_t budget = 64 * 1024;
    const bool reg = op->opts.fft_kernel == 0;
    for (int pr = 0; pr < 2; ++pr) {
        const size_t el = pr == 0 ? sizeof(cd) : sizeof(cf);
        make_line_pass(ms.pass[pr][0], Lx, Nx, Lx, nb, (long long)Ny * Nz, budget, el, reg);
        make_line_pass(ms.pass[pr][1], Ly, Ny, Ly, (long long)Lx * nb, Nz, budget, el, reg);
        make_line_pass(ms.pass[pr][2], Lz, Nz, Nz, (long long)Ly * Lx * nb, 1, budget, el, reg);
        make_line_pass(ms.pass[pr][3], Ly, Ly, Ny, (long long)Lx * nb, Nz, budget, el, reg);
        make_line_pass(ms.pass[pr][4], Lx, Lx, Nx, nb, (long long)Ny * Nz, budget, el, reg);
    }
}

static void tune_fft_tiles(scadyn_operator *op, MatvecSpace &ms);

static Workspace &workspace(scadyn_operator *op, int B) {
    const int m = op->opts.maxit;
    if (op->ws && op->ws->B == B && op->ws->m == m) return *op->ws;
    op->ws.reset(new Workspace());
    Workspace &w = *op->ws;
    w.B = B;
    w.m = m;
    w.n = 3LL * op->g.N;
    const size_t nB = (size_t)w.n * B;
    make_matvec_space(op, B, w.mv);
    tune_fft_tiles(op, w.mv);
    if (f16_basis(op)) {
        w.Vh = w.mem.alloc<ch>((size_t)(m + 1) * nB);
        w.Wf = w.mem.alloc<cf>(nB);
        w.X = w.mem.alloc<cd>(nB);
        w.Xs = w.mem.alloc<cf>(nB);
    } else if (f32_basis(op)) {
        w.Vf = w.mem.alloc<cf>((size_t)(m + 1) * nB);
        w.Wf = w.mem.alloc<cf>(nB);
        w.X = w.mem.alloc<cd>(nB);
    } else {
        w.V = w.mem.alloc<cd>((size_t)(m + 2) * nB);
        w.X = w.V + (size_t)(m + 1) * nB;
        if (op->opts.precision == SCADYN_PREC_MIXED) w.Xs = w.mem.alloc<cf>(nB);
    }
    if (op->opts.precision == SCADYN_PREC_MIXED) {
        w.zoff = w.mem.alloc<long long>(B);
        HIP_CHECK(hipMemsetAsync(w.zoff, 0, sizeof(long long) * B, op->s));
        w.ract = w.mem.alloc<int>(B);
    }
    w.Rhs = w.mem.alloc<cd>(nB);
    w.W = w.mem.alloc<cd>(nB);
    w.ierr = w.mem.alloc<int>(1);
    HIP_CHECK(hipMemsetAsync(w.ierr, 0, sizeof(int), op->s));
    w.st = w.mem.alloc<SlotState>(B);
    HIP_CHECK(hipMemsetAsync(w.st, 0, sizeof(SlotState) * B, op->s));
    w.H = w.mem.alloc<cd>((size_t)B * (m + 1) * m);
    w.cs = w.mem.alloc<cd>((size_t)B * (m + 1));
    w.sn = w.mem.alloc<cd>((size_t)B * (m + 1));
    w.g = w.mem.alloc<cd>((size_t)B * (m + 1));
    w.y = w.mem.alloc<cd>((size_t)B * m);
    {
        // row-sweep workgroups of the Krylov kernels: up to 2048 (8 waves per
        // SIMD), each sweeping at least 16 row groups
        const long long rw = 256 / std::min(B, 256);
        w.nwg = (int)std::max(1LL, std::min(2048LL, (w.n + rw * 16 - 1) / (rw * 16)));
    }
    w.nrm2 = w.mem.alloc<double>(B);
    w.nav2 = w.mem.alloc<double>(B);
    w.pnrm = w.mem.alloc<double>((size_t)w.nwg * B);
    w.pnav = w.mem.alloc<double>((size_t)w.nwg * B);
    w.pdot = w.mem.alloc<cd>((size_t)w.nwg * std::max(m + 1, bmgs_partials()) * B);
    w.flag = w.mem.alloc<int>(B);
    w.cc = w.mem.alloc<cd>((size_t)B * std::max(m + 1, bmgs_partials() + 2));
    w.hprev = w.mem.alloc<cd>((size_t)B * (m + 1));
    w.col = w.mem.alloc<int>(B);
    w.ret = w.mem.alloc<int>(B);
    return w;
}

static void h2d_sync(void *dst, const void *src, size_t bytes, hipStream_t s);

// the narrow FP64 restart batch of a mixed-precision workspace (allocated once)
static void restart_space(scadyn_operator *op, Workspace &w) {
    if (w.rs) return;
    w.rs_cap = std::min(w.B, 8);
    w.rs.reset(new MatvecSpace());
    make_matvec_space(op, w.rs_cap, *w.rs);
    w.rx = w.mem.alloc<cd>((size_t)w.n * w.rs_cap);
    w.ry = w.mem.alloc<cd>((size_t)w.n * w.rs_cap);
    w.rslot = w.mem.alloc<int>(w.rs_cap);
    std::vector<long long> off(w.rs_cap, 0);
    h2d_sync(w.rs->in_off, off.data(), sizeof(long long) * w.rs_cap, op->s);
}

struct ProfScope {
    scadyn_operator *op;
    Profiler::Pending p;
    bool on;
    ProfScope(scadyn_operator *o, int cls, double bytes) : op(o), on(o->prof.on) {
        if (on) op->prof.begin(cls, bytes, op->s, p);
    }
    ~ProfScope() {
        if (on) op->prof.end(p, op->s);
    }
};

template <typename C>
struct OpData;
template <>
struct OpData<cd> {
    static const cd *S4(scadyn_operator *op) { return op->d_S4; }
    static const cd *fg(scadyn_operator *op) { return op->d_fg; }
    static const cd *tw(scadyn_operator *op) { return op->d_tw; }
    static const int pr = 0;
};
template <>
struct OpData<cf> {
    static const cf *S4(scadyn_operator *op) { return op->d_S4f; }
    static const cf *fg(scadyn_operator *op) { return op->d_fgf; }
    static const cf *tw(scadyn_operator *op) { return op->d_twf; }
    static const int pr = 1;
};

// y = A x for the batch: x element (row, b) at xin[in_off[b] + row*B + b];
// grids, S* and FFT(G) in precision C, x and y complex128
template <typename C>
static void run_matvec(scadyn_operator *op, MatvecSpace &w, const C *xin, const int *active, cd *y,
                       const long long *in_off = nullptr) {
    const HostGeometry &g = op->g;
    const int pr = OpData<C>::pr;
    C *g0 = (C *)w.grid0, *g1 = (C *)w.grid1;
    const C *tw = OpData<C>::tw(op);
    MatvecArgs<C> a;
    a.xin = xin;
    a.in_off = in_off ? in_off : w.in_off;
    a.B = w.B;
    a.N = g.N;
    a.Nx = g.Nx;
    a.Ny = g.Ny;
    a.Nz = g.Nz;
    a.ncx = g.ncx;
    a.ncy = g.ncy;
    a.ncz = g.ncz;
    a.cube_start = op->d_cube_start;
    a.tet_cube = op->d_tet_cube;
    a.S4 = OpData<C>::S4(op);
    a.row_start = op->d_row_start;
    a.ent_col = op->d_ent_col;
    a.ent_blk = op->d_ent_blk;
    a.blocks = op->d_blocks;
    a.ne_cubes = op->d_ne_cubes;
    a.n_ne = op->n_ne_cubes;
    a.k2 = op->k * op->k;
    a.slot_active = active;
    a.gather_kernel = op->opts.gather_kernel;
    a.grid_in = g0;
    a.grid_out = g0;
    a.y = y;
    const double c = sizeof(C);
    const int po = pr * SCADYN_KCLASS_COUNT;  // profiler bucket of this precision
    const double nB = 4.0 * w.B;
    const double Lx = op->L[0], Ly = op->L[1], Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    {
        ProfScope ps(op, SCADYN_KCLASS_SCATTER + po, c * (Nx * Ny * Nz * nB + 27.0 * 4 * g.N) + c * 3 * g.N * w.B);
        launch_scatter<C>(a, op->s);
    }
    {
        ProfScope ps(op, SCADYN_KCLASS_FFTX + po, c * nB * (Nx * Ny * Nz + Lx * Ny * Nz));
        launch_line_fft<C>(w.pass[pr][0], g0, g1, tw, -1, op->s);
    }
    {
        ProfScope ps(op, SCADYN_KCLASS_FFTY + po, c * nB * (Lx * Ny * Nz + Lx * Ly * Nz));
        launch_line_fft<C>(w.pass[pr][1], g1, g0, tw + op->L[0], -1, op->s);
    }
    {
        ConvInfo ci;
        ci.Lx = op->L[0];
        ci.Ly = op->L[1];
        ci.Lz = op->L[2];
        ci.Hx = op->Hh[0];
        ci.Hy = op->Hh[1];
        ci.nb = 4LL * w.B;
        ProfScope ps(op, SCADYN_KCLASS_FFTZCONV + po,
                     c * (nB * 2.0 * Lx * Ly * Nz + (double)op->Hh[0] * op->Hh[1] * op->Hh[2]));
        launch_line_conv<C>(w.pass[pr][2], g0, g0, tw + op->L[0] + op->L[1], OpData<C>::fg(op), ci, op->s);
    }
    {
        ProfScope ps(op, SCADYN_KCLASS_IFFTY + po, c * nB * (Lx * Ly * Nz + Lx * Ny * Nz));
        launch_line_fft<C>(w.pass[pr][3], g0, g1, tw + op->L[0], +1, op->s);
    }
    {
        ProfScope ps(op, SCADYN_KCLASS_IFFTX + po, c * nB * (Lx * Ny * Nz + Nx * Ny * Nz));
        launch_line_fft<C>(w.pass[pr][4], g1, g0, tw, +1, op->s);
    }
    {
        ProfScope ps(op, SCADYN_KCLASS_GATHER + po,
                     c * (Nx * Ny * Nz * nB + 27.0 * 4 * g.N) + (c + 16.0) * 3 * g.N * w.B + 80.0 * op->nblocks);
        launch_gather_near<C>(a, op->s);
    }
}

// Autotuner of the line-FFT tiles.  For every register pass of a large batch
// (and both precisions when the operator is mixed), each candidate tile width
// (4 to 32 inner columns per workgroup) is timed on the batch's own grids and
// the fastest kept; the planner's default wins ties within 3%.  Small grids
// keep the default (tuning would cost more than it saves).
template <typename C>
static void run_pass(scadyn_operator *op, MatvecSpace &ms, int q, const LinePass &p) {
    C *g0 = (C *)ms.grid0, *g1 = (C *)ms.grid1;
    const C *tw = OpData<C>::tw(op);
    const C *twq = q == 0 || q == 4 ? tw : (q == 1 || q == 3 ? tw + op->L[0] : tw + op->L[0] + op->L[1]);
    if (q == 2) {
        ConvInfo ci;
        ci.Lx = op->L[0];
        ci.Ly = op->L[1];
        ci.Lz = op->L[2];
        ci.Hx = op->Hh[0];
        ci.Hy = op->Hh[1];
        ci.nb = 4LL * ms.B;
        launch_line_conv<C>(p, g0, g0, twq, OpData<C>::fg(op), ci, op->s);  // in place, as in run_matvec
    } else {
        launch_line_fft<C>(p, g0, g1, twq, q < 2 ? -1 : +1, op->s);
    }
}

static void tune_fft_tiles(scadyn_operator *op, MatvecSpace &ms) {
    const size_t gsz = (size_t)op->L[0] * op->L[1] * op->g.Nz * 4 * ms.B;
    if (op->opts.fft_kernel != 0 || gsz < ((size_t)1 << 24)) return;
    hipStream_t s = op->s;
    HIP_CHECK(hipMemsetAsync(ms.grid0, 0, sizeof(cd) * gsz, s));
    HIP_CHECK(hipMemsetAsync(ms.grid1, 0, sizeof(cd) * gsz, s));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    // candidates may use the device's whole LDS per workgroup (160 KiB on MI355X;
    // the planner's defaults stay within 64 KiB)
    int lds_dev = 0;
    HIP_CHECK(hipDeviceGetAttribute(&lds_dev, hipDeviceAttributeMaxSharedMemoryPerBlock, op->device));
    const size_t budget = std::max<size_t>(64 * 1024, (size_t)lds_dev);
    const int nprec = op->opts.precision == SCADYN_PREC_MIXED ? 2 : 1;
    for (int pr = 0; pr < nprec; ++pr) {
        if (pr == 1 && !op->d_fgf) break;  // no complex64 operator data yet
        const size_t el = pr == 0 ? sizeof(cd) : sizeof(cf);
        for (int q = 0; q < 5; ++q) {
            LinePass &p = ms.pass[pr][q];
            if (!p.reg) continue;
            LinePass keep = p;
            float best = 0.0f;
            const int cand[] = {p.ti, 4, 8, 16, 32};
            for (int ci = 0; ci < 5; ++ci) {
                LinePass c = p;
                if (ci > 0 && (cand[ci] == p.ti || !set_line_tile(c, cand[ci], budget, el))) continue;
                for (int rep = 0; rep < 3; ++rep) {  // first run warms up
                    if (rep == 1) HIP_CHECK(hipEventRecord(e0, s));
                    if (pr == 0)
                        run_pass<cd>(op, ms, q, c);
                    else
                        run_pass<cf>(op, ms, q, c);
                }
                HIP_CHECK(hipEventRecord(e1, s));
                HIP_CHECK(hipEventSynchronize(e1));
                float t = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
                if (ci == 0 || t < 0.97f * best) {
                    best = t;
                    keep = c;
                }
            }
            p = keep;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// pack/unpack between host order [b][3N] (reference tets) and [row][B] (internal)
__global__ void k_pack(const cd *__restrict__ src, int nrhs, int N, const int *__restrict__ perm, int B,
                       cd *__restrict__ dst) {
    const long long tot = 3LL * N * B;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e % B);
        const long long row = e / B;
        const int i = (int)(row / 3), c = (int)(row % 3);
        dst[e] = b < nrhs ? src[(size_t)b * 3 * N + 3 * (size_t)perm[i] + c] : cmk(0, 0);
    }
}
__global__ void k_unpack(const cd *__restrict__ src, int nrhs, int N, const int *__restrict__ perm, int B,
                         cd *__restrict__ dst) {
    const long long tot = 3LL * N * B;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(e % B);
        if (b >= nrhs) continue;
        const long long row = e / B;
        const int i = (int)(row / 3), c = (int)(row % 3);
        dst[(size_t)b * 3 * N + 3 * (size_t)perm[i] + c] = src[e];
    }
}
static unsigned nblk(long long tot) { return (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16); }

// host -> device copy of a host buffer that is reused right after the call
static void h2d_sync(void *dst, const void *src, size_t bytes, hipStream_t s) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

static int auto_batch(scadyn_operator *op, int ncols) {
    if (op->opts.batch > 0) return std::min(op->opts.batch, std::max(ncols, 1));
    // an existing workspace that holds enough slots is reused (no reallocation)
    if (op->ws && op->ws->m == op->opts.maxit && op->ws->B >= std::min(ncols, 64)) return op->ws->B;
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    size_t per = slot_bytes(op, op->opts.maxit);
    size_t cap = (size_t)(0.75 * (double)free_b) / std::max<size_t>(per, 1);
    if (cap == 0) fail(SCADYN_ERR_MEMORY, "not enough device memory for one RHS slot");
    // Up to 64 slots; above 64 columns, split them into the fewest equal
    // generations of at most 128 slots that fill their last 64-slot chunk to at
    // least 48 (the slot-chunk kernels then idle at most a quarter of a chunk)
    // and fit in memory: no nearly empty final generation of refills.
    int B = std::min(64, ncols);
    if (ncols > 64)
        for (int gen = (ncols + 127) / 128; gen <= ncols / 48; ++gen) {
            const int Bg = (ncols + gen - 1) / gen;
            if ((Bg % 64 == 0 || Bg % 64 >= 48) && (size_t)Bg <= cap) {
                B = Bg;
                break;
            }
        }
    // memory, not the column count, limits the batch
    const bool capped = cap < (size_t)std::min(ncols, 128);
    B = (int)std::min<size_t>((size_t)B, cap);
    // a memory-capped batch keeps the slot-chunk kernels: round down to a width
    // whose last 64-slot chunk holds at least 48 slots
    if (B > 64 && B % 64 != 0 && B % 64 < 48) B -= B % 64;
    // a capped batch whose last chunk is nearly full is completed to 64 slots
    // when that still fits within 0.85 of the free memory: a full chunk runs
    // each slot faster (824k-tet sphere, mixed: 64 slots 1.18 columns/s, 59
    // slots 1.11)
    if (capped && B % 64 >= 48) {
        const int Bup = B + (64 - B % 64);
        if (Bup <= ncols && (size_t)Bup * per <= (size_t)(0.85 * (double)free_b)) B = Bup;
    }
    return std::max(1, B);
}

struct SolveSpec {
    int prec = SCADYN_PREC_FP64;  // matvec precision of the Arnoldi steps
    bool x0zero = false;          // x0 = 0 (correction equations) instead of x0 = rhs
    double tol = 1e-5;
    int restart = 4;
    int ortho = 0;                // 0 MGS (reference), 1 CGS + DGKS, 2 block MGS
    int mgs_block = 1;            // MGS: basis vectors per pass (1, 2, 4, 6 or 8)
    // mixed precision: every restart takes the true residual with the FP64
    // matvec and a system is done only when it is <= tol_true ||b||
    bool verify64 = false;
    double tol_true = 0.0;
    // aligned batch: a slot whose restart cycle ends waits (PH_PARK) until no
    // slot of the batch is in its Arnoldi phase; then every waiting slot
    // restarts and free slots are refilled together.  All slots then share
    // one Arnoldi index, so a step's orthogonalisation passes and the mixed
    // solve's FP64 restart matvec serve the whole batch at once, instead of
    // every step running the longest slot's passes for half-idle lanes and a
    // narrow restart matvec.  Per-slot arithmetic is unchanged.
    bool align = true;
};

// Pinned host buffer (page-locked, so async copies never read a reused buffer).
template <class T>
struct Pinned {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        release();
        HIP_CHECK(hipHostMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
    }
    ~Pinned() { release(); }
};

// Batched GMRES over a list of columns, advanced one matvec per step.
// launch() enqueues a step on the operator's stream (refill of free slots,
// input selection, matvec, MGS passes, Givens/least squares, vector update,
// copy of the slot states); finish() waits for that step and retires finished
// slots.  One host synchronisation per step; several solvers on different
// operators (streams) overlap on the device when their launches interleave.
struct ColumnSolver {
    scadyn_operator *op = nullptr;
    SolveSpec spec;
    const cd *src = nullptr;
    long long ld = 0;
    std::vector<int> cols;
    cd *dst = nullptr;
    int32_t *iters = nullptr;
    Workspace *w = nullptr;
    KrylovDims d;
    int B = 0;
    Pinned<SlotState> hs;
    Pinned<int> h_load, h_ret, h_rs, h_ract, h_err;
    std::vector<int> slot_col;
    std::vector<char> col_fin;  // 1 once column c (index into src/dst) has been retired
    int next = 0, nfail = 0, singular = 0, max_cycles = 0;
    long long steps = 0, max_steps = 0;
    // first fills spread over `spread` solver steps: slot b takes its first
    // column no earlier than step b * spread / B (0: all slots at once)
    long long spread = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;

    ~ColumnSolver() {
        if (ev) (void)hipEventDestroy(ev);
    }

    void init(scadyn_operator *o, const cd *s_, long long ld_, const std::vector<int> &c_, cd *dst_, int32_t *it_,
              const SolveSpec &sp) {
        op = o;
        src = s_;
        ld = ld_;
        cols = c_;
        dst = dst_;
        iters = it_;
        spec = sp;
        B = auto_batch(op, (int)cols.size());
        w = &workspace(op, B);
        // a complex64 or binary16 basis serves only the mixed-precision blocked-MGS solve
        if ((w->Vf || w->Vh) && (spec.prec != SCADYN_PREC_MIXED || spec.ortho != 0 || spec.mgs_block <= 1))
            fail(SCADYN_ERR_ARG, "solve settings do not match the operator's reduced-precision Krylov basis");
        d.n = w->n;
        d.B = B;
        d.m = w->m;
        d.nwg = w->nwg;
        hs.alloc(B);
        h_err.alloc(1);
        h_err.p[0] = 0;
        h_load.alloc(B);
        h_ret.alloc(B);
        if (spec.verify64) {
            restart_space(op, *w);
            h_rs.alloc(((B + w->rs_cap - 1) / w->rs_cap) * w->rs_cap);
            h_ract.alloc(B);
        }
        std::memset(hs.p, 0, sizeof(SlotState) * B);
        slot_col.assign(B, -1);
        col_fin.assign((size_t)std::max<long long>(ld, 1), 0);
        const size_t nB = (size_t)w->n * B;
        HIP_CHECK(hipMemsetAsync(w->st, 0, sizeof(SlotState) * B, op->s));
        HIP_CHECK(hipMemsetAsync(w->ierr, 0, sizeof(int), op->s));  // the workspace may be reused
        if (w->Vh) {
            HIP_CHECK(hipMemsetAsync(w->Vh, 0, sizeof(ch) * (size_t)(w->m + 1) * nB, op->s));
            HIP_CHECK(hipMemsetAsync(w->X, 0, sizeof(cd) * nB, op->s));
        } else if (w->Vf) {
            HIP_CHECK(hipMemsetAsync(w->Vf, 0, sizeof(cf) * (size_t)(w->m + 1) * nB, op->s));
            HIP_CHECK(hipMemsetAsync(w->X, 0, sizeof(cd) * nB, op->s));
        } else {
            HIP_CHECK(hipMemsetAsync(w->V, 0, sizeof(cd) * (size_t)(w->m + 2) * nB, op->s));
        }
        // every system finishes within restart cycles of at most m+1 matvecs; bound the loop
        // (an aligned batch adds at most two waiting steps per cycle and one
        // cycle's wait before a refill)
        max_steps = ((long long)spec.restart * (w->m + 3) + w->m + 8) * (((long long)cols.size() + B - 1) / B + 2) +
                    16 + spread;
        HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }

    bool done() const {
        if (next < (int)cols.size()) return false;
        for (int b = 0; b < B; ++b)
            if (slot_col[b] >= 0) return false;
        return true;
    }

    void launch() {
        if (++steps > max_steps) fail(SCADYN_ERR_DEVICE, "GMRES driver made no progress");
        hipStream_t s = op->s;
        // cycle boundary of an aligned batch: no slot in its Arnoldi phase.
        // Free slots are refilled up to two steps into a cycle (a newcomer
        // lags the batch by at most two steps, which the others wait for at
        // the cycle's end), so a system that finishes at a restart is
        // replaced at once rather than a cycle later.
        int maxi = -1;
        for (int b = 0; b < B; ++b)
            if (slot_col[b] >= 0 && hs.p[b].phase == PH_ARNOLDI) maxi = std::max(maxi, hs.p[b].i);
        const bool boundary = !spec.align || maxi < 0;
        const bool refill_ok = !spec.align || maxi < 2;
        if (spec.align && boundary) {
            bool any_park = false;
            for (int b = 0; b < B; ++b)
                if (slot_col[b] >= 0 && hs.p[b].phase == PH_PARK) {
                    hs.p[b].phase = PH_RESTART;
                    any_park = true;
                }
            if (any_park) launch_unpark(d, w->st, s);
        }
        // refill free slots (at a boundary when aligned)
        bool any_load = false;
        for (int b = 0; b < B; ++b) {
            h_load.p[b] = -1;
            if (refill_ok && slot_col[b] < 0 && next < (int)cols.size() && steps - 1 >= (long long)b * spread / B) {
                h_load.p[b] = cols[next++];
                slot_col[b] = h_load.p[b];
                any_load = true;
                hs.p[b].phase = PH_RESTART;  // host view until the state copy returns
                hs.p[b].i = 0;
            }
        }
        if (any_load) {
            HIP_CHECK(hipMemcpyAsync(w->col, h_load.p, sizeof(int) * B, hipMemcpyHostToDevice, s));
            launch_load_rhs(d, w->col, src, ld, w->Rhs, w->X, w->pnrm, spec.x0zero ? 1 : 0, s);
            launch_load_finish(d, w->col, w->st, w->pnrm, s);
        }
        // MGS passes needed: max over busy slots of i+1 (host view of the previous step)
        int K = -1;
        for (int b = 0; b < B; ++b) {
            if (slot_col[b] < 0) continue;
            if (hs.p[b].phase == PH_RESTART) K = std::max(K, 0);
            if (hs.p[b].phase == PH_ARNOLDI) K = std::max(K, hs.p[b].i + 1);
        }
        if (K >= 0) {
            // restarting slots (verify64): true residual from the FP64 matvec of x,
            // packed into the narrow restart batch; the rest take the batch matvec
            int nrst = 0, narn = 0;
            for (int b = 0; b < B; ++b) {
                if (slot_col[b] < 0) continue;
                if (hs.p[b].phase == PH_RESTART && spec.verify64) h_rs.p[nrst++] = b;
                if (hs.p[b].phase == PH_ARNOLDI || (hs.p[b].phase == PH_RESTART && !spec.verify64)) ++narn;
            }
            launch_set_inputs(d, w->st, w->mv.in_off, w->mv.active, spec.verify64 ? 1 : 0,
                              (w->Vf || w->Vh) ? w->m + 1 : w->m + 2, w->ierr, s);
            if (narn > 0) {
                if (spec.prec == SCADYN_PREC_MIXED && w->Vh) {
                    {
                        ProfScope pk(op, SCADYN_KCLASS_KRYLOV, 0.0);
                        launch_stage_basis(d, w->Vh, w->mv.in_off, w->mv.active, w->Xs, s);
                    }
                    run_matvec<cf>(op, w->mv, w->Xs, w->mv.active, w->W, w->zoff);
                } else if (spec.prec == SCADYN_PREC_MIXED && w->Vf) {
                    run_matvec<cf>(op, w->mv, w->Vf, w->mv.active, w->W);
                } else if (spec.prec == SCADYN_PREC_MIXED) {
                    launch_stage_f32(w->V, w->mv.in_off, w->mv.active, w->n, B, w->Xs, s);
                    run_matvec<cf>(op, w->mv, w->Xs, w->mv.active, w->W, w->zoff);
                } else {
                    run_matvec<cd>(op, w->mv, w->V, w->mv.active, w->W);
                }
            }
            // many restarting slots (a synchronised restart of the batch): one
            // full-width FP64 matvec of x over the batch grids, masked to them,
            // costs less than nrst / rs_cap narrow ones (about half per slot)
            if (nrst > 0 && 2 * nrst >= B && w->ract) {
                for (int b = 0; b < B; ++b) h_ract.p[b] = 0;
                for (int j = 0; j < nrst; ++j) h_ract.p[h_rs.p[j]] = 1;
                HIP_CHECK(hipMemcpyAsync(w->ract, h_ract.p, sizeof(int) * B, hipMemcpyHostToDevice, s));
                run_matvec<cd>(op, w->mv, w->X, w->ract, w->W, w->zoff);
                nrst = 0;
            }
            for (int r0 = 0; r0 < nrst; r0 += w->rs_cap) {
                const int nr = std::min(w->rs_cap, nrst - r0);
                for (int j = nr; j < w->rs_cap; ++j) h_rs.p[r0 + j] = -1;
                HIP_CHECK(hipMemcpyAsync(w->rslot, h_rs.p + r0, sizeof(int) * w->rs_cap, hipMemcpyHostToDevice, s));
                launch_pack_slots(d, w->X, w->rslot, w->rs_cap, w->rx, s);
                run_matvec<cd>(op, *w->rs, w->rx, nullptr, w->ry);
                launch_unpack_slots(d, w->ry, w->rslot, w->rs_cap, w->W, s);
            }
            ProfScope ps(op, SCADYN_KCLASS_KRYLOV, 0.0);
            if (spec.ortho == 1) {
                launch_cgs(d, w->W, w->V, w->Rhs, w->st, w->H, w->pdot, w->pnrm, w->pnav, w->nrm2, w->nav2, w->flag,
                           w->cc, w->hprev, std::max(K - 1, 0), s);
            } else if (spec.ortho == 2) {
                launch_bmgs(d, w->W, w->V, w->Rhs, w->st, w->H, w->pdot, w->pnrm, w->pnav, w->cc, w->nrm2, w->nav2,
                            std::max(K - 1, 0), s);
            } else if (spec.mgs_block > 1 && w->Vh) {
                launch_mgsb(d, spec.mgs_block, K, w->W, w->Wf, (const ch *)w->Vh, w->Rhs, w->st, w->H, w->pdot,
                            w->pnrm, w->pnav, w->nrm2, w->nav2, s);
            } else if (spec.mgs_block > 1 && w->Vf) {
                launch_mgsb(d, spec.mgs_block, K, w->W, w->Wf, (const cf *)w->Vf, w->Rhs, w->st, w->H, w->pdot,
                            w->pnrm, w->pnav, w->nrm2, w->nav2, s);
            } else if (spec.mgs_block > 1) {
                launch_mgsb(d, spec.mgs_block, K, w->W, (const cd *)w->V, w->Rhs, w->st, w->H, w->pdot, w->pnrm,
                            w->pnav, w->nrm2, w->nav2, s);
            } else {
                for (int kk = 0; kk <= K; ++kk) {
                    launch_mgs_pass(d, kk, w->W, w->V, w->Rhs, w->st, w->H, w->pdot, w->pnrm, w->pnav, s);
                    launch_mgs_reduce(d, kk, w->st, w->H, w->pdot, w->pnrm, w->pnav, w->nrm2, w->nav2, s);
                }
            }
            launch_slot_update(d, w->st, w->H, w->cs, w->sn, w->g, w->y, w->nrm2, w->nav2, spec.tol, spec.restart,
                               spec.verify64 ? 1 : 0, spec.tol_true, spec.align ? 1 : 0, s);
            if (w->Vh)
                launch_vec_update(d, w->st, (const cf *)w->Wf, w->Vh, w->X, w->y, s);
            else if (w->Vf)
                launch_vec_update(d, w->st, (const cf *)w->Wf, w->Vf, w->X, w->y, s);
            else
                launch_vec_update(d, w->st, w->W, w->V, w->X, w->y, s);
        }
        HIP_CHECK(hipMemcpyAsync(hs.p, w->st, sizeof(SlotState) * B, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(h_err.p, w->ierr, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(ev, s));
        pending = true;
    }

    void finish() {
        if (!pending) return;
        HIP_CHECK(hipEventSynchronize(ev));
        pending = false;
        if (h_err.p[0]) fail(SCADYN_ERR_DEVICE, "internal: a matvec input offset fell outside the Krylov basis");
        if (op->prof.on) op->prof.collect();
        bool any_done = false;
        for (int b = 0; b < B; ++b) {
            h_ret.p[b] = -1;
            if (slot_col[b] >= 0 && hs.p[b].phase == PH_DONE) {
                h_ret.p[b] = slot_col[b];
                if (iters) iters[slot_col[b]] += hs.p[b].ite;
                op->reorth_fired += hs.p[b].reorth;
                max_cycles = std::max(max_cycles, hs.p[b].cycle + 1);
                if (hs.p[b].conv == 0) nfail++;
                if (hs.p[b].conv < 0) singular++;
                col_fin[slot_col[b]] = 1;
                slot_col[b] = -1;
                hs.p[b].phase = PH_IDLE;
                any_done = true;
            }
        }
        if (any_done) {
            HIP_CHECK(hipMemcpyAsync(w->ret, h_ret.p, sizeof(int) * B, hipMemcpyHostToDevice, op->s));
            launch_store_x(d, w->ret, w->X, dst, ld, op->s);
            launch_retire(d, w->ret, w->st, op->s);
            for (int b = 0; b < B; ++b) ndone += h_ret.p[b] >= 0;
        }
        progress();
    }

    // SCADYN_PROGRESS=<seconds>: a line on stderr at most that often while a
    // long solve runs (systems finished, solver steps)
    void progress() {
        static const double every = [] {
            const char *e = std::getenv("SCADYN_PROGRESS");
            return e ? std::atof(e) : 0.0;
        }();
        if (every <= 0.0) return;
        const auto now = std::chrono::steady_clock::now();
        if (std::chrono::duration<double>(now - last_note).count() < every) return;
        last_note = now;
        std::fprintf(stderr, "[scadyn] %d of %zu systems done after %lld steps (batch %d)\n", ndone, cols.size(),
                     steps, B);
        std::fflush(stderr);
    }
    int ndone = 0;
    std::chrono::steady_clock::time_point last_note = std::chrono::steady_clock::now();
};

// Solver settings of the operator's precision mode.  FP64: the reference's
// GMRES.  Mixed: the Arnoldi steps take the FP32 matvec, every restart takes
// the true residual with the FP64 matvec, and a system is done when that
// residual is <= refine_tol ||b|| (an Arnoldi estimate below tol only ends the
// cycle early): iterative refinement with one restart cycle per correction.
static SolveSpec solve_spec(const scadyn_operator *op) {
    SolveSpec spec;
    spec.tol = op->opts.tol;
    spec.restart = op->opts.restart;
    spec.ortho = op->opts.ortho;
    spec.mgs_block = op->opts.mgs_block == 0 ? MGS_BLOCK_AUTO : op->opts.mgs_block;
    if (op->opts.precision == SCADYN_PREC_MIXED) {
        spec.prec = SCADYN_PREC_MIXED;
        spec.verify64 = true;
        spec.tol_true = op->opts.refine_tol;
        if (f16_basis(op)) spec.tol = std::max(spec.tol, F16_CYCLE_TOL);
    }
    return spec;
}

// Solve A X = Bm for all ld columns.
static void solve_all(scadyn_operator *op, const cd *Bm, long long n, int ld, cd *X, int32_t *iters, int *nfail,
                      int32_t *outer_rounds) {
    (void)n;
    std::vector<int> all(ld);
    for (int c = 0; c < ld; ++c) all[c] = c;
    if (iters)
        for (int c = 0; c < ld; ++c) iters[c] = 0;
    const SolveSpec spec = solve_spec(op);
    ColumnSolver cs;
    cs.init(op, Bm, ld, all, X, iters, spec);
    while (!cs.done()) {
        cs.launch();
        cs.finish();
    }
    HIP_CHECK(hipStreamSynchronize(op->s));
    if (cs.singular) fail(SCADYN_ERR_SINGULAR, "Matrix is numerically singular! (GMRES least squares)");
    if (nfail) *nfail += cs.nfail;
    if (outer_rounds) *outer_rounds = cs.max_cycles;
}

// One T-matrix solve (compute_T_matrix, src/T_matrix.f90:240-277) as a
// resumable job: launch()/finish() advance its column solver by one step, so a
// host loop can interleave the steps of several jobs on different streams.
// All 2nm columns run as one continuously refilled batch (solve_spec gives
// the precision mode).
// src[row*ns + j] = mat[row*ncol + sel[j]]: the right-hand sides of a column subset
__global__ void k_select_cols(const cd *__restrict__ mat, long long n, int ncol, const int *__restrict__ sel, int ns,
                              cd *__restrict__ src) {
    const long long tot = n * ns;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
         e += (long long)gridDim.x * blockDim.x) {
        const long long row = e / ns;
        src[e] = mat[row * ncol + sel[e % ns]];
    }
}

// One T-matrix solve (compute_T_matrix, src/T_matrix.f90:240-277) as a
// resumable job: launch()/finish() advance its column solver by one step, so a
// host loop can interleave the steps of several jobs on different streams.
// All 2nm columns (or a given subset of them) run as one continuously
// refilled batch (solve_spec gives the precision mode).
struct TmatJob {
    scadyn_operator *op = nullptr;
    int nmax = 0, nm = 0, ncol = 0, ns = 0;  // ns: columns solved (all, or the subset)
    long long n = 0;
    DevMem tmp;
    cd *mat = nullptr, *src = nullptr, *X = nullptr;
    std::vector<int32_t> iters;
    int rounds = 0, nfail = 0;
    bool finished = false;
    std::unique_ptr<ColumnSolver> cs;

    void start(scadyn_operator *o, int nmax_, const std::vector<int> *subset = nullptr, int spread = 0) {
        op = o;
        nmax = nmax_;
        if (nmax < 1) fail(SCADYN_ERR_ARG, "nmax must be >= 1");
        nm = (nmax + 1) * (nmax + 1) - 1;
        ncol = 2 * nm;
        n = 3LL * op->g.N;
        hipStream_t s = op->s;
        mat = tmp.alloc<cd>((size_t)n * ncol);
        launch_vswf(op->d_tcoord, op->d_vol, op->g.N, nmax, op->k, mat, ncol, s);
        src = mat;
        ns = ncol;
        if (subset) {
            ns = (int)subset->size();
            if (ns < 1) fail(SCADYN_ERR_ARG, "empty column subset");
            for (int c : *subset)
                if (c < 0 || c >= ncol) fail(SCADYN_ERR_ARG, "column index outside 0 .. 2 nm - 1");
            int *d_sel = tmp.upload(*subset, s);
            src = tmp.alloc<cd>((size_t)n * ns);
            const long long tot = n * ns;
            hipLaunchKernelGGL(k_select_cols, dim3((unsigned)std::min<long long>((tot + 255) / 256, 1 << 16)),
                               dim3(256), 0, s, mat, n, ncol, d_sel, ns, src);
            HIP_CHECK(hipGetLastError());
        }
        X = tmp.alloc<cd>((size_t)n * ns);
        HIP_CHECK(hipMemsetAsync(X, 0, sizeof(cd) * (size_t)n * ns, s));  // columns not yet solved read as zero
        iters.assign(ns, 0);
        std::vector<int> all(ns);
        for (int c = 0; c < ns; ++c) all[c] = c;
        cs.reset(new ColumnSolver());
        cs->spread = std::max(spread, 0);
        cs->init(op, src, ns, all, X, iters.data(), solve_spec(op));
    }

    void launch() {
        if (!finished) cs->launch();
    }

    void finish() {
        if (finished) return;
        cs->finish();
        if (!cs->done()) return;
        HIP_CHECK(hipStreamSynchronize(op->s));
        if (cs->singular) fail(SCADYN_ERR_SINGULAR, "Matrix is numerically singular! (GMRES least squares)");
        nfail += cs->nfail;
        rounds = cs->max_cycles;
        finished = true;
    }

    // i k^3 mat^H X: ncol x ns, row-major (i, j), on the host
    std::vector<cd> scaled_T() {
        hipStream_t s = op->s;
        const size_t work = std::max<size_t>((size_t)ncol * ns * 64, 1 << 20);
        DevMem t2;
        cd *wk = t2.alloc<cd>(work);
        cd *T = t2.alloc<cd>((size_t)ncol * ns);
        {
            ProfScope ps(op, SCADYN_KCLASS_TGEMM, 8.0 * (double)ncol * ns * (double)n);
            launch_tgemm(mat, X, n, ncol, ns, T, wk, work, s);
        }
        std::vector<cd> hT((size_t)ncol * ns);
        HIP_CHECK(hipMemcpyAsync(hT.data(), T, sizeof(cd) * hT.size(), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        // sc = i k^3 (src/T_matrix.f90:270-275): (x + i y)(i k3) = -y k3 + i x k3
        const double k3 = std::pow(op->k, 3.0);
        for (auto &v : hT) v = cmk(-v.y * k3, v.x * k3);
        op->last_outer_rounds = rounds;
        return hT;
    }

    // the four nm x nm column-major blocks (all columns solved)
    void result(double *Taa, double *Tab, double *Tba, double *Tbb, int32_t *it_out) {
        const std::vector<cd> hT = scaled_T();
        double *outs[4] = {Taa, Tab, Tba, Tbb};
        for (int q = 0; q < 4; ++q) {
            const int r0 = (q >= 2) ? nm : 0, c0 = (q % 2) ? nm : 0;
            for (int j = 0; j < nm; ++j)
                for (int i = 0; i < nm; ++i) {
                    const cd v = hT[(size_t)(r0 + i) * ncol + (c0 + j)];
                    outs[q][2 * ((size_t)j * nm + i)] = v.x;
                    outs[q][2 * ((size_t)j * nm + i) + 1] = v.y;
                }
        }
        if (it_out) std::copy(iters.begin(), iters.end(), it_out);
    }

    // the solved columns: Tcols (ncol x ns) column-major; a column whose system
    // has not finished yet reads as zero with iteration count -1
    void result_cols(double *Tcols, int32_t *it_out) {
        const std::vector<cd> hT = scaled_T();
        for (int j = 0; j < ns; ++j) {
            const bool fin = cs->col_fin[j] != 0;
            for (int i = 0; i < ncol; ++i) {
                Tcols[2 * ((size_t)j * ncol + i)] = fin ? hT[(size_t)i * ns + j].x : 0.0;
                Tcols[2 * ((size_t)j * ncol + i) + 1] = fin ? hT[(size_t)i * ns + j].y : 0.0;
            }
            if (it_out) it_out[j] = fin ? iters[j] : -1;
        }
    }

    int columns_done() const { return cs ? cs->ndone : 0; }
};

static void tmatrix(scadyn_operator *op, int nmax, double *Taa, double *Tab, double *Tba, double *Tbb,
                    int32_t *iters) {
    TmatJob job;
    job.start(op, nmax);
    while (!job.finished) {
        job.launch();
        job.finish();
    }
    job.result(Taa, Tab, Tba, Tbb, iters);
    if (job.nfail && op->opts.strict)
        fail(SCADYN_ERR_NOCONV, std::to_string(job.nfail) + " T-matrix columns did not reach tol");
}

// ---------------------------------------------------------------- footprint model
// Device bytes one T-matrix request needs at its peak: operator + the larger of
// (setup temporaries) and (solve workspace + right-hand sides/solutions).
struct Footprint {
    size_t op = 0, setup = 0, solve = 0;
    size_t peak() const { return op + std::max(setup, solve); }
};

static Footprint footprint_of(const HostGeometry &g, long long P, const scadyn_opts &o, int nmax, int B) {
    Footprint f;
    const size_t N = g.N, n = 3 * N;
    const int Nax[3] = {g.Nx, g.Ny, g.Nz};
    size_t L[3], H[3];
    for (int c = 0; c < 3; ++c) {
        L[c] = o.fft_pad == 1 ? 2 * Nax[c] : smooth_length(2 * Nax[c] - 1);
        H[c] = L[c] / 2 + 1;
    }
    const size_t cdb = sizeof(cd);
    f.op = N * (96 + 8 + 16 + 4 + 4) + 4 * (g.N_cubes + 1) + 1728 * N + 72 * (size_t)P + 8 * (N + 1) +
           8 * 2 * (size_t)P + cdb * H[0] * H[1] * H[2] + cdb * (L[0] + L[1] + L[2]);
    if (o.precision == SCADYN_PREC_MIXED) f.op += 864 * N + 8 * H[0] * H[1] * H[2];
    f.setup = 2 * cdb * L[0] * L[1] * L[2] + 9 * (size_t)P + cdb * 27 * 27 * 27 + cdb * 98 * 27;
    // solve workspace (workspace(), restart_space()): per slot the Krylov
    // vectors (slot_vec_bytes: the basis storage of the solve) and two
    // complex128-sized grids; the reduction partials of up to 2048 row-sweep
    // workgroups; the mixed solve's narrow FP64 restart batch
    const int m = o.maxit;
    const bool mixed = o.precision == SCADYN_PREC_MIXED;
    const size_t grid = L[0] * L[1] * (size_t)g.Nz * 4;
    const size_t vec = slot_vec_bytes(o, n, m);
    size_t ws = (size_t)B * (vec + 2 * grid * cdb) + (size_t)B * 1024 * (m + 8) +
                (size_t)2048 * (size_t)std::max(m + 1, 36) * B * cdb;
    if (mixed) {
        const size_t rc = (size_t)std::min(B, 8);
        ws += rc * (2 * grid * cdb + 2 * n * cdb);
    }
    // right-hand sides and solutions (mat, X of the T-matrix job) and the GEMM partials
    const size_t ncol = 2 * ((size_t)(nmax + 1) * (nmax + 1) - 1);
    const size_t rhs = 2 * cdb * n * ncol + cdb * ncol * ncol * 65;
    f.solve = ws + rhs;
    return f;
}

// ---------------------------------------------------------------- request loop
enum ReqState : int { REQ_QUEUED = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FAILED = 3 };

struct Request {
    int64_t id = 0;
    std::vector<double> coord, param;
    std::vector<int32_t> etopol;
    int32_t n_node = 0, n_tet = 0;
    scadyn_opts opts;
    double k = 0;
    int nmax = 0;
    size_t fp = 0;
    // state is written only under the server mutex: QUEUED -> RUNNING at
    // admission, DONE / FAILED when the worker retires the request (the same
    // critical section that removes it from the active list and returns its
    // footprint).  A waiter that sees DONE / FAILED may free the request, so
    // the worker never touches it after publishing.
    int state = REQ_QUEUED;
    // the worker's own record of the outcome (REQ_DONE / REQ_FAILED, 0 while
    // solving); read and written by the worker thread only
    int outcome = 0;
    bool set_up = false;
    std::atomic<bool> cancel{false};  // set by scadyn_server_cancel; the worker ends the request
    int stub_steps = 0;  // host-only stand-in solve: steps left
    int err = 0;
    std::string msg;
    double t[4] = {0, 0, 0, 0};  // submit, admit, solve start, done (s since server start); under the mutex
    std::unique_ptr<scadyn_operator> op;
    std::unique_ptr<TmatJob> job;
    std::vector<double> T;  // 4 blocks, column-major complex
    std::vector<int32_t> iters;
};

}  // namespace scadyn

struct scadyn_server {
    int device = 0, max_active = 8;
    bool host_only = false;  // admission and retirement with a host stand-in solve (no device)
    size_t budget = 0, in_use = 0;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<scadyn::Request *> queue;
    std::vector<scadyn::Request *> active;
    std::map<int64_t, std::unique_ptr<scadyn::Request>> all;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
};

namespace scadyn {

// worker thread: record a failure; it is published at retirement
static void fail_request(scadyn_server *, Request *r, int code, const std::string &msg) {
    r->err = code;
    r->msg = msg;
    r->outcome = REQ_FAILED;
    r->job.reset();
    r->op.reset();
}

static void stamp(scadyn_server *sv, Request *r, int i) {
    std::lock_guard<std::mutex> lk(sv->mu);
    r->t[i] = sv->now();
}

}  // namespace scadyn

void scadyn_server::run() {
    using namespace scadyn;
    if (!host_only) (void)hipSetDevice(device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            // admission by device-memory footprint, FIFO
            while (!queue.empty()) {
                Request *r = queue.front();
                if ((int)active.size() >= max_active || in_use + r->fp > budget) break;
                queue.pop_front();
                r->state = REQ_RUNNING;
                r->t[1] = now();
                in_use += r->fp;
                active.push_back(r);
            }
            if (active.empty()) {
                if (stop && queue.empty()) return;
                cv.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
        }
        // The active list changes only in this thread (admission above and
        // retirement below), so it is read here without the lock; a request
        // in it is not freed before this thread retires it.
        for (Request *r : active)
            if (!r->outcome && r->cancel.load()) fail_request(this, r, SCADYN_ERR_ARG, "request cancelled");
        // set up at most one admitted request per pass (host geometry + device setup)
        for (Request *r : active) {
            if (r->set_up || r->outcome) continue;
            try {
                if (host_only) {
                    r->stub_steps = 20;
                } else {
                    scadyn_mesh m{r->n_node, r->n_tet, r->coord.data(), r->etopol.data(), r->param.data()};
                    r->op.reset(new scadyn_operator());
                    create_operator(r->op.get(), m, r->opts, r->k);
                    r->job.reset(new TmatJob());
                    r->job->start(r->op.get(), r->nmax);
                }
                r->set_up = true;
                stamp(this, r, 2);
            } catch (const Error &e) {
                fail_request(this, r, e.code, e.msg);
            }
            break;
        }
        // one step of every solving request: launch all, then complete all
        for (Request *r : active)
            if (r->set_up && !r->outcome && !host_only) {
                try {
                    r->job->launch();
                } catch (const Error &e) {
                    fail_request(this, r, e.code, e.msg);
                }
            }
        for (Request *r : active)
            if (r->set_up && !r->outcome) {
                const int nm = (r->nmax + 1) * (r->nmax + 1) - 1;
                if (host_only) {
                    // stand-in solve for the loop's own tests: no T-matrix is
                    // computed; the iteration counts carry the request id
                    std::this_thread::sleep_for(std::chrono::milliseconds(1));
                    if (--r->stub_steps > 0) continue;
                    r->T.assign((size_t)8 * nm * nm, 0.0);
                    r->iters.assign(2 * nm, (int32_t)r->id);
                    r->outcome = REQ_DONE;
                    continue;
                }
                try {
                    r->job->finish();
                    if (r->job->finished) {
                        r->T.assign((size_t)8 * nm * nm, 0.0);
                        r->iters.assign(2 * nm, 0);
                        const size_t blk = (size_t)2 * nm * nm;
                        r->job->result(r->T.data(), r->T.data() + blk, r->T.data() + 2 * blk,
                                       r->T.data() + 3 * blk, r->iters.data());
                        const int nf = r->job->nfail;
                        r->job.reset();
                        r->op.reset();
                        if (nf && r->opts.strict) {
                            r->err = SCADYN_ERR_NOCONV;
                            r->msg = std::to_string(nf) + " T-matrix columns did not reach tol";
                            r->outcome = REQ_FAILED;
                        } else {
                            r->outcome = REQ_DONE;
                        }
                    }
                } catch (const Error &e) {
                    fail_request(this, r, e.code, e.msg);
                }
            }
        {
            // retire: publish the outcome, return the footprint and drop the
            // request from the active list in one critical section
            std::lock_guard<std::mutex> lk(mu);
            std::vector<Request *> keep;
            for (Request *r : active) {
                if (r->outcome) {
                    in_use -= r->fp;
                    r->t[3] = now();
                    r->state = r->outcome;  // r may be freed by a waiter from here on
                } else {
                    keep.push_back(r);
                }
            }
            if (keep.size() != active.size()) done_cv.notify_all();
            active.swap(keep);
        }
    }
}

namespace scadyn {
}  // namespace scadyn

// ================================================================ no-allocation projection
// Row-chunk partials of the projection GEMM for scadyn_op_project_noalloc: a
// file-scope device array (each device holds its own copy for the life of the
// process), so that entry point calls no allocator.  Host code must not pass
// the array's name to a kernel: it is a host-side stand-in there.  The device
// address is taken once per device with that device current and kept in
// g_scratch_ptr.
namespace scadyn {
constexpr size_t NOALLOC_SCRATCH_ELEMS = (size_t)4 << 20;  // complex128 elements: 64 MiB
static __device__ __align__(256) double scratch[2 * NOALLOC_SCRATCH_ELEMS];
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static cd *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {nullptr};

// the device address of `scratch` on `device` (current), looked up on first use
static cd *scratch_ptr(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) fail(SCADYN_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(SCADYN_ERR_DEVICE, std::string("hipGetSymbolAddress of the projection scratch failed: ") +
                                        hipGetErrorString(e));
        g_scratch_ptr[device] = (cd *)p;
    }
    return g_scratch_ptr[device];
}
}  // namespace scadyn

// ================================================================ C ABI
#define ABI_TRY try {
#define ABI_CATCH                                                   \
    }                                                               \
    catch (const scadyn::Error &e) {                                \
        scadyn::set_error(e.code, e.msg);                           \
        return e.code;                                              \
    }                                                               \
    catch (const std::exception &e) {                               \
        scadyn::set_error(SCADYN_ERR_DEVICE, e.what());             \
        return SCADYN_ERR_DEVICE;                                   \
    }                                                               \
    return SCADYN_OK;

extern "C" {

void scadyn_opts_default(scadyn_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->tol = 1e-5;
    o->maxit = 50;
    o->restart = 4;
    o->M_ex = 2;
    o->near_zone = 1;
    o->grid_size = 1.8;
    o->order = 0;
    o->precision = SCADYN_PREC_FP64;
    o->batch = 0;
    o->strict = 0;
    o->device = 0;
    o->fft_pad = 0;
    o->refine_tol = 1e-9;
}

const char *scadyn_last_error(void) { return g_err.c_str(); }

// Solver settings an operator, a request or a footprint query accepts
// (checked on the host, before anything reaches the device).
static void check_opts(const scadyn_opts &o) {
    if (o.precision != SCADYN_PREC_FP64 && o.precision != SCADYN_PREC_MIXED)
        fail(SCADYN_ERR_ARG, "unknown precision mode");
    if (o.maxit < 1 || o.restart < 1) fail(SCADYN_ERR_ARG, "maxit and restart must be >= 1");
    if (o.ortho < 0 || o.ortho > 2)
        fail(SCADYN_ERR_ARG, "ortho must be 0 (MGS), 1 (CGS + DGKS) or 2 (block MGS)");
    // The first device run of the block-MGS path ended in a GPU memory fault whose
    // cause is not found (its code passes the host self-test under
    // AddressSanitizer); it stays disabled on the device until one is.
    if (o.ortho == 2)
        fail(SCADYN_ERR_UNSUPPORTED, "block MGS (ortho = 2) is disabled on the device: its first run faulted and "
                                     "the cause is not yet found");
    if (o.fft_kernel != 0 && o.fft_kernel != 1) fail(SCADYN_ERR_ARG, "fft_kernel must be 0 (automatic) or 1 (Stockham)");
    if (o.gather_kernel != 0 && o.gather_kernel != 1)
        fail(SCADYN_ERR_ARG, "gather_kernel must be 0 (vector FMA) or 1 (matrix cores)");
    if (o.basis != SCADYN_BASIS_AUTO && o.basis != SCADYN_BASIS_C64 && o.basis != SCADYN_BASIS_F16)
        fail(SCADYN_ERR_ARG, "basis must be SCADYN_BASIS_AUTO, SCADYN_BASIS_C64 or SCADYN_BASIS_F16");
    if (o.mgs_block != 0 && o.mgs_block != 1 && o.mgs_block != 2 && o.mgs_block != 4 && o.mgs_block != 6 &&
        o.mgs_block != 8)
        fail(SCADYN_ERR_ARG, "mgs_block must be 0 (automatic), 1, 2, 4, 6 or 8");
}

int scadyn_op_create(const scadyn_mesh *mesh, const scadyn_opts *opts, double k, scadyn_operator **out) {
    ABI_TRY
    if (!mesh || !out) fail(SCADYN_ERR_ARG, "null argument");
    scadyn_opts o;
    if (opts)
        o = *opts;
    else
        scadyn_opts_default(&o);
    check_opts(o);
    *out = nullptr;
    std::unique_ptr<scadyn_operator> op(new scadyn_operator());
    create_operator(op.get(), *mesh, o, k);
    *out = op.release();
    ABI_CATCH
}

void scadyn_op_destroy(scadyn_operator *op) {
    if (!op) return;
    (void)hipSetDevice(op->device);
    delete op;
}

int scadyn_op_info(const scadyn_operator *op, int64_t info[16], double dinfo[4]) {
    ABI_TRY
    if (!op) fail(SCADYN_ERR_ARG, "null operator");
    if (info) {
        for (int i = 0; i < 16; ++i) info[i] = 0;
        info[0] = op->g.N;
        info[1] = op->g.Nx;
        info[2] = op->g.Ny;
        info[3] = op->g.Nz;
        info[4] = op->g.N_cubes;
        info[5] = op->g.N_tet_cube;
        info[6] = op->nblocks;
        info[7] = op->L[0];
        info[8] = op->L[1];
        info[9] = op->L[2];
        info[10] = op->ws ? op->ws->B : 0;
        info[11] = (int64_t)(op->mem.bytes + (op->ws ? op->ws->mem.bytes : 0));
        info[12] = op->last_outer_rounds;
        info[13] = op->np.cube_rows ? 1 : 0;
        if (op->ws)  // tile widths of the five line-FFT passes, 8 bits each (FP64, FP32)
            for (int q = 0; q < 5; ++q) {
                info[14] |= (int64_t)(op->ws->mv.pass[0][q].ti & 0xff) << (8 * q);
                info[15] |= (int64_t)(op->ws->mv.pass[1][q].ti & 0xff) << (8 * q);
            }
    }
    if (dinfo) {
        dinfo[0] = op->g.delta;
        dinfo[1] = op->g.box_delta;
        dinfo[2] = op->k;
        dinfo[3] = (double)op->reorth_fired;
    }
    ABI_CATCH
}

int scadyn_op_export(const scadyn_operator *opc, int32_t what, void *host, int64_t nbytes) {
    ABI_TRY
    scadyn_operator *op = const_cast<scadyn_operator *>(opc);
    if (!op || !host) fail(SCADYN_ERR_ARG, "null argument");
    HIP_CHECK(hipSetDevice(op->device));
    const int N = op->g.N;
    if (what == SCADYN_EXPORT_S) {
        if (nbytes != (int64_t)sizeof(cd) * 4 * N * 27) fail(SCADYN_ERR_ARG, "size mismatch");
        std::vector<cd> s4((size_t)N * 108);
        HIP_CHECK(hipMemcpy(s4.data(), op->d_S4, sizeof(cd) * s4.size(), hipMemcpyDeviceToHost));
        cd *o = (cd *)host;
        for (int ch = 0; ch < 4; ++ch)
            for (int t = 0; t < N; ++t)
                for (int j = 0; j < 27; ++j)
                    o[((size_t)ch * N + t) * 27 + j] = s4[((size_t)op->g.inv[t] * 27 + j) * 4 + ch];
    } else if (what == SCADYN_EXPORT_BLOCKS) {
        if (nbytes != (int64_t)sizeof(cf) * 9 * op->nblocks) fail(SCADYN_ERR_ARG, "size mismatch");
        HIP_CHECK(hipMemcpy(host, op->d_blocks, nbytes, hipMemcpyDeviceToHost));
    } else if (what == SCADYN_EXPORT_BLKIND) {
        if (nbytes != (int64_t)sizeof(int32_t) * 2 * op->nblocks) fail(SCADYN_ERR_ARG, "size mismatch");
        int32_t *o = (int32_t *)host;
        for (long long p = 0; p < op->nblocks; ++p) {
            o[2 * p] = op->g.perm[op->np.T[p]] + 1;
            o[2 * p + 1] = op->g.perm[op->np.B[p]] + 1;
        }
    } else if (what == SCADYN_EXPORT_FG) {
        size_t n = (size_t)op->Hh[0] * op->Hh[1] * op->Hh[2];
        if (nbytes != (int64_t)(sizeof(cd) * n)) fail(SCADYN_ERR_ARG, "size mismatch");
        HIP_CHECK(hipMemcpy(host, op->d_fg, nbytes, hipMemcpyDeviceToHost));
    } else {
        fail(SCADYN_ERR_ARG, "unknown export id");
    }
    ABI_CATCH
}

int scadyn_op_matvec_prec(scadyn_operator *op, int32_t precision, int32_t nrhs, const double *x, double *y) {
    ABI_TRY
    if (!op || !x || !y || nrhs < 1) fail(SCADYN_ERR_ARG, "bad argument");
    if (precision != SCADYN_PREC_FP64 && precision != SCADYN_PREC_MIXED) fail(SCADYN_ERR_ARG, "bad precision");
    HIP_CHECK(hipSetDevice(op->device));
    if (precision == SCADYN_PREC_MIXED) make_fp32_copies(op);
    const int N = op->g.N;
    Workspace &w = workspace(op, nrhs);
    hipStream_t s = op->s;
    DevMem tmp;
    const size_t nv = (size_t)nrhs * 3 * N;
    cd *hx = tmp.alloc<cd>(nv);
    cd *xin = tmp.alloc<cd>(nv);
    HIP_CHECK(hipMemcpyAsync(hx, x, sizeof(cd) * nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pack, dim3(nblk(3LL * N * nrhs)), dim3(256), 0, s, hx, nrhs, N, op->d_perm, nrhs, xin);
    std::vector<long long> off(nrhs, 0);
    h2d_sync(w.mv.in_off, off.data(), sizeof(long long) * nrhs, s);
    if (precision == SCADYN_PREC_MIXED) {
        cf *xf = tmp.alloc<cf>(nv);  // the complex64 matvec takes complex64 input
        hipLaunchKernelGGL(k_to_f32, dim3(1024), dim3(256), 0, s, xin, xf, nv);
        HIP_CHECK(hipGetLastError());
        run_matvec<cf>(op, w.mv, xf, nullptr, w.W);
    } else {
        run_matvec<cd>(op, w.mv, xin, nullptr, w.W);
    }
    hipLaunchKernelGGL(k_unpack, dim3(nblk(3LL * N * nrhs)), dim3(256), 0, s, w.W, nrhs, N, op->d_perm, nrhs, hx);
    HIP_CHECK(hipMemcpyAsync(y, hx, sizeof(cd) * nv, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (op->prof.on) op->prof.collect();
    ABI_CATCH
}

int scadyn_op_matvec(scadyn_operator *op, int32_t nrhs, const double *x, double *y) {
    return scadyn_op_matvec_prec(op, SCADYN_PREC_FP64, nrhs, x, y);
}

int scadyn_op_solve(scadyn_operator *op, int32_t nrhs, const double *rhs, double *x, int32_t *iters) {
    ABI_TRY
    if (!op || !rhs || !x || nrhs < 1) fail(SCADYN_ERR_ARG, "bad argument");
    HIP_CHECK(hipSetDevice(op->device));
    const int N = op->g.N;
    hipStream_t s = op->s;
    DevMem tmp;
    const size_t nv = (size_t)nrhs * 3 * N;
    cd *hx = tmp.alloc<cd>(nv);
    cd *src = tmp.alloc<cd>(nv);
    cd *dst = tmp.alloc<cd>(nv);
    HIP_CHECK(hipMemcpyAsync(hx, rhs, sizeof(cd) * nv, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pack, dim3(nblk(3LL * N * nrhs)), dim3(256), 0, s, hx, nrhs, N, op->d_perm, nrhs, src);
    int nfail = 0;
    int32_t rounds = 0;
    solve_all(op, src, 3LL * N, nrhs, dst, iters, &nfail, &rounds);
    op->last_outer_rounds = rounds;
    hipLaunchKernelGGL(k_unpack, dim3(nblk(3LL * N * nrhs)), dim3(256), 0, s, dst, nrhs, N, op->d_perm, nrhs, hx);
    HIP_CHECK(hipMemcpyAsync(x, hx, sizeof(cd) * nv, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (nfail && op->opts.strict) fail(SCADYN_ERR_NOCONV, std::to_string(nfail) + " systems did not reach tol");
    ABI_CATCH
}

int scadyn_op_vswf(scadyn_operator *op, int32_t nmax, double *mat) {
    ABI_TRY
    if (!op || !mat || nmax < 1) fail(SCADYN_ERR_ARG, "bad argument");
    HIP_CHECK(hipSetDevice(op->device));
    const int N = op->g.N;
    const int nm = (nmax + 1) * (nmax + 1) - 1, ncol = 2 * nm;
    DevMem tmp;
    cd *d = tmp.alloc<cd>((size_t)3 * N * ncol);
    launch_vswf(op->d_tcoord, op->d_vol, N, nmax, op->k, d, ncol, op->s);
    std::vector<cd> h((size_t)3 * N * ncol);
    HIP_CHECK(hipMemcpyAsync(h.data(), d, sizeof(cd) * h.size(), hipMemcpyDeviceToHost, op->s));
    HIP_CHECK(hipStreamSynchronize(op->s));
    cd *o = (cd *)mat;
    const size_t n = 3 * (size_t)N;
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < ncol; ++j)
                o[(size_t)j * n + 3 * (size_t)op->g.perm[i] + c] = h[(size_t)(3 * i + c) * ncol + j];
    ABI_CATCH
}

int scadyn_op_project(scadyn_operator *op, int32_t nmax, const double *X, double *T) {
    ABI_TRY
    if (!op || !X || !T || nmax < 1) fail(SCADYN_ERR_ARG, "bad argument");
    HIP_CHECK(hipSetDevice(op->device));
    const int N = op->g.N;
    const int nm = (nmax + 1) * (nmax + 1) - 1, ncol = 2 * nm;
    const size_t n = 3 * (size_t)N;
    // X column-major (3N, 2nm) in reference tet order -> internal rows, row-major n x ncol
    const cd *xi = (const cd *)X;
    std::vector<cd> hx(n * ncol);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < ncol; ++j)
                hx[(size_t)(3 * i + c) * ncol + j] = xi[(size_t)j * n + 3 * (size_t)op->g.perm[i] + c];
    DevMem tmp;
    cd *mat = tmp.alloc<cd>(n * ncol);
    launch_vswf(op->d_tcoord, op->d_vol, N, nmax, op->k, mat, ncol, op->s);
    cd *dx = tmp.alloc<cd>(n * ncol);
    h2d_sync(dx, hx.data(), sizeof(cd) * hx.size(), op->s);
    const size_t work = std::max<size_t>((size_t)ncol * ncol * 64, 1 << 20);
    cd *wk = tmp.alloc<cd>(work);
    cd *dT = tmp.alloc<cd>((size_t)ncol * ncol);
    {
        ProfScope ps(op, SCADYN_KCLASS_TGEMM, 8.0 * (double)ncol * ncol * (double)n);
        launch_tgemm(mat, dx, (long long)n, ncol, ncol, dT, wk, work, op->s);
    }
    std::vector<cd> hT((size_t)ncol * ncol);
    HIP_CHECK(hipMemcpyAsync(hT.data(), dT, sizeof(cd) * hT.size(), hipMemcpyDeviceToHost, op->s));
    HIP_CHECK(hipStreamSynchronize(op->s));
    cd *o = (cd *)T;
    for (int i = 0; i < ncol; ++i)
        for (int j = 0; j < ncol; ++j) o[(size_t)j * ncol + i] = hT[(size_t)i * ncol + j];
    ABI_CATCH
}

int scadyn_op_project_noalloc(scadyn_operator *op, int64_t n, int32_t na, int32_t nx, const double *mat,
                              const double *X, double *T) {
    ABI_TRY
    if (!op || !mat || !X || !T) fail(SCADYN_ERR_ARG, "null argument");
    if (n < 1 || na < 1 || nx < 1) fail(SCADYN_ERR_ARG, "n, na and nx must be >= 1");
    if ((size_t)na * (size_t)nx > NOALLOC_SCRATCH_ELEMS)
        fail(SCADYN_ERR_ARG, "na * nx = " + std::to_string((size_t)na * (size_t)nx) +
                                 " exceeds the no-allocation limit of " + std::to_string(NOALLOC_SCRATCH_ELEMS) +
                                 " elements");
    HIP_CHECK(hipSetDevice(op->device));
    cd *work = scratch_ptr(op->device);
    // every partial the reduction reads is written by the GEMM first: nothing to clear
    launch_tgemm((const cd *)mat, (const cd *)X, (long long)n, na, nx, (cd *)T, work, NOALLOC_SCRATCH_ELEMS, op->s);
    HIP_CHECK(hipStreamSynchronize(op->s));
    ABI_CATCH
}

int scadyn_op_tmatrix(scadyn_operator *op, int32_t nmax, double *Taa, double *Tab, double *Tba, double *Tbb,
                      int32_t *iters) {
    ABI_TRY
    if (!op || !Taa || !Tab || !Tba || !Tbb) fail(SCADYN_ERR_ARG, "null argument");
    HIP_CHECK(hipSetDevice(op->device));
    tmatrix(op, nmax, Taa, Tab, Tba, Tbb, iters);
    ABI_CATCH
}

int scadyn_op_tmatrix_cols(scadyn_operator *op, int32_t nmax, int32_t ncols, const int32_t *cols, double *Tcols,
                           int32_t *iters) {
    ABI_TRY
    if (!op || !cols || !Tcols || ncols < 1) fail(SCADYN_ERR_ARG, "bad argument");
    HIP_CHECK(hipSetDevice(op->device));
    std::vector<int> sel(cols, cols + ncols);
    TmatJob job;
    job.start(op, nmax, &sel);
    while (!job.finished) {
        job.launch();
        job.finish();
    }
    job.result_cols(Tcols, iters);
    if (job.nfail && op->opts.strict)
        fail(SCADYN_ERR_NOCONV, std::to_string(job.nfail) + " T-matrix columns did not reach tol");
    ABI_CATCH
}

}  // extern "C"

struct scadyn_tjob {
    scadyn_operator *op = nullptr;
    scadyn::TmatJob job;
};

extern "C" {

int scadyn_tjob_create(scadyn_operator *op, int32_t nmax, int32_t ncols, const int32_t *cols, int32_t spread,
                       scadyn_tjob **out) {
    ABI_TRY
    if (!op || !cols || !out || ncols < 1) fail(SCADYN_ERR_ARG, "bad argument");
    *out = nullptr;
    HIP_CHECK(hipSetDevice(op->device));
    std::unique_ptr<scadyn_tjob> j(new scadyn_tjob());
    j->op = op;
    std::vector<int> sel(cols, cols + ncols);
    j->job.start(op, nmax, &sel, spread);
    *out = j.release();
    ABI_CATCH
}

int scadyn_tjob_advance(scadyn_tjob *j, int32_t want, int32_t *ndone) {
    ABI_TRY
    if (!j) fail(SCADYN_ERR_ARG, "null job");
    HIP_CHECK(hipSetDevice(j->op->device));
    const int target = j->job.columns_done() + std::max(want, 0);
    while (!j->job.finished && j->job.columns_done() < target) {
        j->job.launch();
        j->job.finish();
    }
    HIP_CHECK(hipStreamSynchronize(j->op->s));
    if (ndone) *ndone = j->job.columns_done();
    ABI_CATCH
}

int scadyn_tjob_result(scadyn_tjob *j, double *Tcols, int32_t *iters) {
    ABI_TRY
    if (!j || !Tcols) fail(SCADYN_ERR_ARG, "bad argument");
    HIP_CHECK(hipSetDevice(j->op->device));
    j->job.result_cols(Tcols, iters);
    ABI_CATCH
}

void scadyn_tjob_destroy(scadyn_tjob *j) {
    if (!j) return;
    (void)hipSetDevice(j->op->device);
    (void)hipStreamSynchronize(j->op->s);
    delete j;
}

int scadyn_tmat_solve(const scadyn_mesh *mesh, const scadyn_opts *opts, double k, int32_t nmax, double *Taa,
                      double *Tab, double *Tba, double *Tbb, int32_t *iters) {
    scadyn_operator *op = nullptr;
    int rc = scadyn_op_create(mesh, opts, k, &op);
    if (rc) return rc;
    rc = scadyn_op_tmatrix(op, nmax, Taa, Tab, Tba, Tbb, iters);
    scadyn_op_destroy(op);
    return rc;
}

int scadyn_op_profile(scadyn_operator *op, int32_t enable) {
    ABI_TRY
    if (!op) fail(SCADYN_ERR_ARG, "null operator");
    op->prof.on = enable != 0;
    for (int i = 0; i < 2 * SCADYN_KCLASS_COUNT; ++i) op->prof.ms[i] = op->prof.cnt[i] = op->prof.byt[i] = 0;
    ABI_CATCH
}

int scadyn_op_timings(scadyn_operator *op, double *t) {
    ABI_TRY
    if (!op || !t) fail(SCADYN_ERR_ARG, "null argument");
    op->prof.collect();
    for (int i = 0; i < SCADYN_KCLASS_COUNT; ++i) {
        const int j = i + SCADYN_KCLASS_COUNT;
        t[3 * i] = op->prof.ms[i] + op->prof.ms[j];
        t[3 * i + 1] = op->prof.cnt[i] + op->prof.cnt[j];
        t[3 * i + 2] = op->prof.byt[i] + op->prof.byt[j];
    }
    ABI_CATCH
}

int scadyn_op_timings_prec(scadyn_operator *op, int32_t precision, double *t) {
    ABI_TRY
    if (!op || !t) fail(SCADYN_ERR_ARG, "null argument");
    if (precision != SCADYN_PREC_FP64 && precision != SCADYN_PREC_MIXED) fail(SCADYN_ERR_ARG, "bad precision");
    op->prof.collect();
    const int po = precision == SCADYN_PREC_MIXED ? SCADYN_KCLASS_COUNT : 0;
    for (int i = 0; i < SCADYN_KCLASS_COUNT; ++i) {
        const bool mv = i < SCADYN_KCLASS_KRYLOV;  // matvec classes split by precision
        const int j = i + SCADYN_KCLASS_COUNT;
        t[3 * i] = mv ? op->prof.ms[i + po] : op->prof.ms[i] + op->prof.ms[j];
        t[3 * i + 1] = mv ? op->prof.cnt[i + po] : op->prof.cnt[i] + op->prof.cnt[j];
        t[3 * i + 2] = mv ? op->prof.byt[i + po] : op->prof.byt[i] + op->prof.byt[j];
    }
    ABI_CATCH
}

int scadyn_footprint(const scadyn_mesh *mesh, const scadyn_opts *opts, int32_t nmax, int64_t *bytes) {
    ABI_TRY
    if (!mesh || !bytes) fail(SCADYN_ERR_ARG, "null argument");
    scadyn_opts o;
    if (opts)
        o = *opts;
    else
        scadyn_opts_default(&o);
    if (nmax < 1) fail(SCADYN_ERR_ARG, "nmax must be >= 1");
    check_opts(o);
    HostGeometry g;
    build_geometry(*mesh, o, g);
    NearPairs np;
    build_near_pairs(g, np);
    const int ncol = 2 * ((nmax + 1) * (nmax + 1) - 1);
    const int B = o.batch > 0 ? std::min(o.batch, ncol) : std::min(ncol, 64);
    *bytes = (int64_t)footprint_of(g, (long long)np.T.size(), o, nmax, B).peak();
    ABI_CATCH
}

int scadyn_server_create(const scadyn_server_opts *so, scadyn_server **out) {
    ABI_TRY
    if (!out) fail(SCADYN_ERR_ARG, "null argument");
    std::unique_ptr<scadyn_server> sv(new scadyn_server());
    sv->device = so ? so->device : 0;
    sv->max_active = (so && so->max_active > 0) ? so->max_active : 8;
    sv->host_only = so && so->host_only != 0;
    size_t budget = so ? (size_t)std::max<int64_t>(so->hbm_budget, 0) : 0;
    if (budget == 0 && sv->host_only) fail(SCADYN_ERR_ARG, "a host-only server needs an explicit hbm_budget");
    if (budget == 0) {
        HIP_CHECK(hipSetDevice(sv->device));
        size_t fr = 0, tot = 0;
        HIP_CHECK(hipMemGetInfo(&fr, &tot));
        budget = (size_t)(0.8 * (double)fr);
    }
    sv->budget = budget;
    sv->t0 = std::chrono::steady_clock::now();
    scadyn_server *p = sv.get();
    sv->worker = std::thread([p] { p->run(); });
    *out = sv.release();
    ABI_CATCH
}

int scadyn_server_submit(scadyn_server *sv, const scadyn_mesh *mesh, const scadyn_opts *opts, double k, int32_t nmax,
                         int64_t *id) {
    ABI_TRY
    if (!sv || !mesh || !id) fail(SCADYN_ERR_ARG, "null argument");
    if (nmax < 1) fail(SCADYN_ERR_ARG, "nmax must be >= 1");
    if (!(k > 0.0)) fail(SCADYN_ERR_ARG, "wavenumber must be positive");
    std::unique_ptr<Request> r(new Request());
    if (opts)
        r->opts = *opts;
    else
        scadyn_opts_default(&r->opts);
    const int ncol = 2 * ((nmax + 1) * (nmax + 1) - 1);
    // a fixed slot count makes a request's arithmetic independent of its neighbours
    if (r->opts.batch <= 0) r->opts.batch = std::min(ncol, 32);
    r->opts.device = sv->device;
    // validate and size on the host, before anything reaches the device
    check_opts(r->opts);
    HostGeometry g;
    build_geometry(*mesh, r->opts, g);
    NearPairs np;
    build_near_pairs(g, np);
    r->fp = footprint_of(g, (long long)np.T.size(), r->opts, nmax, std::min(r->opts.batch, ncol)).peak();
    if (r->fp > sv->budget)
        fail(SCADYN_ERR_MEMORY, "request needs " + std::to_string(r->fp) + " device bytes, budget is " +
                                    std::to_string(sv->budget));
    r->n_node = mesh->n_node;
    r->n_tet = mesh->n_tet;
    r->coord.assign(mesh->coord, mesh->coord + 3 * (size_t)mesh->n_node);
    r->etopol.assign(mesh->etopol, mesh->etopol + 4 * (size_t)mesh->n_tet);
    r->param.assign(mesh->param, mesh->param + 2 * (size_t)mesh->n_tet);
    r->k = k;
    r->nmax = nmax;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        r->id = sv->next_id++;
        r->t[0] = sv->now();
        *id = r->id;
        sv->queue.push_back(r.get());
        sv->all[r->id] = std::move(r);
    }
    sv->cv.notify_all();
    ABI_CATCH
}

int scadyn_server_status(scadyn_server *sv, int64_t id, int32_t *state, double times[4]) {
    ABI_TRY
    if (!sv) fail(SCADYN_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) fail(SCADYN_ERR_ARG, "unknown request id");
    if (state) *state = it->second->state;
    if (times)
        for (int i = 0; i < 4; ++i) times[i] = it->second->t[i];
    ABI_CATCH
}

int scadyn_server_wait(scadyn_server *sv, int64_t id, double *Taa, double *Tab, double *Tba, double *Tbb,
                       int32_t *iters, double times[4]) {
    ABI_TRY
    if (!sv) fail(SCADYN_ERR_ARG, "null server");
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->all.find(id);
        if (it == sv->all.end()) fail(SCADYN_ERR_ARG, "unknown request id");
        Request *p = it->second.get();
        sv->done_cv.wait(lk, [p] { return p->state == REQ_DONE || p->state == REQ_FAILED; });
        r = std::move(it->second);
        sv->all.erase(it);
    }
    if (times)
        for (int i = 0; i < 4; ++i) times[i] = r->t[i];
    if (r->state == REQ_FAILED) fail(r->err, r->msg);
    const int nm = (r->nmax + 1) * (r->nmax + 1) - 1;
    const size_t blk = (size_t)2 * nm * nm;
    double *outs[4] = {Taa, Tab, Tba, Tbb};
    for (int q = 0; q < 4; ++q)
        if (outs[q]) std::copy(r->T.begin() + q * blk, r->T.begin() + (q + 1) * blk, outs[q]);
    if (iters) std::copy(r->iters.begin(), r->iters.end(), iters);
    ABI_CATCH
}

int scadyn_server_cancel(scadyn_server *sv, int64_t id) {
    ABI_TRY
    if (!sv) fail(SCADYN_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) fail(SCADYN_ERR_ARG, "unknown request id");
    scadyn::Request *r = it->second.get();
    if (r->state == scadyn::REQ_QUEUED) {
        // not admitted yet: leaves the queue at once
        for (auto q = sv->queue.begin(); q != sv->queue.end(); ++q)
            if (*q == r) {
                sv->queue.erase(q);
                break;
            }
        r->err = SCADYN_ERR_ARG;
        r->msg = "request cancelled";
        r->t[3] = sv->now();
        r->state = scadyn::REQ_FAILED;
        sv->done_cv.notify_all();
    } else if (r->state == scadyn::REQ_RUNNING) {
        r->cancel.store(true);  // the worker ends it before its next step
    }
    ABI_CATCH
}

void scadyn_server_destroy(scadyn_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        // requests still queued are dropped
        for (scadyn::Request *r : sv->queue) {
            r->state = scadyn::REQ_FAILED;
            r->err = SCADYN_ERR_ARG;
            r->msg = "server destroyed";
        }
        sv->queue.clear();
    }
    sv->cv.notify_all();
    if (sv->worker.joinable()) sv->worker.join();
    sv->done_cv.notify_all();
    delete sv;
}

}  // extern "C"
