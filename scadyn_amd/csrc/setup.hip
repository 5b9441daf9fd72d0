// FP64 setup kernels of one wavelength bar:
//   k_projection   pFFT projection coefficients S, Sx, Sy, Sz per tet
//                  (pfft_projection_const, src/projection.f90:103-165)
//   k_green_grid   circulant-embedded Green's function on the padded grid
//                  (build_G, src/build_G.f90:22-61)
//   k_near_blocks  precorrected near-zone 3x3 blocks, FP64 singular integrals,
//                  rounded to complex64 (src/precorrection.f90:182-258,
//                  :729-806; src/singularity_subtraction.f90:15-243)
//   k_vswf         regular VSWFs at tet centroids times sqrt(vol)
//                  (vswf2constant + calc_MN, src/T_matrix.f90:280-305,
//                  src/mie.f90:408-496)
#include <cmath>

#include "setup.hpp"

namespace scadyn {

__constant__ double c_tet24_p[24 * 3];
__constant__ double c_tet24_w[24];
__constant__ double c_tri7_p[7 * 2];
__constant__ double c_tri7_w[7];
__constant__ double c_cube98[98 * 3];  // 2*box_delta * sample points

static const double PI_D = 3.141592653589793;

// ---- quadrature tables (published rules; see the quadrature test) -------
static const double h_tet24_p[24 * 3] = {
#define A_ 3.561913862225449e-01
#define B_ 2.146028712591517e-01
#define C_ 8.779781243961660e-01
#define D_ 4.067395853461134e-02
#define E_ 3.298632957317306e-02
#define F_ 3.223378901422757e-01
#define G_ 2.696723314583159e-01
#define H_ 6.366100187501753e-02
#define I_ 6.030056647916491e-01
    A_, B_, B_, B_, B_, B_, B_, B_, A_, B_, A_, B_,
    C_, D_, D_, D_, D_, D_, D_, D_, C_, D_, C_, D_,
    E_, F_, F_, F_, F_, F_, F_, F_, E_, F_, E_, F_,
    G_, H_, H_, H_, G_, H_, H_, H_, G_, I_, H_, H_,
    H_, I_, H_, H_, H_, I_, H_, G_, I_, G_, I_, H_,
    I_, H_, G_, H_, I_, G_, G_, H_, I_, I_, G_, H_,
#undef A_
#undef B_
#undef C_
#undef D_
#undef E_
#undef F_
#undef G_
#undef H_
#undef I_
};
static const double h_tet24_w[24] = {
    6.653791709694644e-03, 6.653791709694644e-03, 6.653791709694644e-03, 6.653791709694644e-03,
    1.679535175886776e-03, 1.679535175886776e-03, 1.679535175886776e-03, 1.679535175886776e-03,
    9.226196923942397e-03, 9.226196923942397e-03, 9.226196923942397e-03, 9.226196923942397e-03,
    8.035714285714281e-03, 8.035714285714281e-03, 8.035714285714281e-03, 8.035714285714281e-03,
    8.035714285714281e-03, 8.035714285714281e-03, 8.035714285714281e-03, 8.035714285714281e-03,
    8.035714285714281e-03, 8.035714285714281e-03, 8.035714285714281e-03, 8.035714285714281e-03};
static const double h_tri7_p[7 * 2] = {
    3.333333333333330e-01, 3.333333333333330e-01, 5.971587178977000e-02, 4.701420641051150e-01,
    4.701420641051150e-01, 4.701420641051150e-01, 4.701420641051150e-01, 5.971587178977000e-02,
    7.974269853530870e-01, 1.012865073234560e-01, 1.012865073234560e-01, 1.012865073234560e-01,
    1.012865073234560e-01, 7.974269853530870e-01};
static const double h_tri7_w[7] = {1.125000000000000e-01, 6.619707639425300e-02, 6.619707639425300e-02,
                                   6.619707639425300e-02, 6.296959027241350e-02, 6.296959027241350e-02,
                                   6.296959027241350e-02};

const double *host_tet24_p() { return h_tet24_p; }
const double *host_tet24_w() { return h_tet24_w; }
const double *host_tri7_p() { return h_tri7_p; }
const double *host_tri7_w() { return h_tri7_w; }

void upload_tables(const double *cube98_scaled, hipStream_t s) {
    HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tet24_p), h_tet24_p, sizeof(h_tet24_p), 0,
                                     hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tet24_w), h_tet24_w, sizeof(h_tet24_w), 0,
                                     hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tri7_p), h_tri7_p, sizeof(h_tri7_p), 0,
                                     hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_tri7_w), h_tri7_w, sizeof(h_tri7_w), 0,
                                     hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_cube98), cube98_scaled, 98 * 3 * sizeof(double), 0,
                                     hipMemcpyHostToDevice, s));
}

// ---- small vector helpers ---------------------------------------------------
struct v3 {
    double x, y, z;
};
__device__ inline v3 mk3(double a, double b, double c) { return {a, b, c}; }
__device__ inline v3 sub3(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline v3 add3(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline v3 mul3(v3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline v3 cross3(v3 a, v3 b) {
    return {a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x};
}
// norm(r, rp) = sqrt((r1-rp1)^2 + (r2-rp2)^2 + (r3-rp3)^2)  (src/common.f90:219-225)
__device__ inline double dist3(v3 a, v3 b) {
    double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return sqrt(dx * dx + dy * dy + dz * dz);
}
// exp(ikR)/(4 pi R)  (src/common.f90:271-279)
__device__ inline cd green(double R, double k) {
    double s, c;
    sincos(k * R, &s, &c);
    double d = 4.0 * PI_D * R;
    return cmk(c / d, s / d);
}
__device__ inline v3 tet_vertex(const double *tc, int a) { return mk3(tc[3 * a], tc[3 * a + 1], tc[3 * a + 2]); }

// faces opposite each vertex (tetra_face, src/common.f90:296-322)
__device__ inline void tet_face(const double *tc, int f, v3 &p0, v3 &p1, v3 &p2) {
    switch (f) {
        case 0: p0 = tet_vertex(tc, 1); p1 = tet_vertex(tc, 2); p2 = tet_vertex(tc, 3); break;
        case 1: p0 = tet_vertex(tc, 0); p1 = tet_vertex(tc, 3); p2 = tet_vertex(tc, 2); break;
        case 2: p0 = tet_vertex(tc, 0); p1 = tet_vertex(tc, 1); p2 = tet_vertex(tc, 3); break;
        default: p0 = tet_vertex(tc, 0); p1 = tet_vertex(tc, 2); p2 = tet_vertex(tc, 1); break;
    }
}
// tri_n_vectors (src/common.f90:326-339)
__device__ inline v3 tri_normal(v3 p0, v3 p1, v3 p2) {
    v3 n = cross3(sub3(p1, p0), sub3(p2, p0));
    double len = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    return mk3(n.x / len, n.y / len, n.z / len);
}
// linmap_tri area factor ww1 = sqrt(|a|^2|b|^2 - (a.b)^2) (src/common.f90:1380-1408)
__device__ inline double tri_ww1(v3 p0, v3 p1, v3 p2) {
    v3 a = sub3(p1, p0), b = sub3(p2, p0);
    double aa = a.x * a.x + a.y * a.y + a.z * a.z;
    double wwx = aa * b.x, wwy = aa * b.y, wwz = aa * b.z;
    double ab = a.x * b.x + a.y * b.y + a.z * b.z;
    return sqrt((wwx * b.x + wwy * b.y + wwz * b.z) - ab * ab);
}
__device__ inline v3 tri_point(v3 p0, v3 p1, v3 p2, double xi, double eta) {
    v3 a = sub3(p1, p0), b = sub3(p2, p0);
    return add3(add3(p0, mul3(a, xi)), mul3(b, eta));
}
// linmap_tet point (src/common.f90:1354-1376)
__device__ inline v3 tet_point(const double *tc, double x1, double x2, double x3) {
    double N1 = 1 - x1 - x2 - x3;
    v3 p = add3(add3(add3(mul3(tet_vertex(tc, 0), N1), mul3(tet_vertex(tc, 1), x1)),
                     mul3(tet_vertex(tc, 2), x2)),
                mul3(tet_vertex(tc, 3), x3));
    return p;
}

// ---- analytical integrals of 1/R and R over a flat triangle -----------------
// (src/singularity_subtraction.f90:15-108), returns K1 = int 1/R dS, K2 = int R dS
__device__ void analytic_tri(v3 p0, v3 p1, v3 p2, v3 n, v3 r, double &K1, double &K2) {
    v3 p[3] = {p0, p1, p2};
    const double w0 = dot3(sub3(r, p0), n);
    const double aw0 = fabs(w0);
    double k1 = 0.0;
    double tl[3], I1[3];
    const int e1[3] = {1, 2, 0}, e2[3] = {2, 0, 1};
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        v3 pl1 = p[e1[l]], pl2 = p[e2[l]];
        v3 slv = sub3(pl2, pl1);
        double nsl = sqrt(dot3(slv, slv));
        v3 sl = mk3(slv.x / nsl, slv.y / nsl, slv.z / nsl);
        v3 ml = cross3(sl, n);
        v3 pl1mr = sub3(pl1, r);
        double tl0 = dot3(pl1mr, ml);
        v3 pl2mpl1 = sub3(pl2, pl1);
        double ll = sqrt(dot3(pl2mpl1, pl2mpl1));
        double sml = dot3(pl1mr, sl);
        double spl = sml + ll;
        v3 rmplm1 = sub3(r, pl2);
        double Rpl = sqrt(dot3(rmplm1, rmplm1));
        double Rml = sqrt(dot3(pl1mr, pl1mr));
        double R0l = sqrt(tl0 * tl0 + w0 * w0);
        double Km3 = 0.0;
        if (aw0 > 1.0e-10) {
            double Bl = atan((tl0 * spl) / (R0l * R0l + aw0 * Rpl)) -
                        atan((tl0 * sml) / (R0l * R0l + aw0 * Rml));
            Km3 = 1.0 / aw0 * Bl;
        }
        double Il;
        if ((Rml + sml) > (Rpl - spl))
            Il = log(fabs((Rpl + spl) / (Rml + sml)));
        else
            Il = log(fabs((Rml - sml) / (Rpl - spl)));
        k1 = k1 + 1.0 / (-1.0 + 2.0) * (-1.0 * w0 * w0 * Km3 + tl0 * Il);
        tl[l] = tl0;
        I1[l] = 1.0 / (1.0 + 1.0) * (spl * Rpl - sml * Rml + 1.0 * R0l * R0l * Il);
    }
    K1 = k1;
    K2 = 1.0 / (1.0 + 2.0) * (1.0 * w0 * w0 * k1 + (tl[0] * I1[0] + tl[1] * I1[1] + tl[2] * I1[2]));
}

// G - (1/R - k^2 R / 2)/(4 pi), the smooth remainder (singularity_subtraction.f90:146-152)
__device__ inline cd green_remainder(double R, double k) {
    double f = 1.0 / (4 * PI_D) * 1.0 * 1.0 * (1.0 / R);
    f = f + 1.0 / (4 * PI_D) * (k * k) * (-0.5) * R;
    cd G0 = green(R, k);
    return cmk(G0.x - f, G0.y);
}

// int_V G(rf, r') dV' over tet B (singularity_subtraction_int_V_G), 24-point rule
__device__ cd int_V_G(v3 rf, double k, const double *tcB, double vB, double cut) {
    cd intR = cmk(0, 0);
    for (int i = 0; i < 24; ++i) {
        v3 pb = tet_point(tcB, c_tet24_p[3 * i], c_tet24_p[3 * i + 1], c_tet24_p[3 * i + 2]);
        double wb = c_tet24_w[i] * 6.0 * vB;
        double RR = dist3(pb, rf);
        cd II = (RR < cut) ? cmk(0.0, k / (4.0 * PI_D)) : green_remainder(RR, k);
        intR = cadd(intR, cscale(II, wb));
    }
    double r4a = 0.0, r4b = 0.0;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        v3 p0, p1, p2;
        tet_face(tcB, f, p0, p1, p2);
        v3 n = tri_normal(p0, p1, p2);
        double K1, K2;
        analytic_tri(p0, p1, p2, n, rf, K1, K2);
        double a1 = dot3(n, sub3(p0, rf));
        r4a = r4a + a1 * K1 * 0.5;
        r4b = r4b + a1 * K2 * 0.25;
    }
    double f2 = 1.0 / (4 * PI_D) * 1.0 * 1.0 * r4a;
    f2 = f2 + 1.0 / (4 * PI_D) * (k * k) * (-0.5) * r4b;
    return cmk(intR.x + f2, intR.y);
}

// ---- projection coefficients ----------------------------------------------
__global__ __launch_bounds__(128) void k_projection(const double *__restrict__ tcoord,
                                                    const double *__restrict__ vol,
                                                    const int *__restrict__ tet_cube, GridDesc gd,
                                                    double k, const cd *__restrict__ WW,
                                                    cd *__restrict__ S4, int N) {
    __shared__ cd Q[4][98];
    const int t = blockIdx.x;
    if (t >= N) return;
    const double *tc = tcoord + 12 * (size_t)t;
    const int cube = tet_cube[t];
    const int cx = cube % gd.ncx, cy = (cube / gd.ncx) % gd.ncy, cz = cube / (gd.ncx * gd.ncy);
    v3 node1 = mk3(gd.min[0] + gd.delta * (2 * cx), gd.min[1] + gd.delta * (2 * cy),
                   gd.min[2] + gd.delta * (2 * cz));
    v3 cp = mk3(node1.x + gd.box_delta / 2.0, node1.y + gd.box_delta / 2.0, node1.z + gd.box_delta / 2.0);
    const double v = vol[t];
    for (int m = threadIdx.x; m < 98; m += blockDim.x) {
        v3 test = add3(mk3(c_cube98[3 * m], c_cube98[3 * m + 1], c_cube98[3 * m + 2]), cp);
        // 1-point tet rule: centroid, weight W0*6*vol
        v3 pb = tet_point(tc, 0.25, 0.25, 0.25);
        double wb = 1.666666666666667e-01 * 6.0 * v;
        cd q = cscale(green(dist3(test, pb), k), wb);
        cd q2x = cmk(0, 0), q2y = cmk(0, 0), q2z = cmk(0, 0);
        for (int f = 0; f < 4; ++f) {
            v3 p0, p1, p2;
            tet_face(tc, f, p0, p1, p2);
            v3 pt = tri_point(p0, p1, p2, 3.333333333333330e-01, 3.333333333333330e-01);
            double wt = 0.5 * tri_ww1(p0, p1, p2);
            v3 nv = tri_normal(p0, p1, p2);
            cd q2 = cscale(green(dist3(test, pt), k), wt);
            q2x = cadd(q2x, cscale(q2, nv.x));
            q2y = cadd(q2y, cscale(q2, nv.y));
            q2z = cadd(q2z, cscale(q2, nv.z));
        }
        Q[0][m] = q;
        Q[1][m] = q2x;
        Q[2][m] = q2y;
        Q[3][m] = q2z;
    }
    __syncthreads();
    const double sv = sqrt(v);
    for (int e = threadIdx.x; e < 108; e += blockDim.x) {
        const int j = e % 27, ch = e / 27;
        cd acc = cmk(0, 0);
        for (int m = 0; m < 98; ++m) cfma(acc, WW[(size_t)m * 27 + j], Q[ch][m]);
        S4[((size_t)t * 27 + j) * 4 + ch] = cmk(acc.x / sv, acc.y / sv);
    }
}

// ---- Green's function on the padded grid -----------------------------------
// layout [Lz][Ly][Lx]; offsets d with |d| <= N-1 map to index d (d >= 0) or L+d
__global__ void k_green_grid(cd *__restrict__ A, GridDesc gd, int Lx, int Ly, int Lz, double k) {
    const size_t tot = (size_t)Lx * Ly * Lz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
         e += (size_t)gridDim.x * blockDim.x) {
        int x = (int)(e % Lx), y = (int)((e / Lx) % Ly), z = (int)(e / ((size_t)Lx * Ly));
        int dx = x < gd.Nx ? x : (x > Lx - gd.Nx ? Lx - x : -1);
        int dy = y < gd.Ny ? y : (y > Ly - gd.Ny ? Ly - y : -1);
        int dz = z < gd.Nz ? z : (z > Lz - gd.Nz ? Lz - z : -1);
        cd v = cmk(0, 0);
        if (dx >= 0 && dy >= 0 && dz >= 0 && (dx | dy | dz) != 0) {
            // R = norm(nodes(:,1), nodes(:,las)) with nodes = min + delta*(k-1)
            double ax = gd.min[0] - (gd.min[0] + gd.delta * dx);
            double ay = gd.min[1] - (gd.min[1] + gd.delta * dy);
            double az = gd.min[2] - (gd.min[2] + gd.delta * dz);
            double R = sqrt(ax * ax + ay * ay + az * az);
            if (R > 0.0) v = green(R, k);
        }
        A[e] = v;
    }
}

// even octant of FFT(G), scaled by 1/(Lx Ly Lz)
__global__ void k_fg_octant(const cd *__restrict__ F, cd *__restrict__ fg, int Lx, int Ly, int Lz,
                            int Hx, int Hy, int Hz, double scale) {
    const size_t tot = (size_t)Hx * Hy * Hz;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
         e += (size_t)gridDim.x * blockDim.x) {
        int x = (int)(e % Hx), y = (int)((e / Hx) % Hy), z = (int)(e / ((size_t)Hx * Hy));
        cd v = F[((size_t)z * Ly + y) * Lx + x];
        fg[e] = cmk(v.x * scale, v.y * scale);
    }
}

// ---- near-zone G between box nodes for the 27 cube offsets ---------------------
// Gtab[off][i][j] = G(node_i of test box, node_j of basis box), 0 when the
// nodes coincide (compute_pG, src/precorrection.f90:21-46)
__global__ void k_pg_table(cd *__restrict__ Gtab, GridDesc gd, double k) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 27 * 27 * 27) return;
    const int j = e % 27, i = (e / 27) % 27, off = e / 729;
    const int ox = off % 3 - 1, oy = (off / 3) % 3 - 1, oz = off / 9 - 1;
    // test box based at node (0,0,0) of a reference cube (2,2,2), basis box at
    // (2+2ox, 2+2oy, 2+2oz); coordinates as min + delta*index
    int ti[3] = {2 + i % 3, 2 + (i / 3) % 3, 2 + i / 9};
    int bi[3] = {2 + 2 * ox + j % 3, 2 + 2 * oy + (j / 3) % 3, 2 + 2 * oz + j / 9};
    cd v = cmk(0, 0);
    if (!(ti[0] == bi[0] && ti[1] == bi[1] && ti[2] == bi[2])) {
        double d2 = 0;
        for (int c = 0; c < 3; ++c) {
            double a = (gd.min[c] + gd.delta * ti[c]) - (gd.min[c] + gd.delta * bi[c]);
            d2 += a * a;
        }
        v = green(sqrt(d2), k);
    }
    Gtab[e] = v;
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One wave64 per stored block; 4 blocks per workgroup.
__global__ __launch_bounds__(256) void k_near_blocks(
    const int *__restrict__ PT, const int *__restrict__ PB, const uint8_t *__restrict__ POFF,
    long long P, const double *__restrict__ tcoord, const double *__restrict__ vol,
    const cd *__restrict__ param, const cd *__restrict__ S4, const cd *__restrict__ Gtab, double k,
    cf *__restrict__ blocks) {
    __shared__ cd u_s[4][4][27];  // per wave: G * S_B channels
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long p0 = (long long)blockIdx.x * 4 + wv;
    const bool valid = p0 < P;
    const long long p = valid ? p0 : P - 1;  // idle waves recompute the last block, never store
    const int T = PT[p], B = PB[p], off = POFF[p];
    const double *tcT = tcoord + 12 * (size_t)T;
    const double *tcB = tcoord + 12 * (size_t)B;
    const double vT = vol[T], vB = vol[B];
    const cd *ST = S4 + (size_t)T * 108, *SB = S4 + (size_t)B * 108;
    const cd *G = Gtab + (size_t)off * 729;

    // --- correction: u_c = G S_B[c]  (precorrect_const, src/precorrection.f90:50-82)
    for (int e = lane; e < 108; e += 64) {
        const int i = e % 27, ch = e / 27;
        cd acc = cmk(0, 0);
        for (int j = 0; j < 27; ++j) cfma(acc, G[i * 27 + j], SB[j * 4 + ch]);
        u_s[wv][ch][i] = acc;
    }
    __syncthreads();
    // 10 bilinear forms: a = S_T . u_0 ; corr2(a,b) = D_a,T . u_{1+b}
    cd cform = cmk(0, 0);
    if (lane < 10) {
        const int ca = lane == 0 ? 0 : 1 + (lane - 1) / 3;
        const int cb = lane == 0 ? 0 : 1 + (lane - 1) % 3;
        for (int i = 0; i < 27; ++i) cfma(cform, ST[i * 4 + ca], u_s[wv][cb][i]);
    }

    // --- V_V: 24-point rule on both tets (integrate_V_V_G, src/precorrection.f90:729-760)
    cd vv = cmk(0, 0);
    if (lane < 24) {
        double wsum = 0.0;
        for (int i = 0; i < 24; ++i) wsum += c_tet24_w[i] * 6.0 * vB;
        const double cut = 1.0e-10 * pow(wsum, (double)(1.0f / 3.0f));
        v3 rf = tet_point(tcT, c_tet24_p[3 * lane], c_tet24_p[3 * lane + 1], c_tet24_p[3 * lane + 2]);
        double wt = c_tet24_w[lane] * 6.0 * vT;
        vv = cscale(int_V_G(rf, k, tcB, vB, cut), wt);
    }

    // --- dV_dV: face pairs x 7-point rules (integrate_dV_dV_G, :764-806)
    cd dd[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) dd[q] = cmk(0, 0);
    for (int item = lane; item < 112; item += 64) {
        const int ft = item / 28, fb = (item / 7) % 4, i = item % 7;
        v3 t0, t1, t2, b0, b1, b2;
        tet_face(tcT, ft, t0, t1, t2);
        tet_face(tcB, fb, b0, b1, b2);
        v3 Tn = tri_normal(t0, t1, t2);
        v3 Bn = tri_normal(b0, b1, b2);
        double tww = tri_ww1(t0, t1, t2);
        double bww = tri_ww1(b0, b1, b2);
        v3 rf = tri_point(t0, t1, t2, c_tri7_p[2 * i], c_tri7_p[2 * i + 1]);
        double wt = c_tri7_w[i] * tww;
        double wsum = 0.0;
        for (int j = 0; j < 7; ++j) wsum += c_tri7_w[j] * bww;
        const double cut = 1.0e-10 * sqrt(wsum);
        cd intR = cmk(0, 0);
        for (int j = 0; j < 7; ++j) {
            v3 pb = tri_point(b0, b1, b2, c_tri7_p[2 * j], c_tri7_p[2 * j + 1]);
            double R = dist3(pb, rf);
            cd II = (R < cut) ? cmk(0.0, k / (4.0 * PI_D)) : green_remainder(R, k);
            intR = cadd(intR, cscale(II, c_tri7_w[j] * bww));
        }
        double K1, K2;
        analytic_tri(b0, b1, b2, Bn, rf, K1, K2);
        double f2 = 1.0 / (4 * PI_D) * 1.0 * 1.0 * K1;
        f2 = f2 + 1.0 / (4 * PI_D) * (k * k) * (-0.5) * K2;
        cd s = cscale(cmk(intR.x + f2, intR.y), wt);
        const double tn[3] = {Tn.x, Tn.y, Tn.z}, bn[3] = {Bn.x, Bn.y, Bn.z};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) dd[a + 3 * b] = cadd(dd[a + 3 * b], cscale(s, tn[a] * bn[b]));
    }

    // --- reductions over the wave
    double vvr = wave_sum(vv.x), vvi = wave_sum(vv.y);
    double ddr[9], ddi[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        ddr[q] = wave_sum(dd[q].x);
        ddi[q] = wave_sum(dd[q].y);
    }
    double cfr[10], cfi[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        cfr[q] = __shfl(cform.x, q, 64);
        cfi[q] = __shfl(cform.y, q, 64);
    }
    if (lane == 0 && valid) {
        const double sc = sqrt(vT * vB);
        const double inv = 1.0 / sqrt(vB * vT);
        const double k2 = k * k;
        cd epsm1 = cmk(0, 0);
        const bool self = (T == B);
        if (self) {
            cd e = param[T];
            epsm1 = cdiv(cmk(1.0, 0.0), cmk(e.x - 1.0, e.y));
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                cd corr = (a == b) ? cmk(cfr[0] * sc, cfi[0] * sc) : cmk(0, 0);
                cd corr2 = cmk(cfr[1 + 3 * a + b] * sc, cfi[1 + 3 * a + b] * sc);
                cd alok1 = (a == b) ? cmk(vvr, vvi) : cmk(0, 0);
                cd t1 = csub(alok1, corr);
                cd t2 = csub(cmk(ddr[a + 3 * b], ddi[a + 3 * b]), corr2);
                cd inner = cadd(cmk(-k2 * t1.x, -k2 * t1.y), t2);
                cd alok = (self && a == b) ? epsm1 : cmk(0, 0);
                cd val = cadd(alok, cmk(inv * inner.x, inv * inner.y));
                blocks[p * 9 + a + 3 * b] = make_float2((float)val.x, (float)val.y);
            }
    }
}

// ---- regular VSWFs -----------------------------------------------------------
// real spherical Bessel j_0..j_nmax(x) by Miller's downward recurrence,
// normalised with j_0 or j_1 (whichever is larger, as cspherebessel)
__device__ void sph_bessel_j(int nmax, double x, double *j) {
    double s, c;
    sincos(x, &s, &c);
    double j0 = s / x, j1 = (j0 - c) / x;
    j[0] = j0;
    if (nmax >= 1) j[1] = j1;
    if (nmax < 2) return;
    int start = nmax + 20 + (int)x;
    double f1 = 0.0, f0 = 1.0e-200, fa = 0, fb = 0;
    double scale = 1.0;
    // f_{n-1} = (2n+1)/x f_n - f_{n+1}
    for (int n = start; n >= 1; --n) {
        double fm = (2.0 * n + 1.0) / x * f0 - f1;
        f1 = f0;
        f0 = fm;
        if (fabs(f0) > 1e200) {  // renormalise
            f0 *= 1e-200;
            f1 *= 1e-200;
            for (int q = n; q <= nmax; ++q) j[q] *= 1e-200;
        }
        if (n - 1 <= nmax) j[n - 1] = f0;
        if (n <= nmax) j[n] = f1;
    }
    fa = j[0];
    fb = j[1];
    scale = (fabs(j0) > fabs(j1)) ? j0 / fa : j1 / fb;
    for (int n = 0; n <= nmax; ++n) j[n] *= scale;
}

__global__ void k_vswf(const double *__restrict__ tcoord, const double *__restrict__ vol, int N,
                       int Nmax, double k, cd *__restrict__ mat, int ncol) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const double *tc = tcoord + 12 * (size_t)t;
    double r3[3];
    for (int c = 0; c < 3; ++c) r3[c] = (tc[c] + tc[3 + c] + tc[6 + c] + tc[9 + c]) / 4.0;
    const double r = sqrt(r3[0] * r3[0] + r3[1] * r3[1] + r3[2] * r3[2]);
    const double theta = (r != 0.0) ? acos(r3[2] / r) : 0.0;
    const double phi = atan2(r3[1], r3[0]);
    const double kr = k * r;
    double jn[64];
    sph_bessel_j(Nmax + 1, kr, jn);
    const double x = cos(theta), st = sin(theta), ct = x;
    double sp, cp;
    sincos(phi, &sp, &cp);
    const double xq = sqrt(1.0 - x * x);
    const double sv = sqrt(vol[t]);
    const int nm = (Nmax + 1) * (Nmax + 1) - 1;
    cd *row0 = mat + (size_t)(3 * t) * ncol;
    // associated Legendre with Condon-Shortley phase (legendre2, src/sfunctions.f90:615-686)
    double pmm = 1.0;  // P_m^m
    for (int m = 0; m <= Nmax; ++m) {
        if (m > 0) pmm = -(2.0 * m - 1.0) * xq * pmm;
        // P_{n}^m for n = m-1 (0), m, m+1, ... kept as a sliding window
        double Pprev = 0.0, Pcur = pmm;          // P_{m-1}^m = 0, P_m^m
        double Pnext = (2.0 * m + 1.0) * x * pmm;  // P_{m+1}^m
        for (int n = m; n <= Nmax; ++n) {
            if (n >= 1) {
                // factors of src/mie.f90:436-480
                const double alpha = jn[n];
                const double beta = sqrt((double)(n * (n + 1))) / kr * jn[n];
                const double gamma = (n + 1.0) / kr * jn[n] - jn[n + 1];
                const double q = (sqrt(n * (n + 1.0))) / ((n * 2.0 + 1.0) * st);
                double fnm = 1.0, fnp = 1.0;
                for (int i = 1; i <= n - m; ++i) fnm *= i;
                for (int i = 1; i <= n + m; ++i) fnp *= i;
                const double cc = sqrt((2.0 * n + 1.0) * fnm / fnp / (4.0 * PI_D));
                for (int sgn = (m == 0 ? 1 : -1); sgn <= 1; sgn += 2) {
                    const int mm = sgn * m;
                    double se, ce;
                    sincos(mm * phi, &se, &ce);
                    cd Y = cmk(Pcur * ce, Pcur * se);
                    cd Y1 = cmk(Pnext * ce, Pnext * se);
                    cd Y2 = (m == n) ? cmk(0, 0) : cmk(Pprev * ce, Pprev * se);
                    Y1 = cscale(Y1, (n - m + 1.0) / (n + 1.0));
                    Y2 = cscale(Y2, (double)(n + m) / (double)n);
                    cd B2 = csub(Y1, Y2);
                    double fB3 = (double)((float)mm * (2.0f * n + 1.0f)) /
                                 (double)((float)n * ((float)n + 1.0f));
                    cd B3 = cmk(-fB3 * Y.y, fB3 * Y.x);  // i*fB3*Y
                    B2 = cscale(B2, q);
                    B3 = cscale(B3, q);
                    // M = cc*alpha*(0, B3, -B2); N = cc*(beta*Y, gamma*B2, gamma*B3)
                    cd Mt = cscale(B3, cc * alpha), Mp = cscale(B2, -cc * alpha);
                    cd Nr = cscale(Y, cc * beta), Nt = cscale(B2, cc * gamma), Np = cscale(B3, cc * gamma);
                    // spherical -> cartesian (sph2cart_vec, src/common.f90:670-681)
                    const double H[3][3] = {{st * cp, ct * cp, -sp}, {st * sp, ct * sp, cp}, {ct, -st, 0.0}};
                    const int col = n * (n + 1) + mm - 1;
                    for (int c = 0; c < 3; ++c) {
                        cd Mc = cadd(cscale(Mt, H[c][1]), cscale(Mp, H[c][2]));
                        cd Nc = cadd(cadd(cscale(Nr, H[c][0]), cscale(Nt, H[c][1])), cscale(Np, H[c][2]));
                        row0[(size_t)c * ncol + col] = cscale(Mc, sv);
                        row0[(size_t)c * ncol + nm + col] = cscale(Nc, sv);
                    }
                }
            }
            // advance the window: P_{n+2}^m from P_{n+1}^m and P_n^m
            double Pn2 = ((2.0 * (n + 2) - 1.0) * x * Pnext - ((double)(m + n + 2) - 1.0) * Pcur) /
                         (double)(n + 2 - m);
            Pprev = Pcur;
            Pcur = Pnext;
            Pnext = Pn2;
        }
    }
}

// ---- launchers -----------------------------------------------------------------
void launch_projection(const double *tcoord, const double *vol, const int *tet_cube, const GridDesc &gd,
                       double k, const cd *WW, cd *S4, int N, hipStream_t s) {
    hipLaunchKernelGGL(k_projection, dim3(N), dim3(128), 0, s, tcoord, vol, tet_cube, gd, k, WW, S4, N);
    HIP_CHECK(hipGetLastError());
}
void launch_green_grid(cd *A, const GridDesc &gd, int Lx, int Ly, int Lz, double k, hipStream_t s) {
    hipLaunchKernelGGL(k_green_grid, dim3(2048), dim3(256), 0, s, A, gd, Lx, Ly, Lz, k);
    HIP_CHECK(hipGetLastError());
}
void launch_fg_octant(const cd *F, cd *fg, int Lx, int Ly, int Lz, hipStream_t s) {
    int Hx = Lx / 2 + 1, Hy = Ly / 2 + 1, Hz = Lz / 2 + 1;
    double scale = 1.0 / ((double)Lx * Ly * Lz);
    hipLaunchKernelGGL(k_fg_octant, dim3(1024), dim3(256), 0, s, F, fg, Lx, Ly, Lz, Hx, Hy, Hz, scale);
    HIP_CHECK(hipGetLastError());
}
void launch_pg_table(cd *Gtab, const GridDesc &gd, double k, hipStream_t s) {
    hipLaunchKernelGGL(k_pg_table, dim3((27 * 27 * 27 + 255) / 256), dim3(256), 0, s, Gtab, gd, k);
    HIP_CHECK(hipGetLastError());
}
void launch_near_blocks(const int *PT, const int *PB, const uint8_t *POFF, long long P, const double *tcoord,
                        const double *vol, const cd *param, const cd *S4, const cd *Gtab, double k, cf *blocks,
                        hipStream_t s) {
    if (P == 0) return;
    long long nb = (P + 3) / 4;
    const long long maxb = 1LL << 30;
    for (long long b0 = 0; b0 < nb; b0 += maxb) {
        long long cnt = std::min(maxb, nb - b0);
        long long p0 = b0 * 4;
        hipLaunchKernelGGL(k_near_blocks, dim3((unsigned)cnt), dim3(256), 0, s, PT + p0, PB + p0, POFF + p0,
                           P - p0, tcoord, vol, param, S4, Gtab, k, blocks + p0 * 9);
        HIP_CHECK(hipGetLastError());
    }
}
void launch_vswf(const double *tcoord, const double *vol, int N, int Nmax, double k, cd *mat, int ncol,
                 hipStream_t s) {
    if (Nmax + 2 > 64) fail(SCADYN_ERR_UNSUPPORTED, "nmax above 61");
    hipLaunchKernelGGL(k_vswf, dim3((N + 63) / 64), dim3(64), 0, s, tcoord, vol, N, Nmax, k, mat, ncol);
    HIP_CHECK(hipGetLastError());
}

}  // namespace scadyn
