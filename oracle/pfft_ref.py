"""Oracle: CPU restatement of scadyn's JVIE T-matrix solve (precorrected FFT).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the checker
of the HIP path.  The product (scadyn_amd/) never imports it.

It follows, function by function, the reference Fortran (paths relative to
the reference repository):

  geometry     src/geometry.f90:13-377      build_grid / build_box / tetras_in_cubes
  projection   src/projection.f90:17-166    pfft_projection_const (+ possu.f90:395 pinv)
  Green grid   src/build_G.f90:1-66         circulant embedding + forward FFT
  near zone    src/precorrection.f90:21-264, 729-806; singularity_subtraction.f90:15-243
  matvec       src/gmres_module.f90:57-176  compute_Ax2 (order 0)
  GMRES        src/gmres_module.f90:180-325
  VSWF         src/T_matrix.f90:280-305, src/mie.f90:408-496
  T-matrix     src/T_matrix.f90:240-277

Numerical quirks of the reference kept on purpose (they change results):
  * the grid step uses the default-real (float32) exponent 1.0/3.0
    (src/geometry.f90:88);
  * near-zone blocks are rounded to complex64 (src/precorrection.f90:251);
  * the pseudo-inverse drops singular values <= 1e-6 (src/possu.f90:413).

Pinning status of each function (tests/ holds the checks):
  quadrature tables       pinned: exact monomial integration
                          (test_oracle_quadrature.py) and bit-equal to the
                          compiled reference module (test_oracle_vs_ref.py)
  analytical_integration  pinned: brute-force quadrature of 1/R and R
                          (test_oracle_singular.py); equal to the compiled
                          reference to 1e-12 (test_oracle_vs_ref.py)
  int_V_G / int_S_G       pinned: adaptive brute force, near and singular;
                          equal to the compiled reference, with the V_V and
                          dV_dV pair integrals
  geometry                pinned: bit-equal to the compiled reference's
                          build_grid / build_box / tetras_in_cubes
  build_G + FFT conv.     pinned: direct O(n^2) Green's sum on a small grid
                          (test_oracle_operator.py)
  projection S*           pinned: fit residual and grid-translation invariance
  near zone + matvec      pinned: near-pair identity (operator == exact block)
                          and the symmetry defect (self blocks only)
  GMRES                   pinned: dense numpy.linalg.solve; restart semantics
  VSWF                    pinned: curl identities, |M|^2 normalisation, phase
                          convention, closed-form M_10 (test_oracle_vswf.py)
  T-matrix                pinned: the exact Galerkin operator's T converges to
                          Mie at about O(h^2) on faceted balls, planted
                          operator mistakes fail it, the pFFT operator and T
                          stay close to the exact ones, reciprocity on an
                          asymmetric body (test_oracle_tmatrix_pin.py)
"""
import numpy as np

from . import quadrature as quad

PI = 3.141592653589793
# Fortran default-real literal 1.0/3.0 promoted to double (src/geometry.f90:88)
F32_THIRD = float(np.float32(1.0) / np.float32(3.0))


# --------------------------------------------------------------------------
# small geometry helpers (src/common.f90:219-339)
# --------------------------------------------------------------------------

def tetra_volume(c):
    """|det[p2-p1, p3-p1, p4-p1]|/6, c shape (...,4,3) (src/common.f90:255-267)."""
    p41 = c[..., 3, :] - c[..., 0, :]
    p21 = c[..., 1, :] - c[..., 0, :]
    p31 = c[..., 2, :] - c[..., 0, :]
    dt = np.einsum('...i,...i->...', p41, np.cross(p21, p31))
    return np.sqrt(dt * dt) / 6.0


def tetra_det(c):
    J = np.stack([c[..., 1, :] - c[..., 0, :], c[..., 2, :] - c[..., 0, :],
                  c[..., 3, :] - c[..., 0, :]], axis=-1)
    return np.linalg.det(J)


# faces opposite vertex f (src/common.f90:296-322), 0-based vertex ids
FACES = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])


def tri_normal(tri):
    """Unit normal (p2-p1)x(p3-p1) (src/common.f90:326-339); tri (...,3,3)."""
    n = np.cross(tri[..., 1, :] - tri[..., 0, :], tri[..., 2, :] - tri[..., 0, :])
    return n / np.sqrt(np.sum(n * n, axis=-1))[..., None]


def linmap_tet(tet, P0, W0):
    """Map reference-tet points/weights (src/common.f90:1354-1376)."""
    N1 = 1.0 - P0[:, 0] - P0[:, 1] - P0[:, 2]
    Nw = np.stack([N1, P0[:, 0], P0[:, 1], P0[:, 2]], axis=-1)  # (q,4)
    P = np.einsum('qa,...ai->...qi', Nw, tet)
    W = W0 * 6.0 * tetra_volume(tet)[..., None]
    return P, W


def linmap_tri(tri, P0, W0):
    """Map reference-triangle points/weights (src/common.f90:1380-1408)."""
    a = tri[..., 1, :] - tri[..., 0, :]
    b = tri[..., 2, :] - tri[..., 0, :]
    aa = np.sum(a * a, axis=-1)
    ab = np.sum(a * b, axis=-1)
    bb = np.sum(b * b, axis=-1)
    ww1 = np.sqrt(aa * bb - ab * ab)
    P = tri[..., None, 0, :] + a[..., None, :] * P0[:, 0, None] + b[..., None, :] * P0[:, 1, None]
    W = W0 * ww1[..., None]
    return P, W


def Gr(rf, rp, k):
    """exp(ikR)/(4 pi R) (src/common.f90:271-279)."""
    R = np.sqrt(np.sum((rf - rp) ** 2, axis=-1))
    return np.exp(1j * k * R) / (4.0 * PI * R)


# --------------------------------------------------------------------------
# mesh and grid geometry (src/geometry.f90)
# --------------------------------------------------------------------------

class Geometry:
    pass


def build_geometry(coord, etopol, M_ex=2, grid_size=1.8):
    """build_grid + build_box + tetras_in_cubes for M_ex = 2.

    coord (Nnode,3) metres, etopol (N,4) 0-based.
    src/geometry.f90:60-125 (grid), :130-270 (boxes), :330-377 (tet->cube).
    """
    assert M_ex == 2
    g = Geometry()
    N = etopol.shape[0]
    tets = coord[etopol]                       # (N,4,3)
    vol = tetra_volume(tets)
    min_c = coord.min(axis=0)
    max_c = coord.max(axis=0)
    volsum = 0.0
    for v in vol:                              # sequential sum as the Fortran loop
        volsum += v
    delta = grid_size * 1.0 / M_ex * (volsum / float(N)) ** F32_THIRD
    n = [int(np.ceil(abs(max_c[i] - min_c[i]) / delta)) + 1 for i in range(3)]
    n = [v + (v + 1) % 2 + 1 for v in n]       # make even (M_ex 1 or 2)
    g.Nx, g.Ny, g.Nz = n
    g.delta = delta
    g.min_coord = min_c
    g.box_delta = delta * M_ex
    g.ncx, g.ncy, g.ncz = (g.Nx - 2) // 2, (g.Ny - 2) // 2, (g.Nz - 2) // 2
    g.N_cubes = g.ncx * g.ncy * g.ncz
    g.tets = tets
    g.vol = vol
    g.N = N
    # cube of every tet from its centroid (point_in_cube, src/geometry.f90:13-55)
    cp = (tets[:, 0] + tets[:, 1] + tets[:, 2] + tets[:, 3]) / 4
    M = np.floor((cp - min_c) / g.box_delta).astype(np.int64)
    assert (M >= 0).all() and (M[:, 0] < g.ncx).all() and (M[:, 1] < g.ncy).all() \
        and (M[:, 2] < g.ncz).all()
    g.tet_cube = M[:, 0] + g.ncx * M[:, 1] + g.ncx * g.ncy * M[:, 2]
    g.cube_tets = [[] for _ in range(g.N_cubes)]
    for t in range(N):                          # increasing tet order per cube
        g.cube_tets[g.tet_cube[t]].append(t)
    g.N_tet_cube = max(len(c) for c in g.cube_tets)
    return g


def node_xyz(g, ix, iy, iz):
    """Grid node coordinates min + delta*(k-1) (src/geometry.f90:108-118)."""
    return np.stack([g.min_coord[0] + g.delta * ix, g.min_coord[1] + g.delta * iy,
                     g.min_coord[2] + g.delta * iz], axis=-1)


def box_nodes(g, cube):
    """27 node ids (x fastest) of a cube's box, 0-based (src/geometry.f90:185-219)."""
    cx = cube % g.ncx
    cy = (cube // g.ncx) % g.ncy
    cz = cube // (g.ncx * g.ncy)
    o = np.arange(3)
    iz, iy, ix = np.meshgrid(o, o, o, indexing='ij')
    ix = (2 * cx + ix).ravel()
    iy = (2 * cy + iy).ravel()
    iz = (2 * cz + iz).ravel()
    return ix + g.Nx * iy + g.Nx * g.Ny * iz, (ix, iy, iz)


def node_coords_of_ids(g, ids):
    ix = ids % g.Nx
    iy = (ids // g.Nx) % g.Ny
    iz = ids // (g.Nx * g.Ny)
    return node_xyz(g, ix, iy, iz)


def find_near_cubes(g, cube, near_zone=1):
    """Neighbour cubes, z outer / x inner (src/geometry.f90:275-326)."""
    cx = cube % g.ncx
    cy = (cube // g.ncx) % g.ncy
    cz = cube // (g.ncx * g.ncy)
    out = []
    for dz in range(-near_zone, near_zone + 1):
        zz = cz + dz
        if not (0 <= zz < g.ncz):
            continue
        for dy in range(-near_zone, near_zone + 1):
            yy = cy + dy
            if not (0 <= yy < g.ncy):
                continue
            for dx in range(-near_zone, near_zone + 1):
                xx = cx + dx
                if 0 <= xx < g.ncx:
                    out.append(xx + g.ncx * yy + g.ncx * g.ncy * zz)
    return out


# --------------------------------------------------------------------------
# projection coefficients (src/projection.f90:17-166)
# --------------------------------------------------------------------------

def pinv(A, thr=1e-6):
    """SVD pseudo-inverse with absolute threshold (src/possu.f90:395-422)."""
    U, s, Vh = np.linalg.svd(A, full_matrices=True)
    n = A.shape[1]
    SU = np.zeros((n, A.shape[0]), dtype=complex)
    for i in range(n):
        if s[i] > thr:
            SU[i, :] = np.conj(U[:, i]) / s[i]
    return np.conj(Vh).T @ SU


def projection_const(g, k):
    """S, Sx, Sy, Sz (N,27) and indS (N,27) for basis order 0."""
    N = g.N
    P_sph = 2.0 * g.box_delta * quad.CUBE98_P              # (98,3)
    # W from the first tet's cube (src/projection.f90:74-101)
    ids0, _ = box_nodes(g, g.tet_cube[0])
    nodes0 = node_coords_of_ids(g, ids0)
    cp_cube0 = nodes0[0] + g.box_delta / 2.0
    W = Gr((P_sph + cp_cube0)[:, None, :], nodes0[None, :, :], k)   # (98,27)
    WW = pinv(W)                                            # (27,98)
    S = np.zeros((N, 27), complex)
    Sx = np.zeros((N, 27), complex)
    Sy = np.zeros((N, 27), complex)
    Sz = np.zeros((N, 27), complex)
    indS = np.zeros((N, 27), np.int64)
    for t in range(N):
        tet = g.tets[t]
        vol = tetra_volume(tet)
        Pb, wb = linmap_tet(tet, quad.TET1_P, quad.TET1_W)
        ids, _ = box_nodes(g, g.tet_cube[t])
        node1 = node_coords_of_ids(g, ids[:1])[0]
        cp_cube = node1 + g.box_delta / 2.0
        test = P_sph + cp_cube                              # (98,3)
        Q = np.sum(Gr(test[:, None, :], Pb[None, :, :], k) * wb[None, :], axis=1)
        Q2 = np.zeros((98, 3), complex)
        for f in range(4):
            tri = tet[FACES[f]]
            Pt, wt = linmap_tri(tri, quad.TRI1_P, quad.TRI1_W)
            nvec = tri_normal(tri)
            if np.dot(nvec, tri[0] - tet[f]) < 0:
                raise ValueError('inward face normal (node ordering)')
            q2 = np.sum(wt[None, :] * Gr(test[:, None, :], Pt[None, :, :], k), axis=1)
            Q2 += q2[:, None] * nvec[None, :]
        sv = np.sqrt(vol)
        S[t] = WW @ Q / sv
        Sx[t] = WW @ Q2[:, 0] / sv
        Sy[t] = WW @ Q2[:, 1] / sv
        Sz[t] = WW @ Q2[:, 2] / sv
        indS[t] = ids
    return S, Sx, Sy, Sz, indS


# --------------------------------------------------------------------------
# Green's function on the padded grid (src/build_G.f90)
# --------------------------------------------------------------------------

def green_grid(g, k):
    """Circulant-embedded G on (2Nz,2Ny,2Nx) (C order, x fastest), G(0)=0."""
    Nx, Ny, Nz = g.Nx, g.Ny, g.Nz
    A = np.zeros((2 * Nz, 2 * Ny, 2 * Nx), complex)
    iz, iy, ix = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing='ij')
    p = node_xyz(g, ix, iy, iz)
    R = np.sqrt(np.sum((p - g.min_coord) ** 2, axis=-1))
    with np.errstate(divide='ignore', invalid='ignore'):
        G = np.where(R > 0, np.exp(1j * k * R) / (4 * PI * R), 0.0)
    def mir(n, N):
        return np.where(n == 0, 0, 2 * N - n)
    for sz in (False, True):
        for sy in (False, True):
            for sx in (False, True):
                zi = mir(iz, Nz) if sz else iz
                yi = mir(iy, Ny) if sy else iy
                xi = mir(ix, Nx) if sx else ix
                A[zi, yi, xi] = G
    A[0, 0, 0] = 0.0
    return A


def build_Fg(g, k):
    """Unnormalised forward 3D DFT (FFTW sign -1) of the Green grid."""
    return np.fft.fftn(green_grid(g, k))


# --------------------------------------------------------------------------
# singular integrals (src/singularity_subtraction.f90)
# --------------------------------------------------------------------------

def analytical_integration(tri, n_vec, r):
    """[int_S 1/R dS, int_S R dS] over a flat triangle (src/singularity_subtraction.f90:15-108).

    Vectorised: tri (...,3,3), n_vec (...,3), r (...,3) broadcastable.
    """
    p = [tri[..., 0, :], tri[..., 1, :], tri[..., 2, :]]
    w0 = np.sum((r - p[0]) * n_vec, axis=-1)
    aw0 = np.abs(w0)
    K1 = 0.0
    tl = []
    Il = []
    # edges l = 1,2,3 run p(l1) -> p(l2) with (l1,l2) = (2,3),(3,1),(1,2)
    for l1, l2 in ((1, 2), (2, 0), (0, 1)):
        slv = p[l2] - p[l1]
        nsl = np.sqrt(np.sum(slv * slv, axis=-1))
        sl = slv / nsl[..., None]
        ml = np.cross(sl, n_vec)
        pl1mr = p[l1] - r
        tl0 = np.sum(pl1mr * ml, axis=-1)
        ll = nsl
        sml = np.sum(pl1mr * sl, axis=-1)
        spl = sml + ll
        Rpl = np.sqrt(np.sum((r - p[l2]) ** 2, axis=-1))
        Rml = np.sqrt(np.sum(pl1mr * pl1mr, axis=-1))
        R0l = np.sqrt(tl0 * tl0 + w0 * w0)
        with np.errstate(divide='ignore', invalid='ignore'):
            Bl = (np.arctan((tl0 * spl) / (R0l * R0l + aw0 * Rpl))
                  - np.arctan((tl0 * sml) / (R0l * R0l + aw0 * Rml)))
            Km3 = np.where(aw0 > 1.0e-10, Bl / aw0, 0.0)
            I = np.where((Rml + sml) > (Rpl - spl),
                         np.log(np.abs((Rpl + spl) / (Rml + sml))),
                         np.log(np.abs((Rml - sml) / (Rpl - spl))))
        K1 = K1 + (-1.0 * w0 * w0 * Km3 + tl0 * I)      # 1/(n+2) = 1 for n = -1
        tl.append(tl0)
        Il.append((spl, sml, Rpl, Rml, R0l, I))
    # n = 1: I_l <- (spl Rpl - sml Rml + R0l^2 I_l)/2 ; K = (w0^2 K_prev + sum t_l I_l)/3
    s = 0.0
    for t, (spl, sml, Rpl, Rml, R0l, I) in zip(tl, Il):
        I1 = 0.5 * (spl * Rpl - sml * Rml + 1.0 * R0l * R0l * I)
        s = s + t * I1
    K2 = (1.0 * w0 * w0 * K1 + s) / 3.0
    return K1, K2


def _green_minus_static(R, k):
    """G(R) - [1/R - k^2 R/2]/(4 pi); value ik/(4 pi) used at R < cut."""
    f = (1.0 / (4 * PI)) * (1.0 / R) + (1.0 / (4 * PI)) * k ** 2.0 * (-0.5) * R
    return np.exp(1j * k * R) / (4.0 * PI * R) - f


def int_V_G(rf, k, Pb, wb, Bc):
    """int_V G(rf, r') dV' (src/singularity_subtraction.f90:112-189).

    rf (...,q,3); Pb (...,p,3), wb (...,p), Bc (...,4,3) (leading dims broadcast).
    """
    cut = 1.0e-10 * np.sum(wb, axis=-1) ** F32_THIRD
    R = np.sqrt(np.sum((rf[..., :, None, :] - Pb[..., None, :, :]) ** 2, axis=-1))
    with np.errstate(divide='ignore', invalid='ignore'):
        II = np.where(R < cut[..., None, None], 1j * k / (4.0 * PI), _green_minus_static(R, k))
    intR = np.sum(II * wb[..., None, :], axis=-1)
    intR4_1 = 0.0
    intR4_2 = 0.0
    for f in range(4):
        tri = Bc[..., FACES[f], :]
        n = tri_normal(tri)
        K1, K2 = analytical_integration(tri[..., None, :, :], n[..., None, :], rf)
        a1 = np.sum(n[..., None, :] * (tri[..., None, 0, :] - rf), axis=-1)
        intR4_1 = intR4_1 + a1 * K1 * 0.5
        intR4_2 = intR4_2 + a1 * K2 * 0.25
    f2 = (1.0 / (4 * PI)) * intR4_1 + (1.0 / (4 * PI)) * k ** 2.0 * (-0.5) * intR4_2
    return intR + f2


def int_S_G(rf, tri, n, k, Pb, wb):
    """int_S G(rf, r') dS' over a flat triangle (src/singularity_subtraction.f90:193-243)."""
    cut = 1.0e-10 * np.sqrt(np.sum(wb, axis=-1))
    R = np.sqrt(np.sum((rf[..., :, None, :] - Pb[..., None, :, :]) ** 2, axis=-1))
    with np.errstate(divide='ignore', invalid='ignore'):
        II = np.where(R < cut[..., None, None], 1j * k / (4.0 * PI), _green_minus_static(R, k))
    intR = np.sum(II * wb[..., None, :], axis=-1)
    K1, K2 = analytical_integration(tri[..., None, :, :], n[..., None, :], rf)
    f2 = (1.0 / (4 * PI)) * K1 + (1.0 / (4 * PI)) * k ** 2.0 * (-0.5) * K2
    return intR + f2


def integrate_V_V_G(Tc, Bc, k):
    """sum_i wt_i int_V_B G(Pt_i, .)  (src/precorrection.f90:729-760), shape (...)."""
    Pt, wt = linmap_tet(Tc, quad.TET24_P, quad.TET24_W)
    Pb, wb = linmap_tet(Bc, quad.TET24_P, quad.TET24_W)
    intN = int_V_G(Pt, k, Pb, wb, Bc)
    return np.sum(intN * wt, axis=-1)


def integrate_dV_dV_G(Tc, Bc, k):
    """sum over face pairs n_T n_B^T int int G (src/precorrection.f90:764-806), (...,3,3)."""
    out = np.zeros(Tc.shape[:-2] + (3, 3), complex)
    for ft in range(4):
        Ttri = Tc[..., FACES[ft], :]
        Tn = tri_normal(Ttri)
        Pt, wt = linmap_tri(Ttri, quad.TRI7_P, quad.TRI7_W)
        for fb in range(4):
            Btri = Bc[..., FACES[fb], :]
            Bn = tri_normal(Btri)
            Pb, wb = linmap_tri(Btri, quad.TRI7_P, quad.TRI7_W)
            intN = int_S_G(Pt, Btri, Bn, k, Pb, wb)
            s = np.sum(intN * wt, axis=-1)
            out = out + (Tn[..., :, None] * Bn[..., None, :]) * s[..., None, None]
    return out


# --------------------------------------------------------------------------
# near-zone precorrected blocks (src/precorrection.f90:21-264)
# --------------------------------------------------------------------------

def near_pairs(g, near_zone=1):
    """Stored (T,B) pairs with B >= T in the reference's loop order."""
    pairs = []
    for tc in range(g.N_cubes):
        tests = g.cube_tets[tc]
        if not tests:
            continue
        for bc in find_near_cubes(g, tc, near_zone):
            bases = g.cube_tets[bc]
            for T in tests:
                for B in bases:
                    if B >= T:
                        pairs.append((T, B, tc, bc))
    return pairs


def pG_matrix(g, tc, bc, k):
    """27x27 G between box nodes, 0 for coincident nodes (src/precorrection.f90:21-46)."""
    it, _ = box_nodes(g, tc)
    ib, _ = box_nodes(g, bc)
    rt = node_coords_of_ids(g, it)
    rb = node_coords_of_ids(g, ib)
    R = np.sqrt(np.sum((rt[:, None, :] - rb[None, :, :]) ** 2, axis=-1))
    with np.errstate(divide='ignore', invalid='ignore'):
        G = np.exp(1j * k * R) / (4.0 * PI * R)
    G[it[:, None] == ib[None, :]] = 0.0
    return G


def near_zone_const(g, param, S, Sx, Sy, Sz, k, chunk=2048, return_f64=False):
    """Precorrected near-zone 3x3 blocks (complex64) and (T,B) index pairs."""
    pairs = near_pairs(g)
    P = len(pairs)
    T = np.array([p[0] for p in pairs], np.int64)
    B = np.array([p[1] for p in pairs], np.int64)
    TC = np.array([p[2] for p in pairs], np.int64)
    BC = np.array([p[3] for p in pairs], np.int64)
    if (tetra_det(g.tets) <= 0).any():
        raise ValueError('detJ <= 0')
    blocks = np.zeros((P, 3, 3), complex)
    D = np.stack([Sx, Sy, Sz], axis=1)                     # (N,3,27)
    eye = np.eye(3)
    for s0 in range(0, P, chunk):
        sl = slice(s0, min(P, s0 + chunk))
        t = T[sl]
        b = B[sl]
        Tc = g.tets[t]
        Bc = g.tets[b]
        vT = g.vol[t]
        vB = g.vol[b]
        VV = integrate_V_V_G(Tc, Bc, k)
        dVdV = integrate_dV_dV_G(Tc, Bc, k)
        # compute_pG for each pair's cube pair (coincident grid nodes -> 0)
        it = np.stack([box_nodes(g, c)[0] for c in TC[sl]])
        ib = np.stack([box_nodes(g, c)[0] for c in BC[sl]])
        rt = node_coords_of_ids(g, it)
        rb = node_coords_of_ids(g, ib)
        Rn = np.sqrt(np.sum((rt[:, :, None, :] - rb[:, None, :, :]) ** 2, axis=-1))
        with np.errstate(divide='ignore', invalid='ignore'):
            G = np.exp(1j * k * Rn) / (4.0 * PI * Rn)
        G[it[:, :, None] == ib[:, None, :]] = 0.0
        sc = np.sqrt(vT * vB)
        a = np.einsum('pi,pij,pj->p', S[t], G, S[b])
        corr = eye[None] * (a * sc)[:, None, None]
        corr2 = np.einsum('pai,pij,pbj->pab', D[t], G, D[b]) * sc[:, None, None]
        alok = np.where((t == b)[:, None, None],
                        eye[None] * (1.0 / (param[t] - 1.0))[:, None, None], 0.0)
        blocks[sl] = alok + (1.0 / np.sqrt(vB * vT))[:, None, None] * (
            -k ** 2 * (eye[None] * VV[:, None, None] - corr) + (dVdV - corr2))
    b32 = blocks.astype(np.complex64)
    if return_f64:
        return b32, np.stack([T, B], axis=1), blocks
    return b32, np.stack([T, B], axis=1)


# --------------------------------------------------------------------------
# the operator and its matvec (src/gmres_module.f90:57-176)
# --------------------------------------------------------------------------

class Operator:
    pass


def build_operator(coord, etopol, param, k, M_ex=2, grid_size=1.8):
    op = Operator()
    op.g = build_geometry(coord, etopol, M_ex, grid_size)
    op.k = k
    op.param = np.asarray(param, complex)
    op.S, op.Sx, op.Sy, op.Sz, op.indS = projection_const(op.g, k)
    op.Fg = build_Fg(op.g, k)
    op.sp_mat, op.sp_ind = near_zone_const(op.g, op.param, op.S, op.Sx, op.Sy, op.Sz, k)
    return op


def _conv(op, grid_vec):
    """Zero-pad, FFT, x Fg, IFFT, crop and scale 1/(8 NxNyNz) (arr2vec)."""
    g = op.g
    A = np.zeros(op.Fg.shape, complex)
    A[:g.Nz, :g.Ny, :g.Nx] = grid_vec.reshape(g.Nz, g.Ny, g.Nx)
    A = np.fft.ifftn(op.Fg * np.fft.fftn(A))
    return A[:g.Nz, :g.Ny, :g.Nx].ravel()


def _scatter(op, val, x):
    """sparse_T_matmul (src/sparse.f90:13-43): grid[indS(j,t)] += val(j,t) x_t."""
    g = op.g
    out = np.zeros(g.Nx * g.Ny * g.Nz, complex)
    np.add.at(out, op.indS.ravel(), (val * x[:, None]).ravel())
    return out


def _gather(op, val, grid):
    """sparse_eps_matmul (src/sparse.f90:47-77): y_t = sum_j val(j,t) grid[indS(j,t)]."""
    return np.sum(val * grid[op.indS], axis=1)


def near_matvec(op, x):
    """matmul_Acorr_const (src/sparse_mat.f90:96-123), c64 blocks promoted."""
    N = op.g.N
    X = x.reshape(N, 3)
    Y = np.zeros((N, 3), complex)
    blk = op.sp_mat.astype(complex)
    T = op.sp_ind[:, 0]
    B = op.sp_ind[:, 1]
    np.add.at(Y, T, np.einsum('pij,pj->pi', blk, X[B]))
    off = T != B
    np.add.at(Y, B[off], np.einsum('pji,pj->pi', blk[off], X[T[off]]))
    return Y.ravel()


def matvec(op, x):
    """compute_Ax2 for basis order 0 with x interleaved xyz per tet."""
    N = op.g.N
    X = x.reshape(N, 3)
    Y = np.zeros((N, 3), complex)
    for c, Sc in enumerate((op.S, op.S, op.S)):
        Y[:, c] = _gather(op, op.S, _conv(op, _scatter(op, op.S, X[:, c])))
    Y2 = np.zeros((N, 3), complex)
    for c, D in enumerate((op.Sx, op.Sy, op.Sz)):
        grid = _conv(op, _scatter(op, D, X[:, c]))
        for d, E in enumerate((op.Sx, op.Sy, op.Sz)):
            Y2[:, d] += _gather(op, E, grid)
    Yaim = (Y2 - op.k ** 2 * Y).ravel()
    return near_matvec(op, x) + Yaim


def exact_pair_terms(coord, etopol, k, chunk=2048):
    """V_V and dV_dV integrals of every tet pair (T <= B), no grid terms.

    Returns (T, B, VV (P,), dVdV (P,3,3), vol (N,)).  These are the exact
    interactions that the reference evaluates for near pairs
    (src/precorrection.f90:227-246 before the precorrection is subtracted);
    here every pair is integrated, near or far.
    """
    tets = coord[etopol]
    if (tetra_det(tets) <= 0).any():
        raise ValueError('detJ <= 0')
    N = tets.shape[0]
    T, B = np.triu_indices(N)
    VV = np.zeros(T.size, complex)
    dVdV = np.zeros((T.size, 3, 3), complex)
    for s0 in range(0, T.size, chunk):
        sl = slice(s0, min(T.size, s0 + chunk))
        VV[sl] = integrate_V_V_G(tets[T[sl]], tets[B[sl]], k)
        dVdV[sl] = integrate_dV_dV_G(tets[T[sl]], tets[B[sl]], k)
    return T, B, VV, dVdV, tetra_volume(tets)


def exact_matrix(coord, etopol, param, k, terms=None, k2_sign=-1.0, self_term=True):
    """Dense Galerkin matrix of the order-0 JVIE (3N x 3N, rows 3t+c), every
    pair integrated exactly:

        A_TB = delta_TB/(eps_T - 1) I + (-k^2 V_V I + dV_dV) / sqrt(vol_T vol_B)

    the block formula of src/precorrection.f90:227-246 without the grid
    terms, i.e. the operator that the pFFT matvec (grid part + precorrected
    near zone) approximates.  A is complex-symmetric: A_BT = A_TB^T.

    k2_sign and self_term exist so that tests can plant a wrong sign or drop
    the self term and show that the Mie comparison catches it."""
    T, B, VV, dVdV, vol = terms if terms is not None else exact_pair_terms(coord, etopol, k)
    param = np.asarray(param, complex)
    N = vol.size
    eye = np.eye(3)
    blk = (k2_sign * k ** 2 * VV[:, None, None] * eye[None] + dVdV) / np.sqrt(vol[T] * vol[B])[:, None, None]
    if self_term:
        d = T == B
        blk[d] += eye[None] * (1.0 / (param[T[d]] - 1.0))[:, None, None]
    A = np.zeros((N, 3, N, 3), complex)
    A[T, :, B, :] = blk
    off = T != B
    A[B[off], :, T[off], :] = blk[off].transpose(0, 2, 1)
    return A.reshape(3 * N, 3 * N)


def tmatrix_dense(A, mat, k):
    """T blocks of a dense operator: i k^3 mat^H A^-1 mat split as compute_T_matrix
    (src/T_matrix.f90:254-275), solved directly."""
    X = np.linalg.solve(A, mat)
    Tm = 1j * k ** 3 * (np.conj(mat).T @ X)
    nm = Tm.shape[0] // 2
    return Tm[:nm, :nm], Tm[:nm, nm:], Tm[nm:, :nm], Tm[nm:, nm:]


def dense_matrix(op):
    """Dense A by unit vectors (tiny meshes only)."""
    n = 3 * op.g.N
    A = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        A[:, j] = matvec(op, e)
    return A


# --------------------------------------------------------------------------
# restarted GMRES exactly as src/gmres_module.f90:180-325
# --------------------------------------------------------------------------

def gmres(apply, rhs, tol=1e-5, maxit=50, restart=4):
    """Returns (x, inner iterations counted as the reference's `ite`, matvecs)."""
    m = maxit
    n = rhs.size
    x = rhs.copy()
    b_norm = np.sqrt(np.vdot(rhs, rhs).real)
    V = np.zeros((n, m + 1), complex)
    H = np.zeros((m + 1, m), complex)
    cs = np.zeros(m + 1, complex)
    sn = np.zeros(m + 1, complex)
    ite = 0
    nmv = 0
    error = np.inf
    for _cycle in range(restart):
        r = rhs - apply(x)
        nmv += 1
        rn = np.sqrt(np.vdot(r, r))
        V[:, 0] = r / rn
        g = np.zeros(m + 1, complex)
        g[0] = rn
        converged = False
        for i in range(m):
            w = apply(V[:, i])
            nmv += 1
            normav = np.sqrt(np.vdot(w, w)).real
            for kk in range(i + 1):
                H[kk, i] = np.vdot(V[:, kk], w)
                w = w - H[kk, i] * V[:, kk]
            H[i + 1, i] = np.sqrt(np.vdot(w, w))
            normav2 = H[i + 1, i].real
            V[:, i + 1] = w
            if normav + 0.001 * normav2 == normav:
                for j in range(i + 1):
                    hr = np.vdot(V[:, j], V[:, i + 1])
                    H[j, i] += hr
                    V[:, i + 1] -= hr * V[:, j]
                H[i + 1, i] = np.sqrt(np.vdot(V[:, i + 1], V[:, i + 1]))
            if H[i + 1, i] != 0.0:
                V[:, i + 1] = V[:, i + 1] / H[i + 1, i]
            for kk in range(i):
                tmp = cs[kk] * H[kk, i] - sn[kk] * H[kk + 1, i]
                H[kk + 1, i] = sn[kk] * H[kk, i] + np.conj(cs[kk]) * H[kk + 1, i]
                H[kk, i] = tmp
            nu = np.sqrt(np.vdot(H[i:i + 2, i], H[i:i + 2, i])).real
            if nu != 0.0:
                cs[i] = np.conj(H[i, i] / nu)
                sn[i] = -H[i + 1, i] / nu
                H[i, i] = cs[i] * H[i, i] - sn[i] * H[i + 1, i]
                H[i + 1, i] = 0.0
                t1, t2 = g[i], g[i + 1]
                g[i] = cs[i] * t1 - sn[i] * t2
                g[i + 1] = sn[i] * t1 + np.conj(cs[i]) * t2
            error = abs(g[i + 1]) / b_norm
            if error < tol:
                y = np.linalg.inv(H[:i + 1, :i + 1]) @ g[:i + 1]
                x = x + V[:, :i + 1] @ y
                converged = True
                break
            ite += 1
        if converged:
            break
        y = np.linalg.inv(H[:m, :m]) @ g[:m]
        x = x + V[:, :m] @ y
        # the reference spends one more matvec here whose result is unused
    return x, ite, nmv


# --------------------------------------------------------------------------
# regular VSWFs at tet centroids (src/T_matrix.f90:280-305, src/mie.f90:408-496)
# --------------------------------------------------------------------------

def _fact(n):
    r = 1.0
    for i in range(1, n + 1):
        r *= i
    return r


def calc_MN(Nmax, k, r):
    """Regular M_nm, N_nm (cartesian) at point r, columns ind = n(n+1)+m."""
    from scipy.special import spherical_jn, lpmv
    rr = np.sqrt(np.sum(r * r))
    theta = np.arccos(r[2] / rr) if rr != 0 else 0.0
    phi = np.arctan2(r[1], r[0])
    kr = k * rr
    j = spherical_jn(np.arange(Nmax + 2), kr)
    nm = (Nmax + 1) ** 2 - 1
    MM = np.zeros((3, nm), complex)
    NN = np.zeros((3, nm), complex)
    ct, st = np.cos(theta), np.sin(theta)
    cp, sp = np.cos(phi), np.sin(phi)
    H = np.array([[st * cp, ct * cp, -sp], [st * sp, ct * sp, cp], [ct, -st, 0.0]])
    ind = 0
    for n in range(1, Nmax + 1):
        alpha = j[n]
        beta = np.sqrt(n * (n + 1.0)) / kr * j[n]
        gamma = (n + 1.0) / kr * j[n] - j[n + 1]
        q = np.sqrt(n * (n + 1.0)) / ((2.0 * n + 1.0) * st)
        for m in range(-n, n + 1):
            mm = abs(m)
            cc = np.sqrt((2.0 * n + 1.0) * _fact(n - mm) / _fact(n + mm) / (4.0 * PI))
            e = np.exp(1j * m * phi)
            Y = lpmv(mm, n, ct) * e
            Y1 = lpmv(mm, n + 1, ct) * e * ((n - mm + 1.0) / (n + 1.0))
            Y2 = 0.0 if mm == n else lpmv(mm, n - 1, ct) * e * ((n + mm) / n)
            Bt = q * (Y1 - Y2)
            Bp = q * (1j * m * (2 * n + 1.0) / (n * (n + 1.0))) * Y
            M_sph = cc * alpha * np.array([0.0, Bp, -Bt])
            N_sph = cc * np.array([beta * Y, gamma * Bt, gamma * Bp])
            MM[:, ind] = H @ M_sph
            NN[:, ind] = H @ N_sph
            ind += 1
    return MM, NN


def vswf_matrix(coord, etopol, k, Nmax):
    """mat(3N, 2nm): rows 3t+c, columns M then N waves, times sqrt(vol)."""
    tets = coord[etopol]
    N = tets.shape[0]
    nm = (Nmax + 1) ** 2 - 1
    mat = np.zeros((3 * N, 2 * nm), complex)
    vol = tetra_volume(tets)
    for t in range(N):
        r = (tets[t, 0] + tets[t, 1] + tets[t, 2] + tets[t, 3]) / 4.0
        MM, NN = calc_MN(Nmax, k, r)
        sv = np.sqrt(vol[t])
        mat[3 * t:3 * t + 3, :nm] = MM * sv
        mat[3 * t:3 * t + 3, nm:] = NN * sv
    return mat


def tmatrix(op, Nmax, coord, etopol, tol=1e-5, maxit=50, restart=4, cols=None):
    """compute_T_matrix: returns Taa, Tab, Tba, Tbb (nm x nm) and iteration counts."""
    mat = vswf_matrix(coord, etopol, op.k, Nmax)
    n2 = mat.shape[1]
    cols = range(n2) if cols is None else cols
    Tm = np.zeros((n2, n2), complex)
    iters = []
    for c in cols:
        x, ite, _ = gmres(lambda v: matvec(op, v), mat[:, c].copy(), tol, maxit, restart)
        Tm[:, c] = np.conj(mat).T @ x
        iters.append(ite)
    nm = n2 // 2
    sc = 1j * op.k ** 3.0
    return (Tm[:nm, :nm] * sc, Tm[:nm, nm:] * sc, Tm[nm:, :nm] * sc, Tm[nm:, nm:] * sc), iters
