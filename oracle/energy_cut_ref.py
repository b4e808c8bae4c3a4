"""Test infrastructure, not product code: a CPU restatement of the reference's
EnergyCut Hall conductivity in two and three dimensions, the checker for the
library's rtb_hall_energy_cut.  Pure-Python loops: small meshes only.

Followed step by step:
  hall_conductivity, Integration::EnergyCut   src/response/linear/mod.rs:177-185, 226-268
  compute_velocity_kernel                     src/response/primitives.rs:33-137
  build_overlap_matrix, greedy_assign,
  global_band_track, neighbour_indices        src/response/tracking.rs:25-225
  build_triangles_2d_diagavg                  src/response/tracking.rs:284-342
  eval_berry_kernel, eval_berry_band_at_lam   src/response/kernel.rs:54-80, 138-182
  TRI_QUAD_PTS_3 / TRI_QUAD_WTS_3             src/response/quadrature.rs
  triangle_area, clip_triangle, combine_bary,
  triangle_occupied_hybrid,
  integrate_fermi_cut_2d(_t0)                 src/response/energy_cut.rs:59-78, 1102-1409
  3D: build_tetrahedra_3d (CUBE_TETS,
  TET_VOL_FACTOR)                             src/response/tracking.rs:8-23, 344-412
  TET_QUAD_PTS_4 / TET_QUAD_WTS_4             src/response/quadrature.rs
  tet_vol_from_pts, bary_to_phys_3d,
  sub_tet_vol_3d, combine_bary_4,
  sub_tet_k_quad, tetrahedron_occupied_hybrid,
  integrate_fermi_cut_3d(_t0)                 src/response/energy_cut.rs:1446-1850
ndarray's Array1::linspace(a, b, n) is a + i (b - a) / (n - 1).

Pinned (tests/test_energy_cut.py) by: the clipped-triangle area against a
brute-force count of the linearly interpolated occupied region; the band sum
rule (sigma = 0 with every band occupied); the thermal convolution of a
constant and of a linear sigma_0; band tracking recovering vertex data whose
band order was shuffled; greedy assignment of permutation matrices; and the
reference's own checks (EnergyCut within 3e-2 of Direct on the Haldane
model, time reversal sigma(t2) = -sigma(-t2), lib.rs:2640-2760) plus the
Chern plateau in the gap.  3D: the occupied volume of the tetrahedron cut
against a brute-force count, the band sum rule, time reversal on a stacked
Haldane model, and the reference's stacked-QWZ check (plateau 1/(2 pi) within
3e-3 and Direct within 5e-2, lib.rs:3196-3270).
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np

from . import oracle as O

ENERGY_CUT_EPS = 1e-12  # energy_cut.rs:47
FERMI_X_CUT = 18.0  # energy_cut.rs:48
FERMI_X_STEPS = 72  # energy_cut.rs:49
TRI_QUAD_PTS_3 = ((2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0), (1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0),
                  (1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0))
TRI_QUAD_WTS_3 = (1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0)


class Vertex:
    """VertexKernel (types.rs:35-60) without the dir_c fields."""

    def __init__(self, band, evec, k_ab):
        self.band, self.evec, self.k_ab = band, evec, k_ab


def velocity_kernel(model, k, dir_a, dir_b, spin=None) -> Vertex:
    """compute_velocity_kernel (primitives.rs:33-137): eigh of the Atom-gauge
    H, v in the band basis as evec^T v evec^* (primitives.rs:88-104, the spin
    current on dir_a only), and K_nm = v^a_nm v^b_mn.  The reference's evec
    is conj(U) for the true eigenvectors U (see oracle.eigh), so evec^T v
    evec^* = U^H v U; evec is kept as the reference holds it."""
    v, H = O.gen_v_projected(model, np.asarray(k, float), "atom", np.array([dir_a, dir_b], float))
    E, U = O.eigh(H)
    evec = U.conj()
    ut, uc = evec.T, evec.conj()
    va_raw = v[0]
    if model.spin and spin is not None:
        S = O.spin_matrix(model.norb, spin)
        va_raw = (S @ va_raw + va_raw @ S) * 0.5
    va = ut @ (va_raw @ uc)
    vb = ut @ (v[1] @ uc)
    return Vertex(E, evec, va * vb.T)


def build_overlap_matrix(evec_ref, evec_other):
    """tracking.rs:25-42: ov[n, m] = |sum_orb conj(ref[orb, n]) other[orb, m]|^2."""
    s = evec_ref.conj().T @ evec_other
    return s.real * s.real + s.imag * s.imag


def greedy_assign(ov):
    """tracking.rs:44-68: rows in descending order of their best overlap
    (stable), each takes its largest still-free column."""
    n = ov.shape[0]
    rows = [(i, max([-1.0] + [float(ov[i, j]) for j in range(n)])) for i in range(n)]
    rows.sort(key=lambda r: -r[1])
    assigned = [False] * n
    perm = [0] * n
    for i, _ in rows:
        best_j, best_val = 0, -1.0
        for j in range(n):
            if not assigned[j] and ov[i, j] > best_val:
                best_val, best_j = float(ov[i, j]), j
        assigned[best_j] = True
        perm[i] = best_j
    return perm


def permute_vertex(v: Vertex, p) -> Vertex:
    """tracking.rs:70-121: band n of the result is band p[n] of v."""
    p = np.asarray(p)
    return Vertex(v.band[p], v.evec[:, p], v.k_ab[np.ix_(p, p)])


def neighbour_indices(i, k_mesh):
    """tracking.rs:165-224 (2D, 3D): non-periodic neighbours in the order
    +x, -x, +y, -y (+z, -z)."""
    if len(k_mesh) == 2:
        nx, ny = k_mesh
        ix, iy = i // ny, i % ny
        out = []
        for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
            jx, jy = ix + dx, iy + dy
            if 0 <= jx < nx and 0 <= jy < ny:
                out.append(jx * ny + jy)
        return out
    nx, ny, nz = k_mesh
    ix, iy, iz = i // (ny * nz), (i % (ny * nz)) // nz, i % nz
    out = []
    for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        jx, jy, jz = ix + dx, iy + dy, iz + dz
        if 0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz:
            out.append(jx * ny * nz + jy * nz + jz)
    return out


def global_band_track(pts, k_mesh):
    """tracking.rs:123-163: breadth-first from vertex 0; each new vertex is
    permuted by greedy_assign of its overlaps with its visited neighbours,
    summed then divided by their count."""
    nk = len(pts)
    if nk <= 1:
        return
    visited = [False] * nk
    visited[0] = True
    queue = deque([0])
    while queue:
        cur = queue.popleft()
        for nb in neighbour_indices(cur, k_mesh):
            if visited[nb]:
                continue
            ov_sum, cnt = None, 0
            for vn in neighbour_indices(nb, k_mesh):
                if not visited[vn]:
                    continue
                ov = build_overlap_matrix(pts[vn].evec, pts[nb].evec)
                ov_sum = ov if ov_sum is None else ov_sum + ov
                cnt += 1
            if ov_sum is not None:
                pts[nb] = permute_vertex(pts[nb], greedy_assign(ov_sum / float(cnt)))
            visited[nb] = True
            queue.append(nb)


def eval_berry_kernel(band, k_ab, eta):
    """kernel.rs:54-80: Omega_n = -2 Im sum_{m != n} K_nm / (dE^2 + eta^2)."""
    n = band.shape[0]
    eta2 = eta * eta
    out = np.zeros(n)
    for i in range(n):
        re = im = 0.0
        for m in range(n):
            if m == i:
                continue
            de = band[i] - band[m]
            denom = de * de + eta2
            if denom < 1e-30:
                continue
            re += k_ab[i, m].real / denom
            im += k_ab[i, m].imag / denom
        out[i] = -2.0 * im
    return out


def eval_berry_band_at_lam(n, bands, kmats, lam, eta):
    """kernel.rs:138-182: E and row n of K interpolated linearly at the
    barycentric point lam, then the band sum for band n."""
    nsta = bands[0].shape[0]
    e = [0.0] * nsta
    kr = [0.0] * nsta
    ki = [0.0] * nsta
    for v in range(len(bands)):
        lv = lam[v]
        if lv == 0.0:
            continue
        for m in range(nsta):
            e[m] += float(bands[v][m]) * lv
    for v in range(len(kmats)):
        lv = lam[v]
        if lv == 0.0:
            continue
        for m in range(nsta):
            z = kmats[v][n, m]
            kr[m] += z.real * lv
            ki[m] += z.imag * lv
    eta2 = eta * eta
    im = 0.0
    for m in range(nsta):
        if m == n:
            continue
        de = e[n] - e[m]
        denom = de * de + eta2
        if denom < 1e-30:
            continue
        im += ki[m] / denom
    return -2.0 * im


def triangle_area(c):
    """energy_cut.rs:69-78."""
    return 0.5 * abs((c[1][0] - c[0][0]) * (c[2][1] - c[0][1]) - (c[2][0] - c[0][0]) * (c[1][1] - c[0][1]))


def _bary_to_phys(c, lam):
    return [lam[0] * c[0][d] + lam[1] * c[1][d] + lam[2] * c[2][d] for d in (0, 1)]


def sub_tri_area(c, tri):
    """energy_cut.rs:1110-1116."""
    p0, p1, p2 = (_bary_to_phys(c, t) for t in tri)
    return 0.5 * abs((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p2[0] - p0[0]) * (p1[1] - p0[1]))


def clip_triangle(e, mu):
    """energy_cut.rs:1120-1169: the region E <= mu of the linearly
    interpolated triangle as 0-2 sub-triangles in barycentric coordinates."""
    eps = 1e-12
    e_min, e_max = min(e), max(e)
    if mu <= e_min + eps:
        return []
    if mu >= e_max - eps:
        return [((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))]
    a, b, c = sorted(range(3), key=lambda i: e[i])  # stable, as sort_by

    def unit(i):
        lam = [0.0, 0.0, 0.0]
        lam[i] = 1.0
        return tuple(lam)

    def cut(i, j):
        de = e[j] - e[i]
        t = 0.5 if abs(de) < eps else min(max((mu - e[i]) / de, 0.0), 1.0)
        lam = [0.0, 0.0, 0.0]
        lam[i] = 1.0 - t
        lam[j] = t
        return tuple(lam)

    if mu <= e[b] + eps:
        return [(unit(a), cut(a, b), cut(a, c))]
    return [(unit(a), unit(b), cut(a, c)), (unit(b), cut(a, c), cut(b, c))]


def combine_bary(alpha, l0, l1, l2):
    """energy_cut.rs:1172-1178."""
    return [alpha[0] * l0[d] + alpha[1] * l1[d] + alpha[2] * l2[d] for d in range(3)]


def triangle_occupied_hybrid(c, e_v, omega_v, bands, kmats, mu, eta, n):
    """energy_cut.rs:1184-1228: 0 when empty, area x vertex-mean Omega when
    full, else the 3-point rule on each clipped sub-triangle."""
    eps = ENERGY_CUT_EPS
    e_min, e_max = min(e_v), max(e_v)
    area = triangle_area(c)
    if mu <= e_min + eps:
        return 0.0
    if mu >= e_max - eps:
        return area * (omega_v[0] + omega_v[1] + omega_v[2]) / 3.0
    total = 0.0
    for tri in clip_triangle(e_v, mu):
        sa = sub_tri_area(c, tri)
        if sa < 1e-30:
            continue
        for iq in range(3):
            lam = combine_bary(TRI_QUAD_PTS_3[iq], *tri)
            total += sa * TRI_QUAD_WTS_3[iq] * eval_berry_band_at_lam(n, bands, kmats, lam, eta)
    return total


def _partition_point(xs, pred):
    lo = 0
    while lo < len(xs) and pred(xs[lo]):
        lo += 1
    return lo


def integrate_fermi_cut_2d_t0(pts, k_mesh, mu, eta):
    """energy_cut.rs:1231-1345 over the four triangles of each cell
    (both diagonals, tracking.rs:284-342), fractional-coordinate measure."""
    nx, ny = k_mesh
    inv_nx, inv_ny = 1.0 / nx, 1.0 / ny
    n_mu = len(mu)
    acc = np.zeros(n_mu)
    half_vol = inv_nx * inv_ny / 4.0
    for idx in range(nx * ny):
        iy, ix = idx % ny, idx // ny
        ixp, iyp = (ix + 1) % nx, (iy + 1) % ny
        i00, i10, i11, i01 = ix * ny + iy, ixp * ny + iy, ixp * ny + iyp, ix * ny + iyp

        def coord_of(i):
            if i == i00:
                return (ix * inv_nx, iy * inv_ny)
            if i == i10:
                return ((ix + 1) * inv_nx, iy * inv_ny)
            if i == i11:
                return ((ix + 1) * inv_nx, (iy + 1) * inv_ny)
            return (ix * inv_nx, (iy + 1) * inv_ny)

        for tri in ((i00, i10, i01), (i11, i10, i01), (i00, i10, i11), (i00, i11, i01)):
            c = [coord_of(v) for v in tri]
            area = triangle_area(c)
            if area < ENERGY_CUT_EPS:
                continue
            scale = half_vol / area
            vs = [pts[v] for v in tri]
            nsta = vs[0].band.shape[0]
            om = [eval_berry_kernel(v.band, v.k_ab, eta) for v in vs]
            bands = [v.band for v in vs]
            kmats = [v.k_ab for v in vs]
            for n in range(nsta):
                e_v = [float(v.band[n]) for v in vs]
                omega_v = [float(o[n]) for o in om]
                e_min, e_max = min(e_v), max(e_v)
                full_val = area * (omega_v[0] + omega_v[1] + omega_v[2]) / 3.0
                i_partial = _partition_point(mu, lambda x: x <= e_min + ENERGY_CUT_EPS)
                i_full = _partition_point(mu, lambda x: x < e_max - ENERGY_CUT_EPS)
                for im in range(i_partial, i_full):
                    acc[im] += scale * triangle_occupied_hybrid(c, e_v, omega_v, bands, kmats, mu[im], eta, n)
                if i_full < n_mu:
                    add = scale * full_val
                    for im in range(i_full, n_mu):
                        acc[im] += add
    return acc


def fermi_window_x(x):
    """energy_cut.rs:59-66."""
    if abs(x) > 50.0:
        return 0.0
    ex = math.exp(x)
    return ex / ((1.0 + ex) * (1.0 + ex))


def thermal_convolution(mu, width, sigma0_of):
    """energy_cut.rs:1372-1408: sigma_0 on a fine grid (step width / 10,
    +-12 widths beyond the mu range), linearly interpolated and integrated
    against the Fermi window with the 72-point midpoint rule on |x| <= 18.
    Kept as the reference has it: the grid is a linspace of n_ext points,
    whose step is a little below the dmu_fine the index i_f divides by."""
    beta = 1.0 / width
    x_max = 12.0
    mu_lo = min(mu) - x_max / beta
    mu_hi = max(mu) + x_max / beta
    dmu_fine = 0.1 / beta
    n_ext = int(math.ceil((mu_hi - mu_lo) / dmu_fine)) + 1
    step = (mu_hi - mu_lo) / (n_ext - 1) if n_ext > 1 else 0.0
    mu_ext = [mu_lo + step * i for i in range(n_ext)]
    s0 = sigma0_of(mu_ext)
    dx = 2.0 * FERMI_X_CUT / FERMI_X_STEPS
    out = []
    for m in mu:
        s = 0.0
        for iq in range(FERMI_X_STEPS):
            x = -FERMI_X_CUT + (iq + 0.5) * dx
            w = fermi_window_x(x)
            i_f = (m + x / beta - mu_lo) / dmu_fine
            i_lo = max(int(math.floor(i_f)), 0)
            i_hi = min(i_lo + 1, n_ext - 1)
            if i_hi > i_lo:
                t = i_f - i_lo
                s += dx * w * (s0[i_lo] + t * (s0[i_hi] - s0[i_lo]))
        out.append(s)
    return np.array(out)


def integrate_fermi_cut_2d(pts, k_mesh, mu, width, eta):
    """energy_cut.rs:1355-1409."""
    mu = [float(x) for x in mu]
    if width == 0.0:
        return integrate_fermi_cut_2d_t0(pts, k_mesh, mu, eta)
    return thermal_convolution(mu, width, lambda m: integrate_fermi_cut_2d_t0(pts, k_mesh, m, eta))


# ------------------------------------------------------------------ 3D
CUBE_TETS = ((0, 1, 2, 4), (3, 1, 2, 7), (5, 1, 4, 7), (6, 2, 4, 7), (1, 2, 4, 7))  # tracking.rs:8-14
TET_VOL_FACTOR = (1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0)  # tracking.rs:23
_QA, _QB = 0.5854101966249685, 0.1381966011250105
TET_QUAD_PTS_4 = ((_QA, _QB, _QB, _QB), (_QB, _QA, _QB, _QB), (_QB, _QB, _QA, _QB), (_QB, _QB, _QB, _QA))
TET_QUAD_WTS_4 = (0.25, 0.25, 0.25, 0.25)


def tet_vol_from_pts(p0, p1, p2, p3):
    """energy_cut.rs:1448-1458: |det| / 6 of the edge vectors, expanded along
    the first row as the reference writes it."""
    x0, y0, z0 = p0
    x1, y1, z1 = p1
    x2, y2, z2 = p2
    x3, y3, z3 = p3
    det = ((x1 - x0) * ((y2 - y0) * (z3 - z0) - (z2 - z0) * (y3 - y0))
           - (y1 - y0) * ((x2 - x0) * (z3 - z0) - (z2 - z0) * (x3 - x0))
           + (z1 - z0) * ((x2 - x0) * (y3 - y0) - (y2 - y0) * (x3 - x0)))
    return abs(det) / 6.0


def bary_to_phys_3d(c, lam):
    """energy_cut.rs:1461-1478."""
    return [lam[0] * c[0][d] + lam[1] * c[1][d] + lam[2] * c[2][d] + lam[3] * c[3][d] for d in range(3)]


def sub_tet_vol_3d(c, sub):
    """energy_cut.rs:1481-1488."""
    return tet_vol_from_pts(*[bary_to_phys_3d(c, l) for l in sub])


def combine_bary_4(alpha, lam):
    """energy_cut.rs:1492-1500: sum_k alpha_k lam_k."""
    out = [0.0] * 4
    for k in range(4):
        for i in range(4):
            out[i] += alpha[k] * lam[k][i]
    return out


def sub_tet_k_quad(sub, bands, kmats, n, eta, c):
    """energy_cut.rs:1503-1526: the 4-point rule on one sub-tetrahedron."""
    sv = sub_tet_vol_3d(c, sub)
    if sv < 1e-30:
        return 0.0
    total = 0.0
    for iq in range(4):
        lam = combine_bary_4(TET_QUAD_PTS_4[iq], sub)
        total += sv * TET_QUAD_WTS_4[iq] * eval_berry_band_at_lam(n, bands, kmats, lam, eta)
    return total


def tet_cut(e, mu):
    """The occupied part E <= mu of a linearly interpolated tetrahedron as the
    sub-tetrahedra (barycentric rows) of tetrahedron_occupied_hybrid
    (energy_cut.rs:1529-1640), and the sign each enters with (the 3-below case
    is the full tetrahedron minus one); None when mu is outside (e_min, e_max)
    within the cut's eps."""
    eps = ENERGY_CUT_EPS
    if mu <= min(e) + eps or mu >= max(e) - eps:
        return None
    a, b, c, d = sorted(range(4), key=lambda i: e[i])

    def unit(i):
        l = [0.0] * 4
        l[i] = 1.0
        return l

    def cut(i, j):
        de = e[j] - e[i]
        t = 0.5 if abs(de) < eps else min(max((mu - e[i]) / de, 0.0), 1.0)
        l = [0.0] * 4
        l[i] = 1.0 - t
        l[j] = t
        return l

    if mu <= e[b] + eps:
        return [(1.0, [unit(a), cut(a, b), cut(a, c), cut(a, d)])]
    if mu <= e[c] + eps:
        return [(1.0, [unit(a), unit(b), cut(a, c), cut(b, d)]),
                (1.0, [unit(a), cut(a, c), cut(a, d), cut(b, d)]),
                (1.0, [unit(b), cut(b, c), cut(b, d), cut(a, c)])]
    return [(-1.0, [unit(d), cut(a, d), cut(b, d), cut(c, d)])]


def tetrahedron_occupied_hybrid(c, e_v, omega_v, bands, kmats, mu, eta, n):
    """energy_cut.rs:1529-1640: vertex-average Omega for a full tetrahedron,
    the 4-point rule on the sub-tetrahedra of a partial one."""
    full_vol = tet_vol_from_pts(*c)
    if full_vol < ENERGY_CUT_EPS:
        return 0.0
    if mu <= min(e_v) + ENERGY_CUT_EPS:
        return 0.0
    full_val = full_vol * (omega_v[0] + omega_v[1] + omega_v[2] + omega_v[3]) / 4.0
    if mu >= max(e_v) - ENERGY_CUT_EPS:
        return full_val
    subs = tet_cut(e_v, mu)
    if subs[0][0] < 0:
        return full_val - sub_tet_k_quad(subs[0][1], bands, kmats, n, eta, c)
    total = 0.0
    for _, sub in subs:
        total += sub_tet_k_quad(sub, bands, kmats, n, eta, c)
    return total


def build_tetrahedra_3d(ix, iy, iz, nx, ny, nz):
    """tracking.rs:344-412: the five tetrahedra of cell (ix, iy, iz) on the
    0-7 body diagonal, as (vertex indices, corner coordinates, volume)."""
    ixp, iyp, izp = (ix + 1) % nx, (iy + 1) % ny, (iz + 1) % nz

    def idx3(x, y, z):
        return x * ny * nz + y * nz + z

    cidx = [idx3(ix, iy, iz), idx3(ixp, iy, iz), idx3(ix, iyp, iz), idx3(ixp, iyp, iz),
            idx3(ix, iy, izp), idx3(ixp, iy, izp), idx3(ix, iyp, izp), idx3(ixp, iyp, izp)]

    def frac(x, y, z):
        return (x / nx, y / ny, z / nz)

    corners = [frac(ix, iy, iz), frac(ix + 1, iy, iz), frac(ix, iy + 1, iz), frac(ix + 1, iy + 1, iz),
               frac(ix, iy, iz + 1), frac(ix + 1, iy, iz + 1), frac(ix, iy + 1, iz + 1),
               frac(ix + 1, iy + 1, iz + 1)]
    cube_vol = (1.0 / nx) * (1.0 / ny) * (1.0 / nz)
    return [([cidx[l] for l in t], [corners[l] for l in t], cube_vol * TET_VOL_FACTOR[i])
            for i, t in enumerate(CUBE_TETS)]


def integrate_fermi_cut_3d_t0(pts, k_mesh, mu, eta):
    """energy_cut.rs:1650-1797, fractional-coordinate measure."""
    nx, ny, nz = k_mesh
    n_mu = len(mu)
    acc = np.zeros(n_mu)
    for idx in range(nx * ny * nz):
        iz, iy, ix = idx % nz, (idx // nz) % ny, idx // (ny * nz)
        for vidx, c, volume in build_tetrahedra_3d(ix, iy, iz, nx, ny, nz):
            vol = tet_vol_from_pts(*c)
            if vol < ENERGY_CUT_EPS:
                continue
            scale = volume / vol
            vs = [pts[v] for v in vidx]
            nsta = vs[0].band.shape[0]
            om = [eval_berry_kernel(v.band, v.k_ab, eta) for v in vs]
            bands = [v.band for v in vs]
            kmats = [v.k_ab for v in vs]
            for n in range(nsta):
                e_v = [float(v.band[n]) for v in vs]
                omega_v = [float(o[n]) for o in om]
                e_min, e_max = min(e_v), max(e_v)
                full_val = vol * (omega_v[0] + omega_v[1] + omega_v[2] + omega_v[3]) / 4.0
                i_partial = _partition_point(mu, lambda x: x <= e_min + ENERGY_CUT_EPS)
                i_full = _partition_point(mu, lambda x: x < e_max - ENERGY_CUT_EPS)
                for im in range(i_partial, i_full):
                    acc[im] += scale * tetrahedron_occupied_hybrid(c, e_v, omega_v, bands, kmats, mu[im], eta, n)
                if i_full < n_mu:
                    add = scale * full_val
                    for im in range(i_full, n_mu):
                        acc[im] += add
    return acc


def integrate_fermi_cut_3d(pts, k_mesh, mu, width, eta):
    """energy_cut.rs:1798-1848 (the thermal convolution as in 2D)."""
    mu = [float(x) for x in mu]
    if width == 0.0:
        return integrate_fermi_cut_3d_t0(pts, k_mesh, mu, eta)
    return thermal_convolution(mu, width, lambda m: integrate_fermi_cut_3d_t0(pts, k_mesh, m, eta))


def vertices(model, kmesh, direction, spin=None):
    """The mesh's vertex kernels in gen_kmesh order (last axis fastest)."""
    direction = np.asarray(direction, float)
    return [velocity_kernel(model, k, direction[0], direction[1], spin) for k in O.gen_kmesh(kmesh)]


def hall_conductivity_energy_cut(model, kmesh, direction, mu, T=0.0, eta=1e-3, spin=None, pts=None):
    """hall_conductivity with Integration::EnergyCut, 2D and 3D
    (linear/mod.rs:177-185, 226-268): vertex kernels, band tracking, Fermi
    cut, divided by det(lat)."""
    direction = O.validate_direction(direction, 2, model.dim)
    mu = np.atleast_1d(np.asarray(mu, float))
    if any(mu[i] > mu[i + 1] for i in range(mu.size - 1)):  # config.rs:291-303
        raise O.OracleError("InvalidResponseParameter", "mu: must be sorted in ascending order")
    if model.dim not in (2, 3):
        raise O.OracleError("InvalidDimension", str(model.dim))
    if pts is None:
        pts = vertices(model, kmesh, direction, spin)
        global_band_track(pts, list(kmesh))
    width = 0.0 if T <= 0.0 else T * O.BOLTZMANN_CONSTANT_EV_PER_K
    cut = integrate_fermi_cut_2d if model.dim == 2 else integrate_fermi_cut_3d
    return np.asarray(cut(pts, tuple(kmesh), mu, width, eta)) / np.linalg.det(model.lat)
