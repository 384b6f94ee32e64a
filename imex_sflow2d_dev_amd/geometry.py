"""Host-side static fields of the path, computed once per run (numpy).

These restate the reference's setup routines with the same floating-point
operation order, so the fields handed to the device are the ones the
reference builds:
  - bed slopes B_prime_{x,y}_geom: topography_reconstruction,
    src/geometry_2d.f90:555-895 (5-point stencil :748-790, edge copies
    :794-849);
  - cell fractions of a disc: compute_cell_fract, src/geometry_2d.f90:1899-1994;
  - the lateral radial source ring: init_source, src/geometry_2d.f90:907-1323;
  - the source thickness from the mass flow rate: src/inpout_2d.f90:4108-4125,
    :4522-4531, :4732-4747.
"""
from __future__ import annotations

import math

import numpy as np

from .config import RunConfig
from .workloads import cell_centres


def bed_slopes(cfg: RunConfig, B: np.ndarray):
    """B_prime_x_geom, B_prime_y_geom as [k, j] arrays."""
    ny, nx = B.shape
    B = np.asarray(B, dtype=np.float64)
    c1 = (-2.0, -1.0, 0.0, 1.0, 2.0)
    norm1 = 10.0 * cfg.dx
    bx = np.zeros((ny, nx))
    by = np.zeros((ny, nx))
    # interior j, k in 3..n-2 (1-based)
    sx = np.zeros((ny - 4, nx - 4))
    sy = np.zeros((ny - 4, nx - 4))
    for kj in range(5):
        sx = sx + c1[kj] * B[2:ny - 2, kj:nx - 4 + kj]
    for kk in range(5):
        sy = sy + c1[kk] * B[kk:ny - 4 + kk, 2:nx - 2]
    bx[2:ny - 2, 2:nx - 2] = sx / norm1
    by[2:ny - 2, 2:nx - 2] = sy / norm1
    for a in (bx, by):
        a[:, 0] = a[:, 2]
        a[:, 1] = a[:, 2]
        a[:, nx - 2] = a[:, nx - 3]
        a[:, nx - 1] = a[:, nx - 3]
        a[0, :] = a[2, :]
        a[1, :] = a[2, :]
        a[ny - 2, :] = a[ny - 3, :]
        a[ny - 1, :] = a[ny - 3, :]
    return bx, by


def compute_cell_fract(cfg: RunConfig, xs: float, ys: float, rs: float, r2s: float,
                       angle: float = 0.0) -> np.ndarray:
    """Fraction of each cell inside an ellipse, by a 200 x 200 sub-grid count
    rounded in single precision as the reference does ([k, j] array)."""
    x_comp, y_comp, x_stag, y_stag = cell_centres(cfg)
    nx, ny = cfg.nx, cfg.ny
    n = 200
    h = np.arange(1, n + 1, dtype=np.float64)
    sub = (2.0 * h - 1.0) / (2.0 * float(n))
    xsub = (sub - 0.5) * cfg.dx   # x index fastest inside each row of the sub-grid
    ysub = (sub - 0.5) * cfg.dx
    ang = angle / 180.0 * math.atan(1.0) * 4.0
    ca, sa = math.cos(ang), math.sin(ang)
    rmax = max(rs, r2s)
    out = np.zeros((ny, nx))
    jlist = [j for j in range(nx) if not (x_stag[j + 1] < xs - rmax or x_stag[j] > xs + rmax)]
    klist = [k for k in range(ny) if not (y_stag[k + 1] < ys - rmax or y_stag[k] > ys + rmax)]
    XS, YS = np.meshgrid(xsub, ysub)

    def level(X, Y):
        A = X * ca + Y * sa
        Bq = -(X * sa) + Y * ca
        return (A * A) / (rs * rs) + (Bq * Bq) / (r2s * r2s)

    full = float(np.float32(n * n) / np.float32(n * n))
    for k in klist:
        for j in jlist:
            # the ellipse is convex: a cell whose four corner sub-points are
            # well inside has every sub-point inside; one whose sub-rectangle
            # keeps well clear of the circumscribed circle has none (margins
            # of 1e-9 keep the shortcut on the count's side of every rounding)
            cx = np.array([x_comp[j] + xsub[0], x_comp[j] + xsub[-1]]) - xs
            cy = np.array([y_comp[k] + ysub[0], y_comp[k] + ysub[-1]]) - ys
            if (level(cx[:, None], cy[None, :]) <= 1.0 - 1e-9).all():
                out[k, j] = full
                continue
            nxp = min(max(0.0, cx[0]), cx[1]) if cx[0] <= 0.0 <= cx[1] else min(abs(cx[0]), abs(cx[1]))
            nyp = min(max(0.0, cy[0]), cy[1]) if cy[0] <= 0.0 <= cy[1] else min(abs(cy[0]), abs(cy[1]))
            if nxp * nxp + nyp * nyp >= (rmax * rmax) * (1.0 + 1e-9):
                continue
            X = (x_comp[j] + XS) - xs
            Y = (y_comp[k] + YS) - ys
            inside = level(X, Y) <= 1.0
            cnt = int(np.count_nonzero(inside))
            out[k, j] = float(np.float32(cnt) / np.float32(n * n))
    return out


def init_source(cfg: RunConfig):
    """Radial lateral source: source_cell, side bits (E=1, W=2, N=4, S=8),
    cell_arc_perim, cell_arc_n_x, cell_arc_n_y ([k, j] arrays)."""
    x_comp, y_comp, x_stag, y_stag = cell_centres(cfg)
    nx, ny, dx, dy = cfg.nx, cfg.ny, cfg.dx, cfg.dx
    xs, ys, r = cfg.x_source, cfg.y_source, cfg.r_source
    scell = np.zeros((ny, nx), np.int32)
    sides = np.zeros((ny, nx), np.int32)
    vect = {}   # (bit, k, j) -> (vx, vy)
    r2 = r * r
    jlo = max(1, int((xs - r - x_comp[0]) / dx) - 2)
    jhi = min(nx - 2, int((xs + r - x_comp[0]) / dx) + 3)
    klo = max(1, int((ys - r - y_comp[0]) / dy) - 2)
    khi = min(ny - 2, int((ys + r - y_comp[0]) / dy) + 3)

    def sq(a):
        return a * a

    for k in range(klo, khi + 1):
        for j in range(jlo, jhi + 1):
            if sq(x_comp[j] - xs) + sq(y_comp[k] - ys) <= r2:
                scell[k, j] = 1
                if sq(x_comp[j - 1] - xs) + sq(y_comp[k] - ys) >= r2:
                    scell[k, j - 1] = 2
                    sides[k, j - 1] |= 1
                    d = math.sqrt(sq(x_stag[j] - xs) + sq(y_comp[k] - ys))
                    vect[(1, k, j - 1)] = ((x_stag[j] - xs) * r / (d * d),
                                           (y_comp[k] - ys) * r / (d * d))
                elif sq(x_comp[j + 1] - xs) + sq(y_comp[k] - ys) >= r2:
                    scell[k, j + 1] = 2
                    sides[k, j + 1] |= 2
                    d = math.sqrt(sq(x_stag[j + 1] - xs) + sq(y_comp[k] - ys))
                    vect[(2, k, j + 1)] = ((x_stag[j + 1] - xs) * r / (d * d),
                                           (y_comp[k] - ys) * r / (d * d))
                if sq(x_comp[j] - xs) + sq(y_comp[k - 1] - ys) >= r2:
                    scell[k - 1, j] = 2
                    sides[k - 1, j] |= 4
                    d = math.sqrt(sq(x_comp[j] - xs) + sq(y_stag[k] - ys))
                    vect[(4, k - 1, j)] = ((x_comp[j] - xs) * r / (d * d),
                                           (y_stag[k] - ys) * r / (d * d))
                elif sq(x_comp[j] - xs) + sq(y_comp[k + 1] - ys) >= r2:
                    scell[k + 1, j] = 2
                    sides[k + 1, j] |= 8
                    d = math.sqrt(sq(x_comp[j] - xs) + sq(y_stag[k + 1] - ys))
                    vect[(8, k + 1, j)] = ((x_comp[j] - xs) * r / (d * d),
                                           (y_stag[k + 1] - ys) * r / (d * d))
    perim = np.zeros((ny, nx))
    anx = np.zeros((ny, nx))
    any_ = np.zeros((ny, nx))
    ks, js = np.nonzero(scell == 2)
    for k, j in sorted(zip(ks.tolist(), js.tolist())):
        for bit, face_len in ((1, dy), (2, dy), (4, dx), (8, dx)):
            if sides[k, j] & bit:
                vx, vy = vect[(bit, k, j)]
                vmag = math.sqrt(vx * vx + vy * vy)
                if vmag > 0.0:
                    nxv = vx / vmag
                    nyv = vy / vmag
                    perim[k, j] = perim[k, j] + face_len
                    anx[k, j] = anx[k, j] + face_len * nxv
                    any_[k, j] = any_[k, j] + face_len * nyv
        if perim[k, j] > 0.0:
            anx[k, j] = anx[k, j] / perim[k, j]
            any_[k, j] = any_[k, j] / perim[k, j]
            vmag = math.sqrt(anx[k, j] * anx[k, j] + any_[k, j] * any_[k, j])
            if vmag > 0.0:
                anx[k, j] = anx[k, j] / vmag
                any_[k, j] = any_[k, j] / vmag
    total = float(np.sum(perim))   # sum of multiples of dx: exact in any order
    pi_g = 4.0 * math.atan(1.0)
    h_ell = (r - r) ** 2 / (r + r) ** 2
    length = pi_g * (r + r) * (1.0 + 0.25 * h_ell + h_ell ** 2 / 64.0 + h_ell ** 3 / 256.0
                               + h_ell ** 4 * 25.0 / 16384.0 + h_ell ** 5 * 49.0 / 65536.0)
    if total > 0.0:
        perim = perim * (length / total)
    return scell, sides, perim, anx, any_


def source_thickness(cfg: RunConfig) -> float:
    """h_source from MFR_SOURCE and VEL_SOURCE (RI_SOURCE not given):
    mixt_var of the unit-thickness source state, then MFR/(L*rho_m*vel)."""
    from .init import mixt_var_rho_m
    r = cfg.r_source
    pi_g = math.atan(1.0) * 4.0
    h_ell = (r - r) ** 2 / (r + r) ** 2
    length = pi_g * (r + r) * (1.0 + 0.25 * h_ell + h_ell ** 2 / 64.0 + h_ell ** 3 / 256.0
                               + h_ell ** 4 * 25.0 / 16384.0 + h_ell ** 5 * 49.0 / 65536.0)
    nv = cfg.n_vars
    qp = np.zeros(nv + 2)
    qp[0] = 1.0
    qp[1] = 1.0 * cfg.vel_source
    qp[3] = cfg.T_source
    for i in range(cfg.n_solid):
        qp[4 + i] = cfg.alphas_source[i] * 1.0
    qp[nv] = cfg.vel_source
    rho_m = mixt_var_rho_m(cfg, qp)
    return cfg.mfr_source / (length * rho_m * cfg.vel_source * 1.0)
