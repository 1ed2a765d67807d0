// rh_current.hpp -- B of current sources (block and arc currents).
//
// Block current (cpp/src/core/radrec.cpp:184-271 restates the same quantity
// with per-corner logs): for a uniform current density J in a box,
//     B(p) = 1e-4 * J x G(p),   G(p) = integral over the box of (p-x)/|p-x|^3 dV
// (1e-4 = mu0/4pi in mm, A/mm^2 and tesla).  By the divergence theorem
// G = sum_faces n_f * integral over f of dA/|x-p|, and the face potential of
// a planar convex polygon is
//     integral dA/|x-p| = sum_edges (m_e . (v_a - p)) L_e - h Omega,
// with h = (v - p).n (any face vertex v), L_e and Omega as in rh_math.hpp.
//
// Arc current: trapezoid rule over nseg sectors of the azimuthal integral,
// the numerical scheme of radTArcCur::B_compWithTrapeth
// (cpp/src/core/radarccu.cpp:30-234), written here for the device: the
// cosines and sines of the sector angles advance by rotation, and the logs
// and arctangents use the short forms of rh_math.hpp.
#pragma once

#include "rh_math.hpp"

namespace rh {

RH_HD void block_current_B(const double* par, const double p[3], double B[3])
{
    // par: centre[3], dims[3], J[3]
    const double cx = par[0] - p[0], cy = par[1] - p[1], cz = par[2] - p[2];
    const double hx = 0.5 * par[3], hy = 0.5 * par[4], hz = 0.5 * par[5];
    const double X[2] = {cx - hx, cx + hx}, Y[2] = {cy - hy, cy + hy}, Z[2] = {cz - hz, cz + hz};
    double Rc[2][2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++) Rc[i][j][k] = rh_sqrt(X[i] * X[i] + Y[j] * Y[j] + Z[k] * Z[k]);
    // edge logs
    double Lx[2][2], Ly[2][2], Lz[2][2];
    for (int j = 0; j < 2; j++)
        for (int k = 0; k < 2; k++) {
            double num, den;
            edge_ratio(v3(X[0], Y[j], Z[k]), Rc[0][j][k], v3(X[1], Y[j], Z[k]), Rc[1][j][k],
                       v3(1.0, 0.0, 0.0), num, den);
            Lx[j][k] = rh_log_ratio(num, den);
        }
    for (int i = 0; i < 2; i++)
        for (int k = 0; k < 2; k++) {
            double num, den;
            edge_ratio(v3(X[i], Y[0], Z[k]), Rc[i][0][k], v3(X[i], Y[1], Z[k]), Rc[i][1][k],
                       v3(0.0, 1.0, 0.0), num, den);
            Ly[i][k] = rh_log_ratio(num, den);
        }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            double num, den;
            edge_ratio(v3(X[i], Y[j], Z[0]), Rc[i][j][0], v3(X[i], Y[j], Z[1]), Rc[i][j][1],
                       v3(0.0, 0.0, 1.0), num, den);
            Lz[i][j] = rh_log_ratio(num, den);
        }
    // solid angle of a rectangle face at index `side` of axis a (outward normal
    // s*e_a, s = 2*side-1), corners counter-clockwise about the normal
    auto face_omega = [&](int a, int side) -> double {
        int c[4][3];
        const int cc[3][4][3] = {{{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},
                                 {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}},
                                 {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}}};
        for (int u = 0; u < 4; u++)
            for (int w = 0; w < 3; w++) c[u][w] = cc[a][u][w];
        if (side) {
            // opposite face: set the axis index to 1 and reverse the order
            int t[4][3];
            for (int u = 0; u < 4; u++)
                for (int w = 0; w < 3; w++) t[u][w] = c[(4 - u) % 4][w];
            for (int u = 0; u < 4; u++) {
                for (int w = 0; w < 3; w++) c[u][w] = t[u][w];
                c[u][a] = 1;
            }
        }
        V3<double> r[4];
        double rr[4];
        for (int u = 0; u < 4; u++) {
            r[u] = v3(X[c[u][0]], Y[c[u][1]], Z[c[u][2]]);
            rr[u] = Rc[c[u][0]][c[u][1]][c[u][2]];
        }
        double re1, im1, re2, im2;
        tri_factor(r[0], rr[0], r[1], rr[1], r[2], rr[2], re1, im1);
        tri_factor(r[0], rr[0], r[2], rr[2], r[3], rr[3], re2, im2);
        return 2.0 * rh_atan2(re1 * im2 + im1 * re2, re1 * re2 - im1 * im2);
    };
    // face potentials; for face (a, side) the edges run along the other two axes
    double G[3] = {0, 0, 0};
    const double S[2] = {-1.0, 1.0};
    for (int a = 0; a < 3; a++) {
        for (int side = 0; side < 2; side++) {
            double pot = 0.0;
            if (a == 0) {
                int i = side;
                // y-edges at z_k: in-plane outward normal S[k] e_z, distance Z[k]
                for (int k = 0; k < 2; k++) pot += S[k] * Z[k] * Ly[i][k];
                for (int j = 0; j < 2; j++) pot += S[j] * Y[j] * Lz[i][j];
            } else if (a == 1) {
                int j = side;
                for (int k = 0; k < 2; k++) pot += S[k] * Z[k] * Lx[j][k];
                for (int i = 0; i < 2; i++) pot += S[i] * X[i] * Lz[i][j];
            } else {
                int k = side;
                for (int j = 0; j < 2; j++) pot += S[j] * Y[j] * Lx[j][k];
                for (int i = 0; i < 2; i++) pot += S[i] * X[i] * Ly[i][k];
            }
            const double h = S[side] * (a == 0 ? X[side] : (a == 1 ? Y[side] : Z[side]));
            pot -= h * face_omega(a, side);
            G[a] += S[side] * pot;
        }
    }
    const double* J = par + 6;
    B[0] = 1e-4 * (J[1] * G[2] - J[2] * G[1]);
    B[1] = 1e-4 * (J[2] * G[0] - J[0] * G[2]);
    B[2] = 1e-4 * (J[0] * G[1] - J[1] * G[0]);
}

// atan(x)+atan(y) with the branch count of the tangent addition formula.
RH_HD double atan_pair(double x, double y)
{
    double d = 1.0 - x * y;
    if (d == 0.0) d = 1e-50;
    double k = (d > 0) ? 0.0 : ((x < 0) ? -1.0 : 1.0);
    return rh_atan2((x + y) * rh_rcp(d), 1.0) + 3.141592653589793238 * k;
}

RH_HD void arc_current_B(const double* par, const double P[3], double B[3])
{
    // par: centre[3], rmin, rmax, phimin, phimax, height, j, nseg
    const double pi = 3.141592653589793238;
    const double zero_tol = 1e-10, rel = 1e-7, small = 1e-10;
    const double rmin = par[3], rmax = par[4], phimin = par[5], phimax = par[6];
    const double height = par[7], jaz = par[8];
    const int nseg = int(par[9]);
    const double px = P[0] - par[0], py = P[1] - par[1], pz = P[2] - par[2];
    const double r = sqrt(px * px + py * py + small);
    const double phi = (py < 0) ? (2 * pi - acos(px / r)) : acos(px / r);
    const double z1 = -0.5 * height - pz, z2 = z1 + height;
    const double z1z1 = z1 * z1, z2z2 = z2 * z2, rr = r * r;
    const double r1r1 = rmin * rmin, r2r2 = rmax * rmax;
    const double a11 = rr + r1r1 + z1z1, a12 = rr + r1r1 + z2z2;
    const double a21 = rr + r2r2 + z1z1, a22 = rr + r2r2 + z2z2;
    const double tr1 = 2.0 * r * rmin, tr2 = 2.0 * r * rmax;
    const int npt = nseg + 1;
    const double step = (phimax - phimin + small) / (npt - 1);
    const double dphi0 = phimin - phi + small;
    // cos and sin of dphi = dphi0 + i step and of dphi + phi by rotation
    const double cs = cos(step), ss = sin(step);
    double cd = cos(dphi0), sd = sin(dphi0);
    double cp = cos(dphi0 + phi), sp = sin(dphi0 + phi);
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < npt; i++) {
        const double as = fabs(sd);
        double b11 = fabs(a11 - tr1 * cd);
        if (b11 < zero_tol) b11 = zero_tol;
        double b21 = fabs(a21 - tr2 * cd);
        if (b21 < zero_tol) b21 = zero_tol;
        const double D11 = rh_sqrt(b11), D12 = rh_sqrt(fabs(a12 - tr1 * cd));
        const double D21 = rh_sqrt(b21), D22 = rh_sqrt(fabs(a22 - tr2 * cd));
        const double rc = r * cd;
        const double u1 = rmin - rc, u2 = rmax - rc;
        const double q = z1z1 + rr * as * as;
        double q11 = u1 + D11, q12 = u1 + D12, q21 = u2 + D21, q22 = u2 + D22;
        if (fabs(q11) < D11 * rel && q < u1 * u1 * rel) q11 = 0.5 * q / fabs(u1);
        if (fabs(q12) < D12 * rel && q < u1 * u1 * rel) q12 = 0.5 * q / fabs(u1);
        if (fabs(q21) < D21 * rel && q < u2 * u2 * rel) q21 = 0.5 * q / fabs(u2);
        if (fabs(q22) < D22 * rel && q < u2 * u2 * rel) q22 = 0.5 * q / fabs(u2);
        const double l1 = rh_log_ratio(q11, q21), l2 = rh_log_ratio(q22, q12);
        const double ras = r * as;
        const double rr_as = (ras > 0.0) ? rh_rcp(ras) : INFINITY;  // as x/0 would give
        const double w1 = fabs(u1) * rr_as, w2 = fabs(u2) * rr_as;
        const double s1 = (u1 < 0) ? -1.0 : 1.0, s2 = (u2 < 0) ? -1.0 : 1.0;
        const double buf = ras * (s1 * atan_pair(z2 * w1 * rh_rcp(D12), -z1 * w1 * rh_rcp(D11)) +
                                  s2 * atan_pair(z1 * w2 * rh_rcp(D21), -z2 * w2 * rh_rcp(D22))) +
                           z1 * l1 + z2 * l2;
        const double e1 = rr + r1r1 - tr1 * cd, e2 = rr + r2r2 - tr2 * cd;
        double z1p11 = z1 + D11, z1p21 = z1 + D21, z2p12 = z2 + D12, z2p22 = z2 + D22;
        if (fabs(z1p11) < D11 * rel && e1 < z1z1 * rel) z1p11 = 0.5 * e1 / fabs(z1);
        if (fabs(z1p21) < D21 * rel && e2 < z1z1 * rel) z1p21 = 0.5 * e2 / fabs(z1);
        if (fabs(z2p12) < D12 * rel && e1 < z2z2 * rel) z2p12 = 0.5 * e1 / fabs(z2);
        if (fabs(z2p22) < D22 * rel && e2 < z2z2 * rel) z2p22 = 0.5 * e2 / fabs(z2);
        const double fxy = D11 - D21 + D22 - D12 + rc * (l1 + l2);
        const double fz = buf - rc * rh_log_ratio(z1p11 * z2p22, z2p12 * z1p21);
        const double w = (i == 0 || i == npt - 1) ? 0.5 : 1.0;
        sx += w * cp * fxy;
        sy += w * sp * fxy;
        sz += w * fz;
        const double cd1 = cd * cs - sd * ss, sd1 = sd * cs + cd * ss;
        const double cp1 = cp * cs - sp * ss, sp1 = sp * cs + cp * ss;
        cd = cd1;
        sd = sd1;
        cp = cp1;
        sp = sp1;
    }
    const double f = 1e-4 * jaz * step;
    B[0] = f * sx;
    B[1] = f * sy;
    B[2] = f * sz;
}

// Adds the lab-frame B of one current source copy: cp = Minv(9), V(3), F = s*M (9).
RH_HD void current_add(int kind, const double* par, const double* cp, const double x[3], double B[3])
{
    const double d0 = x[0] - cp[9], d1 = x[1] - cp[10], d2 = x[2] - cp[11];
    const double xl[3] = {cp[0] * d0 + cp[1] * d1 + cp[2] * d2, cp[3] * d0 + cp[4] * d1 + cp[5] * d2,
                          cp[6] * d0 + cp[7] * d1 + cp[8] * d2};
    double b[3];
    if (kind == 0) block_current_B(par, xl, b);
    else arc_current_B(par, xl, b);
    const double* F = cp + 12;
    B[0] += F[0] * b[0] + F[1] * b[1] + F[2] * b[2];
    B[1] += F[3] * b[0] + F[4] * b[1] + F[5] * b[2];
    B[2] += F[6] * b[0] + F[7] * b[1] + F[8] * b[2];
}

}  // namespace rh
