/*
 * radora.c -- CPU oracle for the field, interaction-matrix and relaxation paths.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of Radia's
 * CPU algorithms.  It is the checker used by tests/, by the smoke check and by
 * the CPU-baseline leg of the benchmark.  The product library never links,
 * loads or calls it.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the RadiaCUDA repository root).  Lengths are in mm, B/H/M in tesla (Radia's
 * mu0*H convention), current density in A/mm^2.
 *
 * Conventions of this file
 *  - 3x3 matrices are row-major double[9].
 *  - A "model" is a flat description of relaxable elements: per element a type
 *    (0 = rectangular block "RecMag", 1 = polyhedron), the element geometry in
 *    its own frame, and a list of transform copies (copy 0 = the element's
 *    main transform; further copies come from symmetries).  A copy maps a point
 *    x of the element frame to Mc*x + Vc and a field vector h to sc*Mc*h,
 *    following gmTrans (ext/genmath/gmtrans.h:30-85).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Thread count of the OpenMP loops (assembly rows, field points); the
 * reference's assembly and field loops are serial (radintrc.cpp:831-878,
 * radapl3.cpp:170-178), which nthreads = 1 reproduces for timing. */
void ora_set_threads(int nthreads) { omp_set_num_threads(nthreads > 0 ? nthreads : 1); }

#define ORA_PI_REC 3.141592653589793238 /* radrec.cpp:128 */
#define ORA_PI_PLN 3.14159265358979     /* radplnr2.cpp:32 */

/* ------------------------------------------------------------------------- */
/* Randomization ("FldLenRndSw"): core/radcnvrg.h:29-92, radapl2.cpp:2096-2112 */
/* ------------------------------------------------------------------------- */
static double g_abs_rand = 0.0, g_rel_rand = 0.0, g_zero_rand = 0.0;

/* on=1 restores Radia's default magnitudes (radcnvrg.h:39-44); on=0 is the
 * state FldLenRndSw('off') leaves behind (all magnitudes zero). */
void ora_set_randomization(int on)
{
    if (on) { g_abs_rand = 1e-9; g_rel_rand = 1e-11; g_zero_rand = 1e-9; }
    else    { g_abs_rand = 0.0;  g_rel_rand = 0.0;   g_zero_rand = 0.0; }
}

/* radcnvrg.h:85-91 */
static double abs_rand_mag(double a)
{
    double from_rel = g_rel_rand * a;
    double amax = (from_rel < g_abs_rand) ? g_abs_rand : from_rel;
    return (a != 0.) ? amax : ((amax < g_zero_rand) ? g_zero_rand : amax);
}

/* ------------------------------------------------------------------------- */
/* Small 3-vector / 3x3 helpers (ext/genmath/gmvect.h)                        */
/* ------------------------------------------------------------------------- */
static void mv3(const double *m, const double *v, double *o)
{
    double x = m[0]*v[0] + m[1]*v[1] + m[2]*v[2];
    double y = m[3]*v[0] + m[4]*v[1] + m[5]*v[2];
    double z = m[6]*v[0] + m[7]*v[1] + m[8]*v[2];
    o[0] = x; o[1] = y; o[2] = z;
}
static void mm3(const double *a, const double *b, double *o)
{
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            t[3*i+j] = a[3*i]*b[j] + a[3*i+1]*b[3+j] + a[3*i+2]*b[6+j];
    memcpy(o, t, sizeof t);
}
/* gmvect.h:325-341: cofactor inverse */
void ora_inv3(const double *m, double *o)
{
    double det = m[0]*m[4]*m[8] + m[1]*m[5]*m[6] + m[2]*m[3]*m[7]
               - m[2]*m[4]*m[6] - m[0]*m[5]*m[7] - m[1]*m[3]*m[8];
    double c[9] = {
         m[4]*m[8]-m[5]*m[7], -m[1]*m[8]+m[2]*m[7],  m[1]*m[5]-m[2]*m[4],
        -m[3]*m[8]+m[5]*m[6],  m[0]*m[8]-m[2]*m[6], -m[0]*m[5]+m[2]*m[3],
         m[3]*m[7]-m[4]*m[6], -m[0]*m[7]+m[1]*m[6],  m[0]*m[4]-m[1]*m[3]};
    double inv = 1. / det;
    for (int i = 0; i < 9; i++) o[i] = inv * c[i];
}
/* gmvect.h:345-: the two-argument inverse short-cuts diagonal matrices */
static void inv3_diag_aware(const double *m, double *o)
{
    if (m[1] == 0. && m[2] == 0. && m[3] == 0. && m[5] == 0. && m[6] == 0. && m[7] == 0.) {
        memset(o, 0, 9 * sizeof(double));
        o[0] = 1. / m[0]; o[4] = 1. / m[4]; o[8] = 1. / m[8];
        return;
    }
    ora_inv3(m, o);
}

/* ------------------------------------------------------------------------- */
/* RecMag: rectangular block with uniform M or J.  core/radrec.cpp:49-315    */
/* ------------------------------------------------------------------------- */

/* radg3d.h:938-946 */
static double trans_atans(double x, double y, double *pimult)
{
    double buf = 1. - x * y;
    if (buf == 0.) buf = 1.e-50;
    *pimult = ((buf > 0) ? 0. : 1.) * ((x < 0) ? -1. : 1.);
    return (x + y) / buf;
}
/* atan(a)+atan(b)+atan(c)+atan(d) as composed on radrec.cpp:130-141 */
static double sum4_atan(double a, double b, double c, double d)
{
    double p1, p2, p3;
    double t1 = trans_atans(a, b, &p1);
    double t2 = trans_atans(c, d, &p2);
    double t = trans_atans(t1, t2, &p3);
    return atan(t) + ORA_PI_REC * (p1 + p2 + p3);
}

typedef struct {
    double T[3], S[3];     /* Radia's T and S vectors, already x 1/(4 pi) */
    double bufvect[3];     /* current-density branch vector (radrec.cpp:199-215) */
    int inside;
} rec_eval_t;

/* Shared body of radTRecMag::B_comp.  want_cur selects the current-density
 * branch (individual logs, BufVect), otherwise the T/S (magnetic) branch. */
static void recmag_eval(const double P[3], const double C[3], const double D[3],
                        int want_cur, rec_eval_t *out)
{
    double px = P[0] - C[0], py = P[1] - C[1], pz = P[2] - C[2];
    double hx = 0.5 * D[0], hy = 0.5 * D[1], hz = 0.5 * D[2];
    double xs[2], ys[2], zs[2];
    for (int ii = 0; ii <= 1; ii++) {
        int eps = ii * 2 - 1;
        xs[ii] = -px + eps * hx; ys[ii] = -py + eps * hy; zs[ii] = -pz + eps * hz;
        if (xs[ii] == 0.) xs[ii] = abs_rand_mag(hx);
        if (ys[ii] == 0.) ys[ii] = abs_rand_mag(hy);
        if (zs[ii] == 0.) zs[ii] = abs_rand_mag(hz);
    }
    double x0 = xs[0], x1 = xs[1], y0 = ys[0], y1 = ys[1], z0 = zs[0], z1 = zs[1];
    out->inside = (x0 * x1 < 0) && (y0 * y1 < 0) && (z0 * z1 < 0);

    double x0e2 = x0*x0, x1e2 = x1*x1, y0e2 = y0*y0, y1e2 = y1*y1, z0e2 = z0*z0, z1e2 = z1*z1;
    double D000 = sqrt(x0e2+y0e2+z0e2), D100 = sqrt(x1e2+y0e2+z0e2);
    double D010 = sqrt(x0e2+y1e2+z0e2), D110 = sqrt(x1e2+y1e2+z0e2);
    double D001 = sqrt(x0e2+y0e2+z1e2), D101 = sqrt(x1e2+y0e2+z1e2);
    double D011 = sqrt(x0e2+y1e2+z1e2), D111 = sqrt(x1e2+y1e2+z1e2);

    double T0x = sum4_atan(y0*z0/(x0*D000), -y0*z1/(x0*D001), -y1*z0/(x0*D010), y1*z1/(x0*D011));
    double T1x = sum4_atan(-y0*z0/(x1*D100), y0*z1/(x1*D101), y1*z0/(x1*D110), -y1*z1/(x1*D111));
    double T0y = sum4_atan(x0*z0/(y0*D000), -x0*z1/(y0*D001), -x1*z0/(y0*D100), x1*z1/(y0*D101));
    double T1y = sum4_atan(-x0*z0/(y1*D010), x0*z1/(y1*D011), x1*z0/(y1*D110), -x1*z1/(y1*D111));
    double T0z = sum4_atan(x0*y0/(z0*D000), -x1*y0/(z0*D100), -x0*y1/(z0*D010), x1*y1/(z0*D110));
    double T1z = sum4_atan(-x0*y0/(z1*D001), x1*y0/(z1*D101), x0*y1/(z1*D011), -x1*y1/(z1*D111));

    double aD000 = 10.*abs_rand_mag(D000), aD010 = 10.*abs_rand_mag(D010);
    double aD001 = 10.*abs_rand_mag(D001), aD011 = 10.*abs_rand_mag(D011);
    double aD100 = 10.*abs_rand_mag(D100), aD110 = 10.*abs_rand_mag(D110);
    double aD101 = 10.*abs_rand_mag(D101), aD111 = 10.*abs_rand_mag(D111);

    /* radrec.cpp:152-177: cancellation-free R+u when R+u is tiny */
    double z0pD100 = z0+D100; if (z0pD100 < aD100) z0pD100 = 0.5*(x1e2+y0e2)/fabs(z0);
    double z1pD101 = z1+D101; if (z1pD101 < aD101) z1pD101 = 0.5*(x1e2+y0e2)/fabs(z1);
    double z1pD001 = z1+D001; if (z1pD001 < aD001) z1pD001 = 0.5*(x0e2+y0e2)/fabs(z1);
    double z0pD000 = z0+D000; if (z0pD000 < aD000) z0pD000 = 0.5*(x0e2+y0e2)/fabs(z0);
    double z0pD010 = z0+D010; if (z0pD010 < aD010) z0pD010 = 0.5*(x0e2+y1e2)/fabs(z0);
    double z1pD011 = z1+D011; if (z1pD011 < aD011) z1pD011 = 0.5*(x0e2+y1e2)/fabs(z1);
    double z1pD111 = z1+D111; if (z1pD111 < aD111) z1pD111 = 0.5*(x1e2+y1e2)/fabs(z1);
    double z0pD110 = z0+D110; if (z0pD110 < aD110) z0pD110 = 0.5*(x1e2+y1e2)/fabs(z0);

    double y0pD100 = y0+D100; if (y0pD100 < aD100) y0pD100 = 0.5*(x1e2+z0e2)/fabs(y0);
    double y1pD110 = y1+D110; if (y1pD110 < aD110) y1pD110 = 0.5*(x1e2+z0e2)/fabs(y1);
    double y1pD010 = y1+D010; if (y1pD010 < aD010) y1pD010 = 0.5*(x0e2+z0e2)/fabs(y1);
    double y0pD000 = y0+D000; if (y0pD000 < aD000) y0pD000 = 0.5*(x0e2+z0e2)/fabs(y0);
    double y0pD001 = y0+D001; if (y0pD001 < aD001) y0pD001 = 0.5*(x0e2+z1e2)/fabs(y0);
    double y1pD011 = y1+D011; if (y1pD011 < aD011) y1pD011 = 0.5*(x0e2+z1e2)/fabs(y1);
    double y1pD111 = y1+D111; if (y1pD111 < aD111) y1pD111 = 0.5*(x1e2+z1e2)/fabs(y1);
    double y0pD101 = y0+D101; if (y0pD101 < aD101) y0pD101 = 0.5*(x1e2+z1e2)/fabs(y0);

    double x0pD010 = x0+D010; if (x0pD010 < aD010) x0pD010 = 0.5*(y1e2+z0e2)/fabs(x0);
    double x1pD110 = x1+D110; if (x1pD110 < aD110) x1pD110 = 0.5*(y1e2+z0e2)/fabs(x1);
    double x1pD100 = x1+D100; if (x1pD100 < aD100) x1pD100 = 0.5*(y0e2+z0e2)/fabs(x1);
    double x0pD000 = x0+D000; if (x0pD000 < aD000) x0pD000 = 0.5*(y0e2+z0e2)/fabs(x0);
    double x0pD001 = x0+D001; if (x0pD001 < aD001) x0pD001 = 0.5*(y0e2+z1e2)/fabs(x0);
    double x1pD101 = x1+D101; if (x1pD101 < aD101) x1pD101 = 0.5*(y0e2+z1e2)/fabs(x1);
    double x1pD111 = x1+D111; if (x1pD111 < aD111) x1pD111 = 0.5*(y1e2+z1e2)/fabs(x1);
    double x0pD011 = x0+D011; if (x0pD011 < aD011) x0pD011 = 0.5*(y1e2+z1e2)/fabs(x0);

    double rz_a = z0pD100/z1pD101, rz_b = z1pD001/z0pD000, rz_c = z0pD010/z1pD011, rz_d = z1pD111/z0pD110;
    double ry_a = y0pD100/y1pD110, ry_b = y1pD010/y0pD000, ry_c = y0pD001/y1pD011, ry_d = y1pD111/y0pD101;
    double rx_a = x0pD010/x1pD110, rx_b = x1pD100/x0pD000, rx_c = x0pD001/x1pD101, rx_d = x1pD111/x0pD011;

    const double c = 1. / 4. / ORA_PI_REC;
    if (want_cur) {
        /* radrec.cpp:184-215 */
        double lz_a = log(rz_a), lz_b = log(rz_b), lz_c = log(rz_c), lz_d = log(rz_d);
        double ly_a = log(ry_a), ly_b = log(ry_b), ly_c = log(ry_c), ly_d = log(ry_d);
        double lx_a = log(rx_a), lx_b = log(rx_b), lx_c = log(rx_c), lx_d = log(rx_d);
        out->bufvect[0] = x0*T0x + x1*T1x + y0*(lz_a+lz_b) + y1*(lz_c+lz_d)
                        + z0*(ly_a+ly_b) + z1*(ly_c+ly_d);
        out->bufvect[1] = y0*T0y + y1*T1y + x0*(lz_c+lz_b) + x1*(lz_a+lz_d)
                        + z0*(lx_a+lx_b) + z1*(lx_c+lx_d);
        out->bufvect[2] = z0*T0z + z1*T1z + y0*(lx_c+lx_b) + y1*(lx_a+lx_d)
                        + x0*(ly_c+ly_b) + x1*(ly_a+ly_d);
        out->T[0] = out->T[1] = out->T[2] = 0.;
        out->S[0] = out->S[1] = out->S[2] = 0.;
        return;
    }
    /* radrec.cpp:274-290 */
    out->T[0] = c * (T0x + T1x); out->T[1] = c * (T0y + T1y); out->T[2] = c * (T0z + T1z);
    out->S[0] = c * -log(rx_a * rx_b * rx_c * rx_d);
    out->S[1] = c * -log(ry_a * ry_b * ry_c * ry_d);
    out->S[2] = c * -log(rz_a * rz_b * rz_c * rz_d);
    out->bufvect[0] = out->bufvect[1] = out->bufvect[2] = 0.;
}

/* PreRelax interaction block: rows (B, H, A) of radrec.cpp:292-301, so that
 * H = Q*M for a block of magnetization M. */
void ora_recmag_Q(const double P[3], const double C[3], const double D[3], double Q[9])
{
    rec_eval_t e;
    recmag_eval(P, C, D, 0, &e);
    const double *T = e.T, *S = e.S;
    Q[0] = T[0];  Q[1] = -S[2]; Q[2] = -S[1];
    Q[3] = -S[2]; Q[4] = T[1];  Q[5] = -S[0];
    Q[6] = -S[1]; Q[7] = -S[0]; Q[8] = T[2];
}

/* B and H of a magnetized block, radrec.cpp:303-313 (B = H + M inside). */
void ora_recmag_BH(const double P[3], const double C[3], const double D[3],
                   const double M[3], double B[3], double H[3])
{
    rec_eval_t e;
    recmag_eval(P, C, D, 0, &e);
    double Q[9];
    const double *T = e.T, *S = e.S;
    Q[0] = T[0];  Q[1] = -S[2]; Q[2] = -S[1];
    Q[3] = -S[2]; Q[4] = T[1];  Q[5] = -S[0];
    Q[6] = -S[1]; Q[7] = -S[0]; Q[8] = T[2];
    double h[3];
    mv3(Q, M, h);
    for (int k = 0; k < 3; k++) {
        H[k] = h[k];
        B[k] = h[k] + (e.inside ? M[k] : 0.);
    }
}

/* B (= H) of a current-carrying block, radrec.cpp:258-271 (ConstForJ = 1e-4). */
void ora_reccur_BH(const double P[3], const double C[3], const double D[3],
                   const double J[3], double B[3], double H[3])
{
    rec_eval_t e;
    recmag_eval(P, C, D, 1, &e);
    const double *v = e.bufvect;
    const double k = 0.0001;
    double b[3] = {J[1]*v[2] - J[2]*v[1], J[2]*v[0] - J[0]*v[2], J[0]*v[1] - J[1]*v[0]};
    for (int i = 0; i < 3; i++) { B[i] = k * b[i]; H[i] = k * b[i]; }
}

/* ------------------------------------------------------------------------- */
/* Polygon face integrals: core/radplnr2.cpp:29-386                           */
/* ------------------------------------------------------------------------- */
/* v2d: 2*nv vertex coordinates in the face frame (normal = +z), coordz: the
 * face plane's z, centr2d: the polygon's 2D centre (only feeds the jitter
 * magnitudes), P: observation point in the face frame; it is shifted in place
 * exactly as the reference shifts it.  Returns S = (Sx, Sy, Sz) with
 * H_face = -(1/4pi) * Mz * S, and *in_frame = (z > 0). */
void ora_polygon_S(const double *v2d, int nv, double coordz, const double centr2d[2],
                   double P[3], double S[3], int *in_frame)
{
    const double max_k = 1.E+09;
    double arx = abs_rand_mag(v2d[0] - centr2d[0]);
    double ary = abs_rand_mag(v2d[1] - centr2d[1]);
    double arz = abs_rand_mag(coordz);
    if (arz < arx) arz = arx;
    if (arz < ary) arz = ary;
    double rel_rand_magn = abs_rand_mag(1.);

    double z = coordz - P[2];
    if (z == 0.) { z = arz; P[2] -= arz; }
    double ze2 = z * z;

    double f2x = v2d[0] - P[0], f2y = v2d[1] - P[1]; /* First2d */
    double ax = -P[0], ay = -P[1];                    /* Vect2dToAdd */
    double x1 = f2x, y1 = f2y;
    if (x1 == 0.) { x1 = arx; f2x += arx; ax += arx; P[0] -= arx; }
    if (y1 == 0.) { y1 = ary; f2y += ary; ay += ary; P[1] -= ary; }
    double x1e2 = x1 * x1, y1e2 = y1 * y1;
    (void)y1e2;

    double Sx = 0., Sy = 0., sum_atan = 0., arg_logs2 = 1.;
    for (int i = 0; i < nv; i++) {
        double x2, y2;
        if (i != nv - 1) { x2 = v2d[2*(i+1)] + ax; y2 = v2d[2*(i+1)+1] + ay; }
        else { x2 = f2x; y2 = f2y; }
        if (x2 == 0.) { x2 = arx; f2x += arx; ax += arx; P[0] -= arx; }
        if (y2 == 0.) { y2 = ary; f2y += ary; ay += ary; P[1] -= ary; }
        double x2e2 = x2 * x2, y2e2 = y2 * y2;

        double x2mx1 = x2 - x1, y2my1 = y2 - y1;
        double abs_x2mx1 = fabs(x2mx1), abs_y2my1 = fabs(y2my1);
        if (abs_x2mx1 * max_k > abs_y2my1) {
            double k = y2my1 / x2mx1;
            double b = y1 - k * x1;
            if (b == 0.) {
                b = ary;
                y1 += ary; y1e2 = y1 * y1;
                y2my1 -= ary; abs_y2my1 = fabs(y2my1);
                k = y2my1 / x2mx1;
                f2y += ary; ay += ary; P[1] -= ary;
            }
            double bk = b * k, ke2 = k * k, be2 = b * b, twob = 2. * b;
            double ke2p1 = ke2 + 1.;
            if (ke2p1 == 0.) ke2p1 = rel_rand_magn;
            double sqrtke2p1 = sqrt(ke2p1);
            double kx1 = k * x1, kx2 = k * x2;
            double bpkx1 = y1, bpkx2 = y2;
            double bpkx1e2 = bpkx1 * bpkx1, bpkx2e2 = bpkx2 * bpkx2;
            double kx1mb = -b + kx1, kx2mb = -b + kx2;
            double R1 = sqrt(x1e2 + bpkx1e2 + ze2), R2 = sqrt(x2e2 + bpkx2e2 + ze2);
            double x1e2pze2 = x1e2 + ze2, x2e2pze2 = x2e2 + ze2;
            double bkpx1pke2x1 = bk + ke2p1 * x1, bkpx2pke2x2 = bk + ke2p1 * x2;
            double kze2 = k * ze2;
            double ke2ze2 = k * kze2;
            double ke2ze2mbe2 = ke2ze2 - be2, ke2ze2pbe2 = ke2ze2 + be2;
            double bx1 = b * x1, bx2 = b * x2;
            double R1pbpkx1 = bpkx1 + R1, R2pbpkx2 = bpkx2 + R2;

            const double max_rel_tol = 1.E-07;
            double arR2 = 100. * abs_rand_mag(R2);
            double arR1 = 100. * abs_rand_mag(R1);
            double max_arR1 = max_rel_tol * R1, max_arR2 = max_rel_tol * R2;
            if (arR1 > max_arR1) arR1 = max_arR1;
            if (arR2 > max_arR2) arR2 = max_arR2;
            if ((R1pbpkx1 < arR1) && (R1 > 100. * arR1) && ((x1e2 + ze2) < bpkx1e2 * max_rel_tol))
                R1pbpkx1 = 0.5 * (x1e2 + ze2) / fabs(bpkx1);
            if ((R2pbpkx2 < arR2) && (R2 > 100. * arR2) && ((x2e2 + ze2) < bpkx2e2 * max_rel_tol))
                R2pbpkx2 = 0.5 * (x2e2 + ze2) / fabs(bpkx2);

            /* radplnr2.cpp:309-318: direct atan2 accumulation */
            double a1 = -(ke2ze2pbe2 * (bx1 + kze2) * R1pbpkx1 + kze2 * twob * x1e2pze2);
            double a2 = (ke2ze2pbe2 * kx1mb * R1pbpkx1 + ke2ze2mbe2 * x1e2pze2) * z;
            double a3 = ke2ze2pbe2 * (bx2 + kze2) * R2pbpkx2 + kze2 * twob * x2e2pze2;
            double a4 = (ke2ze2pbe2 * kx2mb * R2pbpkx2 + ke2ze2mbe2 * x2e2pze2) * z;
            double num = a1 * a4 + a3 * a2;
            double den = a2 * a4 - a1 * a3;
            sum_atan += atan2(num, den);

            double t1 = bkpx1pke2x1 / sqrtke2p1 + R1;
            double t2 = bkpx2pke2x2 / sqrtke2p1 + R2;
            double be2pze2 = be2 + ze2;
            if ((t1 < arR1) && (R1 > 100. * arR1) && ((be2pze2 + 2 * bk * x1) < x1e2 * ke2p1 * max_rel_tol))
                t1 = 0.5 * be2pze2 / (fabs(x1) * sqrtke2p1);
            if ((t2 < arR2) && (R2 > 100. * arR2) && ((be2pze2 + 2 * bk * x2) < x2e2 * ke2p1 * max_rel_tol))
                t2 = 0.5 * be2pze2 / (fabs(x2) * sqrtke2p1);
            if (t1 == 0.) t1 = 1.e-50;
            if (t2 == 0.) t2 = 1.e-50;

            double sum_logs1 = log(t2 / t1);
            double sl1d = sum_logs1 / sqrtke2p1;
            if (R1pbpkx1 == 0.) R1pbpkx1 = 1.e-50;
            arg_logs2 *= (R2pbpkx2 / R1pbpkx1);
            Sx += -k * sl1d;
            Sy += sl1d;
        }
        x1 = x2; y1 = y2;
        x1e2 = x2e2; y1e2 = y2e2;
    }
    *in_frame = (z > 0.);
    if (arg_logs2 <= 0.) arg_logs2 = 1.e-50;
    Sx += log(arg_logs2);
    S[0] = Sx; S[1] = Sy; S[2] = sum_atan;
}

/* ------------------------------------------------------------------------- */
/* Polyhedron set-up and field: core/radvlpgn.cpp:98-505, radvlpgn.h:82-95   */
/* ------------------------------------------------------------------------- */
/* Inputs: verts[3*nv]; faces as face_len[nf] and face_idx (0-based, flat).
 * Outputs (caller-allocated):
 *   face_idx_out  same size as face_idx, each face re-oriented counter-clockwise
 *                 about its outward normal (radvlpgn.cpp:174-345 on a convex body)
 *   R[9*nf]       face rotation R1 (lab -> face frame), radvlpgn.cpp:357-371
 *   Rinv[9*nf]    its numerical inverse (gmTrans ctor, gmtrans.h:36-39)
 *   v2d[2*sum(face_len)]  face vertices in the face frame
 *   coordz[nf]    z of the face plane in the face frame (last vertex's, :407)
 *   c2d[2*nf]     2D centre of each face (vertex mean; jitter magnitudes only)
 *   centroid[3]   vertex average (radvlpgn.h:321-327 with randomization off)
 * Returns 0 on success, nonzero on a degenerate face. */
int ora_poly_prepare(const double *verts, int nv, const int *face_len, const int *face_idx,
                     int nf, int *face_idx_out, double *R, double *Rinv, double *v2d,
                     double *coordz, double *c2d, double *centroid)
{
    /* tolerances: radvlpgn.cpp:1305-1316 -> RelLengthTol = 100 * max(1e-9, RelRand) */
    double rel_tol = 1e-9;
    if (rel_tol < g_rel_rand) rel_tol = g_rel_rand;
    rel_tol *= 100.;

    /* typical size for the orientation test, radvlpgn.cpp:156-172 */
    double mx = 0, my = 0, mz = 0;
    for (int i = 1; i < nv; i++) {
        double rx = fabs(verts[3*i] - verts[0]), ry = fabs(verts[3*i+1] - verts[1]),
               rz = fabs(verts[3*i+2] - verts[2]);
        if (rx > mx) mx = rx;
        if (ry > my) my = ry;
        if (rz > mz) mz = rz;
    }
    double typ = (mx > my) ? mx : my;
    if (typ < mz) typ = mz;
    double abs_len_tol = rel_tol * typ;

    int off = 0;
    for (int f = 0; f < nf; f++) {
        int L = face_len[f];
        const int *fi = face_idx + off;
        int *fo = face_idx_out + off;
        if (L < 3) return 1;
        const double *P0 = verts + 3 * fi[0], *P1 = verts + 3 * fi[1];
        /* normal via 3 points, radvlpgn.cpp:98-150 (first non-degenerate triple) */
        double n[3] = {0, 0, 0};
        int found = 0;
        double nP0 = sqrt(P0[0]*P0[0] + P0[1]*P0[1] + P0[2]*P0[2]);
        double nP1 = sqrt(P1[0]*P1[0] + P1[1]*P1[1] + P1[2]*P1[2]);
        double tol = rel_tol * ((nP1 > nP0) ? nP1 : nP0);
        for (int j = 2; j < L; j++) {
            const double *P2 = verts + 3 * fi[j];
            double nP2 = rel_tol * sqrt(P2[0]*P2[0] + P2[1]*P2[1] + P2[2]*P2[2]);
            if (nP2 > tol) tol = nP2;
            double a[3] = {P1[0]-P0[0], P1[1]-P0[1], P1[2]-P0[2]};
            double b[3] = {P2[0]-P0[0], P2[1]-P0[1], P2[2]-P0[2]};
            n[0] = a[1]*b[2] - a[2]*b[1];
            n[1] = a[2]*b[0] - a[0]*b[2];
            n[2] = a[0]*b[1] - a[1]*b[0];
            double t2 = tol * tol;
            if (fabs(n[0]) > t2 || fabs(n[1]) > t2 || fabs(n[2]) > t2) { found = 1; break; }
        }
        if (!found) return 2;
        double inv = 1. / sqrt(n[0]*n[0] + n[1]*n[1] + n[2]*n[2]);
        n[0] *= inv; n[1] *= inv; n[2] *= inv;
        /* outward orientation: every vertex must lie on the inner side */
        int flip = 0;
        for (int j = 0; j < nv; j++) {
            double d = (P0[0]-verts[3*j])*n[0] + (P0[1]-verts[3*j+1])*n[1] + (P0[2]-verts[3*j+2])*n[2];
            if (d < -abs_len_tol) { flip = 1; break; }
        }
        if (flip) {
            n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
            for (int k = 0; k < L; k++) fo[k] = fi[L - 1 - k];
        } else {
            for (int k = 0; k < L; k++) fo[k] = fi[k];
        }
        /* face rotation R1, radvlpgn.cpp:357-371 */
        double *r = R + 9 * f;
        if (fabs(n[2] + 1.) > rel_tol) {
            double q = 1. / (n[2] + 1.);
            r[0] = n[1]*n[1]*q + n[2]; r[1] = -n[0]*n[1]*q; r[2] = -n[0];
            r[3] = r[1];               r[4] = n[0]*n[0]*q + n[2]; r[5] = -n[1];
            r[6] = -r[2];              r[7] = -r[5];              r[8] = n[2];
        } else {
            r[0] = 1; r[1] = 0;  r[2] = 0;
            r[3] = 0; r[4] = -1; r[5] = 0;
            r[6] = 0; r[7] = 0;  r[8] = -1;
        }
        ora_inv3(r, Rinv + 9 * f);
        double cx = 0, cy = 0, lz = 0;
        for (int k = 0; k < L; k++) {
            double pl[3];
            mv3(r, verts + 3 * fo[k], pl);
            v2d[2*(off+k)] = pl[0]; v2d[2*(off+k)+1] = pl[1];
            cx += pl[0]; cy += pl[1];
            lz = pl[2];
        }
        coordz[f] = lz;
        c2d[2*f] = cx / L; c2d[2*f+1] = cy / L;
        off += L;
    }
    double s[3] = {0, 0, 0};
    for (int i = 0; i < nv; i++) { s[0] += verts[3*i]; s[1] += verts[3*i+1]; s[2] += verts[3*i+2]; }
    double b = 1. / nv;
    centroid[0] = b * s[0]; centroid[1] = b * s[1]; centroid[2] = b * s[2];
    return 0;
}

/* radTPolyhedron::B_comp_frM, radvlpgn.cpp:441-505.  want_Q selects the
 * PreRelax (interaction block) mode, otherwise B and H for magnetization M. */
static void poly_eval(const double Pin[3], int nf, const int *face_len, const double *R,
                      const double *Rinv, const double *v2d, const double *coordz,
                      const double *c2d, const double *M, int want_Q,
                      double Q[9], double B[3], double H[3])
{
    const double cH = 1. / 4. / ORA_PI_PLN;
    double P[3] = {Pin[0], Pin[1], Pin[2]};
    double q[9] = {0};
    double h[3] = {0, 0, 0};
    int inside = 1, off = 0;
    for (int f = 0; f < nf; f++) {
        const double *r = R + 9 * f, *ri = Rinv + 9 * f;
        double pl[3], prev[3];
        mv3(r, P, pl);
        memcpy(prev, pl, sizeof prev);
        double S[3];
        int in_frame;
        ora_polygon_S(v2d + 2 * off, face_len[f], coordz[f], c2d + 2 * f, pl, S, &in_frame);
        if (pl[0] != prev[0] || pl[1] != prev[1] || pl[2] != prev[2]) mv3(ri, pl, P);
        if (!in_frame) inside = 0;
        if (want_Q) {
            /* local rows (0,0,-cSx),(0,0,-cSy),(0,0,-cSz) -> Rinv * Qloc * R */
            double ql[9] = {0, 0, -cH * S[0], 0, 0, -cH * S[1], 0, 0, -cH * S[2]};
            double t[9];
            mm3(ql, r, t);
            mm3(ri, t, t);
            for (int k = 0; k < 9; k++) q[k] += t[k];
        } else {
            double ml[3];
            mv3(r, M, ml);
            double hl[3] = {-cH * ml[2] * S[0], -cH * ml[2] * S[1], -cH * ml[2] * S[2]};
            double hg[3];
            mv3(ri, hl, hg);
            h[0] += hg[0]; h[1] += hg[1]; h[2] += hg[2];
        }
        off += face_len[f];
    }
    if (want_Q) { memcpy(Q, q, sizeof q); return; }
    for (int k = 0; k < 3; k++) {
        H[k] = h[k];
        B[k] = h[k] + (inside ? M[k] : 0.);
    }
}

void ora_poly_Q(const double P[3], int nf, const int *face_len, const double *R, const double *Rinv,
                const double *v2d, const double *coordz, const double *c2d, double Q[9])
{
    poly_eval(P, nf, face_len, R, Rinv, v2d, coordz, c2d, NULL, 1, Q, NULL, NULL);
}

void ora_poly_BH(const double P[3], int nf, const int *face_len, const double *R, const double *Rinv,
                 const double *v2d, const double *coordz, const double *c2d, const double M[3],
                 double B[3], double H[3])
{
    poly_eval(P, nf, face_len, R, Rinv, v2d, coordz, c2d, M, 0, NULL, B, H);
}

/* ------------------------------------------------------------------------- */
/* Model: relaxable elements with transform copies                           */
/* ------------------------------------------------------------------------- */
/* Element e (0..n-1):
 *   type[e]            0 = RecMag, 1 = polyhedron
 *   rec_c[3e], rec_d[3e]   RecMag centre / full dimensions (element frame)
 *   poly_face_off[e]..[e+1]  its faces in the prepared face arrays
 *   face_vtx_off[f]          start of face f's vertices in v2d (in vertices)
 *   centroid[3e]       collocation point in the element frame
 *   copy_off[e]..[e+1] transform copies; copy 0 is the main transform
 *   cM[9c], cMinv[9c], cV[3c], cS[c]
 */
typedef struct {
    int64_t n;
    const int32_t *type;
    const double *rec_c, *rec_d;
    const int64_t *poly_face_off;
    const int32_t *face_len;
    const int64_t *face_vtx_off;
    const double *R, *Rinv, *v2d, *coordz, *c2d;
    const double *centroid;
    const int64_t *copy_off;
    const double *cM, *cMinv, *cV, *cS;
} ora_model_t;

static void elem_Q(const ora_model_t *m, int64_t j, const double x[3], double Q[9])
{
    if (m->type[j] == 0) {
        ora_recmag_Q(x, m->rec_c + 3 * j, m->rec_d + 3 * j, Q);
    } else {
        int64_t f0 = m->poly_face_off[j];
        int nf = (int)(m->poly_face_off[j + 1] - f0);
        int64_t v0 = m->face_vtx_off[f0];
        ora_poly_Q(x, nf, m->face_len + f0, m->R + 9 * f0, m->Rinv + 9 * f0,
                   m->v2d + 2 * v0, m->coordz + f0, m->c2d + 2 * f0, Q);
    }
}

static void elem_BH(const ora_model_t *m, int64_t j, const double x[3], const double M[3],
                    double B[3], double H[3])
{
    if (m->type[j] == 0) {
        ora_recmag_BH(x, m->rec_c + 3 * j, m->rec_d + 3 * j, M, B, H);
    } else {
        int64_t f0 = m->poly_face_off[j];
        int nf = (int)(m->poly_face_off[j + 1] - f0);
        int64_t v0 = m->face_vtx_off[f0];
        ora_poly_BH(x, nf, m->face_len + f0, m->R + 9 * f0, m->Rinv + 9 * f0,
                    m->v2d + 2 * v0, m->coordz + f0, m->c2d + 2 * f0, M, B, H);
    }
}

/* Collocation point of element i in the lab frame: MainTrans->TrPoint(centroid),
 * radintrc.cpp:852. */
static void colloc_point(const ora_model_t *m, int64_t i, double x[3])
{
    int64_t c = m->copy_off[i];
    mv3(m->cM + 9 * c, m->centroid + 3 * i, x);
    x[0] += m->cV[3*c]; x[1] += m->cV[3*c+1]; x[2] += m->cV[3*c+2];
}

/* Interaction matrix, radintrc.cpp:831-878:
 *   Q_ij = s_i Minv_i * sum_c s_c M_c * Q_j(Minv_c (x_i - V_c))
 * Output Q is (3n x 3n) row-major doubles; rows i0..i1-1 only (block rows). */
void ora_assemble_rows(const ora_model_t *m, int64_t i0, int64_t i1, double *Qout)
{
    int64_t n3 = 3 * m->n;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = i0; i < i1; i++) {
        double x[3];
        colloc_point(m, i, x);
        int64_t ci = m->copy_off[i];
        for (int64_t j = 0; j < m->n; j++) {
            double sub[9] = {0};
            for (int64_t c = m->copy_off[j]; c < m->copy_off[j + 1]; c++) {
                double d[3] = {x[0] - m->cV[3*c], x[1] - m->cV[3*c+1], x[2] - m->cV[3*c+2]};
                double xl[3], q[9], t[9];
                mv3(m->cMinv + 9 * c, d, xl);
                elem_Q(m, j, xl, q);
                mm3(m->cM + 9 * c, q, t);
                for (int k = 0; k < 9; k++) sub[k] += m->cS[c] * t[k];
            }
            double t[9];
            mm3(m->cMinv + 9 * ci, sub, t);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    Qout[(3 * (i - i0) + a) * n3 + 3 * j + b] = m->cS[ci] * t[3 * a + b];
        }
    }
}

/* Field of the relaxable elements at points (lab frame), radg3d.cpp:57-115
 * (NestedFor_B): B(x) = sum_e sum_c s_c M_c B_e(Minv_c (x - V_c)) with the
 * element's magnetization Mloc[3e] in its own frame. */
void ora_field(const ora_model_t *m, const double *Mloc, const double *pts, int64_t np,
               double *Bout, double *Hout)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < np; p++) {
        double b[3] = {0, 0, 0}, h[3] = {0, 0, 0};
        const double *x = pts + 3 * p;
        for (int64_t j = 0; j < m->n; j++) {
            for (int64_t c = m->copy_off[j]; c < m->copy_off[j + 1]; c++) {
                double d[3] = {x[0] - m->cV[3*c], x[1] - m->cV[3*c+1], x[2] - m->cV[3*c+2]};
                double xl[3], bl[3], hl[3], bg[3], hg[3];
                mv3(m->cMinv + 9 * c, d, xl);
                elem_BH(m, j, xl, Mloc + 3 * j, bl, hl);
                mv3(m->cM + 9 * c, bl, bg);
                mv3(m->cM + 9 * c, hl, hg);
                for (int k = 0; k < 3; k++) { b[k] += m->cS[c] * bg[k]; h[k] += m->cS[c] * hg[k]; }
            }
        }
        if (Bout) { Bout[3*p] = b[0]; Bout[3*p+1] = b[1]; Bout[3*p+2] = b[2]; }
        if (Hout) { Hout[3*p] = h[0]; Hout[3*p+1] = h[1]; Hout[3*p+2] = h[2]; }
    }
}

/* Flat-argument wrappers (ctypes-friendly). */
#define ORA_MODEL_ARGS int64_t n, const int32_t *type, const double *rec_c, const double *rec_d, \
    const int64_t *poly_face_off, const int32_t *face_len, const int64_t *face_vtx_off,          \
    const double *R, const double *Rinv, const double *v2d, const double *coordz,               \
    const double *c2d, const double *centroid, const int64_t *copy_off, const double *cM,       \
    const double *cMinv, const double *cV, const double *cS
#define ORA_MODEL_INIT {n, type, rec_c, rec_d, poly_face_off, face_len, face_vtx_off, R, Rinv, \
    v2d, coordz, c2d, centroid, copy_off, cM, cMinv, cV, cS}

void ora_model_assemble(ORA_MODEL_ARGS, int64_t i0, int64_t i1, double *Q)
{
    ora_model_t m = ORA_MODEL_INIT;
    ora_assemble_rows(&m, i0, i1, Q);
}

void ora_model_field(ORA_MODEL_ARGS, const double *Mloc, const double *pts, int64_t np,
                     double *B, double *H)
{
    ora_model_t m = ORA_MODEL_INIT;
    ora_field(&m, Mloc, pts, np, B, H);
}

void ora_model_colloc(ORA_MODEL_ARGS, double *X)
{
    ora_model_t m = ORA_MODEL_INIT;
    for (int64_t i = 0; i < n; i++) colloc_point(&m, i, X + 3 * i);
}

/* ------------------------------------------------------------------------- */
/* Materials: core/radmater.h:164-696, radmater.cpp:72-162                   */
/* ------------------------------------------------------------------------- */
/* Material record (per element), 16 doubles:
 *   [0] kind: 0 linear isotropic (radmater.h:328-412)
 *             1 linear anisotropic (radmater.h:164-303), tensor in [4..12]
 *             2 nonlinear isotropic formula (radmater.h:417-696)
 *   kind 0: [1] ksi
 *   kind 1: [4..12] ksi tensor (row-major)
 *   kind 2: [1] n terms, [2..4] ms, [5..7] ks
 *   [13..15] remanent magnetization (element frame)
 */
#define MAT_STRIDE 16

/* ------------------------------------------------------------------------- */
/* Tabulated isotropic law (MatSatIsoTab): radmater.h:436-451, 611-641,       */
/* radmater.cpp:341-646.  Nodes (x = |H|, y = |M|) with node derivatives d;   */
/* kind 3 records hold [1] first node (index into the table of x,y,d          */
/* triples set by ora_set_tables), [2] node count.                            */
/* ------------------------------------------------------------------------- */
static const double *g_tab = 0;
void ora_set_tables(const double *tab) { g_tab = tab; }

/* Derivative5: radmater.cpp:341-424 (derivative of the 5-point interpolant at
 * node k, clamped to 0 for increasing data). */
static double tab_der5(const double *x, const double *y, int k)
{
    double x0 = x[0], x1 = x[1] - x0, x2 = x[2] - x0, x3 = x[3] - x0, x4 = x[4] - x0;
    double f0 = y[0], f1 = y[1], f2 = y[2], f3 = y[3], f4 = y[4];
    int inc = (x1 > 0. && x2 > x1 && x3 > x2 && x4 > x3) &&
              (f0 <= f1 && f1 <= f2 && f2 <= f3 && f3 <= f4);
    double x1m2 = x1 - x2, x1m3 = x1 - x3, x1m4 = x1 - x4, x2m3 = x2 - x3, x2m4 = x2 - x4, x3m4 = x3 - x4;
    double x1e2 = x1*x1, x2e2 = x2*x2, x3e2 = x3*x3, x4e2 = x4*x4;
    double den = x1*x1m2*x2*x1m3*x2m3*x3*x1m4*x2m4*x3m4*x4, der = 0.;
    if (k == 0) {
        der = (f4*x1e2*x1m2*x2e2*x1m3*x2m3*x3e2 +
               (-f3*x1e2*x1m2*x2e2*x1m4*x2m4 + x3e2*(f2*x1e2*x1m3*x1m4 - f1*x2e2*x2m3*x2m4)*x3m4)*x4e2 -
               f0*x1m2*x1m3*x2m3*x1m4*x2m4*x3m4*(x1*x2*x3 + x2*x3*x4 + x1*(x2 + x3)*x4))/den;
    } else if (k == 1) {
        double x1e3 = x1e2*x1, x1m2e2 = x1m2*x1m2, x1m3e2 = x1m3*x1m3, x1m4e2 = x1m4*x1m4;
        der = (-f4*x1e2*x1m2e2*x2*x1m3e2*x2m3*x3 + f0*x1m2e2*x1m3e2*x2m3*x1m4e2*x2m4*x3m4 +
               x4*(f3*x1e2*x1m2e2*x2*x1m4e2*x2m4 -
                   x3*x3m4*(f2*x1e2*x1m3e2*x1m4e2 + f1*x2*x2m3*x2m4*
                            (-4*x1e3 + x2*x3*x4 + 3*x1e2*(x2 + x3 + x4) - 2*x1*(x3*x4 + x2*(x3 + x4))))))/den;
    } else if (k == 2) {
        double x1m2e2 = x1m2*x1m2, x2m3e2 = x2m3*x2m3, x2m4e2 = x2m4*x2m4;
        der = (f4*x1*x1m2e2*x2e2*x1m3*x2m3e2*x3 - f0*x1m2e2*x1m3*x2m3e2*x1m4*x2m4e2*x3m4 +
               x4*(-f3*x1*x1m2e2*x2e2*x1m4*x2m4e2 +
                   x3*x3m4*(f1*x2e2*x2m3e2*x2m4e2 + f2*x1*x1m3*x1m4*
                            (x1*(3*x2e2 + x3*x4 - 2*x2*(x3 + x4)) + x2*(-4*x2e2 - 2*x3*x4 + 3*x2*(x3 + x4))))))/den;
    }
    if (inc && der < 0.) der = 0.;
    return der;
}

/* Derivative3: radmater.cpp:426-455 */
static double tab_der3(const double *x, const double *y, int k)
{
    double x1 = x[1] - x[0], x2 = x[2] - x[0];
    double f0 = y[0], f1 = y[1], f2 = y[2];
    int inc = (x1 > 0. && x2 > x1) && (f0 <= f1 && f1 <= f2);
    double x1e2 = x1*x1, x2e2 = x2*x2, x1m2 = x1 - x2, der = 0.;
    if (k == 0) der = (f0*x1e2 - f2*x1e2 - f0*x2e2 + f1*x2e2)/(-x1*x2*x1m2);
    else if (k == 1) der = (-f2*x1e2 + f0*x1m2*x1m2 + f1*(2*x1 - x2)*x2)/(x1*x1m2*x2);
    else if (k == 2) der = (f2*x1*(x1 - 2*x2) - f0*x1m2*x1m2 + f1*x2e2)/(x1*x1m2*x1m2*x2);
    if (inc && der < 0.) der = 0.;
    return der;
}

/* Node table from user points (H, M): prepends (0, 0) when both first values
 * exceed 1e-10 (radmater.h:436-451); node derivatives as Compute_dMdH
 * (radmater.cpp:460-550, including its index pattern for 4 to 6 nodes) and the
 * monotone correction CheckAndCorrect_dMdH (radmater.cpp:556-609).  Writes
 * x, y, d (capacity n + 1); returns the node count. */
int ora_tab_prepare(const double *pts, int n, double *x, double *y, double *d)
{
    int L = 0;
    if (pts[0] > 1e-10 && pts[1] > 1e-10) { x[0] = 0.; y[0] = 0.; L = 1; }
    for (int i = 0; i < n; i++) { x[L] = pts[2*i]; y[L] = pts[2*i+1]; L++; }
    if (L < 4) return -1;  /* the node derivatives read 4 nodes */
    int w = 0, t = 0;
    d[w++] = tab_der3(x, y, 0);
    d[w++] = tab_der3(x, y, 1);
    t = 1;
    d[w++] = tab_der3(x + t, y + t, 1);
    for (int i = 3; i < L - 4; i++) { d[w++] = tab_der5(x + t, y + t, 2); t++; }
    if (L >= 7) {
        d[w++] = tab_der5(x + t, y + t, 2); t += 2;
        d[w++] = tab_der3(x + t, y + t, 1); t++;
        d[w++] = tab_der3(x + t, y + t, 1); t++;
        d[w] = (y[t+1] - y[t]) / (x[t+1] - x[t]);
    } else if (L >= 6) {
        d[w++] = tab_der5(x + t, y + t, 2); t += 2;
        d[w++] = tab_der3(x + t, y + t, 1); t++;
        d[w] = (y[t+1] - y[t]) / (x[t+1] - x[t]);
    } else if (L >= 5) {
        d[w++] = tab_der3(x + t, y + t, 1); t++;
        d[w] = (y[t+1] - y[t]) / (x[t+1] - x[t]);
    } else if (L >= 4) {
        d[w] = (y[t+1] - y[t]) / (x[t+1] - x[t]);
    }
    for (int i = L - 2; i >= 0; i--) {
        double x2 = x[i+1] - x[i], f1 = y[i], f2 = y[i+1], fd1 = d[i], fd2 = d[i+1];
        double x2e2 = x2*x2, x2e3 = x2e2*x2, fm = f1 - f2;
        double a1 = fd1, a2 = -((3*fm + (2*fd1 + fd2)*x2)/x2e2), a3 = -((-2*fm - (fd1 + fd2)*x2)/x2e3);
        double D = a2*a2 - 3.*a1*a3, xc1 = 1.E+23, xc2 = 1.E+23;
        if (a3 != 0. && D >= 0.) {
            double R = sqrt(D), b = 1./(3.*a3);
            xc1 = (-a2 - R)*b; xc2 = (-a2 + R)*b;
        }
        int mono = (f1 < f2 && fd1 > 0. && fd2 > 0.) || (f1 > f2 && fd1 < 0. && fd2 < 0.);
        int in1 = 0, in2 = 0;
        if (a3 != 0.) { in1 = (0. < xc1 && xc1 < x2); in2 = (0. < xc2 && xc2 < x2); }
        if (mono && D > 0. && (in1 || in2)) {
            double LD = -3*fd2*x2e3*(4*fm + fd2*x2);
            if (LD >= 0.) {
                double LR = sqrt(LD), b1 = 1./(2*x2e2), b2 = -x2*(6*fm + fd2*x2);
                double fa = (b2 - LR)*b1, fb = (b2 + LR)*b1;
                d[i] = (fabs(fd1 - fa) < fabs(fd1 - fb)) ? fa : fb;
            }
        }
    }
    return L;
}

/* AbsMvsAbsH_Interpol: radmater.cpp:619-646 with CubPln (radmater.h:631-639). */
double ora_tab_eval(const double *tab, int L, double h)
{
    int idx = 0;
    while (idx < L && !(tab[3*idx] > h)) idx++;
    idx--;
    if (idx < 0) return tab[1];
    if (idx >= L - 1) return tab[3*(L-1)+1] + (h - tab[3*(L-1)]) * tab[3*(L-1)+2];
    double arg = h - tab[3*idx], step = tab[3*(idx+1)] - tab[3*idx];
    double f1 = tab[3*idx+1], f2 = tab[3*(idx+1)+1], p1 = tab[3*idx+2], p2 = tab[3*(idx+1)+2];
    double inv = 1./step, s = (f2 - f1)*inv;
    double a2 = (3.*s - 2.*p1 - p2)*inv, a3 = (-2.*s + p1 + p2)*inv*inv;
    return f1 + arg*(p1 + arg*(a2 + arg*a3));
}

/* M(H): radmater.h:224 (aniso), :351 (iso), :644-657 (nonlinear) */
void ora_mat_M(const double *mat, const double H[3], double M[3])
{
    int kind = (int)mat[0];
    const double *mr = mat + 13;
    if (kind == 0) {
        for (int k = 0; k < 3; k++) M[k] = mat[1] * H[k] + mr[k];
    } else if (kind == 1) {
        mv3(mat + 4, H, M);
        for (int k = 0; k < 3; k++) M[k] += mr[k];
    } else {
        int nt = (int)mat[1];
        double absH = sqrt(H[0]*H[0] + H[1]*H[1] + H[2]*H[2]);
        double absM = 0.;
        if (kind == 3) absM = ora_tab_eval(g_tab + 3 * (int64_t)mat[1], (int)mat[2], absH);
        else for (int i = 0; i < nt; i++) {
            double ms = mat[2 + i], ks = mat[5 + i];
            if (ms != 0.) absM += ms * tanh(ks * absH / ms);
        }
        if (absH != 0) for (int k = 0; k < 3; k++) M[k] = (absM / absH) * H[k] + mr[k];
        else for (int k = 0; k < 3; k++) M[k] = mr[k];
    }
}

/* FindNewH: radmater.h:232-240 (aniso), :356-364 (iso), radmater.cpp:72-162 (nonlinear).
 * H in/out; Q = self block (float values widened to double); Hext = quasi-external field. */
static void find_new_H(const double *mat, double H[3], const double Q[9], const double Hext[3],
                       double prec_e2)
{
    int kind = (int)mat[0];
    const double *mr = mat + 13;
    double buf[9], inv[9], qmr[3], rhs[3];
    if (kind == 0 || kind == 1) {
        if (kind == 0) {
            for (int k = 0; k < 9; k++) buf[k] = -mat[1] * Q[k];
        } else {
            mm3(Q, mat + 4, buf);
            for (int k = 0; k < 9; k++) buf[k] = -buf[k];
        }
        buf[0] += 1.; buf[4] += 1.; buf[8] += 1.;
        inv3_diag_aware(buf, inv);
        mv3(Q, mr, qmr);
        for (int k = 0; k < 3; k++) rhs[k] = Hext[k] + qmr[k];
        mv3(inv, rhs, H);
        return;
    }
    const double abs_h_zero_tol = 1.E-10;
    const int max_iter = 15;
    int nt = (int)mat[1];
    double misfit = 3;
    double absH = sqrt(H[0]*H[0] + H[1]*H[1] + H[2]*H[2]);
    double f, inst_ksi = 0.;
    const double *tab = (kind == 3) ? g_tab + 3 * (int64_t)mat[1] : 0;
    for (int it = 0; it < max_iter; it++) {
        f = 0.;
        if (tab) {
            if (absH <= abs_h_zero_tol) { inst_ksi = tab[2]; f = 0; }
            else { f = ora_tab_eval(tab, (int)mat[2], absH); inst_ksi = f / absH; }
        } else if (absH <= abs_h_zero_tol) {
            for (int j = 0; j < nt; j++) inst_ksi += mat[5 + j];
            f = 0;
        } else {
            for (int j = 0; j < nt; j++) {
                double ms = mat[2 + j], ks = mat[5 + j];
                if (ms != 0.) f += ms * tanh(ks * absH / ms);
            }
            inst_ksi = f / absH;
        }
        for (int k = 0; k < 9; k++) buf[k] = -inst_ksi * Q[k];
        buf[0] += 1.; buf[4] += 1.; buf[8] += 1.;
        inv3_diag_aware(buf, inv);
        mv3(inv, Hext, H);
        double new_absH = sqrt(H[0]*H[0] + H[1]*H[1] + H[2]*H[2]);
        double new_misfit = f - new_absH * inst_ksi;
        double am = fabs(misfit), anm = fabs(new_misfit);
        double prob_new = am;
        if (new_misfit * misfit > 0) prob_new += 0.5 * anm;
        double prob_old = anm;
        double alpha = prob_new / (prob_new + prob_old);
        absH = alpha * new_absH + (1 - alpha) * absH;
        misfit = f - absH * inst_ksi;
        if (misfit * misfit <= prec_e2) break;
    }
}

/* ------------------------------------------------------------------------- */
/* Relaxation, method 4 (Gauss-Seidel): radrlmet.cpp:32-67, 496-587, 772-829  */
/* ------------------------------------------------------------------------- */
/* Qf: float interaction matrix (3n x 3n row-major), as Radia stores it
 * (radintrc.h:126).  Hext: external field at the collocation points in element
 * frames.  M, H: in/out (M is reset to the remanent magnetization unless
 * keep_M).  stats[0..2] = misfit, max|M|, max|H|.  Returns the iteration count
 * the reference reports (mIterCount - 1). */
int ora_relax_m4(int64_t n, const float *Qf, const double *mats, const double *Hext,
                 double *M, double *H, double prec, int max_iter, int keep_M, double *stats)
{
    int64_t n3 = 3 * n;
    if (!keep_M) {
        for (int64_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) { M[3*i+k] = mats[MAT_STRIDE*i + 13 + k]; H[3*i+k] = 0.; }
    }
    double desired_e2 = prec * prec;
    double inst_e2 = 1.E+23;
    int iter = 0;
    while (inst_e2 > desired_e2) {
        if (++iter > max_iter) break;
        double norm = 1. / (double)n;
        double best = desired_e2 * norm;
        double loc = (inst_e2 > 1.E+20) ? best : 0.25 * norm * inst_e2;
        if (loc < best) loc = best;
        double mis = 0.;
        for (int64_t i = 0; i < n; i++) {
            double hq[3] = {0, 0, 0};
            for (int64_t j = 0; j < n; j++) {
                if (j == i) continue;
                float mf[3] = {(float)M[3*j], (float)M[3*j+1], (float)M[3*j+2]};
                for (int a = 0; a < 3; a++) {
                    const float *row = Qf + (3 * i + a) * n3 + 3 * j;
                    float d = row[0] * mf[0] + row[1] * mf[1] + row[2] * mf[2];
                    hq[a] += (double)d;
                }
            }
            for (int a = 0; a < 3; a++) hq[a] += Hext[3*i+a];
            double qii[9];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) qii[3*a+b] = (double)Qf[(3*i+a)*n3 + 3*i + b];
            find_new_H(mats + MAT_STRIDE * i, H + 3 * i, qii, hq, loc);
            double mn[3];
            ora_mat_M(mats + MAT_STRIDE * i, H + 3 * i, mn);
            double d0 = mn[0] - M[3*i], d1 = mn[1] - M[3*i+1], d2 = mn[2] - M[3*i+2];
            mis += d0*d0 + d1*d1 + d2*d2;
            M[3*i] = mn[0]; M[3*i+1] = mn[1]; M[3*i+2] = mn[2];
        }
        inst_e2 = mis / n;
    }
    double maxM = 1.E-17, maxH = 1.E-17;
    for (int64_t i = 0; i < n; i++) {
        double am = sqrt(M[3*i]*M[3*i] + M[3*i+1]*M[3*i+1] + M[3*i+2]*M[3*i+2]);
        double ah = sqrt(H[3*i]*H[3*i] + H[3*i+1]*H[3*i+1] + H[3*i+2]*H[3*i+2]);
        if (am > maxM) maxM = am;
        if (ah > maxH) maxH = ah;
    }
    stats[0] = sqrt(inst_e2); stats[1] = maxM; stats[2] = maxH;
    return iter - 1;
}

/* ------------------------------------------------------------------------- */
/* Arc current segment (trapezoid rule, 'man' mode): radarccu.cpp:30-234      */
/* ------------------------------------------------------------------------- */
/* Circle centre c, radii [rmin, rmax], angles [phimin, phimax], height h,
 * azimuthal current density j, nseg sectors; local axis z. */
void ora_arccur_BH(const double P[3], const double c[3], double rmin, double rmax,
                   double phimin, double phimax, double height, double j_azim, int nseg,
                   double B[3], double H[3])
{
    const double Pi = 3.141592653589793238;
    const double zero_tol = 1.E-10;
    const double rel_zero_tol = 1.E-07;
    double px = P[0] - c[0], py = P[1] - c[1], pz = P[2] - c[2];
    const double small = 1.E-10;
    double r = sqrt(px*px + py*py + small);
    double phi = (py < 0) ? (2*Pi - acos(px / r)) : acos(px / r);
    double z1 = -height / 2. - pz;
    double z2 = z1 + height;
    double z1z1 = z1*z1, z2z2 = z2*z2, rr = r*r, r1r1 = rmin*rmin, r2r2 = rmax*rmax;
    double a11 = rr + r1r1 + z1z1, a12 = rr + r1r1 + z2z2, a21 = rr + r2r2 + z1z1, a22 = rr + r2r2 + z2z2;
    double two_rr1 = 2.*r*rmin, two_rr2 = 2.*r*rmax;
    int npt = nseg + 1;
    double step = (phimax - phimin + small) / (npt - 1);
    double dphi = phimin - phi + small;
    double sbx = 0, sby = 0, sbz = 0;
    for (int i = 0; i < npt; i++) {
        double cosd = cos(dphi), abssin = fabs(sin(dphi));
        double t1c = two_rr1 * cosd, t2c = two_rr2 * cosd;
        double bv11 = fabs(a11 - t1c); if (bv11 < zero_tol) bv11 = zero_tol;
        double D11 = sqrt(bv11);
        double D12 = sqrt(fabs(a12 - two_rr1 * cosd));
        double bv21 = fabs(a21 - t2c); if (bv21 < zero_tol) bv21 = zero_tol;
        double D21 = sqrt(bv21);
        double D22 = sqrt(fabs(a22 - two_rr2 * cosd));
        double rcos = r * cosd;
        double r1mrc = rmin - rcos, r2mrc = rmax - rcos;
        double tD11 = D11 * rel_zero_tol, tD12 = D12 * rel_zero_tol, tD21 = D21 * rel_zero_tol,
               tD22 = D22 * rel_zero_tol;
        double re2s2 = rr * abssin * abssin;
        double z1z1pre2 = z1z1 + re2s2;
        double r1e2 = r1mrc * r1mrc;
        double q11 = r1mrc + D11;
        if ((fabs(q11) < tD11) && (z1z1pre2 < r1e2 * rel_zero_tol)) q11 = 0.5 * z1z1pre2 / fabs(r1mrc);
        double q12 = r1mrc + D12;
        if ((fabs(q12) < tD12) && (z1z1pre2 < r1e2 * rel_zero_tol)) q12 = 0.5 * z1z1pre2 / fabs(r1mrc);
        double r2e2 = r2mrc * r2mrc;
        double q21 = r2mrc + D21;
        if ((fabs(q21) < tD21) && (z1z1pre2 < r2e2 * rel_zero_tol)) q21 = 0.5 * z1z1pre2 / fabs(r2mrc);
        double q22 = r2mrc + D22;
        if ((fabs(q22) < tD22) && (z1z1pre2 < r2e2 * rel_zero_tol)) q22 = 0.5 * z1z1pre2 / fabs(r2mrc);
        double ln1 = log(q11 / q21);
        double ln2 = log(q22 / q12);
        double rabssin = r * abssin;
        double ar1 = fabs(r1mrc) / rabssin, ar2 = fabs(r2mrc) / rabssin;
        double pm1 = 0, pm2 = 0;
        double sg1 = (r1mrc < 0.) ? -1. : 1., sg2 = (r2mrc < 0.) ? -1. : 1.;
        double ta1 = trans_atans(z2 * ar1 / D12, -z1 * ar1 / D11, &pm1);
        double ta2 = trans_atans(z1 * ar2 / D21, -z2 * ar2 / D22, &pm2);
        double buf = rabssin * (sg1 * (atan(ta1) + Pi * pm1) + sg2 * (atan(ta2) + Pi * pm2))
                   + z1 * ln1 + z2 * ln2;
        double z1pD11 = z1 + D11;
        double rr_r1 = rr + r1r1 - t1c;
        double z1z1t = z1z1 * rel_zero_tol, z2z2t = z2z2 * rel_zero_tol;
        if ((fabs(z1pD11) < tD11) && (rr_r1 < z1z1t)) z1pD11 = 0.5 * rr_r1 / fabs(z1);
        double z1pD21 = z1 + D21;
        double rr_r2 = rr + r2r2 - t2c;
        if ((fabs(z1pD21) < tD21) && (rr_r2 < z1z1t)) z1pD21 = 0.5 * rr_r2 / fabs(z1);
        double z2pD12 = z2 + D12;
        if ((fabs(z2pD12) < tD12) && (rr_r1 < z2z2t)) z2pD12 = 0.5 * rr_r1 / fabs(z2);
        double z2pD22 = z2 + D22;
        if ((fabs(z2pD22) < tD22) && (rr_r2 < z2z2t)) z2pD22 = 0.5 * rr_r2 / fabs(z2);
        double ra = z1pD11 / z2pD12, rb = z2pD22 / z1pD21;
        double phip = dphi + phi;
        double cp = cos(phip), sp = sin(phip);
        double fbxy = D11 - D21 + D22 - D12 + rcos * (ln1 + ln2);
        double fbz = buf - rcos * log(ra * rb);
        double w = ((i == 0) || (i == npt - 1)) ? 0.5 : 1.0;
        sbx += w * cp * fbxy;
        sby += w * sp * fbxy;
        sbz += w * fbz;
        dphi += step;
    }
    double fact = 0.0001 * j_azim;
    B[0] = fact * step * sbx; B[1] = fact * step * sby; B[2] = fact * step * sbz;
    H[0] = B[0]; H[1] = B[1]; H[2] = B[2];
}
