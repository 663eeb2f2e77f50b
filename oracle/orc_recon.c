/*
 * orc_recon.c -- modal basis and WENO reconstructions (oracle; test
 * infrastructure only).  Restates HyPE-2D/reconstruction.c with the same
 * expression order.  The gradient-only output u_coeffs of the reference is
 * used only under VISCOUS and is not restated.
 */
#include <math.h>
#include "orc_internal.h"

/* hype.h:57-62 */
static const double r1_6 = 1. / 6.;
static const double r13_3 = 13. / 3.;
static const double r7_6 = 7. / 6.;
static const double r11_120 = 11. / 120.;
static const double r1_12 = 1. / 12.;
static const double r41_60 = 41. / 60.;

/* reconstruction.c:11-47 */
double orc_basis(double x, double y, int n)
{
    switch (n) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return y;
    case 3: return x * x - 1. / 12.;
    case 4: return y * y - 1. / 12.;
    case 5: return x * y;
    case 6: return x * (x * x - 3. / 20.);
    case 7: return y * (y * y - 3. / 20.);
    case 8: return y * (x * x - 1. / 12.);
    case 9: return x * (y * y - 1. / 12.);
    default: return 0.0;
    }
}

/* reconstruction.c:53-100 */
void orc_basis_grad(double x, double y, int n, double *gx, double *gy)
{
    switch (n) {
    case 0: *gx = 0.0; *gy = 0.0; break;
    case 1: *gx = 1.0; *gy = 0.0; break;
    case 2: *gx = 0.0; *gy = 1.0; break;
    case 3: *gx = 2.0 * x; *gy = 0.0; break;
    case 4: *gx = 0.0; *gy = 2.0 * y; break;
    case 5: *gx = y; *gy = x; break;
    case 6: *gx = 3.0 * x * x - 3. / 20.; *gy = 0.0; break;
    case 7: *gx = 0.0; *gy = 3.0 * y * y - 3. / 20.; break;
    case 8: *gx = 2.0 * x * y; *gy = (x * x - 1. / 12.); break;
    case 9: *gx = (y * y - 1. / 12.); *gy = 2.0 * x * y; break;
    default: *gx = 0.0; *gy = 0.0;
    }
}

/* reconstruction.c:106-117 */
double orc_minmod(double a, double b)
{
    if (a * b < 0.0) return 0.0;
    if (fabs(a) < fabs(b)) return a;
    return b;
}

/* reconstruction.c:123-127 */
static double pow4(double a)
{
    double a2 = a * a;
    return a2 * a2;
}

/* reconstruction.c:129-142 (N=1: minmod TVD slopes) */
static void weno2(const double *U_x, const double *U_y, const double *U_xy,
                  double *coeffs)
{
    double u_0 = U_xy[0];
    double u_ip1 = U_x[3], u_jp1 = U_y[3];
    double u_im1 = U_x[1], u_jm1 = U_y[1];
    coeffs[0] = u_0;
    coeffs[1] = orc_minmod(u_ip1 - u_0, u_0 - u_im1);
    coeffs[2] = orc_minmod(u_jp1 - u_0, u_0 - u_jm1);
}

/* reconstruction.c:144-221 (N=2: third-order 2D WENO) */
static void weno3(const double *U_x, const double *U_y, const double *U_xy,
                  double *coeffs)
{
    const double central_cell_wt = 100.0;
    const double small_num = ORC_SMALL_NUM;
    double u_0 = U_xy[0];
    double u_ip1 = U_x[3], u_jp1 = U_y[3], u_ip1jp1 = U_xy[1];
    double u_im1 = U_x[1], u_jm1 = U_y[1], u_ip1jm1 = U_xy[2];
    double u_ip2 = U_x[4], u_jp2 = U_y[4], u_im1jp1 = U_xy[3];
    double u_im2 = U_x[0], u_jm2 = U_y[0], u_im1jm1 = U_xy[4];
    double u_x[3], u_xx[3], u_xy[4], wt[4], IS[4], total, temp;

    u_x[0] = -2.0 * u_im1 + 0.5 * u_im2 + 1.5 * u_0;
    u_xx[0] = 0.5 * (u_im2 - 2.0 * u_im1 + u_0);
    u_x[1] = 0.5 * (u_ip1 - u_im1);
    u_xx[1] = 0.5 * (u_im1 - 2.0 * u_0 + u_ip1);
    u_x[2] = -1.5 * u_0 + 2.0 * u_ip1 - 0.5 * u_ip2;
    u_xx[2] = 0.5 * (u_0 - 2.0 * u_ip1 + u_ip2);

    IS[0] = u_x[0] * u_x[0] + r13_3 * u_xx[0] * u_xx[0];
    wt[0] = 1.0 / (pow4(IS[0] + small_num));
    IS[1] = u_x[1] * u_x[1] + r13_3 * u_xx[1] * u_xx[1];
    wt[1] = central_cell_wt / (pow4(IS[1] + small_num));
    IS[2] = u_x[2] * u_x[2] + r13_3 * u_xx[2] * u_xx[2];
    wt[2] = 1.0 / (pow4(IS[2] + small_num));
    total = wt[0] + wt[1] + wt[2];
    wt[0] = wt[0] / total;
    wt[1] = wt[1] / total;
    wt[2] = wt[2] / total;

    coeffs[0] = u_0;
    coeffs[1] = wt[0] * u_x[0] + wt[1] * u_x[1] + wt[2] * u_x[2];
    coeffs[3] = wt[0] * u_xx[0] + wt[1] * u_xx[1] + wt[2] * u_xx[2];

    u_x[0] = -2.0 * u_jm1 + 0.5 * u_jm2 + 1.5 * u_0;
    u_xx[0] = 0.5 * (u_jm2 - 2.0 * u_jm1 + u_0);
    u_x[1] = 0.5 * (u_jp1 - u_jm1);
    u_xx[1] = 0.5 * (u_jm1 - 2.0 * u_0 + u_jp1);
    u_x[2] = -1.5 * u_0 + 2.0 * u_jp1 - 0.5 * u_jp2;
    u_xx[2] = 0.5 * (u_0 - 2.0 * u_jp1 + u_jp2);

    IS[0] = u_x[0] * u_x[0] + r13_3 * u_xx[0] * u_xx[0];
    wt[0] = 1.0 / (pow4(IS[0] + small_num));
    IS[1] = u_x[1] * u_x[1] + r13_3 * u_xx[1] * u_xx[1];
    wt[1] = central_cell_wt / (pow4(IS[1] + small_num));
    IS[2] = u_x[2] * u_x[2] + r13_3 * u_xx[2] * u_xx[2];
    wt[2] = 1.0 / (pow4(IS[2] + small_num));
    total = wt[0] + wt[1] + wt[2];
    wt[0] = wt[0] / total;
    wt[1] = wt[1] / total;
    wt[2] = wt[2] / total;

    coeffs[2] = wt[0] * u_x[0] + wt[1] * u_x[1] + wt[2] * u_x[2];
    coeffs[4] = wt[0] * u_xx[0] + wt[1] * u_xx[1] + wt[2] * u_xx[2];

    u_xy[0] = u_ip1jp1 - u_0 - coeffs[1] - coeffs[2] - coeffs[3] - coeffs[4];
    u_xy[1] = -u_ip1jm1 + u_0 + coeffs[1] - coeffs[2] + coeffs[3] + coeffs[4];
    u_xy[2] = -u_im1jp1 + u_0 - coeffs[1] + coeffs[2] + coeffs[3] + coeffs[4];
    u_xy[3] = u_im1jm1 - u_0 + coeffs[1] + coeffs[2] - coeffs[3] - coeffs[4];

    temp = 4.0 * (coeffs[3] * coeffs[3] + coeffs[4] * coeffs[4]);
    IS[0] = temp + u_xy[0] * u_xy[0];
    wt[0] = 1.0 / (pow4(IS[0] + small_num));
    IS[1] = temp + u_xy[1] * u_xy[1];
    wt[1] = 1.0 / (pow4(IS[1] + small_num));
    IS[2] = temp + u_xy[2] * u_xy[2];
    wt[2] = 1.0 / (pow4(IS[2] + small_num));
    IS[3] = temp + u_xy[3] * u_xy[3];
    wt[3] = 1.0 / (pow4(IS[3] + small_num));
    total = wt[0] + wt[1] + wt[2] + wt[3];
    wt[0] = wt[0] / total;
    wt[1] = wt[1] / total;
    wt[2] = wt[2] / total;
    wt[3] = wt[3] / total;

    coeffs[5] = wt[0] * u_xy[0] + wt[1] * u_xy[1] + wt[2] * u_xy[2] +
                wt[3] * u_xy[3];
}

/* reconstruction.c:223-357 (N=3: fourth-order CWENO, gamma_hi = gamma_lo = 0.85) */
static void weno4(const double *U_x, const double *U_y, const double *U_xy,
                  double *coeffs)
{
    const double small_num = ORC_SMALL_NUM;
    double gammaHi = 0.85;
    double gammaLo = 0.85;
    double u_0 = U_xy[0];
    double u_ip1 = U_x[3], u_jp1 = U_y[3], u_ip1jp1 = U_xy[1];
    double u_im1 = U_x[1], u_jm1 = U_y[1], u_ip1jm1 = U_xy[2];
    double u_ip2 = U_x[4], u_jp2 = U_y[4], u_im1jp1 = U_xy[3];
    double u_im2 = U_x[0], u_jm2 = U_y[0], u_im1jm1 = U_xy[4];
    double u_xR4, u_yR4, u_xxR4, u_yyR4, u_xyR4, u_xxxR4, u_yyyR4, u_xxyR4,
        u_xyyR4;
    double u_xR3[5], u_yR3[5], u_xxR3[5], u_yyR3[5], u_xyR3[5];
    double IS_R4, gamma_R4, w_R4;
    double IS_R3[5], gamma_R3[5], w_R3[5], sum = 0.0, wt_ratio;
    int i;

    gamma_R4 = gammaHi;
    gamma_R3[0] = (1.0 - gammaHi) * gammaLo;
    for (i = 1; i < 5; ++i)
        gamma_R3[i] = 0.25 * (1 - gammaHi) * (1 - gammaLo);

    u_xR4 = r41_60 * (-u_im1 + u_ip1) + r11_120 * (u_im2 - u_ip2);
    u_yR4 = r41_60 * (-u_jm1 + u_jp1) + r11_120 * (u_jm2 - u_jp2);
    u_xxR4 = -u_0 + 0.5 * (u_im1 + u_ip1);
    u_yyR4 = -u_0 + 0.5 * (u_jm1 + u_jp1);
    u_xyR4 = 0.25 * (u_im1jm1 - u_im1jp1 - u_ip1jm1 + u_ip1jp1);
    u_xxxR4 = r1_6 * (u_im1 - u_ip1) + r1_12 * (u_ip2 - u_im2);
    u_yyyR4 = r1_6 * (u_jm1 - u_jp1) + r1_12 * (u_jp2 - u_jm2);
    u_xxyR4 = 0.25 * (-u_im1jm1 + u_im1jp1 - u_ip1jm1 + u_ip1jp1) +
              0.5 * (u_jm1 - u_jp1);
    u_xyyR4 = 0.5 * (u_im1 - u_ip1) +
              0.25 * (u_ip1jp1 - u_im1jm1 - u_im1jp1 + u_ip1jm1);

    IS_R4 = (u_xR4 + 0.1 * u_xxxR4) * (u_xR4 + 0.1 * u_xxxR4) +
            r13_3 * (u_xxR4 * u_xxR4 + u_yyR4 * u_yyR4) +
            (u_yR4 + 0.1 * u_yyyR4) * (u_yR4 + 0.1 * u_yyyR4) +
            39.05 * (u_xxxR4 * u_xxxR4 + u_yyyR4 * u_yyyR4) +
            r7_6 * u_xyR4 * u_xyR4 +
            4.7 * (u_xxyR4 * u_xxyR4 + u_xyyR4 * u_xyyR4);
    w_R4 = gamma_R4 / (pow4(IS_R4 + small_num));
    sum = w_R4;

    u_xR3[0] = 0.5 * (-u_im1 + u_ip1);
    u_yR3[0] = 0.5 * (-u_jm1 + u_jp1);
    u_xxR3[0] = -u_0 + 0.5 * (u_im1 + u_ip1);
    u_yyR3[0] = -u_0 + 0.5 * (u_jm1 + u_jp1);
    u_xyR3[0] = 0.25 * (u_im1jm1 - u_im1jp1 - u_ip1jm1 + u_ip1jp1);

    u_xR3[1] = 1.5 * u_0 - 2.0 * u_im1 + 0.5 * u_im2;
    u_yR3[1] = 0.5 * (-u_jm1 + u_jp1);
    u_xxR3[1] = 0.5 * u_0 - u_im1 + 0.5 * u_im2;
    u_yyR3[1] = -u_0 + 0.5 * u_jm1 + 0.5 * u_jp1;
    u_xyR3[1] = 0.5 * (u_im1jm1 - u_im1jp1 - u_jm1 + u_jp1);

    u_xR3[2] = -1.5 * u_0 + 2.0 * u_ip1 - 0.5 * u_ip2;
    u_yR3[2] = 0.5 * (-u_jm1 + u_jp1);
    u_xxR3[2] = 0.5 * u_0 - u_ip1 + 0.5 * u_ip2;
    u_yyR3[2] = -u_0 + 0.5 * u_jm1 + 0.5 * u_jp1;
    u_xyR3[2] = 0.5 * (-u_ip1jm1 + u_ip1jp1 + u_jm1 - u_jp1);

    u_xR3[3] = 0.5 * (-u_im1 + u_ip1);
    u_yR3[3] = 1.5 * u_0 - 2.0 * u_jm1 + 0.5 * u_jm2;
    u_xxR3[3] = -u_0 + 0.5 * u_im1 + 0.5 * u_ip1;
    u_yyR3[3] = 0.5 * u_0 - u_jm1 + 0.5 * u_jm2;
    u_xyR3[3] = 0.5 * (-u_im1 + u_im1jm1 + u_ip1 - u_ip1jm1);

    u_xR3[4] = 0.5 * (-u_im1 + u_ip1);
    u_yR3[4] = -1.5 * u_0 + 2.0 * u_jp1 - 0.5 * u_jp2;
    u_xxR3[4] = -u_0 + 0.5 * u_im1 + 0.5 * u_ip1;
    u_yyR3[4] = 0.5 * u_0 - u_jp1 + 0.5 * u_jp2;
    u_xyR3[4] = 0.5 * (u_im1 - u_im1jp1 - u_ip1 + u_ip1jp1);

    for (i = 0; i < 5; ++i) {
        IS_R3[i] = u_xR3[i] * u_xR3[i] + u_yR3[i] * u_yR3[i] +
                   r13_3 * (u_xxR3[i] * u_xxR3[i] + u_yyR3[i] * u_yyR3[i]) +
                   r7_6 * u_xyR3[i] * u_xyR3[i];
        w_R3[i] = gamma_R3[i] / (pow4(IS_R3[i] + small_num));
        sum += w_R3[i];
    }
    w_R4 = w_R4 / sum;
    for (i = 0; i < 5; ++i)
        w_R3[i] = w_R3[i] / sum;
    wt_ratio = w_R4 / gamma_R4;

    coeffs[0] = u_0;
    coeffs[1] = wt_ratio * (u_xR4 - gamma_R3[0] * u_xR3[0] - gamma_R3[1] * u_xR3[1] -
                            gamma_R3[2] * u_xR3[2] - gamma_R3[3] * u_xR3[3] -
                            gamma_R3[4] * u_xR3[4]) +
                w_R3[0] * u_xR3[0] + w_R3[1] * u_xR3[1] + w_R3[2] * u_xR3[2] +
                w_R3[3] * u_xR3[3] + w_R3[4] * u_xR3[4];
    coeffs[2] = wt_ratio * (u_yR4 - gamma_R3[0] * u_yR3[0] - gamma_R3[1] * u_yR3[1] -
                            gamma_R3[2] * u_yR3[2] - gamma_R3[3] * u_yR3[3] -
                            gamma_R3[4] * u_yR3[4]) +
                w_R3[0] * u_yR3[0] + w_R3[1] * u_yR3[1] + w_R3[2] * u_yR3[2] +
                w_R3[3] * u_yR3[3] + w_R3[4] * u_yR3[4];
    coeffs[3] = wt_ratio * (u_xxR4 - gamma_R3[0] * u_xxR3[0] - gamma_R3[1] * u_xxR3[1] -
                            gamma_R3[2] * u_xxR3[2] - gamma_R3[3] * u_xxR3[3] -
                            gamma_R3[4] * u_xxR3[4]) +
                w_R3[0] * u_xxR3[0] + w_R3[1] * u_xxR3[1] + w_R3[2] * u_xxR3[2] +
                w_R3[3] * u_xxR3[3] + w_R3[4] * u_xxR3[4];
    coeffs[4] = wt_ratio * (u_yyR4 - gamma_R3[0] * u_yyR3[0] - gamma_R3[1] * u_yyR3[1] -
                            gamma_R3[2] * u_yyR3[2] - gamma_R3[3] * u_yyR3[3] -
                            gamma_R3[4] * u_yyR3[4]) +
                w_R3[0] * u_yyR3[0] + w_R3[1] * u_yyR3[1] + w_R3[2] * u_yyR3[2] +
                w_R3[3] * u_yyR3[3] + w_R3[4] * u_yyR3[4];
    coeffs[5] = wt_ratio * (u_xyR4 - gamma_R3[0] * u_xyR3[0] - gamma_R3[1] * u_xyR3[1] -
                            gamma_R3[2] * u_xyR3[2] - gamma_R3[3] * u_xyR3[3] -
                            gamma_R3[4] * u_xyR3[4]) +
                w_R3[0] * u_xyR3[0] + w_R3[1] * u_xyR3[1] + w_R3[2] * u_xyR3[2] +
                w_R3[3] * u_xyR3[3] + w_R3[4] * u_xyR3[4];
    coeffs[6] = wt_ratio * u_xxxR4;
    coeffs[7] = wt_ratio * u_yyyR4;
    coeffs[8] = wt_ratio * u_xxyR4;
    coeffs[9] = wt_ratio * u_xyyR4;
}

/* reconstruction.c:359-377 */
void orc_reconstruct(const double *ux, const double *uy, const double *uxy,
                     int N, double *coeffs)
{
    switch (N) {
    case 0: coeffs[0] = uxy[0]; break;
    case 1: weno2(ux, uy, uxy, coeffs); break;
    case 2: weno3(ux, uy, uxy, coeffs); break;
    case 3: weno4(ux, uy, uxy, coeffs); break;
    default: break;
    }
}
