/*
 * Oracle: finite-volume diffusion operator on a structured box, FP64.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Output format: 27 coefficients per cell (slot = (oz+1)*9 + (oy+1)*3 + (ox+1),
 * slot 13 = diagonal).  Neighbour directions of the reference (cell.h:21-49)
 * map to lexicographic offsets as FRONT/BACK = +z/-z, TOP/DOWN = +y/-y,
 * RIGHT/LEFT = +x/-x, compound directions being sums (measured with
 * ref_harness dirmap).  A neighbour "exists" when it lies inside the box and is
 * active (fill_discretization_matrix_elements_aniso, assembly_common.c:883-898;
 * fill_discretization_matrix_elements, :911-960).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "oracle.h"

/* ALPHA, monodomain/constants.h:10-11 (the 1e-8f literal is float) */
double orc_alpha(double beta, double cm, double dt, double dx, double dy, double dz) {
    return (((beta * cm) / dt) * 0.00000001f) * ((dx * dy) * dz);
}

static inline int slot_of(int ox, int oy, int oz) { return (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1); }

/* ------------------------------------------------------------------ */
/* isotropic: homogeneous_sigma_assembly_matrix (matrix_assembly.c:334-387) */
/* ------------------------------------------------------------------ */
void orc_assemble_iso(int nx, int ny, int nz, double dx, double dy, double dz, double alpha,
                      const double *sigma3, const uint8_t *active, double *coef27) {
    const int64_t n = (int64_t)nx * ny * nz;
    memset(coef27, 0, sizeof(double) * 27 * n);
    /* face order of the reference loop: BACK, FRONT, TOP, DOWN, RIGHT, LEFT */
    static const int face[6][3] = {{0, 0, -1}, {0, 0, 1}, {0, 1, 0}, {0, -1, 0}, {1, 0, 0}, {-1, 0, 0}};
    for(int64_t c = 0; c < n; c++) {
        if(active && !active[c]) continue;
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        double diag = alpha; /* initialize_diagonal_elements, assembly_common.c:4-39 */
        for(int f = 0; f < 6; f++) {
            const int xx = x + face[f][0], yy = y + face[f][1], zz = z + face[f][2];
            if(xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
            const int64_t nb = xx + (int64_t)nx * (yy + (int64_t)ny * zz);
            if(active && !active[nb]) continue;
            /* harmonic face conductivity, 0 if either side is 0 (:969-991); the pair
             * is evaluated in a canonical order so both rows get the same value */
            const int64_t lo = c < nb ? c : nb, hi = c < nb ? nb : c;
            const int axis = face[f][0] ? 0 : (face[f][1] ? 1 : 2);
            const double s1 = sigma3[3 * lo + axis], s2 = sigma3[3 * hi + axis];
            double s = 0.0;
            if(s1 != 0.0 && s2 != 0.0) s = (2.0f * s1 * s2) / (s1 + s2);
            /* fill_element (:41-76): area/distance multipliers */
            double mult;
            if(axis == 2) mult = ((dx * dy) / dz);
            else if(axis == 1) mult = ((dx * dz) / dy);
            else mult = ((dy * dz) / dx);
            coef27[27 * c + slot_of(face[f][0], face[f][1], face[f][2])] = -s * mult;
            diag += (s * mult);
        }
        coef27[27 * c + 13] = diag;
    }
}

/* ------------------------------------------------------------------ */
/* anisotropic: fill_elements_aniso (assembly_common.c:437-830)        */
/* ------------------------------------------------------------------ */

/* reference direction enum (cell.h:21-49) */
enum { FR, BK, TP, DN, RT, LF, FR_TP, FR_DN, FR_RT, FR_LF, FR_TP_RT, FR_TP_LF, FR_DN_RT, FR_DN_LF,
       BK_TP, BK_DN, BK_RT, BK_LF, BK_TP_RT, BK_TP_LF, BK_DN_RT, BK_DN_LF, TP_RT, TP_LF, DN_RT, DN_LF,
       NDIR };

static const int dir_off[NDIR][3] = {
    {0, 0, 1},  {0, 0, -1}, {0, 1, 0},   {0, -1, 0},  {1, 0, 0},    {-1, 0, 0},   {0, 1, 1},
    {0, -1, 1}, {1, 0, 1},  {-1, 0, 1},  {1, 1, 1},   {-1, 1, 1},   {1, -1, 1},   {-1, -1, 1},
    {0, 1, -1}, {0, -1, -1}, {1, 0, -1}, {-1, 0, -1}, {1, 1, -1},   {-1, 1, -1},  {1, -1, -1},
    {-1, -1, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0},  {-1, -1, 0}};

/* tensor component index in tensor6 */
enum { SXX = 0, SYY = 1, SZZ = 2, SXY = 3, SXZ = 4, SYZ = 5 };

/* offsets -> reference direction (or -1 for the centre) */
static int dir_of(int ox, int oy, int oz) {
    for(int d = 0; d < NDIR; d++)
        if(dir_off[d][0] == ox && dir_off[d][1] == oy && dir_off[d][2] == oz) return d;
    return -1;
}

/* calc_sigmas (assembly_common.c:115-412) for the flux leaving along `axis`
 * with sign `sgn`.  The two cross averages (CALC_PARTIAL_SIGMA, :78-113) run
 * over six neighbour pairs that straddle the flux line along the cross axis b
 * (offsets +1/-1 along b), visited as (a,c) = (0,0), (s,0), (0,+1), (s,+1),
 * (0,-1), (s,-1) where c is the remaining axis; the first average uses the
 * lower-numbered cross axis. */
static void flux_sigmas(const int nbx[NDIR], const double *t6, int64_t self, int axis, int sgn, double *s1,
                        double *s2, double *s3, int *c1, int *c2, int *c3) {
    *s1 = *s2 = *s3 = 0.0;
    int o[3] = {0, 0, 0};
    o[axis] = sgn;
    const int dface = dir_of(o[0], o[1], o[2]);
    if(nbx[dface] < 0) return;
    static const int comp_axis[3] = {SXX, SYY, SZZ};
    *s1 = (t6[6 * (int64_t)nbx[dface] + comp_axis[axis]] + t6[6 * self + comp_axis[axis]]) / 2.0;
    *c1 = 1;
    int others[2], k = 0;
    for(int a = 0; a < 3; a++)
        if(a != axis) others[k++] = a;
    for(int which = 0; which < 2; which++) {
        const int b = others[which], cax = others[1 - which];
        const int lo = axis < b ? axis : b, hi = axis < b ? b : axis;
        const int comp = (lo == 0 && hi == 1) ? SXY : ((lo == 0 && hi == 2) ? SXZ : SYZ);
        double sum = 0.0;
        int cnt = 0;
        static const int cseq[3] = {0, 1, -1};
        for(int ci = 0; ci < 3; ci++) {
            for(int ai = 0; ai < 2; ai++) {
                int p[3] = {0, 0, 0}, q[3] = {0, 0, 0};
                p[axis] = q[axis] = ai ? sgn : 0;
                p[cax] = q[cax] = cseq[ci];
                p[b] = 1;
                q[b] = -1;
                const int dp = dir_of(p[0], p[1], p[2]), dq = dir_of(q[0], q[1], q[2]);
                if(nbx[dp] >= 0 && nbx[dq] >= 0) {
                    sum += t6[6 * (int64_t)nbx[dp] + comp] + t6[6 * (int64_t)nbx[dq] + comp];
                    cnt += 2;
                }
            }
        }
        if(cnt) sum /= cnt;
        if(which == 0) { *s2 = sum; *c2 = cnt; }
        else { *s3 = sum; *c3 = cnt; }
    }
}

/* DIVIDE (assembly_common.c:414-419): the averages are divided by their count
 * a second time, as the reference does */
static inline double Q(double s, int c) { return c != 0 ? s / c : 0.0; }

void orc_assemble_aniso(int nx, int ny, int nz, double dx, double dy, double dz, double alpha,
                        const double *t6, const uint8_t *active, double *coef27) {
    const int64_t n = (int64_t)nx * ny * nz;
    memset(coef27, 0, sizeof(double) * 27 * n);
    const double yz_x = (dy * dz) / dx, xz_y = (dx * dz) / dy, xy_z = (dx * dy) / dz;

    for(int64_t c = 0; c < n; c++) {
        if(active && !active[c]) continue;
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
        int nbx[NDIR];
        for(int d = 0; d < NDIR; d++) {
            const int xx = x + dir_off[d][0], yy = y + dir_off[d][1], zz = z + dir_off[d][2];
            nbx[d] = -1;
            if(xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny || zz >= nz) continue;
            const int64_t k = xx + (int64_t)nx * (yy + (int64_t)ny * zz);
            if(active && !active[k]) continue;
            nbx[d] = (int)k;
        }
        /* the 18 face/cross conductivities (:486-496), names follow the reference:
         * sigma_x_r, sigma_xy_jx_r, sigma_xz_jx_r, ... */
        double x_r, xy_jx_r, xz_jx_r, x_l, xy_jx_l, xz_jx_l;
        double y_t, xy_jy_t, yz_jy_t, y_d, xy_jy_d, yz_jy_d;
        double z_f, xz_jz_f, yz_jz_f, z_b, xz_jz_b, yz_jz_b;
        int cx_r = 0, cxy_jx_r = 0, cxz_jx_r = 0, cx_l = 0, cxy_jx_l = 0, cxz_jx_l = 0;
        int cy_t = 0, cxy_jy_t = 0, cyz_jy_t = 0, cy_d = 0, cxy_jy_d = 0, cyz_jy_d = 0;
        int cz_f = 0, cxz_jz_f = 0, cyz_jz_f = 0, cz_b = 0, cxz_jz_b = 0, cyz_jz_b = 0;
        flux_sigmas(nbx, t6, c, 0, +1, &x_r, &xy_jx_r, &xz_jx_r, &cx_r, &cxy_jx_r, &cxz_jx_r);
        flux_sigmas(nbx, t6, c, 0, -1, &x_l, &xy_jx_l, &xz_jx_l, &cx_l, &cxy_jx_l, &cxz_jx_l);
        flux_sigmas(nbx, t6, c, 1, +1, &y_t, &xy_jy_t, &yz_jy_t, &cy_t, &cxy_jy_t, &cyz_jy_t);
        flux_sigmas(nbx, t6, c, 1, -1, &y_d, &xy_jy_d, &yz_jy_d, &cy_d, &cxy_jy_d, &cyz_jy_d);
        flux_sigmas(nbx, t6, c, 2, +1, &z_f, &xz_jz_f, &yz_jz_f, &cz_f, &cxz_jz_f, &cyz_jz_f);
        flux_sigmas(nbx, t6, c, 2, -1, &z_b, &xz_jz_b, &yz_jz_b, &cz_b, &cxz_jz_b, &cyz_jz_b);
        (void)cx_r; (void)cx_l; (void)cy_t; (void)cy_d; (void)cz_f; (void)cz_b;

        double acc[NDIR];
        for(int d = 0; d < NDIR; d++) acc[d] = 0.0;
        /* UPDATE_OR_ADD_ELEMENT (:421-435): add v to the neighbour's entry */
#define UPD(d, v)                                  \
    do {                                           \
        const double _v = (v);                     \
        if(_v != 0) acc[d] += _v;                  \
    } while(0)

        /* main diagonal (:499-500) */
        const double diag = alpha + (yz_x * x_r + yz_x * x_l + xz_y * y_t + xz_y * y_d + xy_z * z_f + xy_z * z_b);

        /* all neighbours, in enum order (:504-815) */
        for(int d = 0; d < NDIR; d++) {
            if(nbx[d] < 0) continue;
            double v = 0.0, s1, s2, s3;
            switch(d) {
            case FR:
                v += -z_f * xy_z;
                if(nbx[BK] >= 0)
                    v += -Q(yz_jy_t, cyz_jy_t) * dx + Q(yz_jy_d, cyz_jy_d) * dx - Q(xz_jx_r, cxz_jx_r) * dy + Q(xz_jx_l, cxz_jx_l) * dy;
                break;
            case BK:
                v += -z_b * xy_z;
                if(nbx[FR] >= 0)
                    v += Q(yz_jy_t, cyz_jy_t) * dx - Q(yz_jy_d, cyz_jy_d) * dx + Q(xz_jx_r, cxz_jx_r) * dy - Q(xz_jx_l, cxz_jx_l) * dy;
                break;
            case TP:
                v += -y_t * xz_y;
                if(nbx[DN] >= 0)
                    v += -Q(yz_jz_f, cyz_jz_f) * dx + Q(yz_jz_b, cyz_jz_b) * dx - Q(xy_jx_r, cxy_jx_r) * dz + Q(xy_jx_l, cxy_jx_l) * dz;
                break;
            case DN:
                v += -y_d * xz_y;
                if(nbx[TP] >= 0)
                    v += Q(yz_jz_f, cyz_jz_f) * dx - Q(yz_jz_b, cyz_jz_b) * dx + Q(xy_jx_r, cxy_jx_r) * dz - Q(xy_jx_l, cxy_jx_l) * dz;
                break;
            case RT:
                v += -x_r * yz_x;
                if(nbx[LF] >= 0)
                    v += -Q(xz_jz_f, cxz_jz_f) * dy + Q(xz_jz_b, cxz_jz_b) * dy - Q(xy_jy_t, cxy_jy_t) * dz + Q(xy_jy_d, cxy_jy_d) * dz;
                break;
            case LF:
                v += -x_l * yz_x;
                if(nbx[RT] >= 0)
                    v += Q(xz_jz_f, cxz_jz_f) * dy - Q(xz_jz_b, cxz_jz_b) * dy + Q(xy_jy_t, cxy_jy_t) * dz - Q(xy_jy_d, cxy_jy_d) * dz;
                break;

            case FR_TP:
                s1 = 0.0; s2 = 0.0;
                if(nbx[FR_DN] >= 0) {
                    s1 += -Q(xy_jx_r, cxy_jx_r) * dz + Q(xy_jx_l, cxy_jx_l) * dz - Q(yz_jz_f, cyz_jz_f) * dx;
                    UPD(FR_DN, -s1);
                }
                if(nbx[BK_TP] >= 0) {
                    s2 += -Q(yz_jy_t, cyz_jy_t) * dx - Q(xz_jx_r, cxz_jx_r) * dy + Q(xz_jx_l, cxz_jx_l) * dy;
                    UPD(BK_TP, -s2);
                }
                v += s1 + s2;
                break;
            case FR_DN:
                s1 = 0.0;
                if(nbx[BK_DN] >= 0) {
                    s1 += Q(yz_jy_d, cyz_jy_d) * dx - Q(xz_jx_r, cxz_jx_r) * dy + Q(xz_jx_l, cxz_jx_l) * dy;
                    UPD(FR_DN, s1);
                    UPD(BK_DN, -s1);
                }
                break;
            case BK_TP:
                s1 = 0.0;
                if(nbx[BK_DN] >= 0) {
                    s1 += Q(yz_jz_b, cyz_jz_b) * dx - Q(xy_jx_r, cxy_jx_r) * dz + Q(xy_jx_l, cxy_jx_l) * dz;
                    UPD(BK_TP, s1);
                    UPD(BK_DN, -s1);
                }
                break;
            case BK_DN:
                break;

            case FR_RT:
                s1 = 0.0; s2 = 0.0;
                if(nbx[BK_RT] >= 0) {
                    s1 += -Q(xz_jx_r, cxz_jx_r) * dy - Q(yz_jy_t, cyz_jy_t) * dx + Q(yz_jy_d, cyz_jy_d) * dx;
                    UPD(BK_RT, -s1);
                }
                if(nbx[FR_LF] >= 0) {
                    s2 += -Q(xz_jz_f, cxz_jz_f) * dy - Q(xy_jy_t, cxy_jy_t) * dz + Q(xy_jy_d, cxy_jy_d) * dz;
                    UPD(FR_LF, -s2);
                }
                v += s1 + s2;
                break;
            case FR_LF:
                s1 = 0.0;
                if(nbx[BK_LF] >= 0) {
                    s1 += -Q(yz_jy_t, cyz_jy_t) * dx + Q(yz_jy_d, cyz_jy_d) * dx + Q(xz_jx_l, cxz_jx_l) * dy;
                    UPD(FR_LF, s1);
                    UPD(BK_LF, -s1);
                }
                break;
            case BK_RT:
                s1 = 0.0;
                if(nbx[BK_LF] >= 0) {
                    s1 += Q(xz_jz_b, cxz_jz_b) * dy - Q(xy_jy_t, cxy_jy_t) * dz + Q(xy_jy_d, cxy_jy_d) * dz;
                    UPD(BK_RT, s1);
                    UPD(BK_LF, -s1);
                }
                break;
            case BK_LF:
                break;

            case TP_RT:
                s1 = 0; s2 = 0;
                if(nbx[DN_RT] >= 0) {
                    s1 += -Q(xy_jx_r, cxy_jx_r) * dz - Q(yz_jz_f, cyz_jz_f) * dx + Q(yz_jz_b, cyz_jz_b) * dx;
                    UPD(DN_RT, -s1);
                }
                if(nbx[TP_LF] >= 0) {
                    s2 += -Q(xz_jz_f, cxz_jz_f) * dy + Q(xz_jz_b, cxz_jz_b) * dy - Q(xy_jy_t, cxy_jy_t) * dz;
                    UPD(TP_LF, -s2);
                }
                v += s1 + s2;
                break;
            case TP_LF:
                s1 = 0;
                if(nbx[DN_LF] >= 0) {
                    s1 += -Q(yz_jz_f, cyz_jz_f) * dx + Q(yz_jz_b, cyz_jz_b) * dx + Q(xy_jx_l, cxy_jx_l) * dz;
                    UPD(TP_LF, s1);
                    UPD(DN_LF, -s1);
                }
                break;
            case DN_RT:
                s1 = 0;
                if(nbx[DN_LF] >= 0) {
                    s1 += -Q(xz_jz_f, cxz_jz_f) * dy + Q(xz_jz_b, cxz_jz_b) * dy + Q(xy_jy_d, cxy_jy_d) * dz;
                    UPD(DN_RT, s1);
                    UPD(DN_LF, -s1);
                }
                break;
            case DN_LF:
                break;

            case FR_TP_RT:
                s1 = 0; s2 = 0; s3 = 0;
                if(nbx[FR_DN_RT] >= 0) {
                    s1 += -Q(xy_jx_r, cxy_jx_r) * dz - Q(yz_jz_f, cyz_jz_f) * dx;
                    UPD(FR_DN_RT, -s1);
                }
                if(nbx[BK_TP_RT] >= 0) {
                    s2 += -Q(xz_jx_r, cxz_jx_r) * dy - Q(yz_jy_t, cyz_jy_t) * dx;
                    UPD(BK_TP_RT, -s2);
                }
                if(nbx[FR_TP_LF] >= 0) {
                    s3 += -Q(xz_jz_f, cxz_jz_f) * dy - Q(xy_jy_t, cxy_jy_t) * dz;
                    UPD(FR_TP_LF, -s3);
                }
                v += s1 + s2 + s3;
                break;
            case FR_TP_LF:
                s1 = 0; s2 = 0;
                if(nbx[FR_DN_LF] >= 0) {
                    s1 += Q(xy_jx_l, cxy_jx_l) * dz - Q(yz_jz_f, cyz_jz_f) * dx;
                    UPD(FR_TP_LF, s1);
                    UPD(FR_DN_LF, -s1);
                }
                if(nbx[BK_TP_LF] >= 0) {
                    s2 += Q(xz_jx_l, cxz_jx_l) * dy - Q(yz_jy_t, cyz_jy_t) * dx;
                    UPD(FR_TP_LF, s2);
                    UPD(BK_TP_LF, -s2);
                }
                break;
            case FR_DN_RT:
                s1 = 0; s2 = 0;
                if(nbx[BK_DN_RT] >= 0) {
                    s1 += -Q(xz_jx_r, cxz_jx_r) * dy + Q(yz_jy_d, cyz_jy_d) * dx;
                    UPD(FR_DN_RT, s1);
                    UPD(BK_DN_RT, -s1);
                }
                if(nbx[FR_DN_LF] >= 0) {
                    s2 += Q(xy_jy_d, cxy_jy_d) * dz - Q(xz_jz_f, cxz_jz_f) * dy;
                    UPD(FR_DN_RT, s2);
                    UPD(FR_DN_LF, -s2);
                }
                break;
            case FR_DN_LF:
                s1 = 0;
                if(nbx[BK_DN_LF] >= 0) {
                    s1 += Q(xz_jx_l, cxz_jx_l) * dy + Q(yz_jy_d, cyz_jy_d) * dx;
                    UPD(FR_DN_LF, s1);
                    UPD(BK_DN_LF, -s1);
                }
                break;
            case BK_TP_RT:
                s1 = 0; s2 = 0;
                if(nbx[BK_DN_RT] >= 0) {
                    s1 += -Q(xy_jx_r, cxy_jx_r) * dz + Q(yz_jz_b, cyz_jz_b) * dx;
                    UPD(BK_TP_RT, s1);
                    UPD(BK_DN_RT, -s1);
                }
                if(nbx[BK_TP_LF] >= 0) {
                    s2 += -Q(xy_jy_t, cxy_jy_t) * dz + Q(xz_jz_b, cxz_jz_b) * dy;
                    UPD(BK_TP_RT, s2);
                    UPD(BK_TP_LF, -s2);
                }
                break;
            case BK_TP_LF:
                s1 = 0;
                if(nbx[BK_DN_LF] >= 0) {
                    s1 += Q(xy_jx_l, cxy_jx_l) * dz + Q(yz_jz_b, cyz_jz_b) * dx;
                    UPD(BK_TP_LF, s1);
                    UPD(BK_DN_LF, -s1);
                }
                break;
            case BK_DN_RT:
                s1 = 0;
                if(nbx[BK_DN_LF] >= 0) {
                    s1 += Q(xy_jy_d, cxy_jy_d) * dz + Q(xz_jz_b, cxz_jz_b) * dy;
                    UPD(BK_DN_RT, s1);
                    UPD(BK_DN_LF, -s1);
                }
                break;
            default:
                break;
            }
            /* the element for d itself (:817-822) */
            if(v != 0.0) acc[d] += v;
        }
#undef UPD
        double *row = coef27 + 27 * c;
        row[13] = diag;
        for(int d = 0; d < NDIR; d++) row[slot_of(dir_off[d][0], dir_off[d][1], dir_off[d][2])] = acc[d];
    }
}

/* read_fibers' normalisation (assembly_common.c:1141-1146) */
void orc_normalise(double f[3]) {
    const double m = sqrt((f[0] * f[0]) + (f[1] * f[1]) + (f[2] * f[2]));
    f[0] = f[0] / m;
    f[1] = f[1] / m;
    f[2] = f[2] / m;
}

/* calc_tensor2 (assembly_common.c:1121-1139, helpers :1062-1090):
 * D = (sl - st) f f^T + st I, out6 = (sxx, syy, szz, sxy, sxz, syz) */
void orc_fibre_tensor(const double f[3], double sl, double st, double out6[6]) {
    double D[3][3];
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++) D[i][j] = (f[i] * f[j]) * (sl - st) + (i == j ? st : 0.0);
    out6[0] = D[0][0];
    out6[1] = D[1][1];
    out6[2] = D[2][2];
    out6[3] = D[0][1];
    out6[4] = D[0][2];
    out6[5] = D[1][2];
}
