// Device assembly of the finite-volume diffusion operator and its compression.
//
// Isotropic rule: homogeneous_sigma_assembly_matrix (matrix_assembly.c:334-387)
// through fill_discretization_matrix_elements / fill_element
// (assembly_common.c:41-76,911-1048): face conductivity = harmonic mean of the two
// cells (0 if either is 0), off-diagonal = -sigma * area / distance, diagonal =
// alpha + sum of the face terms.
// Anisotropic rule: anisotropic_sigma_assembly_matrix (matrix_assembly.c:389-514)
// through calc_sigmas / fill_elements_aniso (assembly_common.c:115-830): face
// conductivities are arithmetic means; cross-derivative terms are averages over
// the available neighbour pairs, divided once more by their count (DIVIDE,
// :414-419), spread onto edge and corner neighbours in the reference's
// direction order.  Neighbour directions map to offsets FRONT/BACK = +z/-z,
// TOP/DOWN = +y/-y, RIGHT/LEFT = +x/-x (cell.h:21-49).
// A neighbour exists when it lies in the box and is active.
// coefficients are formed exactly as the reference forms them: no FMA
// contraction, also in the mask-row function of aniso_row.h, whose FP64 form
// is checked here bit for bit against the assembled rows (the pragma precedes
// the includes so that it covers the header's function bodies)
#pragma clang fp contract(off)

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "aniso_row.h"
#include "mono_common.h"
#include "operator.h"

namespace mono {

// reference direction enum order (cell.h:21-49) and offsets
enum Dir { FR, BK, TP, DN, RT, LF, FR_TP, FR_DN, FR_RT, FR_LF, FR_TP_RT, FR_TP_LF, FR_DN_RT, FR_DN_LF,
           BK_TP, BK_DN, BK_RT, BK_LF, BK_TP_RT, BK_TP_LF, BK_DN_RT, BK_DN_LF, TP_RT, TP_LF, DN_RT, DN_LF,
           NDIR };

__constant__ int8_t c_dir_off[NDIR][3] = {
    {0, 0, 1},  {0, 0, -1}, {0, 1, 0},   {0, -1, 0},  {1, 0, 0},    {-1, 0, 0},   {0, 1, 1},
    {0, -1, 1}, {1, 0, 1},  {-1, 0, 1},  {1, 1, 1},   {-1, 1, 1},   {1, -1, 1},   {-1, -1, 1},
    {0, 1, -1}, {0, -1, -1}, {1, 0, -1}, {-1, 0, -1}, {1, 1, -1},   {-1, 1, -1},  {1, -1, -1},
    {-1, -1, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0},  {-1, -1, 0}};

__device__ __forceinline__ int dir_index(int ox, int oy, int oz) {
    // inverse of c_dir_off
    const int key = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
    // explicit mapping (z, y, x) -> direction
    switch(key) {
    case 0: return BK_DN_LF;
    case 1: return BK_DN;
    case 2: return BK_DN_RT;
    case 3: return BK_LF;
    case 4: return BK;
    case 5: return BK_RT;
    case 6: return BK_TP_LF;
    case 7: return BK_TP;
    case 8: return BK_TP_RT;
    case 9: return DN_LF;
    case 10: return DN;
    case 11: return DN_RT;
    case 12: return LF;
    case 13: return -1;
    case 14: return RT;
    case 15: return TP_LF;
    case 16: return TP;
    case 17: return TP_RT;
    case 18: return FR_DN_LF;
    case 19: return FR_DN;
    case 20: return FR_DN_RT;
    case 21: return FR_LF;
    case 22: return FR;
    case 23: return FR_RT;
    case 24: return FR_TP_LF;
    case 25: return FR_TP;
    default: return FR_TP_RT;
    }
}

struct Tens {
    double xx, yy, zz, xy, xz, yz;
};

__device__ __forceinline__ Tens tensor_of(const AssemblyParams &p, int64_t c, int z) {
    const double *t;
    if(p.tensor_source == kTensorUniform) t = p.tensor;
    else if(p.tensor_source == kTensorPerLayer) t = p.tensor + 6 * (int64_t)z;
    else t = p.tensor + 6 * c;
    Tens r{t[0], t[1], t[2], t[3], t[4], t[5]};
    if(p.g) {
        const double g = p.g[c];
        r.xx *= g; r.yy *= g; r.zz *= g; r.xy *= g; r.xz *= g; r.yz *= g;
    }
    return r;
}

__device__ __forceinline__ double comp(const Tens &t, int k) {
    switch(k) {
    case 0: return t.xx;
    case 1: return t.yy;
    case 2: return t.zz;
    case 3: return t.xy;
    case 4: return t.xz;
    default: return t.yz;
    }
}

// calc_sigmas (assembly_common.c:115-412) for the flux leaving along `axis`
// with sign `sgn`: s1 = mean of the two cells' diagonal component; the two
// cross averages (CALC_PARTIAL_SIGMA, :78-113) visit six neighbour pairs that
// straddle the flux line along the cross axis b (offsets +1/-1 along b), in the
// order (a,c) = (0,0), (s,0), (0,+1), (s,+1), (0,-1), (s,-1) where a is the
// flux axis and c the remaining axis; the first average uses the lower-numbered
// cross axis.
__device__ void flux_sigmas(const AssemblyParams &p, const int64_t nbi[NDIR], const int nbz[NDIR], int64_t self, int zself,
                            int axis, int sgn, double &s1, double &s2, double &s3, int &c2, int &c3) {
    s1 = s2 = s3 = 0.0;
    int o[3] = {0, 0, 0};
    o[axis] = sgn;
    const int dface = dir_index(o[0], o[1], o[2]);
    if(nbi[dface] < 0) return;
    s1 = (comp(tensor_of(p, nbi[dface], nbz[dface]), axis) + comp(tensor_of(p, self, zself), axis)) / 2.0;
    int others[2], k = 0;
    for(int a = 0; a < 3; a++)
        if(a != axis) others[k++] = a;
    for(int which = 0; which < 2; which++) {
        const int b = others[which], cax = others[1 - which];
        const int lo = axis < b ? axis : b, hi = axis < b ? b : axis;
        const int cmp = (lo == 0 && hi == 1) ? 3 : ((lo == 0 && hi == 2) ? 4 : 5);
        double sum = 0.0;
        int cnt = 0;
        const int cseq[3] = {0, 1, -1};
        for(int ci = 0; ci < 3; ci++) {
            for(int ai = 0; ai < 2; ai++) {
                int pp[3] = {0, 0, 0}, qq[3] = {0, 0, 0};
                pp[axis] = qq[axis] = ai ? sgn : 0;
                pp[cax] = qq[cax] = cseq[ci];
                pp[b] = 1;
                qq[b] = -1;
                const int dp = dir_index(pp[0], pp[1], pp[2]), dq = dir_index(qq[0], qq[1], qq[2]);
                if(nbi[dp] >= 0 && nbi[dq] >= 0) {
                    sum += comp(tensor_of(p, nbi[dp], nbz[dp]), cmp) + comp(tensor_of(p, nbi[dq], nbz[dq]), cmp);
                    cnt += 2;
                }
            }
        }
        if(cnt) sum /= cnt;
        if(which == 0) { s2 = sum; c2 = cnt; }
        else { s3 = sum; c3 = cnt; }
    }
}

__device__ __forceinline__ double Qd(double s, int c) { return c != 0 ? s / c : 0.0; }

__global__ __launch_bounds__(256) void k_assemble(AssemblyParams p, double *coef27) {
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    double row[27];
    for(int k = 0; k < 27; k++) row[k] = 0.0;
    const bool act = !p.active || p.active[c];
    if(act) {
        const int x = (int)(c % p.nx), y = (int)((c / p.nx) % p.ny), z = (int)(c / ((int64_t)p.nx * p.ny));
        int64_t nbi[NDIR];
        int nbz[NDIR];
        for(int d = 0; d < NDIR; d++) {
            const int xx = x + c_dir_off[d][0], yy = y + c_dir_off[d][1], zz = z + c_dir_off[d][2];
            nbi[d] = -1;
            nbz[d] = zz;
            if(xx < 0 || yy < 0 || zz < 0 || xx >= p.nx || yy >= p.ny || zz >= p.nz) continue;
            const int64_t k = xx + (int64_t)p.nx * (yy + (int64_t)p.ny * zz);
            if(p.active && !p.active[k]) continue;
            nbi[d] = k;
        }
        const double dx = p.dx, dy = p.dy, dz = p.dz;
        if(!p.aniso) {
            // faces in the reference loop order BACK, FRONT, TOP, DOWN, RIGHT, LEFT
            const int faces[6] = {BK, FR, TP, DN, RT, LF};
            double diag = p.alpha;
            const Tens ts = tensor_of(p, c, z);
            for(int f = 0; f < 6; f++) {
                const int d = faces[f];
                if(nbi[d] < 0) continue;
                const int axis = c_dir_off[d][0] ? 0 : (c_dir_off[d][1] ? 1 : 2);
                const Tens tn = tensor_of(p, nbi[d], nbz[d]);
                // pair evaluated in canonical (lower index first) order
                const double sa = nbi[d] < c ? comp(tn, axis) : comp(ts, axis);
                const double sb = nbi[d] < c ? comp(ts, axis) : comp(tn, axis);
                double s = 0.0;
                if(sa != 0.0 && sb != 0.0) s = (2.0f * sa * sb) / (sa + sb);
                const double mult = axis == 2 ? ((dx * dy) / dz) : (axis == 1 ? ((dx * dz) / dy) : ((dy * dz) / dx));
                row[slot_of(c_dir_off[d][0], c_dir_off[d][1], c_dir_off[d][2])] = -s * mult;
                diag += (s * mult);
            }
            row[13] = diag;
        } else {
            const double yz_x = (dy * dz) / dx, xz_y = (dx * dz) / dy, xy_z = (dx * dy) / dz;
            double x_r, xy_jx_r, xz_jx_r, x_l, xy_jx_l, xz_jx_l, y_t, xy_jy_t, yz_jy_t, y_d, xy_jy_d, yz_jy_d;
            double z_f, xz_jz_f, yz_jz_f, z_b, xz_jz_b, yz_jz_b;
            int cxy_jx_r = 0, cxz_jx_r = 0, cxy_jx_l = 0, cxz_jx_l = 0, cxy_jy_t = 0, cyz_jy_t = 0;
            int cxy_jy_d = 0, cyz_jy_d = 0, cxz_jz_f = 0, cyz_jz_f = 0, cxz_jz_b = 0, cyz_jz_b = 0;
            flux_sigmas(p, nbi, nbz, c, z, 0, +1, x_r, xy_jx_r, xz_jx_r, cxy_jx_r, cxz_jx_r);
            flux_sigmas(p, nbi, nbz, c, z, 0, -1, x_l, xy_jx_l, xz_jx_l, cxy_jx_l, cxz_jx_l);
            flux_sigmas(p, nbi, nbz, c, z, 1, +1, y_t, xy_jy_t, yz_jy_t, cxy_jy_t, cyz_jy_t);
            flux_sigmas(p, nbi, nbz, c, z, 1, -1, y_d, xy_jy_d, yz_jy_d, cxy_jy_d, cyz_jy_d);
            flux_sigmas(p, nbi, nbz, c, z, 2, +1, z_f, xz_jz_f, yz_jz_f, cxz_jz_f, cyz_jz_f);
            flux_sigmas(p, nbi, nbz, c, z, 2, -1, z_b, xz_jz_b, yz_jz_b, cxz_jz_b, cyz_jz_b);
            // the twelve doubly divided cross terms
            const double Xyr = Qd(xy_jx_r, cxy_jx_r), Xzr = Qd(xz_jx_r, cxz_jx_r);
            const double Xyl = Qd(xy_jx_l, cxy_jx_l), Xzl = Qd(xz_jx_l, cxz_jx_l);
            const double Yxt = Qd(xy_jy_t, cxy_jy_t), Yzt = Qd(yz_jy_t, cyz_jy_t);
            const double Yxd = Qd(xy_jy_d, cxy_jy_d), Yzd = Qd(yz_jy_d, cyz_jy_d);
            const double Zxf = Qd(xz_jz_f, cxz_jz_f), Zyf = Qd(yz_jz_f, cyz_jz_f);
            const double Zxb = Qd(xz_jz_b, cxz_jz_b), Zyb = Qd(yz_jz_b, cyz_jz_b);

            double acc[NDIR];
            for(int d = 0; d < NDIR; d++) acc[d] = 0.0;
#define UPD(d, v)                      \
    do {                               \
        const double _v = (v);         \
        if(_v != 0) acc[d] += _v;      \
    } while(0)
            const double diag = p.alpha + (yz_x * x_r + yz_x * x_l + xz_y * y_t + xz_y * y_d + xy_z * z_f + xy_z * z_b);
            for(int d = 0; d < NDIR; d++) {
                if(nbi[d] < 0) continue;
                double v = 0.0, s1, s2, s3;
                switch(d) {
                case FR:
                    v += -z_f * xy_z;
                    if(nbi[BK] >= 0) v += -Yzt * dx + Yzd * dx - Xzr * dy + Xzl * dy;
                    break;
                case BK:
                    v += -z_b * xy_z;
                    if(nbi[FR] >= 0) v += Yzt * dx - Yzd * dx + Xzr * dy - Xzl * dy;
                    break;
                case TP:
                    v += -y_t * xz_y;
                    if(nbi[DN] >= 0) v += -Zyf * dx + Zyb * dx - Xyr * dz + Xyl * dz;
                    break;
                case DN:
                    v += -y_d * xz_y;
                    if(nbi[TP] >= 0) v += Zyf * dx - Zyb * dx + Xyr * dz - Xyl * dz;
                    break;
                case RT:
                    v += -x_r * yz_x;
                    if(nbi[LF] >= 0) v += -Zxf * dy + Zxb * dy - Yxt * dz + Yxd * dz;
                    break;
                case LF:
                    v += -x_l * yz_x;
                    if(nbi[RT] >= 0) v += Zxf * dy - Zxb * dy + Yxt * dz - Yxd * dz;
                    break;
                case FR_TP:
                    s1 = 0.0; s2 = 0.0;
                    if(nbi[FR_DN] >= 0) { s1 += -Xyr * dz + Xyl * dz - Zyf * dx; UPD(FR_DN, -s1); }
                    if(nbi[BK_TP] >= 0) { s2 += -Yzt * dx - Xzr * dy + Xzl * dy; UPD(BK_TP, -s2); }
                    v += s1 + s2;
                    break;
                case FR_DN:
                    s1 = 0.0;
                    if(nbi[BK_DN] >= 0) { s1 += Yzd * dx - Xzr * dy + Xzl * dy; UPD(FR_DN, s1); UPD(BK_DN, -s1); }
                    break;
                case BK_TP:
                    s1 = 0.0;
                    if(nbi[BK_DN] >= 0) { s1 += Zyb * dx - Xyr * dz + Xyl * dz; UPD(BK_TP, s1); UPD(BK_DN, -s1); }
                    break;
                case FR_RT:
                    s1 = 0.0; s2 = 0.0;
                    if(nbi[BK_RT] >= 0) { s1 += -Xzr * dy - Yzt * dx + Yzd * dx; UPD(BK_RT, -s1); }
                    if(nbi[FR_LF] >= 0) { s2 += -Zxf * dy - Yxt * dz + Yxd * dz; UPD(FR_LF, -s2); }
                    v += s1 + s2;
                    break;
                case FR_LF:
                    s1 = 0.0;
                    if(nbi[BK_LF] >= 0) { s1 += -Yzt * dx + Yzd * dx + Xzl * dy; UPD(FR_LF, s1); UPD(BK_LF, -s1); }
                    break;
                case BK_RT:
                    s1 = 0.0;
                    if(nbi[BK_LF] >= 0) { s1 += Zxb * dy - Yxt * dz + Yxd * dz; UPD(BK_RT, s1); UPD(BK_LF, -s1); }
                    break;
                case TP_RT:
                    s1 = 0; s2 = 0;
                    if(nbi[DN_RT] >= 0) { s1 += -Xyr * dz - Zyf * dx + Zyb * dx; UPD(DN_RT, -s1); }
                    if(nbi[TP_LF] >= 0) { s2 += -Zxf * dy + Zxb * dy - Yxt * dz; UPD(TP_LF, -s2); }
                    v += s1 + s2;
                    break;
                case TP_LF:
                    s1 = 0;
                    if(nbi[DN_LF] >= 0) { s1 += -Zyf * dx + Zyb * dx + Xyl * dz; UPD(TP_LF, s1); UPD(DN_LF, -s1); }
                    break;
                case DN_RT:
                    s1 = 0;
                    if(nbi[DN_LF] >= 0) { s1 += -Zxf * dy + Zxb * dy + Yxd * dz; UPD(DN_RT, s1); UPD(DN_LF, -s1); }
                    break;
                case FR_TP_RT:
                    s1 = 0; s2 = 0; s3 = 0;
                    if(nbi[FR_DN_RT] >= 0) { s1 += -Xyr * dz - Zyf * dx; UPD(FR_DN_RT, -s1); }
                    if(nbi[BK_TP_RT] >= 0) { s2 += -Xzr * dy - Yzt * dx; UPD(BK_TP_RT, -s2); }
                    if(nbi[FR_TP_LF] >= 0) { s3 += -Zxf * dy - Yxt * dz; UPD(FR_TP_LF, -s3); }
                    v += s1 + s2 + s3;
                    break;
                case FR_TP_LF:
                    s1 = 0; s2 = 0;
                    if(nbi[FR_DN_LF] >= 0) { s1 += Xyl * dz - Zyf * dx; UPD(FR_TP_LF, s1); UPD(FR_DN_LF, -s1); }
                    if(nbi[BK_TP_LF] >= 0) { s2 += Xzl * dy - Yzt * dx; UPD(FR_TP_LF, s2); UPD(BK_TP_LF, -s2); }
                    break;
                case FR_DN_RT:
                    s1 = 0; s2 = 0;
                    if(nbi[BK_DN_RT] >= 0) { s1 += -Xzr * dy + Yzd * dx; UPD(FR_DN_RT, s1); UPD(BK_DN_RT, -s1); }
                    if(nbi[FR_DN_LF] >= 0) { s2 += Yxd * dz - Zxf * dy; UPD(FR_DN_RT, s2); UPD(FR_DN_LF, -s2); }
                    break;
                case FR_DN_LF:
                    s1 = 0;
                    if(nbi[BK_DN_LF] >= 0) { s1 += Xzl * dy + Yzd * dx; UPD(FR_DN_LF, s1); UPD(BK_DN_LF, -s1); }
                    break;
                case BK_TP_RT:
                    s1 = 0; s2 = 0;
                    if(nbi[BK_DN_RT] >= 0) { s1 += -Xyr * dz + Zyb * dx; UPD(BK_TP_RT, s1); UPD(BK_DN_RT, -s1); }
                    if(nbi[BK_TP_LF] >= 0) { s2 += -Yxt * dz + Zxb * dy; UPD(BK_TP_RT, s2); UPD(BK_TP_LF, -s2); }
                    break;
                case BK_TP_LF:
                    s1 = 0;
                    if(nbi[BK_DN_LF] >= 0) { s1 += Xyl * dz + Zyb * dx; UPD(BK_TP_LF, s1); UPD(BK_DN_LF, -s1); }
                    break;
                case BK_DN_RT:
                    s1 = 0;
                    if(nbi[BK_DN_LF] >= 0) { s1 += Yxd * dz + Zxb * dy; UPD(BK_DN_RT, s1); UPD(BK_DN_LF, -s1); }
                    break;
                default:
                    break;
                }
                if(v != 0.0) acc[d] += v;
            }
#undef UPD
            row[13] = diag;
            for(int d = 0; d < NDIR; d++) row[slot_of(c_dir_off[d][0], c_dir_off[d][1], c_dir_off[d][2])] = acc[d];
        }
    }
    for(int k = 0; k < 27; k++) coef27[(int64_t)k * n + c] = row[k];
}

hipError_t launch_assemble(const AssemblyParams &p, double *coef27, hipStream_t s) {
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    if(n == 0) return hipSuccess;
    k_assemble<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, coef27);
    return hipGetLastError();
}

// class of a cell: z layer x (x position class) x (y position class)
__host__ __device__ __forceinline__ int pos_class(int i, int len) { return i == 0 ? 0 : (i == len - 1 ? 2 : 1); }

// Count rows that differ (bitwise) from their class row.
__global__ __launch_bounds__(256) void k_verify_classes(const double *coef27, int64_t n, int nx, int ny, const double *table27,
                                                        unsigned long long *mismatch) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
    const int cls = (z * 3 + pos_class(y, ny)) * 3 + pos_class(x, nx);
    bool bad = false;
    for(int k = 0; k < 27; k++) {
        const double a = coef27[(int64_t)k * n + c], b = table27[cls * 27 + k];
        if(__double_as_longlong(a) != __double_as_longlong(b)) bad = true;
    }
    if(bad) atomicAdd(mismatch, 1ull);
}

// Gather the stored slots into SoA FP64 and FP32 arrays (slots < 0: padding).
__global__ __launch_bounds__(256) void k_compact(const double *coef27, int64_t n, int S, const int *slots, double *out64,
                                                 float *out32) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    for(int k = 0; k < S; k++) {
        const double v = slots[k] >= 0 ? coef27[(int64_t)slots[k] * n + c] : 0.0;
        out64[(int64_t)k * n + c] = v;
        if(out32) out32[(int64_t)k * n + c] = (float)v;
    }
}

// rows of the representative cells, [ncls][27]
__global__ void k_gather_rows(const double *coef27, int64_t n, const int64_t *rep, int ncls, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= ncls * 27) return;
    const int c = i / 27, k = i % 27;
    out[i] = rep[c] >= 0 ? coef27[(int64_t)k * n + rep[c]] : 0.0;
}

// union of non-zero slots over all rows
__global__ __launch_bounds__(256) void k_slot_mask(const double *coef27, int64_t n, unsigned *mask) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned m = 0;
    if(c < n) {
        for(int k = 0; k < 27; k++)
            if(coef27[(int64_t)k * n + c] != 0.0) m |= 1u << k;
    }
    // wave-level OR before the atomic
    for(int off = 32; off > 0; off >>= 1) m |= __shfl_xor(m, off);
    if((threadIdx.x & 63) == 0 && m) atomicOr(mask, m);
}

// Row code of each cell (see kCodes): own activity and, for each 7-point
// neighbour, whether it exists in the box and is active.
__global__ __launch_bounds__(256) void k_row_codes(int nx, int ny, int nz, const uint8_t *active, uint8_t *codes) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
    auto on = [&](int xx, int yy, int zz) -> unsigned {
        if(xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) return 0u;
        return !active || active[xx + (int64_t)nx * (yy + (int64_t)ny * zz)] ? 1u : 0u;
    };
    unsigned code = on(x, y, z);
    if(code) {
#pragma unroll
        for(int k = 1; k < 7; k++) {
            const int8_t *o = RowPattern<7>::o[k];
            code |= on(x + o[0], y + o[1], z + o[2]) << k;
        }
    }
    codes[c] = (uint8_t)code;
}

// first cell of each code (reps start at all ones = absent)
__global__ __launch_bounds__(256) void k_code_reps(const uint8_t *codes, int64_t n, unsigned long long *reps) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    const unsigned code = codes[c];
    if((unsigned long long)c < reps[code]) atomicMin(&reps[code], (unsigned long long)c);
}

__global__ __launch_bounds__(256) void k_verify_codes(const double *coef27, int64_t n, const uint8_t *codes, const double *table27,
                                                      unsigned long long *mismatch) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    const int code = codes[c];
    bool bad = false;
    for(int k = 0; k < 27; k++) {
        const double a = coef27[(int64_t)k * n + c], b = table27[code * 27 + k];
        if(__double_as_longlong(a) != __double_as_longlong(b)) bad = true;
    }
    if(bad) atomicAdd(mismatch, 1ull);
}

// Try code mode for a 7-point operator; returns 1 when every row matched its
// code's table row (op->codes and the tables set), 0 when not, -1 on error.
static int try_code_mode(const AssemblyParams &p, const double *coef27, const std::vector<int> &pslots, Operator *op, hipStream_t s) {
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    const unsigned grid = (unsigned)((n + 255) / 256);
    uint8_t *codes = nullptr;
    unsigned long long *dreps = nullptr, *dmis = nullptr;
    int64_t *drep = nullptr;
    double *dtable27 = nullptr;
    auto cleanup = [&]() {
        void *bufs[] = {codes, dreps, dmis, drep, dtable27};
        for(void *b : bufs)
            if(b) (void)hipFree(b);
    };
    if(hipMalloc(&codes, n) != hipSuccess || hipMalloc(&dreps, sizeof(unsigned long long) * kCodes) != hipSuccess ||
       hipMalloc(&dmis, sizeof(unsigned long long)) != hipSuccess || hipMalloc(&drep, sizeof(int64_t) * kCodes) != hipSuccess ||
       hipMalloc(&dtable27, sizeof(double) * kCodes * 27) != hipSuccess) {
        cleanup();
        mono_set_error("out of device memory building the operator");
        return -1;
    }
    k_row_codes<<<grid, 256, 0, s>>>(p.nx, p.ny, p.nz, p.active, codes);
    (void)hipMemsetAsync(dreps, 0xff, sizeof(unsigned long long) * kCodes, s);
    k_code_reps<<<grid, 256, 0, s>>>(codes, n, dreps);
    std::vector<unsigned long long> reps(kCodes);
    std::vector<int64_t> rep(kCodes);
    if(hipMemcpyAsync(reps.data(), dreps, sizeof(unsigned long long) * kCodes, hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        mono_set_error("HIP error building the operator's row codes");
        return -1;
    }
    for(int i = 0; i < kCodes; i++) rep[i] = reps[i] == ~0ull ? -1 : (int64_t)reps[i];
    std::vector<double> table27((size_t)kCodes * 27);
    unsigned long long mismatches = 0;
    if(hipMemcpyAsync(drep, rep.data(), sizeof(int64_t) * kCodes, hipMemcpyHostToDevice, s) != hipSuccess) {
        cleanup();
        mono_set_error("HIP error building the operator's row codes");
        return -1;
    }
    k_gather_rows<<<(kCodes * 27 + 255) / 256, 256, 0, s>>>(coef27, n, drep, kCodes, dtable27);
    (void)hipMemsetAsync(dmis, 0, sizeof(unsigned long long), s);
    k_verify_codes<<<grid, 256, 0, s>>>(coef27, n, codes, dtable27, dmis);
    if(hipGetLastError() != hipSuccess ||
       hipMemcpyAsync(table27.data(), dtable27, sizeof(double) * kCodes * 27, hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipMemcpyAsync(&mismatches, dmis, sizeof(mismatches), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess) {
        cleanup();
        mono_set_error("HIP error verifying the operator's row codes");
        return -1;
    }
    if(mismatches != 0) {
        cleanup();
        return 0;
    }
    std::vector<double> t64((size_t)kCodes * 7);
    std::vector<float> t32((size_t)kCodes * 7);
    for(int c = 0; c < kCodes; c++)
        for(int k = 0; k < 7; k++) {
            t64[(size_t)c * 7 + k] = table27[(size_t)c * 27 + pslots[k]];
            t32[(size_t)c * 7 + k] = (float)t64[(size_t)c * 7 + k];
        }
    (void)hipFree(dreps);
    (void)hipFree(dmis);
    (void)hipFree(drep);
    (void)hipFree(dtable27);
    op->codes = codes;
    if(hipMalloc(&op->ctab64, t64.size() * sizeof(double)) != hipSuccess || hipMalloc(&op->ctab32, t32.size() * sizeof(float)) != hipSuccess ||
       hipMemcpyAsync(op->ctab64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(op->ctab32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess) {
        mono_set_error("HIP error uploading the operator's code table");
        return -1;
    }
    return 1;
}

// Neighbourhood mask of each cell (see aniso_row.h): bit slot_of(o) when the
// neighbour at offset o lies in the box and is active, bit 13 the cell's own
// activity; 0 for an inactive cell.
__global__ __launch_bounds__(256) void k_nb_mask(int nx, int ny, int nz, const uint8_t *active, uint32_t *nbm) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((int64_t)nx * ny));
    uint32_t m = 0;
    if(!active || active[c]) {
        for(int k = 0; k < 27; k++) {
            const int xx = x + k % 3 - 1, yy = y + (k / 3) % 3 - 1, zz = z + k / 9 - 1;
            if(xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) continue;
            if(active && !active[xx + (int64_t)nx * (yy + (int64_t)ny * zz)]) continue;
            m |= 1u << k;
        }
    }
    nbm[c] = m;
}

// Rows re-formed from the masks in FP64 against the assembled rows, bit for bit.
// lstride: 6 for per-layer tensors, 0 for a uniform one.
__global__ __launch_bounds__(256) void k_verify_mask_rows(const double *coef27, int64_t n, int nx, int ny, int nz, const uint32_t *nbm,
                                                          const double *tensor, int lstride, double dx, double dy, double dz,
                                                          double alpha, unsigned long long *mismatch) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(c >= n) return;
    const int z = (int)(c / ((int64_t)nx * ny));
    double tz[3][6];
    for(int l = 0; l < 3; l++) {
        int zl = z - 1 + l;
        zl = zl < 0 ? 0 : (zl >= nz ? nz - 1 : zl);
        for(int k = 0; k < 6; k++) tz[l][k] = tensor[(int64_t)zl * lstride + k];
    }
    double row[27];
    mask_aniso_row<double>(nbm[c], tz, dx, dy, dz, alpha, row);
    bool bad = false;
    for(int k = 0; k < 27; k++)
        if(__double_as_longlong(row[k]) != __double_as_longlong(coef27[(int64_t)k * n + c])) bad = true;
    if(bad) atomicAdd(mismatch, 1ull);
}

// Try mask mode for the FP32 view; returns 1 when every assembled row matched
// its mask-formed row (op->nbm, lt32, geo32 set), 0 when not, -1 on error.
static int try_mask_mode(const AssemblyParams &p, const double *coef27, Operator *op, hipStream_t s) {
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    const unsigned grid = (unsigned)((n + 255) / 256);
    uint32_t *nbm = nullptr;
    unsigned long long *dmis = nullptr;
    if(hipMalloc(&nbm, sizeof(uint32_t) * n) != hipSuccess || hipMalloc(&dmis, sizeof(unsigned long long)) != hipSuccess) {
        if(nbm) (void)hipFree(nbm);
        mono_set_error("out of device memory building the operator");
        return -1;
    }
    const int lstride = p.tensor_source == kTensorPerLayer ? 6 : 0;
    const int nl = lstride ? p.nz : 1;
    std::vector<double> t64((size_t)nl * 6);
    unsigned long long mismatches = 0;
    k_nb_mask<<<grid, 256, 0, s>>>(p.nx, p.ny, p.nz, p.active, nbm);
    (void)hipMemsetAsync(dmis, 0, sizeof(unsigned long long), s);
    k_verify_mask_rows<<<grid, 256, 0, s>>>(coef27, n, p.nx, p.ny, p.nz, nbm, p.tensor, lstride, p.dx, p.dy, p.dz, p.alpha, dmis);
    if(hipGetLastError() != hipSuccess ||
       hipMemcpyAsync(&mismatches, dmis, sizeof(mismatches), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipMemcpyAsync(t64.data(), p.tensor, sizeof(double) * t64.size(), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(nbm);
        (void)hipFree(dmis);
        mono_set_error("HIP error verifying the operator's mask rows");
        return -1;
    }
    (void)hipFree(dmis);
    if(getenv("MONO_TRACE_MASK"))
        fprintf(stderr, "[mono] mask rows: %llu of %lld rows differ from the assembled rows\n", mismatches, (long long)n);
    if(mismatches != 0) {
        (void)hipFree(nbm);
        return 0;
    }
    // layer tables [nz][6] in FP32 and FP64 (a uniform tensor repeated)
    std::vector<float> t32((size_t)p.nz * 6);
    std::vector<double> l64((size_t)p.nz * 6);
    for(int z = 0; z < p.nz; z++)
        for(int k = 0; k < 6; k++) {
            l64[(size_t)z * 6 + k] = t64[(size_t)(lstride ? z : 0) * 6 + k];
            t32[(size_t)z * 6 + k] = (float)l64[(size_t)z * 6 + k];
        }
    if(hipMalloc(&op->lt32, sizeof(float) * t32.size()) != hipSuccess || hipMalloc(&op->lt64, sizeof(double) * l64.size()) != hipSuccess ||
       hipMemcpyAsync(op->lt32, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(op->lt64, l64.data(), sizeof(double) * l64.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(nbm);
        mono_set_error("HIP error uploading the operator's layer tensors");
        return -1;
    }
    op->nbm = nbm;
    op->geo32[0] = (float)p.dx;
    op->geo32[1] = (float)p.dy;
    op->geo32[2] = (float)p.dz;
    op->geo32[3] = (float)p.alpha;
    op->geo64[0] = p.dx;
    op->geo64[1] = p.dy;
    op->geo64[2] = p.dz;
    op->geo64[3] = p.alpha;
    return 1;
}

template <typename T>
static OpView<T> make_view(const Operator &o, const T *table, const T *tint, const T *percell) {
    OpView<T> v;
    memset(&v, 0, sizeof(v));
    v.nx = o.nx;
    v.ny = o.ny;
    v.nz = o.nz;
    v.n = o.n;
    v.pat = o.pat;
    v.table = table;
    v.table_int = tint;
    v.percell = percell;
    v.active = o.active;
    return v;
}

OpView<double> Operator::view64() const {
    if(code_mode) {
        OpView<double> v = make_view<double>(*this, nullptr, nullptr, nullptr);
        v.codes = codes;
        v.ctab = ctab64;
        return v;
    }
    return class_mode ? make_view<double>(*this, table64, tint64, nullptr) : make_view<double>(*this, nullptr, nullptr, percell64);
}
OpView<float> Operator::view32() const {
    if(code_mode) {
        OpView<float> v = make_view<float>(*this, nullptr, nullptr, nullptr);
        v.codes = codes;
        v.ctab = ctab32;
        return v;
    }
    if(mask32) {
        OpView<float> v = make_view<float>(*this, nullptr, nullptr, nullptr);
        v.nbm = nbm;
        v.lt = lt32;
        for(int k = 0; k < 4; k++) v.geo[k] = geo32[k];
        return v;
    }
    return class_mode ? make_view<float>(*this, table32, tint32, nullptr) : make_view<float>(*this, nullptr, nullptr, percell32);
}
OpView<double> Operator::view64_solve() const {
    if(!mask32) return view64();
    OpView<double> v = make_view<double>(*this, nullptr, nullptr, nullptr);
    v.nbm = nbm;
    v.lt = lt64;
    for(int k = 0; k < 4; k++) v.geo[k] = geo64[k];
    return v;
}

int Operator::release() {
    void *bufs[] = {table64, tint64, table32, tint32, percell64, percell32, codes, ctab64, ctab32, nbm, lt32, lt64};
    for(void *p : bufs)
        if(p) MONO_CHECK(hipFree(p));
    table64 = tint64 = percell64 = nullptr;
    table32 = tint32 = percell32 = nullptr;
    codes = nullptr;
    nbm = nullptr;
    lt32 = nullptr;
    lt64 = nullptr;
    mask32 = false;
    ctab64 = nullptr;
    ctab32 = nullptr;
    code_mode = false;
    return 0;
}

// slots (27-point numbering) of a row pattern
template <int P> static std::vector<int> pattern_slots() {
    std::vector<int> v;
    for(int k = 0; k < P; k++) v.push_back(slot_of(RowPattern<P>::o[k][0], RowPattern<P>::o[k][1], RowPattern<P>::o[k][2]));
    return v;
}

// smallest pattern containing every slot of `used`
static int choose_pattern(const std::vector<int> &used, std::vector<int> *slots) {
    const std::vector<int> p7 = pattern_slots<7>(), p19 = pattern_slots<19>(), p27 = pattern_slots<27>();
    auto covers = [&](const std::vector<int> &pat) {
        for(int u : used)
            if(std::find(pat.begin(), pat.end(), u) == pat.end()) return false;
        return true;
    };
    if(covers(p7)) {
        *slots = p7;
        return 7;
    }
    if(covers(p19)) {
        *slots = p19;
        return 19;
    }
    *slots = p27;
    return 27;
}

int build_operator(const AssemblyParams &p, const double *coef27, Operator *op, hipStream_t s) {
    const int64_t n = (int64_t)p.nx * p.ny * p.nz;
    op->nx = p.nx;
    op->ny = p.ny;
    op->nz = p.nz;
    op->n = n;
    op->active = p.active;
    op->alpha = p.alpha;
    const unsigned grid = (unsigned)((n + 255) / 256);

    // stored slots: the diagonal first, then the union of non-zero off-diagonals
    unsigned *dmask = nullptr;
    MONO_CHECK(hipMalloc(&dmask, sizeof(unsigned)));
    MONO_CHECK(hipMemsetAsync(dmask, 0, sizeof(unsigned), s));
    k_slot_mask<<<grid, 256, 0, s>>>(coef27, n, dmask);
    MONO_CHECK(hipGetLastError());
    unsigned mask = 0;
    MONO_CHECK(hipMemcpyAsync(&mask, dmask, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipStreamSynchronize(s));
    MONO_CHECK(hipFree(dmask));
    std::vector<int> slots;
    slots.push_back(13);
    for(int k = 0; k < 27; k++)
        if(k != 13 && (mask >> k) & 1u) slots.push_back(k);
    op->nslots = (int)slots.size();
    std::vector<int> bslots;  // the smallest pattern covering every row (per-cell storage order)
    const int pb = choose_pattern(slots, &bslots);

    // class table candidate: representative cell of each (z, cy, cx) class
    const int ncls = p.nz * 9;
    std::vector<int64_t> rep(ncls, -1);
    for(int z = 0; z < p.nz; z++)
        for(int cy = 0; cy < 3; cy++)
            for(int cx = 0; cx < 3; cx++) {
                const int xs[3] = {0, p.nx > 2 ? 1 : 0, p.nx - 1};
                const int ys[3] = {0, p.ny > 2 ? 1 : 0, p.ny - 1};
                const int x = xs[cx], y = ys[cy];
                if(pos_class(x, p.nx) != cx || pos_class(y, p.ny) != cy) continue;  // class absent
                rep[(z * 3 + cy) * 3 + cx] = x + (int64_t)p.nx * (y + (int64_t)p.ny * z);
            }
    std::vector<double> table27((size_t)ncls * 27, 0.0);
    int64_t *drep = nullptr;
    double *dtable27 = nullptr;
    unsigned long long *dmis = nullptr;
    MONO_CHECK(hipMalloc(&drep, sizeof(int64_t) * ncls));
    MONO_CHECK(hipMalloc(&dtable27, table27.size() * sizeof(double)));
    MONO_CHECK(hipMalloc(&dmis, sizeof(unsigned long long)));
    MONO_CHECK(hipMemcpyAsync(drep, rep.data(), sizeof(int64_t) * ncls, hipMemcpyHostToDevice, s));
    k_gather_rows<<<(ncls * 27 + 255) / 256, 256, 0, s>>>(coef27, n, drep, ncls, dtable27);
    MONO_CHECK(hipGetLastError());
    MONO_CHECK(hipMemcpyAsync(table27.data(), dtable27, table27.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipMemsetAsync(dmis, 0, sizeof(unsigned long long), s));
    k_verify_classes<<<grid, 256, 0, s>>>(coef27, n, p.nx, p.ny, dtable27, dmis);
    MONO_CHECK(hipGetLastError());
    unsigned long long mismatches = 0;
    MONO_CHECK(hipMemcpyAsync(&mismatches, dmis, sizeof(mismatches), hipMemcpyDeviceToHost, s));
    MONO_CHECK(hipStreamSynchronize(s));
    MONO_CHECK(hipFree(drep));
    MONO_CHECK(hipFree(dtable27));
    MONO_CHECK(hipFree(dmis));

    op->class_mode = (mismatches == 0);
    if(op->class_mode) {
        // interior rows (every axis in its fast range: 1..L-2, or the single
        // cell of an axis of extent 1): the union of their non-zero slots
        const int cxF = p.nx == 1 ? 0 : 1, cyF = p.ny == 1 ? 0 : 1;
        const int zlo = p.nz == 1 ? 0 : 1, zhi = p.nz == 1 ? 0 : p.nz - 2;
        std::vector<int> islots;
        islots.push_back(13);
        for(int k = 0; k < 27; k++) {
            if(k == 13) continue;
            bool used = false;
            for(int z = zlo; z <= zhi; z++) {
                const int cls = (z * 3 + cyF) * 3 + cxF;
                if(rep[cls] >= 0 && table27[(size_t)cls * 27 + k] != 0.0) used = true;
            }
            if(used) islots.push_back(k);
        }
        op->nint = (int)islots.size();
        std::vector<int> pslots;
        op->pat = choose_pattern(islots, &pslots);
        const int NI = op->pat;
        std::vector<double> t64(table27), i64((size_t)p.nz * NI, 0.0);
        std::vector<float> t32((size_t)ncls * 27, 0.0f), i32((size_t)p.nz * NI, 0.0f);
        for(size_t i = 0; i < t64.size(); i++) t32[i] = (float)t64[i];
        for(int z = 0; z < p.nz; z++)
            for(int k = 0; k < NI; k++) {
                const double v = table27[((size_t)(z * 3 + cyF) * 3 + cxF) * 27 + pslots[k]];
                i64[(size_t)z * NI + k] = v;
                i32[(size_t)z * NI + k] = (float)v;
            }
        MONO_CHECK(hipMalloc(&op->table64, t64.size() * sizeof(double)));
        MONO_CHECK(hipMalloc(&op->table32, t32.size() * sizeof(float)));
        MONO_CHECK(hipMalloc(&op->tint64, i64.size() * sizeof(double)));
        MONO_CHECK(hipMalloc(&op->tint32, i32.size() * sizeof(float)));
        MONO_CHECK(hipMemcpyAsync(op->table64, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice, s));
        MONO_CHECK(hipMemcpyAsync(op->table32, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice, s));
        MONO_CHECK(hipMemcpyAsync(op->tint64, i64.data(), i64.size() * sizeof(double), hipMemcpyHostToDevice, s));
        MONO_CHECK(hipMemcpyAsync(op->tint32, i32.data(), i32.size() * sizeof(float), hipMemcpyHostToDevice, s));
        MONO_CHECK(hipStreamSynchronize(s));
    } else {
        op->nint = 0;
        std::vector<int> pslots = bslots;
        op->pat = pb;
        if(pb == 7) {
            const int rc = try_code_mode(p, coef27, pslots, op, s);
            if(rc < 0) return -1;
            if(rc == 1) {
                op->code_mode = true;
                return 0;
            }
        }
        if(op->pat == 27 && p.aniso && !p.g && p.tensor_source != kTensorPerCell && p.nz > 1) {
            const char *e = getenv("MONO_MASK_ROWS");
            if(!e || strcmp(e, "0") != 0) {
                const int rc = try_mask_mode(p, coef27, op, s);
                if(rc < 0) return -1;
                op->mask32 = rc == 1;
            }
        }
        int *dslots = nullptr;
        MONO_CHECK(hipMalloc(&dslots, sizeof(int) * op->pat));
        MONO_CHECK(hipMemcpyAsync(dslots, pslots.data(), sizeof(int) * op->pat, hipMemcpyHostToDevice, s));
        MONO_CHECK(hipMalloc(&op->percell64, sizeof(double) * op->pat * n));
        if(!op->mask32) MONO_CHECK(hipMalloc(&op->percell32, sizeof(float) * op->pat * n));
        k_compact<<<grid, 256, 0, s>>>(coef27, n, op->pat, dslots, op->percell64, op->percell32);
        MONO_CHECK(hipGetLastError());
        MONO_CHECK(hipStreamSynchronize(s));
        MONO_CHECK(hipFree(dslots));
    }
    return 0;
}

}  // namespace mono
