// One anisotropic finite-volume row from a cell's neighbourhood mask and the
// conductivity tensors of its three z layers (device, templated on the
// arithmetic type).
//
// The rule is the one k_assemble applies (operator.hip), restated over a
// 27-bit mask: bit slot_of(ox, oy, oz) is set when the neighbour at that
// offset lies in the box and is active, bit 13 when the cell itself is active.
// With per-layer (or uniform) tensors and no per-cell conductivity factor the
// row depends on nothing else, so the stencil kernels can form it on the fly
// from 4 bytes per cell instead of reading 27 stored coefficients.  In FP64
// the operations are those of k_assemble in the same order (both files build
// with FMA contraction off), and the operator builder checks the two bit for
// bit before it hands the mask to a kernel.
//
// Reference rule: calc_sigmas / fill_elements_aniso (assembly_common.c:115-830),
// anisotropic_sigma_assembly_matrix (matrix_assembly.c:389-514).
#pragma once

#include <stdint.h>

#include "mono_common.h"

namespace mono {

// reference direction order (cell.h:21-49): FRONT/BACK = +z/-z, TOP/DOWN =
// +y/-y, RIGHT/LEFT = +x/-x, then the edges and corners
enum MaskDir { kFR, kBK, kTP, kDN, kRT, kLF, kFR_TP, kFR_DN, kFR_RT, kFR_LF, kFR_TP_RT, kFR_TP_LF, kFR_DN_RT, kFR_DN_LF,
               kBK_TP, kBK_DN, kBK_RT, kBK_LF, kBK_TP_RT, kBK_TP_LF, kBK_DN_RT, kBK_DN_LF, kTP_RT, kTP_LF, kDN_RT, kDN_LF,
               kNDir };

struct MaskDirs {
    static constexpr int8_t o[kNDir][3] = {
        {0, 0, 1},  {0, 0, -1}, {0, 1, 0},   {0, -1, 0},  {1, 0, 0},    {-1, 0, 0},   {0, 1, 1},
        {0, -1, 1}, {1, 0, 1},  {-1, 0, 1},  {1, 1, 1},   {-1, 1, 1},   {1, -1, 1},   {-1, -1, 1},
        {0, 1, -1}, {0, -1, -1}, {1, 0, -1}, {-1, 0, -1}, {1, 1, -1},   {-1, 1, -1},  {1, -1, -1},
        {-1, -1, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0},  {-1, -1, 0}};
};

__host__ __device__ constexpr int dir_slot(int d) { return slot_of(MaskDirs::o[d][0], MaskDirs::o[d][1], MaskDirs::o[d][2]); }

__host__ __device__ __forceinline__ bool mask_has(uint32_t m, int ox, int oy, int oz) { return (m >> slot_of(ox, oy, oz)) & 1u; }

// Flux sigmas along `axis` with sign `sgn` (see flux_sigmas in operator.hip):
// s1 the mean of the two cells' diagonal component, s2 / s3 the averages of
// the cross components over the available neighbour pairs straddling the flux
// line along the lower / higher remaining axis, c2 / c3 their pair counts x 2.
// tz[l][k]: component k (xx yy zz xy xz yz) of layer z - 1 + l.
template <typename T>
__host__ __device__ __forceinline__ void mask_flux_sigmas(uint32_t m, const T (&tz)[3][6], int axis, int sgn, T &s1, T &s2, T &s3, int &c2,
                                                 int &c3) {
    s1 = s2 = s3 = T(0);
    c2 = c3 = 0;
    int o[3] = {0, 0, 0};
    o[axis] = sgn;
    if(!mask_has(m, o[0], o[1], o[2])) return;
    s1 = (tz[o[2] + 1][axis] + tz[1][axis]) / T(2);
    int others[2], k = 0;
    for(int a = 0; a < 3; a++)
        if(a != axis) others[k++] = a;
    for(int which = 0; which < 2; which++) {
        const int b = others[which], cax = others[1 - which];
        const int lo = axis < b ? axis : b, hi = axis < b ? b : axis;
        const int cmp = (lo == 0 && hi == 1) ? 3 : ((lo == 0 && hi == 2) ? 4 : 5);
        T sum = T(0);
        int cnt = 0;
        const int cseq[3] = {0, 1, -1};
        for(int ci = 0; ci < 3; ci++) {
            for(int ai = 0; ai < 2; ai++) {
                int pp[3] = {0, 0, 0}, qq[3] = {0, 0, 0};
                pp[axis] = qq[axis] = ai ? sgn : 0;
                pp[cax] = qq[cax] = cseq[ci];
                pp[b] = 1;
                qq[b] = -1;
                if(mask_has(m, pp[0], pp[1], pp[2]) && mask_has(m, qq[0], qq[1], qq[2])) {
                    sum += tz[pp[2] + 1][cmp] + tz[qq[2] + 1][cmp];
                    cnt += 2;
                }
            }
        }
        if(cnt) sum /= T(cnt);
        if(which == 0) {
            s2 = sum;
            c2 = cnt;
        } else {
            s3 = sum;
            c3 = cnt;
        }
    }
}

template <typename T> __host__ __device__ __forceinline__ T mask_qd(T s, int c) { return c != 0 ? s / T(c) : T(0); }

// 1/x: the hardware reciprocal in FP32 (1 ulp); in FP64 the hardware
// reciprocal refined by two Newton steps (within an ulp or two; a division
// costs three times the instructions)
__host__ __device__ __forceinline__ float mask_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
__host__ __device__ __forceinline__ double mask_rcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    r = __builtin_fma(__builtin_fma(-x, r, 1.0), r, r);
    return r;
#else
    return 1.0 / x;
#endif
}

// Bits of the 27-point slots at offset o + t * e_b for the six pairs of the
// flux through face (axis, sgn) along cross axis b (t = -1 gives the lower
// cell of each pair), grouped by the z layer the pair lies in: group g holds
// the pairs whose cells sit in layer z - 1 + g (both cells of a pair share a
// layer unless b is z, when every pair spans layers z-1 and z+1: group 0).
struct PairMasks {
    uint32_t g[3];
};
__host__ __device__ constexpr PairMasks pair_masks(int axis, int sgn, int b) {
    PairMasks pm{{0u, 0u, 0u}};
    int cax = 0;
    for(int a = 0; a < 3; a++)
        if(a != axis && a != b) cax = a;
    const int cseq[3] = {0, 1, -1};
    for(int ci = 0; ci < 3; ci++)
        for(int ai = 0; ai < 2; ai++) {
            int q[3] = {0, 0, 0};
            q[axis] = ai ? sgn : 0;
            q[cax] = cseq[ci];
            q[b] = -1;
            const int grp = b == 2 ? 0 : q[2] + 1;
            pm.g[grp] |= 1u << slot_of(q[0], q[1], q[2]);
        }
    return pm;
}

// Flux sigmas of face (AXIS, SGN) with the two cross terms already divided
// twice by their pair counts (the X values of mask_aniso_row): pair
// availability from the mask shifted by two cells along the cross axis, the
// pairs of one layer counted at once, one reciprocal per cross term.  Same
// rule as mask_flux_sigmas in another summation order: rounding apart.
template <int AXIS, int SGN, typename T>
__host__ __device__ __forceinline__ void mask_flux_fast(uint32_t m, const T (&tz)[3][6], T &s1, T &X2, T &X3) {
    constexpr int oz = AXIS == 2 ? SGN : 0;
    constexpr int face = slot_of(AXIS == 0 ? SGN : 0, AXIS == 1 ? SGN : 0, oz);
    const bool has_face = (m >> face) & 1u;
    s1 = has_face ? (tz[oz + 1][AXIS] + tz[1][AXIS]) * T(0.5) : T(0);
    constexpr int b0 = AXIS == 0 ? 1 : 0, b1 = AXIS == 2 ? 1 : 2;
    T X[2];
#pragma unroll
    for(int which = 0; which < 2; which++) {
        const int b = which == 0 ? b0 : b1;
        const int lo = AXIS < b ? AXIS : b, hi = AXIS < b ? b : AXIS;
        const int cmp = (lo == 0 && hi == 1) ? 3 : ((lo == 0 && hi == 2) ? 4 : 5);
        const int stride = b == 0 ? 1 : (b == 1 ? 3 : 9);
        const uint32_t avail = m & (m >> (2 * stride));  // bit q: cells q and q + 2 e_b both present
        const PairMasks pm = pair_masks(AXIS, SGN, b);
        T sum = T(0);
        int np = 0;
#pragma unroll
        for(int g = 0; g < 3; g++) {
            if(!pm.g[g]) continue;
            const int c = __builtin_popcount(avail & pm.g[g]);
            // both cells of a pair: layers z-1 and z+1 when b is z, else layer z-1+g twice
            const T K = b == 2 ? tz[2][cmp] + tz[0][cmp] : tz[g][cmp] + tz[g][cmp];
            sum += T(c) * K;
            np += c;
        }
        // sum / cnt / cnt with cnt = 2 np (sum is 0 when np is)
        X[which] = has_face ? sum * mask_rcp(T(np ? 4 * np * np : 1)) : T(0);
    }
    X2 = X[0];
    X3 = X[1];
}

// row[slot] (27-point slot order) of a cell with mask m; zero when the cell is
// inactive (bit 13 clear).
// The 26 off-diagonal coefficients of the anisotropic rule (the direction
// cases of mask_aniso_row) with neighbour presence as factors h[d] in {0, 1}:
// a case contributes only when its own neighbour is present (factor g), an
// inner term only when its partner is (factor h of the partner).
template <typename T>
__host__ __device__ __forceinline__ void mask_assemble_fast(const T (&h)[kNDir], T dx, T dy, T dz, T yz_x, T xz_y, T xy_z, T x_r,
                                                            T x_l, T y_t, T y_d, T z_f, T z_b, T Xyr, T Xzr, T Xyl, T Xzl, T Yxt,
                                                            T Yzt, T Yxd, T Yzd, T Zxf, T Zyf, T Zxb, T Zyb, T (&row)[27]) {
    T acc[kNDir];
#pragma unroll
    for(int d = 0; d < kNDir; d++) acc[d] = T(0);
    T g, s1, s2, s3;
    // faces
    acc[kFR] += h[kFR] * (-z_f * xy_z + h[kBK] * (-Yzt * dx + Yzd * dx - Xzr * dy + Xzl * dy));
    acc[kBK] += h[kBK] * (-z_b * xy_z + h[kFR] * (Yzt * dx - Yzd * dx + Xzr * dy - Xzl * dy));
    acc[kTP] += h[kTP] * (-y_t * xz_y + h[kDN] * (-Zyf * dx + Zyb * dx - Xyr * dz + Xyl * dz));
    acc[kDN] += h[kDN] * (-y_d * xz_y + h[kTP] * (Zyf * dx - Zyb * dx + Xyr * dz - Xyl * dz));
    acc[kRT] += h[kRT] * (-x_r * yz_x + h[kLF] * (-Zxf * dy + Zxb * dy - Yxt * dz + Yxd * dz));
    acc[kLF] += h[kLF] * (-x_l * yz_x + h[kRT] * (Zxf * dy - Zxb * dy + Yxt * dz - Yxd * dz));
    // edges
    g = h[kFR_TP];
    s1 = h[kFR_DN] * (-Xyr * dz + Xyl * dz - Zyf * dx);
    s2 = h[kBK_TP] * (-Yzt * dx - Xzr * dy + Xzl * dy);
    acc[kFR_DN] -= g * s1;
    acc[kBK_TP] -= g * s2;
    acc[kFR_TP] += g * (s1 + s2);
    g = h[kFR_DN];
    s1 = h[kBK_DN] * (Yzd * dx - Xzr * dy + Xzl * dy);
    acc[kFR_DN] += g * s1;
    acc[kBK_DN] -= g * s1;
    g = h[kBK_TP];
    s1 = h[kBK_DN] * (Zyb * dx - Xyr * dz + Xyl * dz);
    acc[kBK_TP] += g * s1;
    acc[kBK_DN] -= g * s1;
    g = h[kFR_RT];
    s1 = h[kBK_RT] * (-Xzr * dy - Yzt * dx + Yzd * dx);
    s2 = h[kFR_LF] * (-Zxf * dy - Yxt * dz + Yxd * dz);
    acc[kBK_RT] -= g * s1;
    acc[kFR_LF] -= g * s2;
    acc[kFR_RT] += g * (s1 + s2);
    g = h[kFR_LF];
    s1 = h[kBK_LF] * (-Yzt * dx + Yzd * dx + Xzl * dy);
    acc[kFR_LF] += g * s1;
    acc[kBK_LF] -= g * s1;
    g = h[kBK_RT];
    s1 = h[kBK_LF] * (Zxb * dy - Yxt * dz + Yxd * dz);
    acc[kBK_RT] += g * s1;
    acc[kBK_LF] -= g * s1;
    g = h[kTP_RT];
    s1 = h[kDN_RT] * (-Xyr * dz - Zyf * dx + Zyb * dx);
    s2 = h[kTP_LF] * (-Zxf * dy + Zxb * dy - Yxt * dz);
    acc[kDN_RT] -= g * s1;
    acc[kTP_LF] -= g * s2;
    acc[kTP_RT] += g * (s1 + s2);
    g = h[kTP_LF];
    s1 = h[kDN_LF] * (-Zyf * dx + Zyb * dx + Xyl * dz);
    acc[kTP_LF] += g * s1;
    acc[kDN_LF] -= g * s1;
    g = h[kDN_RT];
    s1 = h[kDN_LF] * (-Zxf * dy + Zxb * dy + Yxd * dz);
    acc[kDN_RT] += g * s1;
    acc[kDN_LF] -= g * s1;
    // corners
    g = h[kFR_TP_RT];
    s1 = h[kFR_DN_RT] * (-Xyr * dz - Zyf * dx);
    s2 = h[kBK_TP_RT] * (-Xzr * dy - Yzt * dx);
    s3 = h[kFR_TP_LF] * (-Zxf * dy - Yxt * dz);
    acc[kFR_DN_RT] -= g * s1;
    acc[kBK_TP_RT] -= g * s2;
    acc[kFR_TP_LF] -= g * s3;
    acc[kFR_TP_RT] += g * (s1 + s2 + s3);
    g = h[kFR_TP_LF];
    s1 = h[kFR_DN_LF] * (Xyl * dz - Zyf * dx);
    s2 = h[kBK_TP_LF] * (Xzl * dy - Yzt * dx);
    acc[kFR_TP_LF] += g * (s1 + s2);
    acc[kFR_DN_LF] -= g * s1;
    acc[kBK_TP_LF] -= g * s2;
    g = h[kFR_DN_RT];
    s1 = h[kBK_DN_RT] * (-Xzr * dy + Yzd * dx);
    s2 = h[kFR_DN_LF] * (Yxd * dz - Zxf * dy);
    acc[kFR_DN_RT] += g * (s1 + s2);
    acc[kBK_DN_RT] -= g * s1;
    acc[kFR_DN_LF] -= g * s2;
    g = h[kFR_DN_LF];
    s1 = h[kBK_DN_LF] * (Xzl * dy + Yzd * dx);
    acc[kFR_DN_LF] += g * s1;
    acc[kBK_DN_LF] -= g * s1;
    g = h[kBK_TP_RT];
    s1 = h[kBK_DN_RT] * (-Xyr * dz + Zyb * dx);
    s2 = h[kBK_TP_LF] * (-Yxt * dz + Zxb * dy);
    acc[kBK_TP_RT] += g * (s1 + s2);
    acc[kBK_DN_RT] -= g * s1;
    acc[kBK_TP_LF] -= g * s2;
    g = h[kBK_TP_LF];
    s1 = h[kBK_DN_LF] * (Xyl * dz + Zyb * dx);
    acc[kBK_TP_LF] += g * s1;
    acc[kBK_DN_LF] -= g * s1;
    g = h[kBK_DN_RT];
    s1 = h[kBK_DN_LF] * (Yxd * dz + Zxb * dy);
    acc[kBK_DN_RT] += g * s1;
    acc[kBK_DN_LF] -= g * s1;
#pragma unroll
    for(int d = 0; d < kNDir; d++) row[dir_slot(d)] = acc[d];
}

// FAST: the flux sigmas by mask_flux_fast and the coefficients by
// mask_assemble_fast; otherwise the operations of k_assemble in its order.
template <typename T, bool FAST = false>
__host__ __device__ __forceinline__ void mask_aniso_row(uint32_t m, const T (&tz)[3][6], T dx, T dy, T dz, T alpha, T (&row)[27]) {
#pragma unroll
    for(int k = 0; k < 27; k++) row[k] = T(0);
    if(!((m >> 13) & 1u)) return;
    const T yz_x = (dy * dz) / dx, xz_y = (dx * dz) / dy, xy_z = (dx * dy) / dz;
    T x_r, x_l, y_t, y_d, z_f, z_b;
    // the twelve doubly divided cross terms
    T Xyr, Xzr, Xyl, Xzl, Yxt, Yzt, Yxd, Yzd, Zxf, Zyf, Zxb, Zyb;
    if constexpr(FAST) {
        mask_flux_fast<0, +1>(m, tz, x_r, Xyr, Xzr);
        mask_flux_fast<0, -1>(m, tz, x_l, Xyl, Xzl);
        mask_flux_fast<1, +1>(m, tz, y_t, Yxt, Yzt);
        mask_flux_fast<1, -1>(m, tz, y_d, Yxd, Yzd);
        mask_flux_fast<2, +1>(m, tz, z_f, Zxf, Zyf);
        mask_flux_fast<2, -1>(m, tz, z_b, Zxb, Zyb);
    } else {
        T xy_jx_r, xz_jx_r, xy_jx_l, xz_jx_l, xy_jy_t, yz_jy_t, xy_jy_d, yz_jy_d;
        T xz_jz_f, yz_jz_f, xz_jz_b, yz_jz_b;
        int cxy_jx_r, cxz_jx_r, cxy_jx_l, cxz_jx_l, cxy_jy_t, cyz_jy_t;
        int cxy_jy_d, cyz_jy_d, cxz_jz_f, cyz_jz_f, cxz_jz_b, cyz_jz_b;
        mask_flux_sigmas(m, tz, 0, +1, x_r, xy_jx_r, xz_jx_r, cxy_jx_r, cxz_jx_r);
        mask_flux_sigmas(m, tz, 0, -1, x_l, xy_jx_l, xz_jx_l, cxy_jx_l, cxz_jx_l);
        mask_flux_sigmas(m, tz, 1, +1, y_t, xy_jy_t, yz_jy_t, cxy_jy_t, cyz_jy_t);
        mask_flux_sigmas(m, tz, 1, -1, y_d, xy_jy_d, yz_jy_d, cxy_jy_d, cyz_jy_d);
        mask_flux_sigmas(m, tz, 2, +1, z_f, xz_jz_f, yz_jz_f, cxz_jz_f, cyz_jz_f);
        mask_flux_sigmas(m, tz, 2, -1, z_b, xz_jz_b, yz_jz_b, cxz_jz_b, cyz_jz_b);
        Xyr = mask_qd(xy_jx_r, cxy_jx_r);
        Xzr = mask_qd(xz_jx_r, cxz_jx_r);
        Xyl = mask_qd(xy_jx_l, cxy_jx_l);
        Xzl = mask_qd(xz_jx_l, cxz_jx_l);
        Yxt = mask_qd(xy_jy_t, cxy_jy_t);
        Yzt = mask_qd(yz_jy_t, cyz_jy_t);
        Yxd = mask_qd(xy_jy_d, cxy_jy_d);
        Yzd = mask_qd(yz_jy_d, cyz_jy_d);
        Zxf = mask_qd(xz_jz_f, cxz_jz_f);
        Zyf = mask_qd(yz_jz_f, cyz_jz_f);
        Zxb = mask_qd(xz_jz_b, cxz_jz_b);
        Zyb = mask_qd(yz_jz_b, cyz_jz_b);
    }

    if constexpr(FAST) {
        // every neighbour test a factor h[d] of 0 or 1: straight-line
        // multiply-adds in place of compares and selects
        T h[kNDir];
#pragma unroll
        for(int d = 0; d < kNDir; d++) h[d] = T((m >> dir_slot(d)) & 1u);
        mask_assemble_fast(h, dx, dy, dz, yz_x, xz_y, xy_z, x_r, x_l, y_t, y_d, z_f, z_b, Xyr, Xzr, Xyl, Xzl, Yxt, Yzt, Yxd, Yzd,
                           Zxf, Zyf, Zxb, Zyb, row);
        row[13] = alpha + (yz_x * x_r + yz_x * x_l + xz_y * y_t + xz_y * y_d + xy_z * z_f + xy_z * z_b);
        return;
    }
    T acc[kNDir];
#pragma unroll
    for(int d = 0; d < kNDir; d++) acc[d] = T(0);
    auto has = [&](int d) { return (m >> dir_slot(d)) & 1u; };
    // the exact form adds only non-zero terms, as k_assemble does (it decides
    // the sign of a zero coefficient); the fast one adds every term
    auto upd = [&](int d, T v) {
        if(FAST || v != T(0)) acc[d] += v;
    };
    const T diag = alpha + (yz_x * x_r + yz_x * x_l + xz_y * y_t + xz_y * y_d + xy_z * z_f + xy_z * z_b);
#pragma unroll
    for(int d = 0; d < kNDir; d++) {
        if(!has(d)) continue;
        T v = T(0), s1, s2, s3;
        switch(d) {
        case kFR:
            v += -z_f * xy_z;
            if(has(kBK)) v += -Yzt * dx + Yzd * dx - Xzr * dy + Xzl * dy;
            break;
        case kBK:
            v += -z_b * xy_z;
            if(has(kFR)) v += Yzt * dx - Yzd * dx + Xzr * dy - Xzl * dy;
            break;
        case kTP:
            v += -y_t * xz_y;
            if(has(kDN)) v += -Zyf * dx + Zyb * dx - Xyr * dz + Xyl * dz;
            break;
        case kDN:
            v += -y_d * xz_y;
            if(has(kTP)) v += Zyf * dx - Zyb * dx + Xyr * dz - Xyl * dz;
            break;
        case kRT:
            v += -x_r * yz_x;
            if(has(kLF)) v += -Zxf * dy + Zxb * dy - Yxt * dz + Yxd * dz;
            break;
        case kLF:
            v += -x_l * yz_x;
            if(has(kRT)) v += Zxf * dy - Zxb * dy + Yxt * dz - Yxd * dz;
            break;
        case kFR_TP:
            s1 = T(0);
            s2 = T(0);
            if(has(kFR_DN)) {
                s1 += -Xyr * dz + Xyl * dz - Zyf * dx;
                upd(kFR_DN, -s1);
            }
            if(has(kBK_TP)) {
                s2 += -Yzt * dx - Xzr * dy + Xzl * dy;
                upd(kBK_TP, -s2);
            }
            v += s1 + s2;
            break;
        case kFR_DN:
            s1 = T(0);
            if(has(kBK_DN)) {
                s1 += Yzd * dx - Xzr * dy + Xzl * dy;
                upd(kFR_DN, s1);
                upd(kBK_DN, -s1);
            }
            break;
        case kBK_TP:
            s1 = T(0);
            if(has(kBK_DN)) {
                s1 += Zyb * dx - Xyr * dz + Xyl * dz;
                upd(kBK_TP, s1);
                upd(kBK_DN, -s1);
            }
            break;
        case kFR_RT:
            s1 = T(0);
            s2 = T(0);
            if(has(kBK_RT)) {
                s1 += -Xzr * dy - Yzt * dx + Yzd * dx;
                upd(kBK_RT, -s1);
            }
            if(has(kFR_LF)) {
                s2 += -Zxf * dy - Yxt * dz + Yxd * dz;
                upd(kFR_LF, -s2);
            }
            v += s1 + s2;
            break;
        case kFR_LF:
            s1 = T(0);
            if(has(kBK_LF)) {
                s1 += -Yzt * dx + Yzd * dx + Xzl * dy;
                upd(kFR_LF, s1);
                upd(kBK_LF, -s1);
            }
            break;
        case kBK_RT:
            s1 = T(0);
            if(has(kBK_LF)) {
                s1 += Zxb * dy - Yxt * dz + Yxd * dz;
                upd(kBK_RT, s1);
                upd(kBK_LF, -s1);
            }
            break;
        case kTP_RT:
            s1 = T(0);
            s2 = T(0);
            if(has(kDN_RT)) {
                s1 += -Xyr * dz - Zyf * dx + Zyb * dx;
                upd(kDN_RT, -s1);
            }
            if(has(kTP_LF)) {
                s2 += -Zxf * dy + Zxb * dy - Yxt * dz;
                upd(kTP_LF, -s2);
            }
            v += s1 + s2;
            break;
        case kTP_LF:
            s1 = T(0);
            if(has(kDN_LF)) {
                s1 += -Zyf * dx + Zyb * dx + Xyl * dz;
                upd(kTP_LF, s1);
                upd(kDN_LF, -s1);
            }
            break;
        case kDN_RT:
            s1 = T(0);
            if(has(kDN_LF)) {
                s1 += -Zxf * dy + Zxb * dy + Yxd * dz;
                upd(kDN_RT, s1);
                upd(kDN_LF, -s1);
            }
            break;
        case kFR_TP_RT:
            s1 = T(0);
            s2 = T(0);
            s3 = T(0);
            if(has(kFR_DN_RT)) {
                s1 += -Xyr * dz - Zyf * dx;
                upd(kFR_DN_RT, -s1);
            }
            if(has(kBK_TP_RT)) {
                s2 += -Xzr * dy - Yzt * dx;
                upd(kBK_TP_RT, -s2);
            }
            if(has(kFR_TP_LF)) {
                s3 += -Zxf * dy - Yxt * dz;
                upd(kFR_TP_LF, -s3);
            }
            v += s1 + s2 + s3;
            break;
        case kFR_TP_LF:
            s1 = T(0);
            s2 = T(0);
            if(has(kFR_DN_LF)) {
                s1 += Xyl * dz - Zyf * dx;
                upd(kFR_TP_LF, s1);
                upd(kFR_DN_LF, -s1);
            }
            if(has(kBK_TP_LF)) {
                s2 += Xzl * dy - Yzt * dx;
                upd(kFR_TP_LF, s2);
                upd(kBK_TP_LF, -s2);
            }
            break;
        case kFR_DN_RT:
            s1 = T(0);
            s2 = T(0);
            if(has(kBK_DN_RT)) {
                s1 += -Xzr * dy + Yzd * dx;
                upd(kFR_DN_RT, s1);
                upd(kBK_DN_RT, -s1);
            }
            if(has(kFR_DN_LF)) {
                s2 += Yxd * dz - Zxf * dy;
                upd(kFR_DN_RT, s2);
                upd(kFR_DN_LF, -s2);
            }
            break;
        case kFR_DN_LF:
            s1 = T(0);
            if(has(kBK_DN_LF)) {
                s1 += Xzl * dy + Yzd * dx;
                upd(kFR_DN_LF, s1);
                upd(kBK_DN_LF, -s1);
            }
            break;
        case kBK_TP_RT:
            s1 = T(0);
            s2 = T(0);
            if(has(kBK_DN_RT)) {
                s1 += -Xyr * dz + Zyb * dx;
                upd(kBK_TP_RT, s1);
                upd(kBK_DN_RT, -s1);
            }
            if(has(kBK_TP_LF)) {
                s2 += -Yxt * dz + Zxb * dy;
                upd(kBK_TP_RT, s2);
                upd(kBK_TP_LF, -s2);
            }
            break;
        case kBK_TP_LF:
            s1 = T(0);
            if(has(kBK_DN_LF)) {
                s1 += Xyl * dz + Zyb * dx;
                upd(kBK_TP_LF, s1);
                upd(kBK_DN_LF, -s1);
            }
            break;
        case kBK_DN_RT:
            s1 = T(0);
            if(has(kBK_DN_LF)) {
                s1 += Yxd * dz + Zxb * dy;
                upd(kBK_DN_RT, s1);
                upd(kBK_DN_LF, -s1);
            }
            break;
        default:
            break;
        }
        if(FAST || v != T(0)) acc[d] += v;
    }
    row[13] = diag;
#pragma unroll
    for(int d = 0; d < kNDir; d++) row[dir_slot(d)] = acc[d];
}

}  // namespace mono
