// Per-site kernels: box set-up, local frames (forward and backward), site
// packing, self energy / polarisation penalty, solver vector algebra and
// deterministic reductions.
#pragma once
#include "admp_common.h"
#include "pair_kernels.h"

namespace admp {

// 10-digit sqrt(3) of the reference's harmonic conversions (dmff/admp/multipole.py:12)
constexpr double RT3_LIT = 1.73205080757;

// A box is usable when its entries are finite and each of its three heights
// (|det| / |a_e x a_f|, the slab width along reciprocal direction d) is at
// least BOX_MIN_HEIGHT A: then every fractional coordinate of a position within
// POS_MAX_ABS A stays far inside int range when it becomes a cell or grid index.
#define BOX_MIN_HEIGHT 0.5
#define POS_MAX_ABS 1.0e5
#define GEOM_BAD_POS 1
#define GEOM_BAD_BOX 2

__device__ __host__ inline bool box_usable(const double* B) {
    for (int k = 0; k < 9; ++k)
        if (!(fabs(B[k]) < 1e300)) return false;  // NaN and infinities fail
    const double c0[3] = {B[4] * B[8] - B[5] * B[7], B[5] * B[6] - B[3] * B[8], B[3] * B[7] - B[4] * B[6]};
    const double c1[3] = {B[7] * B[2] - B[8] * B[1], B[8] * B[0] - B[6] * B[2], B[6] * B[1] - B[7] * B[0]};
    const double c2[3] = {B[1] * B[5] - B[2] * B[4], B[2] * B[3] - B[0] * B[5], B[0] * B[4] - B[1] * B[3]};
    const double det = fabs(B[0] * c0[0] + B[1] * c0[1] + B[2] * c0[2]);
    const double n0 = sqrt(c0[0] * c0[0] + c0[1] * c0[1] + c0[2] * c0[2]);
    const double n1 = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
    const double n2 = sqrt(c2[0] * c2[0] + c2[1] * c2[1] + c2[2] * c2[2]);
    // height along d = det / |a_e x a_f|; compare without dividing by a zero norm
    return det >= BOX_MIN_HEIGHT * n0 && det >= BOX_MIN_HEIGHT * n1 && det >= BOX_MIN_HEIGHT * n2 && det > 0;
}

// Boxes of every frame: inverse, grid maps and reciprocal rows.  An unusable
// box (box_usable) is replaced by a 100 A cube so that nothing downstream
// derives an index from it, and GEOM_BAD_BOX is set in *flags; the entry point
// reports the error once its work has drained.
__device__ inline void box_setup_one(const double* __restrict__ box, int K0, int K1, int K2, int dK0, int dK1,
                                     int dK2, BoxInfo* __restrict__ out, int* __restrict__ flags) {
#pragma clang fp contract(off)
    double B[9];
    for (int k = 0; k < 9; ++k) B[k] = box[k];
    if (!box_usable(B)) {
        for (int k = 0; k < 9; ++k) B[k] = (k % 4 == 0) ? 100.0 : 0.0;
        atomicOr(flags, GEOM_BAD_BOX);
    }
    BoxInfo b;
    for (int k = 0; k < 9; ++k) b.B[k] = B[k];
    const double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                       B[2] * (B[3] * B[7] - B[4] * B[6]);
    const double id = 1.0 / det;
    // inverse: elementwise for a diagonal box (what an LU inverse gives there),
    // cofactors over the determinant otherwise; the oracle's brute-force pair
    // search forms the same values
    b.Binv[0] = (B[4] * B[8] - B[5] * B[7]) * id;
    b.Binv[1] = (B[2] * B[7] - B[1] * B[8]) * id;
    b.Binv[2] = (B[1] * B[5] - B[2] * B[4]) * id;
    b.Binv[3] = (B[5] * B[6] - B[3] * B[8]) * id;
    b.Binv[4] = (B[0] * B[8] - B[2] * B[6]) * id;
    b.Binv[5] = (B[2] * B[3] - B[0] * B[5]) * id;
    b.Binv[6] = (B[3] * B[7] - B[4] * B[6]) * id;
    b.Binv[7] = (B[1] * B[6] - B[0] * B[7]) * id;
    b.Binv[8] = (B[0] * B[4] - B[1] * B[3]) * id;
    b.ortho = (B[1] == 0 && B[2] == 0 && B[3] == 0 && B[5] == 0 && B[6] == 0 && B[7] == 0) ? 1 : 0;
    if (b.ortho) {
        for (int k = 0; k < 9; ++k) b.Binv[k] = 0.0;
        b.Binv[0] = 1.0 / B[0];
        b.Binv[4] = 1.0 / B[4];
        b.Binv[8] = 1.0 / B[8];
    }
    b.V = det;
    const int K[3] = {K0, K1, K2}, dK[3] = {dK0, dK1, dK2};
    for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c) {
            b.A[j * 3 + c] = K[j] * b.Binv[c * 3 + j];
            b.dA[j * 3 + c] = dK[j] * b.Binv[c * 3 + j];
            b.kb[c * 3 + j] = 2.0 * PI * b.Binv[c * 3 + j];
        }
    b.pad = 0;
    *out = b;
}

__global__ void k_box_setup(int F, const double* __restrict__ box, int K0, int K1, int K2, int dK0, int dK1,
                            int dK2, BoxInfo* __restrict__ out, int* __restrict__ flags) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    box_setup_one(box + (size_t)f * 9, K0, K1, K2, dK0, dK1, dK2, out + f, flags);
}

// Positions to double4.  A coordinate that is not finite or exceeds POS_MAX_ABS
// is replaced by 0 and GEOM_BAD_POS is set in *flags (reported by the entry point).
__global__ void k_pack_pos(size_t n, const double* __restrict__ p, double4* __restrict__ o, int* __restrict__ flags) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = p[i * 3], y = p[i * 3 + 1], z = p[i * 3 + 2];
    const bool ok = fabs(x) <= POS_MAX_ABS && fabs(y) <= POS_MAX_ABS && fabs(z) <= POS_MAX_ABS;
    if (!ok) {
        x = fabs(x) <= POS_MAX_ABS ? x : 0.0;
        y = fabs(y) <= POS_MAX_ABS ? y : 0.0;
        z = fabs(z) <= POS_MAX_ABS ? z : 0.0;
        atomicOr(flags, GEOM_BAD_POS);
    }
    o[i] = make_double4(x, y, z, 0.0);
}

// ----------------------------------------------------------------- local frames
struct FrameFwd {
    double vz[3], vx[3], vy[3];        // raw (minimum-image) vectors
    double nvz, nvx, nvy;              // their norms
    double z0[3], x0[3], y0[3];        // normalised
    double zs[3], nzs;                 // un-normalised z sum (bisector / three-fold)
    double xs[3], nxs;                 // un-normalised x sum (z-bisect)
    double z1[3], x1[3];
    double xg[3], nxg, p;
    double X[3], Y[3];
    double sz[3], sx[3], sy[3];        // pbc shifts of the raw vectors
};

__device__ inline void d_norm3(const double v[3], double o[3], double& n) {
    n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    o[0] = v[0] / n; o[1] = v[1] / n; o[2] = v[2] / n;
}

__device__ inline void raw_vec(const BoxInfo& b, const double4* pos, size_t fo, int from, int to, double v[3], double s[3]) {
    double4 a = pos[fo + to], c = pos[fo + from];
    v[0] = a.x - c.x; v[1] = a.y - c.y; v[2] = a.z - c.z;
    pbc_shift(b, v[0], v[1], v[2], s[0], s[1], s[2]);
}

// dmff/admp/spatial.py:77-151 for one site
__device__ inline void frame_forward(int type, const BoxInfo& b, const double4* pos, size_t fo, int i, int zi, int xi,
                                     int yi, FrameFwd& F) {
    for (int k = 0; k < 3; ++k) { F.sz[k] = F.sx[k] = F.sy[k] = 0; F.vy[k] = F.y0[k] = 0; }
    if (type == 5) {  // NoAxisType
        F.X[0] = 1; F.X[1] = 0; F.X[2] = 0;
        F.Y[0] = 0; F.Y[1] = 1; F.Y[2] = 0;
        F.z1[0] = 0; F.z1[1] = 0; F.z1[2] = 1;
        return;
    }
    raw_vec(b, pos, fo, i, zi, F.vz, F.sz);
    d_norm3(F.vz, F.z0, F.nvz);
    if (type == 4) {  // Z-only
        double xz = rint(fabs(F.z0[0]));
        F.x0[0] = 1.0 - xz; F.x0[1] = xz; F.x0[2] = 0.0;
    } else {
        raw_vec(b, pos, fo, i, xi, F.vx, F.sx);
        d_norm3(F.vx, F.x0, F.nvx);
    }
    for (int k = 0; k < 3; ++k) { F.z1[k] = F.z0[k]; F.x1[k] = F.x0[k]; }
    if (type == 2 || type == 3) {
        raw_vec(b, pos, fo, i, yi, F.vy, F.sy);
        d_norm3(F.vy, F.y0, F.nvy);
    }
    if (type == 1) {  // Bisector
        for (int k = 0; k < 3; ++k) F.zs[k] = F.z0[k] + F.x0[k];
        d_norm3(F.zs, F.z1, F.nzs);
    } else if (type == 2) {  // Z-bisect
        for (int k = 0; k < 3; ++k) F.xs[k] = F.x0[k] + F.y0[k];
        d_norm3(F.xs, F.x1, F.nxs);
    } else if (type == 3) {  // Three-fold
        for (int k = 0; k < 3; ++k) F.zs[k] = F.z0[k] + F.x0[k] + F.y0[k];
        d_norm3(F.zs, F.z1, F.nzs);
    }
    F.p = F.x1[0] * F.z1[0] + F.x1[1] * F.z1[1] + F.x1[2] * F.z1[2];
    for (int k = 0; k < 3; ++k) F.xg[k] = F.x1[k] - F.z1[k] * F.p;
    d_norm3(F.xg, F.X, F.nxg);
    F.Y[0] = F.z1[1] * F.X[2] - F.z1[2] * F.X[1];
    F.Y[1] = F.z1[2] * F.X[0] - F.z1[0] * F.X[2];
    F.Y[2] = F.z1[0] * F.X[1] - F.z1[1] * F.X[0];
}

__device__ inline void norm_bwd(const double n[3], double len, const double g[3], double out[3]) {
    double d = n[0] * g[0] + n[1] * g[1] + n[2] * g[2];
    for (int k = 0; k < 3; ++k) out[k] = (g[k] - n[k] * d) / len;
}

// harmonic local multipoles -> Cartesian local (C2_h2c with the rt3 literal)
__device__ inline void harm2cart_local(const double* h, int nh, double& q, double mu[3], double T[3][3]) {
    q = h[0];
    mu[0] = mu[1] = mu[2] = 0;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) T[a][c] = 0;
    if (nh >= 4) { mu[0] = h[2]; mu[1] = h[3]; mu[2] = h[1]; }
    if (nh >= 9) {
        const double r2 = RT3_LIT / 2;
        T[0][0] = -0.5 * h[4] + r2 * h[7];
        T[1][1] = -0.5 * h[4] - r2 * h[7];
        T[2][2] = h[4];
        T[0][1] = T[1][0] = r2 * h[8];
        T[0][2] = T[2][0] = r2 * h[5];
        T[1][2] = T[2][1] = r2 * h[6];
    }
}

// Q_global (Cartesian, 10 words) = rotate(Q_local) with rows of Rm = local axes
__global__ void k_frames_fwd(int N, int nh, int use_frames, const double4* __restrict__ pos,
                             const BoxInfo* __restrict__ boxes, const int* __restrict__ axis_type,
                             const int* __restrict__ axis_idx, const double* __restrict__ Q_local,
                             double* __restrict__ frames, double* __restrict__ Qc) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t fo = (size_t)f * N, a = fo + i;
    double Rm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    if (use_frames) {
        FrameFwd F;
        frame_forward(axis_type[i], boxes[f], pos, fo, i, axis_idx[i * 3], axis_idx[i * 3 + 1], axis_idx[i * 3 + 2], F);
        for (int k = 0; k < 3; ++k) { Rm[0][k] = F.X[k]; Rm[1][k] = F.Y[k]; Rm[2][k] = F.z1[k]; }
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) frames[a * 9 + r * 3 + c] = Rm[r][c];
    double q, mu[3], T[3][3];
    harm2cart_local(Q_local + (size_t)i * nh, nh, q, mu, T);
    double mg[3], Tg[3][3];
    for (int c = 0; c < 3; ++c) mg[c] = Rm[0][c] * mu[0] + Rm[1][c] * mu[1] + Rm[2][c] * mu[2];
    double TR[3][3];
    for (int x = 0; x < 3; ++x)
        for (int d = 0; d < 3; ++d) TR[x][d] = T[x][0] * Rm[0][d] + T[x][1] * Rm[1][d] + T[x][2] * Rm[2][d];
    for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) Tg[c][d] = Rm[0][c] * TR[0][d] + Rm[1][c] * TR[1][d] + Rm[2][c] * TR[2][d];
    double* o = Qc + a * 10;
    o[0] = q; o[1] = mg[0]; o[2] = mg[1]; o[3] = mg[2];
    o[4] = Tg[0][0]; o[5] = Tg[1][1]; o[6] = Tg[2][2]; o[7] = Tg[0][1]; o[8] = Tg[0][2]; o[9] = Tg[1][2];
}

// Backward through rotation and frame construction.  gQc (10): dE/d(q, mu, Theta
// matrix entries).  Writes dE/dQ_local (nh), per-site contributions to the
// gradient of the positions of (site, z, x, y) atoms (12 words) and the frame
// part of dE/dbox (9 words).
__global__ void k_frames_bwd(int N, int nh, int use_frames, const double4* __restrict__ pos,
                             const BoxInfo* __restrict__ boxes, const int* __restrict__ axis_type,
                             const int* __restrict__ axis_idx, const double* __restrict__ Q_local,
                             const double* __restrict__ gQc, double* __restrict__ gQl,
                             double* __restrict__ contrib, double* __restrict__ vir) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t fo = (size_t)f * N, a = fo + i;
    FrameFwd F;
    double Rm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int type = use_frames ? axis_type[i] : 5;
    if (use_frames) {
        frame_forward(type, boxes[f], pos, fo, i, axis_idx[i * 3], axis_idx[i * 3 + 1], axis_idx[i * 3 + 2], F);
        for (int k = 0; k < 3; ++k) { Rm[0][k] = F.X[k]; Rm[1][k] = F.Y[k]; Rm[2][k] = F.z1[k]; }
    }
    double q, mu[3], T[3][3];
    harm2cart_local(Q_local + (size_t)i * nh, nh, q, mu, T);
    const double* g = gQc + a * 10;
    double gmu_g[3] = {g[1], g[2], g[3]};
    double G[3][3];
    G[0][0] = g[4]; G[1][1] = g[5]; G[2][2] = g[6];
    G[0][1] = G[1][0] = g[7]; G[0][2] = G[2][0] = g[8]; G[1][2] = G[2][1] = g[9];
    // local gradients
    double gmu_l[3];
    for (int r = 0; r < 3; ++r) gmu_l[r] = Rm[r][0] * gmu_g[0] + Rm[r][1] * gmu_g[1] + Rm[r][2] * gmu_g[2];
    double RG[3][3], Gl[3][3];
    for (int r = 0; r < 3; ++r)
        for (int d = 0; d < 3; ++d) RG[r][d] = Rm[r][0] * G[0][d] + Rm[r][1] * G[1][d] + Rm[r][2] * G[2][d];
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) Gl[r][s] = RG[r][0] * Rm[s][0] + RG[r][1] * Rm[s][1] + RG[r][2] * Rm[s][2];
    double* ol = gQl + a * nh;
    ol[0] = g[0];
    if (nh >= 4) { ol[1] = gmu_l[2]; ol[2] = gmu_l[0]; ol[3] = gmu_l[1]; }
    if (nh >= 9) {
        const double r2 = RT3_LIT / 2;
        ol[4] = -0.5 * Gl[0][0] - 0.5 * Gl[1][1] + Gl[2][2];
        ol[7] = r2 * (Gl[0][0] - Gl[1][1]);
        ol[8] = 2.0 * r2 * Gl[0][1];
        ol[5] = 2.0 * r2 * Gl[0][2];
        ol[6] = 2.0 * r2 * Gl[1][2];
    }
    double* cb = contrib + a * 12;
    double* vb = vir + a * 9;
    for (int k = 0; k < 12; ++k) cb[k] = 0;
    for (int k = 0; k < 9; ++k) vb[k] = 0;
    if (!use_frames || type == 5) return;
    // dE/dRm[r][c] = mu_l[r] gmu_g[c] + 2 (T Rm G)[r][c]
    double TRG[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) TRG[r][c] = T[r][0] * RG[0][c] + T[r][1] * RG[1][c] + T[r][2] * RG[2][c];
    double gX[3], gY[3], gZ[3];
    for (int c = 0; c < 3; ++c) {
        gX[c] = mu[0] * gmu_g[c] + 2.0 * TRG[0][c];
        gY[c] = mu[1] * gmu_g[c] + 2.0 * TRG[1][c];
        gZ[c] = mu[2] * gmu_g[c] + 2.0 * TRG[2][c];
    }
    // Y = z1 x X
    double gz1[3], gXt[3];
    gz1[0] = F.X[1] * gY[2] - F.X[2] * gY[1];
    gz1[1] = F.X[2] * gY[0] - F.X[0] * gY[2];
    gz1[2] = F.X[0] * gY[1] - F.X[1] * gY[0];
    gXt[0] = gX[0] + (gY[1] * F.z1[2] - gY[2] * F.z1[1]);
    gXt[1] = gX[1] + (gY[2] * F.z1[0] - gY[0] * F.z1[2]);
    gXt[2] = gX[2] + (gY[0] * F.z1[1] - gY[1] * F.z1[0]);
    for (int k = 0; k < 3; ++k) gz1[k] += gZ[k];
    double gxg[3];
    norm_bwd(F.X, F.nxg, gXt, gxg);
    // xg = x1 - z1 (x1.z1)
    const double gz = gxg[0] * F.z1[0] + gxg[1] * F.z1[1] + gxg[2] * F.z1[2];
    double gx1[3];
    for (int k = 0; k < 3; ++k) {
        gx1[k] = gxg[k] - F.z1[k] * gz;
        gz1[k] += -F.p * gxg[k] - F.x1[k] * gz;
    }
    double gz0[3] = {0, 0, 0}, gx0[3] = {0, 0, 0}, gy0[3] = {0, 0, 0};
    if (type == 1) {
        double gzs[3];
        norm_bwd(F.z1, F.nzs, gz1, gzs);
        for (int k = 0; k < 3; ++k) { gz0[k] = gzs[k]; gx0[k] = gx1[k] + gzs[k]; }
    } else if (type == 2) {
        double gxs[3];
        norm_bwd(F.x1, F.nxs, gx1, gxs);
        for (int k = 0; k < 3; ++k) { gz0[k] = gz1[k]; gx0[k] = gxs[k]; gy0[k] = gxs[k]; }
    } else if (type == 3) {
        double gzs[3];
        norm_bwd(F.z1, F.nzs, gz1, gzs);
        for (int k = 0; k < 3; ++k) { gz0[k] = gzs[k]; gx0[k] = gx1[k] + gzs[k]; gy0[k] = gzs[k]; }
    } else {
        for (int k = 0; k < 3; ++k) { gz0[k] = gz1[k]; gx0[k] = gx1[k]; }
    }
    double gvz[3], gvx[3] = {0, 0, 0}, gvy[3] = {0, 0, 0};
    norm_bwd(F.z0, F.nvz, gz0, gvz);
    if (type != 4) norm_bwd(F.x0, F.nvx, gx0, gvx);
    if (type == 2 || type == 3) norm_bwd(F.y0, F.nvy, gy0, gvy);
    // raw vector v = pos[other] - pos[i] - n B
    for (int k = 0; k < 3; ++k) {
        cb[0 + k] = -(gvz[k] + gvx[k] + gvy[k]);
        cb[3 + k] = gvz[k];
        cb[6 + k] = gvx[k];
        cb[9 + k] = gvy[k];
    }
    for (int p = 0; p < 3; ++p)
        for (int c = 0; c < 3; ++c) vb[p * 3 + c] = -(F.sz[p] * gvz[c] + F.sx[p] * gvx[c] + F.sy[p] * gvy[c]);
}

// dpos[j] += own contribution + contributions of sites whose frame uses j
// (rev_off/rev_ent: static reverse map, entry = site*4 + slot).
// gpos = (gpos + add) + the frame contributions of the sites that use atom j
// (add: e.g. the reciprocal part, may be null)
__global__ void k_frames_gather(int N, const int* __restrict__ rev_off, const int* __restrict__ rev_ent,
                                const double* __restrict__ contrib, const double* __restrict__ add,
                                double* __restrict__ gpos) {
    const int f = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const size_t fo = (size_t)f * N;
    double s0 = contrib[(fo + j) * 12 + 0], s1 = contrib[(fo + j) * 12 + 1], s2 = contrib[(fo + j) * 12 + 2];
    for (int e = rev_off[j]; e < rev_off[j + 1]; ++e) {
        const int ent = rev_ent[e];
        const int site = ent >> 2, slot = ent & 3;
        const double* c = contrib + (fo + site) * 12 + slot * 3;
        s0 += c[0]; s1 += c[1]; s2 += c[2];
    }
    double* g = gpos + (fo + j) * 3;
    if (add) {
        const double* ad = add + (fo + j) * 3;
        g[0] = (g[0] + ad[0]) + s0;
        g[1] = (g[1] + ad[1]) + s1;
        g[2] = (g[2] + ad[2]) + s2;
    } else {
        g[0] += s0;
        g[1] += s1;
        g[2] += s2;
    }
}

// ----------------------------------------------------------------- site packing
// site record for the pair kernels (see pair_kernels.h)
template <class R>
__global__ void k_pack_site(int N, const double* __restrict__ Qc, const double* __restrict__ pol,
                            const double* __restrict__ thole, const double* __restrict__ U, R* __restrict__ site) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    R* s = site + a * SITE;
    const double* c = Qc + a * 10;
    for (int k = 0; k < 10; ++k) s[k] = (R)c[k];
    const double al = pol ? pol[i] : 0.0;
    s[10] = (R)(al > 0 ? pow(al, 1.0 / 6.0) : 0.0);
    s[11] = (R)(thole ? thole[i] : 0.0);
    s[12] = (R)al;
    s[13] = (R)(U ? U[a * 3 + 0] : 0.0);
    s[14] = (R)(U ? U[a * 3 + 1] : 0.0);
    s[15] = (R)(U ? U[a * 3 + 2] : 0.0);
}

template <class R> __global__ void k_cast(size_t n, const double* __restrict__ in, R* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (R)in[i];
}

// Self energy (pme.py:1183-1222) on Q + U and polarisation penalty (pme.py:1225-1240):
// per-site energy, and their gradients added into gQc / gpol.
// gq_add (may be null): added to gQc after the self terms (the reciprocal part)
__global__ void k_self(int N, double kappa, int lpme, int lpol, const double* __restrict__ Qc,
                       const double* __restrict__ U, const double* __restrict__ pol, double* __restrict__ e_self,
                       double* __restrict__ e_pol, double* __restrict__ gQc, double* __restrict__ gpol,
                       const double* __restrict__ gq_add) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    const double* c = Qc + a * 10;
    double mu[3] = {c[1], c[2], c[3]};
    if (lpol) { mu[0] += U[a * 3]; mu[1] += U[a * 3 + 1]; mu[2] += U[a * 3 + 2]; }
    double es = 0;
    if (lpme) {
        const double f0 = kappa / SQRT_PI;
        const double f1 = f0 * 2.0 * kappa * kappa / 3.0;
        const double f2 = f0 * 4.0 * kappa * kappa * kappa * kappa / 15.0;
        const double ThTh = c[4] * c[4] + c[5] * c[5] + c[6] * c[6] + 2.0 * (c[7] * c[7] + c[8] * c[8] + c[9] * c[9]);
        const double m2 = mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
        es = -DIELECTRIC * (f0 * c[0] * c[0] + f1 * m2 + f2 * (2.0 / 3.0) * ThTh);
        double* g = gQc + a * 10;
        g[0] += -2.0 * DIELECTRIC * f0 * c[0];
        for (int k = 0; k < 3; ++k) g[1 + k] += -2.0 * DIELECTRIC * f1 * mu[k];
        for (int k = 0; k < 6; ++k) g[4 + k] += -(4.0 / 3.0) * DIELECTRIC * f2 * c[4 + k];
    }
    if (gq_add) {
        double* g = gQc + a * 10;
        const double* ad = gq_add + a * 10;
        for (int k = 0; k < 10; ++k) g[k] = g[k] + ad[k];
    }
    e_self[a] = es;
    if (lpol) {
        const double al = pol[i];
        const double ap = al < 1e-8 ? 1e-8 : al;
        const double u2 = U[a * 3] * U[a * 3] + U[a * 3 + 1] * U[a * 3 + 1] + U[a * 3 + 2] * U[a * 3 + 2];
        e_pol[a] = DIELECTRIC * 0.5 * u2 / ap;
        if (gpol && al >= 1e-8) gpol[a] += -DIELECTRIC * 0.5 * u2 / (al * al);
    } else {
        e_pol[a] = 0.0;
    }
}

// ----------------------------------------------------------------- reductions
// out[f*ncol + c] = sum_i in[(f*N + i)*ncol + c]   (one block per (column, frame), fixed order)
// one block (RED_BLOCK threads) per (column, frame); four independent partial
// sums per thread keep several loads in flight, combined in a fixed order
constexpr int RED_BLOCK = 1024;
__global__ void __launch_bounds__(RED_BLOCK) k_reduce_cols(int N, int ncol, const double* __restrict__ in,
                                                           double* __restrict__ out, int accumulate) {
    __shared__ double sh[16];
    const int c = blockIdx.x, f = blockIdx.y;
    const double* col = in + (size_t)f * N * ncol + c;
    const int st = blockDim.x;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int i = threadIdx.x;
    for (; i + 3 * st < N; i += 4 * st) {
        s0 += col[(size_t)i * ncol];
        s1 += col[(size_t)(i + st) * ncol];
        s2 += col[(size_t)(i + 2 * st) * ncol];
        s3 += col[(size_t)(i + 3 * st) * ncol];
    }
    for (; i < N; i += st) s0 += col[(size_t)i * ncol];
    double s = (s0 + s1) + (s2 + s3);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        if (accumulate) out[(size_t)f * ncol + c] += s;
        else out[(size_t)f * ncol + c] = s;
    }
}

// the virial column sums of the final pass in one launch: block (c, f, z) writes
// the sum of column c of frame f of in_z into out_z (each in the order of
// k_reduce_cols)
__device__ __forceinline__ double reduce_col(int N, int ncol, const double* __restrict__ col, double* sh) {
    const int st = blockDim.x;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int i = threadIdx.x;
    for (; i + 3 * st < N; i += 4 * st) {
        s0 += col[(size_t)i * ncol];
        s1 += col[(size_t)(i + st) * ncol];
        s2 += col[(size_t)(i + 2 * st) * ncol];
        s3 += col[(size_t)(i + 3 * st) * ncol];
    }
    for (; i < N; i += st) s0 += col[(size_t)i * ncol];
    double s = (s0 + s1) + (s2 + s3);
    return block_sum(s, sh);
}
struct Vir3 {
    const double* in[3];
    double* out[3];
};
__global__ void __launch_bounds__(RED_BLOCK) k_reduce_vir(int N, Vir3 v) {
    __shared__ double sh[16];
    const int c = blockIdx.x, f = blockIdx.y, z = blockIdx.z;
    const double s = reduce_col(N, 9, v.in[z] + (size_t)f * N * 9 + c, sh);
    if (threadIdx.x == 0) v.out[z][(size_t)f * 9 + c] = s;
}

// three single-column reductions in one launch: block (k, f) sums in[k] of frame
// f (N values) in the order of k_reduce_cols into out[k * F + f]
struct Reduce3 {
    const double* in[3];
};
__global__ void __launch_bounds__(RED_BLOCK) k_reduce3(int N, Reduce3 r, double* __restrict__ out) {
    __shared__ double sh[16];
    const int k = blockIdx.x, f = blockIdx.y;
    const double* col = r.in[k] + (size_t)f * N;
    const int st = blockDim.x;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int i = threadIdx.x;
    for (; i + 3 * st < N; i += 4 * st) {
        s0 += col[i];
        s1 += col[i + st];
        s2 += col[i + 2 * st];
        s3 += col[i + 3 * st];
    }
    for (; i < N; i += st) s0 += col[i];
    double s = (s0 + s1) + (s2 + s3);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[(size_t)k * gridDim.y + f] = s;
}

// ----------------------------------------------------------------- solver algebra (FP64)
// y = sum of terms: out = real + recip + diag * v   where diag_i = D/max(pol,1e-8) - 2 D f1
__global__ void k_apply_diag(int N, double f1, const double* __restrict__ pol, const double* __restrict__ v,
                             const double* __restrict__ real, const double* __restrict__ recip, double* __restrict__ out) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    const double ap = pol[i] < 1e-8 ? 1e-8 : pol[i];
    const double d = DIELECTRIC / ap - 2.0 * DIELECTRIC * f1;
    for (int k = 0; k < 3; ++k) {
        double r = real ? real[a * 3 + k] : 0.0;
        double q = recip ? recip[a * 3 + k] : 0.0;
        out[a * 3 + k] = r + q + d * v[a * 3 + k];
    }
}

// per-frame partial dot products / max norms: part[(f*nb + b)*3 + {0,1,2}]
// = (sum x.y, sum x.z, max over polarisable sites |x component|)
__global__ void k_dots(int N, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                       const double* __restrict__ pol, double* __restrict__ part) {
    __shared__ double sh[16];
    __shared__ double shm[16];
    const int f = blockIdx.y;
    double sxy = 0, sxz = 0, mx = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const size_t a = (size_t)f * N + i;
        for (int k = 0; k < 3; ++k) {
            const double xv = x[a * 3 + k];
            if (y) sxy += xv * y[a * 3 + k];
            if (z) sxz += xv * z[a * 3 + k];
            if (pol[i] > 0.001) mx = fmax(mx, fabs(xv));
        }
    }
    sxy = block_sum(sxy, sh);
    sxz = block_sum(sxz, sh);
    // block max
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) shm[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = fmax(m, shm[k]);
        double* p = part + ((size_t)f * gridDim.x + blockIdx.x) * 3;
        p[0] = sxy; p[1] = sxz; p[2] = m;
    }
}

__global__ void k_dots_final(int F, int nb, const double* __restrict__ part, double* __restrict__ out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double a = 0, b = 0, m = 0;
    for (int k = 0; k < nb; ++k) {
        const double* p = part + ((size_t)f * nb + k) * 3;
        a += p[0]; b += p[1]; m = fmax(m, p[2]);
    }
    out[f * 3 + 0] = a; out[f * 3 + 1] = b; out[f * 3 + 2] = m;
}

// generic elementwise: out = a*x + b*y (+ c*z)
__global__ void k_lincomb(size_t n, double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                          double* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

// Jacobi update (pme.py:395): U -= g * pol / DIELECTRIC  for active frames
__global__ void k_jacobi(int N, const int* __restrict__ active, const double* __restrict__ pol,
                         const double* __restrict__ g, double* __restrict__ U) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || !active[f]) return;
    const size_t a = (size_t)f * N + i;
    for (int k = 0; k < 3; ++k) U[a * 3 + k] -= g[a * 3 + k] * pol[i] / DIELECTRIC;
}

// minus-in-place
__global__ void k_sub(size_t n, const double* __restrict__ b, double* __restrict__ x) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] -= b[i];
}

// dE/dbox assembly: vir_direct (+ vir_direct2, may be null) (F,9) hold the direct
// sums (real pairs, frames), dEdA the reciprocal A term;
// here: dbox = vir_direct - Binv^T H Binv^T + (-E_rec Binv^T + 2 Binv^T W)
// H = (diag(K) dEdA)^T.
__global__ void k_box_grad(int F, const BoxInfo* __restrict__ boxes, const double* __restrict__ vir_direct,
                           const double* __restrict__ vir_direct2, const double* __restrict__ dEdA, const double* __restrict__ erec,
                           const double* __restrict__ W6, int K0, int K1, int K2, int use_dA, double* __restrict__ dbox,
                           int accumulate) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const BoxInfo& b = boxes[f];
    double BiT[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) BiT[p][q] = b.Binv[q * 3 + p];
    double out[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
            out[p][q] = vir_direct ? vir_direct[f * 9 + p * 3 + q] : 0.0;
            if (vir_direct2) out[p][q] += vir_direct2[f * 9 + p * 3 + q];
        }
    if (dEdA) {
        const int K[3] = {K0, K1, K2};
        double H[3][3];  // H[c][j] = dEdA[j][c] K_j
        for (int c = 0; c < 3; ++c)
            for (int j = 0; j < 3; ++j) H[c][j] = dEdA[f * 9 + j * 3 + c] * K[j];
        double HB[3][3];
        for (int c = 0; c < 3; ++c)
            for (int q = 0; q < 3; ++q) HB[c][q] = H[c][0] * BiT[0][q] + H[c][1] * BiT[1][q] + H[c][2] * BiT[2][q];
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) out[p][q] -= BiT[p][0] * HB[0][q] + BiT[p][1] * HB[1][q] + BiT[p][2] * HB[2][q];
    }
    if (W6) {
        const double* w = W6 + f * 7;  // e, w00, w11, w22, w01, w02, w12
        const double E = w[0];
        double Wm[3][3];
        Wm[0][0] = w[1]; Wm[1][1] = w[2]; Wm[2][2] = w[3];
        Wm[0][1] = Wm[1][0] = w[4]; Wm[0][2] = Wm[2][0] = w[5]; Wm[1][2] = Wm[2][1] = w[6];
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q)
                out[p][q] += -E * BiT[p][q] + 2.0 * (BiT[p][0] * Wm[0][q] + BiT[p][1] * Wm[1][q] + BiT[p][2] * Wm[2][q]);
    }
    for (int k = 0; k < 9; ++k) {
        if (accumulate) dbox[f * 9 + k] += out[k / 3][k % 3];
        else dbox[f * 9 + k] = out[k / 3][k % 3];
    }
}


// b += 2 D f1 mu_perm (self-energy part of -dE/dU at U = 0)
// b = (b - recip) + 2 D f1 mu per component: the reciprocal field of the
// permanent multipoles taken off and the self term added in one pass
__global__ void k_rhs_finish(int N, double f1, const double* __restrict__ Qc, const double* __restrict__ recip,
                             double* __restrict__ b) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    for (int k = 0; k < 3; ++k) {
        double v = b[a * 3 + k];
        v -= recip[a * 3 + k];
        v += 2.0 * DIELECTRIC * f1 * Qc[a * 10 + 1 + k];
        b[a * 3 + k] = v;
    }
}

// energy[f] = sum of the per-frame components in red[f*16 + 0..3] (+ rec[f*7])
__global__ void k_energy_sum(int F, const double* __restrict__ red, const double* __restrict__ rec, double* __restrict__ e) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    e[f] = red[f * 16 + 0] + red[f * 16 + 1] + red[f * 16 + 2] + (rec ? rec[f * 7] : 0.0);
}

// out = x + y (+ z)
__global__ void k_add3(size_t n, const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                       double* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x[i] + (y ? y[i] : 0.0) + (z ? z[i] : 0.0);
}

// strided copy of one column-block: out[a*w + k] = in[a*win + off + k]
// o[2i] = a[i], o[2i+1] = b[i]
__global__ void k_interleave2(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    o[2 * (size_t)i] = a[i];
    o[2 * (size_t)i + 1] = b[i];
}

__global__ void k_copy_cols(size_t n, int w, int win, int off, const double* __restrict__ in, double* __restrict__ out) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    for (int k = 0; k < w; ++k) out[a * w + k] = in[a * win + off + k];
}

// Qd[a*10] = (c[a*3 + col], 0 ...)  : a dispersion coefficient as a charge-only source
__global__ void k_disp_source(int N, int col, const double* __restrict__ c, double* __restrict__ Qd) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    Qd[a * 10] = c[(size_t)i * 3 + col];
    for (int k = 1; k < 10; ++k) Qd[a * 10 + k] = 0.0;
}

// dispersion self energy (disp_pme.py:264-289) per site and its c-gradient
__global__ void k_disp_self(int N, double kappa, int pmax, const double* __restrict__ c, double* __restrict__ e_self,
                            double* __restrict__ gc) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    const double k2 = kappa * kappa, k6 = k2 * k2 * k2, k8 = k6 * k2, k10 = k8 * k2;
    const double c6 = c[i * 3], c8 = c[i * 3 + 1], c10 = c[i * 3 + 2];
    double e = -k6 / 12.0 * c6 * c6;
    gc[a * 3 + 0] += -k6 / 6.0 * c6;
    if (pmax >= 8) { e += -k8 / 48.0 * c8 * c8; gc[a * 3 + 1] += -k8 / 24.0 * c8; }
    if (pmax >= 10) { e += -k10 / 240.0 * c10 * c10; gc[a * 3 + 2] += -k10 / 120.0 * c10; }
    e_self[a] = e;
}

// gc[a*3 + col] += gq[a*10]   (recip charge gradient of one dispersion grid)
__global__ void k_disp_collect(size_t n, int col, const double* __restrict__ gq, double* __restrict__ gc) {
    size_t a = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) gc[a * 3 + col] += gq[a * 10];
}

// generic per-frame summed vector add: out[f*w+k] += in[f*w+k]
__global__ void k_vadd(size_t n, const double* __restrict__ in, double* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += in[i];
}

// ----------------------------------------------------------------- device-resident PCG


// ---------------------------------------------------------------------------
// Device-resident preconditioned CG in the Chronopoulos-Gear form (one global
// reduction point per iteration; Chronopoulos & Gear 1989):
//   w = A z,  delta = z.w,  gamma = r.z
//   beta = gamma / gamma_prev,  alpha = gamma / (delta - beta gamma / alpha_prev)
//   p = z + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,  z = M^-1 r
// (s = A p by recurrence; the first step after a start takes beta = 0).  Every
// per-site kernel runs on blocks of PCG_BLOCK sites and leaves one partial per
// block; each consumer block reduces the partials of its frame in a fixed
// order, so all blocks see the same bits.  Partials of frame f (CgParts): nw
// slots of z.w (one per block of the kernel that forms w: the gather epilogue
// or k_cg_w), then r.z buffers 0 and 1 and max|r| buffers 0 and 1, nb each.
// The update of parity c reads buffers c and writes buffers c^1.  A frame is
// flagged converged by the update that finds max|r| < tol (the product before
// it is spent); max|r| runs over sites with pol > 0.001 A^3 (pme.py:393-395).
// ---------------------------------------------------------------------------
constexpr int PCG_BLOCK = 128;  // 96 blocks at 12,288 sites: 0.7 % faster per evaluation than 256 (64: slower)


__device__ inline double wave_allsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_allmax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// three totals at once (max a, sum b, sum c) of partial arrays; valid in every
// thread; sh holds >= 48 doubles
__device__ inline void parts_total3(const double* __restrict__ a, const double* __restrict__ b,
                                    const double* __restrict__ c, int nb, double* sh, double& ma, double& sb,
                                    double& sc) {
    double v0 = 0, v1 = 0, v2 = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        v0 = fmax(v0, a[k]);
        v1 += b[k];
        v2 += c[k];
    }
    v0 = wave_allmax(v0); v1 = wave_allsum(v1); v2 = wave_allsum(v2);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[w] = v0; sh[16 + w] = v1; sh[32 + w] = v2; }
    __syncthreads();
    ma = sh[0]; sb = sh[16]; sc = sh[32];
    const int nw = (int)(blockDim.x >> 6);
    for (int k = 1; k < nw; ++k) { ma = fmax(ma, sh[k]); sb += sh[16 + k]; sc += sh[32 + k]; }
    __syncthreads();
}

// The CG update's three totals in one pass: max of the max|r| partials and sum
// of the r.z partials (nb slots each), sum of the z.w partials (nw slots); each
// total is formed in the same order as parts_total3 forms it; sh >= 48 doubles
__device__ inline void parts_total_cg(const double* __restrict__ mxp, const double* __restrict__ rzp, int nb,
                                      const double* __restrict__ zwp, int nw, double* sh, double& res,
                                      double& gamma, double& delta) {
    double v0 = 0, v1 = 0, v2 = 0;
    const int n = nb > nw ? nb : nw;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        if (k < nb) {
            v0 = fmax(v0, mxp[k]);
            v1 += rzp[k];
        }
        if (k < nw) v2 += zwp[k];
    }
    v0 = wave_allmax(v0); v1 = wave_allsum(v1); v2 = wave_allsum(v2);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[w] = v0; sh[16 + w] = v1; sh[32 + w] = v2; }
    __syncthreads();
    res = sh[0]; gamma = sh[16]; delta = sh[32];
    const int nwv = (int)(blockDim.x >> 6);
    for (int k = 1; k < nwv; ++k) { res = fmax(res, sh[k]); gamma += sh[16 + k]; delta += sh[32 + k]; }
    __syncthreads();
}

// block reduction of one value into every thread
template <bool MAX>
__device__ inline double block_red(double v, double* sh) {
    v = MAX ? wave_allmax(v) : wave_allsum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = sh[0];
    const int nw = (int)(blockDim.x >> 6);
    for (int k = 1; k < nw; ++k) t = MAX ? fmax(t, sh[k]) : t + sh[k];
    __syncthreads();
    return t;
}

__device__ inline double inv_diag(double pol) { return (pol < 1e-8 ? 1e-8 : pol) / DIELECTRIC; }

// the product inputs derived from z: its R copy for the pair kernel and the
// fractional dipoles A^T z (slots 1..3 of the 10-wide fractional multipoles)
template <class R>
__device__ inline void cg_derived(const BoxInfo* __restrict__ boxes, int f, size_t a, const double* u, R* zR, R* fq) {
    if (zR)
        for (int k = 0; k < 3; ++k) zR[a * 3 + k] = (R)u[k];
    if (!fq) return;
    const double* A = boxes[f].A;
    R* o = fq + a * 10;
    for (int j = 0; j < 3; ++j) o[1 + j] = (R)(A[0 * 3 + j] * u[0] + A[1 * 3 + j] * u[1] + A[2 * 3 + j] * u[2]);
}

// z = M^-1 r at site a: its three rows of the cluster inverse (pd: 3 rows of 12,
// column blocks a, then the partners pj[0..2]; see k_pc_clusters) or, without
// a cluster (pj[0] < 0), the diagonal m = alpha / DIELECTRIC.  rj: the partners'
// r (zero for unused partners, whose coefficients are zero too).
constexpr int PC_ROW = 12;
__device__ inline void pc_apply(const int* pj, const double* __restrict__ pd, double m, const double* ri,
                                const double (*rj)[3], double* u) {
    if (pj[0] < 0) {
        for (int k = 0; k < 3; ++k) u[k] = m * ri[k];
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double* row = pd + k * PC_ROW;
        double v = row[0] * ri[0] + row[1] * ri[1] + row[2] * ri[2];
#pragma unroll
        for (int t = 0; t < 3; ++t)
            v += row[3 * t + 3] * rj[t][0] + row[3 * t + 4] * rj[t][1] + row[3 * t + 5] * rj[t][2];
        u[k] = v;
    }
}

// r = b - Ax (Ax null: x = 0, r = b), z = M^-1 r and the product inputs derived from z (with the spline
// coordinates wfrac of the solve's precision R); r into buffer 0 of its two,
// partials into buffers 0.  pcj/pcd: block preconditioner (pcj null: diagonal).
// xzero (may be null): the iterate, set to zero here (a cold start).
template <class R>
__global__ void __launch_bounds__(PCG_BLOCK) k_cg_start(int N, const BoxInfo* __restrict__ boxes,
                                                        const double* __restrict__ wd, const double* __restrict__ pol,
                                                        const double* __restrict__ b, const double* __restrict__ Ax,
                                                        double* __restrict__ r, double* __restrict__ z,
                                                        R* __restrict__ zR, R* __restrict__ wfrac, R* __restrict__ fq,
                                                        const int* __restrict__ pcj, const double* __restrict__ pcd,
                                                        CgParts parts, double* __restrict__ xzero) {
    __shared__ double sh[16];
    const int f = blockIdx.y;
    const int i = blockIdx.x * PCG_BLOCK + threadIdx.x;
    double rz = 0, mx = 0;
    if (i < N) {
        const size_t a = (size_t)f * N + i;
        double rv[3], rj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, u[3];
        if (xzero)
            for (int k = 0; k < 3; ++k) xzero[a * 3 + k] = 0.0;
        for (int k = 0; k < 3; ++k) {
            rv[k] = Ax ? b[a * 3 + k] - Ax[a * 3 + k] : b[a * 3 + k];
            r[a * 3 + k] = rv[k];
            if (pol[i] > 0.001) mx = fmax(mx, fabs(rv[k]));
            if (wd) wfrac[a * 3 + k] = (R)wd[a * 3 + k];
        }
        int pj[3] = {-1, -1, -1};
        if (pcj)
            for (int t = 0; t < 3; ++t) pj[t] = pcj[a * 3 + t];
        for (int t = 0; t < 3; ++t)
            if (pj[t] >= 0) {
                const size_t aj = (size_t)f * N + pj[t];
                for (int k = 0; k < 3; ++k) rj[t][k] = Ax ? b[aj * 3 + k] - Ax[aj * 3 + k] : b[aj * 3 + k];
            }
        pc_apply(pj, pcd + a * 3 * PC_ROW, inv_diag(pol[i]), rv, rj, u);
        for (int k = 0; k < 3; ++k) {
            z[a * 3 + k] = u[k];
            rz += rv[k] * u[k];
        }
        if (fq) {
            R* o = fq + a * 10;
            o[0] = 0;
            for (int k = 4; k < 10; ++k) o[k] = 0;
        }
        cg_derived<R>(boxes, f, a, u, zR, fq);
    }
    rz = block_red<false>(rz, sh);
    mx = block_red<true>(mx, sh);
    if (threadIdx.x == 0) {
        parts.rz(f, 0)[blockIdx.x] = rz;
        parts.mx(f, 0)[blockIdx.x] = mx;
    }
}

// state after k_cg_start (one block per frame): residual and convergence; mode 1
// (fresh start) also resets the iteration count and restarts the direction for
// the next update (of parity 0), mode 2 (residual replaced) keeps both
__global__ void __launch_bounds__(PCG_BLOCK) k_cg_publish(int mode, double tol, CgParts parts,
                                                          PcgState* __restrict__ st) {
    __shared__ double sh[48];
    const int f = blockIdx.x;
    double mx, g, unused;
    parts_total3(parts.mx(f, 0), parts.rz(f, 0), parts.rz(f, 0), parts.nb, sh, mx, g, unused);
    if (threadIdx.x != 0) return;
    PcgState s = st[f];
    if (mode == 1) {
        s.iter = 0;
        s.gam[1] = INFINITY;
        s.alp[1] = 0;
    }
    s.res = mx;
    s.done = mx < tol ? 1 : 0;
    st[f] = s;
}

// w = A z = real + recip + d z (d: the diagonal of the polarisation matrix); partial z.w
__global__ void __launch_bounds__(PCG_BLOCK) k_cg_w(int N, double f1, const double* __restrict__ pol,
                                                    const double* __restrict__ z, const double* __restrict__ real,
                                                    const double* __restrict__ recip, double* __restrict__ w,
                                                    CgParts parts, const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    __shared__ double sh[16];
    const int f = blockIdx.y;
    const int i = blockIdx.x * PCG_BLOCK + threadIdx.x;
    double zw = 0;
    if (i < N) {
        const size_t a = (size_t)f * N + i;
        const double ap = pol[i] < 1e-8 ? 1e-8 : pol[i];
        const double d = DIELECTRIC / ap - 2.0 * DIELECTRIC * f1;
        for (int k = 0; k < 3; ++k) {
            const double zv = z[a * 3 + k];
            const double wv = real[a * 3 + k] + (recip ? recip[a * 3 + k] : 0.0) + d * zv;
            w[a * 3 + k] = wv;
            zw += zv * wv;
        }
    }
    zw = block_red<false>(zw, sh);
    if (threadIdx.x == 0) parts.zw(f)[blockIdx.x] = zw;
}

// One CG step of parity cur (see above); nwu: z.w partials written by the
// product.  A frame whose residual (buffer cur) is below tol is flagged
// converged and left unchanged.
template <class R>
__global__ void __launch_bounds__(PCG_BLOCK) k_cg_update(int N, int nwu, int cur, double tol,
                                                         const BoxInfo* __restrict__ boxes,
                                                         const double* __restrict__ pol, double* __restrict__ x,
                                                         double* __restrict__ r, double* __restrict__ z,
                                                         double* __restrict__ p, double* __restrict__ s,
                                                         const double* __restrict__ w, R* __restrict__ zR,
                                                         R* __restrict__ fq, const int* __restrict__ pcj,
                                                         const double* __restrict__ pcd, CgParts parts,
                                                         PcgState* __restrict__ st) {
    __shared__ double sh[48];
    const int f = blockIdx.y;
    const int nx = cur ^ 1;
    const int i = blockIdx.x * PCG_BLOCK + threadIdx.x;
    // r and s have two buffers: this update reads buffer cur and writes cur ^ 1,
    // so a site can form its partner's new residual from values no thread writes
    const size_t nf3 = (size_t)gridDim.y * N * 3;
    const double* rc = r + (size_t)cur * nf3;
    double* rn = r + (size_t)nx * nf3;
    const double* sc = s + (size_t)cur * nf3;
    double* sn = s + (size_t)nx * nf3;
    // Every load that does not depend on the step lengths is issued before the
    // reductions that give them, so the wave waits for memory once, not four
    // times in a row (state, partial sums, site vectors, partner vectors).
    const double gp = st[f].gam[nx], ap = st[f].alp[nx];
    const int done = st[f].done;
    const size_t a = (size_t)f * N + (i < N ? i : 0);
    double zv[3] = {0, 0, 0}, pv0[3] = {0, 0, 0}, wv[3] = {0, 0, 0}, sv0[3] = {0, 0, 0}, rv0[3] = {0, 0, 0};
    double wj[3][3] = {}, sj[3][3] = {}, rj0[3][3] = {};
    double poli = 0.0;
    int pj[3] = {-1, -1, -1};
    if (i < N) {
        for (int k = 0; k < 3; ++k) {
            const size_t e = a * 3 + k;
            zv[k] = z[e];
            pv0[k] = p[e];
            wv[k] = w[e];
            sv0[k] = sc[e];
            rv0[k] = rc[e];
        }
        poli = pol[i];
        if (pcj)
            for (int t = 0; t < 3; ++t) pj[t] = pcj[a * 3 + t];
        for (int t = 0; t < 3; ++t)
            if (pj[t] >= 0) {
                const size_t aj = (size_t)f * N + pj[t];
                for (int k = 0; k < 3; ++k) {
                    const size_t e = aj * 3 + k;
                    wj[t][k] = w[e];
                    sj[t][k] = sc[e];
                    rj0[t][k] = rc[e];
                }
            }
    }
    if (done) return;   // the same value in every thread of the block
    double res, gamma, delta;
    parts_total_cg(parts.mx(f, cur), parts.rz(f, cur), parts.nb, parts.zw(f), nwu, sh, res, gamma, delta);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (res < tol) {
        if (lead) {
            st[f].res = res;
            st[f].done = 1;
        }
        return;
    }
    const bool restart = !(gp < INFINITY) || gp == 0.0 || ap == 0.0;
    const double beta = restart ? 0.0 : gamma / gp;
    const double den = restart ? delta : delta - beta * gamma / ap;
    const double alpha = den != 0.0 ? gamma / den : 0.0;
    double lrz = 0, lmx = 0;
    if (i < N) {
        double u[3], rv[3], rj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 0; k < 3; ++k) {
            const size_t e = a * 3 + k;
            const double pv = restart ? zv[k] : zv[k] + beta * pv0[k];
            const double sv = restart ? wv[k] : wv[k] + beta * sv0[k];
            p[e] = pv;
            sn[e] = sv;
            x[e] += alpha * pv;
            rv[k] = rv0[k] - alpha * sv;
            rn[e] = rv[k];
            if (poli > 0.001) lmx = fmax(lmx, fabs(rv[k]));
        }
        for (int t = 0; t < 3; ++t)
            if (pj[t] >= 0)
                for (int k = 0; k < 3; ++k) {
                    const double sv = restart ? wj[t][k] : wj[t][k] + beta * sj[t][k];
                    rj[t][k] = rj0[t][k] - alpha * sv;
                }
        pc_apply(pj, pcd + a * 3 * PC_ROW, inv_diag(poli), rv, rj, u);
        for (int k = 0; k < 3; ++k) {
            z[a * 3 + k] = u[k];
            lrz += rv[k] * u[k];
        }
        cg_derived<R>(boxes, f, a, u, zR, fq);
    }
    lrz = block_red<false>(lrz, sh);
    lmx = block_red<true>(lmx, sh);
    if (threadIdx.x == 0) {
        parts.rz(f, nx)[blockIdx.x] = lrz;
        parts.mx(f, nx)[blockIdx.x] = lmx;
    }
    if (lead) {
        st[f].gam[cur] = gamma;
        st[f].alp[cur] = alpha;
        st[f].res = res;
        st[f].iter += 1;
    }
}

// per-type sums of a per-site array: block (t * d + c, f) of 1024 threads; each
// thread keeps four accumulators over four strided site sequences (four loads
// in flight), then a fixed-order block reduction: the summation order is fixed
constexpr int TR_BLOCK = 1024;
__global__ void __launch_bounds__(TR_BLOCK) k_type_reduce(const double* __restrict__ in, int N, int d,
                                                          const int* __restrict__ tmap, double* __restrict__ out,
                                                          int n_types) {
    __shared__ double sh[16];
    const int f = blockIdx.y, t = blockIdx.x / d, c = blockIdx.x - t * d;
    const double* src = in + (size_t)f * N * d + c;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i0 = threadIdx.x; i0 < N; i0 += 4 * TR_BLOCK) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * TR_BLOCK;
            if (i < N && tmap[i] == t) a[u] += src[(size_t)i * d];
        }
    }
    const double acc = block_sum((a[0] + a[1]) + (a[2] + a[3]), sh);
    if (threadIdx.x == 0) out[((size_t)f * n_types + t) * d + c] = acc;
}

// energy of frame f: ((real + self) + polarisation) + reciprocal, from the
// per-frame sums red[f], red[F + f], red[2F + f] and rec[7 f]
__global__ void k_energy_total(int F, const double* __restrict__ red, const double* __restrict__ rec,
                               double* __restrict__ e) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < F) e[f] = red[f] + red[F + f] + red[2 * F + f] + rec[7 * f];
}

// up to six device-to-device copies in one launch (grid.y selects the copy)
struct CopyJobs {
    const double* src[6];
    double* dst[6];
    size_t n[6];
};
__global__ void __launch_bounds__(256) k_copy4(CopyJobs j) {
    const int c = blockIdx.y;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < j.n[c]) j.dst[c][i] = j.src[c][i];
}
}  // namespace admp

namespace admp {

// Electrostatic potential at probe points from global Cartesian multipoles
// (dmff/admp/pme.py:236-296, esp_kernel): for probe g and site i with
// R = r_i - g (nm), r = sqrt(R.R + 1e-16),
//   phi = q/r - (mu/10 + U).R / r^3 + 3 (Theta/300):RR / r^5,  times DIELECTRIC/10
// (mu in e A, Theta in e A^2 as the library holds them, U in e nm).  One block
// per probe; the sites are summed in a fixed order (per-thread strided partial
// sums, then the block sum).
__global__ void k_esp(int N, int G, const double* __restrict__ pos, const double* __restrict__ Qc,
                      const double* __restrict__ U, const double* __restrict__ grid, double* __restrict__ out) {
    __shared__ double sh[16];
    const int g = blockIdx.x;
    if (g >= G) return;
    const double gx = grid[g * 3], gy = grid[g * 3 + 1], gz = grid[g * 3 + 2];
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const double Rx = pos[i * 3] - gx, Ry = pos[i * 3 + 1] - gy, Rz = pos[i * 3 + 2] - gz;
        const double r2 = Rx * Rx + Ry * Ry + Rz * Rz + 1e-16;
        const double r = sqrt(r2);
        const double rr1 = 1.0 / r, rr2 = rr1 * rr1, rr3 = rr1 * rr2, rr5 = 3.0 * rr3 * rr2;
        const double* c = Qc + (size_t)i * 10;
        double phi = c[0] * rr1;
        double scd = (c[1] * Rx + c[2] * Ry + c[3] * Rz) / 10.0;
        const double scu = U ? (U[i * 3] * Rx + U[i * 3 + 1] * Ry + U[i * 3 + 2] * Rz) : 0.0;
        phi -= (scd + scu) * rr3;
        const double qxx = c[4] / 300.0, qyy = c[5] / 300.0, qzz = c[6] / 300.0;
        const double qxy = c[7] / 300.0, qxz = c[8] / 300.0, qyz = c[9] / 300.0;
        double scq = Rx * (qxx * Rx + qxy * Ry + qxz * Rz);
        scq += Ry * (qxy * Rx + qyy * Ry + qyz * Rz);
        scq += Rz * (qxz * Rx + qyz * Ry + qzz * Rz);
        phi += scq * rr5;
        acc += phi * DIELECTRIC * 0.1;
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[g] = s;
}

}  // namespace admp

namespace admp {

// OpenMM-style force accumulation (backend/openmm_dmff_plugin/platforms/cuda/src/
// kernels/DMFFForce.cu:1-10): force = -scale * dE/dpos of the atom that
// atom_index maps to, added to the fixed-point buffers (2^32 units) laid out as
// [x..., y..., z...] with row length padded.  One thread per buffer slot, so
// each slot is written by one thread (no atomics).
__global__ void k_add_forces_fixed(int N, int padded, double scale, const double* __restrict__ dEdpos,
                                   const int* __restrict__ atom_index, long long* __restrict__ buf) {
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += blockDim.x * gridDim.x) {
        const int src = atom_index ? atom_index[a] : a;
        for (int c = 0; c < 3; ++c) {
            const double f = -scale * dEdpos[(size_t)src * 3 + c];
            buf[a + (size_t)c * padded] += (long long)(f * 4294967296.0);
        }
    }
}

}  // namespace admp
