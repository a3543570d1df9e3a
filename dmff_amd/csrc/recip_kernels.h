// Reciprocal-space kernels: order-6 B-spline spreading of multipoles, a
// hand-written mixed-radix 3-D FFT for any grid size, the influence function
// (energy and box derivative), and the gather of potential derivatives.
//
// Fractional formulation.  R = A r with A[j][c] = K_j Binv[c][j] is the grid
// coordinate of a site (dmff/admp/recip.py:30-67); m = ceil(R) and w = m - R.
// The reference spreads q theta + mu.grad(theta) + Theta:grad grad(theta)/3 with
// the Cartesian derivatives written as -A[i][j] d/du_j (recip.py:207,251); in
// terms of derivatives with respect to R this is a fractional multipole
//   fmu = T mu,  fTheta = T Theta T^T / 3,  T = A^T,
// so spreading and gathering only need 1-D spline values per dimension.
// (For orthorhombic boxes A is diagonal and T = A; for triclinic boxes T = A^T
// reproduces the reference's contraction exactly.)
#pragma once
#include "admp_common.h"

namespace admp {

// Cardinal B-spline of order 6 at u = w + k, k = 0..5, and its derivatives with
// respect to R = -u (orders 0..3), by the recursion
//   M_n(u) = (u M_{n-1}(u) + (n-u) M_{n-1}(u-1)) / (n-1),
//   M_n' (u) = M_{n-1}(u) - M_{n-1}(u-1).
// Piece k of the reference's piecewise quintics (recip.py:74-177) is M_6(w+k).
template <class R>
__device__ __forceinline__ void bspline6(R w, R M[6], R d1[6], R d2[6], R d3[6]) {
    R a[6];
    // order 3 (need M_3 for the third derivative)
    R a2_0 = w, a2_1 = (R)1 - w;
    R a3[3];
    a3[0] = (w * a2_0) * (R)0.5;
    a3[1] = ((w + (R)1) * a2_1 + ((R)2 - w) * a2_0) * (R)0.5;
    a3[2] = (((R)1 - w) * a2_1) * (R)0.5;
    R a4[4];
    for (int k = 0; k < 4; ++k) {
        R cur = k < 3 ? a3[k] : (R)0;
        R prv = k > 0 ? a3[k - 1] : (R)0;
        a4[k] = ((w + (R)k) * cur + ((R)4 - w - (R)k) * prv) * (R)(1.0 / 3.0);
    }
    R a5[5];
    for (int k = 0; k < 5; ++k) {
        R cur = k < 4 ? a4[k] : (R)0;
        R prv = k > 0 ? a4[k - 1] : (R)0;
        a5[k] = ((w + (R)k) * cur + ((R)5 - w - (R)k) * prv) * (R)0.25;
    }
    for (int k = 0; k < 6; ++k) {
        R cur = k < 5 ? a5[k] : (R)0;
        R prv = k > 0 ? a5[k - 1] : (R)0;
        a[k] = ((w + (R)k) * cur + ((R)6 - w - (R)k) * prv) * (R)0.2;
    }
    for (int k = 0; k < 6; ++k) {
        M[k] = a[k];
        R c5 = k < 5 ? a5[k] : (R)0, p5 = k > 0 ? a5[k - 1] : (R)0;
        d1[k] = -(c5 - p5);                                  // dM/dR = -M'
        R c4 = k < 4 ? a4[k] : (R)0, p4 = (k >= 1 && k - 1 < 4) ? a4[k - 1] : (R)0,
          q4 = (k >= 2 && k - 2 < 4) ? a4[k - 2] : (R)0;
        d2[k] = c4 - (R)2 * p4 + q4;                          // M''
        R c3 = k < 3 ? a3[k] : (R)0, p3 = (k >= 1 && k - 1 < 3) ? a3[k - 1] : (R)0,
          q3 = (k >= 2 && k - 2 < 3) ? a3[k - 2] : (R)0, s3 = (k >= 3 && k - 3 < 3) ? a3[k - 3] : (R)0;
        d3[k] = -(c3 - (R)3 * p3 + (R)3 * q3 - s3);          // -M'''
    }
}

__device__ __forceinline__ int wrapi(int a, int K) {
    int r = a % K;
    return r < 0 ? r + K : r;
}

// Per-site set-up: grid base m, offset w, and fractional multipoles (10 words:
// q, fmu0..2, fTheta 00 11 22 01 02 12).  Sources: Qc (F,N,10) Cartesian
// [q mu Theta(xx yy zz xy xz yz)] and/or U (F,N,3) added to mu.  mode bit 0: use
// Qc, bit 1: add U.  use_dA selects the dispersion grid's A.
template <class R>
__global__ void k_spline_prep(int N, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
                              const double* __restrict__ Qc, const double* __restrict__ U, int mode, int use_dA,
                              int4* __restrict__ mbase, R* __restrict__ wfrac, R* __restrict__ fq) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const BoxInfo& b = boxes[f];
    const double* A = use_dA ? b.dA : b.A;
    const size_t a = (size_t)f * N + i;
    const double4 p = pos[a];
    double Rf[3];
    int4 m;
    for (int j = 0; j < 3; ++j) Rf[j] = A[j * 3 + 0] * p.x + A[j * 3 + 1] * p.y + A[j * 3 + 2] * p.z;
    m.x = (int)ceil(Rf[0]); m.y = (int)ceil(Rf[1]); m.z = (int)ceil(Rf[2]); m.w = 0;
    mbase[a] = m;
    wfrac[a * 3 + 0] = (R)((double)m.x - Rf[0]);
    wfrac[a * 3 + 1] = (R)((double)m.y - Rf[1]);
    wfrac[a * 3 + 2] = (R)((double)m.z - Rf[2]);
    double q = 0, mu[3] = {0, 0, 0}, Th[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (mode & 1) {
        const double* c = Qc + a * 10;
        q = c[0]; mu[0] = c[1]; mu[1] = c[2]; mu[2] = c[3];
        Th[0][0] = c[4]; Th[1][1] = c[5]; Th[2][2] = c[6];
        Th[0][1] = Th[1][0] = c[7]; Th[0][2] = Th[2][0] = c[8]; Th[1][2] = Th[2][1] = c[9];
    }
    if (mode & 2) {
        mu[0] += U[a * 3 + 0]; mu[1] += U[a * 3 + 1]; mu[2] += U[a * 3 + 2];
    }
    // T = A^T: fmu_j = sum_c A[c][j] mu_c ; fTheta = T Theta T^T / 3
    double fm[3], TT[3][3];
    for (int j = 0; j < 3; ++j) fm[j] = A[0 * 3 + j] * mu[0] + A[1 * 3 + j] * mu[1] + A[2 * 3 + j] * mu[2];
    double TA[3][3];  // Theta A  : TA[c][k] = sum_d Theta[c][d] A[d][k]
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) TA[c][k] = Th[c][0] * A[0 * 3 + k] + Th[c][1] * A[1 * 3 + k] + Th[c][2] * A[2 * 3 + k];
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
            TT[j][k] = (A[0 * 3 + j] * TA[0][k] + A[1 * 3 + j] * TA[1][k] + A[2 * 3 + j] * TA[2][k]) * (1.0 / 3.0);
    R* o = fq + a * 10;
    o[0] = (R)q; o[1] = (R)fm[0]; o[2] = (R)fm[1]; o[3] = (R)fm[2];
    o[4] = (R)TT[0][0]; o[5] = (R)TT[1][1]; o[6] = (R)TT[2][2];
    o[7] = (R)TT[0][1]; o[8] = (R)TT[0][2]; o[9] = (R)TT[1][2];
}

// Per-site grid base m and offset w only (positions are fixed during a solve).
// Also zeroes n_zero ints of zero_a and zero_b (the counters of the site sort
// that follows; null when there is none), grid-stride over all threads.
__global__ void k_grid_base(int N, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes, int use_dA,
                            int4* __restrict__ mbase, double* __restrict__ wd, int* __restrict__ zero_a,
                            int* __restrict__ zero_b, size_t n_zero) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (zero_a) {
        const size_t nthreads = (size_t)gridDim.x * blockDim.x * gridDim.y;
        for (size_t k = (size_t)f * gridDim.x * blockDim.x + i; k < n_zero; k += nthreads) {
            zero_a[k] = 0;
            zero_b[k] = 0;
        }
    }
    if (i >= N) return;
    const BoxInfo& b = boxes[f];
    const double* A = use_dA ? b.dA : b.A;
    const size_t a = (size_t)f * N + i;
    const double4 p = pos[a];
    int mm[3];
    for (int j = 0; j < 3; ++j) {
        const double Rf = A[j * 3 + 0] * p.x + A[j * 3 + 1] * p.y + A[j * 3 + 2] * p.z;
        mm[j] = (int)ceil(Rf);
        wd[a * 3 + j] = (double)mm[j] - Rf;
    }
    mbase[a] = make_int4(mm[0], mm[1], mm[2], 0);
}

// Fractional multipoles from Cartesian sources (mode bit 0: Qc, bit 1: add U);
// also writes the R-precision copy of w.
template <class R>
__global__ void k_frac_mult(int N, const BoxInfo* __restrict__ boxes, int use_dA, const double* __restrict__ Qc,
                            const double* __restrict__ U, int mode, const double* __restrict__ wd,
                            R* __restrict__ wfrac, R* __restrict__ fq) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const BoxInfo& b = boxes[f];
    const double* A = use_dA ? b.dA : b.A;
    const size_t a = (size_t)f * N + i;
    for (int j = 0; j < 3; ++j) wfrac[a * 3 + j] = (R)wd[a * 3 + j];
    double q = 0, mu[3] = {0, 0, 0}, Th[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (mode & 1) {
        const double* c = Qc + a * 10;
        q = c[0]; mu[0] = c[1]; mu[1] = c[2]; mu[2] = c[3];
        Th[0][0] = c[4]; Th[1][1] = c[5]; Th[2][2] = c[6];
        Th[0][1] = Th[1][0] = c[7]; Th[0][2] = Th[2][0] = c[8]; Th[1][2] = Th[2][1] = c[9];
    }
    if (mode & 2) { mu[0] += U[a * 3 + 0]; mu[1] += U[a * 3 + 1]; mu[2] += U[a * 3 + 2]; }
    double fm[3], TA[3][3];
    for (int j = 0; j < 3; ++j) fm[j] = A[0 * 3 + j] * mu[0] + A[1 * 3 + j] * mu[1] + A[2 * 3 + j] * mu[2];
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) TA[c][k] = Th[c][0] * A[0 * 3 + k] + Th[c][1] * A[1 * 3 + k] + Th[c][2] * A[2 * 3 + k];
    R* o = fq + a * 10;
    o[0] = (R)q; o[1] = (R)fm[0]; o[2] = (R)fm[1]; o[3] = (R)fm[2];
    const int pr[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int k = 0; k < 6; ++k) {
        const int j = pr[k][0], l = pr[k][1];
        o[4 + k] = (R)((A[0 * 3 + j] * TA[0][l] + A[1 * 3 + j] * TA[1][l] + A[2 * 3 + j] * TA[2][l]) * (1.0 / 3.0));
    }
}

// Scatter spread with device atomics: one thread per (site, k0, k1), six points
// along the contiguous axis.  grid is the real part of a complex array.
template <class R>
__global__ void k_spread_atomic(int N, int K0, int K1, int K2, const int4* __restrict__ mbase,
                                const R* __restrict__ wfrac, const R* __restrict__ fq, int lmax_eff,
                                Cplx<R>* __restrict__ grid) {
    const int f = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / 36, kk = t % 36;
    if (i >= N) return;
    const int k0 = kk / 6, k1 = kk % 6;
    const size_t a = (size_t)f * N + i;
    const int4 m = mbase[a];
    R M0[6], D0[6], E0[6], F0[6], M1[6], D1[6], E1[6], F1[6], M2[6], D2[6], E2[6], F2[6];
    bspline6<R>(wfrac[a * 3 + 0], M0, D0, E0, F0);
    bspline6<R>(wfrac[a * 3 + 1], M1, D1, E1, F1);
    bspline6<R>(wfrac[a * 3 + 2], M2, D2, E2, F2);
    const R* c = fq + a * 10;
    const R P = M0[k0] * M1[k1], Px = D0[k0] * M1[k1], Py = M0[k0] * D1[k1];
    R c0 = c[0] * P, c1 = 0, c2 = 0;
    if (lmax_eff >= 1) {
        c0 += c[1] * Px + c[2] * Py;
        c1 += c[3] * P;
    }
    if (lmax_eff >= 2) {
        const R Pxx = E0[k0] * M1[k1], Pyy = M0[k0] * E1[k1], Pxy = D0[k0] * D1[k1];
        c0 += c[4] * Pxx + c[5] * Pyy + (R)2 * c[7] * Pxy;
        c1 += (R)2 * (c[8] * Px + c[9] * Py);
        c2 += c[6] * P;
    }
    const int g0 = wrapi(m.x + k0 - 3, K0), g1 = wrapi(m.y + k1 - 3, K1);
    Cplx<R>* row = grid + (size_t)f * K0 * K1 * K2 + ((size_t)g0 * K1 + g1) * K2;
    for (int k2 = 0; k2 < 6; ++k2) {
        const R wv = c0 * M2[k2] + c1 * D2[k2] + c2 * E2[k2];
        atomicAdd(&row[wrapi(m.z + k2 - 3, K2)].x, wv);
    }
}

template <class R> __global__ void k_zero_cplx(size_t n, Cplx<R>* __restrict__ g) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] = Cplx<R>{(R)0, (R)0};
}

// ---------------------------------------------------------------------------
// 3-D FFT: one pass per axis.  A block loads LB lines of length K into LDS and
// runs a Stockham autosort transform, one radix per stage; each thread computes
// whole outputs of a radix-f butterfly directly (f complex MACs), so any factor
// (2, 3, 4, 5, ... or a large prime such as 19, 61) is handled by the same code.
// ---------------------------------------------------------------------------
struct FFTPlan {
    int K;
    int nst;
    int fac[12];
};

template <class R>
__global__ void k_fft_pass(Cplx<R>* __restrict__ data, int K0, int K1, int K2, int dim, FFTPlan plan,
                           const Cplx<R>* __restrict__ tw, int inverse, int LB) {
    extern __shared__ unsigned char smem_raw[];
    Cplx<R>* buf0 = reinterpret_cast<Cplx<R>*>(smem_raw);
    const int K = plan.K;
    Cplx<R>* buf1 = buf0 + (size_t)LB * K;
    const size_t G = (size_t)K0 * K1 * K2;
    Cplx<R>* d = data + (size_t)blockIdx.y * G;
    const int nlines = (int)(G / K);
    const int l0 = blockIdx.x * LB;
    const int nl = min(LB, nlines - l0);
    if (nl <= 0) return;
    const int nt = blockDim.x;
    // load
    if (dim == 2) {
        for (int idx = threadIdx.x; idx < nl * K; idx += nt) {
            int li = idx / K, e = idx % K;
            buf0[e * LB + li] = d[(size_t)(l0 + li) * K + e];
        }
    } else {
        const size_t stride = dim == 1 ? (size_t)K2 : (size_t)K1 * K2;
        for (int idx = threadIdx.x; idx < nl * K; idx += nt) {
            int e = idx / nl, li = idx % nl;
            int l = l0 + li;
            size_t base = dim == 1 ? (size_t)(l / K2) * K1 * K2 + (l % K2) : (size_t)l;
            buf0[e * LB + li] = d[base + e * stride];
        }
    }
    __syncthreads();
    Cplx<R>* in = buf0;
    Cplx<R>* out = buf1;
    int Ns = 1;
    for (int s = 0; s < plan.nst; ++s) {
        const int fr = plan.fac[s];
        const int Kf = K / fr;
        const int step = K / (Ns * fr);
        for (int idx = threadIdx.x; idx < nl * K; idx += nt) {
            const int li = idx % nl, t = idx / nl;
            const int j = t / fr, k = t % fr;
            const int kj = j % Ns;
            const int inc = ((kj + k * Ns) * step) % K;
            int acc = 0;
            R sx = 0, sy = 0;
            for (int r = 0; r < fr; ++r) {
                Cplx<R> x = in[(j + r * Kf) * LB + li];
                Cplx<R> w = tw[acc];
                if (inverse) w.y = -w.y;
                sx += x.x * w.x - x.y * w.y;
                sy += x.x * w.y + x.y * w.x;
                acc += inc;
                if (acc >= K) acc -= K;
            }
            const int o = (j / Ns) * Ns * fr + kj + k * Ns;
            out[o * LB + li] = Cplx<R>{sx, sy};
        }
        __syncthreads();
        Cplx<R>* tmp = in; in = out; out = tmp;
        Ns *= fr;
    }
    // store
    if (dim == 2) {
        for (int idx = threadIdx.x; idx < nl * K; idx += nt) {
            int li = idx / K, e = idx % K;
            d[(size_t)(l0 + li) * K + e] = in[e * LB + li];
        }
    } else {
        const size_t stride = dim == 1 ? (size_t)K2 : (size_t)K1 * K2;
        for (int idx = threadIdx.x; idx < nl * K; idx += nt) {
            int e = idx / nl, li = idx % nl;
            int l = l0 + li;
            size_t base = dim == 1 ? (size_t)(l / K2) * K1 * K2 + (l % K2) : (size_t)l;
            d[base + e * stride] = in[e * LB + li];
        }
    }
}

// ---------------------------------------------------------------------------
// Influence function.  kind 0: electrostatic C_k = 2 pi/(V k^2) exp(-k^2/4 kappa^2)
// with k = 0 excluded and a DIELECTRIC factor (recip.py:453-466,474); kinds 6, 8,
// 10: dispersion C_k (recip.py:477-508) with k = 0 included.  For every grid
// point: E += pref C |S/theta|^2, W += pref |S/theta|^2 (-dC/dk^2) k k^T, and the
// spectrum is replaced by 2 pref C/theta^2 S so that the unnormalised inverse
// transform is dE/dQ_mesh.  Per-block partial sums (7 doubles) in `part`.
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_influence(int K0, int K1, int K2, const BoxInfo* __restrict__ boxes, int use_dA,
                            const double* __restrict__ bmod0, const double* __restrict__ bmod1,
                            const double* __restrict__ bmod2, double kappa, int kind, double pref,
                            Cplx<R>* __restrict__ grid, double* __restrict__ part) {
    __shared__ double sh[16];
    const int f = blockIdx.y;
    const size_t G = (size_t)K0 * K1 * K2;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const BoxInfo& b = boxes[f];
    double e = 0, w00 = 0, w11 = 0, w22 = 0, w01 = 0, w02 = 0, w12 = 0;
    if (p < G) {
        const int i0 = (int)(p / ((size_t)K1 * K2));
        const int i1 = (int)((p / K2) % K1);
        const int i2 = (int)(p % K2);
        // FFT-order integers (recip.py:371-381)
        const int m0 = i0 < (K0 + 1) / 2 ? i0 : i0 - K0;
        const int m1 = i1 < (K1 + 1) / 2 ? i1 : i1 - K1;
        const int m2 = i2 < (K2 + 1) / 2 ? i2 : i2 - K2;
        // k = 2 pi Binv m
        const double kx = b.kb[0] * m0 + b.kb[1] * m1 + b.kb[2] * m2;
        const double ky = b.kb[3] * m0 + b.kb[4] * m1 + b.kb[5] * m2;
        const double kz = b.kb[6] * m0 + b.kb[7] * m1 + b.kb[8] * m2;
        const double ksq = kx * kx + ky * ky + kz * kz;
        const double th = bmod0[i0] * bmod1[i1] * bmod2[i2];
        const double ith2 = 1.0 / (th * th);
        Cplx<R>* g = grid + (size_t)f * G + p;
        Cplx<R> S = *g;
        const double s2 = (double)S.x * (double)S.x + (double)S.y * (double)S.y;
        double C = 0, dC = 0;  // C_k and dC/dk^2
        const double V = b.V;
        if (kind == 0) {
            if (p != 0) {
                C = 2.0 * PI / V / ksq * exp(-ksq / (4.0 * kappa * kappa));
                dC = -C * (1.0 / ksq + 1.0 / (4.0 * kappa * kappa));
            }
        } else {
            const double kq = ksq < 1e-16 ? 1e-16 : ksq;
            const double x2 = kq / (4.0 * kappa * kappa), x = sqrt(x2);
            const double ex = exp(-x2), ec = erfc(x);
            const double spi = SQRT_PI;
            double fv, fd, pw;
            if (kind == 6) {
                fv = (1.0 - 2.0 * x2) * ex + 2.0 * x2 * x * spi * ec;
                fd = -6.0 * x * ex + 6.0 * x2 * spi * ec;
                pw = kappa * kappa * kappa / 3.0;
            } else if (kind == 8) {
                fv = (3.0 - 2.0 * x2 + 4.0 * x2 * x2) * ex - 4.0 * x2 * x2 * x * spi * ec;
                fd = (-10.0 * x + 20.0 * x2 * x) * ex - 20.0 * x2 * x2 * spi * ec;
                pw = kappa * kappa * kappa * kappa * kappa / 45.0;
            } else {
                const double x4 = x2 * x2;
                fv = (15.0 - 6.0 * x2 + 4.0 * x4 - 8.0 * x4 * x2) * ex + 8.0 * x4 * x2 * x * spi * ec;
                fd = (-42.0 * x + 28.0 * x2 * x - 56.0 * x4 * x) * ex + 56.0 * x4 * x2 * spi * ec;
                pw = pow(kappa, 7.0) / 1260.0;
            }
            const double c0 = spi * PI / 2.0 / V * pw;
            C = c0 * fv;
            // dC/dk^2 = c0 f'(x) dx/dk^2,  dx/dk^2 = 1/(8 kappa^2 x); zero at the floored k = 0
            dC = ksq < 1e-16 ? 0.0 : c0 * fd / (8.0 * kappa * kappa * x);
        }
        const double wgt = pref * s2 * ith2;
        e = wgt * C;
        const double wd = -wgt * dC;
        w00 = wd * kx * kx; w11 = wd * ky * ky; w22 = wd * kz * kz;
        w01 = wd * kx * ky; w02 = wd * kx * kz; w12 = wd * ky * kz;
        const double scale = 2.0 * pref * C * ith2;
        *g = Cplx<R>{(R)(S.x * scale), (R)(S.y * scale)};
    }
    double vals[7] = {e, w00, w11, w22, w01, w02, w12};
    for (int k = 0; k < 7; ++k) {
        double t = block_sum(vals[k], sh);
        if (threadIdx.x == 0) part[((size_t)f * gridDim.x + blockIdx.x) * 7 + k] = t;
    }
}

// ---------------------------------------------------------------------------
// Gather: fractional potential derivatives fphi (20: 0 | 1 2 3 | 11 22 33 12 13
// 23 | 111 222 333 112 113 122 133 223 233 123) at each site from the
// potential grid phi (real part), then
//   dE/dq = fphi0, dE/dmu = A fphi1, dE/dTheta = A fphi2 A^T / 3,
//   dE/dr = A^T g, g_j = q fphi_j + fmu_k fphi_kj + fTheta_kl fphi_klj,
//   dE/dA += g r^T + mu fphi1^T + 2/3 Theta A fphi2.
// mode 0: full outputs (energy gradients); mode 1: only dE/dmu into out3.
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_gather(int N, int K0, int K1, int K2, const double4* __restrict__ pos,
                         const BoxInfo* __restrict__ boxes, int use_dA, const int4* __restrict__ mbase,
                         const R* __restrict__ wfrac, const R* __restrict__ fq, const double* __restrict__ Qc,
                         const double* __restrict__ U, int qmode, const Cplx<R>* __restrict__ grid, int mode,
                         double* __restrict__ out3, double* __restrict__ g_q, double* __restrict__ g_pos,
                         double* __restrict__ vir_atom) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    const BoxInfo& b = boxes[f];
    const double* A = use_dA ? b.dA : b.A;
    const int4 m = mbase[a];
    R M[3][4][6];
    bspline6<R>(wfrac[a * 3 + 0], M[0][0], M[0][1], M[0][2], M[0][3]);
    bspline6<R>(wfrac[a * 3 + 1], M[1][0], M[1][1], M[1][2], M[1][3]);
    bspline6<R>(wfrac[a * 3 + 2], M[2][0], M[2][1], M[2][2], M[2][3]);
    const Cplx<R>* gbase = grid + (size_t)f * K0 * K1 * K2;
    // fp[c][a][b]: derivative orders (x=c, y=a, z=b), c+a+b <= 3
    double fp[4][4][4];
    for (int c = 0; c < 4; ++c)
        for (int aa = 0; aa < 4; ++aa)
            for (int bb = 0; bb < 4; ++bb) fp[c][aa][bb] = 0.0;
    int gz[6];
    for (int k2 = 0; k2 < 6; ++k2) gz[k2] = wrapi(m.z + k2 - 3, K2);
    for (int k0 = 0; k0 < 6; ++k0) {
        const int g0 = wrapi(m.x + k0 - 3, K0);
        R u[4][4];
        for (int aa = 0; aa < 4; ++aa)
            for (int bb = 0; bb < 4; ++bb) u[aa][bb] = 0;
        for (int k1 = 0; k1 < 6; ++k1) {
            const int g1 = wrapi(m.y + k1 - 3, K1);
            const Cplx<R>* row = gbase + ((size_t)g0 * K1 + g1) * K2;
            R s[4] = {0, 0, 0, 0};
            for (int k2 = 0; k2 < 6; ++k2) {
                const R ph = row[gz[k2]].x;
                s[0] += ph * M[2][0][k2]; s[1] += ph * M[2][1][k2];
                s[2] += ph * M[2][2][k2]; s[3] += ph * M[2][3][k2];
            }
            for (int aa = 0; aa < 4; ++aa)
                for (int bb = 0; aa + bb < 4; ++bb) u[aa][bb] += M[1][aa][k1] * s[bb];
        }
        for (int c = 0; c < 4; ++c)
            for (int aa = 0; c + aa < 4; ++aa)
                for (int bb = 0; c + aa + bb < 4; ++bb) fp[c][aa][bb] += (double)(M[0][c][k0] * u[aa][bb]);
    }
    // first and second derivative tensors
    double p1[3] = {fp[1][0][0], fp[0][1][0], fp[0][0][1]};
    double p2[3][3];
    p2[0][0] = fp[2][0][0]; p2[1][1] = fp[0][2][0]; p2[2][2] = fp[0][0][2];
    p2[0][1] = p2[1][0] = fp[1][1][0]; p2[0][2] = p2[2][0] = fp[1][0][1]; p2[1][2] = p2[2][1] = fp[0][1][1];
    // dE/dmu = A fphi1
    double dmu[3];
    for (int c = 0; c < 3; ++c) dmu[c] = A[c * 3 + 0] * p1[0] + A[c * 3 + 1] * p1[1] + A[c * 3 + 2] * p1[2];
    if (mode == 1) {
        out3[a * 3 + 0] = dmu[0]; out3[a * 3 + 1] = dmu[1]; out3[a * 3 + 2] = dmu[2];
        return;
    }
    // third derivatives p3[j][k][l]
    auto p3 = [&](int j, int k, int l) -> double {
        int o[3] = {0, 0, 0};
        o[j]++; o[k]++; o[l]++;
        return fp[o[0]][o[1]][o[2]];
    };
    const R* c = fq + a * 10;
    const double q = (double)c[0];
    const double fm[3] = {(double)c[1], (double)c[2], (double)c[3]};
    double fT[3][3];
    fT[0][0] = c[4]; fT[1][1] = c[5]; fT[2][2] = c[6];
    fT[0][1] = fT[1][0] = c[7]; fT[0][2] = fT[2][0] = c[8]; fT[1][2] = fT[2][1] = c[9];
    double g[3];
    for (int j = 0; j < 3; ++j) {
        double s = q * p1[j];
        for (int k = 0; k < 3; ++k) s += fm[k] * p2[k][j];
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l) s += fT[k][l] * p3(k, l, j);
        g[j] = s;
    }
    // dE/dTheta = A p2 A^T / 3
    double Ap2[3][3];
    for (int x = 0; x < 3; ++x)
        for (int k = 0; k < 3; ++k) Ap2[x][k] = A[x * 3 + 0] * p2[0][k] + A[x * 3 + 1] * p2[1][k] + A[x * 3 + 2] * p2[2][k];
    double GT[3][3];
    for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y)
            GT[x][y] = (Ap2[x][0] * A[y * 3 + 0] + Ap2[x][1] * A[y * 3 + 1] + Ap2[x][2] * A[y * 3 + 2]) * (1.0 / 3.0);
    double* gq = g_q + a * 10;
    gq[0] = fp[0][0][0];
    gq[1] = dmu[0]; gq[2] = dmu[1]; gq[3] = dmu[2];
    gq[4] = GT[0][0]; gq[5] = GT[1][1]; gq[6] = GT[2][2]; gq[7] = GT[0][1]; gq[8] = GT[0][2]; gq[9] = GT[1][2];
    const double4 pp = pos[a];
    for (int cc = 0; cc < 3; ++cc) g_pos[a * 3 + cc] = A[0 * 3 + cc] * g[0] + A[1 * 3 + cc] * g[1] + A[2 * 3 + cc] * g[2];
    // dE/dA += g r^T + mu p1^T + 2/3 Theta A p2
    double mu[3] = {0, 0, 0}, Th[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (qmode & 1) {
        const double* qc = Qc + a * 10;
        mu[0] = qc[1]; mu[1] = qc[2]; mu[2] = qc[3];
        Th[0][0] = qc[4]; Th[1][1] = qc[5]; Th[2][2] = qc[6];
        Th[0][1] = Th[1][0] = qc[7]; Th[0][2] = Th[2][0] = qc[8]; Th[1][2] = Th[2][1] = qc[9];
    }
    if (qmode & 2) { mu[0] += U[a * 3]; mu[1] += U[a * 3 + 1]; mu[2] += U[a * 3 + 2]; }
    const double r[3] = {pp.x, pp.y, pp.z};
    double* v = vir_atom + a * 9;
    for (int j = 0; j < 3; ++j)
        for (int cc = 0; cc < 3; ++cc) {
            double TAp = 0;  // (Theta A p2)[j][cc]
            for (int x = 0; x < 3; ++x) {
                double Ax = A[x * 3 + 0] * p2[0][cc] + A[x * 3 + 1] * p2[1][cc] + A[x * 3 + 2] * p2[2][cc];
                TAp += Th[j][x] * Ax;
            }
            v[j * 3 + cc] = g[j] * r[cc] + mu[j] * p1[cc] + (2.0 / 3.0) * TAp;
        }
}

}  // namespace admp
