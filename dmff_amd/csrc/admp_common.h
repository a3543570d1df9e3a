// Shared device definitions for the ADMP multipolar PME kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace admp {

// dmff/common/constants.py:3
constexpr double DIELECTRIC = 1389.35455846;
constexpr double SQRT_PI = 1.7724538509055160273;
constexpr double TWO_OVER_SQRT_PI = 1.1283791670955125739;
constexpr double PI = 3.14159265358979323846;
// 2/sqrt(3): weight of the m=1 quadrupole components in the invariant form
constexpr double K1_INV = 1.1547005383792515290;
constexpr double SQRT3 = 1.7320508075688772935;

constexpr int NB_SHIFT = 28;                 // neighbour word: j | nbond << 28
constexpr int NB_MASK = (1 << NB_SHIFT) - 1;

// Per-frame box quantities (all double).  B rows are lattice vectors.
struct BoxInfo {
    double B[9];      // B[p*3+q]
    double Binv[9];   // inverse of B
    double A[9];      // A[j*3+c] = K_j * Binv[c*3+j]  (fractional grid coordinate R = A r)
    double V;         // det(B)
    double kb[9];     // k = kb m : kb[c*3+j] = 2*pi*Binv[c*3+j]
    double dA[9];     // same as A for the dispersion grid
    int ortho;        // 1 if B is diagonal
    int pad;
};

template <class R> struct Cplx { R x, y; };

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  Work items that share data (neighbouring bricks, sites of one
// grid region) are therefore renumbered so that each XCD receives one contiguous
// range: physical block p runs logical item (p % 8) * ceil(n / 8) + p / 8.  The
// grid is launched with 8 * ceil(n / 8) blocks; logical items >= n return.
constexpr int N_XCD = 8;
__host__ __device__ inline int xcd_blocks(int n) { return ((n + N_XCD - 1) / N_XCD) * N_XCD; }
__device__ inline int xcd_item(int p, int n) { return (p % N_XCD) * ((n + N_XCD - 1) / N_XCD) + p / N_XCD; }

// Per-frame scalars of the device-resident CG solve.  gam/alp are kept in two
// slots by iteration parity: the update of parity c reads slot c^1 (written by
// the previous update) and block 0 writes slot c, so no block reads a value
// written in the same launch.
struct PcgState {
    double gam[2];  // r.z used by the update of that parity (+inf: restart the direction)
    double alp[2];  // step length of the update of that parity
    double res;     // max |r| over polarisable sites
    int done, iter, pad0, pad1;
};

// Partial sums of the CG solve, per frame: nw slots of z.w, then r.z buffers 0
// and 1 and max|r| buffers 0 and 1 of nb slots each.
struct CgParts {
    double* base;
    int nb, nw;  // blocks of the per-site kernels; z.w slots per frame
    __device__ double* frame(int f) const { return base + (size_t)f * (nw + 4 * nb); }
    __device__ double* zw(int f) const { return frame(f); }
    __device__ double* rz(int f, int b) const { return frame(f) + nw + b * nb; }
    __device__ double* mx(int f, int b) const { return frame(f) + nw + (2 + b) * nb; }
};

// Convergence test at the start of a CG product: the max|r| partials of parity
// cur (written by the previous update, read by the next one) are reduced by the
// product's first kernel; when they are below tol the frame is flagged
// converged and the product (whose result that update would discard) is skipped.
struct CgCheck {
    CgParts parts;
    PcgState* st;
    double tol;
    int cur, on;
};


template <class R> __device__ __forceinline__ Cplx<R> cmul(Cplx<R> a, Cplx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Minimum image exactly as dmff/admp/spatial.py:13-32, in double:
//   s = dr . Binv ; s -= floor(s + 0.5) ; dr = s . B   (returns n = floor(s+0.5))
__device__ __forceinline__ void pbc_shift(const BoxInfo& b, double& dx, double& dy, double& dz,
                                          double& n0, double& n1, double& n2) {
    double s0 = dx * b.Binv[0] + dy * b.Binv[3] + dz * b.Binv[6];
    double s1 = dx * b.Binv[1] + dy * b.Binv[4] + dz * b.Binv[7];
    double s2 = dx * b.Binv[2] + dy * b.Binv[5] + dz * b.Binv[8];
    n0 = floor(s0 + 0.5);
    n1 = floor(s1 + 0.5);
    n2 = floor(s2 + 0.5);
    s0 -= n0;
    s1 -= n1;
    s2 -= n2;
    dx = s0 * b.B[0] + s1 * b.B[3] + s2 * b.B[6];
    dy = s0 * b.B[1] + s1 * b.B[4] + s2 * b.B[7];
    dz = s0 * b.B[2] + s1 * b.B[5] + s2 * b.B[8];
}

// scale factor for a pair with bonded distance nbond: S[nbond-1] with the
// reference's Python-style wrap for nbond = 0 (dmff/admp/pme.py:1111-1125)
__device__ __forceinline__ int scale_index(int nb, int nscale) {
    int idx = nb - 1;
    if (idx < 0) idx += nscale;
    if (idx >= nscale) idx = nscale - 1;
    if (idx < 0) idx = 0;
    return idx;
}

// sum over groups of `width` consecutive lanes (width = power of two <= 64)
template <class T> __device__ __forceinline__ T group_sum(T v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, width);
    return v;
}

// block-wide sum of one double (blockDim.x multiple of 64, <= 1024); result valid in thread 0
__device__ inline double block_sum(double v, double* sh) {
    v = group_sum(v, 64);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        for (int k = 0; k < nw; ++k) t += sh[k];
    }
    __syncthreads();
    return t;
}

}  // namespace admp
