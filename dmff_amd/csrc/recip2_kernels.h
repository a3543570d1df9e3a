// Reciprocal space, second generation: atomic-free spreading, real-to-complex
// FFT passes with the influence function fused into the middle pass, and a
// multi-lane gather.
//
//   spline_prep   per site: grid base m, offset w = m - R, fractional multipoles
//   grid_sort     sites bucketed by base grid point (counting sort, fixed order)
//   spread_brick  one wavefront per 8x4x2 brick of grid points; the sites whose
//                 6^3 support touches the brick are staged in LDS in chunks and
//                 every lane accumulates its own grid point: each grid value is
//                 written exactly once (no atomics, no zero-fill)
//   fft_z_r2c     two real z-lines packed as one complex FFT, split into halves
//   fft_y         complex passes along y (forward / inverse)
//   fft_x_infl    forward x FFT, influence function (energy, box term,
//                 S <- 2 pref C/theta^2 S), inverse x FFT, in one pass
//   fft_z_c2r     two half spectra -> one complex inverse FFT -> two real lines
//   gather_ml     eight lanes per site (one per x-offset), shuffled reduction
#pragma once
#include "admp_common.h"
#include "pair_kernels.h"
#include "recip_kernels.h"

namespace admp {

// Piece k (= support offset) of the order-6 cardinal B-spline in the local
// coordinate t = w in [0,1): value, d/dR (= -d/du) and d2/dR2.
template <class R>
__device__ __forceinline__ void bs_piece(int k, R t, R& v, R& d1, R& d2) {
    const R c = (R)(1.0 / 120.0);
    R p, dp, ddp;
    switch (k) {
        case 0: {
            R t2 = t * t, t3 = t2 * t;
            p = t3 * t2 * c; dp = t2 * t2 * (R)(1.0 / 24.0); ddp = t3 * (R)(1.0 / 6.0);
        } break;
        case 1:
            p = ((((-(R)5 * t + (R)5) * t + (R)10) * t + (R)10) * t + (R)5) * t + (R)1;
            dp = (((-(R)25 * t + (R)20) * t + (R)30) * t + (R)20) * t + (R)5;
            ddp = ((-(R)100 * t + (R)60) * t + (R)60) * t + (R)20;
            p *= c; dp *= c; ddp *= c;
            break;
        case 2:
            p = (((((R)10 * t - (R)20) * t - (R)20) * t + (R)20) * t + (R)50) * t + (R)26;
            dp = ((((R)50 * t - (R)80) * t - (R)60) * t + (R)40) * t + (R)50;
            ddp = (((R)200 * t - (R)240) * t - (R)120) * t + (R)40;
            p *= c; dp *= c; ddp *= c;
            break;
        case 3:
            p = ((((-(R)10 * t + (R)30) * t) * t - (R)60) * t) * t + (R)66;
            dp = (((-(R)50 * t + (R)120) * t) * t - (R)120) * t;
            ddp = ((-(R)200 * t + (R)360) * t) * t - (R)120;
            p *= c; dp *= c; ddp *= c;
            break;
        case 4:
            p = (((((R)5 * t - (R)20) * t + (R)20) * t + (R)20) * t - (R)50) * t + (R)26;
            dp = ((((R)25 * t - (R)80) * t + (R)60) * t + (R)40) * t - (R)50;
            ddp = (((R)100 * t - (R)240) * t + (R)120) * t + (R)40;
            p *= c; dp *= c; ddp *= c;
            break;
        default: {
            R s = (R)1 - t, s2 = s * s;
            p = s2 * s2 * s * c; dp = -s2 * s2 * (R)(1.0 / 24.0); ddp = s2 * s * (R)(1.0 / 6.0);
        } break;
    }
    v = p;
    d1 = -dp;
    d2 = ddp;
}

// cell of each site = its base grid point (m mod K), for the counting sort
__global__ void k_grid_cell(int N, int K0, int K1, int K2, const int4* __restrict__ mbase, int* __restrict__ cell,
                            int* __restrict__ ccount) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t a = (size_t)f * N + i;
    const int4 m = mbase[a];
    const int id = (wrapi(m.x, K0) * K1 + wrapi(m.y, K1)) * K2 + wrapi(m.z, K2);
    cell[a] = id;
    atomicAdd(&ccount[(size_t)f * K0 * K1 * K2 + id], 1);
}


constexpr int BRX = 4, BRY = 4, BRZ = 4;   // 64 grid points per wavefront
constexpr int SP_CHUNK = 32;               // sites staged in LDS per round
// the solver's column spread of a single frame (dipoles only) stages 48 (two
// rounds for the ~72 candidates of a brick at liquid density): 1.5 % faster per
// evaluation at 12,288 sites, while 64-frame batches run 6 % slower with it and
// the quadrupole spreads (three times the tables per site) 4 us slower
constexpr int SP_CHUNK_ONE_FRAME = 48;

// ---------------------------------------------------------------------------
// Per-brick candidate lists.  The candidates of a brick are the sites whose
// base grid point lies in the brick grown by 2 below and 3 above in each axis
// (their 6-point support can touch the brick).  They are enumerated column by
// column ((x, y) pairs of the grown region, z ranges of the cell-sorted sites),
// so the order is fixed.  Positions do not change during a solve, so the lists
// are built once per evaluation and every spread of the solve reuses them.
// Entry: {site, (rx + 5) | (ry + 5) << 4 | (rz + 5) << 8} with r the support
// start relative to the brick origin.  Per axis at most ceil((B + 5) / B) + 1
// brick origins fit a window of B + 5 cells, so a site is a candidate of at
// most 4 x 4 x 4 bricks and 64 N entries bound a list.
// ---------------------------------------------------------------------------
constexpr int BRICK_MAXCAND = ((BRX + 5 + BRX - 1) / BRX + 1) * ((BRY + 5 + BRY - 1) / BRY + 1) *
                              ((BRZ + 5 + BRZ - 1) / BRZ + 1);
constexpr int BRICK_SEG = (BRX + 5) * (BRY + 5) * 2;   // column segments (two when z wraps)
constexpr int BRICK_SEG_LANE = (BRICK_SEG + 63) / 64;   // segments per lane of the wavefront

template <bool FILL>
__global__ void __launch_bounds__(64)
k_brick_cands(int N, int K0, int K1, int K2, const int4* __restrict__ mbase, const int* __restrict__ cstart,
              const int* __restrict__ catoms, int* __restrict__ bcount, const int* __restrict__ bstart,
              int2* __restrict__ blist, long cap) {
    constexpr int NC = BRICK_SEG;
    constexpr int SL = BRICK_SEG_LANE;
    __shared__ int col_lo[NC];
    __shared__ int col_off[NC + 1];
    const int f = blockIdx.y;
    const int nbz = (K2 + BRZ - 1) / BRZ, nby = (K1 + BRY - 1) / BRY, nbx = (K0 + BRX - 1) / BRX;
    const int nbr = nbx * nby * nbz;
    const int b = blockIdx.x;
    const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
    const int X0 = bx * BRX, Y0 = by * BRY, Z0 = bz * BRZ;
    const int t = threadIdx.x;
    const size_t G = (size_t)K0 * K1 * K2;
    const int* cs = cstart + (size_t)f * (G + 1);
    const int* ca = catoms + (size_t)f * N;
    const int nx = min(BRX + 5, K0), ny = min(BRY + 5, K1), nzr = min(BRZ + 5, K2);
    const int nseg = nx * ny * 2;
    const int zs = wrapi(Z0 - 2, K2);
    const int zlen1 = min(nzr, K2 - zs), zlen2 = nzr - zlen1;
    // each lane owns segments SL t .. SL t + SL - 1
    int lo2[SL], n2[SL];
#pragma unroll
    for (int k = 0; k < SL; ++k) {
        lo2[k] = 0;
        n2[k] = 0;
        const int c = SL * t + k;
        if (c >= nseg) continue;
        const int cc = c >> 1, seg = c & 1;
        const int x = wrapi(X0 - 2 + cc / ny, K0), y = wrapi(Y0 - 2 + cc % ny, K1);
        const size_t base = ((size_t)x * K1 + y) * K2;
        if (seg == 0) {
            lo2[k] = cs[base + zs];
            n2[k] = cs[base + zs + zlen1] - lo2[k];
        } else if (zlen2 > 0) {
            lo2[k] = cs[base];
            n2[k] = cs[base + zlen2] - lo2[k];
        }
    }
    int mine = 0;
#pragma unroll
    for (int k = 0; k < SL; ++k) mine += n2[k];
    const int incl = wave_incl_scan(mine);
    const int total = __shfl(incl, 63, 64);
    if (!FILL) {
        if (t == 0) bcount[(size_t)f * nbr + b] = total;
        return;
    }
    int ex = incl - mine;
#pragma unroll
    for (int k = 0; k < SL; ++k) {
        const int c = SL * t + k;
        if (c < NC) { col_lo[c] = lo2[k]; col_off[c] = ex; }
        ex += n2[k];
    }
    if (t == 0) col_off[NC] = total;
    __syncthreads();
    // each frame owns cap entries of the list buffer
    const long base = bstart[(size_t)f * (nbr + 1) + b];
    int2* fl = blist + (size_t)f * cap;
    for (int g = t; g < total; g += 64) {
        int lo = 0, hi = NC;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (col_off[mid] <= g) lo = mid; else hi = mid;
        }
        const int atom = ca[col_lo[lo] + (g - col_off[lo])];
        const int4 m = mbase[(size_t)f * N + atom];
        const int rx = wrapi(m.x - 3 - X0 + 5, K0);
        const int ry = wrapi(m.y - 3 - Y0 + 5, K1);
        const int rz = wrapi(m.z - 3 - Z0 + 5, K2);
        if (base + g < cap) fl[base + g] = make_int2(atom, rx | (ry << 4) | (rz << 8));
    }
}

// Spread from the per-brick lists: one wavefront per brick, one grid point per
// lane, candidates staged 64 at a time (spline tables and sources in LDS).
template <class R, int LM>
__global__ void __launch_bounds__(64)
k_spread_list(int N, int K0, int K1, int K2, const R* __restrict__ wfrac, const R* __restrict__ fq,
              const int* __restrict__ bstart, const int2* __restrict__ blist, long cap, R* __restrict__ rgrid, const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    constexpr int NQ = LM == 0 ? 1 : (LM == 1 ? 4 : 10);
    constexpr int NT = LM == 0 ? 1 : (LM == 1 ? 2 : 3);
    __shared__ int s_r[SP_CHUNK];   // packed support starts (+5), 4 bits per axis
    __shared__ R s_t[SP_CHUNK][3][NT][6];
    __shared__ R s_q[SP_CHUNK][NQ];
    const int f = blockIdx.y;
    const int nbz = (K2 + BRZ - 1) / BRZ, nby = (K1 + BRY - 1) / BRY, nbx = (K0 + BRX - 1) / BRX;
    const int nbr = nbx * nby * nbz;
    const int b = xcd_item(blockIdx.x, nbr);
    if (b >= nbr) return;
    const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
    const int t = threadIdx.x;
    const int lx = t / (BRY * BRZ), ly = (t / BRZ) % BRY, lz = t % BRZ;
    const int px = bx * BRX + lx, py = by * BRY + ly, pz = bz * BRZ + lz;
    const bool own = px < K0 && py < K1 && pz < K2;
    const size_t G = (size_t)K0 * K1 * K2;
    const int lo = bstart[(size_t)f * (nbr + 1) + b];
    const int total = (int)min((long)bstart[(size_t)f * (nbr + 1) + b + 1], cap) - lo;
    R acc = 0;
    for (int c0 = 0; c0 < total; c0 += SP_CHUNK) {
        const int nchunk = min(SP_CHUNK, total - c0);
        if (t < nchunk) {
            const int2 e = blist[(size_t)f * cap + lo + c0 + t];
            const size_t a = (size_t)f * N + e.x;
            s_r[t] = e.y;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const R w = wfrac[a * 3 + d];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    R v, d1, d2;
                    bs_piece<R>(k, w, v, d1, d2);
                    s_t[t][d][0][k] = v;
                    if (NT > 1) s_t[t][d][NT > 1 ? 1 : 0][k] = d1;
                    if (NT > 2) s_t[t][d][NT > 2 ? 2 : 0][k] = d2;
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) s_q[t][q] = fq[a * 10 + q];
        }
        __syncthreads();
        if (own) {
            for (int s2 = 0; s2 < nchunk; ++s2) {
                const int e = s_r[s2];
                const int k0 = lx + 5 - (e & 15), k1 = ly + 5 - ((e >> 4) & 15), k2 = lz + 5 - ((e >> 8) & 15);
                if ((unsigned)k0 > 5u || (unsigned)k1 > 5u || (unsigned)k2 > 5u) continue;
                const R v0 = s_t[s2][0][0][k0], v1 = s_t[s2][1][0][k1], v2 = s_t[s2][2][0][k2];
                R wgt = s_q[s2][0] * v0 * v1 * v2;
                if (LM >= 1) {
                    const R a0 = s_t[s2][0][NT > 1 ? 1 : 0][k0], a1 = s_t[s2][1][NT > 1 ? 1 : 0][k1],
                            a2 = s_t[s2][2][NT > 1 ? 1 : 0][k2];
                    wgt += s_q[s2][NQ > 1 ? 1 : 0] * a0 * v1 * v2 + s_q[s2][NQ > 2 ? 2 : 0] * v0 * a1 * v2 +
                           s_q[s2][NQ > 3 ? 3 : 0] * v0 * v1 * a2;
                    if (LM >= 2) {
                        const R b0 = s_t[s2][0][NT > 2 ? 2 : 0][k0], b1 = s_t[s2][1][NT > 2 ? 2 : 0][k1],
                                b2 = s_t[s2][2][NT > 2 ? 2 : 0][k2];
                        wgt += s_q[s2][NQ > 4 ? 4 : 0] * b0 * v1 * v2 + s_q[s2][NQ > 5 ? 5 : 0] * v0 * b1 * v2 +
                               s_q[s2][NQ > 6 ? 6 : 0] * v0 * v1 * b2 +
                               (R)2 * (s_q[s2][NQ > 7 ? 7 : 0] * a0 * a1 * v2 + s_q[s2][NQ > 8 ? 8 : 0] * a0 * v1 * a2 +
                                       s_q[s2][NQ > 9 ? 9 : 0] * v0 * a1 * a2);
                    }
                }
                acc += wgt;
            }
        }
        __syncthreads();
    }
    if (own) rgrid[(size_t)f * G + ((size_t)px * K1 + py) * K2 + pz] = acc;
}

// ---------------------------------------------------------------------------
// Column spread: four bricks per wavefront (16 lanes each); a lane owns one
// (x, y) column of the brick's 4 z points.  Per candidate the lane forms the
// x-y factors once, as coefficients of the z value, first and second
// derivative tables, and updates its 4 points from them:
//   w(x,y,z) = C0 vz + C1 az + C2 bz
//   C0 = q vx vy + mx ax vy + my vx ay + Txx bx vy + Tyy vx by + 2 Txy ax ay
//   C1 = mz vx vy + 2 Txz ax vy + 2 Tyz vx ay,   C2 = Tzz vx vy
// (v: value, a: first, b: second derivative tables per axis).  Candidates of
// every brick are staged SPC at a time (one per lane); the order of the sum
// over candidates is fixed, so results are bitwise reproducible.
// ---------------------------------------------------------------------------
constexpr int SPC = 16;   // candidates per brick per staging round

template <class R, int LM>
__global__ void __launch_bounds__(64)
k_spread_col(int N, int K0, int K1, int K2, const R* __restrict__ wfrac, const R* __restrict__ fq,
             const int* __restrict__ bstart, const int2* __restrict__ blist, long cap, R* __restrict__ rgrid,
             const PcgState* __restrict__ skip) {
    static_assert(BRX == 4 && BRY == 4 && BRZ == 4, "column spread assumes 4x4x4 bricks");
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    constexpr int NQ = LM == 0 ? 1 : (LM == 1 ? 4 : 10);
    constexpr int NT = LM == 0 ? 1 : (LM == 1 ? 2 : 3);
    // FP64 tables stage half as many candidates per round: LDS per wavefront sets the occupancy
    constexpr int SP = sizeof(R) == 8 ? SPC / 2 : SPC;
    __shared__ int s_r[4][SP];
    __shared__ R s_t[4][SP][3][NT][6];
    __shared__ R s_q[4][SP][NQ];
    const int f = blockIdx.y;
    const int nbz = (K2 + BRZ - 1) / BRZ, nby = (K1 + BRY - 1) / BRY, nbx = (K0 + BRX - 1) / BRX;
    const int nbr = nbx * nby * nbz;
    const int nq4 = (nbr + 3) / 4;
    const int qb = xcd_item(blockIdx.x, nq4);
    if (qb >= nq4) return;
    const int t = threadIdx.x;
    const int g = t >> 4, c = t & 15;
    const int b = 4 * qb + g;                      // this lane's brick
    const bool live = b < nbr;
    const int bb = live ? b : 0;
    const int bz = bb % nbz, by = (bb / nbz) % nby, bx = bb / (nbz * nby);
    const int lx = c >> 2, ly = c & 3;
    const int px = bx * BRX + lx, py = by * BRY + ly, pz0 = bz * BRZ;
    const int lo = live ? bstart[(size_t)f * (nbr + 1) + b] : 0;
    const int total = live ? (int)min((long)bstart[(size_t)f * (nbr + 1) + b + 1], cap) - lo : 0;
    // rounds needed by the busiest brick of the wavefront
    int rounds = (total + SP - 1) / SP;
    for (int o = 16; o < 64; o <<= 1) rounds = max(rounds, __shfl_xor(rounds, o, 64));
    R acc[4] = {0, 0, 0, 0};
    for (int rd = 0; rd < rounds; ++rd) {
        const int c0 = rd * SP;
        const int nch = max(0, min(SP, total - c0));
        if (c < nch) {
            const int2 e = blist[(size_t)f * cap + lo + c0 + c];
            const size_t a = (size_t)f * N + e.x;
            s_r[g][c] = e.y;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const R w = wfrac[a * 3 + d];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    R v, d1, d2;
                    bs_piece<R>(k, w, v, d1, d2);
                    s_t[g][c][d][0][k] = v;
                    if (NT > 1) s_t[g][c][d][NT > 1 ? 1 : 0][k] = d1;
                    if (NT > 2) s_t[g][c][d][NT > 2 ? 2 : 0][k] = d2;
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) s_q[g][c][q] = fq[a * 10 + q];
        }
        __syncthreads();
        for (int s2 = 0; s2 < nch; ++s2) {
            const int e = s_r[g][s2];
            const int k0 = lx + 5 - (e & 15), k1 = ly + 5 - ((e >> 4) & 15);
            if ((unsigned)k0 > 5u || (unsigned)k1 > 5u) continue;
            const int kz = 5 - ((e >> 8) & 15);   // z table index of the brick's first point
            const R vx = s_t[g][s2][0][0][k0], vy = s_t[g][s2][1][0][k1];
            R C0 = s_q[g][s2][0] * vx * vy, C1 = 0, C2 = 0;
            if (LM >= 1) {
                const R ax = s_t[g][s2][0][NT > 1 ? 1 : 0][k0], ay = s_t[g][s2][1][NT > 1 ? 1 : 0][k1];
                C0 += s_q[g][s2][NQ > 1 ? 1 : 0] * ax * vy + s_q[g][s2][NQ > 2 ? 2 : 0] * vx * ay;
                C1 = s_q[g][s2][NQ > 3 ? 3 : 0] * vx * vy;
                if (LM >= 2) {
                    const R bxv = s_t[g][s2][0][NT > 2 ? 2 : 0][k0], byv = s_t[g][s2][1][NT > 2 ? 2 : 0][k1];
                    C0 += s_q[g][s2][NQ > 4 ? 4 : 0] * bxv * vy + s_q[g][s2][NQ > 5 ? 5 : 0] * vx * byv +
                          (R)2 * s_q[g][s2][NQ > 7 ? 7 : 0] * ax * ay;
                    C1 += (R)2 * (s_q[g][s2][NQ > 8 ? 8 : 0] * ax * vy + s_q[g][s2][NQ > 9 ? 9 : 0] * vx * ay);
                    C2 = s_q[g][s2][NQ > 6 ? 6 : 0] * vx * vy;
                }
            }
#pragma unroll
            for (int lz = 0; lz < 4; ++lz) {
                const int k2 = kz + lz;
                if ((unsigned)k2 > 5u) continue;
                R w = C0 * s_t[g][s2][2][0][k2];
                if (LM >= 1) w += C1 * s_t[g][s2][2][NT > 1 ? 1 : 0][k2];
                if (LM >= 2) w += C2 * s_t[g][s2][2][NT > 2 ? 2 : 0][k2];
                acc[lz] += w;
            }
        }
        __syncthreads();
    }
    if (!live || px >= K0 || py >= K1) return;
    R* col = rgrid + (size_t)f * K0 * K1 * K2 + ((size_t)px * K1 + py) * K2;
#pragma unroll
    for (int lz = 0; lz < 4; ++lz)
        if (pz0 + lz < K2) col[pz0 + lz] = acc[lz];
}

// ---------------------------------------------------------------------------
// Column spread with one brick per wavefront: lane = (quarter g, column c); the
// 16 columns of the brick each hold 4 z points, and the four lanes of a column
// take every fourth candidate of each staged round (g, g+4, ...), forming the
// x-y factors of a candidate once per column as in k_spread_col.  The four
// quarter sums of a point are combined by a fixed shuffle tree, so the result is
// bitwise reproducible.  Per candidate a lane reads 4 x-y and up to 12 z table
// values for 4 points, where the one-point-per-lane spread reads 9 for one.
// ---------------------------------------------------------------------------
// CH: sites staged in LDS per round (a multiple of four, so each candidate keeps
// its quarter of the wavefront and the sums their order)
template <class R, int LM, int CH = SP_CHUNK>
__global__ void __launch_bounds__(64)
k_spread_cw(int N, int K0, int K1, int K2, const R* __restrict__ wfrac, const R* __restrict__ fq,
            const int* __restrict__ bstart, const int2* __restrict__ blist, long cap, R* __restrict__ rgrid,
            const PcgState* __restrict__ skip, CgCheck chk) {
    static_assert(CH % 4 == 0 && CH <= 64, "one staged site per lane, whole quarters");
    static_assert(BRX == 4 && BRY == 4 && BRZ == 4, "column spread assumes 4x4x4 bricks");
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    constexpr int NQ = LM == 0 ? 1 : (LM == 1 ? 4 : 10);
    constexpr int NT = LM == 0 ? 1 : (LM == 1 ? 2 : 3);
    const int f = blockIdx.y;
    const int nbz = (K2 + BRZ - 1) / BRZ, nby = (K1 + BRY - 1) / BRY, nbx = (K0 + BRX - 1) / BRX;
    const int nbr = nbx * nby * nbz;
    const int b = xcd_item(blockIdx.x, nbr);
    if (b >= nbr) return;
    const int t = threadIdx.x;
    const int lo = bstart[(size_t)f * (nbr + 1) + b];
    const int total = (int)min((long)bstart[(size_t)f * (nbr + 1) + b + 1], cap) - lo;
    // A lane's candidate for the next staging round is fetched into registers one
    // round ahead (the first one before the convergence test), so its global
    // loads are in flight while the wavefront works on the current round.
    int2 ne = make_int2(0, 0);
    R nw[3] = {0, 0, 0}, nq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) nq[q] = 0;
    auto fetch = [&](int c0) {
        if (t < min(CH, total - c0)) {
            ne = blist[(size_t)f * cap + lo + c0 + t];
            const size_t a = (size_t)f * N + ne.x;
#pragma unroll
            for (int d = 0; d < 3; ++d) nw[d] = wfrac[a * 3 + d];
#pragma unroll
            for (int q = 0; q < NQ; ++q) nq[q] = fq[a * 10 + q];
        }
    };
    if (total > 0) fetch(0);
    if (chk.on) {
        // every block reduces the same partials in the same order: one decision
        const double* mxp = chk.parts.mx(f, chk.cur);
        double m = 0.0;
        for (int k = threadIdx.x; k < chk.parts.nb; k += 64) m = fmax(m, mxp[k]);
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if (m < chk.tol) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                chk.st[f].res = m;
                chk.st[f].done = 1;
            }
            return;
        }
    }
    __shared__ int s_r[CH];
    __shared__ R s_t[CH][3][NT][6];
    __shared__ R s_q[CH][NQ];
    const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
    const int g = t >> 4, c = t & 15;
    const int lx = c >> 2, ly = c & 3;
    const int px = bx * BRX + lx, py = by * BRY + ly, pz0 = bz * BRZ;
    R acc[4] = {0, 0, 0, 0};
    for (int c0 = 0; c0 < total; c0 += CH) {
        const int nchunk = min(CH, total - c0);
        if (t < nchunk) {
            s_r[t] = ne.y;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const R w = nw[d];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    R v, d1, d2;
                    bs_piece<R>(k, w, v, d1, d2);
                    s_t[t][d][0][k] = v;
                    if (NT > 1) s_t[t][d][NT > 1 ? 1 : 0][k] = d1;
                    if (NT > 2) s_t[t][d][NT > 2 ? 2 : 0][k] = d2;
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) s_q[t][q] = nq[q];
        }
        __syncthreads();
        if (c0 + CH < total) fetch(c0 + CH);
        // the next candidate's packed offsets are read one step ahead, and the z
        // table is read at a clamped index for all four points at once (a point
        // outside the support takes a zero weight), so a candidate waits for LDS
        // twice instead of six times
        int en = g < nchunk ? s_r[g] : 0;
        for (int s2 = g; s2 < nchunk; s2 += 4) {
            const int e = en;
            if (s2 + 4 < nchunk) en = s_r[s2 + 4];
            const int k0 = lx + 5 - (e & 15), k1 = ly + 5 - ((e >> 4) & 15);
            if ((unsigned)k0 > 5u || (unsigned)k1 > 5u) continue;
            const int kz = 5 - ((e >> 8) & 15);   // z table index of the brick's first point
            R tz[4][NT];
#pragma unroll
            for (int lz = 0; lz < 4; ++lz) {
                const int kc = min(max(kz + lz, 0), 5);
#pragma unroll
                for (int j = 0; j < NT; ++j) tz[lz][j] = s_t[s2][2][j][kc];
            }
            const R vx = s_t[s2][0][0][k0], vy = s_t[s2][1][0][k1];
            R C0 = s_q[s2][0] * vx * vy, C1 = 0, C2 = 0;
            if (LM >= 1) {
                const R ax = s_t[s2][0][NT > 1 ? 1 : 0][k0], ay = s_t[s2][1][NT > 1 ? 1 : 0][k1];
                C0 += s_q[s2][NQ > 1 ? 1 : 0] * ax * vy + s_q[s2][NQ > 2 ? 2 : 0] * vx * ay;
                C1 = s_q[s2][NQ > 3 ? 3 : 0] * vx * vy;
                if (LM >= 2) {
                    const R bxv = s_t[s2][0][NT > 2 ? 2 : 0][k0], byv = s_t[s2][1][NT > 2 ? 2 : 0][k1];
                    C0 += s_q[s2][NQ > 4 ? 4 : 0] * bxv * vy + s_q[s2][NQ > 5 ? 5 : 0] * vx * byv +
                          (R)2 * s_q[s2][NQ > 7 ? 7 : 0] * ax * ay;
                    C1 += (R)2 * (s_q[s2][NQ > 8 ? 8 : 0] * ax * vy + s_q[s2][NQ > 9 ? 9 : 0] * vx * ay);
                    C2 = s_q[s2][NQ > 6 ? 6 : 0] * vx * vy;
                }
            }
#pragma unroll
            for (int lz = 0; lz < 4; ++lz) {
                R w = C0 * tz[lz][0];
                if (LM >= 1) w += C1 * tz[lz][NT > 1 ? 1 : 0];
                if (LM >= 2) w += C2 * tz[lz][NT > 2 ? 2 : 0];
                if ((unsigned)(kz + lz) <= 5u) acc[lz] += w;
            }
        }
        __syncthreads();
    }
    // quarter sums of each point: (g0 + g1) + (g2 + g3) in every lane
#pragma unroll
    for (int lz = 0; lz < 4; ++lz) {
        acc[lz] += __shfl_xor(acc[lz], 16, 64);
        acc[lz] += __shfl_xor(acc[lz], 32, 64);
    }
    // lane group g writes z point g of its column
    R mine = acc[0];
#pragma unroll
    for (int lz = 1; lz < 4; ++lz)
        if (g == lz) mine = acc[lz];
    if (px >= K0 || py >= K1 || pz0 + g >= K2) return;
    rgrid[(size_t)f * K0 * K1 * K2 + ((size_t)px * K1 + py) * K2 + pz0 + g] = mine;
}

// ---------------------------------------------------------------------------
// Line FFT engine.  L lines (compile time) live in LDS as [element][line]; the
// block's threads split into L lanes (one per line) times T = blockDim / L
// butterfly slots.  One Stockham autosort stage per factor r of K: butterfly b
// (0 <= b < K/r) reads x_q = in[b + q K/r], multiplies x_q by the stage twiddle
// W_{Ns r}^{q (b mod Ns)} and writes the r-point DFT to
// out[(b div Ns) Ns r + (b mod Ns) + k Ns].  Radices 2, 3, 4, 5 are written
// out; other factors use a direct sum per output.  tw[t] = e^{-2 pi i t/K}
// (conjugated for the inverse) is staged in LDS by the caller.
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ Cplx<R> cadd(Cplx<R> a, Cplx<R> b) { return Cplx<R>{a.x + b.x, a.y + b.y}; }
template <class R>
__device__ __forceinline__ Cplx<R> csub(Cplx<R> a, Cplx<R> b) { return Cplx<R>{a.x - b.x, a.y - b.y}; }
// multiply by -i (sg = +1) or +i (sg = -1)
template <class R>
__device__ __forceinline__ Cplx<R> mul_mi(Cplx<R> a, R sg) { return Cplx<R>{sg * a.y, -sg * a.x}; }

// One radix-r butterfly of a Stockham stage over a line whose element e sits at
// x[e * es] (in and out point at element 0 of the line).  b: butterfly index
// (0 <= b < nb = K / r), Ns: product of the radices of the earlier stages.
// floor(a / d) for 0 <= a < 2^22 and inv = 1.0f / d: the float product is within
// one of the quotient and the remainder corrects it (cheaper than an integer
// division on CDNA, which has none in hardware)
__device__ __forceinline__ int qdiv(int a, int d, float inv) {
    int q = (int)((float)a * inv);
    const int rem = a - q * d;
    q += (rem >= d) - (rem < 0);
    return q;
}

template <class R>
__device__ __forceinline__ void fft_bfly(const Cplx<R>* in, Cplx<R>* out, int es, int K, int r, int nb, int Ns,
                                         float invNs, int b, const Cplx<R>* tw, R sg) {
    const int step = K / (Ns * r);
    const int jq = qdiv(b, Ns, invNs), kj = b - jq * Ns;
    const int t1 = kj * step;  // twiddle index of q = 1 (< K)
    const int ob = jq * Ns * r + kj;
    auto twq = [&](int q) -> Cplx<R> {
        int t = q * t1;
        while (t >= K) t -= K;
        Cplx<R> w = tw[t];
        w.y *= sg;
        return w;
    };
    auto I = [&](int e) -> Cplx<R> { return in[e * es]; };
    auto O = [&](int e, Cplx<R> v) { out[e * es] = v; };
    if (r == 2) {
        const Cplx<R> x0 = I(b);
        const Cplx<R> x1 = cmul(I(b + nb), twq(1));
        O(ob, cadd(x0, x1));
        O(ob + Ns, csub(x0, x1));
    } else if (r == 4) {
        const Cplx<R> x0 = I(b);
        const Cplx<R> x1 = cmul(I(b + nb), twq(1));
        const Cplx<R> x2 = cmul(I(b + 2 * nb), twq(2));
        const Cplx<R> x3 = cmul(I(b + 3 * nb), twq(3));
        const Cplx<R> a = cadd(x0, x2), bb = csub(x0, x2), c = cadd(x1, x3), d = mul_mi(csub(x1, x3), sg);
        O(ob, cadd(a, c));
        O(ob + Ns, cadd(bb, d));
        O(ob + 2 * Ns, csub(a, c));
        O(ob + 3 * Ns, csub(bb, d));
    } else if (r == 3) {
        const R h = (R)0.86602540378443864676;  // sin(2 pi / 3)
        const Cplx<R> x0 = I(b);
        const Cplx<R> x1 = cmul(I(b + nb), twq(1));
        const Cplx<R> x2 = cmul(I(b + 2 * nb), twq(2));
        const Cplx<R> t1c = cadd(x1, x2);
        const Cplx<R> t2 = Cplx<R>{x0.x - (R)0.5 * t1c.x, x0.y - (R)0.5 * t1c.y};
        const Cplx<R> d = csub(x1, x2);
        const Cplx<R> t3 = mul_mi(Cplx<R>{h * d.x, h * d.y}, sg);
        O(ob, cadd(x0, t1c));
        O(ob + Ns, cadd(t2, t3));
        O(ob + 2 * Ns, csub(t2, t3));
    } else if (r == 5) {
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        const Cplx<R> x0 = I(b);
        const Cplx<R> x1 = cmul(I(b + nb), twq(1));
        const Cplx<R> x2 = cmul(I(b + 2 * nb), twq(2));
        const Cplx<R> x3 = cmul(I(b + 3 * nb), twq(3));
        const Cplx<R> x4 = cmul(I(b + 4 * nb), twq(4));
        const Cplx<R> t1c = cadd(x1, x4), t2 = cadd(x2, x3), t3 = csub(x1, x4), t4 = csub(x2, x3);
        const Cplx<R> a1 = Cplx<R>{x0.x + c1 * t1c.x + c2 * t2.x, x0.y + c1 * t1c.y + c2 * t2.y};
        const Cplx<R> a2 = Cplx<R>{x0.x + c2 * t1c.x + c1 * t2.x, x0.y + c2 * t1c.y + c1 * t2.y};
        const Cplx<R> b1 = mul_mi(Cplx<R>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, sg);
        const Cplx<R> b2 = mul_mi(Cplx<R>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, sg);
        O(ob, Cplx<R>{x0.x + t1c.x + t2.x, x0.y + t1c.y + t2.y});
        O(ob + Ns, cadd(a1, b1));
        O(ob + 4 * Ns, csub(a1, b1));
        O(ob + 2 * Ns, cadd(a2, b2));
        O(ob + 3 * Ns, csub(a2, b2));
    } else if (r == 10) {
        // 10 points as two 5-point transforms of the even and the odd inputs
        // (Cooley-Tukey with q = 2a + b): E_k, O_k for k < 5, then
        // y_k = E_k + W10^k O_k and y_{k+5} = E_k - W10^k O_k, W10 = tw[K / 10]
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        Cplx<R> x[10];
        x[0] = I(b);
#pragma unroll
        for (int q = 1; q < 10; ++q) x[q] = cmul(I(b + q * nb), twq(q));
        auto dft5 = [&](const Cplx<R>& a0, const Cplx<R>& a1, const Cplx<R>& a2, const Cplx<R>& a3,
                        const Cplx<R>& a4, Cplx<R>* y) {
            const Cplx<R> t1c = cadd(a1, a4), t2 = cadd(a2, a3), t3 = csub(a1, a4), t4 = csub(a2, a3);
            const Cplx<R> u1 = Cplx<R>{a0.x + c1 * t1c.x + c2 * t2.x, a0.y + c1 * t1c.y + c2 * t2.y};
            const Cplx<R> u2 = Cplx<R>{a0.x + c2 * t1c.x + c1 * t2.x, a0.y + c2 * t1c.y + c1 * t2.y};
            const Cplx<R> v1 = mul_mi(Cplx<R>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, sg);
            const Cplx<R> v2 = mul_mi(Cplx<R>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, sg);
            y[0] = Cplx<R>{a0.x + t1c.x + t2.x, a0.y + t1c.y + t2.y};
            y[1] = cadd(u1, v1);
            y[4] = csub(u1, v1);
            y[2] = cadd(u2, v2);
            y[3] = csub(u2, v2);
        };
        Cplx<R> E[5], Od[5];
        dft5(x[0], x[2], x[4], x[6], x[8], E);
        dft5(x[1], x[3], x[5], x[7], x[9], Od);
        const int k10 = K / 10;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            Cplx<R> t = Od[k];
            if (k > 0) {
                Cplx<R> w = tw[k * k10];
                w.y *= sg;
                t = cmul(t, w);
            }
            O(ob + k * Ns, cadd(E[k], t));
            O(ob + (k + 5) * Ns, csub(E[k], t));
        }
    } else {
        // direct r-point sum per output: y_k = sum_q x_q W_{Ns r}^{q kj} W_r^{q k}
        const int kr = K / r;  // W_r = tw[K / r]
        for (int k = 0; k < r; ++k) {
            R sx = 0, sy = 0;
            for (int q = 0; q < r; ++q) {
                int t = q * t1 + ((q * k) % r) * kr;
                while (t >= K) t -= K;
                Cplx<R> w = tw[t];
                w.y *= sg;
                const Cplx<R> x = I(b + q * nb);
                sx += x.x * w.x - x.y * w.y;
                sy += x.x * w.y + x.y * w.x;
            }
            O(ob + k * Ns, Cplx<R>{sx, sy});
        }
    }
}

// L lines of length K stored [element][line] (element e of line li at e * L + li);
// blockDim.x / L butterfly slots per line.  Ping-pongs between in and out and
// returns the buffer that holds the result.
template <class R, int L>
__device__ Cplx<R>* fft_lines(Cplx<R>* in, Cplx<R>* out, int K, const FFTPlan& plan, const Cplx<R>* tw, int inverse) {
    const int T = (int)blockDim.x / L;
    const int li = (int)threadIdx.x % L, slot = (int)threadIdx.x / L;
    const R sg = inverse ? (R)-1 : (R)1;  // sign of the exponent's -i
    int Ns = 1;
    for (int s = 0; s < plan.nst; ++s) {
        const int r = plan.fac[s];
        const int nb = K / r;
        const float invNs = 1.0f / (float)Ns;
        for (int b = slot; b < nb; b += T) fft_bfly<R>(in + li, out + li, L, K, r, nb, Ns, invNs, b, tw, sg);
        __syncthreads();
        Cplx<R>* t = in; in = out; out = t;
        Ns *= r;
    }
    return in;
}

// L lines of length K with element e of line li at li * ls + e * es, the
// (line, butterfly) work items of each stage spread over the whole block: line
// index fastest across threads when ls == 1, butterfly index fastest otherwise,
// so that neighbouring threads touch neighbouring LDS words.
template <class R>
__device__ Cplx<R>* fft_strided(Cplx<R>* in, Cplx<R>* out, int K, int L, int es, int ls, const FFTPlan& plan,
                                const Cplx<R>* tw, int inverse) {
    const R sg = inverse ? (R)-1 : (R)1;
    int Ns = 1;
    for (int s = 0; s < plan.nst; ++s) {
        const int r = plan.fac[s];
        const int nb = K / r;
        const int items = L * nb;
        const float invNs = 1.0f / (float)Ns;
        const int dv = ls == 1 ? L : nb;
        const float invdv = 1.0f / (float)dv;
        for (int t = threadIdx.x; t < items; t += blockDim.x) {
            const int qt = qdiv(t, dv, invdv), rt = t - qt * dv;
            const int li = ls == 1 ? rt : qt, b = ls == 1 ? qt : rt;
            fft_bfly<R>(in + li * ls, out + li * ls, es, K, r, nb, Ns, invNs, b, tw, sg);
        }
        __syncthreads();
        Cplx<R>* t = in; in = out; out = t;
        Ns *= r;
    }
    return in;
}

// block shape of the y and x line passes
constexpr int FL = 4;      // lines per block
constexpr int FTH = 64;    // threads per block (16 butterfly slots per line)

template <class R>
__device__ inline void stage_tw(Cplx<R>* dst, const Cplx<R>* __restrict__ src, int K) {
    for (int k = threadIdx.x; k < K; k += blockDim.x) dst[k] = src[k];
}

// Half-spectrum layouts.  Line-major: (K0, K1, H), kz fastest, for the separate
// y and x line passes.  Plane-major: (H, K0, K1), one contiguous (x, y) plane per
// kz, for the fused plane pass.  H = K2 / 2 + 1.
__device__ __forceinline__ size_t hs_index(int pm, int row, int kz, int H, int K0K1) {
    return pm ? (size_t)kz * K0K1 + row : (size_t)row * H + kz;
}

// real rows (K0 K1 of length K2) -> half spectrum; FLT complex lines = 2 FLT rows
// per block of 16 FLT threads
template <class R, int FLT>
__global__ void __launch_bounds__(16 * FLT) k_fz_r2c(int K0, int K1, int K2, FFTPlan plan,
                                                     const Cplx<R>* __restrict__ twg, const R* __restrict__ rgrid,
                                                     Cplx<R>* __restrict__ hgrid, int pm,
                                                     const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    extern __shared__ unsigned char smem_raw[];
    Cplx<R>* b0 = reinterpret_cast<Cplx<R>*>(smem_raw);
    Cplx<R>* b1 = b0 + FLT * K2;
    Cplx<R>* tw = b1 + FLT * K2;
    const int f = blockIdx.y, H = K2 / 2 + 1;
    const int nrl = K0 * K1;
    const int r0 = 2 * blockIdx.x * FLT;  // first real row of the block
    stage_tw(tw, twg, K2);
    const R* rg = rgrid + (size_t)f * nrl * K2 + (size_t)r0 * K2;
    Cplx<R>* hg = hgrid + (size_t)f * nrl * H;
    // 2 FLT consecutive rows are contiguous: row rr -> line rr / 2, real (even) or imaginary (odd) part
    R* b0r = reinterpret_cast<R*>(b0);
    for (int idx = threadIdx.x; idx < 2 * FLT * K2; idx += blockDim.x) {
        const int rr = idx / K2, e = idx - rr * K2;
        b0r[(e * FLT + (rr >> 1)) * 2 + (rr & 1)] = r0 + rr < nrl ? rg[idx] : (R)0;
    }
    __syncthreads();
    Cplx<R>* res = fft_lines<R, FLT>(b0, b1, K2, plan, tw, 0);
    // unpack the two real transforms of each complex line: row (even) a = (Z + conj Z~) / 2, (odd) b = (Z - conj Z~) / 2i
    for (int idx = threadIdx.x; idx < 2 * FLT * H; idx += blockDim.x) {
        int rr, k;
        if (pm) { k = idx / (2 * FLT); rr = idx - k * 2 * FLT; }
        else { rr = idx / H; k = idx - rr * H; }
        const int row = r0 + rr;
        if (row >= nrl) continue;
        const int li = rr >> 1;
        const Cplx<R> Z = res[k * FLT + li];
        const Cplx<R> Zc = res[((K2 - k) % K2) * FLT + li];
        const Cplx<R> v = (rr & 1) ? Cplx<R>{(Z.y + Zc.y) * (R)0.5, (Zc.x - Z.x) * (R)0.5}
                                   : Cplx<R>{(Z.x + Zc.x) * (R)0.5, (Z.y - Zc.y) * (R)0.5};
        hg[hs_index(pm, row, k, H, nrl)] = v;
    }
}

// half spectrum -> real rows (inverse, unnormalised)
template <class R, int FLT>
__global__ void __launch_bounds__(16 * FLT) k_fz_c2r(int K0, int K1, int K2, FFTPlan plan,
                                                     const Cplx<R>* __restrict__ twg, const Cplx<R>* __restrict__ hgrid,
                                                     R* __restrict__ rgrid, int pm, const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    extern __shared__ unsigned char smem_raw[];
    Cplx<R>* b0 = reinterpret_cast<Cplx<R>*>(smem_raw);
    Cplx<R>* b1 = b0 + FLT * K2;
    Cplx<R>* tw = b1 + FLT * K2;
    const int f = blockIdx.y, H = K2 / 2 + 1;
    const int nrl = K0 * K1;
    const int r0 = 2 * blockIdx.x * FLT;
    stage_tw(tw, twg, K2);
    R* rg = rgrid + (size_t)f * nrl * K2 + (size_t)r0 * K2;
    const Cplx<R>* hg = hgrid + (size_t)f * nrl * H;
    for (int idx = threadIdx.x; idx < FLT * K2; idx += blockDim.x) {
        int li, k;
        if (pm) { k = idx / FLT; li = idx - k * FLT; }
        else { li = idx / K2; k = idx - li * K2; }
        const int la = r0 + 2 * li, lb = la + 1;
        const int kk = k < H ? k : K2 - k;
        Cplx<R> A = {0, 0}, B = {0, 0};
        if (la < nrl) A = hg[hs_index(pm, la, kk, H, nrl)];
        if (lb < nrl) B = hg[hs_index(pm, lb, kk, H, nrl)];
        if (k >= H) { A.y = -A.y; B.y = -B.y; }
        b0[k * FLT + li] = Cplx<R>{A.x - B.y, A.y + B.x};  // Z = A + i B
    }
    __syncthreads();
    const Cplx<R>* res = fft_lines<R, FLT>(b0, b1, K2, plan, tw, 1);
    const R* resr = reinterpret_cast<const R*>(res);
    for (int idx = threadIdx.x; idx < 2 * FLT * K2; idx += blockDim.x) {
        const int rr = idx / K2, e = idx - rr * K2;
        if (r0 + rr < nrl) rg[idx] = resr[(e * FLT + (rr >> 1)) * 2 + (rr & 1)];
    }
}

// Influence function at half-spectrum point (i0, i1, i2) for `kind` (0:
// electrostatic, recip.py:474-475; 6/8/10: dispersion, pme.py:1063-1082):
// C_k and dC/d(k^2) with the reciprocal vector k; C = 0 at k = 0 for kind 0.
struct InflPt {
    double C, dC, kx, ky, kz;
};
__device__ inline InflPt infl_point(const BoxInfo& b, int K0, int K1, int K2, int i0, int i1, int i2, double kappa,
                                    int kind) {
    InflPt o;
    const int m0 = i0 < (K0 + 1) / 2 ? i0 : i0 - K0;
    const int m1 = i1 < (K1 + 1) / 2 ? i1 : i1 - K1;
    const int m2 = i2 < (K2 + 1) / 2 ? i2 : i2 - K2;
    o.kx = b.kb[0] * m0 + b.kb[1] * m1 + b.kb[2] * m2;
    o.ky = b.kb[3] * m0 + b.kb[4] * m1 + b.kb[5] * m2;
    o.kz = b.kb[6] * m0 + b.kb[7] * m1 + b.kb[8] * m2;
    const double ksq = o.kx * o.kx + o.ky * o.ky + o.kz * o.kz;
    const double V = b.V;
    o.C = 0;
    o.dC = 0;
    if (kind == 0) {
        if (!(i0 == 0 && i1 == 0 && i2 == 0)) {
            o.C = 2.0 * PI / V / ksq * exp(-ksq / (4.0 * kappa * kappa));
            o.dC = -o.C * (1.0 / ksq + 1.0 / (4.0 * kappa * kappa));
        }
    } else {
        const double kq = ksq < 1e-16 ? 1e-16 : ksq;
        const double x2 = kq / (4.0 * kappa * kappa), x = sqrt(x2);
        const double ex = exp(-x2), ec = erfc(x);
        double fv, fd, pw;
        if (kind == 6) {
            fv = (1.0 - 2.0 * x2) * ex + 2.0 * x2 * x * SQRT_PI * ec;
            fd = -6.0 * x * ex + 6.0 * x2 * SQRT_PI * ec;
            pw = kappa * kappa * kappa / 3.0;
        } else if (kind == 8) {
            fv = (3.0 - 2.0 * x2 + 4.0 * x2 * x2) * ex - 4.0 * x2 * x2 * x * SQRT_PI * ec;
            fd = (-10.0 * x + 20.0 * x2 * x) * ex - 20.0 * x2 * x2 * SQRT_PI * ec;
            pw = kappa * kappa * kappa * kappa * kappa / 45.0;
        } else {
            const double x4 = x2 * x2;
            fv = (15.0 - 6.0 * x2 + 4.0 * x4 - 8.0 * x4 * x2) * ex + 8.0 * x4 * x2 * x * SQRT_PI * ec;
            fd = (-42.0 * x + 28.0 * x2 * x - 56.0 * x4 * x) * ex + 56.0 * x4 * x2 * SQRT_PI * ec;
            pw = pow(kappa, 7.0) / 1260.0;
        }
        const double c0 = SQRT_PI * PI / 2.0 / V * pw;
        o.C = c0 * fv;
        o.dC = ksq < 1e-16 ? 0.0 : c0 * fd / (8.0 * kappa * kappa * x);
    }
    return o;
}

// Fused plane pass on the plane-major half spectrum: one block per (kz, frame)
// holds the K0 x K1 plane in LDS and does forward y and x transforms, the
// influence function, inverse x and y transforms, then writes the plane back:
// one read and one write of the half spectrum instead of three of each.
// ACC: energy and box-derivative partials per block (multiplicity 2 for
// 0 < kz < K2/2, the conjugate half).  tab (solver products): tabulated
// 2 pref C_k / theta_k^2 in the same plane-major order.
// threads of a solver plane-pass workgroup (one per kz plane); the variant that
// also sums energy and box derivatives keeps 512 (it spills at 1024)
constexpr int PLANE_THREADS = 1024;

// plane i2 of frame f (one workgroup); nplane: planes per frame (energy partials)
template <class R, bool ACC>
__device__ __forceinline__ void fxy_plane_body(int f, int i2, int nplane, int K0, int K1, int K2, FFTPlan plan0,
                                               FFTPlan plan1, const Cplx<R>* __restrict__ tw0g,
                                               const Cplx<R>* __restrict__ tw1g, Cplx<R>* __restrict__ hgrid,
                                               const BoxInfo* __restrict__ boxes, const double* __restrict__ bmod0,
                                               const double* __restrict__ bmod1, const double* __restrict__ bmod2,
                                               double kappa, int kind, double pref, double* __restrict__ part,
                                               const double* __restrict__ tab) {
    extern __shared__ unsigned char smem_raw[];
    __shared__ double sh[16];
    const int H = K2 / 2 + 1, P = K0 * K1;
    Cplx<R>* b0 = reinterpret_cast<Cplx<R>*>(smem_raw);
    Cplx<R>* b1 = b0 + P;
    Cplx<R>* tw0 = b1 + P;
    Cplx<R>* tw1 = tw0 + K0;
    Cplx<R>* pl = hgrid + ((size_t)f * H + i2) * P;
    for (int i = threadIdx.x; i < P; i += blockDim.x) b0[i] = pl[i];
    stage_tw(tw0, tw0g, K0);
    stage_tw(tw1, tw1g, K1);
    __syncthreads();
    // y lines: line x at x * K1, element stride 1; x lines: line y at y, element stride K1
    Cplx<R>* res = fft_strided<R>(b0, b1, K1, K0, 1, K1, plan1, tw1, 0);
    res = fft_strided<R>(res, res == b0 ? b1 : b0, K0, K1, K1, 1, plan0, tw0, 0);
    const BoxInfo& b = boxes[f];
    const double* tp = tab ? tab + ((size_t)f * H + i2) * P : nullptr;
    const double mult = (i2 == 0 || 2 * i2 == K2) ? 1.0 : 2.0;
    double e = 0, w00 = 0, w11 = 0, w22 = 0, w01 = 0, w02 = 0, w12 = 0;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const Cplx<R> S = res[i];
        double scale;
        if (!ACC && tp) {
            scale = tp[i];
        } else {
            const int i0 = i / K1, i1 = i - i0 * K1;
            const InflPt q = infl_point(b, K0, K1, K2, i0, i1, i2, kappa, kind);
            const double th = bmod0[i0] * bmod1[i1] * bmod2[i2];
            const double ith2 = 1.0 / (th * th);
            if (ACC) {
                const double s2 = (double)S.x * (double)S.x + (double)S.y * (double)S.y;
                const double wgt = pref * s2 * ith2;
                e += mult * wgt * q.C;
                const double wd = -mult * wgt * q.dC;
                w00 += wd * q.kx * q.kx; w11 += wd * q.ky * q.ky; w22 += wd * q.kz * q.kz;
                w01 += wd * q.kx * q.ky; w02 += wd * q.kx * q.kz; w12 += wd * q.ky * q.kz;
            }
            scale = 2.0 * pref * q.C * ith2;
        }
        res[i] = Cplx<R>{(R)(S.x * scale), (R)(S.y * scale)};
    }
    __syncthreads();
    res = fft_strided<R>(res, res == b0 ? b1 : b0, K0, K1, K1, 1, plan0, tw0, 1);
    res = fft_strided<R>(res, res == b0 ? b1 : b0, K1, K0, 1, K1, plan1, tw1, 1);
    for (int i = threadIdx.x; i < P; i += blockDim.x) pl[i] = res[i];
    if (!ACC) return;
    double vals[7] = {e, w00, w11, w22, w01, w02, w12};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        double tsum = block_sum(vals[k], sh);
        if (threadIdx.x == 0) part[((size_t)f * nplane + i2) * 7 + k] = tsum;
    }
}

template <class R, bool ACC>
__global__ void __launch_bounds__(ACC ? 512 : PLANE_THREADS) k_fxy_plane(int K0, int K1, int K2, FFTPlan plan0, FFTPlan plan1,
                                                    const Cplx<R>* __restrict__ tw0g, const Cplx<R>* __restrict__ tw1g,
                                                    Cplx<R>* __restrict__ hgrid, const BoxInfo* __restrict__ boxes,
                                                    const double* __restrict__ bmod0, const double* __restrict__ bmod1,
                                                    const double* __restrict__ bmod2, double kappa, int kind,
                                                    double pref, double* __restrict__ part,
                                                    const double* __restrict__ tab, const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    fxy_plane_body<R, ACC>(blockIdx.y, blockIdx.x, gridDim.x, K0, K1, K2, plan0, plan1, tw0g, tw1g, hgrid, boxes,
                           bmod0, bmod1, bmod2, kappa, kind, pref, part, tab);
}

// complex lines along y of the half spectrum: lines (x, kz), stride H
template <class R>
__global__ void __launch_bounds__(FTH) k_fy(int K0, int K1, int H, FFTPlan plan, const Cplx<R>* __restrict__ twg,
                                           Cplx<R>* __restrict__ hgrid, int inverse, const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    extern __shared__ unsigned char smem_raw[];
    Cplx<R>* b0 = reinterpret_cast<Cplx<R>*>(smem_raw);
    Cplx<R>* b1 = b0 + FL * K1;
    Cplx<R>* tw = b1 + FL * K1;
    const int f = blockIdx.y;
    const int nlines = K0 * H;
    const int l0 = blockIdx.x * FL;
    stage_tw(tw, twg, K1);
    Cplx<R>* hg = hgrid + (size_t)f * K0 * K1 * H;
    // thread -> (element e, line li) with li fastest: consecutive kz in memory
    for (int idx = threadIdx.x; idx < FL * K1; idx += FTH) {
        const int e = idx / FL, li = idx % FL;
        const int l = l0 + li;
        Cplx<R> v = {0, 0};
        if (l < nlines) v = hg[(size_t)(l / H) * K1 * H + (l % H) + (size_t)e * H];
        b0[e * FL + li] = v;
    }
    __syncthreads();
    Cplx<R>* res = fft_lines<R, FL>(b0, b1, K1, plan, tw, inverse);
    for (int idx = threadIdx.x; idx < FL * K1; idx += FTH) {
        const int e = idx / FL, li = idx % FL;
        const int l = l0 + li;
        if (l < nlines) hg[(size_t)(l / H) * K1 * H + (l % H) + (size_t)e * H] = res[e * FL + li];
    }
}

// forward x FFT, influence function, inverse x FFT; lines (y, kz), stride K1*H.
// ACC: accumulate energy and box-derivative partials per block (multiplicity 2
// for 0 < kz < K2/2, the conjugate half).
template <class R, bool ACC>
__global__ void __launch_bounds__(FTH) k_fx_infl(int K0, int K1, int K2, FFTPlan plan, const Cplx<R>* __restrict__ twg,
                                                Cplx<R>* __restrict__ hgrid, const BoxInfo* __restrict__ boxes,
                                                const double* __restrict__ bmod0, const double* __restrict__ bmod1,
                                                const double* __restrict__ bmod2, double kappa, int kind, double pref,
                                                double* __restrict__ part, const double* __restrict__ tab,
                                                const PcgState* __restrict__ skip) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    extern __shared__ unsigned char smem_raw[];
    __shared__ double sh[16];
    Cplx<R>* b0 = reinterpret_cast<Cplx<R>*>(smem_raw);
    Cplx<R>* b1 = b0 + FL * K0;
    Cplx<R>* tw = b1 + FL * K0;
    const int f = blockIdx.y;
    const int H = K2 / 2 + 1;
    const int nlines = K1 * H;
    const int l0 = blockIdx.x * FL;
    stage_tw(tw, twg, K0);
    Cplx<R>* hg = hgrid + (size_t)f * K0 * K1 * H;
    const size_t stride = (size_t)K1 * H;
    for (int idx = threadIdx.x; idx < FL * K0; idx += FTH) {
        const int e = idx / FL, li = idx % FL;
        const int l = l0 + li;
        b0[e * FL + li] = l < nlines ? hg[(size_t)l + (size_t)e * stride] : Cplx<R>{0, 0};
    }
    __syncthreads();
    Cplx<R>* res = fft_lines<R, FL>(b0, b1, K0, plan, tw, 0);
    const BoxInfo& b = boxes[f];
    double e = 0, w00 = 0, w11 = 0, w22 = 0, w01 = 0, w02 = 0, w12 = 0;
    for (int idx = threadIdx.x; idx < FL * K0; idx += FTH) {
        const int i0 = idx / FL, li = idx % FL;
        const int l = l0 + li;
        if (l >= nlines) continue;
        const int i1 = l / H, i2 = l % H;
        const Cplx<R> S = res[i0 * FL + li];
        if (!ACC && tab) {
            // tabulated 2 pref C_k / theta_k^2 (electrostatic solver products)
            const double scale = tab[(size_t)f * K0 * nlines + (size_t)i0 * nlines + l];
            res[i0 * FL + li] = Cplx<R>{(R)(S.x * scale), (R)(S.y * scale)};
            continue;
        }
        const InflPt q = infl_point(b, K0, K1, K2, i0, i1, i2, kappa, kind);
        const double th = bmod0[i0] * bmod1[i1] * bmod2[i2];
        const double ith2 = 1.0 / (th * th);
        if (ACC) {
            const double s2 = (double)S.x * (double)S.x + (double)S.y * (double)S.y;
            const double mult = (i2 == 0 || 2 * i2 == K2) ? 1.0 : 2.0;
            const double wgt = pref * s2 * ith2;
            e += mult * wgt * q.C;
            const double wd = -mult * wgt * q.dC;
            w00 += wd * q.kx * q.kx; w11 += wd * q.ky * q.ky; w22 += wd * q.kz * q.kz;
            w01 += wd * q.kx * q.ky; w02 += wd * q.kx * q.kz; w12 += wd * q.ky * q.kz;
        }
        const double scale = 2.0 * pref * q.C * ith2;
        res[i0 * FL + li] = Cplx<R>{(R)(S.x * scale), (R)(S.y * scale)};
    }
    __syncthreads();
    Cplx<R>* other = (res == b0) ? b1 : b0;
    Cplx<R>* res2 = fft_lines<R, FL>(res, other, K0, plan, tw, 1);
    for (int idx = threadIdx.x; idx < FL * K0; idx += FTH) {
        const int el = idx / FL, li = idx % FL;
        const int l = l0 + li;
        if (l < nlines) hg[(size_t)l + (size_t)el * stride] = res2[el * FL + li];
    }
    if (!ACC) return;
    double vals[7] = {e, w00, w11, w22, w01, w02, w12};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        double tsum = block_sum(vals[k], sh);
        if (threadIdx.x == 0) part[((size_t)f * gridDim.x + blockIdx.x) * 7 + k] = tsum;
    }
}

// 2 pref C_k / theta_k^2 for the electrostatic influence function (kind 0,
// recip.py:474-475; k = 0 excluded) on the half spectrum, laid out as the pass
// that reads it: plane-major [frame][kz][i0][i1] for the plane pass (pm), else
// [frame][i0][line], line = ky * H + kz, for the x line pass.  Depends on the
// box only.
__global__ void k_infl_table(int K0, int K1, int K2, const BoxInfo* __restrict__ boxes,
                             const double* __restrict__ bmod0, const double* __restrict__ bmod1,
                             const double* __restrict__ bmod2, double kappa, double pref, int pm,
                             double* __restrict__ tab) {
    const int f = blockIdx.y;
    const int H = K2 / 2 + 1;
    const size_t P = (size_t)K0 * K1, nt = P * H;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nt) return;
    int i0, i1, i2;
    if (pm) {
        i2 = (int)(idx / P);
        const int r = (int)(idx - (size_t)i2 * P);
        i0 = r / K1; i1 = r - i0 * K1;
    } else {
        const int nlines = K1 * H;
        i0 = (int)(idx / nlines);
        const int l = (int)(idx - (size_t)i0 * nlines);
        i1 = l / H; i2 = l - i1 * H;
    }
    const InflPt q = infl_point(boxes[f], K0, K1, K2, i0, i1, i2, kappa, 0);
    const double th = bmod0[i0] * bmod1[i1] * bmod2[i2];
    tab[(size_t)f * nt + idx] = 2.0 * pref * q.C / (th * th);
}

// ---------------------------------------------------------------------------
// Gather with eight lanes per site (lanes 0..5 own one x-offset each).
// MODE 1: dE/dmu only (solver products) into out3.  MODE 0: everything the
// backward needs: dE/dQ (fractional potential and its first/second derivatives
// mapped to Cartesian), dE/dpos and the per-atom virial.  Loop bounds are
// compile-time constants so every array stays in registers.
// ---------------------------------------------------------------------------
// MODE 2: the CG product epilogue (solver products): w = real + dE/dmu + d z
// into out3 with z = U (d: the diagonal of the polarisation matrix, real: the
// real-space product), and one z.w partial per block.
struct CgEpi {
    const double* real;
    const double* pol;
    double f1;
    CgParts parts;
};

// eight lanes per site; 16 sites per block measured within ±1 % of 32, 8 slower
constexpr int GATHER_SITES_PER_BLOCK = 32;

template <class R, int MODE>
__global__ void __launch_bounds__(8 * GATHER_SITES_PER_BLOCK)
k_gather_ml(int N, int K0, int K1, int K2, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
            int use_dA, const int4* __restrict__ mbase, const R* __restrict__ wfrac, const R* __restrict__ fq,
            const double* __restrict__ Qc, const double* __restrict__ U, int qmode, const R* __restrict__ rgrid,
            double* __restrict__ out3, double* __restrict__ g_q, double* __restrict__ g_pos,
            double* __restrict__ vir_atom, const int* __restrict__ order, const PcgState* __restrict__ skip,
            CgEpi epi) {
    if (skip && skip[blockIdx.y].done) return;  // frame already converged
    constexpr int ND = MODE == 0 ? 4 : 2;  // derivative orders needed
    const int f = blockIdx.y;
    // GATHER_SITES_PER_BLOCK sites per block, taken in cell-sorted order (order[])
    // and XCD-aware
    constexpr int GS = GATHER_SITES_PER_BLOCK;
    const int nsb = (N + GS - 1) / GS;
    const int sb = xcd_item(blockIdx.x, nsb);
    const int k = sb * GS + (int)(threadIdx.x >> 3), lane = threadIdx.x & 7;
    const bool active = sb < nsb && k < N;
    const int i = active ? (order ? order[(size_t)f * N + k] : k) : 0;
    const int ii = i;
    const size_t a = (size_t)f * N + ii;
    const BoxInfo& b = boxes[f];
    const double* A = use_dA ? b.dA : b.A;
    const int4 m = mbase[a];
    const R* gbase = rgrid + (size_t)f * K0 * K1 * K2;
    // the CG epilogue's inputs do not depend on the grid: fetched before it
    double e_pol = 0.0, e_real[3] = {0, 0, 0}, e_z[3] = {0, 0, 0};
    if constexpr (MODE == 2) {
        if (active && lane == 0) {
            e_pol = epi.pol[ii];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                e_real[c] = epi.real[a * 3 + c];
                e_z[c] = U[a * 3 + c];
            }
        }
    }
    // fp[c][aa][bb]: d^(c,aa,bb) of the fractional potential, c + aa + bb < ND
    double fp[ND][ND][ND];
#pragma unroll
    for (int c = 0; c < ND; ++c)
#pragma unroll
        for (int aa = 0; aa < ND; ++aa)
#pragma unroll
            for (int bb = 0; bb < ND; ++bb) fp[c][aa][bb] = 0.0;
    if (active && lane < 6) {
        R M1[4][6], M2[4][6], M0[4][6];
        bspline6<R>(wfrac[a * 3 + 1], M1[0], M1[1], M1[2], M1[3]);
        bspline6<R>(wfrac[a * 3 + 2], M2[0], M2[1], M2[2], M2[3]);
        bspline6<R>(wfrac[a * 3 + 0], M0[0], M0[1], M0[2], M0[3]);
        const int k0 = lane;
        const int g0 = wrapi(m.x + k0 - 3, K0);
        int gz[6];
#pragma unroll
        for (int k2 = 0; k2 < 6; ++k2) gz[k2] = wrapi(m.z + k2 - 3, K2);
        R u[ND][ND];
#pragma unroll
        for (int aa = 0; aa < ND; ++aa)
#pragma unroll
            for (int bb = 0; bb < ND; ++bb) u[aa][bb] = 0;
#pragma unroll
        for (int k1 = 0; k1 < 6; ++k1) {
            const int g1 = wrapi(m.y + k1 - 3, K1);
            const R* row = gbase + ((size_t)g0 * K1 + g1) * K2;
            R ph[6];
#pragma unroll
            for (int k2 = 0; k2 < 6; ++k2) ph[k2] = row[gz[k2]];
            R sz[ND];
#pragma unroll
            for (int o = 0; o < ND; ++o) {
                sz[o] = 0;
#pragma unroll
                for (int k2 = 0; k2 < 6; ++k2) sz[o] += ph[k2] * M2[o][k2];
            }
#pragma unroll
            for (int aa = 0; aa < ND; ++aa)
#pragma unroll
                for (int bb = 0; bb < ND; ++bb)
                    if (aa + bb < ND) u[aa][bb] += M1[aa][k1] * sz[bb];
        }
        // select this lane's x row (static indexing: unrolled over k0 candidates)
        R mx[ND];
#pragma unroll
        for (int c = 0; c < ND; ++c) {
            R v = M0[c][0];
#pragma unroll
            for (int t = 1; t < 6; ++t) v = (k0 == t) ? M0[c][t] : v;
            mx[c] = v;
        }
#pragma unroll
        for (int c = 0; c < ND; ++c)
#pragma unroll
            for (int aa = 0; aa < ND; ++aa)
#pragma unroll
                for (int bb = 0; bb < ND; ++bb)
                    if (c + aa + bb < ND) fp[c][aa][bb] = (double)(mx[c] * u[aa][bb]);
    }
#pragma unroll
    for (int c = 0; c < ND; ++c)
#pragma unroll
        for (int aa = 0; aa < ND; ++aa)
#pragma unroll
            for (int bb = 0; bb < ND; ++bb)
                if (c + aa + bb < ND) fp[c][aa][bb] = group_sum(fp[c][aa][bb], 8);
    if constexpr (MODE == 2) {
        __shared__ double shz[GS / 8];
        double zw = 0;
        if (active && lane == 0) {
            const double q1[3] = {fp[1][0][0], fp[0][1][0], fp[0][0][1]};
            const double ap = e_pol < 1e-8 ? 1e-8 : e_pol;
            const double d = DIELECTRIC / ap - 2.0 * DIELECTRIC * epi.f1;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double dm = A[c * 3 + 0] * q1[0] + A[c * 3 + 1] * q1[1] + A[c * 3 + 2] * q1[2];
                const double zv = e_z[c];
                const double wv = e_real[c] + dm + d * zv;
                out3[a * 3 + c] = wv;
                zw += zv * wv;
            }
        }
        zw = group_sum(zw, 64);
        if ((threadIdx.x & 63) == 0) shz[threadIdx.x >> 6] = zw;
        __syncthreads();
        if (threadIdx.x == 0 && sb < nsb) {
            double t = shz[0];
#pragma unroll
            for (int w = 1; w < GS / 8; ++w) t += shz[w];
            epi.parts.zw(f)[sb] = t;
        }
        return;
    }
    if (!active || lane != 0) return;
    const double p1[3] = {fp[1][0][0], fp[0][1][0], fp[0][0][1]};
    double dmu[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dmu[c] = A[c * 3 + 0] * p1[0] + A[c * 3 + 1] * p1[1] + A[c * 3 + 2] * p1[2];
    if constexpr (MODE == 1) {
        out3[a * 3 + 0] = dmu[0]; out3[a * 3 + 1] = dmu[1]; out3[a * 3 + 2] = dmu[2];
        return;
    } else {
        double p2[3][3];
        p2[0][0] = fp[2][0][0]; p2[1][1] = fp[0][2][0]; p2[2][2] = fp[0][0][2];
        p2[0][1] = p2[1][0] = fp[1][1][0]; p2[0][2] = p2[2][0] = fp[1][0][1]; p2[1][2] = p2[2][1] = fp[0][1][1];
        // third derivatives d3[j][k][l] (symmetric), from the ten independent entries
        double d3[3][3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    const int ox = (j == 0) + (k == 0) + (l == 0);
                    const int oy = (j == 1) + (k == 1) + (l == 1);
                    const int oz = (j == 2) + (k == 2) + (l == 2);
                    double v = 0;
                    if (ox == 3) v = fp[3][0][0];
                    else if (oy == 3) v = fp[0][3][0];
                    else if (oz == 3) v = fp[0][0][3];
                    else if (ox == 2 && oy == 1) v = fp[2][1][0];
                    else if (ox == 2 && oz == 1) v = fp[2][0][1];
                    else if (oy == 2 && ox == 1) v = fp[1][2][0];
                    else if (oy == 2 && oz == 1) v = fp[0][2][1];
                    else if (oz == 2 && ox == 1) v = fp[1][0][2];
                    else if (oz == 2 && oy == 1) v = fp[0][1][2];
                    else v = fp[1][1][1];
                    d3[j][k][l] = v;
                }
        const R* c = fq + a * 10;
        const double q = (double)c[0];
        const double fm[3] = {(double)c[1], (double)c[2], (double)c[3]};
        double fT[3][3];
        fT[0][0] = c[4]; fT[1][1] = c[5]; fT[2][2] = c[6];
        fT[0][1] = fT[1][0] = c[7]; fT[0][2] = fT[2][0] = c[8]; fT[1][2] = fT[2][1] = c[9];
        double g[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double sj = q * p1[j];
#pragma unroll
            for (int k = 0; k < 3; ++k) sj += fm[k] * p2[k][j];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) sj += fT[k][l] * d3[k][l][j];
            g[j] = sj;
        }
        double Ap2[3][3];
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                Ap2[x][k] = A[x * 3 + 0] * p2[0][k] + A[x * 3 + 1] * p2[1][k] + A[x * 3 + 2] * p2[2][k];
        double GT[3][3];
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
            for (int y = 0; y < 3; ++y)
                GT[x][y] = (Ap2[x][0] * A[y * 3 + 0] + Ap2[x][1] * A[y * 3 + 1] + Ap2[x][2] * A[y * 3 + 2]) * (1.0 / 3.0);
        double* gq = g_q + a * 10;
        gq[0] = fp[0][0][0];
        gq[1] = dmu[0]; gq[2] = dmu[1]; gq[3] = dmu[2];
        gq[4] = GT[0][0]; gq[5] = GT[1][1]; gq[6] = GT[2][2]; gq[7] = GT[0][1]; gq[8] = GT[0][2]; gq[9] = GT[1][2];
        const double4 pp = pos[a];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
            g_pos[a * 3 + cc] = A[0 * 3 + cc] * g[0] + A[1 * 3 + cc] * g[1] + A[2 * 3 + cc] * g[2];
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
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                double TAp = 0;
#pragma unroll
                for (int x = 0; x < 3; ++x) TAp += Th[j][x] * Ap2[x][cc];
                v[j * 3 + cc] = g[j] * r[cc] + mu[j] * p1[cc] + (2.0 / 3.0) * TAp;
            }
    }
}

// ---------------------------------------------------------------------------
// The plane pass of a solver product and the real-space product of the same z
// in one launch.  The plane pass runs one workgroup per kz plane (26 for a 50^3
// grid), far too few to fill the GPU, and the real-space product needs only z:
// its workgroups fill the idle CUs, and no second stream (fork and join) is
// needed.  Blocks [0, nplane_blocks) are planes (plane-major over frames;
// blocks past H F return), the rest are real-space items of 512 threads (32
// sites per 512 threads) in XCD order.  Every block is given the plane pass's dynamic LDS.
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(PLANE_THREADS)
k_plane_matvec(int K0, int K1, int K2, FFTPlan plan0, FFTPlan plan1, const Cplx<R>* __restrict__ tw0g,
               const Cplx<R>* __restrict__ tw1g, Cplx<R>* __restrict__ hgrid, const BoxInfo* __restrict__ boxes,
               const double* __restrict__ bmod0, const double* __restrict__ bmod1, const double* __restrict__ bmod2,
               double kappa, int kind, double pref, const double* __restrict__ tab, int F, int nplane_blocks,
               int N, int nbf, int cap, const double4* __restrict__ pos, const int* __restrict__ nbr,
               const int* __restrict__ ncnt, const double* __restrict__ pcoef, const double* __restrict__ v,
               double* __restrict__ out, const PcgState* __restrict__ skip) {
    const int H = K2 / 2 + 1;
    const int b = blockIdx.x;
    if (b < nplane_blocks) {
        if (b >= H * F) return;
        const int f = b / H, i2 = b - f * H;
        if (skip && skip[f].done) return;  // frame already converged
        fxy_plane_body<R, false>(f, i2, H, K0, K1, K2, plan0, plan1, tw0g, tw1g, hgrid, boxes, bmod0, bmod1, bmod2,
                                 kappa, kind, pref, nullptr, tab);
        return;
    }
    const int item = xcd_item(b - nplane_blocks, nbf * F);
    if (item >= nbf * F) return;
    real_matvec_body<double, double, PLANE_THREADS>(item, N, nbf, cap, pos, boxes, nbr, ncnt, pcoef, v, out, skip);
}

}  // namespace admp
