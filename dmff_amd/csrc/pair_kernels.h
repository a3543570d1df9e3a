// Real-space multipole / Thole pair kernels.
//
// The reference evaluates each pair in a quasi-internal (QI) frame: both sites'
// harmonic multipoles are rotated so z points along r_ij, and the energy is a
// sum of radial coefficients times products of QI components
// (dmff/admp/pme.py:632-1048).  Every QI product is a rotational invariant of
// the global Cartesian multipoles and z = r_ij/|r_ij|, e.g.
//   Q21_I . Q21_J = 4/3 ((Theta_I z).(Theta_J z) - zz_I zz_J),
//   Q22_I . Q22_J = 2/3 Theta_I:Theta_J + 1/3 zz_I zz_J - 4/3 (Theta_I z).(Theta_J z),
// so the kernel works directly with Cartesian invariants: no per-pair rotation
// matrices, and the backward (forces, torques, parameter gradients) is a short
// closed form.  The radial coefficients are the reference's, term for term.
//
// Thread mapping: TPA lanes cooperate on one atom i (full neighbour list, both
// directions), striding over its neighbours; per-atom sums are reduced across
// the TPA lanes with cross-lane shuffles.  No atomics: every atom's outputs are
// written once, so results are deterministic.
#pragma once
#include "admp_common.h"
#include "sr_kernels.h"

namespace admp {

constexpr int TPA = 16;           // lanes per atom
constexpr int PAIR_BLOCK = 256;   // 16 atoms per block
constexpr double PC_RPAIR = 2.0;  // A: block preconditioner neighbour range

// site record (R precision), 16 words per atom
// 0 q | 1-3 mu | 4-9 Theta xx yy zz xy xz yz | 10 alpha^(1/6) | 11 thole | 12 alpha | 13-15 U
constexpr int SITE = 16;

template <class R> __device__ __forceinline__ R r_exp(R x) { return exp(x); }
template <class R> __device__ __forceinline__ R r_erf(R x) { return erf(x); }

// Radial pieces at distance r (dmff/admp/pme.py:656-675): P_n = D/r^n, the bVec
// recursion b2..b4 and Y_k = (kappa r)^k X with X = 2 exp(-kappa^2 r^2)/sqrt(pi),
// plus their r-derivatives (db_n/dr = -2^(n-1)/(2n-3)!! kappa x^(2n-2) X).
template <class R> struct Radial {
    R P1, P2, P3, P4, P5, dP1, dP2, dP3, dP4, dP5;
    R b2, b3, b4, db2, db3, db4;
    R Y1, Y3, Y5, dY1, dY3, dY5;
    R x, kap;
};

template <class R> __device__ __forceinline__ void radial(R r, R kappa, Radial<R>& q) {
    const R D = (R)DIELECTRIC;
    R ri = (R)1 / r;
    q.P1 = D * ri; q.P2 = q.P1 * ri; q.P3 = q.P2 * ri; q.P4 = q.P3 * ri; q.P5 = q.P4 * ri;
    q.dP1 = -q.P1 * ri; q.dP2 = -(R)2 * q.P2 * ri; q.dP3 = -(R)3 * q.P3 * ri;
    q.dP4 = -(R)4 * q.P4 * ri; q.dP5 = -(R)5 * q.P5 * ri;
    R x = kappa * r, x2 = x * x;
    R X = (R)TWO_OVER_SQRT_PI * r_exp<R>(-x2);
    R x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x, x6 = x4 * x2;
    q.x = x; q.kap = kappa;
    q.Y1 = x * X; q.Y3 = x3 * X; q.Y5 = x5 * X;
    q.dY1 = kappa * ((R)1 - (R)2 * x2) * X;
    q.dY3 = kappa * ((R)3 * x2 - (R)2 * x4) * X;
    q.dY5 = kappa * ((R)5 * x4 - (R)2 * x6) * X;
    q.b2 = -r_erf<R>(x) + q.Y1;
    q.b3 = q.b2 + (R)(2.0 / 3.0) * q.Y3;
    q.b4 = q.b3 + (R)(4.0 / 15.0) * q.Y5;
    q.db2 = -(R)2 * kappa * x2 * X;
    q.db3 = -(R)(4.0 / 3.0) * kappa * x4 * X;
    q.db4 = -(R)(8.0 / 15.0) * kappa * x6 * X;
}

// The five shape functions shared by permanent and induced coefficients
// (calc_e_perm pme.py:677-700 / calc_e_ind pme.py:832-879, the latter halved):
//   cd(s)  = P2 (s + b2)
//   dd0(s) = -2/3 P3 (3 (s + b3) + Y3)
//   dd1(s) = P3 (s + b3 - 2/3 Y3)
//   dq0(s) = P4 (3 (s + b3) + 4/3 Y5)
//   dq1(s) = -sqrt3 P4 (s + b3)
// with s the scale (mscale, or pscale*thole_f(au), or dscale*thole_f(au)).
template <class R> __device__ __forceinline__ void f_cd(const Radial<R>& q, R s, R ds, R& v, R& d) {
    v = q.P2 * (s + q.b2);
    d = q.dP2 * (s + q.b2) + q.P2 * (ds + q.db2);
}
template <class R> __device__ __forceinline__ void f_dd0(const Radial<R>& q, R s, R ds, R& v, R& d) {
    R a = (R)3 * (s + q.b3) + q.Y3;
    v = -(R)(2.0 / 3.0) * q.P3 * a;
    d = -(R)(2.0 / 3.0) * (q.dP3 * a + q.P3 * ((R)3 * (ds + q.db3) + q.dY3));
}
template <class R> __device__ __forceinline__ void f_dd1(const Radial<R>& q, R s, R ds, R& v, R& d) {
    R a = s + q.b3 - (R)(2.0 / 3.0) * q.Y3;
    v = q.P3 * a;
    d = q.dP3 * a + q.P3 * (ds + q.db3 - (R)(2.0 / 3.0) * q.dY3);
}
template <class R> __device__ __forceinline__ void f_dq0(const Radial<R>& q, R s, R ds, R& v, R& d) {
    R a = (R)3 * (s + q.b3) + (R)(4.0 / 3.0) * q.Y5;
    v = q.P4 * a;
    d = q.dP4 * a + q.P4 * ((R)3 * (ds + q.db3) + (R)(4.0 / 3.0) * q.dY5);
}
template <class R> __device__ __forceinline__ void f_dq1(const Radial<R>& q, R s, R ds, R& v, R& d) {
    R a = s + q.b3;
    v = -(R)SQRT3 * q.P4 * a;
    d = -(R)SQRT3 * (q.dP4 * a + q.P4 * (ds + q.db3));
}

// Thole damping functions of au (pme.py:797-818) and their derivatives;
// e^{-au} = 0 for au >= 50, and the 1e8 power caps never bind below that.
template <class R> struct Thole { R c, d0, q0, q1, dc, dd0, dq0, dq1; };
template <class R> __device__ __forceinline__ void thole_f(R au, Thole<R>& t) {
    R e = au < (R)50 ? r_exp<R>(-au) : (R)0;
    R a2 = au * au, a3 = a2 * au, a4 = a3 * au;
    R base = (R)1 + au + (R)0.5 * a2;
    t.c = (R)1 - e * base;
    t.d0 = (R)1 - e * (base + a3 * (R)0.25);
    t.q1 = (R)1 - e * (base + a3 * (R)(1.0 / 6.0));
    t.q0 = (R)1 - e * (base + a3 * (R)(1.0 / 6.0) + a4 * (R)(1.0 / 18.0));
    t.dc = e * (R)0.5 * a2;
    t.dd0 = e * (a3 - a2) * (R)0.25;
    t.dq1 = e * a3 * (R)(1.0 / 6.0);
    t.dq0 = e * (a4 - a3) * (R)(1.0 / 18.0);
}

struct PairParams {
    double ms[8], ps[8], ds[8];
    int nscale;
    double kappa;
};

// ---------------------------------------------------------------------------
// Final pass: energy and every gradient, U held fixed.
// ---------------------------------------------------------------------------
template <class R, bool LPOL>
__global__ void __launch_bounds__(PAIR_BLOCK)
k_real_final(int N, int cap, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
             const R* __restrict__ site, const int* __restrict__ nbr, const int* __restrict__ ncnt,
             PairParams pp, double* __restrict__ e_atom, double* __restrict__ g_pos,
             double* __restrict__ g_q, double* __restrict__ vir_atom, double* __restrict__ g_pol,
             double* __restrict__ g_thole) {
    const int f = blockIdx.y;
    const int gt = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    const int i = gt / TPA, sub = gt % TPA;
    const bool active = i < N;
    const int ii = active ? i : 0;
    const BoxInfo& bx = boxes[f];
    const size_t fo = (size_t)f * N;
    const double4 pi_ = pos[fo + ii];
    const R* si = site + (fo + ii) * SITE;
    const R qI = si[0], mIx = si[1], mIy = si[2], mIz_ = si[3];
    const R TIxx = si[4], TIyy = si[5], TIzz = si[6], TIxy = si[7], TIxz = si[8], TIyz = si[9];
    const R a6I = si[10], thI = si[11], alI = si[12];
    R UIx = 0, UIy = 0, UIz = 0;
    if (LPOL) { UIx = si[13]; UIy = si[14]; UIz = si[15]; }
    const R kap = (R)pp.kappa;
    double E = 0, gx = 0, gy = 0, gz = 0;
    double gq = 0, gmx = 0, gmy = 0, gmz = 0;
    double gTxx = 0, gTyy = 0, gTzz = 0, gTxy = 0, gTxz = 0, gTyz = 0;
    double v00 = 0, v01 = 0, v02 = 0, v10 = 0, v11 = 0, v12 = 0, v20 = 0, v21 = 0, v22 = 0;
    double gpol = 0, gth = 0;
    const int cnt = active ? ncnt[fo + ii] : 0;
    const int* nl = nbr + (fo + ii) * (size_t)cap;
    for (int k = sub; k < cnt; k += TPA) {
        const int w = nl[k];
        const int j = w & NB_MASK, nb = w >> NB_SHIFT;
        const double4 pj = pos[fo + j];
        double dx = pi_.x - pj.x, dy = pi_.y - pj.y, dz = pi_.z - pj.z, n0, n1, n2;
        pbc_shift(bx, dx, dy, dz, n0, n1, n2);
        const double rd = sqrt(dx * dx + dy * dy + dz * dz);
        const R r = (R)rd, rinv = (R)(1.0 / rd);
        const R zx = (R)dx * rinv, zy = (R)dy * rinv, zz = (R)dz * rinv;
        const R* sj = site + (fo + j) * SITE;
        const R qJ = sj[0], mJx = sj[1], mJy = sj[2], mJz_ = sj[3];
        const R TJxx = sj[4], TJyy = sj[5], TJzz = sj[6], TJxy = sj[7], TJxz = sj[8], TJyz = sj[9];
        const int sidx = scale_index(nb, pp.nscale);
        const R ms = (R)pp.ms[sidx];
        // invariants
        const R mIz = mIx * zx + mIy * zy + mIz_ * zz;
        const R mJz = mJx * zx + mJy * zy + mJz_ * zz;
        const R tIx = TIxx * zx + TIxy * zy + TIxz * zz;
        const R tIy = TIxy * zx + TIyy * zy + TIyz * zz;
        const R tIz = TIxz * zx + TIyz * zy + TIzz * zz;
        const R tJx = TJxx * zx + TJxy * zy + TJxz * zz;
        const R tJy = TJxy * zx + TJyy * zy + TJyz * zz;
        const R tJz = TJxz * zx + TJyz * zy + TJzz * zz;
        const R zzI = tIx * zx + tIy * zy + tIz * zz;
        const R zzJ = tJx * zx + tJy * zy + tJz * zz;
        const R tImJ = tIx * mJx + tIy * mJy + tIz * mJz_;
        const R tJmI = tJx * mIx + tJy * mIy + tJz * mIz_;
        const R tItJ = tIx * tJx + tIy * tJy + tIz * tJz;
        const R mImJ = mIx * mJx + mIy * mJy + mIz_ * mJz_;
        const R ThTh = TIxx * TJxx + TIyy * TJyy + TIzz * TJzz +
                       (R)2 * (TIxy * TJxy + TIxz * TJxz + TIyz * TJyz);
        Radial<R> rq;
        radial<R>(r, kap, rq);
        // permanent coefficients (calc_e_perm)
        R cc, dcc, cd, dcd, dd0, ddd0, dd1, ddd1, cq, dcq, dq0, ddq0, dq1, ddq1;
        {
            R a = ms + rq.b2 - rq.Y1;
            cc = rq.P1 * a;
            dcc = rq.dP1 * a + rq.P1 * (rq.db2 - rq.dY1);
        }
        f_cd<R>(rq, ms, (R)0, cd, dcd);
        f_dd0<R>(rq, ms, (R)0, dd0, ddd0);
        f_dd1<R>(rq, ms, (R)0, dd1, ddd1);
        cq = rq.P3 * (ms + rq.b3);
        dcq = rq.dP3 * (ms + rq.b3) + rq.P3 * rq.db3;
        f_dq0<R>(rq, ms, (R)0, dq0, ddq0);
        f_dq1<R>(rq, ms, (R)0, dq1, ddq1);
        R qq0, dqq0, qq1, dqq1, qq2, dqq2;
        {
            const R x = rq.x, x2 = x * x;
            R a0 = (R)6 * (ms + rq.b4) + (R)(4.0 / 45.0) * ((R)-3 + (R)10 * x2) * rq.Y5;
            qq0 = rq.P5 * a0;
            dqq0 = rq.dP5 * a0 + rq.P5 * ((R)6 * rq.db4 + (R)(4.0 / 45.0) * ((R)20 * x * kap * rq.Y5 +
                                                                    ((R)-3 + (R)10 * x2) * rq.dY5));
            R a1 = (R)15 * (ms + rq.b4) + rq.Y5;
            qq1 = -(R)(4.0 / 15.0) * rq.P5 * a1;
            dqq1 = -(R)(4.0 / 15.0) * (rq.dP5 * a1 + rq.P5 * ((R)15 * rq.db4 + rq.dY5));
            R a2 = ms + rq.b4 - (R)(4.0 / 15.0) * rq.Y5;
            qq2 = rq.P5 * a2;
            dqq2 = rq.dP5 * a2 + rq.P5 * (rq.db4 - (R)(4.0 / 15.0) * rq.dY5);
        }
        const R k1 = (R)K1_INV;
        // energy, permanent part
        R e = cc * qI * qJ + cd * (qI * mJz - qJ * mIz) + dd0 * mIz * mJz + dd1 * (mImJ - mIz * mJz) +
              cq * (qI * zzJ + qJ * zzI) + dq0 * (zzI * mJz - zzJ * mIz) +
              dq1 * k1 * ((tImJ - zzI * mJz) - (tJmI - zzJ * mIz)) + qq0 * zzI * zzJ +
              qq1 * (R)(4.0 / 3.0) * (tItJ - zzI * zzJ) +
              qq2 * ((R)(2.0 / 3.0) * ThTh + (R)(1.0 / 3.0) * zzI * zzJ - (R)(4.0 / 3.0) * tItJ);
        R dEdr = dcc * qI * qJ + dcd * (qI * mJz - qJ * mIz) + ddd0 * mIz * mJz + ddd1 * (mImJ - mIz * mJz) +
                 dcq * (qI * zzJ + qJ * zzI) + ddq0 * (zzI * mJz - zzJ * mIz) +
                 ddq1 * k1 * ((tImJ - zzI * mJz) - (tJmI - zzJ * mIz)) + dqq0 * zzI * zzJ +
                 dqq1 * (R)(4.0 / 3.0) * (tItJ - zzI * zzJ) +
                 dqq2 * ((R)(2.0 / 3.0) * ThTh + (R)(1.0 / 3.0) * zzI * zzJ - (R)(4.0 / 3.0) * tItJ);
        // building-block partials
        R e_mIz = -cd * qJ + (dd0 - dd1) * mJz + (k1 * dq1 - dq0) * zzJ;
        R e_mJz = cd * qI + (dd0 - dd1) * mIz + (dq0 - k1 * dq1) * zzI;
        const R czz = qq0 - (R)(4.0 / 3.0) * qq1 + (R)(1.0 / 3.0) * qq2;
        R e_zzI = cq * qJ + (dq0 - k1 * dq1) * mJz + czz * zzJ;
        R e_zzJ = cq * qI + (k1 * dq1 - dq0) * mIz + czz * zzI;
        const R e_tImJ = k1 * dq1, e_tJmI = -k1 * dq1;
        const R e_tItJ = (R)(4.0 / 3.0) * (qq1 - qq2);
        const R e_ThTh = (R)(2.0 / 3.0) * qq2;
        R e_uIz = 0, e_uJz = 0;
        // multipole partials (atom I)
        R gqI = cc * qJ + cd * mJz + cq * zzJ;
        // vector accumulators for dE/dmu_I (beyond e_mIz z) and dE/dz
        R gmIx = dd1 * mJx + e_tJmI * tJx, gmIy = dd1 * mJy + e_tJmI * tJy, gmIz = dd1 * mJz_ + e_tJmI * tJz;
        // sym(z (x) v) coefficient vector for dE/dTheta_I
        R vx = e_tImJ * mJx + e_tItJ * tJx, vy = e_tImJ * mJy + e_tItJ * tJy, vz = e_tImJ * mJz_ + e_tItJ * tJz;
        // dE/dz partial sums
        R Zx = 0, Zy = 0, Zz = 0;
        // Theta_I mu_J and Theta_J mu_I, Theta_I t_J, Theta_J t_I
        {
            R TImJx = TIxx * mJx + TIxy * mJy + TIxz * mJz_;
            R TImJy = TIxy * mJx + TIyy * mJy + TIyz * mJz_;
            R TImJz = TIxz * mJx + TIyz * mJy + TIzz * mJz_;
            R TJmIx = TJxx * mIx + TJxy * mIy + TJxz * mIz_;
            R TJmIy = TJxy * mIx + TJyy * mIy + TJyz * mIz_;
            R TJmIz = TJxz * mIx + TJyz * mIy + TJzz * mIz_;
            R TItJx = TIxx * tJx + TIxy * tJy + TIxz * tJz;
            R TItJy = TIxy * tJx + TIyy * tJy + TIyz * tJz;
            R TItJz = TIxz * tJx + TIyz * tJy + TIzz * tJz;
            R TJtIx = TJxx * tIx + TJxy * tIy + TJxz * tIz;
            R TJtIy = TJxy * tIx + TJyy * tIy + TJyz * tIz;
            R TJtIz = TJxz * tIx + TJyz * tIy + TJzz * tIz;
            Zx += e_tImJ * TImJx + e_tJmI * TJmIx + e_tItJ * (TItJx + TJtIx);
            Zy += e_tImJ * TImJy + e_tJmI * TJmIy + e_tItJ * (TItJy + TJtIy);
            Zz += e_tImJ * TImJz + e_tJmI * TJmIz + e_tItJ * (TItJz + TJtIz);
        }
        R gthe = 0, gdmp = 0;
        if (LPOL) {
            const R UJx = sj[13], UJy = sj[14], UJz = sj[15];
            const R ps = (R)pp.ps[sidx], dsc = (R)pp.ds[sidx];
            const R uIz = UIx * zx + UIy * zy + UIz * zz;
            const R uJz = UJx * zx + UJy * zy + UJz * zz;
            const R tIUJ = tIx * UJx + tIy * UJy + tIz * UJz;
            const R tJUI = tJx * UIx + tJy * UIy + tJz * UIz;
            const R mJUI = mJx * UIx + mJy * UIy + mJz_ * UIz;
            const R mIUJ = mIx * UJx + mIy * UJy + mIz_ * UJz;
            const R UIUJ = UIx * UJx + UIy * UJy + UIz * UJz;
            // Thole width and argument (pme.py:793-801)
            const R a = ps > (R)0 ? (R)5.0 : thI + sj[11];
            R dmp = a6I * sj[10];
            const bool dmp_cl = dmp < (R)1e-8;
            if (dmp_cl) dmp = (R)1e-8;
            R u = r / dmp;
            const bool u_cl = !(u < (R)1e8);
            if (u_cl) u = (R)1e8;
            const R au = a * u;
            const R dau_dr = u_cl ? (R)0 : a / dmp;
            Thole<R> th;
            thole_f<R>(au, th);
            R cud, dcud, dud0, ddud0, dud1, ddud1, udq0, dudq0, udq1, dudq1, ud0, dud0d, ud1, dud1d;
            f_cd<R>(rq, ps * th.c, ps * th.dc * dau_dr, cud, dcud);
            f_dd0<R>(rq, ps * th.d0, ps * th.dd0 * dau_dr, dud0, ddud0);
            f_dd1<R>(rq, ps * th.c, ps * th.dc * dau_dr, dud1, ddud1);
            f_dq0<R>(rq, ps * th.q0, ps * th.dq0 * dau_dr, udq0, dudq0);
            f_dq1<R>(rq, ps * th.q1, ps * th.dq1 * dau_dr, udq1, dudq1);
            f_dd0<R>(rq, dsc * th.d0, dsc * th.dd0 * dau_dr, ud0, dud0d);
            f_dd1<R>(rq, dsc * th.c, dsc * th.dc * dau_dr, ud1, dud1d);
            const R I_cud = qI * uJz - qJ * uIz;
            const R I_dud0 = mJz * uIz + mIz * uJz;
            const R I_dud1 = (mJUI - mJz * uIz) + (mIUJ - mIz * uJz);
            const R I_udq0 = zzI * uJz - zzJ * uIz;
            const R I_udq1 = k1 * ((tIUJ - zzI * uJz) - (tJUI - zzJ * uIz));
            const R I_u0 = uIz * uJz;
            const R I_u1 = UIUJ - uIz * uJz;
            e += cud * I_cud + dud0 * I_dud0 + dud1 * I_dud1 + udq0 * I_udq0 + udq1 * I_udq1 + ud0 * I_u0 +
                 ud1 * I_u1;
            dEdr += dcud * I_cud + ddud0 * I_dud0 + ddud1 * I_dud1 + dudq0 * I_udq0 + dudq1 * I_udq1 +
                    dud0d * I_u0 + dud1d * I_u1;
            // dE/dau through the Thole functions (shape-function slopes in s)
            const R dEdau = rq.P2 * ps * I_cud * th.dc +
                            (-(R)2 * rq.P3) * (ps * I_dud0 + dsc * I_u0) * th.dd0 +
                            rq.P3 * (ps * I_dud1 + dsc * I_u1) * th.dc +
                            (R)3 * rq.P4 * ps * I_udq0 * th.dq0 - (R)SQRT3 * rq.P4 * ps * I_udq1 * th.dq1;
            gdmp = (dmp_cl || u_cl) ? (R)0 : dEdau * (-au / dmp);
            gthe = (ps > (R)0) ? (R)0 : dEdau * u;
            e_mIz += (dud0 - dud1) * uJz;
            e_mJz += (dud0 - dud1) * uIz;
            e_zzI += (udq0 - k1 * udq1) * uJz;
            e_zzJ += (k1 * udq1 - udq0) * uIz;
            e_uIz = -cud * qJ + (dud0 - dud1) * mJz + (k1 * udq1 - udq0) * zzJ + (ud0 - ud1) * uJz;
            e_uJz = cud * qI + (dud0 - dud1) * mIz + (udq0 - k1 * udq1) * zzI + (ud0 - ud1) * uIz;
            gqI += cud * uJz;
            gmIx += dud1 * UJx; gmIy += dud1 * UJy; gmIz += dud1 * UJz;
            const R e_tIUJ = k1 * udq1, e_tJUI = -k1 * udq1;
            vx += e_tIUJ * UJx; vy += e_tIUJ * UJy; vz += e_tIUJ * UJz;
            // dz parts: e_uIz UI + e_uJz UJ + e_tIUJ Theta_I UJ + e_tJUI Theta_J UI
            Zx += e_uIz * UIx + e_uJz * UJx;
            Zy += e_uIz * UIy + e_uJz * UJy;
            Zz += e_uIz * UIz + e_uJz * UJz;
            Zx += e_tIUJ * (TIxx * UJx + TIxy * UJy + TIxz * UJz) + e_tJUI * (TJxx * UIx + TJxy * UIy + TJxz * UIz);
            Zy += e_tIUJ * (TIxy * UJx + TIyy * UJy + TIyz * UJz) + e_tJUI * (TJxy * UIx + TJyy * UIy + TJyz * UIz);
            Zz += e_tIUJ * (TIxz * UJx + TIyz * UJy + TIzz * UJz) + e_tJUI * (TJxz * UIx + TJyz * UIy + TJzz * UIz);
        }
        // dE/dz: e_mIz mu_I + e_mJz mu_J + 2 e_zzI t_I + 2 e_zzJ t_J + (tensor terms above)
        Zx += e_mIz * mIx + e_mJz * mJx + (R)2 * (e_zzI * tIx + e_zzJ * tJx);
        Zy += e_mIz * mIy + e_mJz * mJy + (R)2 * (e_zzI * tIy + e_zzJ * tJy);
        Zz += e_mIz * mIz_ + e_mJz * mJz_ + (R)2 * (e_zzI * tIz + e_zzJ * tJz);
        const R zZ = Zx * zx + Zy * zy + Zz * zz;
        const R fx = dEdr * zx + (Zx - zZ * zx) * rinv;
        const R fy = dEdr * zy + (Zy - zZ * zy) * rinv;
        const R fz = dEdr * zz + (Zz - zZ * zz) * rinv;
        E += (double)e;
        gx += (double)fx; gy += (double)fy; gz += (double)fz;
        v00 -= n0 * (double)fx; v01 -= n0 * (double)fy; v02 -= n0 * (double)fz;
        v10 -= n1 * (double)fx; v11 -= n1 * (double)fy; v12 -= n1 * (double)fz;
        v20 -= n2 * (double)fx; v21 -= n2 * (double)fy; v22 -= n2 * (double)fz;
        gq += (double)gqI;
        gmx += (double)(gmIx + e_mIz * zx);
        gmy += (double)(gmIy + e_mIz * zy);
        gmz += (double)(gmIz + e_mIz * zz);
        // dE/dTheta_I = e_zzI z z^T + sym(z v^T) + e_ThTh Theta_J   (matrix entries)
        gTxx += (double)(e_zzI * zx * zx + zx * vx + e_ThTh * TJxx);
        gTyy += (double)(e_zzI * zy * zy + zy * vy + e_ThTh * TJyy);
        gTzz += (double)(e_zzI * zz * zz + zz * vz + e_ThTh * TJzz);
        gTxy += (double)(e_zzI * zx * zy + (R)0.5 * (zx * vy + zy * vx) + e_ThTh * TJxy);
        gTxz += (double)(e_zzI * zx * zz + (R)0.5 * (zx * vz + zz * vx) + e_ThTh * TJxz);
        gTyz += (double)(e_zzI * zy * zz + (R)0.5 * (zy * vz + zz * vy) + e_ThTh * TJyz);
        if (LPOL) {
            gth += (double)gthe;
            if (alI > (R)0) gpol += (double)(gdmp * (a6I * sj[10]) / ((R)6 * alI));
        }
    }
    E = group_sum(E, TPA);
    gx = group_sum(gx, TPA); gy = group_sum(gy, TPA); gz = group_sum(gz, TPA);
    gq = group_sum(gq, TPA);
    gmx = group_sum(gmx, TPA); gmy = group_sum(gmy, TPA); gmz = group_sum(gmz, TPA);
    gTxx = group_sum(gTxx, TPA); gTyy = group_sum(gTyy, TPA); gTzz = group_sum(gTzz, TPA);
    gTxy = group_sum(gTxy, TPA); gTxz = group_sum(gTxz, TPA); gTyz = group_sum(gTyz, TPA);
    v00 = group_sum(v00, TPA); v01 = group_sum(v01, TPA); v02 = group_sum(v02, TPA);
    v10 = group_sum(v10, TPA); v11 = group_sum(v11, TPA); v12 = group_sum(v12, TPA);
    v20 = group_sum(v20, TPA); v21 = group_sum(v21, TPA); v22 = group_sum(v22, TPA);
    if (LPOL) { gpol = group_sum(gpol, TPA); gth = group_sum(gth, TPA); }
    if (active && sub == 0) {
        const size_t a = fo + i;
        e_atom[a] = 0.5 * E;
        g_pos[a * 3 + 0] = gx; g_pos[a * 3 + 1] = gy; g_pos[a * 3 + 2] = gz;
        double* q = g_q + a * 10;
        q[0] = gq; q[1] = gmx; q[2] = gmy; q[3] = gmz;
        q[4] = gTxx; q[5] = gTyy; q[6] = gTzz; q[7] = gTxy; q[8] = gTxz; q[9] = gTyz;
        double* v = vir_atom + a * 9;
        v[0] = 0.5 * v00; v[1] = 0.5 * v01; v[2] = 0.5 * v02;
        v[3] = 0.5 * v10; v[4] = 0.5 * v11; v[5] = 0.5 * v12;
        v[6] = 0.5 * v20; v[7] = 0.5 * v21; v[8] = 0.5 * v22;
        if (LPOL) { g_pol[a] = gpol; g_thole[a] = gth; }
    }
}

// ---------------------------------------------------------------------------
// Solver set-up: real-space part of -dE/dU at U = 0 (the permanent field felt
// by each induced dipole, p-scaled and Thole damped), and the two coefficients
// of the induced-induced coupling of each directed pair, stored for the matvec:
//   dE_dd/dU_I = sum_J c1 U_J + c2 (z.U_J) z,  c1 = udud_m1, c2 = udud_m0 - udud_m1,
// with z = d / r for the minimum-image separation d; the second coefficient is
// stored as c2 / r^2, so that the product forms c2 (z.U) z = (c2 / r^2)(d.U) d
// without a square root or a division.
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(PAIR_BLOCK)
k_real_permfield(int N, int cap, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
                 const R* __restrict__ site, const int* __restrict__ nbr, const int* __restrict__ ncnt,
                 PairParams pp, double* __restrict__ field, R* __restrict__ pcoef, int* __restrict__ pnear,
                 int* __restrict__ nchild) {
    const int f = blockIdx.y;
    const int gt = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    const int i = gt / TPA, sub = gt % TPA;
    const bool active = i < N;
    const int ii = active ? i : 0;
    const BoxInfo& bx = boxes[f];
    const size_t fo = (size_t)f * N;
    const double4 pi_ = pos[fo + ii];
    const R* si = site + (fo + ii) * SITE;
    const R a6I = si[10], thI = si[11];
    const R kap = (R)pp.kappa;
    double bx_ = 0, by_ = 0, bz_ = 0;
    const int cnt = active ? ncnt[fo + ii] : 0;
    const int* nl = nbr + (fo + ii) * (size_t)cap;
    R* pc = pcoef + (fo + ii) * (size_t)cap * 2;
    // nearest neighbour with induced-induced coupling (d-scale > 0) inside
    // PC_RPAIR: the nearest-neighbour link of the block preconditioner
    double nr = PC_RPAIR;
    int nj = -1, nk = -1;
    for (int k = sub; k < cnt; k += TPA) {
        const int w = nl[k];
        const int j = w & NB_MASK, nb = w >> NB_SHIFT;
        const double4 pj = pos[fo + j];
        double dx = pi_.x - pj.x, dy = pi_.y - pj.y, dz = pi_.z - pj.z, n0, n1, n2;
        pbc_shift(bx, dx, dy, dz, n0, n1, n2);
        const double rd = sqrt(dx * dx + dy * dy + dz * dz);
        const R r = (R)rd, rinv = (R)(1.0 / rd);
        const R zx = (R)dx * rinv, zy = (R)dy * rinv, zz = (R)dz * rinv;
        const R* sj = site + (fo + j) * SITE;
        const R qJ = sj[0], mJx = sj[1], mJy = sj[2], mJz_ = sj[3];
        const R TJxx = sj[4], TJyy = sj[5], TJzz = sj[6], TJxy = sj[7], TJxz = sj[8], TJyz = sj[9];
        const int sidx = scale_index(nb, pp.nscale);
        const R ps = (R)pp.ps[sidx], dsc = (R)pp.ds[sidx];
        if (pp.ds[sidx] > 0.0 && (rd < nr || (rd == nr && nj >= 0 && j < nj))) { nr = rd; nj = j; nk = k; }
        const R mJz = mJx * zx + mJy * zy + mJz_ * zz;
        const R tJx = TJxx * zx + TJxy * zy + TJxz * zz;
        const R tJy = TJxy * zx + TJyy * zy + TJyz * zz;
        const R tJz = TJxz * zx + TJyz * zy + TJzz * zz;
        const R zzJ = tJx * zx + tJy * zy + tJz * zz;
        Radial<R> rq;
        radial<R>(r, kap, rq);
        const R a = ps > (R)0 ? (R)5.0 : thI + sj[11];
        R dmp = a6I * sj[10];
        if (dmp < (R)1e-8) dmp = (R)1e-8;
        R u = r / dmp;
        if (!(u < (R)1e8)) u = (R)1e8;
        Thole<R> th;
        thole_f<R>(a * u, th);
        R cud, dud0, dud1, udq0, udq1, ud0, ud1, t;
        f_cd<R>(rq, ps * th.c, (R)0, cud, t);
        f_dd0<R>(rq, ps * th.d0, (R)0, dud0, t);
        f_dd1<R>(rq, ps * th.c, (R)0, dud1, t);
        f_dq0<R>(rq, ps * th.q0, (R)0, udq0, t);
        f_dq1<R>(rq, ps * th.q1, (R)0, udq1, t);
        f_dd0<R>(rq, dsc * th.d0, (R)0, ud0, t);
        f_dd1<R>(rq, dsc * th.c, (R)0, ud1, t);
        const R k1 = (R)K1_INV;
        // dE/dU_I at U = 0
        const R s = -cud * qJ + (dud0 - dud1) * mJz + (k1 * udq1 - udq0) * zzJ;
        const R gxI = s * zx + dud1 * mJx - k1 * udq1 * tJx;
        const R gyI = s * zy + dud1 * mJy - k1 * udq1 * tJy;
        const R gzI = s * zz + dud1 * mJz_ - k1 * udq1 * tJz;
        bx_ -= (double)gxI; by_ -= (double)gyI; bz_ -= (double)gzI;
        pc[2 * k] = ud1;
        pc[2 * k + 1] = (ud0 - ud1) * rinv * rinv;
    }
    bx_ = group_sum(bx_, TPA); by_ = group_sum(by_, TPA); bz_ = group_sum(bz_, TPA);
    // smallest (distance, atom index) over the lanes of the atom
    for (int o = TPA >> 1; o > 0; o >>= 1) {
        const double r2 = __shfl_xor(nr, o, TPA);
        const int j2 = __shfl_xor(nj, o, TPA), k2 = __shfl_xor(nk, o, TPA);
        if (j2 >= 0 && (nj < 0 || r2 < nr || (r2 == nr && j2 < nj))) { nr = r2; nj = j2; nk = k2; }
    }
    if (active && sub == 0) {
        const size_t a = fo + i;
        field[a * 3 + 0] = bx_; field[a * 3 + 1] = by_; field[a * 3 + 2] = bz_;
        if (pnear) pnear[a] = nk;
        if (nchild) nchild[a] = 0;   // counters of k_pc_children, which runs next
    }
}

// ---------------------------------------------------------------------------
// Block preconditioner (set-up once per solve; positions are fixed).
// Every site has a nearest coupled neighbour n(i) inside PC_RPAIR (or none).
// Two sites with n(i) = j and n(j) = i form a mutual pair (p < q); a site c
// whose n(c) lies in a mutual pair, c outside it, is a child of that site.  A
// cluster is a mutual pair plus up to maxm - 2 children of its two sites, the
// nearest to their parent first (site index on ties); a site whose parent has
// more than PC_MAXCH children takes no child of it (the recorded ones depend
// on atomic order).  In water the clusters are an oxygen with the hydrogens
// hydrogen-bonded to it inside PC_RPAIR.  The cluster's block of the
// polarisation matrix is approximated by
//   diagonal blocks d_i I, d = DIELECTRIC / alpha,
//   for the pair and each child-parent link: B = c1 I + c2 z z^T  (the
//   Thole-damped dipole-dipole tensor without Ewald screening, d-scaled, with
//   the Thole width of the induced-induced product and the bonded distance of
//   the link's neighbour-list entry); the other couplings inside a cluster are
//   left out (numpy on the dense matrix: the same iteration count),
// a tree, whose inverse is exact in 3x3 blocks: eliminating the children
// leaves the pair with P_x = d_x I - sum_{c -> x} B_c B_c / d_c, then
//   S = P_q - B_pq P_p^-1 B_pq,  K_qq = S^-1,  K_pq = -P_p^-1 B_pq S^-1,
//   K_pp = P_p^-1 + P_p^-1 B_pq S^-1 B_pq P_p^-1,
// and for children (parent x, y any pair site, c' another child):
//   M^-1_cy = -B_c K_xy / d_c,  M^-1_cc' = delta I / d_c + B_c K_xx' B_c' / (d_c d_c').
// Every member evaluates the same expressions in the same order, so every
// member holds the same inverse.  |det| of P_p or S below 1e-3 d^3 (a nearly
// singular block) makes the cluster fall back to its pair, and the pair to the
// diagonal 1/d.  The blocks of close contacts may be indefinite; they carry the
// local modes that slow a diagonal-preconditioned solve most.  Output per site:
// the partner sites pcj[3] (-1: unused; pcj[0] = -1: diagonal) and its three
// rows of the inverse, pcd[3][12] (column blocks: the site itself, then the
// partners in pcj order).
// ---------------------------------------------------------------------------
constexpr int PC_MAXM = 4;    // sites per cluster
constexpr int PC_MAXCH = 4;   // children recorded per pair site

// n(i): the nearest coupled neighbour of site i inside PC_RPAIR, -1 if none
__device__ __forceinline__ int pc_nn(const int* __restrict__ nbr, const int* __restrict__ pnear, size_t fo,
                                     int cap, int i) {
    const int k = pnear[fo + i];
    return k < 0 ? -1 : (nbr[(fo + i) * (size_t)cap + k] & NB_MASK);
}

// records each child c of a mutual-pair site p in child[p][slot]; nchild
// (cleared by k_real_permfield) counts them, also past PC_MAXCH
__global__ void __launch_bounds__(256)
k_pc_children(int N, int cap, const int* __restrict__ nbr, const int* __restrict__ pnear, int* __restrict__ nchild,
              int* __restrict__ child) {
    const int f = blockIdx.y, c = blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    const size_t fo = (size_t)f * N;
    const int p = pc_nn(nbr, pnear, fo, cap, c);
    if (p < 0) return;
    const int q = pc_nn(nbr, pnear, fo, cap, p);
    if (q < 0 || q == c || pc_nn(nbr, pnear, fo, cap, q) != p) return;
    const int slot = atomicAdd(nchild + fo + p, 1);
    if (slot < PC_MAXCH) child[(fo + p) * PC_MAXCH + slot] = c;
}

// 16 lanes per site: the site's cluster block is solved by Gauss-Jordan
// elimination with one matrix column per lane (lanes 0..11) and the three unit
// right-hand sides of the site's own rows in lanes 12..14; pivot columns are
// broadcast by shuffles.
constexpr int PC_LANES = 16;

template <class R>
__global__ void __launch_bounds__(256)
k_pc_clusters(int N, int cap, int maxm, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
              const R* __restrict__ site, const double* __restrict__ pol, const int* __restrict__ nbr,
              const int* __restrict__ ncnt, const int* __restrict__ pnear, const int* __restrict__ nchild,
              const int* __restrict__ child, PairParams pp, int* __restrict__ pcj, double* __restrict__ pcd) {
    static_assert(PC_MAXM == 4 && 256 % PC_LANES == 0, "four sites per cluster, whole lane groups per block");
    constexpr int NR = 3 * PC_MAXM;
    const int f = blockIdx.y;
    const int gt = blockIdx.x * 256 + threadIdx.x;
    const int a = gt / PC_LANES, lane = gt % PC_LANES;
    if (a >= N) return;  // the whole lane group
    const size_t fo = (size_t)f * N, ga = fo + a;
    const BoxInfo& bx = boxes[f];
    int outj[3] = {-1, -1, -1};
    // Every condition below depends on the site only, so the 16 lanes of a
    // group take the same branches (the shuffles need the whole group).
    [&]() {
        // the mutual pair (p < q) of a, as one of its sites or as a child
        int p, q;
        {
            const int j = pc_nn(nbr, pnear, fo, cap, a);
            if (j < 0) return;
            const int jj = pc_nn(nbr, pnear, fo, cap, j);
            if (jj == a) { p = a; q = j; }
            else if (jj >= 0 && pc_nn(nbr, pnear, fo, cap, jj) == j) { p = j; q = jj; }
            else return;
            if (p > q) { const int t = p; p = q; q = t; }
        }
        // children: up to maxm - 2, nearest to their parent first (index on ties),
        // then in ascending index order
        int ch0 = -1, ch1 = -1, nch = 0;
        if (maxm > 2) {
            int cand[2 * PC_MAXCH];
            double cd[2 * PC_MAXCH];
            int nc = 0;
            for (int s2 = 0; s2 < 2; ++s2) {
                const int par = s2 ? q : p;
                const int n = nchild[fo + par];
                if (n > PC_MAXCH) continue;
                for (int t = 0; t < n; ++t) {
                    const int c = child[(fo + par) * PC_MAXCH + t];
                    const double4 pa = pos[fo + par], pc = pos[fo + c];
                    double dx = pa.x - pc.x, dy = pa.y - pc.y, dz = pa.z - pc.z, n0, n1, n2;
                    pbc_shift(bx, dx, dy, dz, n0, n1, n2);
                    cand[nc] = c;
                    cd[nc] = dx * dx + dy * dy + dz * dz;
                    ++nc;
                }
            }
            for (int pick = 0; pick < maxm - 2 && nc > 0; ++pick) {
                int b = 0;
                for (int t = 1; t < nc; ++t)
                    if (cd[t] < cd[b] || (cd[t] == cd[b] && cand[t] < cand[b])) b = t;
                if (pick == 0) ch0 = cand[b]; else ch1 = cand[b];
                ++nch;
                cand[b] = cand[nc - 1];
                cd[b] = cd[nc - 1];
                --nc;
            }
            if (nch == 2 && ch1 < ch0) { const int t = ch0; ch0 = ch1; ch1 = t; }
        }
        int self = a == p ? 0 : a == q ? 1 : a == ch0 ? 2 : a == ch1 ? 3 : -1;
        if (self < 0) return;  // a child left out of its pair's cluster
        const int mem[PC_MAXM] = {p, q, ch0, ch1};
        // bonded distance of every member pair (t < u) from the neighbour-list
        // entry the product uses (-1: no entry, uncoupled): the pnear entry
        // for a nearest-neighbour link, else a scan of the row shared by the lanes
        int nbp[6];
        {
            int e = 0;
#pragma unroll
            for (int t = 0; t < PC_MAXM; ++t)
#pragma unroll
                for (int u = t + 1; u < PC_MAXM; ++u, ++e) {
                    // every group runs the same sequence of shuffles (n = 0 where no scan is needed)
                    int found = -1, n = 0;
                    const int* row = nbr;
                    const int x = mem[t], y = mem[u];
                    if (u < 2 + nch) {
                        const int kx = pnear[fo + x], ky = pnear[fo + y];
                        const int wx = kx >= 0 ? nbr[(fo + x) * (size_t)cap + kx] : -1;
                        const int wy = ky >= 0 ? nbr[(fo + y) * (size_t)cap + ky] : -1;
                        if (kx >= 0 && (wx & NB_MASK) == y) found = wx >> NB_SHIFT;
                        else if (ky >= 0 && (wy & NB_MASK) == x) found = wy >> NB_SHIFT;
                        else { row = nbr + (fo + x) * (size_t)cap; n = ncnt[fo + x]; }
                    }
                    for (int k = lane; k < n; k += PC_LANES)
                        if ((row[k] & NB_MASK) == y) found = row[k] >> NB_SHIFT;
                    for (int o = PC_LANES >> 1; o > 0; o >>= 1) found = max(found, __shfl_xor(found, o, PC_LANES));
                    nbp[e] = found;
                }
        }
        auto dinv = [&](int x) {
            const double al = pol[x];
            return DIELECTRIC / (al < 1e-8 ? 1e-8 : al);
        };
        // (k, l) element of the coupling block of members x, y with bonded distance nbl
        auto coupling = [&](int x, int y, int nbl, int k, int l) {
            const double4 px = pos[fo + x], py = pos[fo + y];
            double dx = px.x - py.x, dy = px.y - py.y, dz = px.z - py.z, n0, n1, n2;
            pbc_shift(bx, dx, dy, dz, n0, n1, n2);
            const double r = sqrt(dx * dx + dy * dy + dz * dz), ri = 1.0 / r;
            const int sidx = scale_index(nbl, pp.nscale);
            const double dsc = pp.ds[sidx];
            const R* sx = site + (fo + x) * SITE;
            const R* sy = site + (fo + y) * SITE;
            const double th_a = pp.ps[sidx] > 0.0 ? 5.0 : (double)sx[11] + (double)sy[11];
            double dmp = (double)sx[10] * (double)sy[10];
            if (dmp < 1e-8) dmp = 1e-8;
            double u = r / dmp;
            if (!(u < 1e8)) u = 1e8;
            Thole<double> th;
            thole_f<double>(th_a * u, th);
            const double P3 = DIELECTRIC * ri * ri * ri;
            const double c1 = P3 * dsc * th.c, c2 = P3 * dsc * (-2.0 * th.d0 - th.c);
            const double z[3] = {dx * ri, dy * ri, dz * ri};
            return (k == l ? c1 : 0.0) + c2 * z[k] * z[l];
        };
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1) {
                if (nch == 0 || self >= 2) return;  // the pair block failed, or a child of a failed cluster
                nch = 0;
            }
            const int ncl = 2 + nch;
            double dref[PC_MAXM];
#pragma unroll
            for (int t = 0; t < PC_MAXM; ++t) dref[t] = t < ncl ? dinv(mem[t]) : 1.0;
            // this lane's column: matrix column j = 3 u + l, or a unit right-hand side
            double col[NR];
            if (lane < NR) {
                const int u = lane / 3, l = lane % 3;
#pragma unroll
                for (int t = 0; t < PC_MAXM; ++t) {
                    const int e = t < u ? t * (2 * PC_MAXM - t - 1) / 2 + (u - t - 1)
                                        : u * (2 * PC_MAXM - u - 1) / 2 + (t - u - 1);
                    const bool cpl = t != u && t < ncl && u < ncl && nbp[t == u ? 0 : e] >= 0;
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        col[3 * t + k] = t == u ? (k == l ? dref[t] : 0.0)
                                                : (cpl ? coupling(mem[t], mem[u], nbp[e], k, l) : 0.0);
                }
            } else {
#pragma unroll
                for (int r = 0; r < NR; ++r) col[r] = (lane - NR < 3 && r == 3 * self + (lane - NR)) ? 1.0 : 0.0;
            }
            // Gauss-Jordan without pivoting (the diagonal blocks dominate unless
            // the block is nearly singular, which the pivot test catches)
            bool ok = true;
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                double ck[NR];
#pragma unroll
                for (int r = 0; r < NR; ++r) ck[r] = __shfl(col[r], k, PC_LANES);
                const double piv = ck[k];
                ok = ok && fabs(piv) >= 1e-3 * dref[k / 3];
                const double fk = col[k] / piv;
#pragma unroll
                for (int r = 0; r < NR; ++r) col[r] = r == k ? fk : col[r] - ck[r] * fk;
            }
            if (!ok) continue;
            // lanes 12..14: column r of the inverse = a's row r (the block is symmetric)
            if (lane >= NR && lane < NR + 3) {
                const int r = lane - NR;
                double* o = pcd + ga * (3 * NR) + r * NR;
                int slot = 0;
#pragma unroll
                for (int t = 0; t < PC_MAXM; ++t) {
                    const int blk = t == self ? 0 : ++slot;
#pragma unroll
                    for (int c = 0; c < 3; ++c) o[3 * blk + c] = t < ncl ? col[3 * t + c] : 0.0;
                }
            }
            int slot = 0;
            for (int t = 0; t < PC_MAXM; ++t)
                if (t != self) { if (t < ncl) outj[slot] = mem[t]; ++slot; }
            return;
        }
    }();
    if (lane == 0) {
        int* oj = pcj + ga * 3;
        oj[0] = outj[0]; oj[1] = outj[1]; oj[2] = outj[2];
    }
}

// out_I = sum_J c1 V_J + c2 (z.V_J) z   (induced-induced real-space coupling),
// formed as c1 V_J + (c2 / r^2)(d.V_J) d from the stored (c1, c2 / r^2) and the
// minimum-image separation d (no square root or division per pair).
// RC: stored coefficient type, R: arithmetic type of the products (double for the
// FP64 residual of the mixed-precision solve)
// One-dimensional grid over all (frame, atom block) items, renumbered so that
// each XCD works through a contiguous range of frames and atoms (its L2 then
// holds the positions and v of few frames).  nbf: atom blocks per frame.
// item: (frame, atom block of BS threads) pair, numbered frame-major; nbf atom
// blocks per frame
template <bool ORTHO, class RC, class R>
__device__ __forceinline__ void real_matvec_pairs(int sub, int cnt, const BoxInfo& bx, const double4& pi_,
                                                  const double4* __restrict__ posf, const R* __restrict__ vf,
                                                  const int* __restrict__ nl, const RC* __restrict__ pc, double& ox,
                                                  double& oy, double& oz) {
    // one neighbour per trip: the loop is bound by the random position and
    // direction loads, and the fewer registers of a plain loop beat two or four
    // neighbours per trip (64-frame batch: 2.3 % and 5 % faster).  posf and vf
    // are the frame's bases, so the per-neighbour offsets are 32-bit.
    auto term = [&](int j, RC c1r, RC c2r) {
        const double4 pj = posf[j];
        double dx = pi_.x - pj.x, dy = pi_.y - pj.y, dz = pi_.z - pj.z;
        if constexpr (ORTHO) {
            // the general minimum image of pbc_shift with zero off-diagonal
            // terms (which add exact zeros there): the same value
            const double t0 = dx * bx.Binv[0], t1 = dy * bx.Binv[4], t2 = dz * bx.Binv[8];
            dx = (t0 - floor(t0 + 0.5)) * bx.B[0];
            dy = (t1 - floor(t1 + 0.5)) * bx.B[4];
            dz = (t2 - floor(t2 + 0.5)) * bx.B[8];
        } else {
            double n0, n1, n2;
            pbc_shift(bx, dx, dy, dz, n0, n1, n2);
        }
        const R ex = (R)dx, ey = (R)dy, ez = (R)dz;
        const R* vj = vf + 3 * j;
        const R vx = vj[0], vy = vj[1], vz = vj[2];
        const R c1 = (R)c1r, c2 = (R)c2r;
        const R p = c2 * (ex * vx + ey * vy + ez * vz);
        ox += (double)(c1 * vx + p * ex);
        oy += (double)(c1 * vy + p * ey);
        oz += (double)(c1 * vz + p * ez);
    };
    for (int k = sub; k < cnt; k += TPA) term(nl[k] & NB_MASK, pc[2 * k], pc[2 * k + 1]);
}

template <class RC, class R, int BS>
__device__ __forceinline__ void real_matvec_body(int item, int N, int nbf, int cap, const double4* __restrict__ pos,
                                                 const BoxInfo* __restrict__ boxes, const int* __restrict__ nbr,
                                                 const int* __restrict__ ncnt, const RC* __restrict__ pcoef,
                                                 const R* __restrict__ v, double* __restrict__ out,
                                                 const PcgState* __restrict__ skip) {
    const int f = item / nbf, bxi = item - f * nbf;
    if (skip && skip[f].done) return;  // frame already converged
    const int gt = bxi * BS + threadIdx.x;
    const int i = gt / TPA, sub = gt % TPA;
    const bool active = i < N;
    const int ii = active ? i : 0;
    const BoxInfo& bx = boxes[f];
    const size_t fo = (size_t)f * N;
    const double4* posf = pos + fo;
    const R* vf = v + fo * 3;
    const double4 pi_ = posf[ii];
    double ox = 0, oy = 0, oz = 0;
    const int cnt = active ? ncnt[fo + ii] : 0;
    const int* nl = nbr + (fo + ii) * (size_t)cap;
    const RC* pc = pcoef + (fo + ii) * (size_t)cap * 2;
    if (bx.ortho)
        real_matvec_pairs<true>(sub, cnt, bx, pi_, posf, vf, nl, pc, ox, oy, oz);
    else
        real_matvec_pairs<false>(sub, cnt, bx, pi_, posf, vf, nl, pc, ox, oy, oz);
    ox = group_sum(ox, TPA); oy = group_sum(oy, TPA); oz = group_sum(oz, TPA);
    if (active && sub == 0) {
        const size_t a = fo + i;
        out[a * 3 + 0] = ox; out[a * 3 + 1] = oy; out[a * 3 + 2] = oz;
    }
}

template <class RC, class R>
__global__ void __launch_bounds__(PAIR_BLOCK)
k_real_matvec(int N, int nbf, int F, int cap, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
              const int* __restrict__ nbr, const int* __restrict__ ncnt, const RC* __restrict__ pcoef,
              const R* __restrict__ v, double* __restrict__ out, const PcgState* __restrict__ skip) {
    const int item = xcd_item(blockIdx.x, nbf * F);
    if (item >= nbf * F) return;
    real_matvec_body<RC, R, PAIR_BLOCK>(item, N, nbf, cap, pos, boxes, nbr, ncnt, pcoef, v, out, skip);
}

}  // namespace admp

namespace admp {

// ---------------------------------------------------------------------------
// Short-range scalar pair terms on the same full lists:
//   MODE 0: dispersion real space, sum_p (m + g_p(kappa^2 r^2) - 1) c_i^p c_j^p / r^p
//           (disp_pme.py:180-226; g_p of :229-261)
//   MODE SR_*: the generate_pairwise_interaction kernels of sr_kernels.h
//           (pairwise.py:98-169)
// prm: MODE 0 -> c (N,3); SR kinds -> (N, sr_nprm(MODE)) in the kernel's
// parameter order.  Outputs per site: energy (half of each pair), dE/dr_i,
// virial (half), and dE/dparams of site i (row stride gstride).
// ---------------------------------------------------------------------------
template <class R, int MODE>
__global__ void __launch_bounds__(PAIR_BLOCK)
k_pair_short(int N, int cap, const double4* __restrict__ pos, const BoxInfo* __restrict__ boxes,
             const double* __restrict__ prm, const int* __restrict__ nbr, const int* __restrict__ ncnt,
             PairParams pp, int pmax, double* __restrict__ e_atom, double* __restrict__ g_pos,
             double* __restrict__ vir_atom, double* __restrict__ g_prm, int gstride) {
    const int f = blockIdx.y;
    const int gt = blockIdx.x * PAIR_BLOCK + threadIdx.x;
    const int i = gt / TPA, sub = gt % TPA;
    const bool active = i < N;
    const int ii = active ? i : 0;
    const BoxInfo& bx = boxes[f];
    const size_t fo = (size_t)f * N;
    const double4 pi_ = pos[fo + ii];
    constexpr int np = MODE == 0 ? 3 : sr_nprm(MODE);
    R pI[SR_NPRM_MAX] = {0, 0, 0, 0};
    for (int k = 0; k < np; ++k) pI[k] = (R)prm[(size_t)ii * np + k];
    const R kap2 = (R)(pp.kappa * pp.kappa);
    double E = 0, gx = 0, gy = 0, gz = 0;
    double gp[SR_NPRM_MAX] = {0, 0, 0, 0};
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int cnt = active ? ncnt[fo + ii] : 0;
    const int* nl = nbr + (fo + ii) * (size_t)cap;
    for (int k = sub; k < cnt; k += TPA) {
        const int w = nl[k];
        const int j = w & NB_MASK, nb = w >> NB_SHIFT;
        const double4 pj = pos[fo + j];
        double dx = pi_.x - pj.x, dy = pi_.y - pj.y, dz = pi_.z - pj.z, n0, n1, n2;
        pbc_shift(bx, dx, dy, dz, n0, n1, n2);
        const double rd = sqrt(dx * dx + dy * dy + dz * dz);
        const R r = (R)rd, rinv = (R)(1.0 / rd);
        const R m = (R)pp.ms[scale_index(nb, pp.nscale)];
        R e = 0, dEdr = 0;
        R dI[SR_NPRM_MAX] = {0, 0, 0, 0};
        if constexpr (MODE == 0) {
            const R r2 = r * r, x2 = kap2 * r2, ex = r_exp<R>(-x2);
            const R x4 = x2 * x2;
            const R g6 = ((R)1 + x2 + (R)0.5 * x4) * ex;
            const R dg6 = -(R)0.5 * x4 * ex;                  // d g6 / d x2
            const R g8 = g6 + x4 * x2 * (R)(1.0 / 6.0) * ex;
            const R dg8 = -x4 * x2 * (R)(1.0 / 6.0) * ex;
            const R g10 = g8 + x4 * x4 * (R)(1.0 / 24.0) * ex;
            const R dg10 = -x4 * x4 * (R)(1.0 / 24.0) * ex;
            const R dx2dr = (R)2 * kap2 * r;
            const R ir2 = rinv * rinv, ir6 = ir2 * ir2 * ir2, ir8 = ir6 * ir2, ir10 = ir8 * ir2;
            const R cj6 = (R)prm[(size_t)j * 3], cj8 = (R)prm[(size_t)j * 3 + 1], cj10 = (R)prm[(size_t)j * 3 + 2];
            R s6 = m + g6 - (R)1;
            e = s6 * pI[0] * cj6 * ir6;
            dEdr = pI[0] * cj6 * (dg6 * dx2dr * ir6 - (R)6 * s6 * ir6 * rinv);
            dI[0] = s6 * cj6 * ir6;
            if (pmax >= 8) {
                R s8 = m + g8 - (R)1;
                e += s8 * pI[1] * cj8 * ir8;
                dEdr += pI[1] * cj8 * (dg8 * dx2dr * ir8 - (R)8 * s8 * ir8 * rinv);
                dI[1] = s8 * cj8 * ir8;
            }
            if (pmax >= 10) {
                R s10 = m + g10 - (R)1;
                e += s10 * pI[2] * cj10 * ir10;
                dEdr += pI[2] * cj10 * (dg10 * dx2dr * ir10 - (R)10 * s10 * ir10 * rinv);
                dI[2] = s10 * cj10 * ir10;
            }
        } else {
            R pJ[SR_NPRM_MAX] = {0, 0, 0, 0}, dJ[SR_NPRM_MAX];
            for (int q = 0; q < np; ++q) pJ[q] = (R)prm[(size_t)j * np + q];
            sr_pair<R, MODE>(r, m, pI, pJ, e, dEdr, dI, dJ);
        }
        const R fx = dEdr * (R)dx * rinv, fy = dEdr * (R)dy * rinv, fz = dEdr * (R)dz * rinv;
        E += (double)e;
        gx += (double)fx; gy += (double)fy; gz += (double)fz;
        for (int q = 0; q < np; ++q) gp[q] += (double)dI[q];
        const double nn[3] = {n0, n1, n2}, ff[3] = {(double)fx, (double)fy, (double)fz};
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) v[p * 3 + q] -= nn[p] * ff[q];
    }
    E = group_sum(E, TPA);
    gx = group_sum(gx, TPA); gy = group_sum(gy, TPA); gz = group_sum(gz, TPA);
    for (int q = 0; q < np; ++q) gp[q] = group_sum(gp[q], TPA);
    for (int k = 0; k < 9; ++k) v[k] = group_sum(v[k], TPA);
    if (active && sub == 0) {
        const size_t a = fo + i;
        e_atom[a] = 0.5 * E;
        g_pos[a * 3] = gx; g_pos[a * 3 + 1] = gy; g_pos[a * 3 + 2] = gz;
        for (int k = 0; k < 9; ++k) vir_atom[a * 9 + k] = 0.5 * v[k];
        for (int q = 0; q < np; ++q) g_prm[a * gstride + q] = gp[q];
    }
}

}  // namespace admp
