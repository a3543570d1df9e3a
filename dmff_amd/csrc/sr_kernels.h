// Short-range pair terms of the ADMP force fields (dmff/admp/pairwise.py), the
// kernels behind generate_pairwise_interaction (pairwise.py:47-95):
//
//   SR_SLATER       slater_sr_kernel (:156-169)            params (A, B)
//                   m A_i A_j (1/3 x^2 + x + 1) e^{-x},  x = b r,  b = sqrt(B_i B_j)
//   SR_TT_QQ_C6     TT_damping_qq_c6_kernel (:98-117)     params (A, B, Q, C)
//                   m [2625.5 a e^{-x} - 2625.5 e^{-x} (1 + x) q / t + e^{-x} S6(x) c / r^6],
//                   t = 1.889726878 r (bohr), x = b t, a = sqrt(A_i A_j), b = sqrt(B_i B_j),
//                   q = Q_i Q_j, c = C_i C_j, S_n(x) = sum_{k<=n} x^k / k!
//   SR_TT_QQ        TT_damping_qq_kernel (:120-128)       params (B, Q)
//                   -m DIELECTRIC e^{-x} (1 + x) q / r,  x = b r
//   SR_SLATER_DAMP  slater_disp_damping_kernel (:131-153) params (B, C6, C8, C10)
//                   m sum_{n=6,8,10} e^{-y} S_n(y) c_n / r^n,
//                   y = x - (2 x^2 + 3 x) / (x^2 + 3 x + 3),  x = b r,  c_n = c_n,i c_n,j
//
// sr_pair() returns the energy of one pair (topological scale m included), its
// derivative with respect to the distance r and the derivatives with respect to
// the parameters of site i and of site j (sqrt-combined parameters:
// d/dP_i = d/dp * P_j / (2 p), zero where p = 0; product-combined: d/dP_i = d/dp * P_j).
#pragma once
#include "admp_common.h"

namespace admp {

enum SrKind { SR_SLATER = 1, SR_TT_QQ_C6 = 2, SR_TT_QQ = 3, SR_SLATER_DAMP = 4 };
constexpr int SR_KIND_MAX = 4;
constexpr int SR_NPRM_MAX = 4;

__host__ __device__ constexpr int sr_nprm(int kind) {
    return kind == SR_SLATER ? 2 : kind == SR_TT_QQ_C6 ? 4 : kind == SR_TT_QQ ? 2 : kind == SR_SLATER_DAMP ? 4 : 0;
}

template <class R> __device__ __forceinline__ R sr_exp(R x) { return exp(x); }

// d/dP_i of a sqrt-combined parameter p = sqrt(P_i P_j)
template <class R> __device__ __forceinline__ R sqrt_chain(R dEdp, R p, R Pother) {
    return p > (R)0 ? dEdp * Pother / ((R)2 * p) : (R)0;
}

template <class R, int KIND>
__device__ __forceinline__ void sr_pair(R r, R m, const R* pi, const R* pj, R& e, R& dedr, R* dpi, R* dpj) {
    if (KIND == SR_SLATER) {
        const R a = pi[0] * pj[0];
        const R b = sqrt(pi[1] * pj[1]);
        const R x = b * r, ex = sr_exp<R>(-x);
        const R P = x * x * (R)(1.0 / 3.0) + x + (R)1;
        const R dPe = -(x * (R)(1.0 / 3.0)) * ((R)1 + x) * ex;  // d(P e^-x)/dx
        e = m * a * P * ex;
        dedr = m * a * b * dPe;
        const R dEda = m * P * ex, dEdb = m * a * r * dPe;
        dpi[0] = dEda * pj[0];
        dpj[0] = dEda * pi[0];
        dpi[1] = sqrt_chain(dEdb, b, pj[1]);
        dpj[1] = sqrt_chain(dEdb, b, pi[1]);
    } else if (KIND == SR_TT_QQ) {
        const R b = sqrt(pi[0] * pj[0]);
        const R q = pi[1] * pj[1];
        const R x = b * r, ex = sr_exp<R>(-x);
        const R rinv = (R)1 / r;
        const R g = ((R)1 + x) * ex * rinv;                           // e^{-br}(1+br)/r
        const R dgdr = -b * b * ex - ((R)1 + x) * ex * rinv * rinv;
        const R D = (R)DIELECTRIC;
        e = -D * m * q * g;
        dedr = -D * m * q * dgdr;
        const R dEdb = D * m * q * b * r * ex;                        // -D m q dg/db, dg/db = -b r e^{-br}
        const R dEdq = -D * m * g;
        dpi[0] = sqrt_chain(dEdb, b, pj[0]);
        dpj[0] = sqrt_chain(dEdb, b, pi[0]);
        dpi[1] = dEdq * pj[1];
        dpj[1] = dEdq * pi[1];
    } else if (KIND == SR_TT_QQ_C6) {
        const R H = (R)2625.5, KB = (R)1.889726878;
        const R a = sqrt(pi[0] * pj[0]);
        const R b = sqrt(pi[1] * pj[1]);
        const R q = pi[2] * pj[2];
        const R c = pi[3] * pj[3];
        const R t = r * KB, x = b * t, ex = sr_exp<R>(-x);
        const R x2 = x * x, x3 = x2 * x, x6 = x3 * x3;
        const R S6 = (R)1 + x + x2 * (R)0.5 + x3 * (R)(1.0 / 6.0) + x2 * x2 * (R)(1.0 / 24.0) +
                     x3 * x2 * (R)(1.0 / 120.0) + x6 * (R)(1.0 / 720.0);
        const R rinv = (R)1 / r, ir2 = rinv * rinv, ir6 = ir2 * ir2 * ir2;
        const R tinv = (R)1 / t;
        const R f = H * a * ex - H * ex * ((R)1 + x) * q * tinv + ex * S6 * c * ir6;
        const R t6 = x6 * (R)(1.0 / 720.0) * ex;                     // -(d/dx)(e^{-x} S6)
        const R dfdr = -H * a * b * KB * ex + H * q * KB * (b * b * ex + ((R)1 + x) * ex * tinv * tinv) +
                       c * (-b * KB * t6 * ir6 - (R)6 * ex * S6 * ir6 * rinv);
        const R dfdb = -H * a * t * ex + H * q * b * t * ex - c * ir6 * t * t6;
        const R dfda = H * ex, dfdq = -H * ex * ((R)1 + x) * tinv, dfdc = ex * S6 * ir6;
        e = m * f;
        dedr = m * dfdr;
        dpi[0] = sqrt_chain(m * dfda, a, pj[0]);
        dpj[0] = sqrt_chain(m * dfda, a, pi[0]);
        dpi[1] = sqrt_chain(m * dfdb, b, pj[1]);
        dpj[1] = sqrt_chain(m * dfdb, b, pi[1]);
        dpi[2] = m * dfdq * pj[2];
        dpj[2] = m * dfdq * pi[2];
        dpi[3] = m * dfdc * pj[3];
        dpj[3] = m * dfdc * pi[3];
    } else {  // SR_SLATER_DAMP
        const R b = sqrt(pi[0] * pj[0]);
        const R c6 = pi[1] * pj[1], c8 = pi[2] * pj[2], c10 = pi[3] * pj[3];
        const R x = b * r;
        const R Dn = x * x + (R)3 * x + (R)3, Nm = (R)2 * x * x + (R)3 * x;
        const R y = x - Nm / Dn;
        const R dydx = (R)1 - (((R)4 * x + (R)3) * Dn - Nm * ((R)2 * x + (R)3)) / (Dn * Dn);
        const R ey = sr_exp<R>(-y);
        const R y2 = y * y, y3 = y2 * y, y4 = y2 * y2, y5 = y4 * y, y6 = y3 * y3, y7 = y6 * y, y8 = y4 * y4;
        const R y9 = y8 * y, y10 = y5 * y5;
        const R s6 = (R)1 + y + y2 * (R)0.5 + y3 * (R)(1.0 / 6.0) + y4 * (R)(1.0 / 24.0) + y5 * (R)(1.0 / 120.0) +
                     y6 * (R)(1.0 / 720.0);
        const R s8 = s6 + y7 * (R)(1.0 / 5040.0) + y8 * (R)(1.0 / 40320.0);
        const R s10 = s8 + y9 * (R)(1.0 / 362880.0) + y10 * (R)(1.0 / 3628800.0);
        const R f6 = ey * s6, f8 = ey * s8, f10 = ey * s10;
        // d f_n / dy = -e^{-y} y^n / n!
        const R g6 = -ey * y6 * (R)(1.0 / 720.0), g8 = -ey * y8 * (R)(1.0 / 40320.0),
                g10 = -ey * y10 * (R)(1.0 / 3628800.0);
        const R rinv = (R)1 / r, ir2 = rinv * rinv, ir6 = ir2 * ir2 * ir2, ir8 = ir6 * ir2, ir10 = ir8 * ir2;
        e = m * (f6 * c6 * ir6 + f8 * c8 * ir8 + f10 * c10 * ir10);
        const R dy = (g6 * c6 * ir6 + g8 * c8 * ir8 + g10 * c10 * ir10) * dydx;   // d/dx at fixed r
        dedr = m * (dy * b - ((R)6 * f6 * c6 * ir6 + (R)8 * f8 * c8 * ir8 + (R)10 * f10 * c10 * ir10) * rinv);
        const R dEdb = m * dy * r;
        dpi[0] = sqrt_chain(dEdb, b, pj[0]);
        dpj[0] = sqrt_chain(dEdb, b, pi[0]);
        dpi[1] = m * f6 * ir6 * pj[1];
        dpj[1] = m * f6 * ir6 * pi[1];
        dpi[2] = m * f8 * ir8 * pj[2];
        dpj[2] = m * f8 * ir8 * pi[2];
        dpi[3] = m * f10 * ir10 * pj[3];
        dpj[3] = m * f10 * ir10 * pi[3];
    }
}

}  // namespace admp
