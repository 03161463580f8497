// Learned banded proposal for the rest-frame momentum, and the network that
// shapes it.  (one translation unit: included by iss_hip.hip)
//
// Target.  For one hadron of species s emitted by cell c, the reference's
// rejection scheme (src/FSSW.cpp:1852-1966 with src/MomentumSamplerBase.cpp:
// 20-58) produces rest-frame momenta with density, in t = (E - mu)/T and the
// direction n,
//     pi(t, n) ~ p E f0(E) w1(p) w2(p),      t in [x, t_max], x = (m - mu)/T,
// w1 = clamp(p.dsigma / (E (|ds0| + |ds|)), 0, 1) and
// w2 = clamp((1 + df_shear + df_bulk) / 2, 0, 1); t_max is the top of the
// reference's |p| table, so the support is the reference's.
//
// Proposal.  t - x is cut into 9 bands: 8 bands with edges
// 2 (-log(1 - k/8)) and a far tail.  In band k the |p| part uses an
// exponential envelope exp(c_k + s_k (t - x)) of G(t) = log(p E f0 e^x): the
// tangent of G at the band midpoint when G is concave on the band (checked
// per band through an upper bound of G''), else the bound
// sup(p E) sup(e^t f0) e^-t.  Directions follow
//     q(n) = (1 + kappa_k n.pi.n) (1 + beta_k n.d) / (4 pi),
// sampled exactly as a mixture of the three n_i^2 laws in the eigenframe of
// the shear tensor (weights (1 + kappa lambda_i)/3, cos = cbrt(2u - 1))
// followed by an antipodal flip that keeps n with probability
// (1 + beta n.d)/2.  Per band, S1 >= w1 / (1 + beta y) and
// S2 >= w2 / (1 + kappa W) are verified suprema over the band's momenta and
// all directions (interval bounds of the delta-f terms, endpoint and
// clamp-crossing candidates of piecewise Moebius functions).  Band k is
// chosen with probability proportional to its envelope mass and a draw is
// accepted with probability pi / envelope <= 1, so the accepted momenta
// follow pi exactly.  Hadrons whose envelope cannot be formed (a
// Bose-Einstein species at its condensation point, non-finite bounds, or the
// reference's Fermi-Dirac table quirk in force) go to the plain sampler.
//
// Network.  A 14 -> 64 -> 64 -> 8 perceptron with tanh-GELU, BF16 weights and
// activations, FP32 accumulation on v_mfma_f32_32x32x16_bf16, maps cell state
// and species to kappa_k = sigmoid(z_k) kappa0_k (band k), where kappa0_k is
// the unclamped optimum at the band midpoint.  It is trained on the
// (cell, species) of hadrons accepted by the plain sampler, minimising the
// mean log envelope mass, i.e. maximising the acceptance of the exact
// sampler.  With zero weights every kappa is kappa0 / 2.
#pragma once

#include "kernels.h"
#include "sample_kernels.h"

namespace issdev {

// ============================================================== cell aux
// eigen-decomposition of the rest-frame shear tensor, the dsigma direction
// in its eigenframe and |dsigma|; [kAuxFields][Nc] floats
enum AuxField {
    A_L1 = 0, A_L2, A_L3,                       // eigenvalues, descending
    A_R00, A_R01, A_R02, A_R10, A_R11, A_R12,   // R[r][c]: component r of eigenvector c
    A_R20, A_R21, A_R22,
    A_DH0, A_DH1, A_DH2,                        // unit dsigma in the eigenframe
    A_DN,                                       // |dsigma| (spatial)
    kAuxFields
};

// cyclic Jacobi rotations of a symmetric 3x3 matrix; A becomes diagonal, V
// holds the eigenvectors as columns
__device__ inline void jacobi3(double A[3][3], double V[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) V[i][j] = (i == j) ? 1. : 0.;
    for (int sweep = 0; sweep < 16; sweep++) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-30 * dia || off == 0.) break;
        for (int pq = 0; pq < 3; pq++) {
            const int p = (pq == 2) ? 1 : 0, q = (pq == 0) ? 1 : 2;
            const double apq = A[p][q];
            if (apq == 0.) continue;
            const double theta = (A[q][q] - A[p][p]) / (2. * apq);
            const double t = (theta >= 0. ? 1. : -1.) / (fabs(theta) + sqrt(theta * theta + 1.));
            const double cs = 1. / sqrt(t * t + 1.), sn = t * cs;
            A[p][p] -= t * apq;
            A[q][q] += t * apq;
            A[p][q] = A[q][p] = 0.;
            const int r = 3 - p - q;
            const double arp = A[r][p], arq = A[r][q];
            A[r][p] = A[p][r] = cs * arp - sn * arq;
            A[r][q] = A[q][r] = sn * arp + cs * arq;
            for (int k = 0; k < 3; k++) {
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = cs * vkp - sn * vkq;
                V[k][q] = sn * vkp + cs * vkq;
            }
        }
    }
}

__global__ void k_cell_aux(const float* __restrict__ cells, int64_t Nc, float* __restrict__ aux) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= Nc) return;
    const double xx = cells[F_PIXX * Nc + c], xy = cells[F_PIXY * Nc + c], xz = cells[F_PIXZ * Nc + c];
    const double yy = cells[F_PIYY * Nc + c], yz = cells[F_PIYZ * Nc + c];
    double A[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, -xx - yy}};
    double V[3][3];
    jacobi3(A, V);
    double lam[3] = {A[0][0], A[1][1], A[2][2]};
    int o[3] = {0, 1, 2};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2 - i; j++)
            if (lam[o[j]] < lam[o[j + 1]]) { const int t = o[j]; o[j] = o[j + 1]; o[j + 1] = t; }
    const double d1 = cells[F_DA1 * Nc + c], d2 = cells[F_DA2 * Nc + c], d3 = cells[F_DA3 * Nc + c];
    const double dn = sqrt(d1 * d1 + d2 * d2 + d3 * d3);
    for (int i = 0; i < 3; i++) {
        aux[(A_L1 + i) * Nc + c] = static_cast<float>(lam[o[i]]);
        for (int r = 0; r < 3; r++) aux[(A_R00 + 3 * r + i) * Nc + c] = static_cast<float>(V[r][o[i]]);
        const double proj = d1 * V[0][o[i]] + d2 * V[1][o[i]] + d3 * V[2][o[i]];
        aux[(A_DH0 + i) * Nc + c] = static_cast<float>(dn > 0. ? proj / dn : 0.);
    }
    aux[A_DN * Nc + c] = static_cast<float>(dn);
}

// ======================================================= per-hadron params
struct LParams {
    float T, mu, m, sgn;      // sgn: -1 Bose-Einstein, +1 Fermi-Dirac, 0 Boltzmann
    float x, tmax;            // t = (E - mu)/T range of the reference's |p| table
    float d0, dn, D;          // rest-frame dsigma_0, |dsigma|, |dsigma_0| + |dsigma|
    float l1, l2, l3;         // shear eigenvalues [GeV/fm^3]
    float dh0, dh1, dh2;      // unit dsigma in the eigenframe
    float ps0;                // shear: df_s = omf p^2 W Ps(E), Ps = ps0 (smode 1) or ps0/E (smode 2)
    float qi, q0, q1, q2;     // bulk: df_b = omf (qi/E + q0 + q1 E + q2 E^2)
    int smode;
    int flags;                // 1: m/T - mu/T outside the table (hole); 2: plain-sampler fallback
};

__device__ __forceinline__ float ps_of(const LParams& P, float E) {
    return P.smode == 2 ? P.ps0 * __builtin_amdgcn_rcpf(E) : (P.smode == 1 ? P.ps0 : 0.f);
}
__device__ __forceinline__ float qb_of(const LParams& P, float E) {
    return P.qi * __builtin_amdgcn_rcpf(E) + P.q0 + E * (P.q1 + P.q2 * E);
}

// The delta-f forms of the plain sampler's trial loop, rewritten as
// polynomials in E: shear (src/FSSW.cpp:1897-1913), bulk (src/FSSW.cpp:
// 1916-1928, 1795-1849).
__device__ inline void lparams_of(const SampleArgs& a, const float* __restrict__ aux, int64_t c, int s,
                                  LParams& P) {
    const int64_t Nc = a.Nc;
    PState q;
    pstate_of(a, c, s, q);
    const int sign = a.sp.sign[s];
    const int tb = ((sign == -1) ? 0 : 3) + ((q.x < 30.) ? 0 : (q.x < 50. ? 1 : 2));
    const int i0 = static_cast<int>((q.x - a.pt.Et0[tb]) / a.pt.dEt[tb]);
    P.flags = (i0 < 0 || i0 >= a.pt.len[tb] - 1) ? 1 : 0;
    P.T = q.T;
    P.mu = static_cast<float>(q.mu);
    P.m = q.fm;
    P.sgn = static_cast<float>(sign);
    P.x = static_cast<float>(q.x);
    P.tmax = static_cast<float>(a.pt.Et0[tb] + (a.pt.len[tb] - 1) * a.pt.dEt[tb]);
    P.d0 = a.cells[F_DA0 * Nc + c];
    P.dn = aux[A_DN * Nc + c];
    P.D = fabsf(P.d0) + P.dn;
    P.l1 = aux[A_L1 * Nc + c]; P.l2 = aux[A_L2 * Nc + c]; P.l3 = aux[A_L3 * Nc + c];
    P.dh0 = aux[A_DH0 * Nc + c]; P.dh1 = aux[A_DH1 * Nc + c]; P.dh2 = aux[A_DH2 * Nc + c];
    const float T = q.T;
    const float Ed = a.cells[F_E * Nc + c], Pd = a.cells[F_P * Nc + c];
    float cf0 = 0.f, cf1 = 0.f, cf2 = 0.f, cf3 = 0.f, cf4 = 0.f, cf5 = 0.f;
    if (a.bulk != BULK_NONE || a.neos >= 0) {
        cf0 = a.coef[c]; cf1 = a.coef[Nc + c]; cf2 = a.coef[2 * Nc + c];
        cf3 = a.coef[3 * Nc + c]; cf4 = a.coef[4 * Nc + c]; cf5 = a.coef[5 * Nc + c];
    }
    P.smode = 0;
    P.ps0 = 0.f;
    if (a.shear) {
        if (a.neos == 1) { P.smode = 2; P.ps0 = 1.f / (2.f * cf2) / T; }
        else if (a.neos == 0) { P.smode = 1; P.ps0 = cf0; }
        else { P.smode = 1; P.ps0 = 1.0f / (2.0f * T * T * (Ed + Pd)); }
    }
    P.qi = P.q0 = P.q1 = P.q2 = 0.f;
    if (a.bulk != BULK_NONE) {
        float Pi = a.cells[F_PI * Nc + c];
        if (a.bulk == BULK_RTA) Pi = Pi / static_cast<float>(kHbarC);
        const float m = P.m;
        const float Bq = a.sp.B[s], Sq = a.sp.S[s], Qq = a.sp.Q[s];
        if (a.bulk == BULK_RTA || a.bulk == BULK_CE) {
            P.qi = -cf0 * Pi * m * m / (3.f * T);
            P.q1 = cf0 * cf1 * Pi / T;
        } else if (a.bulk == BULK_14MOM) {
            P.q0 = Pi * cf0 * m * m;
            P.q1 = Pi * cf1 * Bq;
            P.q2 = Pi * cf2;
        } else if (a.bulk == BULK_22MOM) {
            P.q0 = Pi * m * m * cf2;
            P.q1 = Pi * (Bq * cf3 + Sq * cf4 + Qq * cf5);
            P.q2 = Pi * (cf1 - cf2);
        }
    }
    // the envelope does not bound the diffusion term: those hadrons take the
    // plain sampler
    if ((sign < 0 && q.x < 1e-3) || (a.pt.quirk && tb >= 3) || a.diff) P.flags |= 2;
}

// The same parameters for the trial kernel, in FP32, given x (from the
// set-up record): mu' = min(m, B muB + S muS + Q muQ) with the float sum of
// pstate_of.
__device__ __forceinline__ void lparams_fast(const SampleArgs& a, const float* __restrict__ aux, int64_t c, int s, float x,
                                    LParams& P) {
    const int64_t Nc = a.Nc;
    const float T = a.cells[F_T * Nc + c];
    const float muB = a.cells[F_MUB * Nc + c], muS = a.cells[F_MUS * Nc + c], muQ = a.cells[F_MUQ * Nc + c];
    const float Bq = a.sp.B[s], Sq = a.sp.S[s], Qq = a.sp.Q[s];
    const float m = static_cast<float>(a.sp.mass[s]);
    const int sign = a.sp.sign[s];
    P.T = T;
    P.mu = fminf(m, Bq * muB + Sq * muS + Qq * muQ);
    P.m = m;
    P.sgn = static_cast<float>(sign);
    P.x = x;
    const int tb = ((sign == -1) ? 0 : 3) + ((x < 30.f) ? 0 : (x < 50.f ? 1 : 2));
    P.tmax = static_cast<float>(a.pt.Et0[tb] + (a.pt.len[tb] - 1) * a.pt.dEt[tb]);
    P.flags = 0;
    P.d0 = a.cells[F_DA0 * Nc + c];
    P.dn = aux[A_DN * Nc + c];
    P.D = fabsf(P.d0) + P.dn;
    P.l1 = aux[A_L1 * Nc + c]; P.l2 = aux[A_L2 * Nc + c]; P.l3 = aux[A_L3 * Nc + c];
    P.dh0 = aux[A_DH0 * Nc + c]; P.dh1 = aux[A_DH1 * Nc + c]; P.dh2 = aux[A_DH2 * Nc + c];
    const float Ed = a.cells[F_E * Nc + c], Pd = a.cells[F_P * Nc + c];
    float cf0 = 0.f, cf1 = 0.f, cf2 = 0.f, cf3 = 0.f, cf4 = 0.f, cf5 = 0.f;
    if (a.bulk != BULK_NONE || a.neos >= 0) {
        cf0 = a.coef[c]; cf1 = a.coef[Nc + c]; cf2 = a.coef[2 * Nc + c];
        cf3 = a.coef[3 * Nc + c]; cf4 = a.coef[4 * Nc + c]; cf5 = a.coef[5 * Nc + c];
    }
    P.smode = 0;
    P.ps0 = 0.f;
    if (a.shear) {
        if (a.neos == 1) { P.smode = 2; P.ps0 = 1.f / (2.f * cf2) / T; }
        else if (a.neos == 0) { P.smode = 1; P.ps0 = cf0; }
        else { P.smode = 1; P.ps0 = 1.0f / (2.0f * T * T * (Ed + Pd)); }
    }
    P.qi = P.q0 = P.q1 = P.q2 = 0.f;
    if (a.bulk != BULK_NONE) {
        float Pi = a.cells[F_PI * Nc + c];
        if (a.bulk == BULK_RTA) Pi = Pi / static_cast<float>(kHbarC);
        if (a.bulk == BULK_RTA || a.bulk == BULK_CE) {
            P.qi = -cf0 * Pi * m * m / (3.f * T);
            P.q1 = cf0 * cf1 * Pi / T;
        } else if (a.bulk == BULK_14MOM) {
            P.q0 = Pi * cf0 * m * m;
            P.q1 = Pi * cf1 * Bq;
            P.q2 = Pi * cf2;
        } else if (a.bulk == BULK_22MOM) {
            P.q0 = Pi * m * m * cf2;
            P.q1 = Pi * (Bq * cf3 + Sq * cf4 + Qq * cf5);
            P.q2 = Pi * (cf1 - cf2);
        }
    }
}

// 14 network inputs (two zero pads): T, muB, the five rest-frame shear
// components and the bulk pressure relative to e + P, mass, degeneracy,
// baryon number, statistics, and the shear and bulk delta-f at the thermal
// energy E = m + 2T
constexpr int kIn = 16, kHid = 64, kOut = 8;
__device__ inline void features_of(const SampleArgs& a, int64_t c, int s, const LParams& P, float f[kIn]) {
    const int64_t Nc = a.Nc;
    const float epp = a.cells[F_E * Nc + c] + a.cells[F_P * Nc + c];
    const float ie = epp > 0.f ? 5.f / epp : 0.f;
    f[0] = (P.T / 0.15f - 1.f) * 10.f;
    f[1] = a.cells[F_MUB * Nc + c] / 0.3f;
    f[2] = a.cells[F_PIXX * Nc + c] * ie;
    f[3] = a.cells[F_PIXY * Nc + c] * ie;
    f[4] = a.cells[F_PIXZ * Nc + c] * ie;
    f[5] = a.cells[F_PIYY * Nc + c] * ie;
    f[6] = a.cells[F_PIYZ * Nc + c] * ie;
    f[7] = a.cells[F_PI * Nc + c] * ie;
    f[8] = P.m;
    f[9] = static_cast<float>(a.sp.g[s]) * 0.25f;
    f[10] = static_cast<float>(a.sp.B[s]);
    f[11] = P.sgn;
    const float Eth = P.m + 2.f * P.T;
    f[12] = (Eth * Eth - P.m * P.m) * ps_of(P, Eth) * (P.l1 - P.l3);
    f[13] = qb_of(P, Eth);
    f[14] = 0.f;
    f[15] = 0.f;
}

// ============================================================ BF16 MFMA
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even: the hardware conversion
// (v_cvt_pk_bf16_f32), one instruction per pair
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f)); }
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
// two values as one packed word (a in the low half)
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    const f32x2_t v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ f32x16 mfma32(uint4 a, uint4 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                     c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 zero16() {
    f32x16 z;
#pragma unroll
    for (int i = 0; i < 16; i++) z[i] = 0.f;
    return z;
}

// Tiles are 32 x 32 in the layout of v_mfma_f32_32x32x16_bf16: operand A
// (rows m0.., 16 k per step) comes from lane l as row l%32, k = 8 (l/32)..+7;
// operand B^T likewise with rows n0..; the accumulator register i of lane l
// is C[m0 + 8 (i/4) + 4 (l/32) + i%4][n0 + l%32].
//
// A from pre-packed fragments in global memory (16 bytes per lane and step),
// B^T from LDS rows (bf16, row stride ldb elements)
__device__ __forceinline__ f32x16 tile_gl(const uint4* __restrict__ afrag, const uint16_t* Bt, int ldb, int n0,
                                          int nk, f32x16 acc) {
    const int l = threadIdx.x & 63, r = l & 31, g = l >> 5;
    for (int kt = 0; kt < nk; kt++) {
        const uint4 av = afrag[kt * 64 + l];
        const uint4 bv = *reinterpret_cast<const uint4*>(Bt + (n0 + r) * ldb + kt * 16 + 8 * g);
        acc = mfma32(av, bv, acc);
    }
    return acc;
}
// both operands from LDS; rows past mvalid / nvalid read as zero
__device__ __forceinline__ f32x16 tile_ll(const uint16_t* A, int lda, int m0, int mvalid, const uint16_t* Bt,
                                          int ldb, int n0, int nvalid, int nk, f32x16 acc) {
    const int l = threadIdx.x & 63, r = l & 31, g = l >> 5;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (int kt = 0; kt < nk; kt++) {
        const uint4 av = (m0 + r < mvalid) ? *reinterpret_cast<const uint4*>(A + (m0 + r) * lda + kt * 16 + 8 * g) : z;
        const uint4 bv = (n0 + r < nvalid) ? *reinterpret_cast<const uint4*>(Bt + (n0 + r) * ldb + kt * 16 + 8 * g) : z;
        acc = mfma32(av, bv, acc);
    }
    return acc;
}

// Hardware reciprocal, exp2 and log2 (1-2 ulp).  Every bound they enter
// carries a relative margin of 1e-4, far above their error.
// (fexp, the exp2-based exponential, comes with the yield kernels)
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float flog(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
// log1p(u (e^w - 1)): s (tau - a) of a draw from exp(s tau) on a band [a, b]
// of width w / s (inverse CDF at u), with hardware exp2 / log2 and Taylor
// series where either argument is below 1/8 (truncation below 1e-8 relative)
__device__ __forceinline__ float trunc_exp_inv(float u, float w) {
    const float em1 = fabsf(w) < 0.125f
                          ? w * (1.f + w * (0.5f + w * (1.f / 6.f + w * (1.f / 24.f + w * (1.f / 120.f + w * (1.f / 720.f))))))
                          : __builtin_amdgcn_exp2f(w * 1.4426950408889634f) - 1.f;
    const float y = u * em1;
    return fabsf(y) < 0.125f
               ? y * (1.f - y * (0.5f - y * (1.f / 3.f - y * (0.25f - y * (0.2f - y * (1.f / 6.f - y * (1.f / 7.f - y * 0.125f)))))))
               : __builtin_amdgcn_logf(1.f + y) * 0.6931471805599453f;
}

// cube root with sign, through exp2(log2(|x|) / 3)
__device__ __forceinline__ float fcbrt(float x) {
    const float r = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fabsf(x)) * (1.f / 3.f));
    return x < 0.f ? -r : r;
}

// tanh-form GELU, 0.5 x (1 + tanh(u)) = x sigmoid(2u), and its derivative
// with u = sqrt(2/pi) (x + 0.044715 x^3): sigmoid(2u) = 1 / (1 + 2^(x (A + B x^2))),
// A = -2 log2(e) sqrt(2/pi), B = 0.044715 A (constants folded: three VALU
// operations and two transcendentals)
__device__ __forceinline__ float gelu_sig(float x) {
    constexpr float A = -2.8853900817779268f * 0.7978845608028654f, B = A * 0.044715f;
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * fmaf(B, x * x, A)));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sig(x); }
// the logistic function with the hardware exp2 and reciprocal (1 ulp each)
__device__ __forceinline__ float sigmoid_fast(float z) {
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float gelu_d(float x) {
    const float sg = gelu_sig(x);
    return sg + 2.f * x * sg * (1.f - sg) * 0.7978845608028654f * (1.f + 3.f * 0.044715f * x * x);
}

// parameters, FP32 master layout (row-major [out][in])
constexpr int P_W1 = 0;                       // [64][16]
constexpr int P_B1 = P_W1 + kHid * kIn;       // [64]
constexpr int P_W2 = P_B1 + kHid;             // [64][64]
constexpr int P_B2 = P_W2 + kHid * kHid;      // [64]
constexpr int P_W3 = P_B2 + kHid;             // [8][64]
constexpr int P_B3 = P_W3 + kOut * kHid;      // [8]
constexpr int kNumParams = P_B3 + kOut;       // 5768

// packed A fragments: [mt][kt][64 lanes] x 16 bytes
constexpr int F_W1 = 0;                       // W1   (64 x 16): 2 x 1
constexpr int F_W2 = F_W1 + 2 * 1 * 64;       // W2   (64 x 64): 2 x 4
constexpr int F_W3 = F_W2 + 2 * 4 * 64;       // W3   (8->32 x 64): 1 x 4
constexpr int F_W3T = F_W3 + 1 * 4 * 64;      // W3^T (64 x 8->16): 2 x 1
constexpr int F_W2T = F_W3T + 2 * 1 * 64;     // W2^T (64 x 64): 2 x 4
constexpr int kNumFrags = F_W2T + 2 * 4 * 64;

struct MlpDev {
    const uint4* frag;   // [kNumFrags]
    const float* par;    // FP32 master parameters (biases read from here)
};

// one thread per 16-byte fragment
__global__ void k_mlp_pack(const float* __restrict__ par, uint4* __restrict__ frag) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= kNumFrags) return;
    int base, M, K, nkt, src, trans;
    if (f < F_W2) { base = F_W1; M = kHid; K = kIn; nkt = 1; src = P_W1; trans = 0; }
    else if (f < F_W3) { base = F_W2; M = kHid; K = kHid; nkt = 4; src = P_W2; trans = 0; }
    else if (f < F_W3T) { base = F_W3; M = kOut; K = kHid; nkt = 4; src = P_W3; trans = 0; }
    else if (f < F_W2T) { base = F_W3T; M = kHid; K = kOut; nkt = 1; src = P_W3; trans = 1; }
    else { base = F_W2T; M = kHid; K = kHid; nkt = 4; src = P_W2; trans = 1; }
    const int idx = f - base;
    const int lane = idx & 63, kt = (idx >> 6) % nkt, mt = (idx >> 6) / nkt;
    const int row = mt * 32 + (lane & 31);
    uint16_t v[8];
    for (int j = 0; j < 8; j++) {
        const int k = kt * 16 + 8 * (lane >> 5) + j;
        float w = 0.f;
        if (row < M && k < K) {
            // A = W: A[row][k] = W[row][k]; A = W^T: A[row][k] = W[k][row]
            w = trans ? par[src + k * M + row] : par[src + row * K + k];
        }
        v[j] = f2bf(w);
    }
    frag[f] = make_uint4(v[0] | (uint32_t(v[1]) << 16), v[2] | (uint32_t(v[3]) << 16),
                         v[4] | (uint32_t(v[5]) << 16), v[6] | (uint32_t(v[7]) << 16));
}

// LDS strides (elements)
constexpr int kXld = 24, kHld = 72, kZld = 68;

// store a 32x32 accumulator tile, after bias and GELU, into a sample-major
// bf16 array H[sample][hid] (rows n0.., columns m0..); bv: the lane's 16
// biases of the tile's rows m0 + 8 q + 4 g + e (bias_tile)
__device__ __forceinline__ void store_act_s(const f32x16& acc, const float4 (&bv)[4], uint16_t* H, int m0, int n0) {
    const int l = threadIdx.x & 63, r = l & 31, g = l >> 5;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int h0 = m0 + 8 * q + 4 * g;
        const float v0 = gelu(acc[4 * q + 0] + bv[q].x), v1 = gelu(acc[4 * q + 1] + bv[q].y);
        const float v2 = gelu(acc[4 * q + 2] + bv[q].z), v3 = gelu(acc[4 * q + 3] + bv[q].w);
        *reinterpret_cast<uint2*>(H + (n0 + r) * kHld + h0) = make_uint2(pack_bf16(v0, v1), pack_bf16(v2, v3));
    }
}

// the lane's biases for store_act_s of the tile with rows m0.. (16-byte loads,
// issued before the tile's MFMAs so that their latency overlaps them)
__device__ __forceinline__ void bias_tile(const float* __restrict__ bias, int m0, float4 (&bv)[4]) {
    const int g = (threadIdx.x & 63) >> 5;
#pragma unroll
    for (int q = 0; q < 4; q++) bv[q] = *reinterpret_cast<const float4*>(bias + m0 + 8 * q + 4 * g);
}

// Forward pass of one wave's 64 samples: X [64][kXld] (bf16) -> O [64][8]
// (FP32, aliasing X).  H [64][kHld] holds the hidden activations.
__device__ void mlp_forward_wave(const MlpDev& W, uint16_t* Xs, uint16_t* Hs, float* Os) {
    const int l = threadIdx.x & 63, r = l & 31, g = l >> 5;
    const float* b1 = W.par + P_B1;
    const float* b2 = W.par + P_B2;
    const float* b3 = W.par + P_B3;
    float4 bv0[4], bv1[4];
    bias_tile(b1, 0, bv0);
    bias_tile(b1, 32, bv1);
    wave_lds_sync();
    for (int nt = 0; nt < 2; nt++) {
        const f32x16 a0 = tile_gl(W.frag + F_W1 + 0 * 64, Xs, kXld, nt * 32, 1, zero16());
        store_act_s(a0, bv0, Hs, 0, nt * 32);
        const f32x16 a1 = tile_gl(W.frag + F_W1 + 1 * 64, Xs, kXld, nt * 32, 1, zero16());
        store_act_s(a1, bv1, Hs, 32, nt * 32);
    }
    bias_tile(b2, 0, bv0);
    bias_tile(b2, 32, bv1);
    wave_lds_sync();
    for (int nt = 0; nt < 2; nt++) {
        const f32x16 a0 = tile_gl(W.frag + F_W2 + 0 * 4 * 64, Hs, kHld, nt * 32, 4, zero16());
        const f32x16 a1 = tile_gl(W.frag + F_W2 + 1 * 4 * 64, Hs, kHld, nt * 32, 4, zero16());
        wave_lds_sync();
        store_act_s(a0, bv0, Hs, 0, nt * 32);
        store_act_s(a1, bv1, Hs, 32, nt * 32);
        wave_lds_sync();
    }
    for (int nt = 0; nt < 2; nt++) {
        const f32x16 acc = tile_gl(W.frag + F_W3, Hs, kHld, nt * 32, 4, zero16());
#pragma unroll
        for (int e = 0; e < 4; e++) Os[(nt * 32 + r) * kOut + 4 * g + e] = acc[e] + b3[4 * g + e];
    }
    wave_lds_sync();
}

// ========================================================== band envelope
constexpr int kBands = 9;
constexpr float kLogMargin = 1e-4f;
// band edges in t - x: 2 (-log(1 - k/8)), k = 0..7, then +4 for the far tail
__device__ __forceinline__ float band_edge(int k) {
    switch (k) {
        case 0: return 0.f;
        case 1: return 0.26706278524904526f;
        case 2: return 0.57536414490356180f;
        case 3: return 0.94000725849147116f;
        case 4: return 1.38629436111989062f;
        case 5: return 1.96165850602344528f;
        case 6: return 2.77258872223978124f;
        case 7: return 4.15888308335967186f;
        default: return 8.15888308335967186f;
    }
}

// One band of the envelope, 8 bytes: log envelope c + s (t - x) (fp16,
// rounded so that the stored line lies above the exact one on the band),
// kappa (fp16) and beta (u8, beta = b / 255), both quantised before the
// bounds are computed.
struct BandTable {
    uint16_t Q[8];        // band k is chosen when Q[k-1] <= word16 < Q[k] (Q[-1] = 0, 16-bit words)
    uint2 band[kBands];   // {c | s << 16, kappa | beta << 16}
    int klast;            // last band of positive mass (chosen for word >= Q[klast - 1])
};

__device__ __forceinline__ uint16_t h_bits(_Float16 h) { return __builtin_bit_cast(uint16_t, h); }
__device__ __forceinline__ float h2f(uint16_t b) { return static_cast<float>(__builtin_bit_cast(_Float16, b)); }
__device__ __forceinline__ uint16_t f2h(float x) { return h_bits(static_cast<_Float16>(x)); }
// smallest fp16 >= x (x finite and within the fp16 range)
__device__ __forceinline__ uint16_t f2h_up(float x) {
    uint16_t b = f2h(x);
    if (h2f(b) < x) {
        if (b == 0x8000u) b = 0x0001u;
        else if (b & 0x8000u) b = static_cast<uint16_t>(b - 1u);
        else b = static_cast<uint16_t>(b + 1u);
    }
    return b;
}
__device__ __forceinline__ void band_unpack(uint2 v, float& c, float& s, float& kap, float& beta) {
    c = h2f(static_cast<uint16_t>(v.x & 0xFFFFu));
    s = h2f(static_cast<uint16_t>(v.x >> 16));
    kap = h2f(static_cast<uint16_t>(v.y & 0xFFFFu));
    beta = static_cast<float>((v.y >> 16) & 0xFFu) * (1.f / 255.f);
}

__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.f), 1.f); }

// w2 / (1 + kappa W) <= S2 over W in [l3, l1]: the numerator's bound is
// linear on W <= 0 (slope cmin) and W >= 0 (slope cmax), then clamped, so
// the supremum is at an end, at 0 or where the clamp starts.
__device__ __forceinline__ void s2_try(float W, float Amax, float cmin, float cmax, float kap, float& S2,
                                       float& Wst, float& Nst) {
    const float num = clamp01(0.5f * (Amax + (W >= 0.f ? cmax * W : cmin * W)));
    const float v = fdiv(num, 1.f + kap * W);
    if (v > S2) { S2 = v; Wst = W; Nst = num; }
}

// Envelope of one hadron.  rho[k] in (0, 1) scales the band-midpoint
// optimum kappa0_k.  Returns false when no envelope can be formed.  With
// GRAD also returns d log(M) / d z_k for rho = sigmoid(z) (M = sum of the
// band masses: the expected proposals per accepted hadron times the
// target's mass).
template <bool GRAD>
__device__ bool band_setup(const LParams& P, const float rho[kOut], BandTable& Bt, float* logM, float* dz) {
    const float T = P.T, mu = P.mu, m = P.m, sg = P.sgn, x = P.x;
    const float tau_max = P.tmax - x;
    const float rD = P.D > 0.f ? __builtin_amdgcn_rcpf(P.D) : 0.f;
    const float rd0 = P.d0 > 0.f ? __builtin_amdgcn_rcpf(P.d0) : 0.f;
    const bool shear = P.smode != 0 && P.l1 > 0.f;
    const float klo = shear ? fdiv(-0.98f, P.l1) : 0.f, khi = (shear && P.l3 < 0.f) ? fdiv(0.98f, -P.l3) : 0.f;
    // quantities at the band ends: edge j < 9 is band_edge(j), edge 9 is the table top
    float te[kBands + 1], Ee[kBands + 1], pe[kBands + 1], ve[kBands + 1], ee[kBands + 1], ome[kBands + 1];
    float qe[kBands + 1], pse[kBands + 1];
#pragma unroll
    for (int j = 0; j <= kBands; j++) {
        const float tau = (j < kBands) ? fminf(band_edge(j), tau_max) : tau_max;
        te[j] = tau;
        const float E = T * (x + tau) + mu;
        const float rE = __builtin_amdgcn_rcpf(E);
        Ee[j] = E;
        pe[j] = __builtin_amdgcn_sqrtf(fmaxf(E * E - m * m, 0.f));
        ve[j] = E > 0.f ? pe[j] * rE : 0.f;
        ee[j] = fexp(-(x + tau));
        ome[j] = __builtin_amdgcn_rcpf(1.f + sg * ee[j]);   // 1 - sign f0 = e^t f0
        qe[j] = P.qi * rE + P.q0 + E * (P.q1 + P.q2 * E);
        pse[j] = P.smode == 2 ? P.ps0 * rE : (P.smode == 1 ? P.ps0 : 0.f);
    }
    float mk[kBands];
    float dS2[kOut], Mt_S1[kOut];
    bool ok = tau_max > 0.f;
#pragma unroll
    for (int k = 0; k < kBands; k++) {
        mk[k] = 0.f;
        Bt.band[k] = make_uint2(0u, 0u);
        if (k < kOut) { dS2[k] = 0.f; Mt_S1[k] = 0.f; }
        const int jb = k + 1;
        const float a = te[k], b = te[jb];
        const float w = b - a;
        const float Elo = Ee[k], Ehi = Ee[jb], plo = pe[k], phi = pe[jb], vlo = ve[k], vhi = ve[jb];
        if (!(w > 0.f) || !(phi > 0.f) || !(Ehi > 0.f)) continue;
        // w1 / (1 + beta y) <= S1
        float beta = 0.f;
        if (P.dn > 0.f) beta = P.d0 > 0.f ? fminf(0.98f, vlo * P.dn * rd0) : 0.98f;
        const uint32_t beta_b = static_cast<uint32_t>(beta * 255.f);
        beta = static_cast<float>(beta_b) * (1.f / 255.f);
        const float S1 = fmaxf(clamp01((P.d0 + vhi * P.dn) * rD) * __builtin_amdgcn_rcpf(1.f + beta),
                               clamp01((P.d0 - vlo * P.dn) * rD) * __builtin_amdgcn_rcpf(1.f - beta));
        const float om_lo = fminf(ome[k], ome[jb]), om_hi = fmaxf(ome[k], ome[jb]);
        // bulk: sup of omf Q(E) over the band
        float Qmax = fmaxf(qe[k], qe[jb]);
        if (P.qi == 0.f && P.q2 != 0.f) {
            const float Es = -P.q1 * __builtin_amdgcn_rcpf(2.f * P.q2);
            if (Es > Elo && Es < Ehi) Qmax = fmaxf(Qmax, qb_of(P, Es));
        } else if (P.qi != 0.f && P.q2 == 0.f && P.q1 != 0.f && P.qi * P.q1 > 0.f) {
            const float Es = __builtin_amdgcn_sqrtf(P.qi * __builtin_amdgcn_rcpf(P.q1));
            if (Es > Elo && Es < Ehi) Qmax = fmaxf(Qmax, qb_of(P, Es));
        }
        const float Amax = 1.f + (Qmax >= 0.f ? om_hi * Qmax : om_lo * Qmax);
        // shear: c(E) = omf p^2 Ps(E) in [cmin, cmax]
        float cmin = 0.f, cmax = 0.f;
        if (shear) {
            const float psmin = fminf(pse[k], pse[jb]), psmax = fmaxf(pse[k], pse[jb]);
            if (psmin >= 0.f) { cmin = om_lo * plo * plo * psmin; cmax = om_hi * phi * phi * psmax; }
            else if (psmax <= 0.f) { cmin = om_hi * phi * phi * psmin; cmax = om_lo * plo * plo * psmax; }
            else { cmin = om_hi * phi * phi * psmin; cmax = om_hi * phi * phi * psmax; }
        }
        // the band's interior point: midpoint (tail: a + 1) for kappa0 and the tangent
        const float tau_s = (k < kBands - 1) ? 0.5f * (a + b) : fminf(a + 1.f, b);
        const float ts = x + tau_s;
        const float Es = T * ts + mu;
        const float ps2 = Es * Es - m * m;
        const float es = fexp(-ts);
        const float oms = __builtin_amdgcn_rcpf(1.f + sg * es);
        const float rEs = __builtin_amdgcn_rcpf(Es);
        // kappa: the network's share of the band-midpoint optimum
        float kap = 0.f, kap0 = 0.f;
        bool clipped = true;
        if (k < kOut && shear && P.l3 < 0.f) {
            const float cm = oms * fmaxf(ps2, 0.f) * (P.smode == 2 ? P.ps0 * rEs : P.ps0);
            const float am = 1.f + oms * (P.qi * rEs + P.q0 + Es * (P.q1 + P.q2 * Es));
            kap0 = am > 0.f ? cm * __builtin_amdgcn_rcpf(am) : 0.f;
            const float raw = rho[k] * kap0;
            kap = h2f(f2h(fminf(fmaxf(raw, klo), khi)));
            if (1.f + kap * P.l1 < 0.01f || 1.f + kap * P.l3 < 0.01f) kap = 0.f;
            clipped = !(raw > klo && raw < khi);
        }
        float S2 = 0.f, Wst = 0.f, Nst = 0.f;
        s2_try(0.f, Amax, cmin, cmax, kap, S2, Wst, Nst);
        if (shear) {
            s2_try(P.l3, Amax, cmin, cmax, kap, S2, Wst, Nst);
            s2_try(P.l1, Amax, cmin, cmax, kap, S2, Wst, Nst);
            if (cmax != 0.f) {
                const float wc = (2.f - Amax) * __builtin_amdgcn_rcpf(cmax);
                if (wc > 0.f && wc < P.l1) s2_try(wc, Amax, cmin, cmax, kap, S2, Wst, Nst);
            }
            if (cmin != 0.f) {
                const float wc = (2.f - Amax) * __builtin_amdgcn_rcpf(cmin);
                if (wc < 0.f && wc > P.l3) s2_try(wc, Amax, cmin, cmax, kap, S2, Wst, Nst);
            }
        }
        if (!(S1 * S2 > 0.f)) continue;
        // |p| envelope exp(cc + ss tau): the tangent of G at tau_s when G is
        // concave on the band (G'' <= U <= 0), else sup(p E) sup(e^t f0) e^-tau
        const float rphi2 = __builtin_amdgcn_rcpf(phi * phi), rEhi = __builtin_amdgcn_rcpf(Ehi);
        const float ea = ee[k];
        const float U = -T * T * (Ehi * Ehi + m * m) * rphi2 * rphi2 - T * T * rEhi * rEhi +
                        (sg < 0.f ? ea * __builtin_amdgcn_rcpf((1.f - ea) * (1.f - ea)) : 0.f);
        float cc, ss;
        if (U <= 0.f && ps2 > 0.f) {
            const float Gs = flog(__builtin_amdgcn_sqrtf(ps2) * Es * oms) - tau_s;
            ss = T * Es * __builtin_amdgcn_rcpf(ps2) + T * rEs - 1.f + sg * es * oms;
            cc = Gs - ss * tau_s;
        } else {
            ss = -1.f;
            cc = flog(phi * Ehi * om_hi);
        }
        if (!isfinite(cc) || !isfinite(ss)) { ok = false; continue; }
        // stored line: fp16 slope, intercept raised to stay above the exact
        // line at both band ends, then rounded up
        const float c0 = cc + flog(S1 * S2) + kLogMargin;
        const uint16_t sb = f2h(ss);
        const float s16 = h2f(sb);
        const float c1 = c0 + fmaxf(0.f, fmaxf((ss - s16) * a, (ss - s16) * b));
        if (!(fabsf(c1) < 6.0e4f) || !(fabsf(ss) < 6.0e4f)) { ok = false; continue; }
        const uint16_t cb = f2h_up(c1);
        const float c16 = h2f(cb);
        const float sw16 = s16 * w;
        // integral of exp(c16 + s16 tau) over [a, b]
        const float em1 = fabsf(sw16) > 1e-2f ? fexp(sw16) - 1.f : sw16 * (1.f + sw16 * (0.5f + sw16 * (1.f / 6.f)));
        mk[k] = fexp(c16 + s16 * a) * (fabsf(sw16) > 1e-6f ? em1 * __builtin_amdgcn_rcpf(s16) : w);
        Bt.band[k] = make_uint2(static_cast<uint32_t>(cb) | (static_cast<uint32_t>(sb) << 16),
                                static_cast<uint32_t>(f2h(kap)) | (beta_b << 16));
        if (GRAD && k < kOut) {
            const float sw = ss * w;
            const float e1 = fabsf(sw) > 1e-2f ? fexp(sw) - 1.f : sw * (1.f + sw * (0.5f + sw * (1.f / 6.f)));
            const float Mt = fexp(cc + ss * a) * (fabsf(sw) > 1e-6f ? e1 * __builtin_amdgcn_rcpf(ss) : w);
            Mt_S1[k] = Mt * S1 * fexp(kLogMargin);
            dS2[k] = (clipped || !shear) ? 0.f : -fdiv(Nst * Wst, (1.f + kap * Wst) * (1.f + kap * Wst)) * kap0;
        }
    }
    float M = 0.f;
#pragma unroll
    for (int k = 0; k < kBands; k++) M += mk[k];
    if (!ok || !(M > 0.f) || !isfinite(M)) {
        if (logM) *logM = 0.f;
        if (GRAD)
            for (int k = 0; k < kOut; k++) dz[k] = 0.f;
        return false;
    }
    if (logM) *logM = flog(M);
    if (GRAD) {
#pragma unroll
        for (int k = 0; k < kOut; k++) dz[k] = Mt_S1[k] * dS2[k] / M * rho[k] * (1.f - rho[k]);
    }
    // selection thresholds (16-bit words): every band of positive mass keeps
    // a non-empty range and the last one reaches 2^16
    int klast = 0;
#pragma unroll
    for (int k = 0; k < kBands; k++)
        if (mk[k] > 0.f) klast = k;
    constexpr uint32_t two16 = 1u << 16;
    uint32_t Qk[kBands];
    uint32_t Qp = 0;
    float cum = 0.f;
    const float iM = 1.f / M;
#pragma unroll
    for (int k = 0; k < kBands; k++) {
        cum += mk[k];
        uint32_t q;
        if (k >= klast) {
            q = two16;
        } else {
            q = static_cast<uint32_t>(fminf(cum * iM, 1.f) * 65536.f);
            if (mk[k] > 0.f) q = q > Qp + 1 ? q : Qp + 1;
            else q = Qp;
            int after = 0;
#pragma unroll
            for (int j = k + 1; j < kBands; j++) after += (j <= klast && mk[j] > 0.f) ? 1 : 0;
            const uint32_t cap = two16 - static_cast<uint32_t>(after);
            if (q > cap) q = cap;
        }
        Qk[k] = q;
        Qp = q;
    }
    Bt.klast = klast;
    // the chosen probabilities are (Q_k - Q_{k-1}) / 2^16; rescale each
    // band's envelope by g_k = m_k / (c_sel_k R), R = max_j m_j / c_sel_j
    float R = 0.f;
    Qp = 0;
#pragma unroll
    for (int k = 0; k < kBands; k++) {
        const float cs = static_cast<float>(Qk[k] - Qp) * (1.f / 65536.f);
        if (mk[k] > 0.f && cs > 0.f) R = fmaxf(R, fdiv(mk[k], cs));
        Qp = Qk[k];
    }
    Qp = 0;
#pragma unroll
    for (int k = 0; k < kBands; k++) {
        const float cs = static_cast<float>(Qk[k] - Qp) * (1.f / 65536.f);
        if (mk[k] > 0.f && cs > 0.f) {
            const float c16 = h2f(static_cast<uint16_t>(Bt.band[k].x & 0xFFFFu));
            const float cg = c16 - flog(fdiv(mk[k], cs * R));
            Bt.band[k].x = (Bt.band[k].x & 0xFFFF0000u) | f2h_up(fmaxf(cg, c16));
        }
        Qp = Qk[k];
        if (k < kBands - 1) Bt.Q[k] = static_cast<uint16_t>(Qk[k] < 0xFFFFu ? Qk[k] : 0xFFFFu);
    }
    return true;
}

// ======================================================== learned sampler
struct LSlot;
struct LearnArgs {
    SampleArgs a;
    const float* aux;
    MlpDev W;
    int64_t* fb_list;                 // slots handed to the plain sampler
    unsigned long long* fb_count;
    unsigned long long* stats2;       // [0] fallback hadrons, [1] envelope violations
    // (cell, species) pairs of the pass ([*P] compact list; the count lives on
    // the device) and each slot's cell, counter and pair
    const uint2* pkey;
    const int64_t* P;
    const LSlot* slot;
    // with routing, the number of slots sampled by the plain sampler (the
    // learned slot list holds the others); null: every slot is listed
    const int64_t* nsingle;
};

// Per-hadron record of the learned sampler (96 bytes): the band table with
// 16-bit selection thresholds, fp16 envelope lines and kappas, u8 betas.
struct alignas(16) LRec {
    uint32_t cell;
    uint32_t j;
    uint16_t species;
    uint8_t flags;        // 1: hole (table range), 2: plain-sampler fallback
    uint8_t klast;
    uint16_t Q[8];
    uint32_t cs[kBands];  // c | s << 16 (fp16)
    uint16_t kap[8];      // fp16 (the far tail has kappa = 0)
    uint8_t beta[kBands];
    uint8_t pad[3];
    float x;              // (m - mu) / T, the lower end of t
};
static_assert(sizeof(LRec) == 96, "LRec layout");

// Per-slot record of the learned sampler: the hadron's cell, its index in
// the cell (its random streams) and its (cell, species) pair, whose band
// table every hadron of that pair shares.
struct alignas(16) LSlot {
    uint32_t cell;
    uint32_t j;
    uint32_t pair;
    uint16_t species;
    uint16_t pad;
};
static_assert(sizeof(LSlot) == 16, "LSlot layout");

// The pairs of a pass.  One hadron in 1.5 shares its (cell, species) with
// another in the 200k-cell workload, so the network and the band envelope are
// evaluated once per pair.  k_pick (the species of every slot) also leaves each
// cell's species bitmap and its number of pairs (with routing: the species of
// two or more slots, and the number of the other slots); then
//   scan           the cells' first pair indices (and first routed slots)
//   k_lpair_fill   the compact pair list (cell, species) in bitmap order,
//                  the learned slots' records with their pairs, in slot order
//                  (one wave per cell), and with routing the routed slots'
//                  indices, in slot order, at the head of the plain list
// Every position comes from the scans, so the lists do not depend on timing.
constexpr int kPairCellsPerBlock = 4;
constexpr int kPairWords = kMaxSpecies / 32;

__global__ __launch_bounds__(64 * kPairCellsPerBlock) void k_lpair_fill(SampleArgs a, int Wb,
                                                                          const uint32_t* __restrict__ gbits,
                                                                          const int64_t* __restrict__ pairoff,
                                                                          const uint16_t* __restrict__ sp16,
                                                                          LSlot* __restrict__ slot,
                                                                          uint2* __restrict__ pkey,
                                                                          const int64_t* __restrict__ soff,
                                                                          int64_t* __restrict__ plain) {
    __shared__ uint32_t s_bits[kPairCellsPerBlock][kPairWords];
    __shared__ uint32_t s_pre[kPairCellsPerBlock][kPairWords];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kPairCellsPerBlock + w;
    if (c >= a.Nc) return;
    const int64_t o0 = a.off[c];
    int64_t n = a.off[c + 1] - o0;
    // (batched rounds: only the slots with set-up records)
    if (a.setup_cap > 0) n = max<int64_t>(0, min<int64_t>(n, a.setup_cap - o0));
    if (n == 0) return;
    uint32_t* bits = s_bits[w];
    uint32_t* pre = s_pre[w];
    const uint32_t wd = lane < Wb ? gbits[c * Wb + lane] : 0u;
    const uint32_t pc = __popc(wd);
    const uint32_t incl = wave_inclusive_scan(pc);
    if (lane < Wb) {
        bits[lane] = wd;
        pre[lane] = incl - pc;
    }
    wave_lds_sync();
    const int64_t base = pairoff[c];
    auto rank = [&](int s) { return pre[s >> 5] + __popc(bits[s >> 5] & ((1u << (s & 31)) - 1u)); };
    for (int s = lane; s < a.Ns; s += 64)
        if (bits[s >> 5] & (1u << (s & 31))) pkey[base + rank(s)] = make_uint2(static_cast<uint32_t>(c), s);
    // this cell's first learned and first routed positions
    const int64_t s0 = soff ? soff[c] : 0;
    int64_t lpos = o0 - s0, spos = s0;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int64_t r0 = 0; r0 < n; r0 += 64) {
        const int64_t r = r0 + lane;
        const int s = r < n ? static_cast<int>(sp16[o0 + r]) : 0;
        const bool pair = r < n && ((bits[s >> 5] >> (s & 31)) & 1u);
        const bool routed = r < n && !pair;
        const unsigned long long mp = __ballot(pair), mr = __ballot(routed);
        if (pair) {
            LSlot rec;
            rec.cell = static_cast<uint32_t>(c);
            rec.j = static_cast<uint32_t>(r);
            rec.pair = static_cast<uint32_t>(base + rank(s));
            rec.species = static_cast<uint16_t>(s);
            rec.pad = 0;
            slot[lpos + __popcll(mp & below)] = rec;
        }
        if (routed) plain[spos + __popcll(mr & below)] = o0 + r;
        lpos += __popcll(mp);
        spos += __popcll(mr);
    }
}

// the plain list's length before the learned trials append their fallbacks:
// the routed slots (none without routing); acc (batched rounds) sums them
__global__ void k_plain_count(const int64_t* __restrict__ nrouted, unsigned long long* __restrict__ count,
                              unsigned long long* __restrict__ acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const unsigned long long n = nrouted ? static_cast<unsigned long long>(*nrouted) : 0ull;
    *count = n;
    if (acc) *acc += n;
}


constexpr int kSetupThreads = 256;
constexpr int kMlpBytes = (64 * kXld + 64 * kHld) * 2;

// Set-up of the learned sampler, one wave per 64 consecutive (cell, species)
// pairs: the network inputs, the network on MFMA and the band envelope.
__global__ __launch_bounds__(kSetupThreads) void k_learn_setup(LearnArgs L, LRec* __restrict__ rec) {
    extern __shared__ __align__(16) char lds[];
    const SampleArgs& a = L.a;
    const int lane = threadIdx.x & 63;
    char* wlds = lds + (threadIdx.x >> 6) * kMlpBytes;
    uint16_t* Xs = reinterpret_cast<uint16_t*>(wlds);
    uint16_t* Hs = Xs + 64 * kXld;
    float* Os = reinterpret_cast<float*>(wlds);
    const int64_t s0 = (static_cast<int64_t>(blockIdx.x) * (kSetupThreads / 64) + (threadIdx.x >> 6)) * 64;
    const int64_t npairs = *L.P;
    if (s0 >= npairs) return;  // whole waves only: the MFMA below needs every lane
    const int64_t i = s0 + lane;
    const bool live = i < npairs;
    int64_t c = 0;
    int s = 0;
    const uint32_t jj = 0;
    LParams P{};
    float f[kIn];
    for (int k = 0; k < kIn; k++) f[k] = 0.f;
    if (live) {
        const uint2 pk = L.pkey[i];
        c = pk.x;
        s = static_cast<int>(pk.y);
        lparams_of(a, L.aux, c, s, P);
        features_of(a, c, s, P, f);
    }
    for (int k = 0; k < kIn; k += 4)
        *reinterpret_cast<uint2*>(Xs + lane * kXld + k) = make_uint2(pack_bf16(f[k], f[k + 1]), pack_bf16(f[k + 2], f[k + 3]));
    mlp_forward_wave(L.W, Xs, Hs, Os);
    if (!live) return;
    float rho[kOut];
    for (int k = 0; k < kOut; k++) rho[k] = sigmoid_fast(Os[lane * kOut + k]);
    BandTable B;
    int flags = P.flags & 3;
    if (!flags && !band_setup<false>(P, rho, B, nullptr, nullptr)) flags = 2;
    LRec r;
    r.cell = static_cast<uint32_t>(c);
    r.j = jj;
    r.species = static_cast<uint16_t>(s);
    r.flags = static_cast<uint8_t>(flags);
    r.klast = static_cast<uint8_t>(flags ? 0 : B.klast);
    for (int k = 0; k < 8; k++) r.Q[k] = flags ? 0 : B.Q[k];
    for (int k = 0; k < kBands; k++) {
        r.cs[k] = flags ? 0u : B.band[k].x;
        r.beta[k] = flags ? 0 : static_cast<uint8_t>((B.band[k].y >> 16) & 0xFFu);
        if (k < 8) r.kap[k] = flags ? 0 : static_cast<uint16_t>(B.band[k].y & 0xFFFFu);
    }
    for (int k = 0; k < 3; k++) r.pad[k] = 0;
    r.x = P.x;
    rec[i] = r;
}

// Persistent trial kernel of the learned sampler with per-lane refill (as
// k_trials).  A lane keeps the hadron's parameters and band thresholds in
// registers (thresholds, lines, kappas, betas: 20 words).  Proposal
// `it` of hadron (c, j) uses the Philox block (c, j, TAG_LEARN, it).
__global__ __launch_bounds__(256) void k_learn_trials(LearnArgs L, const LRec* __restrict__ rec) {
    __shared__ unsigned long long s_stats[6];
    if (threadIdx.x < 6) s_stats[threadIdx.x] = 0ull;
    __syncthreads();
    const SampleArgs& a = L.a;
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    const int64_t wid = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
    // (batched rounds: the round index and hadron count live on the device)
    const uint32_t chunk = a.chunk_dev ? *a.chunk_dev : a.chunk;
    int64_t Kn = a.Kdev ? static_cast<int64_t>(*a.Kdev) : a.K;
    if (a.setup_cap > 0 && Kn > a.setup_cap) Kn = a.setup_cap;
    // (the learned slot list: every slot but the routed ones)
    if (L.nsingle) Kn = max<int64_t>(0, Kn - *L.nsingle);
    const int64_t base = Kn / nwaves, rem = Kn % nwaves;
    const int64_t beg = wid * base + min(wid, rem);
    const int64_t end = beg + base + (wid < rem ? 1 : 0);
    const int64_t Nc = a.Nc;
    const uint32_t key0 = static_cast<uint32_t>(a.seed), key1 = static_cast<uint32_t>(a.seed >> 32);
    const uint32_t tw1 = tag_word(TAG_LEARN, chunk);
    // one proposal is one outer trial of the reference (src/FSSW.cpp:1872)
    const int max_it = a.max_trials;
    // launch-uniform forms of the shear and bulk terms (scalar registers):
    // shear Ps(E) = ps0 (1) or ps0 / E (2); bulk qi / E + q1 E (RTA, CE) or
    // q0 + q1 E + q2 E^2, so qi and q0 share the lane's register qa
    const int smode_u = a.shear ? (a.neos == 1 ? 2 : 1) : 0;
    const bool bulk_inv = a.bulk == BULK_RTA || a.bulk == BULK_CE;

    int64_t next = beg;
    // per-lane counters (32 bits: a lane's share of one launch), summed in 64
    uint32_t n_it = 0, n_fail = 0, n_range = 0, n_fb = 0, n_viol = 0;

    bool active = false;
    uint32_t c = 0;
    int64_t cur = 0;
    uint32_t jj = 0;
    int s = 0, it = 0, klast = 0;
    uint64_t outp = 0;
    // band table in registers: thresholds (2 x u16 per word), lines, kappas (2 x fp16), betas (4 x u8)
    uint32_t Qw[4], CS[kBands], KW[4], BW[3];
    for (int k = 0; k < 4; k++) { Qw[k] = 0u; KW[k] = 0u; }
    for (int k = 0; k < kBands; k++) CS[k] = 0u;
    for (int k = 0; k < 3; k++) BW[k] = 0u;
    LParams P{};

    while (true) {
        const unsigned long long need = __ballot(!active);
        if (need && next < end && (__popcll(need) >= a.refill_min || need == ~0ull)) {
            const int64_t navail = end - next;
            const int rank = __popcll(need & ((1ull << lane) - 1));
            if (!active && rank < navail) {
                const LSlot sl = L.slot[next + rank];
                const uint4* r = reinterpret_cast<const uint4*>(rec + sl.pair);
                const uint4 h0 = r[0], h1 = r[1], h2 = r[2], h3 = r[3], h4 = r[4], h5 = r[5];
                c = sl.cell;
                jj = sl.j;
                cur = a.off[c] + jj;  // the slot (its output position)
                s = static_cast<int>(h0.z & 0xFFFFu);
                const int flags = static_cast<int>((h0.z >> 16) & 0xFFu);
                klast = static_cast<int>(h0.z >> 24);
                outp = a.perms ? batch_out(a, cur, c) : a.perm(static_cast<uint64_t>(cur));
                if (flags & 1) {
                    // m/T - mu/T outside the table: a hole, as in the plain sampler
                    n_range++;
                    write_hole(a, outp);
                } else if (flags & 2) {
                    n_fb++;
                    const unsigned long long pos = atomicAdd(L.fb_count, 1ull);
                    L.fb_list[pos] = cur;
                } else {
                    // bytes 12..27 Q, 28..63 cs, 64..79 kappa, 80..88 beta
                    Qw[0] = h0.w; Qw[1] = h1.x; Qw[2] = h1.y; Qw[3] = h1.z;
                    CS[0] = h1.w; CS[1] = h2.x; CS[2] = h2.y; CS[3] = h2.z; CS[4] = h2.w;
                    CS[5] = h3.x; CS[6] = h3.y; CS[7] = h3.z; CS[8] = h3.w;
                    KW[0] = h4.x; KW[1] = h4.y; KW[2] = h4.z; KW[3] = h4.w;
                    BW[0] = h5.x; BW[1] = h5.y; BW[2] = h5.z;
                    lparams_fast(a, L.aux, c, s, __uint_as_float(h5.w), P);
                    P.qi += P.q0;  // (one of the two is zero: qa of the bulk form)
                    P.D = P.D > 0.f ? __builtin_amdgcn_rcpf(P.D) : 0.f;  // (its reciprocal)

                    it = 0;
                    active = true;
                }
            }
            next += min<int64_t>(__popcll(need), navail);
        }
        if (!__any(active)) {
            if (next >= end) break;
            continue;
        }
        if (!active) continue;
        // ---- one proposal: one Philox block gives the band (16 bits), the
        // flip (16), t, the axis and cos, phi (24 each) and the acceptance
        // (the three low bytes)
        const u32x4 w1r = philox4x32(u32x4{static_cast<uint32_t>(c), jj, tw1, static_cast<uint32_t>(it)}, key0, key1);
        it++;
        n_it++;
        const uint32_t w16 = w1r.x >> 16;
        int k = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            k += (w16 >= (Qw[j] & 0xFFFFu)) ? 1 : 0;
            k += (w16 >= (Qw[j] >> 16)) ? 1 : 0;
        }
        k = min(k, klast);
        uint32_t csk = CS[0], kwk = KW[0], bwk = BW[0];
#pragma unroll
        for (int j = 1; j < kBands; j++)
            if (k == j) { csk = CS[j]; kwk = KW[min(j >> 1, 3)]; bwk = BW[j >> 2]; }
        const float ck = h2f(static_cast<uint16_t>(csk & 0xFFFFu));
        const float sk = h2f(static_cast<uint16_t>(csk >> 16));
        const float kk = (k < kBands - 1) ? h2f(static_cast<uint16_t>((kwk >> ((k & 1) * 16)) & 0xFFFFu)) : 0.f;
        const float bk = static_cast<float>((bwk >> ((k & 3) * 8)) & 0xFFu) * (1.f / 255.f);
        const float tau_max = P.tmax - P.x;
        // the band's edges by selects (a switch on the lane's band would branch)
        float ea = 0.f, eb = band_edge(1);
#pragma unroll
        for (int j = 1; j < kBands; j++) {
            ea = k >= j ? band_edge(j) : ea;
            eb = k >= j ? band_edge(j + 1) : eb;
        }
        const float ba = fminf(ea, tau_max);
        const float bb = (k < kBands - 1) ? fminf(eb, tau_max) : tau_max;
        const float bw = bb - ba;
        // tau from exp(s tau) on [ba, bb]
        const float ut = u24(w1r.y);
        const float sw = sk * bw;
        float tau = fabsf(sw) > 1e-5f ? ba + trunc_exp_inv(ut, sw) * __builtin_amdgcn_rcpf(sk) : ba + ut * bw;
        tau = fminf(fmaxf(tau, ba), bb);
        const float t = P.x + tau;
        const float E = P.T * t + P.mu;
        const float p2 = fmaxf(E * E - P.m * P.m, 0.f);
        const float p = __builtin_amdgcn_sqrtf(p2);
        const float et = fexp(-t);
        const float fe = fdiv(1.f, 1.f + P.sgn * et);     // e^t f0 = 1 - sign f0
        // the target over the envelope line, exp(log(p E fe) - tau - (c + s tau)),
        // as a product (the log envelope's intercept c is an fp16 of magnitude
        // below 6e4, so the exponent is finite; p = 0 gives 0)
        const float pEf = p * E * fe;
        // direction in the shear eigenframe
        const float P0 = (1.f + kk * P.l1) * (1.f / 3.f), P1 = P0 + (1.f + kk * P.l2) * (1.f / 3.f);
        const float uy = u24(w1r.z);
        // the axis with probabilities P0, P1 - P0, 1 - P1 and the position
        // within its interval (one division for every lane)
        const bool ax0 = uy < P0, ax1 = !ax0 && uy < P1;
        const int ax = ax0 ? 0 : (ax1 ? 1 : 2);
        const float ylo = ax0 ? 0.f : (ax1 ? P0 : P1), yhi = ax0 ? P0 : (ax1 ? P1 : 1.f);
        float rr = fdiv(uy - ylo, yhi - ylo);
        rr = fminf(fmaxf(rr, 0.f), 1.f);
        const float cz = fcbrt(2.f * rr - 1.f);
        const float sz = __builtin_amdgcn_sqrtf(fmaxf(1.f - cz * cz, 0.f));
        float sph, cph;
        __sincosf(2.f * static_cast<float>(kPi) * u24(w1r.w), &sph, &cph);
        // (cz along the axis, then sz cos, sz sin along the next two, cyclically)
        const float nc = sz * cph, ns = sz * sph;
        float n0 = ax == 0 ? cz : (ax == 1 ? ns : nc);
        float n1 = ax == 0 ? nc : (ax == 1 ? cz : ns);
        float n2 = ax == 0 ? ns : (ax == 1 ? nc : cz);
        float y = n0 * P.dh0 + n1 * P.dh1 + n2 * P.dh2;
        const float uf = (static_cast<float>(w1r.x & 0xFFFFu) + 0.5f) * (1.f / 65536.f);
        if (uf >= 0.5f * (1.f + bk * y)) { n0 = -n0; n1 = -n1; n2 = -n2; y = -y; }
        const float W = P.l1 * n0 * n0 + P.l2 * n1 * n1 + P.l3 * n2 * n2;
        // the exact Cooper-Frye weights of the plain sampler
        const float rE = __builtin_amdgcn_rcpf(E);
        const float w1 = clamp01((P.d0 + p * rE * P.dn * y) * P.D);
        const float dfs = fe * p2 * W * (smode_u == 2 ? P.ps0 * rE : (smode_u == 1 ? P.ps0 : 0.f));
        const float dfb = fe * (bulk_inv ? P.qi * rE + P.q1 * E : P.qi + E * (P.q1 + P.q2 * E));
        const float w2 = clamp01(0.5f * (1.f + dfs + dfb));
        const float racc = fdiv(pEf * fexp(-(tau + ck + sk * tau)) * w1 * w2, (1.f + kk * W) * (1.f + bk * y));
        if (racc > 1.f + 1e-5f) n_viol++;
        bool done = false;
        const uint32_t ua = (w1r.y & 0xFFu) | ((w1r.z & 0xFFu) << 8) | ((w1r.w & 0xFFu) << 16);
        if ((static_cast<float>(ua) + 0.5f) * (1.0f / 16777216.0f) < racc) {
            const float R00 = L.aux[A_R00 * Nc + c], R01 = L.aux[A_R01 * Nc + c], R02 = L.aux[A_R02 * Nc + c];
            const float R10 = L.aux[A_R10 * Nc + c], R11 = L.aux[A_R11 * Nc + c], R12 = L.aux[A_R12 * Nc + c];
            const float R20 = L.aux[A_R20 * Nc + c], R21 = L.aux[A_R21 * Nc + c], R22 = L.aux[A_R22 * Nc + c];
            const float px = p * (R00 * n0 + R01 * n1 + R02 * n2);
            const float py = p * (R10 * n0 + R11 * n1 + R12 * n2);
            const float pz = p * (R20 * n0 + R21 * n1 + R22 * n2);
            if (a.ecell) {
                const float4* ec = a.ecell + 2 * static_cast<int64_t>(c);
                emit_cell_lab(a, ec[0], ec[1], a.sp.pid[s], P.m, px, py, pz, E, outp);
            }
            else emit_lab(a, chunk, c, jj, a.sp.pid[s], P.m, px, py, pz, E, outp);
            done = true;
        } else if (it >= max_it) {
            n_fail++;
            write_hole(a, outp);
            done = true;
        }
        if (done) active = false;
    }
    const unsigned long long t_it = wave_sum(static_cast<unsigned long long>(n_it));
    const unsigned long long t_fail = wave_sum(static_cast<unsigned long long>(n_fail));
    const unsigned long long t_range = wave_sum(static_cast<unsigned long long>(n_range));
    const unsigned long long t_fb = wave_sum(static_cast<unsigned long long>(n_fb));
    const unsigned long long t_viol = wave_sum(static_cast<unsigned long long>(n_viol));
    if (lane == 0) {
        atomicAdd(&s_stats[0], t_it);
        atomicAdd(&s_stats[1], t_it);
        atomicAdd(&s_stats[2], t_fail);
        atomicAdd(&s_stats[3], t_range);
        atomicAdd(&s_stats[4], t_fb);
        atomicAdd(&s_stats[5], t_viol);
    }
    __syncthreads();
    if (threadIdx.x < 4 && s_stats[threadIdx.x]) atomicAdd(&a.stats[threadIdx.x], s_stats[threadIdx.x]);
    if (threadIdx.x >= 4 && threadIdx.x < 6 && s_stats[threadIdx.x])
        atomicAdd(&L.stats2[threadIdx.x - 4], s_stats[threadIdx.x]);
}

// =============================================================== training
struct TrainArgs {
    SampleArgs a;                  // surface, species, tables, delta-f switches
    const float* aux;
    MlpDev W;
    const uint32_t* cell;          // training set: (cell, species) of accepted hadrons
    const uint16_t* species;
    int64_t n;                     // training-set size
    const int64_t* pick;           // explicit sample indices (or nullptr)
    int direct;                    // sample i of the step is training sample i (else Philox draws)
    int64_t npick;                 // samples in this step
    uint64_t seed;
    uint32_t step;
    float* partial;                // [gridDim.x][kNumParams]
    float* loss_part;              // [gridDim.x]: sum of log M, [gridDim.x + b]: valid samples
    float* out_loss;               // optional per-sample log M
    float* out_dz;                 // optional per-sample dlogM/dz [8]
    float* out_x;                  // optional per-sample inputs [16] (after bf16 rounding)
};

constexpr int kTrainLds =
    (64 * kXld + kIn * kHld + 4 * 64 * kHld + 64 * kXld + kOut * kHld + 3 * 64 * kHld) * 2 +
    (2 * 64 * kZld + 64 * kOut) * 4;

// bias gradient: sum over the 32 columns (samples) of one lane half
__device__ __forceinline__ float half_sum(float v) {
#pragma unroll
    for (int d = 16; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Forward and backward pass for 64 samples per block (one wave); writes the
// block's gradient of sum_s log M_s and its loss.
__global__ __launch_bounds__(64) void k_mlp_train(TrainArgs t) {
    extern __shared__ __align__(16) char lds[];
    uint16_t* Xs = reinterpret_cast<uint16_t*>(lds);            // [64][kXld]
    uint16_t* Xt = Xs + 64 * kXld;                              // [16][kHld]
    uint16_t* H1s = Xt + kIn * kHld;                            // [64][kHld]
    uint16_t* H1t = H1s + 64 * kHld;
    uint16_t* H2s = H1t + 64 * kHld;
    uint16_t* H2t = H2s + 64 * kHld;
    uint16_t* dOs = H2t + 64 * kHld;                            // [64][kXld]
    uint16_t* dOt = dOs + 64 * kXld;                            // [8][kHld]
    uint16_t* dZ2s = dOt + kOut * kHld;                         // [64][kHld]
    uint16_t* dZ2t = dZ2s + 64 * kHld;
    uint16_t* dZ1t = dZ2t + 64 * kHld;
    float* Z1 = reinterpret_cast<float*>(dZ1t + 64 * kHld);     // [64 hid][kZld]
    float* Z2 = Z1 + 64 * kZld;
    float* Os = Z2 + 64 * kZld;                                 // [64][8]
    const int l = threadIdx.x & 63, r = l & 31, g = l >> 5;
    const SampleArgs& a = t.a;
    const float* par = t.W.par;
    float* gp = t.partial + static_cast<int64_t>(blockIdx.x) * kNumParams;

    // ---- inputs
    const int64_t si = static_cast<int64_t>(blockIdx.x) * 64 + l;
    const bool live = si < t.npick && t.n > 0;
    float f[kIn];
    for (int i = 0; i < kIn; i++) f[i] = 0.f;
    LParams P{};
    bool valid = false;
    if (live) {
        int64_t idx;
        if (t.pick) idx = t.pick[si];
        else if (t.direct) idx = si;
        else {
            const u32x4 w = philox4x32(u32x4{static_cast<uint32_t>(si), t.step, tag_word(TAG_TRAIN, 0), 0u},
                                       static_cast<uint32_t>(t.seed), static_cast<uint32_t>(t.seed >> 32));
            idx = static_cast<int64_t>((static_cast<uint64_t>(w.x) << 32 | w.y) % static_cast<uint64_t>(t.n));
        }
        const int64_t c = t.cell[idx];
        const int s = t.species[idx];
        lparams_of(a, t.aux, c, s, P);
        valid = !(P.flags & 3);
        if (valid) features_of(a, c, s, P, f);
    }
    for (int i = 0; i < kIn; i++) {
        const uint16_t h = f2bf(f[i]);
        Xs[l * kXld + i] = h;
        Xt[i * kHld + l] = h;
        if (t.out_x && live) t.out_x[si * kIn + i] = bf2f(h);
    }
    wave_lds_sync();
    // ---- forward
    for (int nt = 0; nt < 2; nt++)
        for (int mt = 0; mt < 2; mt++) {
            const f32x16 acc = tile_gl(t.W.frag + F_W1 + mt * 64, Xs, kXld, nt * 32, 1, zero16());
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int h = mt * 32 + 8 * (i >> 2) + 4 * g + (i & 3), s = nt * 32 + r;
                const float z = acc[i] + par[P_B1 + h];
                Z1[h * kZld + s] = z;
                const uint16_t hb = f2bf(gelu(z));
                H1s[s * kHld + h] = hb;
                H1t[h * kHld + s] = hb;
            }
        }
    wave_lds_sync();
    for (int nt = 0; nt < 2; nt++)
        for (int mt = 0; mt < 2; mt++) {
            const f32x16 acc = tile_gl(t.W.frag + F_W2 + mt * 4 * 64, H1s, kHld, nt * 32, 4, zero16());
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int h = mt * 32 + 8 * (i >> 2) + 4 * g + (i & 3), s = nt * 32 + r;
                const float z = acc[i] + par[P_B2 + h];
                Z2[h * kZld + s] = z;
                const uint16_t hb = f2bf(gelu(z));
                H2s[s * kHld + h] = hb;
                H2t[h * kHld + s] = hb;
            }
        }
    wave_lds_sync();
    for (int nt = 0; nt < 2; nt++) {
        const f32x16 acc = tile_gl(t.W.frag + F_W3, H2s, kHld, nt * 32, 4, zero16());
#pragma unroll
        for (int e = 0; e < 4; e++) Os[(nt * 32 + r) * kOut + 4 * g + e] = acc[e] + par[P_B3 + 4 * g + e];
    }
    wave_lds_sync();
    // ---- loss log M and its gradient at the outputs
    float dz[kOut], lm = 0.f;
    for (int k = 0; k < kOut; k++) dz[k] = 0.f;
    if (valid) {
        float rho[kOut];
        for (int k = 0; k < kOut; k++) rho[k] = sigmoid_fast(Os[l * kOut + k]);
        BandTable B;
        if (!band_setup<true>(P, rho, B, &lm, dz)) { valid = false; lm = 0.f; }
    }
    if (t.out_loss && live) t.out_loss[si] = valid ? lm : 0.f;
    if (t.out_dz && live)
        for (int k = 0; k < kOut; k++) t.out_dz[si * kOut + k] = dz[k];
    for (int k = 0; k < kIn; k++) dOs[l * kXld + k] = k < kOut ? f2bf(dz[k]) : 0;
    for (int k = 0; k < kOut; k++) dOt[k * kHld + l] = f2bf(dz[k]);
    {
        const float lsum = wave_sum(valid ? lm : 0.f);
        const float nval = wave_sum(valid ? 1.f : 0.f);
        for (int k = 0; k < kOut; k++) {
            const float gb = wave_sum(dz[k]);
            if (l == 0) gp[P_B3 + k] = gb;
        }
        if (l == 0) {
            t.loss_part[blockIdx.x] = lsum;
            t.loss_part[gridDim.x + blockIdx.x] = nval;
        }
    }
    wave_lds_sync();
    // ---- backward
    // dW3[o][h] = sum_s dO[s][o] H2[s][h]
    for (int nt = 0; nt < 2; nt++) {
        const f32x16 acc = tile_ll(dOt, kHld, 0, kOut, H2t, kHld, nt * 32, 64, 4, zero16());
#pragma unroll
        for (int e = 0; e < 4; e++) gp[P_W3 + (4 * g + e) * kHid + nt * 32 + r] = acc[e];
    }
    // dH2^T = W3^T dO^T; dZ2 = dH2 * gelu'(Z2)
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++) {
            const f32x16 acc = tile_gl(t.W.frag + F_W3T + mt * 64, dOs, kXld, nt * 32, 1, zero16());
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int h = mt * 32 + 8 * (i >> 2) + 4 * g + (i & 3), s = nt * 32 + r;
                const float d = acc[i] * gelu_d(Z2[h * kZld + s]);
                const uint16_t db = f2bf(d);
                dZ2s[s * kHld + h] = db;
                dZ2t[h * kHld + s] = db;
                Z2[h * kZld + s] = d;  // keep FP32 dZ2 for the bias gradient
            }
        }
    wave_lds_sync();
    // bias 2: sum over samples
    for (int h = 0; h < kHid; h++) {
        const float v = Z2[h * kZld + l];
        const float sv = wave_sum(v);
        if (l == 0) gp[P_B2 + h] = sv;
    }
    // dW2[ho][hi] = sum_s dZ2[ho][s] H1[s][hi]
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++) {
            const f32x16 acc = tile_ll(dZ2t, kHld, mt * 32, 64, H1t, kHld, nt * 32, 64, 4, zero16());
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int ho = mt * 32 + 8 * (i >> 2) + 4 * g + (i & 3), hi = nt * 32 + r;
                gp[P_W2 + ho * kHid + hi] = acc[i];
            }
        }
    // dH1^T = W2^T dZ2; dZ1 = dH1 * gelu'(Z1)
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++) {
            const f32x16 acc = tile_gl(t.W.frag + F_W2T + mt * 4 * 64, dZ2s, kHld, nt * 32, 4, zero16());
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int h = mt * 32 + 8 * (i >> 2) + 4 * g + (i & 3), s = nt * 32 + r;
                const float d = acc[i] * gelu_d(Z1[h * kZld + s]);
                dZ1t[h * kHld + s] = f2bf(d);
                Z1[h * kZld + s] = d;
            }
        }
    wave_lds_sync();
    for (int h = 0; h < kHid; h++) {
        const float v = Z1[h * kZld + l];
        const float sv = wave_sum(v);
        if (l == 0) gp[P_B1 + h] = sv;
    }
    // dW1[h][i] = sum_s dZ1[h][s] X[s][i]
    for (int mt = 0; mt < 2; mt++) {
        const f32x16 acc = tile_ll(dZ1t, kHld, mt * 32, 64, Xt, kHld, 0, kIn, 4, zero16());
        if (r < kIn) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int h = mt * 32 + 8 * (i >> 2) + 4 * g + (i & 3);
                gp[P_W1 + h * kIn + r] = acc[i];
            }
        }
    }
}

// Sum the blocks' gradients; with lr > 0 also take one Adam step on the
// mean gradient (betas 0.9, 0.999, eps 1e-8).
__global__ void k_mlp_reduce_adam(const float* __restrict__ partial, int nblk, const float* __restrict__ loss_part,
                                  float* __restrict__ par, float* __restrict__ adam_m, float* __restrict__ adam_v,
                                  float* __restrict__ grad_out, float lr, int t_step) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    float gsum = 0.f, nval = 0.f;
    for (int b = 0; b < nblk; b++) {
        gsum += partial[static_cast<int64_t>(b) * kNumParams + p];
        nval += loss_part[nblk + b];
    }
    if (grad_out) grad_out[p] = gsum;
    if (lr > 0.f && nval > 0.f) {
        const float gm = gsum / nval;
        const float m = 0.9f * adam_m[p] + 0.1f * gm;
        const float v = 0.999f * adam_v[p] + 0.001f * gm * gm;
        adam_m[p] = m;
        adam_v[p] = v;
        const float mh = m / (1.f - powf(0.9f, static_cast<float>(t_step)));
        const float vh = v / (1.f - powf(0.999f, static_cast<float>(t_step)));
        par[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
    }
}

// Training set: the (cell, species) of accepted hadrons of a plain pass, in
// slot order; more than cap of them are thinned evenly.
__global__ void k_collect_flag(const HSetup* __restrict__ rec, int64_t K, const Hadron* __restrict__ out,
                               uint32_t* __restrict__ flag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= K) return;
    flag[i] = out[rec[i].out].pid != 0 ? 1u : 0u;
}
__global__ void k_collect_write(const HSetup* __restrict__ rec, int64_t K, const uint32_t* __restrict__ flag,
                                const int64_t* __restrict__ off, int64_t cap, uint32_t* __restrict__ cell,
                                uint16_t* __restrict__ spec) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= K || !flag[i]) return;
    const int64_t o = off[i], n = off[K];
    int64_t pos = o;
    if (n > cap) {
        const int64_t lo = o * cap / n, hi = (o + 1) * cap / n;
        if (hi == lo) return;
        pos = lo;
    }
    cell[pos] = rec[i].cell;
    spec[pos] = rec[i].species;
}

// Band tables for given (cell, species, rho): diagnostics and tests.
// out[i]: [0] ok, [1] log M, [2..10] c, [11..19] s, [20..28] kappa,
// [29..37] beta, [38..45] Q / 2^16, [46] last band of positive mass
constexpr int kBandDiag = 47;
__global__ void k_band_diag(SampleArgs a, const float* __restrict__ aux, const int32_t* __restrict__ cell,
                            const int32_t* __restrict__ spec, const float* __restrict__ rho, int64_t n,
                            float* __restrict__ out, float* __restrict__ lp) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    LParams P;
    lparams_of(a, aux, cell[i], spec[i], P);
    float rr[kOut];
    for (int k = 0; k < kOut; k++) rr[k] = rho[i * kOut + k];
    BandTable B{};
    float lm = 0.f;
    const bool ok = !(P.flags & 3) && band_setup<false>(P, rr, B, &lm, nullptr);
    float* o = out + i * kBandDiag;
    o[0] = ok ? 1.f : 0.f;
    o[1] = lm;
    for (int k = 0; k < kBands; k++) {
        float cc, ss, kk, bb;
        band_unpack(ok ? B.band[k] : make_uint2(0u, 0u), cc, ss, kk, bb);
        o[2 + k] = cc; o[11 + k] = ss; o[20 + k] = kk; o[29 + k] = bb;
    }
    for (int k = 0; k < kBands - 1; k++) o[38 + k] = ok ? static_cast<float>(B.Q[k]) * (1.f / 65536.f) : 0.f;
    o[46] = ok ? static_cast<float>(B.klast) : 0.f;
    // the hadron parameters the tables were built from
    float* q = lp + i * 24;
    q[0] = P.T; q[1] = P.mu; q[2] = P.m; q[3] = P.sgn; q[4] = P.x; q[5] = P.tmax;
    q[6] = P.d0; q[7] = P.dn; q[8] = P.D; q[9] = P.l1; q[10] = P.l2; q[11] = P.l3;
    q[12] = P.dh0; q[13] = P.dh1; q[14] = P.dh2; q[15] = P.ps0; q[16] = P.qi; q[17] = P.q0;
    q[18] = P.q1; q[19] = P.q2; q[20] = static_cast<float>(P.smode); q[21] = static_cast<float>(P.flags);
    q[22] = 0.f; q[23] = 0.f;
}

}  // namespace issdev
