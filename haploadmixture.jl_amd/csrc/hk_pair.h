// hk_pair.h — per-(individual, SNP block) algebra shared by the genotype-pass kernels (hk_pass.hip, hk_ll.hip).
#pragma once
#include "hk_device.h"

namespace hk {

// -----------------------------------------------------------------------------------------------
// Per-pair algebra (shared by every pass)
//
// Per-nibble compatibility coefficients, 128 entries in LDS (each workgroup fills them with threads
// 0..127): lut[4 nib + f] = (h s1_00, h s1_01, h s1_11, e)[f] and lut[64 + 4 nib + f] =
// (s2_00, s2_01, s2_11, 2 h s1_01)[f], with S1, S2 the compatibility matrices of SNPs 2j and 2j+1
// (snp_s), h = 1/2 when either SNP is heterozygous (else 1), and e = 1 for the double-missing nibble,
// whose S1 is 0.  With Y = S1' X S2 (S1' = h S1) and L' = <X, Y> + e:
//   likelihood pass:  L' = L / 2^[a SNP is het]   (1 for a double-missing pair)
//   gradient passes:  w_h / 2 = Y_h / L'          (0 for a double-missing pair)
//   curvature passes: v_h = Y_h / (L' X_h) = c_h / (2 qp_h^2)
// (the weights do not depend on the scale h), so no per-pair test or branch is needed for any class.
// -----------------------------------------------------------------------------------------------

template <typename TP>
__device__ __forceinline__ void fill_pair_lut(TP* lut) {
  const int t = threadIdx.x;
  if (t < 128) {
    const int tab = t >> 6, nib = (t >> 2) & 15, f = t & 3;
    TP s00, s01, s11, h = ((nib & 3) == 2 || (nib >> 2) == 2) ? (TP)0.5 : (TP)1;
    snp_s<TP>(tab ? (unsigned)nib >> 2 : (unsigned)nib & 3u, s00, s01, s11);
    TP v;
    if (tab == 0) {
      v = f == 0 ? h * s00 : (f == 1 ? h * s01 : (f == 2 ? h * s11 : (TP)0));
      if (nib == (int)HK_NIB_DM) v = f == 3 ? (TP)1 : (TP)0;
    } else {
      if (f < 3) {
        v = f == 0 ? s00 : (f == 1 ? s01 : s11);
      } else {
        TP a00, a01, a11;  // 2 h s1_01 of SNP 2j (0 for the double-missing nibble)
        snp_s<TP>((unsigned)nib & 3u, a00, a01, a11);
        v = nib == (int)HK_NIB_DM ? (TP)0 : (TP)2 * h * a01;
      }
    }
    lut[t] = v;
  }
}

// Y = S1' X S2 and L' = <X, Y> + e from the two table entries of a nibble.
template <typename TP>
struct PairForm {
  typedef TP T4 __attribute__((ext_vector_type(4)));
  TP Y[4], L;
  __device__ __forceinline__ PairForm(const TP (&X)[4], const TP* __restrict__ lut, unsigned nib) {
    const T4 a = *(const T4*)(lut + 4 * nib);
    const T4 b = *(const T4*)(lut + 64 + 4 * nib);
    const TP z00 = hk_fma(X[1], b[1], X[0] * b[0]), z01 = hk_fma(X[1], b[2], X[0] * b[1]);
    const TP z10 = hk_fma(X[3], b[1], X[2] * b[0]), z11 = hk_fma(X[3], b[2], X[2] * b[1]);
    Y[0] = hk_fma(a[1], z10, a[0] * z00);
    Y[1] = hk_fma(a[1], z11, a[0] * z01);
    Y[2] = hk_fma(a[2], z10, a[1] * z00);
    Y[3] = hk_fma(a[2], z11, a[1] * z01);
    L = hk_fma(X[0], Y[0], a[3]);
    L = hk_fma(X[1], Y[1], L);
    L = hk_fma(X[2], Y[2], L);
    L = hk_fma(X[3], Y[3], L);
  }
};

// Pair likelihood L' without Y: with Z = X S2 and G_aa' = sum_b X_ab Z_a'b (G_01 = G_10, S2 symmetric),
// L' = h s1_00 G_00 + 2 h s1_01 G_01 + h s1_11 G_11 + e.
template <typename TP>
__device__ __forceinline__ TP pair_lik(const TP (&X)[4], const TP* __restrict__ lut, unsigned nib) {
  typedef TP T4 __attribute__((ext_vector_type(4)));
  const T4 a = *(const T4*)(lut + 4 * nib);
  const T4 b = *(const T4*)(lut + 64 + 4 * nib);
  const TP z00 = hk_fma(X[1], b[1], X[0] * b[0]), z01 = hk_fma(X[1], b[2], X[0] * b[1]);
  const TP z10 = hk_fma(X[3], b[1], X[2] * b[0]), z11 = hk_fma(X[3], b[2], X[2] * b[1]);
  const TP g00 = hk_fma(X[1], z01, X[0] * z00);
  const TP g01 = hk_fma(X[1], z11, X[0] * z10);
  const TP g11 = hk_fma(X[3], z11, X[2] * z10);
  return hk_fma(b[3], g01, hk_fma(a[2], g11, hk_fma(a[0], g00, a[3])));
}

// Factored pair likelihood (the matrix-core likelihood pass): the L' of pair_lik in fewer operations.  With X
// indexed 2a + b (a the allele of SNP 2j, b of SNP 2j+1), the compatibility matrix of a homozygous or missing SNP
// code is u u' (u = (1,0), (0,1), or (1,1) for missing) and that of a heterozygous one e0 e1' + e1 e0', so
// L' = h sum S1[a,a'] S2[b,b'] X_ab X_a'b' + e factors per nibble as
//   L' = (alpha . X)(beta . X) + dh (X_00 X_11 + X_01 X_10) + e:
// alpha = beta = u1 (x) u2 when neither SNP is heterozygous; alpha = e0 (x) u2, beta = e1 (x) u2 when SNP 2j is
// (h = 1/2 cancels the 2 of the two cross terms), and alpha = u1 (x) e0, beta = u1 (x) e1 when SNP 2j+1 is;
// dh = 1 (alpha = beta = 0) when both are; e = 1 (all else 0) for the double-missing nibble.  12 operations per
// pair against 17 for the general form.  Table (160 entries): lut[4 nib + f] = alpha_f, lut[64 + 4 nib + f] =
// beta_f, lut[128 + 2 nib] = (dh, e); threads 0..159 fill it.
template <typename TP>
__device__ __forceinline__ void fill_lik_lut(TP* lut) {
  const int t = threadIdx.x;
  if (t < 128) {
    const int tab = t >> 6, nib = (t >> 2) & 15, f = t & 3;
    const unsigned c1 = (unsigned)nib & 3u, c2 = (unsigned)nib >> 2;
    const int a = f >> 1, b = f & 1;
    // u_x of a homozygous or missing code (codes: 0 hom, 1 missing, 2 het, 3 hom alt)
    auto u = [](unsigned c, int x) { return c == 1u ? 1 : (c == 0u ? (int)(x == 0) : (c == 3u ? (int)(x == 1) : 0)); };
    int v;
    if (nib == (int)HK_NIB_DM || (c1 == 2u && c2 == 2u)) v = 0;
    else if (c1 == 2u) v = (int)(a == tab) * u(c2, b);
    else if (c2 == 2u) v = u(c1, a) * (int)(b == tab);
    else v = u(c1, a) * u(c2, b);
    lut[t] = (TP)v;
  } else if (t < 160) {
    const int nib = (t - 128) >> 1;
    lut[t] = (t & 1) ? (nib == (int)HK_NIB_DM ? (TP)1 : (TP)0) : (nib == 10 ? (TP)1 : (TP)0);
  }
}

template <typename TP>
__device__ __forceinline__ TP lik_factored(const TP (&X)[4], const TP* __restrict__ lut, unsigned nib) {
  typedef TP T4 __attribute__((ext_vector_type(4)));
  typedef TP T2 __attribute__((ext_vector_type(2)));
  const T4 al = *(const T4*)(lut + 4 * nib);
  const T4 be = *(const T4*)(lut + 64 + 4 * nib);
  const T2 de = *(const T2*)(lut + 128 + 2 * nib);
  const TP A = hk_fma(al[3], X[3], hk_fma(al[2], X[2], hk_fma(al[1], X[1], al[0] * X[0])));
  const TP B = hk_fma(be[3], X[3], hk_fma(be[2], X[2], hk_fma(be[1], X[1], be[0] * X[0])));
  const TP D = hk_fma(X[1], X[2], X[0] * X[3]);
  return hk_fma(A, B, hk_fma(de[0], D, de[1]));
}

// Curvature weights v_h = Y_h / (L X_h) (the factor 2 of c_h / qp_h^2 is in the outer-product tables).
// Two reciprocals per pair: v_0, v_1 share 1 / (L' X_0 X_1) and v_2, v_3 share 1 / (L' X_2 X_3) (in FP32
// range for every X_h >= ~3e-10, far below the 1e-5 floor the projections keep).
template <typename TP>
__device__ __forceinline__ void pair_v(const TP (&X)[4], const TP* __restrict__ lut, unsigned nib, TP (&v)[4]) {
  const PairForm<TP> f(X, lut, nib);
  const TP r01 = hk_rcp(f.L * X[0] * X[1]), r23 = hk_rcp(f.L * X[2] * X[3]);
  v[0] = f.Y[0] * X[1] * r01;
  v[1] = f.Y[1] * X[0] * r01;
  v[2] = f.Y[2] * X[3] * r23;
  v[3] = f.Y[3] * X[2] * r23;
}

// Genotype nibble at bit position sh of a word (one bit-field extract).
__device__ __forceinline__ unsigned nib_at(uint32_t word, unsigned sh) { return __builtin_amdgcn_ubfe(word, sh, 4); }


// Matrix-core step of the curvature passes for the parameter type (16 x 16 x 4, FP32 or FP64).
template <typename TP>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return g + 4 * v; }
};

__host__ __device__ constexpr int pow2_at_least(int v) { return v <= 1 ? 1 : 2 * pow2_at_least((v + 1) / 2); }

// Staging table rows of the curvature passes: row c < nrow of out (RG entries) holds, for s < sub, the QS packed
// products 2 x[k,m] x[k',m] of column m = sub c + s, then the sub K values x[:, sub c + s], then zeros.
// Rows c >= ncol / sub repeat the last column group.
template <int K, typename TP>
__global__ void k_stage_rows(const TP* __restrict__ x, int64_t ncol, int64_t nrow, int sub, int QS, int RG,
                             TP* __restrict__ out, const int* __restrict__ gate) {
  if (gate_off(gate)) return;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nrow * RG;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = idx / RG;
    const int e = (int)(idx - c * RG);
    const int64_t last = ncol / sub - 1;
    const int64_t cc = c < last ? c : last;
    TP v = (TP)0;
    if (e < sub * QS) {
      const int sgrp = e / QS, n = e - sgrp * QS;
      if (n < K * (K + 1) / 2) {
        int k = 0;
        while ((k + 1) * (k + 2) / 2 <= n) ++k;
        const int k2 = n - k * (k + 1) / 2;
        const TP* xc = x + (int64_t)K * (sub * cc + sgrp);
        v = (TP)2 * (xc[k] * xc[k2]);
      }
    } else if (e < sub * (QS + K)) {
      const int t = e - sub * QS, sgrp = t / K, k = t - sgrp * K;
      v = x[k + (int64_t)K * (sub * cc + sgrp)];
    }
    out[idx] = v;
  }
}

}  // namespace hk
