// hk_device.h — gfx950 device helpers shared by the HaploADMIXTURE pass, QP and fit kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HK_MAXK 16         // largest K with compiled pass and step kernels
#define HK_MAXK_REG 8      // largest K whose kernels keep a pair's K x K products in one lane's registers
#define HK_WAVE 64

// Gate of a fit's iteration kernels: a fit whose device-side stopping rule has fired (*gate == 0) skips every
// kernel of later iterations (they return at once), so a graph holding several fits' iterations can be
// replayed until all of them stopped.  nullptr = always run.
__device__ __forceinline__ bool gate_off(const int* gate) { return gate != nullptr && *gate == 0; }

// XCD-grouped workgroup order.  The dispatcher hands workgroup L (linear id in a 2-D grid) to XCD L % 8, and each
// XCD has an L2 of its own.  For a pass whose grid is (row groups, chunks), with every workgroup of a chunk reading
// the same staging-table slice, this bijection gives XCD x the contiguous range of logical ids
// [x q + min(x, r), ...) (q = N / 8, r = N % 8): the workgroups of a chunk run back to back on one XCD, and the
// slice is fetched into one L2 instead of eight.  Returns the logical (x, y) of this workgroup.
#define HK_XCDS 8
__device__ __forceinline__ void xcd_grouped_block(unsigned& bx, unsigned& by) {
  const unsigned N = gridDim.x * gridDim.y, L = blockIdx.x + gridDim.x * blockIdx.y;
  const unsigned q = N / HK_XCDS, r = N % HK_XCDS, x = L % HK_XCDS;
  const unsigned lg = x * q + (x < r ? x : r) + L / HK_XCDS;
  bx = lg % gridDim.x;
  by = lg / gridDim.x;
}

// ---------------------------------------------------------------------------------------------
// Genotype nibble: bits 0-1 = BED code of SNP 2j, bits 2-3 = BED code of SNP 2j+1.
// BED codes: 0 hom allele-1, 1 missing, 2 het, 3 hom allele-2 (src/loops.jl:63 maps them to
// g = 0, 3(missing), 1, 2).  Nibble 0b0101 = both SNPs missing ("double missing").
// ---------------------------------------------------------------------------------------------
#define HK_NIB_DM 5u

// Compatibility matrix of one SNP code, S = [[s00, s01], [s01, s11]]: S[a][a'] = 1 when alleles
// (a, a') on the two haplotype copies fit the genotype.  hom allele-1: s00; hom allele-2: s11;
// het: s01; missing: all ones.  With hi/lo the code bits: s00 = !hi, s11 = lo, s01 = hi ^ lo.
template <typename T>
__device__ __forceinline__ void snp_s(unsigned c, T& s00, T& s01, T& s11) {
  const unsigned hi = c >> 1, lo = c & 1u;
  s00 = (T)(hi ^ 1u);
  s11 = (T)lo;
  s01 = (T)(hi ^ lo);
}

// Type-generic fused multiply-add (__builtin_fma is the double-precision builtin).
__device__ __forceinline__ float hk_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double hk_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------------------------------------
// BF16 terms of FP32 values for the BF16 matrix instructions.  split3: x = h + m + l, each term the rounded
// remainder of the previous ones, so the three carry the 24 bits of x (each remainder is exact in FP32: it has at
// most 16, then 8 significant bits), and every BF16 x BF16 product is exact in an FP32 accumulator.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ void split3(float x, uint16_t& h, uint16_t& m, uint16_t& l) {
  h = f32_to_bf16(x);
  const float r1 = x - bf16_to_f32(h);
  m = f32_to_bf16(r1);
  l = f32_to_bf16(r1 - bf16_to_f32(m));
}

// ---------------------------------------------------------------------------------------------
// Reciprocals.  FP32: the hardware v_rcp_f32 (1 ulp).  FP64: v_rcp_f64 seed + two Newton steps.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float hk_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double hk_rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// ---------------------------------------------------------------------------------------------
// Quad (4-lane) DPP broadcasts and sums.  quad_perm control = p0 | p1<<2 | p2<<4 | p3<<6.
// ---------------------------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
}
template <int LANE>
__device__ __forceinline__ float quad_bcast(float v) { return dpp_f<LANE * 0x55>(v); }
template <int LANE>
__device__ __forceinline__ double quad_bcast(double v) { return dpp_f<LANE * 0x55>(v); }

// Sum over the 4 lanes of a quad; every lane gets the same (bitwise) value.
template <typename T>
__device__ __forceinline__ T quad_sum(T v) {
  v = v + dpp_f<0xB1>(v);  // quad_perm [1,0,3,2]
  v = v + dpp_f<0x4E>(v);  // quad_perm [2,3,0,1]
  return v;
}

// Wave-uniform value (forces SGPR residency of loop-invariant indices).
__device__ __forceinline__ int hk_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---------------------------------------------------------------------------------------------
// Block reduction of doubles (blockDim.x multiple of 64), fixed order: result valid in thread 0.
// ---------------------------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  // wave-level tree via LDS for determinism and simplicity
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] += sh[t + s];
    __syncthreads();
  }
  return sh[0];
}
