// MI355X (gfx950) kernels for the point-set LDDMM hot path of py-lddmm.
//
// One segmented pair-sum engine serves every kernel sum of the path:
//   applyK / applyDiffKT / applyDDiffK11and12   (py-lddmm/kernelFunctions.f90:1-114,430-496)
//   the Euler flow step and the covector step    (py-lddmm/pointEvolution.py:694-697,746-760)
//   the RKHS products of the matching driver      (py-lddmm/surfaceMatching.py:353-377)
//   the currents norm and its gradient            (py-lddmm/surfaces.py:1009-1089, curves.py:539-613)
//
// Work decomposition.  A launch holds nseg independent segments (one kernel
// sum each: a registration, a time step, a call).  A segment's targets are cut
// into row tiles of ROWS rows (BLOCK threads x RPT rows per thread, rows held
// in registers) and its sources into chunks of whole LDS tiles.  One work item
// = (segment, row tile, source chunk) = one workgroup.  The workgroup stages
// TS sources at a time into LDS (structure of arrays, all lanes read the same
// source: LDS broadcast) and every lane accumulates its RPT rows.  Chunk
// partials go to the workspace in FP64; a finalize kernel combines them in
// chunk order and applies the operation's epilogue (Euler update, covector
// update, row products); FP64 totals are reduced in a fixed tree.  The chunk
// length of a segment depends only on that segment's sizes, so a segment's
// result is bit-identical alone and inside any batch.  Segments with a processing
// order and energy-only sums have their source tiles built once per launch by
// k_stage into the workspace; their workgroups copy the tiles instead of building
// them (same values).
//
// Precision.  FP64 mode evaluates every pair in double (the pair kernels' own
// exp2_tab256 / exp2_nonpos; the dense kernel-matrix kernel uses the device math
// library's exp) with the reference's diagonal cut-off.  FP32 mode prescales positions so that
// the Gaussian is a bare v_exp_f32 of -r'^2, accumulates each run of 32 or 64
// sources in FP32 and folds it into a Kahan-compensated (sum, comp) pair;
// partials leave the kernel as double(sum) - double(comp).  Segments with a
// processing order (spatially compact tiles) evaluate the FP32 Gaussian exponent
// of compact blocks as 2<x, y> - |y|^2 - |x|^2 in tile coordinates.

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <stdlib.h>

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>
#include <thread>
#include <atomic>

#include "lddmm.h"

namespace {

constexpr int BLOCK = 256;       // threads per pair-sum workgroup
constexpr int TS = 256;          // sources per LDS tile (= BLOCK: one source per thread)
constexpr int KF = 32;           // FP32 Kahan fold interval (sources)
constexpr int FIN_BLOCK = 64;    // rows per finalize workgroup
constexpr int TOT_BLOCK = 256;   // threads of the totals kernel
// Chunking target per segment (work items = row tiles x source chunks).  The K-sum kernels
// hold 4 waves per SIMD, so the chip takes ~1024 workgroups at a time; large sums need many
// more items than that for the last ones to spread evenly over the CUs.  Measured on the
// 40,962-point flow step (4,374 items) and the 81,920-face FP64 currents energy (6,400
// items): 8% and 9% faster than with a 2,048 target (1,863 and 2,560 items).
constexpr int64_t ITEMS_PER_SEGMENT = 6144;
// Energy-only sums (no per-row outputs) evaluate each unordered pair of self tiles once:
// the workgroups of a row tile skip every source tile below it, so their items differ widely
// in size and balance needs more of them.  The 81,920-face FP64 currents energy: 3.17 ms at
// 6,144 items per segment, 2.94 at 12,288, 2.81 at 24,576.
constexpr int64_t ITEMS_ENERGY = 24576;
constexpr int MAX_NRED = 2;      // FP64 totals per segment

// Chunking targets set at run time (lddmm_set_items_target), per operation, precision,
// energy-only form and size class of the segment's rows; 0 = the built-in value above.
constexpr int N_OPS = LDDMM_OP_DIFF_K + 1;
int64_t g_items_target[N_OPS][2][2][LDDMM_SIZE_CLASSES] = {};

int size_class(int64_t ny) {
  int c = 0;
  for (int64_t lim = 1024; c < LDDMM_SIZE_CLASSES - 1 && ny >= lim; lim *= 4) ++c;
  return c;
}

int64_t items_target(int op, int prec, bool energy_only, int64_t ny) {
  const int64_t t = g_items_target[op][prec == LDDMM_PREC_FP32C][energy_only][size_class(ny)];
  return t > 0 ? t : (energy_only ? ITEMS_ENERGY : ITEMS_PER_SEGMENT);
}

constexpr double LOG2E = 1.4426950408889634074;
// The reference's cut-off literal 1e-8 is single precision widened to double.
constexpr double CUTOFF = (double)1e-8f;

int g_last_hip_error = 0;

template <typename R>
constexpr int rows_per_thread() { return sizeof(R) == 4 ? 2 : 1; }

int rows_per_item(int prec) {
  return BLOCK * (prec == LDDMM_PREC_FP32C ? rows_per_thread<float>()
                                           : rows_per_thread<double>());
}

// Laplacian polynomial tables, float32 literals of kernelFunctions.f90 widened
// to double.  P_o: c_(o+1,:) (:25-26); Q_o: c1_(o,:) (:85-86); R_o: c2_(o-1,:) (:461).
__constant__ double LAP_P[5][5] = {
    {1.0, 0.0, 0.0, 0.0, 0.0},
    {1.0, 1.0, 0.0, 0.0, 0.0},
    {1.0, 1.0, (double)(1.0f / 3.0f), 0.0, 0.0},
    {1.0, 1.0, (double)0.4f, (double)(1.0f / 15.0f), 0.0},
    {1.0, 1.0, (double)(3.0f / 7.0f), (double)(2.0f / 21.0f), (double)(1.0f / 105.0f)}};
__constant__ double LAP_Q[5][4] = {
    {0.0, 0.0, 0.0, 0.0},  // order 0 has no first derivative (rejected on the host)
    {1.0, 0.0, 0.0, 0.0},
    {(double)(1.0f / 3.0f), (double)(1.0f / 3.0f), 0.0, 0.0},
    {(double)(1.0f / 5.0f), (double)(1.0f / 5.0f), (double)(1.0f / 15.0f), 0.0},
    {(double)(1.0f / 7.0f), (double)(1.0f / 7.0f), (double)(2.0f / 35.0f), (double)(1.0f / 105.0f)}};
__constant__ double LAP_R[5][3] = {
    {0.0, 0.0, 0.0},  // orders 0, 1 have no second derivative (rejected on the host)
    {0.0, 0.0, 0.0},
    {(double)(1.0f / 3.0f), 0.0, 0.0},
    {(double)(1.0f / 15.0f), (double)(1.0f / 15.0f), 0.0},
    {(double)(1.0f / 35.0f), (double)(1.0f / 35.0f), (double)(1.0f / 105.0f)}};

// Coordinate prescale s applied to positions before the pair loop.
//   FP32 Gaussian:  s = sqrt(log2(e)/2)/sigma, so K = exp2(-|s dx|^2)
//   FP32 Laplacian: s = log2(e)/sigma, so |s dx| = log2(e) r/sigma and e^-u = exp2(-|s dx|)
//   FP64:           s = 1 (positions untouched, the kernel scales r itself)
// Folding the constants into the FP64 prescale keeps FP32-rounded constants out
// of the exponents (a rounded log2(e) would bias every kernel value by a fixed
// relative amount proportional to the exponent).
__host__ __device__ inline double coord_scale(int prec, int kernel, double sigma) {
  if (prec != LDDMM_PREC_FP32C) return 1.0;
  return kernel == LDDMM_KERNEL_GAUSS ? sqrt(0.5 * LOG2E) / sigma : LOG2E / sigma;
}

// ---------------------------------------------------------------------------
// Pair evaluation: k0 = K, k1 = -2 sigma^2 K'(r^2), k2 = 4 sigma^4 K''(r^2).
// ---------------------------------------------------------------------------
template <typename R, int KT>
struct KEval;

// R2_0: start value of the squared-distance accumulation in the pair loops.
template <>
struct KEval<float, LDDMM_KERNEL_GAUSS> {
  static constexpr float R2_0 = 0.0f;
  float dummy;
  __device__ void init(int, double) {}
  __device__ inline void operator()(float r2, float &k0, float &k1, float &k2) const {
    float e = __builtin_amdgcn_exp2f(-r2);
    k0 = e;
    k1 = e;
    k2 = e;
  }
};

template <>
struct KEval<float, LDDMM_KERNEL_LAPLACIAN> {
  static constexpr float R2_0 = 1e-30f;
  static constexpr float LN2F = 0.69314718055994530942f;
  float p0, p1, p2, p3, p4, q0, q1, q2, q3, c0, c1, c2;
  __device__ void init(int order, double) {
    p0 = (float)LAP_P[order][0]; p1 = (float)LAP_P[order][1]; p2 = (float)LAP_P[order][2];
    p3 = (float)LAP_P[order][3]; p4 = (float)LAP_P[order][4];
    q0 = (float)LAP_Q[order][0]; q1 = (float)LAP_Q[order][1]; q2 = (float)LAP_Q[order][2];
    q3 = (float)LAP_Q[order][3];
    c0 = (float)LAP_R[order][0]; c1 = (float)LAP_R[order][1]; c2 = (float)LAP_R[order][2];
  }
  // r2 = (log2(e) r / sigma)^2, at least R2_0 > 0.  The distance v = sqrt(r2) is
  // v_rsq_f32 followed by one Newton correction with an FMA residual: unbiased
  // and within one ulp.  (v_sqrt_f32 alone is low by 1.5e-8 on average on
  // gfx950; such a bias shifts every kernel value the same way and accumulates
  // along a flow.)  R2_0 keeps rsq finite on coincident points.
  __device__ inline void operator()(float r2, float &k0, float &k1, float &k2) const {
    const float y = __builtin_amdgcn_rsqf(r2);
    float v = r2 * y;
    const float t = fmaf(-v, v, r2);
    v = fmaf(t * y, 0.5f, v);
    const float e = __builtin_amdgcn_exp2f(-v);
    const float u = v * LN2F;
    k0 = fmaf(fmaf(fmaf(fmaf(p4, u, p3), u, p2), u, p1), u, p0) * e;
    k1 = fmaf(fmaf(fmaf(q3, u, q2), u, q1), u, q0) * e;
    k2 = fmaf(fmaf(c2, u, c1), u, c0) * e;
  }
};

// 2^t in double for t <= 0, written for the pair loops (all VALU, no library call):
// t = n + f with n = rint(t) (exact: the 1.5 * 2^52 shift rounds to an integer), |f| <= 1/2;
// 2^f = sum_k (f ln 2)^k / k! truncated after degree 13 (remainder < 4e-18 on |f| <= 1/2,
// Horner rounding ~2 ulp); 2^n is added to the exponent field.  t is clamped at -1000
// (2^-1000 ~ 9e-302), which keeps every result a normal double; a NaN t gives NaN.  (Estrin's form, depth 4
// instead of 13 for three more multiplies, measured 6-9% slower in the FP64 energy sums:
// they are bound by instruction issue, not by this chain's latency.)
struct Exp2Coef {
  double c[14];
  constexpr Exp2Coef() : c() {
    const double ln2 = 0.69314718055994530942;
    double v = 1.0;
    for (int k = 0; k < 14; ++k) {
      c[k] = v;
      v = v * ln2 / (double)(k + 1);
    }
  }
};
constexpr Exp2Coef EXP2C{};

__device__ inline double exp2_nonpos(double t) {
  t = t < -1000.0 ? -1000.0 : t;  // a select, not fmax: a NaN exponent stays NaN
  const double shifter = 6755399441055744.0;  // 1.5 * 2^52
  const double k = t + shifter;
  const double n = k - shifter;
  const double f = t - n;
  double p = EXP2C.c[13];
#pragma unroll
  for (int i = 12; i >= 0; --i) p = fma(p, f, EXP2C.c[i]);
  const int ni = __double2loint(k);
  return __hiloint2double(__double2hiint(p) + (ni << 20), __double2loint(p));
}

// 2^(j/256), j = 0..255, each the double nearest to the exact value (generated with 200-bit
// arithmetic; every entry lies at least 0.0035 ulp from a rounding midpoint).
__constant__ double EXP2_TAB256[256] = {
    0x1.0000000000000p+0, 0x1.00b1afa5abcbfp+0, 0x1.0163da9fb3335p+0, 0x1.02168143b0281p+0,
    0x1.02c9a3e778061p+0, 0x1.037d42e11bbccp+0, 0x1.04315e86e7f85p+0, 0x1.04e5f72f654b1p+0,
    0x1.059b0d3158574p+0, 0x1.0650a0e3c1f89p+0, 0x1.0706b29ddf6dep+0, 0x1.07bd42b72a836p+0,
    0x1.0874518759bc8p+0, 0x1.092bdf66607e0p+0, 0x1.09e3ecac6f383p+0, 0x1.0a9c79b1f3919p+0,
    0x1.0b5586cf9890fp+0, 0x1.0c0f145e46c85p+0, 0x1.0cc922b7247f7p+0, 0x1.0d83b23395decp+0,
    0x1.0e3ec32d3d1a2p+0, 0x1.0efa55fdfa9c5p+0, 0x1.0fb66affed31bp+0, 0x1.1073028d7233ep+0,
    0x1.11301d0125b51p+0, 0x1.11edbab5e2ab6p+0, 0x1.12abdc06c31ccp+0, 0x1.136a814f204abp+0,
    0x1.1429aaea92de0p+0, 0x1.14e95934f312ep+0, 0x1.15a98c8a58e51p+0, 0x1.166a45471c3c2p+0,
    0x1.172b83c7d517bp+0, 0x1.17ed48695bbc0p+0, 0x1.18af9388c8deap+0, 0x1.1972658375d2fp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1af99f8138a1cp+0, 0x1.1bbe084045cd4p+0, 0x1.1c82f95281c6bp+0,
    0x1.1d4873168b9aap+0, 0x1.1e0e75eb44027p+0, 0x1.1ed5022fcd91dp+0, 0x1.1f9c18438ce4dp+0,
    0x1.2063b88628cd6p+0, 0x1.212be3578a819p+0, 0x1.21f49917ddc96p+0, 0x1.22bdda27912d1p+0,
    0x1.2387a6e756238p+0, 0x1.2451ffb82140ap+0, 0x1.251ce4fb2a63fp+0, 0x1.25e85711ece75p+0,
    0x1.26b4565e27cddp+0, 0x1.2780e341ddf29p+0, 0x1.284dfe1f56381p+0, 0x1.291ba7591bb70p+0,
    0x1.29e9df51fdee1p+0, 0x1.2ab8a66d10f13p+0, 0x1.2b87fd0dad990p+0, 0x1.2c57e39771b2fp+0,
    0x1.2d285a6e4030bp+0, 0x1.2df961f641589p+0, 0x1.2ecafa93e2f56p+0, 0x1.2f9d24abd886bp+0,
    0x1.306fe0a31b715p+0, 0x1.31432edeeb2fdp+0, 0x1.32170fc4cd831p+0, 0x1.32eb83ba8ea32p+0,
    0x1.33c08b26416ffp+0, 0x1.3496266e3fa2dp+0, 0x1.356c55f929ff1p+0, 0x1.36431a2de883bp+0,
    0x1.371a7373aa9cbp+0, 0x1.37f26231e754ap+0, 0x1.38cae6d05d866p+0, 0x1.39a401b7140efp+0,
    0x1.3a7db34e59ff7p+0, 0x1.3b57fbfec6cf4p+0, 0x1.3c32dc313a8e5p+0, 0x1.3d0e544ede173p+0,
    0x1.3dea64c123422p+0, 0x1.3ec70df1c5175p+0, 0x1.3fa4504ac801cp+0, 0x1.40822c367a024p+0,
    0x1.4160a21f72e2ap+0, 0x1.423fb2709468ap+0, 0x1.431f5d950a897p+0, 0x1.43ffa3f84b9d4p+0,
    0x1.44e086061892dp+0, 0x1.45c2042a7d232p+0, 0x1.46a41ed1d0057p+0, 0x1.4786d668b3237p+0,
    0x1.486a2b5c13cd0p+0, 0x1.494e1e192aed2p+0, 0x1.4a32af0d7d3dep+0, 0x1.4b17dea6db7d7p+0,
    0x1.4bfdad5362a27p+0, 0x1.4ce41b817c114p+0, 0x1.4dcb299fddd0dp+0, 0x1.4eb2d81d8abffp+0,
    0x1.4f9b2769d2ca7p+0, 0x1.508417f4531eep+0, 0x1.516daa2cf6642p+0, 0x1.5257de83f4eefp+0,
    0x1.5342b569d4f82p+0, 0x1.542e2f4f6ad27p+0, 0x1.551a4ca5d920fp+0, 0x1.56070dde910d2p+0,
    0x1.56f4736b527dap+0, 0x1.57e27dbe2c4cfp+0, 0x1.58d12d497c7fdp+0, 0x1.59c0827ff07ccp+0,
    0x1.5ab07dd485429p+0, 0x1.5ba11fba87a03p+0, 0x1.5c9268a5946b7p+0, 0x1.5d84590998b93p+0,
    0x1.5e76f15ad2148p+0, 0x1.5f6a320dceb71p+0, 0x1.605e1b976dc09p+0, 0x1.6152ae6cdf6f4p+0,
    0x1.6247eb03a5585p+0, 0x1.633dd1d1929fdp+0, 0x1.6434634ccc320p+0, 0x1.652b9febc8fb7p+0,
    0x1.6623882552225p+0, 0x1.671c1c70833f6p+0, 0x1.68155d44ca973p+0, 0x1.690f4b19e9538p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6b052fa75173ep+0, 0x1.6c012750bdabfp+0, 0x1.6cfdcddd47645p+0,
    0x1.6dfb23c651a2fp+0, 0x1.6ef9298593ae5p+0, 0x1.6ff7df9519484p+0, 0x1.70f7466f42e87p+0,
    0x1.71f75e8ec5f74p+0, 0x1.72f8286ead08ap+0, 0x1.73f9a48a58174p+0, 0x1.74fbd35d7cbfdp+0,
    0x1.75feb564267c9p+0, 0x1.77024b1ab6e09p+0, 0x1.780694fde5d3fp+0, 0x1.790b938ac1cf6p+0,
    0x1.7a11473eb0187p+0, 0x1.7b17b0976cfdbp+0, 0x1.7c1ed0130c132p+0, 0x1.7d26a62ff86f0p+0,
    0x1.7e2f336cf4e62p+0, 0x1.7f3878491c491p+0, 0x1.80427543e1a12p+0, 0x1.814d2add106d9p+0,
    0x1.82589994cce13p+0, 0x1.8364c1eb941f7p+0, 0x1.8471a4623c7adp+0, 0x1.857f4179f5b21p+0,
    0x1.868d99b4492edp+0, 0x1.879cad931a436p+0, 0x1.88ac7d98a6699p+0, 0x1.89bd0a478580fp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8be05bad61778p+0, 0x1.8cf3216b5448cp+0, 0x1.8e06a5e0866d9p+0,
    0x1.8f1ae99157736p+0, 0x1.902fed0282c8ap+0, 0x1.9145b0b91ffc6p+0, 0x1.925c353aa2fe2p+0,
    0x1.93737b0cdc5e5p+0, 0x1.948b82b5f98e5p+0, 0x1.95a44cbc8520fp+0, 0x1.96bdd9a7670b3p+0,
    0x1.97d829fde4e50p+0, 0x1.98f33e47a22a2p+0, 0x1.9a0f170ca07bap+0, 0x1.9b2bb4d53fe0dp+0,
    0x1.9c49182a3f090p+0, 0x1.9d674194bb8d5p+0, 0x1.9e86319e32323p+0, 0x1.9fa5e8d07f29ep+0,
    0x1.a0c667b5de565p+0, 0x1.a1e7aed8eb8bbp+0, 0x1.a309bec4a2d33p+0, 0x1.a42c980460ad8p+0,
    0x1.a5503b23e255dp+0, 0x1.a674a8af46052p+0, 0x1.a799e1330b358p+0, 0x1.a8bfe53c12e59p+0,
    0x1.a9e6b5579fdbfp+0, 0x1.ab0e521356ebap+0, 0x1.ac36bbfd3f37ap+0, 0x1.ad5ff3a3c2774p+0,
    0x1.ae89f995ad3adp+0, 0x1.afb4ce622f2ffp+0, 0x1.b0e07298db666p+0, 0x1.b20ce6c9a8952p+0,
    0x1.b33a2b84f15fbp+0, 0x1.b468415b749b1p+0, 0x1.b59728de5593ap+0, 0x1.b6c6e29f1c52ap+0,
    0x1.b7f76f2fb5e47p+0, 0x1.b928cf22749e4p+0, 0x1.ba5b030a1064ap+0, 0x1.bb8e0b79a6f1fp+0,
    0x1.bcc1e904bc1d2p+0, 0x1.bdf69c3f3a207p+0, 0x1.bf2c25bd71e09p+0, 0x1.c06286141b33dp+0,
    0x1.c199bdd85529cp+0, 0x1.c2d1cd9fa652cp+0, 0x1.c40ab5fffd07ap+0, 0x1.c544778fafb22p+0,
    0x1.c67f12e57d14bp+0, 0x1.c7ba88988c933p+0, 0x1.c8f6d9406e7b5p+0, 0x1.ca3405751c4dbp+0,
    0x1.cb720dcef9069p+0, 0x1.ccb0f2e6d1675p+0, 0x1.cdf0b555dc3fap+0, 0x1.cf3155b5bab74p+0,
    0x1.d072d4a07897cp+0, 0x1.d1b532b08c968p+0, 0x1.d2f87080d89f2p+0, 0x1.d43c8eacaa1d6p+0,
    0x1.d5818dcfba487p+0, 0x1.d6c76e862e6d3p+0, 0x1.d80e316c98398p+0, 0x1.d955d71ff6075p+0,
    0x1.da9e603db3285p+0, 0x1.dbe7cd63a8315p+0, 0x1.dd321f301b460p+0, 0x1.de7d5641c0658p+0,
    0x1.dfc97337b9b5fp+0, 0x1.e11676b197d17p+0, 0x1.e264614f5a129p+0, 0x1.e3b333b16ee12p+0,
    0x1.e502ee78b3ff6p+0, 0x1.e653924676d76p+0, 0x1.e7a51fbc74c83p+0, 0x1.e8f7977cdb740p+0,
    0x1.ea4afa2a490dap+0, 0x1.eb9f4867cca6ep+0, 0x1.ecf482d8e67f1p+0, 0x1.ee4aaa2188510p+0,
    0x1.efa1bee615a27p+0, 0x1.f0f9c1cb6412ap+0, 0x1.f252b376bba97p+0, 0x1.f3ac948dd7274p+0,
    0x1.f50765b6e4540p+0, 0x1.f6632798844f8p+0, 0x1.f7bfdad9cbe14p+0, 0x1.f91d802243c89p+0,
    0x1.fa7c1819e90d8p+0, 0x1.fbdba3692d514p+0, 0x1.fd3c22b8f71f1p+0, 0x1.fe9d96b2a23d9p+0,
};

// (ln 2 / 256)^k / k!, k = 0..4.
struct Exp2T256Coef {
  double c[5];
  constexpr Exp2T256Coef() : c() {
    const double ln2 = 0.69314718055994530942;
    double v = 1.0;
    for (int k = 0; k < 5; ++k) {
      c[k] = v;
      v = v * ln2 / 256.0 / (double)(k + 1);
    }
  }
};
constexpr Exp2T256Coef EXP2T256C{};

// 2^(t / 256) in double for t <= 0, with the 256-entry table (in LDS) of 2^(j/256):
// t = 256 m + j + g with rint (the 1.5 * 2^52 shift), j = 0..255 and |g| <= 1/2;
// 2^(g/256) - 1 = q = sum_{k>=1} (g ln 2 / 256)^k / k! truncated after degree 4 (remainder
// < 4e-17 relative: |g ln 2 / 256| < 0.0014), the result ldexp(fma(T[j], q, T[j]), m).  4 FMA
// + 1 multiply instead of exp2_nonpos's 13 FMA.  The integer m must fit 32 bits: CLAMP
// clamps t near -256000 (2^-1000) by selecting the high word of -256000 (the low word
// stays, so the clamped t is -256000 within 2^-20): one select that keeps a NaN t NaN,
// where fmax would turn it into the clamp value.  Without CLAMP the caller guarantees
// t >= -2^30 (the result then underflows correctly through ldexp).
template <bool CLAMP>
__device__ inline double exp2_tab256(double t, const double *__restrict__ tab) {
  if constexpr (CLAMP)
    t = __hiloint2double(t < -256000.0 ? (int)0xC10F4000 : __double2hiint(t), __double2loint(t));
  const double shifter = 6755399441055744.0;  // 1.5 * 2^52
  const double k = t + shifter;
  // The table read first: its LDS latency overlaps the polynomial.
  const int ni = __double2loint(k);
  const double tj = tab[ni & 255];
  const double n = k - shifter;
  const double g = t - n;
  double q = EXP2T256C.c[4];
#pragma unroll
  for (int i = 3; i >= 1; --i) q = fma(q, g, EXP2T256C.c[i]);
  q *= g;
  const double r = fma(tj, q, tj);
  return ldexp(r, ni >> 8);  // exact: r in [1, 2), the result a normal double
}

template <>
struct KEval<double, LDDMM_KERNEL_GAUSS> {
  static constexpr double R2_0 = 0.0;
  double h, h256, cut_r2;
  __device__ void init(int, double sigma) {
    h = LOG2E / (2.0 * sigma * sigma);  // exp(-r^2/(2 sigma^2)) = 2^(-h r^2)
    h256 = 256.0 * h;                   // exact
    cut_r2 = CUTOFF * sigma * sigma;    // (r/sigma)^2 < 1e-8  <=>  r^2 < 1e-8 sigma^2
  }
  __device__ inline void operator()(double r2, double &k0, double &k1, double &k2) const {
    double e = exp2_nonpos(-h * r2);
    double k = r2 < cut_r2 ? 1.0 : e;
    k0 = k;
    k1 = k;
    k2 = k;
  }
  // The same with the LDS table form of the exponential (tab: EXP2_TAB256 in LDS).  The
  // cut-off zeroes the high word of r^2 instead of selecting the result: r^2 then is below
  // 2^-1042, and exp2_tab256 of -h256 times it is exactly 1 (q rounds away in fma(1, q, 1)).
  // CLAMP: see exp2_tab256.
  template <bool CLAMP>
  __device__ inline void eval_tab(double r2, const double *__restrict__ tab, double &k0,
                                  double &k1, double &k2) const {
    const double r2c = __hiloint2double(r2 < cut_r2 ? 0 : __double2hiint(r2), __double2loint(r2));
    const double k = exp2_tab256<CLAMP>(-h256 * r2c, tab);
    k0 = k;
    k1 = k;
    k2 = k;
  }
};

template <>
struct KEval<double, LDDMM_KERNEL_LAPLACIAN> {
  static constexpr double R2_0 = 0.0;
  double inv_s, p0, p1, p2, p3, p4, q0, q1, q2, q3, c0, c1, c2;
  __device__ void init(int order, double sigma) {
    inv_s = 1.0 / sigma;
    p0 = LAP_P[order][0]; p1 = LAP_P[order][1]; p2 = LAP_P[order][2];
    p3 = LAP_P[order][3]; p4 = LAP_P[order][4];
    q0 = LAP_Q[order][0]; q1 = LAP_Q[order][1]; q2 = LAP_Q[order][2]; q3 = LAP_Q[order][3];
    c0 = LAP_R[order][0]; c1 = LAP_R[order][1]; c2 = LAP_R[order][2];
  }
  __device__ inline void operator()(double r2, double &k0, double &k1, double &k2) const {
    double u = sqrt(r2) * inv_s;
    double e = exp2_nonpos(-LOG2E * u);
    bool cut = u < CUTOFF;
    double pk = (((p4 * u + p3) * u + p2) * u + p1) * u + p0;
    double qk = ((q3 * u + q2) * u + q1) * u + q0;
    double rk = (c2 * u + c1) * u + c0;
    k0 = cut ? 1.0 : pk * e;
    k1 = cut ? q0 : qk * e;
    k2 = cut ? c0 : rk * e;
  }
};

// ---------------------------------------------------------------------------
// Tile culling.  A wave skips a staged source tile when the distance between
// the bounding box of its target rows and the tile's bounding box makes every
// kernel value of the tile smaller than EPS_K (2^-24 in FP32, 2^-56 in FP64).
// A skipped term is below half (FP32) or 1/8 (FP64) of an ulp of a unit
// kernel value; even all of a row's skipped terms together (each below
// EPS_K |w|) stay within the rounding error of its kept near-field terms, and
// in an FP64 data-term energy of F faces they move the total by at most
// F^2 EPS_K max|w|^2 (1.5e-6 max|w|^2 at 327,680 faces, whose self-energy alone
// is of order F |w|^2).  For the Laplacian
// the bound uses P, Q, R <= (1+u)^4, and (1+u)^4 e^-u < EPS_K for u > U_CUT.
// Returns the squared cut radius in the kernel's coordinates (prescaled in
// FP32, raw in FP64).
// ---------------------------------------------------------------------------
__host__ __device__ inline double cull_r2(int prec, int kernel, double sigma) {
  const bool f32 = prec == LDDMM_PREC_FP32C;
  if (kernel == LDDMM_KERNEL_GAUSS) {
    // FP32: K = 2^-r'^2 with prescaled r'.  FP64: K = 2^(-h r^2), h = log2(e)/(2 sigma^2).
    // The FP32 bound, 2^-24, is FP32's unit roundoff: a source beyond it adds less than half
    // an ulp of its own weight (on the large surface pair the flow step runs 13% faster than
    // at 2^-32 for the same energy error, 2.4e-8).
    // FP64E (the energies of FP32C registrations): 2^-40; see far_r2.
    if (f32) return 24.0;
    return (prec == LDDMM_PREC_FP64E ? 40.0 : 56.0) * 2.0 * sigma * sigma / LOG2E;
  }
  // (1+u)^4 e^-u = 2^-24 at u~30.4, 2^-40 at u~42.8, 2^-56 at u~54.9; FP32 coordinates carry
  // log2(e)/sigma.
  const double u_cut =
      f32 ? 30.5 * LOG2E : (prec == LDDMM_PREC_FP64E ? 43.0 : 55.0) * sigma;
  return u_cut * u_cut;
}

// Mixed precision by magnitude (FP64 Gaussian energy sums): blocks whose box distance puts
// every kernel value below 2^-26 (FP64) or 2^-16 (FP64E) take the FP32 exponential of the
// FP64 exponent, a relative error of at most ~2e-6 per term.  FP64E is FP32C's energy form:
// with the cull at 2^-40 it keeps the data term of the 81,920-face surface pair within
// 2.2e-10 of the FP64 value (measured at a data term 6.4e-7 of the single sums, where the
// culled terms of the shape's self sum and of the cross sum cancel), for 24% less time.
__host__ __device__ inline double far_r2(int prec, double sigma) {
  return (prec == LDDMM_PREC_FP64E ? 16.0 : 26.0) * 2.0 * sigma * sigma / LOG2E;
}

// Dot-product exponent (FP32 Gaussian K-sums of segments with a processing order).
// In tile-origin coordinates, -r'^2 = (2 x.y - |y|^2) - |x|^2: the pair loop seeds the
// exponent with -|y|^2 (staged per source, next to 2y in place of y) and adds 2 x.y with
// three FMAs, and the row factor 2^-|x|^2 multiplies each 32-source run at its fold, which
// saves the three differences and the square of the direct form.  Its rounding error grows
// with the offsets from the origin, so a 64-source quarter takes it only when every source
// lies within R'^2 <= EXP_R2MAX of the origin.  Rows with |x|'^2 > EXP_X2MAX drop their runs
// over such a quarter at the fold: by the triangle inequality their r'^2 exceeds
// (sqrt(60) - sqrt(2))^2 > 41, so every kernel value is below 2^-41, under the culling bound.  For
// the other rows 2^(2x.y - |y|^2) <= 2^|x|^2 <= 2^60, which cannot overflow for weights below
// 2^60; a dropped row's run may overflow, and the fold discards it.
constexpr float EXP_R2MAX = 2.0f;
constexpr float EXP_X2MAX = 60.0f;

template <typename R>
__device__ inline R wave_min(R v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename R>
__device__ inline R wave_max(R v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
// A wave-uniform double (every lane holds the same value, e.g. after a wave reduction) moved
// to scalar registers, which frees the two VGPRs the compiler would otherwise keep it in.
__device__ inline double wave_uniform(double v) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}

// ---------------------------------------------------------------------------
// Operand resolution per operation (uniform per workgroup).
// ---------------------------------------------------------------------------
struct SrcView {
  const double *pos;
  const double *w[4];
  int64_t n;
  double sign;
};

struct TgtView {
  const double *pos;
  const double *w[4];
  const double *w0b;  // tw[0] += tcoef * w0b[k] on load (covector)
  double tcoef;
};

__device__ inline const double *or_y(const double *p, const lddmm_seg_t *s) {
  return p ? p : s->y;
}

__device__ inline SrcView resolve_src(int op, const lddmm_seg_t *s, int set) {
  SrcView v;
  v.w[0] = v.w[1] = v.w[2] = v.w[3] = nullptr;
  v.sign = 1.0;
  v.n = s->nx0;
  switch (op) {
    case LDDMM_OP_APPLY_K:
      v.pos = or_y(s->x0, s);
      v.w[0] = s->w0;
      break;
    case LDDMM_OP_DIFF_KT:
      v.pos = or_y(s->x0, s);
      for (int j = 0; j < 4; ++j) v.w[j] = s->sw[j];
      break;
    case LDDMM_OP_COVECTOR_STEP:  // a2 = [a, p]
      v.pos = s->y;
      v.w[0] = s->w0;
      v.w[1] = s->w1;
      break;
    case LDDMM_OP_CURRENTS:  // set 0: (c1, cr1); set 1: (c2, -cr2)
    case LDDMM_OP_MEASURE:   // set 0: (c1, a1);  set 1: (c2, -a2)
      if (set == 0) {
        v.pos = s->y;
        v.w[0] = s->w0;
      } else {
        v.pos = s->x1;
        v.w[0] = s->w1;
        v.n = s->nx1;
        v.sign = -1.0;
      }
      break;
    default:  // FLOW_STEP, KDOT, KNORM: targets = sources = y, weights w0
      v.pos = s->y;
      v.w[0] = s->w0;
      break;
  }
  return v;
}

__device__ inline TgtView resolve_tgt(int op, const lddmm_seg_t *s) {
  TgtView v;
  v.pos = s->y;
  v.w[0] = v.w[1] = v.w[2] = v.w[3] = nullptr;
  v.w0b = nullptr;
  v.tcoef = 0.0;
  switch (op) {
    case LDDMM_OP_DIFF_KT:
      for (int j = 0; j < 4; ++j) v.w[j] = s->tw[j];
      break;
    case LDDMM_OP_COVECTOR_STEP:  // a1 = [p - 2 lambda a, a]
      v.w[0] = s->w1;
      v.w0b = s->w0;
      v.tcoef = -2.0 * s->coef;
      v.w[1] = s->w0;
      break;
    case LDDMM_OP_CURRENTS:  // a1 = [cr1]
    case LDDMM_OP_MEASURE:   // a1 = [a1]
      v.w[0] = s->w0;
      break;
    default:
      break;
  }
  return v;
}

// Row stride of the weight / output arrays of an operation.
__host__ __device__ inline int weight_ld(int op, const lddmm_seg_t *s, int d) {
  if (op == LDDMM_OP_MEASURE) return 1;
  return (op == LDDMM_OP_APPLY_K || op == LDDMM_OP_DIFF_KT) ? s->m : d;
}

__device__ inline int find_segment(const lddmm_seg_t *__restrict__ segs, int nseg,
                                   int64_t key, int which) {
  // Largest i with start(i) <= key; which: 0 -> item0, 1 -> fin0, 2 -> stage0.
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    int64_t st = which == 0 ? segs[mid].item0 : (which == 1 ? segs[mid].fin0 : segs[mid].stage0);
    if (st <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// The segment-local work item (row tile, source chunk) that workgroup `item` of the launch
// runs: its own position in the segment, or what the segment's dispatch order puts there.
__device__ inline int64_t local_item(const lddmm_seg_t *sg, int64_t item) {
  const int64_t j = item - sg->item0;
  return sg->item_order ? (int64_t)sg->item_order[j] : j;
}

// ---------------------------------------------------------------------------
// Source tiles in LDS form.  A tile holds TS consecutive sources of one set, in the
// segment's processing order for set 0 when it has one.  FP32 evaluates each tile in a
// frame centred at the tile (origin below), so the staged positions depend only on the
// tile, not on the rows that read it.  k_pairs builds a tile in LDS itself, or, for
// ordered segments, copies one that k_stage built once per launch in the workspace
// (the same values: both paths call the functions below), so the gathers through the
// order, the FP64 reads and conversions and the box reductions run once per tile instead
// of once per row tile.
// ---------------------------------------------------------------------------

// FP64 Gaussian K-sums without derivative terms (applyK, flow steps, RKHS products and
// norms, data-term energies) evaluate blocks of pairs near the tile origin with the expanded
// exponent -h r^2 = -h|x'|^2 - h|y'|^2 + 2h <x', y'> in tile coordinates (see k_pairs).
template <typename R, int KT, bool KSUM, int NA>
constexpr bool xform_kernel() {
  return std::is_same<R, double>::value && KT == LDDMM_KERNEL_GAUSS && KSUM && NA == 0;
}

// Tile origin: FP32 mode evaluates the tile in a frame centred at its middle, so rounding
// to FP32 costs relative precision of the local distances, not of the absolute
// coordinates; so does the FP64 expanded exponent, whose rounding error grows with the
// offsets from the origin.  The origin is the mean of the tile's four quarter-middle
// sources (sources 32, 96, 160, 224, clamped to the tile), close to the centre of the
// tile's extent for spatially ordered sources.  Other FP64 sums keep the raw frame (origin 0).
template <typename R, int D, bool CENTRE>
__device__ inline void tile_origin(const SrcView &src, const int32_t *__restrict__ sperm,
                                   int64_t base, int64_t tlen, double scale, double org[D]) {
#pragma unroll
  for (int c = 0; c < D; ++c) org[c] = 0.0;
  if constexpr (CENTRE) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t ls = base + (32 + 64 * q < tlen ? 32 + 64 * q : tlen - 1);
      if (sperm) ls = sperm[ls];
#pragma unroll
      for (int c = 0; c < D; ++c) org[c] += src.pos[ls * D + c];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) org[c] *= 0.25 * scale;
  }
}

// One source of a tile, as staged: tile-frame position (doubled in a compact quarter of
// the dot-product form), signed weights, -|v|^2, and the wave-reduced box and largest
// |v|^2 of its 64-source quarter.
template <typename R, int D, int M, int NSW>
struct SrcStage {
  R v[D];
  R w[NSW > 0 ? NSW : 1][M];
  R n2;
  R lo[D], hi[D];
  R q2;
};

// Stage source l (< l1: a real source; else zero-weight padding with an empty box).  Every
// lane of the wave calls it (wave reductions).  DOT: the dot-product form is possible for
// this segment (its quarter is staged doubled when max |v|^2 <= EXP_R2MAX).  QMAX: the
// quarter's max |v|^2 is wanted without the dot-product form (the FP64 expanded exponent).
template <typename R, int D, int M, int NSW>
__device__ inline void stage_source(const SrcView &src, const int32_t *__restrict__ sperm,
                                    int ld, int64_t l, int64_t l1, double scale,
                                    const double org[D], bool dot, bool qmax,
                                    SrcStage<R, D, M, NSW> &o) {
  R n2 = 0;
  if (l < l1) {
    const int64_t lg = sperm ? (int64_t)sperm[l] : l;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      o.v[c] = (R)(src.pos[lg * D + c] * scale - org[c]);
      o.lo[c] = o.v[c];
      o.hi[c] = o.v[c];
      n2 = o.v[c] * o.v[c] + n2;
    }
#pragma unroll
    for (int j = 0; j < NSW; ++j)
#pragma unroll
      for (int c = 0; c < M; ++c) o.w[j][c] = (R)(src.sign * src.w[j][lg * ld + c]);
  } else {
    // Padding sources carry zero weight: their terms are exact zeros.
#pragma unroll
    for (int c = 0; c < D; ++c) {
      o.v[c] = (R)0;
      o.lo[c] = (R)INFINITY;
      o.hi[c] = (R)-INFINITY;
    }
#pragma unroll
    for (int j = 0; j < NSW; ++j)
#pragma unroll
      for (int c = 0; c < M; ++c) o.w[j][c] = (R)0;
  }
#pragma unroll
  for (int c = 0; c < D; ++c) {
    o.lo[c] = wave_min(o.lo[c]);
    o.hi[c] = wave_max(o.hi[c]);
  }
  o.n2 = -n2;
  o.q2 = 0;
  if (dot || qmax) {
    o.q2 = wave_max(n2);
    if (dot && o.q2 <= (R)EXP_R2MAX) {
#pragma unroll
      for (int c = 0; c < D; ++c) o.v[c] = (R)2 * o.v[c];
    }
  }
}

// Staged tile record in the workspace: origin (D doubles), quarter boxes [4][2D], quarter
// max |v|^2 [4], then positions [D][TS], weights [NSW][M][TS] and -|v|^2 [TS].
constexpr int STAGE_HEAD = 256;
__host__ __device__ inline size_t stage_rec_bytes(int rsize, int d, int m, int nsw) {
  return STAGE_HEAD + (size_t)(d + nsw * m + 1) * TS * rsize;
}

// Source sets of a segment whose tiles are staged once per launch: set 0 of segments with a
// processing order (flow and covector steps, RKHS products); both sets of the energy-only
// sums of k_pairs (data-term energies and RKHS norms without per-row outputs), whose every
// row tile reads every source tile; none otherwise.  na: the launch's derivative pair count.
__host__ __device__ inline int stage_sets(int op, const lddmm_seg_t &s, int na) {
  if (s.perm) return 1;
  if (!s.rowdot && !s.out0 && na == 0 &&
      (op == LDDMM_OP_KNORM || op == LDDMM_OP_CURRENTS || op == LDDMM_OP_MEASURE))
    return 2;
  return 0;
}

// Staged tiles of a planned segment (nx0, nx1 as the plan sets them).
__host__ __device__ inline int64_t stage_tiles(int op, const lddmm_seg_t &s, int na) {
  const int sets = stage_sets(op, s, na);
  if (sets == 0) return 0;
  return (s.nx0 + TS - 1) / TS + (sets == 2 ? (s.nx1 + TS - 1) / TS : 0);
}

// Sum of one segment's per-finalize-block partial reductions (`red`, nred values each) into
// its totals: one workgroup of TOT_BLOCK threads, in a fixed order.
__device__ inline void segment_totals(const lddmm_seg_t *sg, int nred,
                                      const double *__restrict__ red) {
  __shared__ double sh[TOT_BLOCK];
  const int tid = threadIdx.x;
  for (int j = 0; j < nred; ++j) {
    double v = 0.0;
    for (int b = tid; b < sg->fin_blocks; b += TOT_BLOCK) v += red[sg->red0 + (int64_t)j * sg->fin_blocks + b];
    sh[tid] = v;
    __syncthreads();
    for (int s = TOT_BLOCK / 2; s > 0; s >>= 1) {
      if (tid < s) sh[tid] += sh[tid + s];
      __syncthreads();
    }
    if (tid == 0 && sg->total) sg->total[j] = sh[0];
    __syncthreads();
  }
}

// Staged source tiles, one workgroup per tile (n_stage of them).  In a stepped sequence the
// launch also carries the previous step's totals (tot_segs, one more workgroup per segment,
// after the tiles): they read the reductions that step's finalize left, before this step's
// finalize overwrites them, and spare the sequence a totals launch per step.
template <typename R, int KT, int D, int M, bool KSUM, int NA>
__global__ __launch_bounds__(BLOCK) void k_stage(const lddmm_seg_t *__restrict__ segs, int nseg,
                                                int op, int prec, char *__restrict__ stage,
                                                int64_t n_stage,
                                                const lddmm_seg_t *__restrict__ tot_segs,
                                                int nred, const double *__restrict__ red) {
  static_assert(BLOCK == TOT_BLOCK, "the totals workgroups run with the staging ones");
  if ((int64_t)blockIdx.x >= n_stage) {
    segment_totals(tot_segs + (blockIdx.x - n_stage), nred, red);
    return;
  }
  constexpr int NSW = (NA > (KSUM ? 1 : 0)) ? NA : (KSUM ? 1 : 0);
  constexpr bool DOT = std::is_same<R, float>::value && KT == LDDMM_KERNEL_GAUSS && KSUM && NA == 0;
  constexpr bool XFORM = xform_kernel<R, KT, KSUM, NA>();
  const int tid = threadIdx.x;
  const int64_t tile = blockIdx.x;
  const lddmm_seg_t *sg = segs + find_segment(segs, nseg, tile, 2);
  // Tiles of set 0, then (energy-only sums) of set 1.
  const int64_t t = tile - sg->stage0;
  const int64_t tiles0 = (sg->nx0 + TS - 1) / TS;
  const int set = t < tiles0 ? 0 : 1;
  const SrcView src = resolve_src(op, sg, set);
  const int32_t *__restrict__ sperm = set == 0 ? sg->perm : nullptr;
  const double scale = coord_scale(prec, KT, sg->sigma);
  const int64_t base = (set ? t - tiles0 : t) * TS;
  const int64_t tlen = src.n - base < TS ? src.n - base : TS;
  double org[D];
  tile_origin<R, D, std::is_same<R, float>::value || XFORM>(src, sperm, base, tlen, scale, org);
  SrcStage<R, D, M, NSW> ss;
  stage_source<R, D, M, NSW>(src, sperm, weight_ld(op, sg, D), base + tid, src.n, scale, org,
                             DOT && sperm != nullptr, XFORM, ss);
  char *rec = stage + tile * stage_rec_bytes(sizeof(R), D, M, NSW);
  R *rp = (R *)(rec + STAGE_HEAD);
#pragma unroll
  for (int c = 0; c < D; ++c) rp[c * TS + tid] = ss.v[c];
#pragma unroll
  for (int j = 0; j < NSW; ++j)
#pragma unroll
    for (int c = 0; c < M; ++c) rp[(D + j * M + c) * TS + tid] = ss.w[j][c];
  rp[(D + NSW * M) * TS + tid] = ss.n2;
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) {
    R *box = (R *)(rec + 32);
#pragma unroll
    for (int c = 0; c < D; ++c) {
      box[wave * 2 * D + c] = ss.lo[c];
      box[wave * 2 * D + D + c] = ss.hi[c];
    }
    ((R *)(rec + 224))[wave] = ss.q2;
  }
  if (tid < D) ((double *)rec)[tid] = org[tid];
}

// ---------------------------------------------------------------------------
// Pair-sum kernel: K-weighted sums (KSUM) and/or K'-weighted derivative sums
// with NA target/source weight pairs.  Writes W = KSUM*M + (NA>0)*D doubles per
// row and chunk.
// ---------------------------------------------------------------------------
template <typename R, int KT, int D, int M, bool KSUM, int NA>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(xform_kernel<R, KT, KSUM, NA>() ? 4 : 1)))
void k_pairs(const lddmm_seg_t *__restrict__ segs,
                                                int nseg, int op, int prec,
                                                double *__restrict__ part,
                                                const char *__restrict__ stage) {
  constexpr int RPT = rows_per_thread<R>();
  constexpr int NSW = (NA > (KSUM ? 1 : 0)) ? NA : (KSUM ? 1 : 0);
  constexpr int W = (KSUM ? M : 0) + (NA > 0 ? D : 0);
  constexpr bool KAHAN = std::is_same<R, float>::value;
  constexpr int NKAH = KAHAN ? 1 : 0;
  // Sources per unrolled step of the pair loop, measured on gfx950: 4 (8 costs the FP32 flow
  // step 4%, 2 costs it 8%; with the table exponential the FP64 energy is 1.8% faster at 4
  // than at 8), and 2 for sums with two or more derivative weight pairs (the covector step:
  // 126 instead of 154 VGPRs, 4 waves per SIMD instead of 3, 1.2% faster).
  constexpr int PAIR_UNROLL = NA >= 2 ? 2 : 4;
  // Sources per FP32 run before its Kahan fold: 64 for the plain K-sums (one culling
  // quarter), which halves the fold cost of the 32 used elsewhere; measured on the
  // 40,962-point flow step: 2-3% faster, the step's error against FP64 from 1.6e-7 to 2.3e-7
  // of its largest move.  Sums with derivative terms (the currents gradient: a difference of
  // large sums) keep 32.
  constexpr int KFP = (KAHAN && NA == 0) ? 64 : KF;

  // Dot-product exponent: FP32 Gaussian K-sums whose segments carry a processing order.
  constexpr bool EXPAND = KAHAN && KT == LDDMM_KERNEL_GAUSS && KSUM && NA == 0;
  // FP64 Gaussian K-sums: the expanded exponent in tile coordinates,
  //   t = -h256 r^2 = (-h256 |y'|^2 - h256 |x'|^2) + <2 h256 x', y'>   (one add, three FMAs
  // in place of three differences, a square and a scale).  Its rounding error is about
  // eps h256 (|x'| + |y'|)^2, so a (wave, quarter) block takes it only when
  // h (|x'|max + |y'|max)^2 <= XF_HMAX: the exponent's error is then at most 64 eps bits,
  // no more than the direct form's 2 eps h |x - y| |x| in absolute coordinates on a unit
  // sphere, and t >= -256 XF_HMAX needs no clamp.
  constexpr bool XFORM = xform_kernel<R, KT, KSUM, NA>();
  constexpr double XF_HMAX = 64.0;

  __shared__ R s_pos[D][TS];
  __shared__ R s_w[NSW][M][TS];
  __shared__ R s_box[TS / 64][2 * D];  // bounding box of each 64-source quarter
  // -|y|^2 of each source (FP32 dot-product exponent), -h256 |y'|^2 (FP64 expanded exponent)
  __shared__ R s_n2[(EXPAND || XFORM) ? TS : 1];
  __shared__ R s_r2max[TS / 64];       // max |y|^2 of each quarter
  // FP64 Gaussian sums evaluate the exponential with the 2^(j/256) table (exp2_tab256).
  constexpr bool TAB = std::is_same<R, double>::value && KT == LDDMM_KERNEL_GAUSS;
  __shared__ double s_e2[TAB ? 256 : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t item = blockIdx.x;
  const int si = find_segment(segs, nseg, item, 0);
  const lddmm_seg_t *sg = segs + si;

  const int64_t ny = sg->ny;
  const int row_tiles = sg->row_tiles;
  const int64_t it = local_item(sg, item);
  const int rt = (int)(it % row_tiles);
  const int cg = (int)(it / row_tiles);
  const int set = cg < sg->chunks0 ? 0 : 1;
  const int ch = set ? cg - sg->chunks0 : cg;
  const int ld = weight_ld(op, sg, D);

  const SrcView src = resolve_src(op, sg, set);
  const TgtView tgt = resolve_tgt(op, sg);
  const double scale = coord_scale(prec, KT, sg->sigma);
  // Processing order of the rows and of the set-0 sources (the same point set): row i
  // of the sums is point perm[i]; partials are written at the point's own index.
  const int32_t *__restrict__ perm = sg->perm;
  const int32_t *__restrict__ sperm = set == 0 ? perm : nullptr;
  const bool expand_seg = EXPAND && perm != nullptr;
  // Ordered set-0 tiles, and both sets of energy-only sums, come staged when the launch has
  // a staging area; this set's first record.
  const bool staged = stage != nullptr && (sperm != nullptr || stage_sets(op, *sg, NA) == 2);
  const int64_t stage_t0 = sg->stage0 + (set ? (sg->nx0 + TS - 1) / TS : 0);

  KEval<R, KT> kev;
  kev.init(sg->order, sg->sigma);
  if constexpr (TAB) {
    // Read only after the tile loop's first barrier.
    static_assert(BLOCK == 256, "one table entry per thread");
    s_e2[tid] = EXP2_TAB256[tid];
  }

  // Target rows held in registers; each wave owns RPT x 64 consecutive rows.
  // Positions stay in FP64 (yd, prescaled); FP32 mode rounds them per tile,
  // after moving the origin to the tile (see the tile loop).
  double yd[RPT][D];
  R ty[RPT][D];
  R tw[RPT][NA > 0 ? NA : 1][M];
  int64_t rows[RPT];
  int64_t orow[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = (int64_t)rt * (BLOCK * RPT) + wave * (64 * RPT) + r * 64 + lane;
    rows[r] = k;
    int64_t kk = k < ny ? k : ny - 1;
    if (perm) kk = perm[kk];
    orow[r] = kk;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      yd[r][c] = tgt.pos[kk * D + c] * scale;
      ty[r][c] = (R)yd[r][c];
    }
    if constexpr (NA > 0) {
#pragma unroll
      for (int j = 0; j < NA; ++j) {
#pragma unroll
        for (int c = 0; c < M; ++c) {
          double v = tgt.w[j][kk * ld + c];
          if (j == 0 && tgt.w0b) v += tgt.tcoef * tgt.w0b[kk * ld + c];
          tw[r][j][c] = (R)v;
        }
      }
    }
  }

  // Bounding box of this wave's target rows (clamped duplicates are inside it),
  // FP64, prescaled absolute coordinates.
  double tlo[D], thi[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double lo = yd[0][c], hi = yd[0][c];
#pragma unroll
    for (int r = 1; r < RPT; ++r) { lo = fmin(lo, yd[r][c]); hi = fmax(hi, yd[r][c]); }
    tlo[c] = wave_uniform(wave_min(lo));
    thi[c] = wave_uniform(wave_max(hi));
  }
  const R cut2 = (R)cull_r2(prec, KT, sg->sigma);
  // The workgroup's box (union of its waves' boxes): a staged tile whose quarters all lie
  // beyond the cut from it is skipped before it is copied (no loads, no barriers).
  __shared__ double s_wgbox[BLOCK / 64][2 * D];
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      s_wgbox[wave][c] = tlo[c];
      s_wgbox[wave][D + c] = thi[c];
    }
  }
  __syncthreads();
  double wlo[D], whi[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    wlo[c] = s_wgbox[0][c];
    whi[c] = s_wgbox[0][D + c];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) {
      wlo[c] = fmin(wlo[c], s_wgbox[w][c]);
      whi[c] = fmax(whi[c], s_wgbox[w][D + c]);
    }
    wlo[c] = wave_uniform(wlo[c]);
    whi[c] = wave_uniform(whi[c]);
  }

  // Accumulators: running sum (s), Kahan compensation (c).
  R sK[RPT][KSUM ? M : 1], cK[RPT][KSUM ? M : 1];
  R sD[RPT][D], cD[RPT][D];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
#pragma unroll
    for (int c = 0; c < (KSUM ? M : 1); ++c) { sK[r][c] = 0; cK[r][c] = 0; }
#pragma unroll
    for (int c = 0; c < D; ++c) { sD[r][c] = 0; cD[r][c] = 0; }
  }

  const int64_t l0 = (int64_t)ch * sg->chunk_len;
  int64_t l1 = l0 + sg->chunk_len;
  if (l1 > src.n) l1 = src.n;

  // Symmetric self-sums.  When only the total <w, K(y, y) w> is asked for (KNORM, and
  // the self set of an energy-only CURRENTS or MEASURE segment: no per-row outputs), each
  // unordered pair of TS-blocks is evaluated once: a wave whose rows lie in block
  // rb takes source tile sb with weight 2 if sb > rb, 1 if sb == rb, 0 if sb < rb.
  // Tiles start at multiples of TS and a wave's rows never straddle a TS block,
  // so the weights are wave-uniform and every unordered pair counts twice.
  const bool energy_only =
      !sg->rowdot && !sg->out0 &&
      (op == LDDMM_OP_KNORM || ((op == LDDMM_OP_CURRENTS || op == LDDMM_OP_MEASURE) && NA == 0));
  const bool sym = energy_only && set == 0;
  const int64_t wave_blk = ((int64_t)rt * (BLOCK * RPT) + wave * (64 * RPT)) / TS;
  const int64_t wg_blk = ((int64_t)rt * (BLOCK * RPT)) / TS;
  // Mixed precision by magnitude (FP64 Gaussian energy sums only: segments without per-row
  // outputs, both source sets): runs of 64 sources whose bounding box is so far from the
  // wave's rows that every kernel value is below far_r2's bound are evaluated with an FP32
  // exponential.
  constexpr bool MIXED = std::is_same<R, double>::value && KT == LDDMM_KERNEL_GAUSS && KSUM &&
                         NA == 0;
  const bool mixed = MIXED && energy_only;
  const R far2 = MIXED ? (R)far_r2(prec, sg->sigma) : (R)0;

  for (int64_t base = l0; base < l1; base += TS) {
    if (sym && base / TS < wg_blk) continue;  // workgroup-uniform: below every wave's block
    const int sym_w = !sym ? 1 : (base / TS > wave_blk ? 2 : (base / TS == wave_blk ? 1 : 0));
    double org[D];
    const int64_t tlen = l1 - base < TS ? l1 - base : TS;
    if (staged) {
      // Copy the tile k_stage built (chunks start at multiples of TS: tile base / TS).
      const char *rec = stage + (stage_t0 + base / TS) * stage_rec_bytes(sizeof(R), D, M, NSW);
#pragma unroll
      for (int c = 0; c < D; ++c) org[c] = ((const double *)rec)[c];
      if constexpr (!KAHAN) {
        // Workgroup-level cull from the record's quarter boxes (tile coordinates); FP64
        // only: the FP32 flow step measured 1.3% slower with it (few of its tiles were
        // wholly culled, measured at the earlier 2^-32 cut).
        const R *qb = (const R *)(rec + 32);
        R d2 = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          R lo = qb[c], hi = qb[D + c];
#pragma unroll
          for (int q = 1; q < TS / 64; ++q) {
            lo = fmin(lo, qb[q * 2 * D + c]);
            hi = fmax(hi, qb[q * 2 * D + D + c]);
          }
          const R g = fmax(fmax((R)(wlo[c] - org[c]) - hi, lo - (R)(whi[c] - org[c])), (R)0);
          d2 = g * g + d2;
        }
        if (d2 > cut2) continue;  // workgroup-uniform: no quarter of this tile is live
      }
      __syncthreads();
      const R *rp = (const R *)(rec + STAGE_HEAD);
#pragma unroll
      for (int c = 0; c < D; ++c) s_pos[c][tid] = rp[c * TS + tid];
#pragma unroll
      for (int j = 0; j < NSW; ++j)
#pragma unroll
        for (int c = 0; c < M; ++c) s_w[j][c][tid] = rp[(D + j * M + c) * TS + tid];
      if constexpr (EXPAND) {
        if (expand_seg) {
          s_n2[tid] = rp[(D + NSW * M) * TS + tid];
          if (tid < TS / 64) s_r2max[tid] = ((const R *)(rec + 224))[tid];
        }
      }
      if constexpr (XFORM) {
        s_n2[tid] = kev.h256 * rp[(D + NSW * M) * TS + tid];
        if (tid < TS / 64) s_r2max[tid] = ((const R *)(rec + 224))[tid];
      }
      if (tid < (TS / 64) * 2 * D) (&s_box[0][0])[tid] = ((const R *)(rec + 32))[tid];
    } else {
      tile_origin<R, D, KAHAN || XFORM>(src, sperm, base, tlen, scale, org);
      __syncthreads();
      SrcStage<R, D, M, NSW> ss;
      stage_source<R, D, M, NSW>(src, sperm, ld, base + tid, l1, scale, org, expand_seg, XFORM,
                                 ss);
      if constexpr (XFORM) {
        s_n2[tid] = kev.h256 * ss.n2;
        if (lane == 0) s_r2max[wave] = ss.q2;
      }
#pragma unroll
      for (int c = 0; c < D; ++c) s_pos[c][tid] = ss.v[c];
#pragma unroll
      for (int j = 0; j < NSW; ++j)
#pragma unroll
        for (int c = 0; c < M; ++c) s_w[j][c][tid] = ss.w[j][c];
      if constexpr (EXPAND) {
        if (expand_seg) {
          s_n2[tid] = ss.n2;
          if (lane == 0) s_r2max[wave] = ss.q2;
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          s_box[wave][c] = ss.lo[c];
          s_box[wave][D + c] = ss.hi[c];
        }
      }
    }
    if constexpr (KAHAN) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) ty[r][c] = (R)(yd[r][c] - org[c]);
    }
    // Expanded exponent: rows in tile coordinates x', 2 h256 x' and -h256 |x'|^2; the wave's
    // largest |x'|^2 decides which blocks take the form.
    R xX[XFORM ? RPT : 1][D], xc[XFORM ? RPT : 1];
    R x2max = 0;
    if constexpr (XFORM) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        R x2 = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          ty[r][c] = yd[r][c] - org[c];
          xX[r][c] = (2.0 * kev.h256) * ty[r][c];
          x2 = ty[r][c] * ty[r][c] + x2;
        }
        xc[r] = -kev.h256 * x2;
        x2max = fmax(x2max, x2);
      }
      x2max = wave_uniform(wave_max(x2max));
    }
    __syncthreads();
    // Which 64-source quarters of the tile hold a pair above the cut for this
    // wave (squared distance between the target box and each quarter's box).
    unsigned live = 0, far = 0, expq = 0, xq = 0;
#pragma unroll
    for (int q = 0; q < TS / 64; ++q) {
      if constexpr (EXPAND) {
        if (expand_seg && s_r2max[q] <= (R)EXP_R2MAX) expq |= 1u << q;
      }
      if constexpr (XFORM) {
        // (FP32 square roots: a threshold test, with a margin for their rounding)
        const float xy = __builtin_sqrtf((float)x2max) + __builtin_sqrtf((float)s_r2max[q]);
        if ((float)kev.h * xy * xy <= (float)(0.98 * XF_HMAX)) xq |= 1u << q;
      }

      R d2 = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        R g = fmax(fmax((R)(tlo[c] - org[c]) - s_box[q][D + c], s_box[q][c] - (R)(thi[c] - org[c])),
                   (R)0);
        d2 += g * g;
      }
      if (!(d2 > cut2)) live |= 1u << q;
      if constexpr (MIXED) {
        if (d2 > far2) far |= 1u << q;
      }
    }
    if (!live || sym_w == 0) continue;  // wave-uniform: nothing of this tile to add
    const R wfac = (R)sym_w;
    // Run factor of the dot-product form for this tile: wfac * 2^-|x|^2, zero for rows
    // beyond EXP_X2MAX (their runs are dropped at the fold).
    R xfac[EXPAND ? RPT : 1];
    if constexpr (EXPAND) {
      if (expq & live) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          R x2 = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) x2 = ty[r][c] * ty[r][c] + x2;
          xfac[r] = x2 <= (R)EXP_X2MAX ? wfac * __builtin_amdgcn_exp2f(-x2) : (R)0;
        }
      }
    }

    for (int j0 = 0; j0 < TS; j0 += KFP) {
      if (!((live >> (j0 / 64)) & 1u)) continue;
      R tK[RPT][KSUM ? M : 1], tD[RPT][D];
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int c = 0; c < (KSUM ? M : 1); ++c) tK[r][c] = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) tD[r][c] = 0;
      }
      if constexpr (EXPAND) {
        if ((expq >> (j0 / 64)) & 1u) {
          // exponent = -|y|^2 + x.(2y) (three FMAs); the row factor is applied at the fold.
#pragma unroll PAIR_UNROLL
          for (int j = j0; j < j0 + KFP; ++j) {
            R sx[D], sw[M];
#pragma unroll
            for (int c = 0; c < D; ++c) sx[c] = s_pos[c][j];
#pragma unroll
            for (int c = 0; c < M; ++c) sw[c] = s_w[0][c][j];
            const R n2j = s_n2[j];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
              R e = n2j;
#pragma unroll
              for (int c = 0; c < D; ++c) e = ty[r][c] * sx[c] + e;
              const R k = __builtin_amdgcn_exp2f(e);
#pragma unroll
              for (int c = 0; c < M; ++c) tK[r][c] = k * sw[c] + tK[r][c];
            }
          }
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            const bool keep = xfac[r] > (R)0;
#pragma unroll
            for (int c = 0; c < M; ++c) {
              R y = keep ? tK[r][c] * xfac[r] - cK[r][c] : -cK[r][c];
              R t = sK[r][c] + y;
              cK[r][c] = (t - sK[r][c]) - y;
              sK[r][c] = t;
            }
          }
          continue;
        }
      }
      if constexpr (XFORM) {
        if ((xq >> (j0 / 64)) & 1u) {
          if (mixed && ((far >> (j0 / 64)) & 1u)) {
            // Every pair of this run has K below far_r2's bound (see below): v_exp_f32 of
            // the expanded FP64 exponent.
#pragma unroll PAIR_UNROLL
            for (int j = j0; j < j0 + KFP; ++j) {
              R sx[D], sw[M];
#pragma unroll
              for (int c = 0; c < D; ++c) sx[c] = s_pos[c][j];
#pragma unroll
              for (int c = 0; c < M; ++c) sw[c] = s_w[0][c][j];
              const R a = s_n2[j];
#pragma unroll
              for (int r = 0; r < RPT; ++r) {
                R t = a + xc[r];
#pragma unroll
                for (int c = 0; c < D; ++c) t = xX[r][c] * sx[c] + t;
                const R k = (R)__builtin_amdgcn_exp2f((float)(t * (1.0 / 256.0)));
#pragma unroll
                for (int c = 0; c < M; ++c) tK[r][c] = k * sw[c] + tK[r][c];
              }
            }
          } else {
            // The reference's cut-off (K = 1 for r^2 < 1e-8 sigma^2) on t: zeroing its high
            // word makes the exponential exactly 1, as for the direct form.  The loop runs
            // exp2_tab256 in stages over XU sources (exponents, then table reads, then
            // polynomials), so each table read's LDS latency overlaps the other sources'
            // arithmetic; written as one chain per source, the compiler issued the chains
            // one after another and waited on every table read.
            const R tcut = -kev.h256 * kev.cut_r2;
            constexpr int XU = 4;
            static_assert(RPT == 1, "one row per thread in FP64");
#pragma unroll 1
            for (int j = j0; j < j0 + KFP; j += XU) {
              R sw[XU][M], t[XU];
#pragma unroll
              for (int u = 0; u < XU; ++u) {
#pragma unroll
                for (int c = 0; c < M; ++c) sw[u][c] = s_w[0][c][j + u];
                R e = s_n2[j + u] + xc[0];
#pragma unroll
                for (int c = 0; c < D; ++c) e = xX[0][c] * s_pos[c][j + u] + e;
                t[u] = __hiloint2double(e > tcut ? 0 : __double2hiint(e), __double2loint(e));
              }
              const double shifter = 6755399441055744.0;  // 1.5 * 2^52, as in exp2_tab256
              int ni[XU];
              R tj[XU], g[XU];
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                const double kk = t[u] + shifter;
                ni[u] = __double2loint(kk);
                tj[u] = s_e2[ni[u] & 255];
                g[u] = t[u] - (kk - shifter);
              }
#pragma unroll
              for (int u = 0; u < XU; ++u) {
                double q = EXP2T256C.c[4];
#pragma unroll
                for (int i = 3; i >= 1; --i) q = fma(q, g[u], EXP2T256C.c[i]);
                q *= g[u];
                const R k = ldexp(fma(tj[u], q, tj[u]), ni[u] >> 8);
#pragma unroll
                for (int c = 0; c < M; ++c) tK[0][c] = k * sw[u][c] + tK[0][c];
              }
            }
          }
          goto fold;
        }
      }
      if constexpr (MIXED) {
        if (mixed && ((far >> (j0 / 64)) & 1u)) {
          // Every pair of this run has K below far_r2's bound (2^-26 in FP64): v_exp_f32 of
          // the FP64 exponent (relative error ~4e-6 from rounding an exponent of magnitude
          // < 80 to FP32) changes each term by less than 1e-13 of a unit kernel value (FP64;
          // 2e-11 in FP64E); the products accumulate in FP64.
#pragma unroll 8
          for (int j = j0; j < j0 + KFP; ++j) {
            R sx[D], sw[M];
#pragma unroll
            for (int c = 0; c < D; ++c) sx[c] = s_pos[c][j];
#pragma unroll
            for (int c = 0; c < M; ++c) sw[c] = s_w[0][c][j];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
              R r2 = 0;
#pragma unroll
              for (int c = 0; c < D; ++c) {
                const R dx = ty[r][c] - sx[c];
                r2 = dx * dx + r2;
              }
              const R k = (R)__builtin_amdgcn_exp2f((float)(-kev.h * r2));
#pragma unroll
              for (int c = 0; c < M; ++c) tK[r][c] = k * sw[c] + tK[r][c];
            }
          }
          goto fold;
        }
      }
      {
#pragma unroll PAIR_UNROLL
          for (int j = j0; j < j0 + KFP; ++j) {
            R sx[D], sw[NSW][M];
#pragma unroll
            for (int c = 0; c < D; ++c) sx[c] = s_pos[c][j];
#pragma unroll
            for (int q = 0; q < NSW; ++q)
#pragma unroll
              for (int c = 0; c < M; ++c) sw[q][c] = s_w[q][c][j];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
              R dx[D];
              R r2 = KEval<R, KT>::R2_0;
#pragma unroll
              for (int c = 0; c < D; ++c) {
                dx[c] = ty[r][c] - sx[c];
                r2 = dx[c] * dx[c] + r2;
              }
              R k0, k1, k2;
              if constexpr (TAB) kev.template eval_tab<true>(r2, s_e2, k0, k1, k2);
              else kev(r2, k0, k1, k2);
              if constexpr (KSUM) {
#pragma unroll
                for (int c = 0; c < M; ++c) tK[r][c] = k0 * sw[0][c] + tK[r][c];
              }
              if constexpr (NA > 0) {
                R wsum = 0;
#pragma unroll
                for (int q = 0; q < NA; ++q)
#pragma unroll
                  for (int c = 0; c < M; ++c) wsum = tw[r][q][c] * sw[q][c] + wsum;
                R f = k1 * wsum;
#pragma unroll
                for (int c = 0; c < D; ++c) tD[r][c] = f * dx[c] + tD[r][c];
              }
            }
          }
      }
    fold:
      // Fold the run into the compensated row sums.
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if constexpr (KSUM) {
#pragma unroll
          for (int c = 0; c < M; ++c) {
            tK[r][c] *= wfac;  // exact: 1 or 2
            if constexpr (KAHAN) {
              R y = tK[r][c] - cK[r][c];
              R t = sK[r][c] + y;
              cK[r][c] = (t - sK[r][c]) - y;
              sK[r][c] = t;
            } else {
              sK[r][c] += tK[r][c];
            }
          }
        }
        if constexpr (NA > 0) {
#pragma unroll
          for (int c = 0; c < D; ++c) {
            if constexpr (KAHAN) {
              R y = tD[r][c] - cD[r][c];
              R t = sD[r][c] + y;
              cD[r][c] = (t - sD[r][c]) - y;
              sD[r][c] = t;
            } else {
              sD[r][c] += tD[r][c];
            }
          }
        }
      }
    }
  }
  (void)NKAH;

  // Chunk partials, FP64.
  double *pp = part + sg->part0 + (int64_t)cg * ny * W;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if (rows[r] >= ny) continue;
    double *o = pp + orow[r] * W;
    int q = 0;
    if constexpr (KSUM) {
#pragma unroll
      for (int c = 0; c < M; ++c) o[q++] = (double)sK[r][c] - (double)cK[r][c];
    }
    if constexpr (NA > 0) {
#pragma unroll
      for (int c = 0; c < D; ++c) o[q++] = (double)sD[r][c] - (double)cD[r][c];
    }
  }
}

// ---------------------------------------------------------------------------
// applyDDiffK11and12 pair kernel: targets = sources = y; a1 = w0, a2 = w1, p = w2.
//   acc += <a1_k, a2_l> (k2 g <dx', dp> dx' - k1 dp),  out = acc / sigma^2.
// Form (seg.na): 0 = applykdiff11and12, dp = p_k - p_l (Fortran tables,
// kernelFunctions.f90:430-496); 1 = applyDDiffK11, dp = p_k, and 2 = applyDDiffK12,
// dp = -p_l (the dense siblings, kernelFunctions.py:459-506, whose K' and K''
// matrices use the exact-double Laplacian polynomials lapPolDiff / lapPolDiff2,
// :75-104).  Forms 1 and 2 add up to form 0.
// ---------------------------------------------------------------------------
__constant__ double LAP_R_EXACT[5][3] = {
    {0.0, 0.0, 0.0},  // orders 0, 1 have no second derivative (rejected on the host)
    {0.0, 0.0, 0.0},
    {1.0 / 3.0, 0.0, 0.0},
    {1.0 / 15.0, 1.0 / 15.0, 0.0},
    {3.0 / 105.0, 3.0 / 105.0, 1.0 / 105.0}};
__constant__ double LAP_Q_EXACT[5][4] = {
    {0.0, 0.0, 0.0, 0.0},  // order 0 has no first derivative (rejected on the host)
    {1.0, 0.0, 0.0, 0.0},
    {1.0 / 3.0, 1.0 / 3.0, 0.0, 0.0},
    {3.0 / 15.0, 3.0 / 15.0, 1.0 / 15.0, 0.0},
    {15.0 / 105.0, 15.0 / 105.0, 6.0 / 105.0, 1.0 / 105.0}};

// The dense paths' exact-double derivative polynomials in place of the Fortran tables.
template <typename R, int KT>
__device__ inline void use_exact_tables(KEval<R, KT> &kev, int order) {
  if constexpr (KT == LDDMM_KERNEL_LAPLACIAN) {
    kev.q0 = (R)LAP_Q_EXACT[order][0];
    kev.q1 = (R)LAP_Q_EXACT[order][1];
    kev.q2 = (R)LAP_Q_EXACT[order][2];
    kev.q3 = (R)LAP_Q_EXACT[order][3];
    kev.c0 = (R)LAP_R_EXACT[order][0];
    kev.c1 = (R)LAP_R_EXACT[order][1];
    kev.c2 = (R)LAP_R_EXACT[order][2];
  }
}
template <typename R, int KT, int D>
__global__ __launch_bounds__(BLOCK) void k_pairs_ddk(const lddmm_seg_t *__restrict__ segs,
                                                    int nseg, int prec,
                                                    double *__restrict__ part) {
  constexpr int RPT = rows_per_thread<R>();
  constexpr bool KAHAN = std::is_same<R, float>::value;
  __shared__ R s_pos[D][TS];
  __shared__ R s_a[D][TS];
  __shared__ R s_p[D][TS];

  const int tid = threadIdx.x;
  const int64_t item = blockIdx.x;
  const int si = find_segment(segs, nseg, item, 0);
  const lddmm_seg_t *sg = segs + si;
  const int64_t ny = sg->ny;
  const int row_tiles = sg->row_tiles;
  const int64_t it = local_item(sg, item);
  const int rt = (int)(it % row_tiles);
  const int cg = (int)(it / row_tiles);
  const double scale = coord_scale(prec, KT, sg->sigma);
  // g = 1/(s^2 sigma^2) with s the effective coordinate scale (1 in FP64).
  const R g = (R)(prec == LDDMM_PREC_FP32C ? 1.0 / (scale * scale * sg->sigma * sg->sigma)
                                           : 1.0 / (sg->sigma * sg->sigma));
  KEval<R, KT> kev;
  kev.init(sg->order, sg->sigma);
  const int form = sg->na;
  if (form != 0) use_exact_tables(kev, sg->order);
  // dp = tp - sp with tp = alpha p_k, sp = beta p_l (alpha, beta in {0, 1} by form)
  const double alpha = form == 2 ? 0.0 : 1.0, beta = form == 1 ? 0.0 : 1.0;

  R ty[RPT][D], ta[RPT][D], tp[RPT][D];
  int64_t rows[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = (int64_t)rt * (BLOCK * RPT) + r * BLOCK + tid;
    rows[r] = k;
    int64_t kk = k < ny ? k : ny - 1;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      ty[r][c] = (R)(sg->y[kk * D + c] * scale);
      ta[r][c] = (R)sg->w0[kk * D + c];
      tp[r][c] = (R)(alpha * sg->w2[kk * D + c]);
    }
  }
  R sD[RPT][D], cD[RPT][D];
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) { sD[r][c] = 0; cD[r][c] = 0; }

  const int64_t l0 = (int64_t)cg * sg->chunk_len;
  int64_t l1 = l0 + sg->chunk_len;
  if (l1 > sg->ny) l1 = sg->ny;
  for (int64_t base = l0; base < l1; base += TS) {
    __syncthreads();
    {
      int64_t l = base + tid;
      bool ok = l < l1;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        s_pos[c][tid] = ok ? (R)(sg->y[l * D + c] * scale) : (R)0;
        s_a[c][tid] = ok ? (R)sg->w1[l * D + c] : (R)0;
        s_p[c][tid] = ok ? (R)(beta * sg->w2[l * D + c]) : (R)0;
      }
    }
    __syncthreads();
    for (int j0 = 0; j0 < TS; j0 += KF) {
      R tD[RPT][D];
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) tD[r][c] = 0;
#pragma unroll 4
      for (int j = j0; j < j0 + KF; ++j) {
        R sx[D], sa[D], sp[D];
#pragma unroll
        for (int c = 0; c < D; ++c) { sx[c] = s_pos[c][j]; sa[c] = s_a[c][j]; sp[c] = s_p[c][j]; }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          R dx[D], dp[D];
          R r2 = KEval<R, KT>::R2_0, xp = 0, aa = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) {
            dx[c] = ty[r][c] - sx[c];
            dp[c] = tp[r][c] - sp[c];
            r2 = dx[c] * dx[c] + r2;
            xp = dx[c] * dp[c] + xp;
            aa = ta[r][c] * sa[c] + aa;
          }
          R k0, k1, k2;
          kev(r2, k0, k1, k2);
          R f2 = aa * k2 * g * xp;
          R f1 = aa * k1;
#pragma unroll
          for (int c = 0; c < D; ++c) tD[r][c] = f2 * dx[c] - f1 * dp[c] + tD[r][c];
        }
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
          if constexpr (KAHAN) {
            R y = tD[r][c] - cD[r][c];
            R t = sD[r][c] + y;
            cD[r][c] = (t - sD[r][c]) - y;
            sD[r][c] = t;
          } else {
            sD[r][c] += tD[r][c];
          }
        }
    }
  }
  double *pp = part + sg->part0 + (int64_t)cg * ny * D;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = rows[r];
    if (k >= ny) continue;
#pragma unroll
    for (int c = 0; c < D; ++c) pp[k * D + c] = (double)sD[r][c] - (double)cD[r][c];
  }
}

// ---------------------------------------------------------------------------
// Varifold data term of triangulated surfaces (d = 3): pair weight
// a_k a_l (1 + g^2) with g = <u_k, u_l>, u = n / a the unit normals and
// a = sqrt(|n|^2 + 1e-10) the face weights.  Per target row k and source l
// (weight sign -1 for the target-shape set):
//   E += K a_l (1 + g^2)                      (energy; total = sum_k a_k E_k)
//   P += K a_l (1 - g^2),  V += K a_l g u_l   (z1 = 2 V + P u_k)
//   D += k1 a_l (1 + g^2) dx                  (dz, scaled by a_k in finalize)
// The reference forms the (F, F) pair weights as dense arrays and sums them with
// applykmat / applykdiffmat (surfaces.py:1184-1276, kernelFunctions.f90:499-610);
// here they are formed per pair from O(F) face data.  Same work decomposition,
// culling, Kahan folds and symmetric energy-only path as k_pairs.  Partials:
// W = 1 (E) or 8 (E, P, V[3], D[3]) doubles per row and chunk.
// ---------------------------------------------------------------------------
template <typename R, int KT, bool GRAD>
__global__ __launch_bounds__(BLOCK) void k_pairs_varifold(const lddmm_seg_t *__restrict__ segs,
                                                         int nseg, int prec,
                                                         double *__restrict__ part) {
  constexpr int D = 3;
  constexpr int RPT = rows_per_thread<R>();
  constexpr int NACC = GRAD ? 5 : 1;
  constexpr int W = NACC + (GRAD ? D : 0);
  constexpr bool KAHAN = std::is_same<R, float>::value;

  __shared__ R s_pos[D][TS];
  __shared__ R s_u[D][TS];
  __shared__ R s_a[TS];
  __shared__ R s_box[TS / 64][2 * D];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int64_t item = blockIdx.x;
  const int si = find_segment(segs, nseg, item, 0);
  const lddmm_seg_t *sg = segs + si;

  const int64_t ny = sg->ny;
  const int row_tiles = sg->row_tiles;
  const int64_t it = local_item(sg, item);
  const int rt = (int)(it % row_tiles);
  const int cg = (int)(it / row_tiles);
  const int set = cg < sg->chunks0 ? 0 : 1;
  const int ch = set ? cg - sg->chunks0 : cg;

  // Sources: set 0 = the shape itself (y, w0, w2), set 1 = the target (x1, w1, sw[0]).
  const double *spos = set ? sg->x1 : sg->y;
  const double *swt = set ? sg->w1 : sg->w0;
  const double *snr = set ? sg->sw[0] : sg->w2;
  const int64_t sn = set ? sg->nx1 : sg->ny;
  const double sgn = set ? -1.0 : 1.0;
  const double scale = coord_scale(prec, KT, sg->sigma);

  KEval<R, KT> kev;
  kev.init(sg->order, sg->sigma);

  double yd[RPT][D];
  R ty[RPT][D], tu[RPT][D];
  int64_t rows[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = (int64_t)rt * (BLOCK * RPT) + wave * (64 * RPT) + r * 64 + lane;
    rows[r] = k;
    int64_t kk = k < ny ? k : ny - 1;
    const double a = sg->w0[kk];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      yd[r][c] = sg->y[kk * D + c] * scale;
      ty[r][c] = (R)yd[r][c];
      tu[r][c] = (R)(sg->w2[kk * D + c] / a);
    }
  }
  double tlo[D], thi[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    double lo = yd[0][c], hi = yd[0][c];
#pragma unroll
    for (int r = 1; r < RPT; ++r) { lo = fmin(lo, yd[r][c]); hi = fmax(hi, yd[r][c]); }
    tlo[c] = wave_min(lo);
    thi[c] = wave_max(hi);
  }
  const R cut2 = (R)cull_r2(prec, KT, sg->sigma);

  R sA[RPT][NACC], cA[RPT][NACC], sD[RPT][D], cD[RPT][D];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
#pragma unroll
    for (int c = 0; c < NACC; ++c) { sA[r][c] = 0; cA[r][c] = 0; }
#pragma unroll
    for (int c = 0; c < D; ++c) { sD[r][c] = 0; cD[r][c] = 0; }
  }

  const int64_t l0 = (int64_t)ch * sg->chunk_len;
  int64_t l1 = l0 + sg->chunk_len;
  if (l1 > sn) l1 = sn;

  // Energy-only self set: each unordered pair of TS blocks once (see k_pairs).
  const bool energy_only = !GRAD && !sg->rowdot && !sg->out0;
  const bool sym = energy_only && set == 0;
  // FP64 Gaussian energies: 64-source runs whose every kernel value is below far_r2's bound
  // take the FP32 exponential of the FP64 exponent, as in k_pairs.
  constexpr bool MIXED = std::is_same<R, double>::value && KT == LDDMM_KERNEL_GAUSS && !GRAD;
  const bool mixed = MIXED && energy_only;
  const R far2 = MIXED ? (R)far_r2(prec, sg->sigma) : (R)0;
  const int64_t wave_blk = ((int64_t)rt * (BLOCK * RPT) + wave * (64 * RPT)) / TS;
  const int64_t wg_blk = ((int64_t)rt * (BLOCK * RPT)) / TS;

  for (int64_t base = l0; base < l1; base += TS) {
    if (sym && base / TS < wg_blk) continue;
    const int sym_w = !sym ? 1 : (base / TS > wave_blk ? 2 : (base / TS == wave_blk ? 1 : 0));
    double org[D];
    const int64_t mid = base + ((l1 - base < TS ? l1 - base : TS) >> 1);
#pragma unroll
    for (int c = 0; c < D; ++c) org[c] = KAHAN ? spos[mid * D + c] * scale : 0.0;
    if constexpr (KAHAN) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) ty[r][c] = (R)(yd[r][c] - org[c]);
    }
    __syncthreads();
    {
      int64_t l = base + tid;
      R lo[D], hi[D];
      if (l < l1) {
        const double a = swt[l];
        s_a[tid] = (R)(sgn * a);
#pragma unroll
        for (int c = 0; c < D; ++c) {
          R v = (R)(spos[l * D + c] * scale - org[c]);
          s_pos[c][tid] = v;
          s_u[c][tid] = (R)(snr[l * D + c] / a);
          lo[c] = v;
          hi[c] = v;
        }
      } else {
        s_a[tid] = (R)0;  // padding sources: zero weight, exact zero terms
#pragma unroll
        for (int c = 0; c < D; ++c) {
          s_pos[c][tid] = (R)0;
          s_u[c][tid] = (R)0;
          lo[c] = (R)INFINITY;
          hi[c] = (R)-INFINITY;
        }
      }
#pragma unroll
      for (int c = 0; c < D; ++c) {
        lo[c] = wave_min(lo[c]);
        hi[c] = wave_max(hi[c]);
      }
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < D; ++c) {
          s_box[wave][c] = lo[c];
          s_box[wave][D + c] = hi[c];
        }
      }
    }
    __syncthreads();
    unsigned live = 0, far = 0;
#pragma unroll
    for (int q = 0; q < TS / 64; ++q) {
      R d2 = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        R g = fmax(fmax((R)(tlo[c] - org[c]) - s_box[q][D + c], s_box[q][c] - (R)(thi[c] - org[c])),
                   (R)0);
        d2 += g * g;
      }
      if (!(d2 > cut2)) live |= 1u << q;
      if constexpr (MIXED) {
        if (d2 > far2) far |= 1u << q;
      }
    }
    if (!live || sym_w == 0) continue;
    const R wfac = (R)sym_w;

    for (int j0 = 0; j0 < TS; j0 += KF) {
      if (!((live >> (j0 / 64)) & 1u)) continue;
      R tA[RPT][NACC], tD[RPT][D];
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int c = 0; c < NACC; ++c) tA[r][c] = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) tD[r][c] = 0;
      }
      if constexpr (MIXED) {
        if (mixed && ((far >> (j0 / 64)) & 1u)) {
#pragma unroll 4
          for (int j = j0; j < j0 + KF; ++j) {
            R sx[D], su[D];
#pragma unroll
            for (int c = 0; c < D; ++c) { sx[c] = s_pos[c][j]; su[c] = s_u[c][j]; }
            const R sa = s_a[j];
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
              R r2 = 0, g = 0;
#pragma unroll
              for (int c = 0; c < D; ++c) {
                const R dx = ty[r][c] - sx[c];
                r2 = dx * dx + r2;
                g = tu[r][c] * su[c] + g;
              }
              const R w = (R)__builtin_amdgcn_exp2f((float)(-kev.h * r2)) * sa;
              tA[r][0] = (w * (g * g) + w) + tA[r][0];
            }
          }
          goto vfold;
        }
      }
#pragma unroll 4
      for (int j = j0; j < j0 + KF; ++j) {
        R sx[D], su[D];
#pragma unroll
        for (int c = 0; c < D; ++c) { sx[c] = s_pos[c][j]; su[c] = s_u[c][j]; }
        const R sa = s_a[j];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          R dx[D];
          R r2 = KEval<R, KT>::R2_0, g = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) {
            dx[c] = ty[r][c] - sx[c];
            r2 = dx[c] * dx[c] + r2;
            g = tu[r][c] * su[c] + g;
          }
          R k0, k1, k2;
          kev(r2, k0, k1, k2);
          const R g2 = g * g;
          const R w = k0 * sa;
          tA[r][0] = (w * g2 + w) + tA[r][0];
          if constexpr (GRAD) {
            tA[r][1] = (w - w * g2) + tA[r][1];
            const R wg = w * g;
#pragma unroll
            for (int c = 0; c < D; ++c) tA[r][2 + c] = wg * su[c] + tA[r][2 + c];
            const R f0 = k1 * sa;
            const R f = f0 * g2 + f0;
#pragma unroll
            for (int c = 0; c < D; ++c) tD[r][c] = f * dx[c] + tD[r][c];
          }
        }
      }
    vfold:
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
#pragma unroll
        for (int c = 0; c < NACC; ++c) {
          tA[r][c] *= wfac;  // exact: 1 or 2
          if constexpr (KAHAN) {
            R y = tA[r][c] - cA[r][c];
            R t = sA[r][c] + y;
            cA[r][c] = (t - sA[r][c]) - y;
            sA[r][c] = t;
          } else {
            sA[r][c] += tA[r][c];
          }
        }
        if constexpr (GRAD) {
#pragma unroll
          for (int c = 0; c < D; ++c) {
            if constexpr (KAHAN) {
              R y = tD[r][c] - cD[r][c];
              R t = sD[r][c] + y;
              cD[r][c] = (t - sD[r][c]) - y;
              sD[r][c] = t;
            } else {
              sD[r][c] += tD[r][c];
            }
          }
        }
      }
    }
  }

  double *pp = part + sg->part0 + (int64_t)cg * ny * W;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = rows[r];
    if (k >= ny) continue;
    double *o = pp + k * W;
    int q = 0;
#pragma unroll
    for (int c = 0; c < NACC; ++c) o[q++] = (double)sA[r][c] - (double)cA[r][c];
    if constexpr (GRAD) {
#pragma unroll
      for (int c = 0; c < D; ++c) o[q++] = (double)sD[r][c] - (double)cD[r][c];
    }
  }
}

// ---------------------------------------------------------------------------
// applyDiffK family (kernelFunctions.f90:116-290), targets = sources = y,
// a1 = w0, a2 = w1 (n, d):
//   out_k = sum_l 2K'(r^2) <y_k - y_l, b_kl> a2_l,  b_kl = a1_k (form 1, applykdiff1),
//   -a1_l (form 2, applykdiff2), a1_k - a1_l (form 3, applykdiff1and2); form = seg.na.
// Fortran tables and cut-off (standard KEval).  The reference's applykdiff1/2 add
// into their f2py output array (f = f + df), which f2py zero-fills; out is the sum.
// ---------------------------------------------------------------------------
template <typename R, int KT, int D>
__global__ __launch_bounds__(BLOCK) void k_pairs_dk(const lddmm_seg_t *__restrict__ segs,
                                                   int nseg, int prec,
                                                   double *__restrict__ part) {
  constexpr int RPT = rows_per_thread<R>();
  constexpr bool KAHAN = std::is_same<R, float>::value;
  __shared__ R s_pos[D][TS];
  __shared__ R s_b[D][TS];   // source part of b: a1_l (forms 2, 3), else 0
  __shared__ R s_w[D][TS];   // a2_l

  const int tid = threadIdx.x;
  const int64_t item = blockIdx.x;
  const int si = find_segment(segs, nseg, item, 0);
  const lddmm_seg_t *sg = segs + si;
  const int64_t ny = sg->ny;
  const int row_tiles = sg->row_tiles;
  const int64_t it = local_item(sg, item);
  const int rt = (int)(it % row_tiles);
  const int cg = (int)(it / row_tiles);
  const double scale = coord_scale(prec, KT, sg->sigma);
  KEval<R, KT> kev;
  kev.init(sg->order, sg->sigma);
  const double tb = (sg->na & 1) ? 1.0 : 0.0, sb = (sg->na & 2) ? 1.0 : 0.0;

  double yd[RPT][D];
  R ty[RPT][D], tbv[RPT][D];
  int64_t rows[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = (int64_t)rt * (BLOCK * RPT) + r * BLOCK + tid;
    rows[r] = k;
    int64_t kk = k < ny ? k : ny - 1;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      yd[r][c] = sg->y[kk * D + c] * scale;
      ty[r][c] = (R)yd[r][c];
      tbv[r][c] = (R)(tb * sg->w0[kk * D + c]);
    }
  }
  R sD[RPT][D], cD[RPT][D];
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) { sD[r][c] = 0; cD[r][c] = 0; }

  const int64_t l0 = (int64_t)cg * sg->chunk_len;
  int64_t l1 = l0 + sg->chunk_len;
  if (l1 > ny) l1 = ny;
  for (int64_t base = l0; base < l1; base += TS) {
    double org[D];
    const int64_t mid = base + ((l1 - base < TS ? l1 - base : TS) >> 1);
#pragma unroll
    for (int c = 0; c < D; ++c) org[c] = KAHAN ? sg->y[mid * D + c] * scale : 0.0;
    if constexpr (KAHAN) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) ty[r][c] = (R)(yd[r][c] - org[c]);
    }
    __syncthreads();
    {
      int64_t l = base + tid;
      bool ok = l < l1;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        s_pos[c][tid] = ok ? (R)(sg->y[l * D + c] * scale - org[c]) : (R)0;
        s_b[c][tid] = ok ? (R)(sb * sg->w0[l * D + c]) : (R)0;
        s_w[c][tid] = ok ? (R)sg->w1[l * D + c] : (R)0;
      }
    }
    __syncthreads();
    for (int j0 = 0; j0 < TS; j0 += KF) {
      R tD[RPT][D];
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) tD[r][c] = 0;
#pragma unroll 4
      for (int j = j0; j < j0 + KF; ++j) {
        R sx[D], sbv[D], sw[D];
#pragma unroll
        for (int c = 0; c < D; ++c) { sx[c] = s_pos[c][j]; sbv[c] = s_b[c][j]; sw[c] = s_w[c][j]; }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          R r2 = KEval<R, KT>::R2_0, dot = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) {
            const R dx = ty[r][c] - sx[c];
            r2 = dx * dx + r2;
            dot = dx * (tbv[r][c] - sbv[c]) + dot;
          }
          R k0, k1, k2;
          kev(r2, k0, k1, k2);
          const R f = k1 * dot;
#pragma unroll
          for (int c = 0; c < D; ++c) tD[r][c] = f * sw[c] + tD[r][c];
        }
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
          if constexpr (KAHAN) {
            R y = tD[r][c] - cD[r][c];
            R t = sD[r][c] + y;
            cD[r][c] = (t - sD[r][c]) - y;
            sD[r][c] = t;
          } else {
            sD[r][c] += tD[r][c];
          }
        }
    }
  }
  double *pp = part + sg->part0 + (int64_t)cg * ny * D;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = rows[r];
    if (k >= ny) continue;
#pragma unroll
    for (int c = 0; c < D; ++c) pp[k * D + c] = (double)sD[r][c] - (double)cD[r][c];
  }
}

// ---------------------------------------------------------------------------
// Divergence of the velocity field v = K(., x) a at the targets y:
//   div_k = sum_l 2K'(r^2) <a_l, y_k - x_l>
// (Kernel.applyDivergence, kernelFunctions.py:510-530).  The reference forms the
// dense K' matrix (precompute(diff=True): kernelMatrixGauss :10-55,
// kernelMatrixLaplacian :152-183), whose Laplacian polynomials lapPolDiff
// (:75-88) are exact double expressions, not the Fortran's float32 tables; the
// first-derivative coefficients below are those doubles.  Its diagonal fill does
// not matter here: the self term carries y_k - x_k = 0.  One FP64 partial per
// row and chunk; no culling (an output-only path).
// ---------------------------------------------------------------------------
template <typename R, int KT, int D>
__global__ __launch_bounds__(BLOCK) void k_pairs_div(const lddmm_seg_t *__restrict__ segs,
                                                    int nseg, int prec,
                                                    double *__restrict__ part) {
  constexpr int RPT = rows_per_thread<R>();
  constexpr bool KAHAN = std::is_same<R, float>::value;
  __shared__ R s_pos[D][TS];
  __shared__ R s_a[D][TS];

  const int tid = threadIdx.x;
  const int64_t item = blockIdx.x;
  const int si = find_segment(segs, nseg, item, 0);
  const lddmm_seg_t *sg = segs + si;
  const int64_t ny = sg->ny;
  const int row_tiles = sg->row_tiles;
  const int64_t it = local_item(sg, item);
  const int rt = (int)(it % row_tiles);
  const int cg = (int)(it / row_tiles);
  const double *xs = sg->x0 ? sg->x0 : sg->y;
  const double scale = coord_scale(prec, KT, sg->sigma);
  KEval<R, KT> kev;
  kev.init(sg->order, sg->sigma);
  use_exact_tables(kev, sg->order);

  double yd[RPT][D];
  R ty[RPT][D];
  int64_t rows[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = (int64_t)rt * (BLOCK * RPT) + r * BLOCK + tid;
    rows[r] = k;
    int64_t kk = k < ny ? k : ny - 1;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      yd[r][c] = sg->y[kk * D + c] * scale;
      ty[r][c] = (R)yd[r][c];
    }
  }
  R sA[RPT], cA[RPT];
#pragma unroll
  for (int r = 0; r < RPT; ++r) { sA[r] = 0; cA[r] = 0; }

  const int64_t l0 = (int64_t)cg * sg->chunk_len;
  int64_t l1 = l0 + sg->chunk_len;
  if (l1 > sg->nx0) l1 = sg->nx0;
  for (int64_t base = l0; base < l1; base += TS) {
    double org[D];
    const int64_t mid = base + ((l1 - base < TS ? l1 - base : TS) >> 1);
#pragma unroll
    for (int c = 0; c < D; ++c) org[c] = KAHAN ? xs[mid * D + c] * scale : 0.0;
    if constexpr (KAHAN) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) ty[r][c] = (R)(yd[r][c] - org[c]);
    }
    __syncthreads();
    {
      int64_t l = base + tid;
      bool ok = l < l1;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        s_pos[c][tid] = ok ? (R)(xs[l * D + c] * scale - org[c]) : (R)0;
        s_a[c][tid] = ok ? (R)sg->w0[l * D + c] : (R)0;
      }
    }
    __syncthreads();
    for (int j0 = 0; j0 < TS; j0 += KF) {
      R tA[RPT];
#pragma unroll
      for (int r = 0; r < RPT; ++r) tA[r] = 0;
#pragma unroll 4
      for (int j = j0; j < j0 + KF; ++j) {
        R sx[D], sa[D];
#pragma unroll
        for (int c = 0; c < D; ++c) { sx[c] = s_pos[c][j]; sa[c] = s_a[c][j]; }
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          R r2 = KEval<R, KT>::R2_0, dot = 0;
#pragma unroll
          for (int c = 0; c < D; ++c) {
            const R dx = ty[r][c] - sx[c];
            r2 = dx * dx + r2;
            dot = dx * sa[c] + dot;
          }
          R k0, k1, k2;
          kev(r2, k0, k1, k2);
          tA[r] = k1 * dot + tA[r];
        }
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if constexpr (KAHAN) {
          R y = tA[r] - cA[r];
          R t = sA[r] + y;
          cA[r] = (t - sA[r]) - y;
          sA[r] = t;
        } else {
          sA[r] += tA[r];
        }
      }
    }
  }
  double *pp = part + sg->part0 + (int64_t)cg * ny;
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    int64_t k = rows[r];
    if (k < ny) pp[k] = (double)sA[r] - (double)cA[r];
  }
}

// ---------------------------------------------------------------------------
// Finalize: combine chunk partials in chunk order and apply the epilogue.
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = FIN_BLOCK / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

template <int OP, int D, int M, bool KSUM, int NA>
__global__ __launch_bounds__(FIN_BLOCK) void k_finalize(const lddmm_seg_t *__restrict__ segs,
                                                       int nseg, int prec, int kernel,
                                                       const double *__restrict__ part,
                                                       double *__restrict__ red) {
  constexpr int W = (KSUM ? M : 0) + (NA > 0 ? D : 0);
  __shared__ double sh[FIN_BLOCK];
  const int tid = threadIdx.x;
  const int64_t fb = blockIdx.x;
  const int si = find_segment(segs, nseg, fb, 1);
  const lddmm_seg_t *sg = segs + si;
  const int b = (int)(fb - sg->fin0);
  const int64_t ny = sg->ny;
  const int64_t k = (int64_t)b * FIN_BLOCK + tid;
  const bool valid = k < ny;
  const int64_t kk = valid ? k : 0;
  const int nch = sg->chunks0 + sg->chunks1;
  const double *pp = part + sg->part0;

  double accK0[KSUM ? M : 1], accK1[KSUM ? M : 1], accD[D];
#pragma unroll
  for (int c = 0; c < (KSUM ? M : 1); ++c) { accK0[c] = 0; accK1[c] = 0; }
#pragma unroll
  for (int c = 0; c < D; ++c) accD[c] = 0;
  if (valid) {
    // Chunk partials in chunk order, set 0 then set 1; unrolled so that the loads of
    // several chunks are in flight together (one HBM round trip per group, not per chunk).
    const int nch0 = sg->chunks0;
#pragma unroll 8
    for (int cgi = 0; cgi < nch0; ++cgi) {
      const double *o = pp + ((int64_t)cgi * ny + kk) * W;
      if constexpr (KSUM) {
#pragma unroll
        for (int c = 0; c < M; ++c) accK0[c] += o[c];
      }
      if constexpr (NA > 0) {
#pragma unroll
        for (int c = 0; c < D; ++c) accD[c] += o[(KSUM ? M : 0) + c];
      }
    }
#pragma unroll 8
    for (int cgi = nch0; cgi < nch; ++cgi) {
      const double *o = pp + ((int64_t)cgi * ny + kk) * W;
      if constexpr (KSUM) {
#pragma unroll
        for (int c = 0; c < M; ++c) accK1[c] += o[c];
      }
      if constexpr (NA > 0) {
#pragma unroll
        for (int c = 0; c < D; ++c) accD[c] += o[(KSUM ? M : 0) + c];
      }
    }
  }
  const double sigma = sg->sigma;
  const double s = coord_scale(prec, kernel, sigma);
  // sum_l 2 K'(r^2) (y_k - x_l) w  =  -acc / (sigma^2 s)
  const double dscale = -1.0 / (sigma * sigma * s);
  double rd0 = 0.0, rd1 = 0.0;

  if constexpr (OP == LDDMM_OP_APPLY_K) {
    if (valid) {
      const int ld = sg->m;
#pragma unroll
      for (int c = 0; c < M; ++c) sg->out0[k * ld + c] = accK0[c];
    }
  } else if constexpr (OP == LDDMM_OP_DIFF_KT) {
    if (valid) {
#pragma unroll
      for (int c = 0; c < D; ++c) sg->out0[k * D + c] = dscale * accD[c];
    }
  } else if constexpr (OP == LDDMM_OP_FLOW_STEP) {
    if (valid) {
      const double dt = sg->dt;
      const double *A = sg->w1, *bv = sg->w2;  // optional affine part (d x d, d)
      double yk[D];
#pragma unroll
      for (int c = 0; c < D; ++c) yk[c] = sg->y[k * D + c];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double v = accK0[c];
        double x1 = yk[c] + dt * v;
        if (A) {
          // x_{t+1} += dt (A x_t + b), added after the kernel step as the reference does
          double ax = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) ax += yk[j] * A[c * D + j];
          x1 += dt * (ax + bv[c]);
        }
        sg->out0[k * D + c] = x1;
        if (sg->out1) sg->out1[k * D + c] = v;
        rd0 += sg->w0[k * D + c] * v;
      }
      if (sg->rowdot) sg->rowdot[k] = rd0;
    }
  } else if constexpr (OP == LDDMM_OP_COVECTOR_STEP) {
    if (valid) {
      const double dt = sg->dt;
      const double *A = sg->w2;  // optional affine part (d x d)
      double pk[D];
#pragma unroll
      for (int c = 0; c < D; ++c) pk[c] = sg->w1[k * D + c];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double p0 = pk[c] + dt * (dscale * accD[c]);
        if (A) {
          // p_t += dt p_{t+1} A, after the kernel term as the reference does
          double pa = 0.0;
#pragma unroll
          for (int j = 0; j < D; ++j) pa += pk[j] * A[j * D + c];
          p0 += dt * pa;
        }
        sg->out0[k * D + c] = p0;
      }
    }
  } else if constexpr (OP == LDDMM_OP_KDOT) {
    if (valid) {
#pragma unroll
      for (int c = 0; c < D; ++c) {
        rd0 += sg->g[0][k * D + c] * accK0[c];
        if (sg->na > 1) rd1 += sg->g[1][k * D + c] * accK0[c];
      }
    }
  } else if constexpr (OP == LDDMM_OP_KNORM) {
    if (valid) {
#pragma unroll
      for (int c = 0; c < D; ++c) rd0 += sg->w0[k * D + c] * accK0[c];
      if (sg->rowdot) sg->rowdot[k] = rd0;
    }
  } else if constexpr (OP == LDDMM_OP_CURRENTS) {
    if (valid) {
      // accK1 already carries the sign of set 1: accK1 = -K(c1,c2) cr2.
      double e0 = 0.0, e1 = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double w = sg->w0[k * D + c];
        e0 += w * accK0[c];
        e1 += w * accK1[c];
      }
      rd0 = e0 + 2.0 * e1;
      if (sg->rowdot) sg->rowdot[k] = rd0;
      if (sg->out0) {
#pragma unroll
        for (int c = 0; c < D; ++c) sg->out0[k * D + c] = accK0[c] + accK1[c];
      }
      if constexpr (NA > 0) {
        if (sg->out1) {
#pragma unroll
          for (int c = 0; c < D; ++c) sg->out1[k * D + c] = sg->coef * (dscale * accD[c]);
        }
      }
    }
  } else if constexpr (OP == LDDMM_OP_MEASURE) {
    if (valid) {
      // One weight column: accK0 = (K(c1,c1) a1)_k, accK1 = -(K(c1,c2) a2)_k.
      const double a = sg->w0[k];
      rd0 = a * accK0[0] + 2.0 * (a * accK1[0]);
      if (sg->rowdot) sg->rowdot[k] = rd0;
      if (sg->out0) {
        const double zs = accK0[0] + accK1[0];
#pragma unroll
        for (int c = 0; c < D; ++c) sg->out0[k * D + c] = zs * (sg->w2[k * D + c] / a);
      }
      if constexpr (NA > 0) {
        if (sg->out1) {
#pragma unroll
          for (int c = 0; c < D; ++c) sg->out1[k * D + c] = sg->coef * (dscale * accD[c]);
        }
      }
    }
  } else if constexpr (OP == LDDMM_OP_VARIFOLD) {
    if (valid) {
      // Column 0: E (energy); gradient form adds P, V[3] and the derivative sum.
      const double a = sg->w0[k];
      rd0 = a * accK0[0] + 2.0 * (a * accK1[0]);
      if (sg->rowdot) sg->rowdot[k] = rd0;
      if constexpr (M == 5 && NA > 0) {
        if (sg->out0) {
          const double pk = accK0[1] + accK1[1];
#pragma unroll
          for (int c = 0; c < D; ++c)
            sg->out0[k * D + c] = 2.0 * (accK0[2 + c] + accK1[2 + c]) + pk * (sg->w2[k * D + c] / a);
        }
        if (sg->out1) {
#pragma unroll
          for (int c = 0; c < D; ++c) sg->out1[k * D + c] = sg->coef * (a * (dscale * accD[c]));
        }
      }
    }
  } else if constexpr (OP == LDDMM_OP_DIVERGENCE) {
    if (valid) sg->out0[k] = dscale * accK0[0];
  } else if constexpr (OP == LDDMM_OP_DIFF_K) {
    if (valid) {
#pragma unroll
      for (int c = 0; c < D; ++c) sg->out0[k * D + c] = dscale * accD[c];
    }
  } else if constexpr (OP == LDDMM_OP_DDIFF_K11_12) {
    if (valid) {
      const double inv = 1.0 / (sigma * sigma);
#pragma unroll
      for (int c = 0; c < D; ++c) sg->out0[k * D + c] = inv * accD[c];
    }
  }

  constexpr int NRED = (OP == LDDMM_OP_FLOW_STEP || OP == LDDMM_OP_KNORM ||
                        OP == LDDMM_OP_CURRENTS || OP == LDDMM_OP_MEASURE ||
                        OP == LDDMM_OP_VARIFOLD)
                           ? 1 : (OP == LDDMM_OP_KDOT ? 2 : 0);
  if constexpr (NRED > 0) {
    double t0 = block_sum(rd0, sh);
    double t1 = 0.0;
    if constexpr (NRED > 1) t1 = block_sum(rd1, sh);
    if (tid == 0) {
      red[sg->red0 + b] = t0;
      if constexpr (NRED > 1) red[sg->red0 + sg->fin_blocks + b] = t1;
    }
  }
}

// Per-segment FP64 totals from the finalize blocks' partials (fixed tree).
__global__ __launch_bounds__(TOT_BLOCK) void k_totals(const lddmm_seg_t *__restrict__ segs,
                                                     int nred, const double *__restrict__ red) {
  segment_totals(segs + blockIdx.x, nred, red);
}

// ---------------------------------------------------------------------------
// Geometry and vertex gathers.
// ---------------------------------------------------------------------------
// Geometry and gathers evaluate each expression in the reference's NumPy order without
// FMA contraction, so their FP64 results are the reference's bit for bit.
// Measure weight of a face: sqrt(|n|^2 + 1e-10), the sum of squares in index order
// (surfaces.py:1098,1130; curves.py:658).
__device__ inline double measure_weight(const double *n, int d) {
#pragma clang fp contract(off)
  double s = n[0] * n[0];
  for (int q = 1; q < d; ++q) s = s + n[q] * n[q];
  return sqrt(s + 1e-10);
}

__device__ inline void surface_face(const double *__restrict__ v, const int32_t *__restrict__ f,
                                    int64_t i, double *__restrict__ cen, double *__restrict__ nrm,
                                    double *__restrict__ wts = nullptr) {
#pragma clang fp contract(off)
  const double *a = v + 3 * (int64_t)f[3 * i + 0];
  const double *b = v + 3 * (int64_t)f[3 * i + 1];
  const double *c = v + 3 * (int64_t)f[3 * i + 2];
#pragma unroll
  for (int q = 0; q < 3; ++q) cen[3 * i + q] = (a[q] + b[q] + c[q]) / 3.0;
  double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                 e1[0] * e2[1] - e1[1] * e2[0]};
#pragma unroll
  for (int q = 0; q < 3; ++q) nrm[3 * i + q] = n[q];
  if (wts) wts[i] = measure_weight(n, 3);
}

__device__ inline void curve_edge(const double *__restrict__ v, const int32_t *__restrict__ e,
                                  int64_t i, int d, double *__restrict__ cen,
                                  double *__restrict__ len, double *__restrict__ wts = nullptr) {
#pragma clang fp contract(off)
  const double *a = v + (int64_t)d * e[2 * i + 0];
  const double *b = v + (int64_t)d * e[2 * i + 1];
  double l[8];  // d <= 8, checked by lddmm_curve_geometry and lddmm_geometry_batch
  for (int q = 0; q < d; ++q) {
    cen[d * i + q] = (a[q] + b[q]) / 2.0;
    l[q] = b[q] - a[q];
    len[d * i + q] = l[q];
  }
  if (wts) wts[i] = measure_weight(l, d);
}

__device__ inline void surface_gather(const double *__restrict__ v, const int32_t *__restrict__ f,
                                      const double *__restrict__ z1, const double *__restrict__ dz,
                                      const int32_t *__restrict__ ptr,
                                      const int32_t *__restrict__ inc, int64_t i, double scale,
                                      double *__restrict__ px) {
#pragma clang fp contract(off)
  double acc[3] = {0.0, 0.0, 0.0};
  for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) {
    int32_t code = inc[q];
    int32_t fi = code / 3, j = code - 3 * fi;
    const double *xa = v + 3 * (int64_t)f[3 * fi + (j + 2) % 3];
    const double *xb = v + 3 * (int64_t)f[3 * fi + (j + 1) % 3];
    double e[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
    const double *z = z1 + 3 * (int64_t)fi;
    double cr[3] = {e[1] * z[2] - e[2] * z[1], e[2] * z[0] - e[0] * z[2], e[0] * z[1] - e[1] * z[0]};
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] = (acc[c] + dz[3 * (int64_t)fi + c]) - cr[c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) px[3 * i + c] = scale * acc[c];
}

__device__ inline void curve_gather(const double *__restrict__ z1, const double *__restrict__ dz,
                                    const int32_t *__restrict__ ptr,
                                    const int32_t *__restrict__ inc, int64_t i, int d,
                                    double scale, double *__restrict__ px) {
#pragma clang fp contract(off)
  double acc[8];
  for (int c = 0; c < d; ++c) acc[c] = 0.0;
  for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) {
    int32_t code = inc[q];
    int32_t ei = code / 2, j = code - 2 * ei;
    for (int c = 0; c < d; ++c) {
      double z = z1[(int64_t)d * ei + c];
      double t = acc[c] + dz[(int64_t)d * ei + c];
      acc[c] = j ? t + z : t - z;
    }
  }
  for (int c = 0; c < d; ++c) px[(int64_t)d * i + c] = scale * acc[c];
}

__global__ void k_surface_geometry(const double *__restrict__ v, const int32_t *__restrict__ f,
                                   int64_t nf, double *__restrict__ cen, double *__restrict__ nrm) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nf) surface_face(v, f, i, cen, nrm);
}

__global__ void k_curve_geometry(const double *__restrict__ v, const int32_t *__restrict__ e,
                                 int64_t ne, int d, double *__restrict__ cen,
                                 double *__restrict__ len) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ne) curve_edge(v, e, i, d, cen, len);
}

__global__ void k_surface_vertex_gradient(const double *__restrict__ v, const int32_t *__restrict__ f,
                                          const double *__restrict__ z1, const double *__restrict__ dz,
                                          const int32_t *__restrict__ ptr, const int32_t *__restrict__ inc,
                                          int64_t nv, double scale, double *__restrict__ px) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) surface_gather(v, f, z1, dz, ptr, inc, i, scale, px);
}

__global__ void k_curve_vertex_gradient(const double *__restrict__ z1, const double *__restrict__ dz,
                                        const int32_t *__restrict__ ptr, const int32_t *__restrict__ inc,
                                        int64_t nv, int d, double scale, double *__restrict__ px) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) curve_gather(z1, dz, ptr, inc, i, d, scale, px);
}

// Batched forms: block -> entry by binary search on blk0.
template <typename T>
__device__ inline int find_blk(const T *__restrict__ e, int n, int64_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (e[mid].blk0 <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_geometry_batch(const lddmm_geom_t *__restrict__ g, int n) {
  const int64_t b = blockIdx.x;
  const lddmm_geom_t *e = g + find_blk(g, n, b);
  const int64_t i = (b - e->blk0) * 256 + threadIdx.x;
  if (i >= e->nf) return;
  if (e->k == 3)
    surface_face(e->vertices, e->faces, i, e->centers, e->normals, e->weights);
  else
    curve_edge(e->vertices, e->faces, i, e->d, e->centers, e->normals, e->weights);
}

__global__ __launch_bounds__(256) void k_gather_batch(const lddmm_gather_t *__restrict__ g, int n) {
  const int64_t b = blockIdx.x;
  const lddmm_gather_t *e = g + find_blk(g, n, b);
  const int64_t i = (b - e->blk0) * 256 + threadIdx.x;
  if (i >= e->nv) return;
  if (e->k == 3)
    surface_gather(e->vertices, e->faces, e->z1, e->dz, e->incid_ptr, e->incid, i, e->scale, e->px);
  else
    curve_gather(e->z1, e->dz, e->incid_ptr, e->incid, i, e->d, e->scale, e->px);
}

// Landmark data term: one workgroup per entry, strided FP64 partials, fixed tree.
__global__ __launch_bounds__(1024) void k_landmark_batch(const lddmm_lmk_t *__restrict__ g) {
  __shared__ double sh[1024];
  const lddmm_lmk_t *e = g + blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t n = e->n;
  const double scale = e->scale;
  double v = 0.0;
  for (int64_t i = tid; i < n; i += 1024) {
    const double u = e->x[i] - e->y[i];
    v += u * u;
    if (e->grad) e->grad[i] = scale * u;
  }
  if (!e->total) return;
  sh[tid] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  if (tid == 0) *e->total = sh[0];
}

__global__ __launch_bounds__(1024) void k_sum_f64(const double *__restrict__ x, int64_t n,
                                                 double *__restrict__ out) {
  __shared__ double sh[1024];
  const int tid = threadIdx.x;
  double v = 0.0;
  for (int64_t i = tid; i < n; i += 1024) v += x[i];
  sh[tid] = v;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if (tid < s) sh[tid] += sh[tid + s];
    __syncthreads();
  }
  if (tid == 0) *out = sh[0];
}

// Descriptor upload through kernel arguments: a by-value copy of up to STAMP
// segments is written to device memory.  Fully stream-ordered; no host buffer
// has to outlive the call.
constexpr int STAMP = 4;
struct SegPack {
  lddmm_seg_t s[STAMP];
};
__global__ void k_stamp(SegPack pack, int n, lddmm_seg_t *__restrict__ dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = pack.s[i];
}

// out_i = a_i + (s_i d_i) for up to AXPY_PACK arrays per launch (the line search's trial
// momenta at - eps dir of a whole batch of registrations), item by blockIdx.y, elements by
// a grid-stride loop; the product is rounded before the sum, as two elementwise passes do.
constexpr int AXPY_PACK = 32;
struct AxpyPack {
  double *out[AXPY_PACK];
  const double *a[AXPY_PACK];
  const double *d[AXPY_PACK];
  double s[AXPY_PACK];
  int64_t len[AXPY_PACK];
};
__global__ __launch_bounds__(256) void k_axpy_batch(AxpyPack pk) {
#pragma clang fp contract(off)
  const int j = blockIdx.y;
  const int64_t n = pk.len[j];
  double *__restrict__ out = pk.out[j];
  const double *__restrict__ a = pk.a[j];
  const double *__restrict__ d = pk.d[j];
  const double s = pk.s[j];
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double p = s * d[k];
    out[k] = a[k] + p;
  }
}

// ---------------------------------------------------------------------------
// Host side: planning, dispatch.
// ---------------------------------------------------------------------------
int hip_check(hipError_t e) {
  if (e != hipSuccess) {
    g_last_hip_error = (int)e;
    return LDDMM_EHIP;
  }
  return LDDMM_OK;
}

int op_nred(int op) {
  switch (op) {
    case LDDMM_OP_FLOW_STEP:
    case LDDMM_OP_KNORM:
    case LDDMM_OP_CURRENTS:
    case LDDMM_OP_MEASURE:
    case LDDMM_OP_VARIFOLD:
      return 1;
    case LDDMM_OP_KDOT:
      return 2;
    default:
      return 0;
  }
}

bool op_valid(int op) { return op >= LDDMM_OP_APPLY_K && op <= LDDMM_OP_DIFF_K; }

// Launch-uniform shape of a planned operation.
struct Shape {
  int M, NA;
  bool KSUM;
  int W;
};

int op_shape(int op, int d, const lddmm_seg_t &s0, Shape *sh) {
  switch (op) {
    case LDDMM_OP_APPLY_K:
      if (s0.m < 1 || s0.m > 3) return LDDMM_EINVAL;
      *sh = {s0.m, 0, true, s0.m};
      return LDDMM_OK;
    case LDDMM_OP_DIFF_KT:
      if (!(s0.m == 1 || s0.m == d) || s0.na < 1 || s0.na > 3) return LDDMM_EINVAL;
      *sh = {s0.m, s0.na, false, d};
      return LDDMM_OK;
    case LDDMM_OP_FLOW_STEP:
    case LDDMM_OP_KDOT:
    case LDDMM_OP_KNORM:
      *sh = {d, 0, true, d};
      return LDDMM_OK;
    case LDDMM_OP_COVECTOR_STEP:
      *sh = {d, 2, false, d};
      return LDDMM_OK;
    case LDDMM_OP_CURRENTS: {
      bool grad = s0.out1 != nullptr;
      *sh = {d, grad ? 1 : 0, true, grad ? 2 * d : d};
      return LDDMM_OK;
    }
    case LDDMM_OP_MEASURE: {
      bool grad = s0.out1 != nullptr;
      *sh = {1, grad ? 1 : 0, true, grad ? 1 + d : 1};
      return LDDMM_OK;
    }
    case LDDMM_OP_VARIFOLD: {
      if (d != 3) return LDDMM_EINVAL;  // triangulated surfaces only
      bool grad = s0.out1 != nullptr;
      *sh = {grad ? 5 : 1, grad ? 1 : 0, true, grad ? 8 : 1};
      return LDDMM_OK;
    }
    case LDDMM_OP_DDIFF_K11_12:
      *sh = {d, 0, false, d};
      return LDDMM_OK;
    case LDDMM_OP_DIVERGENCE:
      *sh = {1, 0, true, 1};
      return LDDMM_OK;
    case LDDMM_OP_DIFF_K:
      *sh = {d, 1, false, d};
      return LDDMM_OK;
  }
  return LDDMM_EINVAL;
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Descriptor head of the workspace (a device copy of the segment array).
size_t desc_bytes(int nseg) {
  size_t b = (size_t)nseg * sizeof(lddmm_seg_t);
  return (b + 255) & ~(size_t)255;
}

// Staging area of the ordered segments' source tiles: after the descriptor head and
// `doubles` partials, 256-byte aligned; one record of stage_rec bytes per tile.
size_t stage_offset(int nseg, int64_t doubles) {
  size_t b = desc_bytes(nseg) + (size_t)doubles * sizeof(double);
  return (b + 255) & ~(size_t)255;
}

size_t stage_rec(int prec, int d, const Shape &sh) {
  const int nsw = sh.NA > (sh.KSUM ? 1 : 0) ? sh.NA : (sh.KSUM ? 1 : 0);
  return stage_rec_bytes(prec == LDDMM_PREC_FP32C ? 4 : 8, d, sh.M, nsw);
}

}  // namespace

extern "C" int lddmm_segments_plan(int op, int prec, int d, lddmm_seg_t *segs, int nseg,
                                   int64_t *n_items, int64_t *n_fin_blocks,
                                   size_t *work_bytes) {
  if (!op_valid(op) || nseg < 1 || !segs || (d != 2 && d != 3)) return LDDMM_EINVAL;
  if (prec != LDDMM_PREC_FP64 && prec != LDDMM_PREC_FP32C && prec != LDDMM_PREC_FP64E)
    return LDDMM_EINVAL;
  Shape shp;
  int rc = op_shape(op, d, segs[0], &shp);
  if (rc) return rc;
  const int rows = rows_per_item(prec);
  const int nred = op_nred(op);
  const int kern = segs[0].kernel;
  int64_t item = 0, part = 0, fin = 0, redo = 0;
  int64_t stage = 0;  // staged source tiles of the ordered segments
  for (int i = 0; i < nseg; ++i) {
    lddmm_seg_t &s = segs[i];
    if (s.kernel != kern) return LDDMM_EINVAL;  // one kernel family per launch
    if (s.kernel != LDDMM_KERNEL_GAUSS && s.kernel != LDDMM_KERNEL_LAPLACIAN) return LDDMM_EINVAL;
    if (s.kernel == LDDMM_KERNEL_LAPLACIAN) {
      if (s.order < 0 || s.order > 4) return LDDMM_EINVAL;
      bool deriv = op == LDDMM_OP_DIFF_KT || op == LDDMM_OP_COVECTOR_STEP ||
                   op == LDDMM_OP_DIVERGENCE || op == LDDMM_OP_DIFF_K ||
                   ((op == LDDMM_OP_CURRENTS || op == LDDMM_OP_MEASURE ||
                     op == LDDMM_OP_VARIFOLD) && s.out1);
      if (deriv && s.order < 1) return LDDMM_EINVAL;
      if (op == LDDMM_OP_DDIFF_K11_12 && s.order < 2) return LDDMM_EINVAL;
    }
    if (!(s.sigma > 0.0) || !isfinite(s.sigma)) return LDDMM_EINVAL;
    if (s.ny < 1 || !s.y) return LDDMM_EINVAL;
    Shape si;
    if (op_shape(op, d, s, &si) || si.M != shp.M || si.NA != shp.NA || si.W != shp.W)
      return LDDMM_EINVAL;
    int64_t nx0 = s.nx0, nx1 = 0;
    switch (op) {
      case LDDMM_OP_APPLY_K:
        if (!s.w0 || !s.out0) return LDDMM_EINVAL;
        if (!s.x0) nx0 = s.ny;
        break;
      case LDDMM_OP_DIFF_KT:
        if (!s.out0) return LDDMM_EINVAL;
        for (int j = 0; j < s.na; ++j)
          if (!s.tw[j] || !s.sw[j]) return LDDMM_EINVAL;
        if (!s.x0) nx0 = s.ny;
        break;
      case LDDMM_OP_FLOW_STEP:
        if (!s.w0 || !s.out0) return LDDMM_EINVAL;
        if (!s.w1 != !s.w2) return LDDMM_EINVAL;  // affine part: both A and b, or neither
        nx0 = s.ny;
        break;
      case LDDMM_OP_COVECTOR_STEP:
        if (!s.w0 || !s.w1 || !s.out0) return LDDMM_EINVAL;
        nx0 = s.ny;
        break;
      case LDDMM_OP_KDOT:
        if (!s.w0 || s.na < 1 || s.na > 2 || !s.g[0] || (s.na > 1 && !s.g[1])) return LDDMM_EINVAL;
        nx0 = s.ny;
        break;
      case LDDMM_OP_KNORM:
        if (!s.w0) return LDDMM_EINVAL;
        nx0 = s.ny;
        break;
      case LDDMM_OP_CURRENTS:
        if (!s.w0 || !s.x1 || !s.w1 || s.nx1 < 1) return LDDMM_EINVAL;
        nx0 = s.ny;
        nx1 = s.nx1;
        break;
      case LDDMM_OP_MEASURE:
        if (!s.w0 || s.nx1 < 0 || (s.nx1 > 0 && (!s.x1 || !s.w1))) return LDDMM_EINVAL;
        if (s.out0 && !s.w2) return LDDMM_EINVAL;  // z1 needs the normals
        nx0 = s.ny;
        nx1 = s.nx1;
        break;
      case LDDMM_OP_VARIFOLD:
        if (!s.w0 || !s.w2 || s.nx1 < 0) return LDDMM_EINVAL;
        if (s.nx1 > 0 && (!s.x1 || !s.w1 || !s.sw[0])) return LDDMM_EINVAL;
        if (s.out0 && !s.out1) return LDDMM_EINVAL;  // z1 only in the gradient form
        nx0 = s.ny;
        nx1 = s.nx1;
        break;
      case LDDMM_OP_DDIFF_K11_12:
        if (!s.w0 || !s.w1 || !s.w2 || !s.out0) return LDDMM_EINVAL;
        if (s.na < 0 || s.na > 2) return LDDMM_EINVAL;  // form
        nx0 = s.ny;
        break;
      case LDDMM_OP_DIFF_K:
        if (!s.w0 || !s.w1 || !s.out0 || s.na < 1 || s.na > 3) return LDDMM_EINVAL;
        nx0 = s.ny;
        break;
      case LDDMM_OP_DIVERGENCE:
        if (!s.w0 || !s.out0) return LDDMM_EINVAL;
        if (!s.x0) nx0 = s.ny;
        break;
    }
    // A processing order: the flow and covector steps and the RKHS products (rows = set-0
    // sources).
    if (s.perm && op != LDDMM_OP_FLOW_STEP && op != LDDMM_OP_COVECTOR_STEP && op != LDDMM_OP_KDOT)
      return LDDMM_EINVAL;
    if (nx0 < 1) return LDDMM_EINVAL;
    s.nx0 = nx0;
    if (op != LDDMM_OP_CURRENTS && op != LDDMM_OP_MEASURE && op != LDDMM_OP_VARIFOLD) s.nx1 = 0;
    const int64_t row_tiles = ceil_div(s.ny, rows);
    const int64_t src_tiles = ceil_div(nx0, TS) + (nx1 ? ceil_div(nx1, TS) : 0);
    const bool energy_only =
        !s.rowdot && !s.out0 && shp.NA == 0 &&
        (op == LDDMM_OP_KNORM || op == LDDMM_OP_CURRENTS || op == LDDMM_OP_MEASURE ||
         op == LDDMM_OP_VARIFOLD);
    if (prec == LDDMM_PREC_FP64E && !energy_only) return LDDMM_EINVAL;  // energies only
    int64_t chunk_tiles =
        ceil_div(src_tiles * row_tiles, items_target(op, prec, energy_only, s.ny));
    if (chunk_tiles < 1) chunk_tiles = 1;
    const int64_t chunk_len = chunk_tiles * TS;
    const int64_t c0 = ceil_div(nx0, chunk_len), c1 = nx1 ? ceil_div(nx1, chunk_len) : 0;
    if (row_tiles > INT32_MAX || chunk_len > INT32_MAX) return LDDMM_EINVAL;
    // A dispatch order covers exactly this plan's items.
    if (s.item_order && s.n_order != row_tiles * (c0 + c1)) return LDDMM_EINVAL;
    s.row_tiles = (int32_t)row_tiles;
    s.chunk_len = (int32_t)chunk_len;
    s.chunks0 = (int32_t)c0;
    s.chunks1 = (int32_t)c1;
    s.item0 = item;
    item += row_tiles * (c0 + c1);
    s.part0 = part;
    part += s.ny * (c0 + c1) * shp.W;
    s.fin_blocks = (int32_t)ceil_div(s.ny, FIN_BLOCK);
    s.fin0 = fin;
    fin += s.fin_blocks;
    s.red0 = redo;
    redo += (int64_t)s.fin_blocks * (nred > 0 ? nred : 0);
    if (stage > INT32_MAX) return LDDMM_EINVAL;
    s.stage0 = (int32_t)stage;
    stage += stage_tiles(op, s, shp.NA);
  }
  if (n_items) *n_items = item;
  if (n_fin_blocks) *n_fin_blocks = fin;
  if (work_bytes)
    *work_bytes = stage_offset(nseg, part + redo) + (size_t)stage * stage_rec(prec, d, shp);
  return LDDMM_OK;
}

extern "C" int lddmm_set_items_target(int op, int prec, int energy_only, int size_class,
                                      int64_t items) {
  if (!op_valid(op) || (prec != LDDMM_PREC_FP64 && prec != LDDMM_PREC_FP32C) ||
      (energy_only != 0 && energy_only != 1) || size_class < 0 ||
      size_class >= LDDMM_SIZE_CLASSES || items > ((int64_t)1 << 30))
    return LDDMM_EINVAL;
  g_items_target[op][prec == LDDMM_PREC_FP32C][energy_only][size_class] = items > 0 ? items : 0;
  return LDDMM_OK;
}

extern "C" int64_t lddmm_get_items_target(int op, int prec, int energy_only, int size_class) {
  if (!op_valid(op) || (prec != LDDMM_PREC_FP64 && prec != LDDMM_PREC_FP32C) ||
      (energy_only != 0 && energy_only != 1) || size_class < 0 ||
      size_class >= LDDMM_SIZE_CLASSES)
    return LDDMM_EINVAL;
  const int64_t t = g_items_target[op][prec == LDDMM_PREC_FP32C][energy_only][size_class];
  return t > 0 ? t : (energy_only ? ITEMS_ENERGY : ITEMS_PER_SEGMENT);
}

extern "C" int lddmm_size_class(int64_t ny) { return ny < 1 ? LDDMM_EINVAL : size_class(ny); }

namespace {

// Per-launch options of the K-sum kernels: the staging area of the source tiles built once
// per launch (n_stage tiles; none when 0).
struct PairOpts {
  char *stage;
  int64_t n_stage;
  // Totals the staging launch also reduces: the previous step's segments (nseg of them),
  // or none.
  const lddmm_seg_t *tot_segs;
  int nred;
  const double *red;
};

template <typename R, int KT, int D, int M, bool KSUM, int NA>
void launch_pairs(const lddmm_seg_t *segs, int nseg, int op, int prec, int64_t n_items,
                  double *part, hipStream_t st, const PairOpts &po) {
  const char *stage = nullptr;
  // Staged segments: K-sums without derivative terms (flow steps, RKHS products and norms,
  // data-term energies) and the covector step (M = D, two derivative weight pairs).
  if constexpr ((KSUM && NA == 0) || (M == D && !KSUM && NA == 2)) {
    if (po.n_stage > 0) {
      const int64_t extra = po.tot_segs ? nseg : 0;
      hipLaunchKernelGGL((k_stage<R, KT, D, M, KSUM, NA>), dim3((unsigned)(po.n_stage + extra)),
                         dim3(BLOCK), 0, st, segs, nseg, op, prec, po.stage, po.n_stage,
                         po.tot_segs, po.nred, po.red);
      stage = po.stage;
    }
  }
  hipLaunchKernelGGL((k_pairs<R, KT, D, M, KSUM, NA>), dim3((unsigned)n_items), dim3(BLOCK), 0,
                     st, segs, nseg, op, prec, part, stage);
}

template <int D, int M, bool KSUM, int NA>
int dispatch_pairs(int prec, int kern, const lddmm_seg_t *segs, int nseg, int op,
                   int64_t n_items, double *part, hipStream_t st, const PairOpts &po) {
  if (prec == LDDMM_PREC_FP32C) {
    if (kern == LDDMM_KERNEL_GAUSS)
      launch_pairs<float, LDDMM_KERNEL_GAUSS, D, M, KSUM, NA>(segs, nseg, op, prec, n_items, part, st, po);
    else
      launch_pairs<float, LDDMM_KERNEL_LAPLACIAN, D, M, KSUM, NA>(segs, nseg, op, prec, n_items, part, st, po);
  } else {
    if (kern == LDDMM_KERNEL_GAUSS)
      launch_pairs<double, LDDMM_KERNEL_GAUSS, D, M, KSUM, NA>(segs, nseg, op, prec, n_items, part, st, po);
    else
      launch_pairs<double, LDDMM_KERNEL_LAPLACIAN, D, M, KSUM, NA>(segs, nseg, op, prec, n_items, part, st, po);
  }
  return LDDMM_OK;
}

template <int D>
int dispatch_pairs_ddk(int prec, int kern, const lddmm_seg_t *segs, int nseg, int64_t n_items,
                       double *part, hipStream_t st) {
  dim3 g((unsigned)n_items), b(BLOCK);
  if (prec == LDDMM_PREC_FP32C) {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_ddk<float, LDDMM_KERNEL_GAUSS, D>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_ddk<float, LDDMM_KERNEL_LAPLACIAN, D>), g, b, 0, st, segs, nseg, prec, part);
  } else {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_ddk<double, LDDMM_KERNEL_GAUSS, D>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_ddk<double, LDDMM_KERNEL_LAPLACIAN, D>), g, b, 0, st, segs, nseg, prec, part);
  }
  return LDDMM_OK;
}

template <int D>
void dispatch_pairs_dk(int prec, int kern, const lddmm_seg_t *segs, int nseg, int64_t n_items,
                       double *part, hipStream_t st) {
  dim3 g((unsigned)n_items), b(BLOCK);
  if (prec == LDDMM_PREC_FP32C) {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_dk<float, LDDMM_KERNEL_GAUSS, D>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_dk<float, LDDMM_KERNEL_LAPLACIAN, D>), g, b, 0, st, segs, nseg, prec, part);
  } else {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_dk<double, LDDMM_KERNEL_GAUSS, D>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_dk<double, LDDMM_KERNEL_LAPLACIAN, D>), g, b, 0, st, segs, nseg, prec, part);
  }
}

template <int D>
void dispatch_pairs_div(int prec, int kern, const lddmm_seg_t *segs, int nseg, int64_t n_items,
                        double *part, hipStream_t st) {
  dim3 g((unsigned)n_items), b(BLOCK);
  if (prec == LDDMM_PREC_FP32C) {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_div<float, LDDMM_KERNEL_GAUSS, D>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_div<float, LDDMM_KERNEL_LAPLACIAN, D>), g, b, 0, st, segs, nseg, prec, part);
  } else {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_div<double, LDDMM_KERNEL_GAUSS, D>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_div<double, LDDMM_KERNEL_LAPLACIAN, D>), g, b, 0, st, segs, nseg, prec, part);
  }
}

template <bool GRAD>
void dispatch_pairs_varifold(int prec, int kern, const lddmm_seg_t *segs, int nseg,
                             int64_t n_items, double *part, hipStream_t st) {
  dim3 g((unsigned)n_items), b(BLOCK);
  if (prec == LDDMM_PREC_FP32C) {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_varifold<float, LDDMM_KERNEL_GAUSS, GRAD>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_varifold<float, LDDMM_KERNEL_LAPLACIAN, GRAD>), g, b, 0, st, segs, nseg, prec, part);
  } else {
    if (kern == LDDMM_KERNEL_GAUSS)
      hipLaunchKernelGGL((k_pairs_varifold<double, LDDMM_KERNEL_GAUSS, GRAD>), g, b, 0, st, segs, nseg, prec, part);
    else
      hipLaunchKernelGGL((k_pairs_varifold<double, LDDMM_KERNEL_LAPLACIAN, GRAD>), g, b, 0, st, segs, nseg, prec, part);
  }
}

template <int OP, int D, int M, bool KSUM, int NA>
void launch_fin(const lddmm_seg_t *segs, int nseg, int prec, int kern, int64_t n_fin,
                const double *part, double *red, hipStream_t st) {
  hipLaunchKernelGGL((k_finalize<OP, D, M, KSUM, NA>), dim3((unsigned)n_fin), dim3(FIN_BLOCK), 0, st,
                     segs, nseg, prec, kern, part, red);
}

template <int D>
int run_d(int op, int prec, int kern, const Shape &sh, const lddmm_seg_t *segs, int nseg,
          int64_t n_items, int64_t n_fin, double *part, double *red, hipStream_t st,
          const PairOpts &po) {
  switch (op) {
    case LDDMM_OP_APPLY_K:
      switch (sh.M) {
        case 1:
          dispatch_pairs<D, 1, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
          launch_fin<LDDMM_OP_APPLY_K, D, 1, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
          break;
        case 2:
          dispatch_pairs<D, 2, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
          launch_fin<LDDMM_OP_APPLY_K, D, 2, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
          break;
        case 3:
          dispatch_pairs<D, 3, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
          launch_fin<LDDMM_OP_APPLY_K, D, 3, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
          break;
        default:
          return LDDMM_EINVAL;
      }
      return LDDMM_OK;
    case LDDMM_OP_DIFF_KT: {
#define LDDMM_DKT(MM, NN)                                                                  \
  if (sh.M == MM && sh.NA == NN) {                                                         \
    dispatch_pairs<D, MM, false, NN>(prec, kern, segs, nseg, op, n_items, part, st, po);       \
    launch_fin<LDDMM_OP_DIFF_KT, D, MM, false, NN>(segs, nseg, prec, kern, n_fin, part, red, st); \
    return LDDMM_OK;                                                                       \
  }
      LDDMM_DKT(1, 1) LDDMM_DKT(1, 2) LDDMM_DKT(1, 3)
      LDDMM_DKT(D, 1) LDDMM_DKT(D, 2) LDDMM_DKT(D, 3)
#undef LDDMM_DKT
      return LDDMM_EINVAL;
    }
    case LDDMM_OP_FLOW_STEP:
      dispatch_pairs<D, D, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
      launch_fin<LDDMM_OP_FLOW_STEP, D, D, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
    case LDDMM_OP_KDOT:
      dispatch_pairs<D, D, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
      launch_fin<LDDMM_OP_KDOT, D, D, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
    case LDDMM_OP_KNORM:
      dispatch_pairs<D, D, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
      launch_fin<LDDMM_OP_KNORM, D, D, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
    case LDDMM_OP_COVECTOR_STEP:
      dispatch_pairs<D, D, false, 2>(prec, kern, segs, nseg, op, n_items, part, st, po);
      launch_fin<LDDMM_OP_COVECTOR_STEP, D, D, false, 2>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
    case LDDMM_OP_CURRENTS:
      if (sh.NA > 0) {
        dispatch_pairs<D, D, true, 1>(prec, kern, segs, nseg, op, n_items, part, st, po);
        launch_fin<LDDMM_OP_CURRENTS, D, D, true, 1>(segs, nseg, prec, kern, n_fin, part, red, st);
      } else {
        dispatch_pairs<D, D, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
        launch_fin<LDDMM_OP_CURRENTS, D, D, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
      }
      return LDDMM_OK;
    case LDDMM_OP_MEASURE:
      if (sh.NA > 0) {
        dispatch_pairs<D, 1, true, 1>(prec, kern, segs, nseg, op, n_items, part, st, po);
        launch_fin<LDDMM_OP_MEASURE, D, 1, true, 1>(segs, nseg, prec, kern, n_fin, part, red, st);
      } else {
        dispatch_pairs<D, 1, true, 0>(prec, kern, segs, nseg, op, n_items, part, st, po);
        launch_fin<LDDMM_OP_MEASURE, D, 1, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
      }
      return LDDMM_OK;
    case LDDMM_OP_VARIFOLD:
      if constexpr (D == 3) {
        if (sh.NA > 0) {
          dispatch_pairs_varifold<true>(prec, kern, segs, nseg, n_items, part, st);
          launch_fin<LDDMM_OP_VARIFOLD, 3, 5, true, 1>(segs, nseg, prec, kern, n_fin, part, red, st);
        } else {
          dispatch_pairs_varifold<false>(prec, kern, segs, nseg, n_items, part, st);
          launch_fin<LDDMM_OP_VARIFOLD, 3, 1, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
        }
        return LDDMM_OK;
      }
      return LDDMM_EINVAL;
    case LDDMM_OP_DIFF_K:
      dispatch_pairs_dk<D>(prec, kern, segs, nseg, n_items, part, st);
      launch_fin<LDDMM_OP_DIFF_K, D, D, false, 1>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
    case LDDMM_OP_DIVERGENCE:
      dispatch_pairs_div<D>(prec, kern, segs, nseg, n_items, part, st);
      launch_fin<LDDMM_OP_DIVERGENCE, D, 1, true, 0>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
    case LDDMM_OP_DDIFF_K11_12:
      dispatch_pairs_ddk<D>(prec, kern, segs, nseg, n_items, part, st);
      launch_fin<LDDMM_OP_DDIFF_K11_12, D, D, false, 1>(segs, nseg, prec, kern, n_fin, part, red, st);
      return LDDMM_OK;
  }
  return LDDMM_EINVAL;
}

}  // namespace

namespace {

// One launch of planned segments.  tot_prev: device descriptors of the previous launch of
// a stepped sequence whose totals the staging launch reduces (NULL: none); skip_totals: the
// next launch of the sequence reduces this one's totals.
int segments_run_impl(int op, int prec, int d, const lddmm_seg_t *segs_host,
                      const lddmm_seg_t *segs_dev, int nseg, int64_t n_items,
                      int64_t n_fin_blocks, void *work, size_t work_bytes, void *stream,
                      const lddmm_seg_t *tot_prev, bool skip_totals) {
  if (!op_valid(op) || nseg < 1 || !segs_host || (d != 2 && d != 3) || !work) return LDDMM_EINVAL;
  if (prec != LDDMM_PREC_FP64 && prec != LDDMM_PREC_FP32C && prec != LDDMM_PREC_FP64E)
    return LDDMM_EINVAL;
  if (n_items < 1 || n_items > INT32_MAX || n_fin_blocks < 1 || n_fin_blocks > INT32_MAX)
    return LDDMM_EINVAL;
  const lddmm_seg_t &last = segs_host[nseg - 1];
  Shape sh;
  int rc = op_shape(op, d, segs_host[0], &sh);
  if (rc) return rc;
  const int64_t parts = last.part0 + last.ny * (int64_t)(last.chunks0 + last.chunks1) * sh.W;
  const int nred = op_nred(op);
  const int64_t reds = last.red0 + (int64_t)last.fin_blocks * nred;
  // Staged tiles: the plan numbers the ordered segments' tiles consecutively.
  int64_t n_stage = 0;
  for (int i = 0; i < nseg; ++i) {
    if (segs_host[i].stage0 != n_stage) return LDDMM_EINVAL;  // not this order's plan
    n_stage += stage_tiles(op, segs_host[i], sh.NA);
  }
  const size_t need = stage_offset(nseg, parts + reds) + (size_t)n_stage * stage_rec(prec, d, sh);
  if (work_bytes < need) return LDDMM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  char *wb = (char *)work;
  const lddmm_seg_t *segs = segs_dev;
  if (!segs) {
    lddmm_seg_t *dst = (lddmm_seg_t *)wb;
    for (int i0 = 0; i0 < nseg; i0 += STAMP) {
      SegPack pack;
      memset(&pack, 0, sizeof(pack));
      int n = nseg - i0 < STAMP ? nseg - i0 : STAMP;
      for (int i = 0; i < n; ++i) pack.s[i] = segs_host[i0 + i];
      hipLaunchKernelGGL(k_stamp, dim3(1), dim3(STAMP), 0, st, pack, n, dst + i0);
    }
    segs = dst;
  }
  double *part = (double *)(wb + desc_bytes(nseg));
  double *red = part + parts;
  const int kern = segs_host[0].kernel;
  PairOpts po;
  po.stage = wb + stage_offset(nseg, parts + reds);
  po.n_stage = n_stage;
  po.tot_segs = (n_stage > 0 && nred > 0) ? tot_prev : nullptr;
  po.nred = nred;
  po.red = red;
  rc = d == 2 ? run_d<2>(op, prec, kern, sh, segs, nseg, n_items, n_fin_blocks, part, red, st, po)
              : run_d<3>(op, prec, kern, sh, segs, nseg, n_items, n_fin_blocks, part, red, st, po);
  if (rc) return rc;
  if (nred > 0 && !skip_totals)
    hipLaunchKernelGGL(k_totals, dim3((unsigned)nseg), dim3(TOT_BLOCK), 0, st, segs, nred, red);
  return hip_check(hipGetLastError());
}

}  // namespace

extern "C" int lddmm_segments_run(int op, int prec, int d, const lddmm_seg_t *segs_host,
                                  const lddmm_seg_t *segs_dev, int nseg, int64_t n_items,
                                  int64_t n_fin_blocks, void *work, size_t work_bytes,
                                  void *stream) {
  return segments_run_impl(op, prec, d, segs_host, segs_dev, nseg, n_items, n_fin_blocks, work,
                           work_bytes, stream, nullptr, false);
}

extern "C" int lddmm_segments_run_steps(int op, int prec, int d, const lddmm_seg_t *segs_host,
                                        const lddmm_seg_t *segs_dev, int n_steps, int nseg,
                                        int64_t n_items, int64_t n_fin_blocks, void *work,
                                        size_t work_bytes, void *stream) {
  if (n_steps < 1 || nseg < 1 || !segs_host || !segs_dev) return LDDMM_EINVAL;
  if (!op_valid(op) || (d != 2 && d != 3)) return LDDMM_EINVAL;
  // Flow steps that stage tiles (ordered segments) reduce the previous step's totals in
  // their staging launch (launch_pairs), sparing a totals launch per step.
  Shape sh;
  if (op_shape(op, d, segs_host[0], &sh)) return LDDMM_EINVAL;
  int64_t n_stage = 0;
  for (int i = 0; i < nseg; ++i) n_stage += stage_tiles(op, segs_host[i], sh.NA);
  const bool fold = op == LDDMM_OP_FLOW_STEP && n_stage > 0;
  for (int t = 0; t < n_steps; ++t) {
    const lddmm_seg_t *h = segs_host + (int64_t)t * nseg;
    // Steps share step 0's plan: same sizes and scheduling fields, segment by segment.
    for (int i = 0; t > 0 && i < nseg; ++i) {
      const lddmm_seg_t &a = h[i], &b = segs_host[i];
      if (a.ny != b.ny || a.nx0 != b.nx0 || a.nx1 != b.nx1 || a.item0 != b.item0 ||
          a.part0 != b.part0 || a.fin0 != b.fin0 || a.red0 != b.red0 ||
          a.row_tiles != b.row_tiles || a.chunks0 != b.chunks0 || a.chunks1 != b.chunks1 ||
          a.chunk_len != b.chunk_len || a.fin_blocks != b.fin_blocks || a.stage0 != b.stage0 ||
          !a.perm != !b.perm || !a.rowdot != !b.rowdot || !a.out0 != !b.out0 ||
          !a.item_order != !b.item_order || a.n_order != b.n_order)
        return LDDMM_EINVAL;
    }
    int rc = segments_run_impl(op, prec, d, h, segs_dev + (int64_t)t * nseg, nseg, n_items,
                               n_fin_blocks, work, work_bytes, stream,
                               fold && t > 0 ? segs_dev + (int64_t)(t - 1) * nseg : nullptr,
                               fold && t < n_steps - 1);
    if (rc) return rc;
  }
  return LDDMM_OK;
}

// ---------------------------------------------------------------------------
// One-call forms of the reference's kernel sums.
// ---------------------------------------------------------------------------
namespace {

lddmm_seg_t blank_seg() {
  lddmm_seg_t s;
  memset(&s, 0, sizeof(s));
  return s;
}

int plan_one(int op, int prec, int d, lddmm_seg_t *s, int64_t *items, int64_t *fins,
             size_t *bytes) {
  return lddmm_segments_plan(op, prec, d, s, 1, items, fins, bytes);
}

// Workspace of a one-call form, whatever the precision: plan both and take the max.
size_t ws_query(int op, int d, lddmm_seg_t s) {
  size_t best = 0;
  for (int prec = 0; prec < 2; ++prec) {
    lddmm_seg_t t = s;
    int64_t a, b;
    size_t wb = 0;
    if (plan_one(op, prec, d, &t, &a, &b, &wb) == LDDMM_OK && wb > best) best = wb;
  }
  return best;
}

// Dummy non-null device-side pointers for workspace queries (never dereferenced).
const double *const QP = (const double *)(uintptr_t)256;

}  // namespace

extern "C" size_t lddmm_apply_k_workspace_bytes(int64_t ny, int64_t nx, int d, int m) {
  if (ny < 1 || nx < 1 || m < 1) return 0;
  size_t best = 0;
  for (int c0 = 0; c0 < m; c0 += 3) {
    lddmm_seg_t s = blank_seg();
    s.y = QP; s.x0 = QP; s.w0 = QP; s.out0 = (double *)QP;
    s.ny = ny; s.nx0 = nx; s.sigma = 1.0; s.kernel = LDDMM_KERNEL_GAUSS;
    s.m = (m - c0) < 3 ? (m - c0) : 3;
    size_t b = ws_query(LDDMM_OP_APPLY_K, d < 2 ? 2 : d, s);
    if (b > best) best = b;
  }
  return best;
}

namespace {

// lddmm_apply_k's body, for a caller-given workspace or the allocation-free form's scratch.
int apply_k_impl(const double *y, const double *x, const double *w, int64_t ny, int64_t nx,
                 int d, int m, int kernel, int order, double sigma, int prec, double *out,
                 void *work, size_t work_bytes, void *stream) {
  if (!y || !x || !w || !out || ny < 1 || nx < 1 || m < 1 || d < 2 || d > 3) return LDDMM_EINVAL;
  // Columns are processed in groups of up to 3: the weight/output row stride
  // stays m while the kernel's column count is the group width.
  for (int c0 = 0; c0 < m; c0 += 3) {
    lddmm_seg_t s = blank_seg();
    s.y = y; s.x0 = x; s.w0 = w + c0; s.out0 = out + c0;
    s.ny = ny; s.nx0 = nx; s.sigma = sigma; s.kernel = kernel; s.order = order;
    s.m = m;
    // The kernel width of this group travels in `na` (unused by APPLY_K).
    int width = (m - c0) < 3 ? (m - c0) : 3;
    lddmm_seg_t plan = s;
    plan.m = width;  // plan with the group width so the partial width matches
    int64_t items, fins;
    size_t need;
    int rc = plan_one(LDDMM_OP_APPLY_K, prec, d, &plan, &items, &fins, &need);
    if (rc) return rc;
    if (need > work_bytes) return LDDMM_EWORKSPACE;
    // The stored descriptor keeps the full row stride m for addressing.
    plan.m = m;
    lddmm_seg_t shape = plan;
    shape.m = width;
    // Launch shape from `shape`, device descriptor from `plan`.
    SegPack pack;
    memset(&pack, 0, sizeof(pack));
    pack.s[0] = plan;
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(STAMP), 0, (hipStream_t)stream, pack, 1,
                       (lddmm_seg_t *)work);
    rc = lddmm_segments_run(LDDMM_OP_APPLY_K, prec, d, &shape, (const lddmm_seg_t *)work, 1,
                            items, fins, work, work_bytes, stream);
    if (rc) return rc;
  }
  return LDDMM_OK;
}

// Workspace bytes apply_k_impl needs at this precision (the largest column group's plan).
int apply_k_need(int64_t ny, int64_t nx, int d, int m, int kernel, int order, double sigma,
                 int prec, size_t *need) {
  *need = 0;
  for (int c0 = 0; c0 < m; c0 += 3) {
    lddmm_seg_t s = blank_seg();
    s.y = QP; s.x0 = QP; s.w0 = QP; s.out0 = (double *)QP;
    s.ny = ny; s.nx0 = nx; s.sigma = sigma; s.kernel = kernel; s.order = order;
    s.m = (m - c0) < 3 ? (m - c0) : 3;
    int64_t items, fins;
    size_t b = 0;
    int rc = plan_one(LDDMM_OP_APPLY_K, prec, d, &s, &items, &fins, &b);
    if (rc) return rc;
    if (b > *need) *need = b;
  }
  return LDDMM_OK;
}

// Scratch of the allocation-free entry point: one fixed array per GPU, in the code
// object's zero-filled data, held until the process ends.  Its host-side name is a
// stand-in, never a device address: host code uses the address hipGetSymbolAddress gives
// for the current device, looked up once per device on that device's first call.
constexpr size_t NOALLOC_SCRATCH = LDDMM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[MAX_DEVICES];  // device address of `scratch` on each GPU (null: not looked up)

int scratch_on_current_device(void **ptr) {
  int dev = -1;
  int rc = hip_check(hipGetDevice(&dev));
  if (rc) return rc;
  if (dev < 0 || dev >= MAX_DEVICES) return LDDMM_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (rc) return rc;
    if (!p) return LDDMM_EHIP;
    g_scratch_dev[dev] = p;
  }
  *ptr = g_scratch_dev[dev];
  return LDDMM_OK;
}

}  // namespace

extern "C" int lddmm_apply_k(const double *y, const double *x, const double *w, int64_t ny,
                             int64_t nx, int d, int m, int kernel, int order, double sigma,
                             int prec, double *out, void *work, size_t work_bytes,
                             void *stream) {
  return apply_k_impl(y, x, w, ny, nx, d, m, kernel, order, sigma, prec, out, work, work_bytes,
                      stream);
}

extern "C" int lddmm_apply_k_noalloc(const double *y, const double *x, const double *w,
                                     int64_t ny, int64_t nx, int d, int m, int kernel,
                                     int order, double sigma, int prec, double *out,
                                     void *stream) {
  if (!y || !x || !w || !out || ny < 1 || nx < 1 || m < 1 || d < 2 || d > 3) return LDDMM_EINVAL;
  // The whole request is checked against the fixed scratch before anything is launched.
  size_t need = 0;
  int rc = apply_k_need(ny, nx, d, m, kernel, order, sigma, prec, &need);
  if (rc) return rc;
  if (need > NOALLOC_SCRATCH) return LDDMM_ESCRATCH;
  void *ptr = nullptr;
  rc = scratch_on_current_device(&ptr);
  if (rc) return rc;
  // Nothing in the scratch has to start at zero: the descriptor is stamped and every
  // chunk partial is written before it is read, by the same kernels as lddmm_apply_k's.
  return apply_k_impl(y, x, w, ny, nx, d, m, kernel, order, sigma, prec, out, ptr,
                      NOALLOC_SCRATCH, stream);
}

extern "C" size_t lddmm_apply_diff_kt_workspace_bytes(int64_t ny, int64_t nx, int d, int m,
                                                      int na) {
  if (ny < 1 || nx < 1 || m < 1 || na < 1) return 0;
  lddmm_seg_t s = blank_seg();
  s.y = QP; s.x0 = QP; s.out0 = (double *)QP;
  for (int j = 0; j < 4; ++j) { s.tw[j] = QP; s.sw[j] = QP; }
  s.ny = ny; s.nx0 = nx; s.sigma = 1.0; s.kernel = LDDMM_KERNEL_GAUSS;
  s.m = m; s.na = na < 3 ? na : 3;
  return ws_query(LDDMM_OP_DIFF_KT, d, s);
}

extern "C" int lddmm_apply_diff_kt(const double *y, const double *x, const double *a1,
                                   const double *a2, int64_t ny, int64_t nx, int d, int m,
                                   int na, int kernel, int order, double sigma, int prec,
                                   double *out, void *work, size_t work_bytes, void *stream) {
  if (!y || !x || !a1 || !a2 || !out || ny < 1 || nx < 1 || d < 2 || d > 3) return LDDMM_EINVAL;
  if (!(m == 1 || m == d) || na < 1 || na > 3) return LDDMM_EUNSUPPORTED;
  lddmm_seg_t s = blank_seg();
  s.y = y; s.x0 = x; s.out0 = out;
  for (int j = 0; j < na; ++j) {
    s.tw[j] = a1 + (int64_t)j * ny * m;
    s.sw[j] = a2 + (int64_t)j * nx * m;
  }
  s.ny = ny; s.nx0 = nx; s.sigma = sigma; s.kernel = kernel; s.order = order;
  s.m = m; s.na = na;
  int64_t items, fins;
  size_t need;
  int rc = plan_one(LDDMM_OP_DIFF_KT, prec, d, &s, &items, &fins, &need);
  if (rc) return rc;
  if (need > work_bytes) return LDDMM_EWORKSPACE;
  return lddmm_segments_run(LDDMM_OP_DIFF_KT, prec, d, &s, nullptr, 1, items, fins, work,
                            work_bytes, stream);
}

extern "C" size_t lddmm_apply_ddiff_k11and12_workspace_bytes(int64_t n, int d) {
  if (n < 1) return 0;
  lddmm_seg_t s = blank_seg();
  s.y = QP; s.w0 = QP; s.w1 = QP; s.w2 = QP; s.out0 = (double *)QP;
  s.ny = n; s.sigma = 1.0; s.kernel = LDDMM_KERNEL_GAUSS;
  return ws_query(LDDMM_OP_DDIFF_K11_12, d, s);
}

extern "C" int lddmm_apply_ddiff_k11and12(const double *x, const double *a1, const double *a2,
                                          const double *p, int64_t n, int d, int kernel,
                                          int order, double sigma, int prec, double *out,
                                          void *work, size_t work_bytes, void *stream) {
  if (!x || !a1 || !a2 || !p || !out || n < 1 || d < 2 || d > 3) return LDDMM_EINVAL;
  lddmm_seg_t s = blank_seg();
  s.y = x; s.w0 = a1; s.w1 = a2; s.w2 = p; s.out0 = out;
  s.ny = n; s.sigma = sigma; s.kernel = kernel; s.order = order;
  int64_t items, fins;
  size_t need;
  int rc = plan_one(LDDMM_OP_DDIFF_K11_12, prec, d, &s, &items, &fins, &need);
  if (rc) return rc;
  if (need > work_bytes) return LDDMM_EWORKSPACE;
  return lddmm_segments_run(LDDMM_OP_DDIFF_K11_12, prec, d, &s, nullptr, 1, items, fins, work,
                            work_bytes, stream);
}

// ---------------------------------------------------------------------------
// Matrix-weight sums (applykmat / applykdiffmat, kernelFunctions.f90:499-610):
// one weight per (target, source) pair, beta row-major (ny, nx[, m]).  Reading
// beta (8 ny nx m bytes) bounds them, so each target row is one wave whose lanes
// stride over the sources (coalesced beta reads), everything in FP64 with the
// Fortran tables and cut-off, and the 64 lane sums are combined by a fixed
// butterfly (deterministic).
// ---------------------------------------------------------------------------
template <int KT, int D, bool DIFF>
__global__ __launch_bounds__(256) void k_kmat(const double *__restrict__ y, const double *__restrict__ x,
                                             const double *__restrict__ beta, int64_t ny, int64_t nx,
                                             int m, int order, double sigma, double *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= ny) return;  // wave-uniform; no barriers below
  KEval<double, KT> kev;
  kev.init(order, sigma);
  double yk[D];
#pragma unroll
  for (int c = 0; c < D; ++c) yk[c] = y[k * D + c];
  double acc[3] = {0.0, 0.0, 0.0};
  const int mw = DIFF ? 1 : m;
  const double *brow = beta + k * nx * mw;
  for (int64_t l = lane; l < nx; l += 64) {
    double dx[D], r2 = 0.0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dx[c] = yk[c] - x[l * D + c];
      r2 += dx[c] * dx[c];
    }
    double k0, k1, k2;
    kev(r2, k0, k1, k2);
    if constexpr (DIFF) {
      const double f = k1 * brow[l];
#pragma unroll
      for (int c = 0; c < D; ++c) acc[c] += f * dx[c];
    } else {
      for (int c = 0; c < mw; ++c) acc[c] += k0 * brow[l * mw + c];
    }
  }
  const int nout = DIFF ? D : m;
  for (int c = 0; c < nout; ++c) {
    double v = acc[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    // 2K'(r^2) = -k1 / sigma^2 in the FP64 kernel evaluation
    if (lane == 0) out[k * nout + c] = DIFF ? -v / (sigma * sigma) : v;
  }
}

namespace {
template <bool DIFF>
int launch_kmat(const double *y, const double *x, const double *beta, int64_t ny, int64_t nx,
                int d, int m, int kernel, int order, double sigma, double *out, void *stream) {
  if (!y || !x || !beta || !out || ny < 1 || nx < 1 || (d != 2 && d != 3)) return LDDMM_EINVAL;
  if (!DIFF && (m < 1 || m > 3)) return LDDMM_EINVAL;
  if (!(sigma > 0.0) || !isfinite(sigma)) return LDDMM_EINVAL;
  if (kernel == LDDMM_KERNEL_LAPLACIAN) {
    if (order < (DIFF ? 1 : 0) || order > 4) return LDDMM_EINVAL;
  } else if (kernel != LDDMM_KERNEL_GAUSS) {
    return LDDMM_EINVAL;
  }
  const int64_t blocks = ceil_div(ny, 4);
  if (blocks > INT32_MAX) return LDDMM_EINVAL;
  dim3 g((unsigned)blocks), b(256);
  hipStream_t st = (hipStream_t)stream;
#define LDDMM_KMAT(KT, DD) \
  hipLaunchKernelGGL((k_kmat<KT, DD, DIFF>), g, b, 0, st, y, x, beta, ny, nx, m, order, sigma, out)
  if (kernel == LDDMM_KERNEL_GAUSS) {
    if (d == 2) LDDMM_KMAT(LDDMM_KERNEL_GAUSS, 2); else LDDMM_KMAT(LDDMM_KERNEL_GAUSS, 3);
  } else {
    if (d == 2) LDDMM_KMAT(LDDMM_KERNEL_LAPLACIAN, 2); else LDDMM_KMAT(LDDMM_KERNEL_LAPLACIAN, 3);
  }
#undef LDDMM_KMAT
  return hip_check(hipGetLastError());
}
}  // namespace

extern "C" int lddmm_apply_k_mat(const double *y, const double *x, const double *beta, int64_t ny,
                                 int64_t nx, int d, int m, int kernel, int order, double sigma,
                                 double *out, void *stream) {
  return launch_kmat<false>(y, x, beta, ny, nx, d, m, kernel, order, sigma, out, stream);
}

extern "C" int lddmm_apply_diff_k_mat(const double *y, const double *x, const double *beta,
                                      int64_t ny, int64_t nx, int d, int kernel, int order,
                                      double sigma, double *out, void *stream) {
  return launch_kmat<true>(y, x, beta, ny, nx, d, 1, kernel, order, sigma, out, stream);
}

// ---------------------------------------------------------------------------
// Dense kernel matrices (Kernel.precompute, kernelFunctions.py:10-55,152-198,285-298):
// out[k, l] = K(y_k, x_l) (deriv 0), K'(r^2) (1) or K''(r^2) (2), row-major (ny, nx),
// FP64 with the dense path's exact-double Laplacian polynomials lapPol / lapPolDiff /
// lapPolDiff2 (:58-104) and no cut-off.  self_set (firstVar and grid absent, pdist
// + squareform there) fills the diagonal as the reference does: Gaussian K = 1
// before the derivative scaling; Laplacian 1, -1/((2 order - 1) 2 sigma^2) and
// 1/(35 * 4 sigma^4).  One thread per entry, consecutive threads along a row.
// ---------------------------------------------------------------------------
__device__ inline double lap_exact(int order, int deriv, double u) {
  if (deriv == 0) {
    switch (order) {
      case 0: return 1.0;
      case 1: return 1.0 + u;
      case 2: return (3.0 + 3.0 * u + u * u) / 3.0;
      case 3: return (15.0 + 15.0 * u + 6.0 * u * u + u * u * u) / 15.0;
      default: return (105.0 + 105.0 * u + 45.0 * u * u + 10.0 * u * u * u + u * u * u * u) / 105.0;
    }
  }
  if (deriv == 1) {
    switch (order) {
      case 1: return 1.0;
      case 2: return (1.0 + u) / 3.0;
      case 3: return (3.0 + 3.0 * u + u * u) / 15.0;
      default: return (15.0 + 15.0 * u + 6.0 * u * u + u * u * u) / 105.0;
    }
  }
  switch (order) {
    case 2: return 1.0 / 3.0;
    case 3: return (1.0 + u) / 15.0;
    default: return (3.0 + 3.0 * u + u * u) / 105.0;
  }
}

__global__ __launch_bounds__(256) void k_kernel_matrix(const double *__restrict__ y,
                                                      const double *__restrict__ x, int64_t ny,
                                                      int64_t nx, int d, int kernel, int order,
                                                      double sigma, int deriv, int self_set,
                                                      double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= ny * nx) return;
  const int64_t k = i / nx, l = i - k * nx;
  const double s2 = 2.0 * sigma * sigma;
  double r2 = 0.0;
  for (int c = 0; c < d; ++c) {
    const double t = y[k * d + c] - x[l * d + c];
    r2 += t * t;
  }
  double v;
  if (kernel == LDDMM_KERNEL_GAUSS) {
    const double K = (self_set && k == l) ? 1.0 : exp(-r2 / s2);
    v = deriv == 0 ? K : (deriv == 1 ? -K / s2 : K / (s2 * s2));
  } else if (self_set && k == l) {
    v = deriv == 0 ? 1.0
                   : (deriv == 1 ? -1.0 / ((2 * order - 1) * s2) : 1.0 / (35.0 * 4.0 * sigma * sigma * sigma * sigma));
  } else {
    const double u = sqrt(r2) / sigma;
    const double e = exp(-u);
    const double p = lap_exact(order, deriv, u);
    v = deriv == 0 ? p * e : (deriv == 1 ? -p * e / s2 : p * e / (4.0 * sigma * sigma * sigma * sigma));
  }
  out[i] = v;
}

extern "C" int lddmm_kernel_matrix(const double *y, const double *x, int64_t ny, int64_t nx, int d,
                                   int kernel, int order, double sigma, int deriv, int self_set,
                                   double *out, void *stream) {
  if (!y || !x || !out || ny < 1 || nx < 1 || d < 1 || d > 8) return LDDMM_EINVAL;
  if (deriv < 0 || deriv > 2 || !(sigma > 0.0) || !isfinite(sigma)) return LDDMM_EINVAL;
  if (self_set && ny != nx) return LDDMM_EINVAL;
  if (kernel == LDDMM_KERNEL_LAPLACIAN) {
    if (order < deriv || order > 4 || (deriv == 2 && order < 2)) return LDDMM_EINVAL;
  } else if (kernel != LDDMM_KERNEL_GAUSS) {
    return LDDMM_EINVAL;
  }
  const int64_t n = ny * nx;
  if (nx > 0 && n / nx != ny) return LDDMM_EINVAL;
  const int64_t blocks = ceil_div(n, 256);
  if (blocks > INT32_MAX) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_kernel_matrix, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     y, x, ny, nx, d, kernel, order, sigma, deriv, self_set, out);
  return hip_check(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Geometry entry points.
// ---------------------------------------------------------------------------
extern "C" int lddmm_surface_geometry(const double *vertices, const int32_t *faces, int64_t nf,
                                      double *centers, double *surfel, void *stream) {
  if (!vertices || !faces || !centers || !surfel || nf < 1) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_surface_geometry, dim3((unsigned)ceil_div(nf, 256)), dim3(256), 0,
                     (hipStream_t)stream, vertices, faces, nf, centers, surfel);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_curve_geometry(const double *vertices, const int32_t *edges, int64_t ne,
                                    int d, double *centers, double *lenel, void *stream) {
  if (!vertices || !edges || !centers || !lenel || ne < 1 || d < 1 || d > 8) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_curve_geometry, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0,
                     (hipStream_t)stream, vertices, edges, ne, d, centers, lenel);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_surface_vertex_gradient(const double *vertices, const int32_t *faces,
                                             const double *z1, const double *dz,
                                             const int32_t *incid_ptr, const int32_t *incid,
                                             int64_t nv, double scale, double *px,
                                             void *stream) {
  if (!vertices || !faces || !z1 || !dz || !incid_ptr || !incid || !px || nv < 1)
    return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_surface_vertex_gradient, dim3((unsigned)ceil_div(nv, 256)), dim3(256), 0,
                     (hipStream_t)stream, vertices, faces, z1, dz, incid_ptr, incid, nv, scale,
                     px);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_curve_vertex_gradient(const double *z1, const double *dz,
                                           const int32_t *incid_ptr, const int32_t *incid,
                                           int64_t nv, int d, double scale, double *px,
                                           void *stream) {
  if (!z1 || !dz || !incid_ptr || !incid || !px || nv < 1 || d < 1 || d > 8) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_curve_vertex_gradient, dim3((unsigned)ceil_div(nv, 256)), dim3(256), 0,
                     (hipStream_t)stream, z1, dz, incid_ptr, incid, nv, d, scale, px);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_geometry_batch(const lddmm_geom_t *geoms_host, const lddmm_geom_t *geoms_dev,
                                    int n, void *stream) {
  if (!geoms_host || !geoms_dev || n < 1) return LDDMM_EINVAL;
  for (int i = 0; i < n; ++i) {
    const lddmm_geom_t &e = geoms_host[i];
    if (!e.vertices || !e.faces || !e.centers || !e.normals || e.nf < 1) return LDDMM_EINVAL;
    if (!((e.k == 3 && e.d == 3) || (e.k == 2 && e.d >= 1 && e.d <= 8))) return LDDMM_EINVAL;
    int64_t expect = i == 0 ? 0 : geoms_host[i - 1].blk0 + ceil_div(geoms_host[i - 1].nf, 256);
    if (e.blk0 != expect) return LDDMM_EINVAL;
  }
  const int64_t blocks = geoms_host[n - 1].blk0 + ceil_div(geoms_host[n - 1].nf, 256);
  if (blocks > INT32_MAX) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_geometry_batch, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     geoms_dev, n);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_vertex_gradient_batch(const lddmm_gather_t *g_host, const lddmm_gather_t *g_dev,
                                           int n, void *stream) {
  if (!g_host || !g_dev || n < 1) return LDDMM_EINVAL;
  for (int i = 0; i < n; ++i) {
    const lddmm_gather_t &e = g_host[i];
    if (!e.z1 || !e.dz || !e.incid_ptr || !e.incid || !e.px || e.nv < 1) return LDDMM_EINVAL;
    if (!((e.k == 3 && e.d == 3 && e.vertices && e.faces) || (e.k == 2 && e.d >= 1 && e.d <= 8)))
      return LDDMM_EINVAL;
    int64_t expect = i == 0 ? 0 : g_host[i - 1].blk0 + ceil_div(g_host[i - 1].nv, 256);
    if (e.blk0 != expect) return LDDMM_EINVAL;
  }
  const int64_t blocks = g_host[n - 1].blk0 + ceil_div(g_host[n - 1].nv, 256);
  if (blocks > INT32_MAX) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     g_dev, n);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_landmark_batch(const lddmm_lmk_t *host, const lddmm_lmk_t *dev, int n,
                                    void *stream) {
  if (!host || !dev || n < 1) return LDDMM_EINVAL;
  for (int i = 0; i < n; ++i) {
    const lddmm_lmk_t &e = host[i];
    if (!e.x || !e.y || e.n < 1 || (!e.total && !e.grad)) return LDDMM_EINVAL;
  }
  hipLaunchKernelGGL(k_landmark_batch, dim3((unsigned)n), dim3(1024), 0, (hipStream_t)stream, dev);
  return hip_check(hipGetLastError());
}

extern "C" int lddmm_sum_f64(const double *x, int64_t n, double *out, void *stream) {
  if (!x || !out || n < 1) return LDDMM_EINVAL;
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, (hipStream_t)stream, x, n, out);
  return hip_check(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Processing order (host).  Iterative bisection: each range [lo, lo + len) of
// `order` is split at h (a multiple of 256 for ranges above a tile, of 64
// above a quarter) along the longest side of its bounding box by nth_element;
// ties on the coordinate break by point index, so the result is deterministic.
// ---------------------------------------------------------------------------
namespace {

struct BisectRange {
  int64_t lo, len;
};

// One split of a range of the processing order: the points order[lo, lo + len) are
// partitioned along the longest side of their bounding box at a multiple of 256 (then 64)
// near the middle; false for a leaf (64 points or fewer).  The partition depends on the
// range's contents only (ties broken by index), so ranges may be split in any order.
bool bisect_split(const double *points, int d, int32_t *order, const BisectRange &r,
                  BisectRange *lo_half, BisectRange *hi_half) {
  if (r.len <= 64) return false;
  const int64_t unit = r.len > TS ? TS : 64;
  int64_t h = (int64_t)llround((double)r.len / (2.0 * (double)unit)) * unit;
  if (h < unit) h = unit;
  if (h > r.len - 1) h = r.len - 1;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = r.lo; i < r.lo + r.len; ++i) {
    const double *p = points + (int64_t)order[i] * d;
    for (int c = 0; c < d; ++c) {
      lo[c] = p[c] < lo[c] ? p[c] : lo[c];
      hi[c] = p[c] > hi[c] ? p[c] : hi[c];
    }
  }
  int ax = 0;
  for (int c = 1; c < d; ++c)
    if (hi[c] - lo[c] > hi[ax] - lo[ax]) ax = c;
  int32_t *first = order + r.lo;
  std::nth_element(first, first + h, first + r.len, [&](int32_t a, int32_t b) {
    const double pa = points[(int64_t)a * d + ax], pb = points[(int64_t)b * d + ax];
    return pa < pb || (pa == pb && a < b);
  });
  *lo_half = {r.lo, h};
  *hi_half = {r.lo + h, r.len - h};
  return true;
}

void bisect_all(const double *points, int d, int32_t *order, BisectRange r0) {
  std::vector<BisectRange> stack{r0};
  while (!stack.empty()) {
    const BisectRange r = stack.back();
    stack.pop_back();
    BisectRange a, b;
    if (!bisect_split(points, d, order, r, &a, &b)) continue;
    stack.push_back(b);
    stack.push_back(a);
  }
}

}  // namespace

extern "C" int lddmm_spatial_order(const double *points, int64_t n, int d, int32_t *order) {
  if (n < 0 || n > INT32_MAX || d < 1 || d > 3 || (n > 0 && (!points || !order)))
    return LDDMM_EINVAL;
  for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
  // Large sets: split the top of the tree breadth-first, then finish the subtrees on host
  // threads (disjoint ranges; the same permutation as one thread).
  std::vector<BisectRange> front{{0, n}};
  const int workers = n >= 32768 ? 8 : 1;
  while ((int)front.size() < 2 * workers) {
    std::vector<BisectRange> next;
    bool split = false;
    for (const BisectRange &r : front) {
      BisectRange a, b;
      if (bisect_split(points, d, order, r, &a, &b)) {
        next.push_back(a);
        next.push_back(b);
        split = true;
      } else {
        next.push_back(r);
      }
    }
    front.swap(next);
    if (!split) break;
  }
  if (workers == 1) {
    for (const BisectRange &r : front) bisect_all(points, d, order, r);
    return LDDMM_OK;
  }
  std::vector<std::thread> pool;
  std::atomic<size_t> next_range{0};
  for (int w = 0; w < workers; ++w)
    pool.emplace_back([&]() {
      for (size_t i = next_range++; i < front.size(); i = next_range++)
        bisect_all(points, d, order, front[i]);
    });
  for (std::thread &t : pool) t.join();
  return LDDMM_OK;
}

extern "C" double lddmm_order_compactness(const double *points, int64_t n, int d,
                                          const int32_t *order, double scale, double r2max) {
  if (n < 1 || d < 1 || d > 3 || !points || !order) return -1.0;
  int64_t ok = 0, total = 0;
  for (int64_t base = 0; base < n; base += TS) {
    const int64_t len = n - base < TS ? n - base : TS;
    double org[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 4; ++q) {
      const int64_t k = 32 + 64 * q < len ? 32 + 64 * q : len - 1;
      const double *p = points + (int64_t)order[base + k] * d;
      for (int c = 0; c < d; ++c) org[c] += p[c];
    }
    for (int c = 0; c < d; ++c) org[c] *= 0.25;
    for (int64_t q0 = 0; q0 < len; q0 += 64) {
      double m = 0.0;
      for (int64_t i = q0; i < len && i < q0 + 64; ++i) {
        const double *p = points + (int64_t)order[base + i] * d;
        double r2 = 0.0;
        for (int c = 0; c < d; ++c) {
          const double v = (p[c] - org[c]) * scale;
          r2 += v * v;
        }
        m = r2 > m ? r2 : m;
      }
      ok += m <= r2max;
      ++total;
    }
  }
  return (double)ok / (double)total;
}

// Host geometry of a triangulated surface in the reference's NumPy expression order
// (surfaces.py:62-70 and the volume's triple products), FMA contraction off, so that the
// results are those of the NumPy expressions bit for bit: centers = ((x1 + x2) + x3) / 3,
// surfel = (x2 - x1) x (x3 - x1) with each component a product minus a product, and
// det = x1 . (x2 x x3) summed as ((p0 + p1) + p2).  Face indices below zero count from the
// end, as NumPy indexing does; others out of range are an error.
extern "C" int lddmm_axpy_batch(int n, double *const *out, const double *const *a,
                                const double *const *d, const double *s, const int64_t *len,
                                void *stream) {
  if (n < 0 || (n > 0 && (!out || !a || !d || !s || !len))) return LDDMM_EINVAL;
  for (int i = 0; i < n; ++i)   // every item checked before anything is enqueued
    if (len[i] < 0 || (len[i] > 0 && (!out[i] || !a[i] || !d[i]))) return LDDMM_EINVAL;
  bool launched = false;
  for (int i0 = 0; i0 < n; i0 += AXPY_PACK) {
    AxpyPack pk;
    memset(&pk, 0, sizeof(pk));
    const int m = n - i0 < AXPY_PACK ? n - i0 : AXPY_PACK;
    int64_t longest = 0;
    for (int j = 0; j < m; ++j) {
      const int i = i0 + j;
      pk.out[j] = out[i];
      pk.a[j] = a[i];
      pk.d[j] = d[i];
      pk.s[j] = s[i];
      pk.len[j] = len[i];
      longest = len[i] > longest ? len[i] : longest;
    }
    if (longest == 0) continue;
    // enough workgroups for one element per thread up to ~1 M elements, then grid-stride
    int64_t blocks = (longest + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_axpy_batch, dim3((unsigned)blocks, (unsigned)m), dim3(256), 0,
                       (hipStream_t)stream, pk);
    launched = true;
  }
  return launched ? hip_check(hipGetLastError()) : LDDMM_OK;
}

extern "C" int lddmm_surface_geometry_host(const double *v, int64_t nv, const int64_t *f,
                                           int64_t nf, double *centers, double *surfel,
                                           double *det) {
#pragma clang fp contract(off)
  if (nv < 1 || nf < 0 || !v || (nf > 0 && !f)) return LDDMM_EINVAL;
  for (int64_t i = 0; i < nf; ++i) {
    const double *x[3];
    for (int k = 0; k < 3; ++k) {
      int64_t j = f[3 * i + k];
      if (j < 0) j += nv;
      if (j < 0 || j >= nv) return LDDMM_EINVAL;
      x[k] = v + 3 * j;
    }
    if (centers)
      for (int c = 0; c < 3; ++c) centers[3 * i + c] = (x[0][c] + x[1][c] + x[2][c]) / 3.0;
    if (surfel) {
      const double a[3] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2]};
      const double b[3] = {x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2]};
      surfel[3 * i + 0] = a[1] * b[2] - a[2] * b[1];
      surfel[3 * i + 1] = a[2] * b[0] - a[0] * b[2];
      surfel[3 * i + 2] = a[0] * b[1] - a[1] * b[0];
    }
    if (det) {
      const double *p = x[0], *q = x[1], *r = x[2];
      const double p0 = p[0] * (q[1] * r[2] - q[2] * r[1]);
      const double p1 = p[1] * (q[2] * r[0] - q[0] * r[2]);
      const double p2 = p[2] * (q[0] * r[1] - q[1] * r[0]);
      det[i] = p0 + p1 + p2;
    }
  }
  return LDDMM_OK;
}

extern "C" double lddmm_order_live_fraction(const double *points, int64_t n, int d,
                                            const int32_t *order, int rows, double cut2,
                                            int64_t max_pairs) {
  if (n < 1 || d < 1 || d > 3 || !points || !order || rows < 64 || rows % 64 ||
      !(cut2 >= 0.0) || max_pairs < 1)
    return -1.0;
  // Boxes of the 64-point quarters; a row block is `rows / 64` consecutive quarters.  Rows
  // past the end repeat the last point, which the last quarter's box already holds.
  const int64_t nq = (n + 63) / 64;
  std::vector<double> lo((size_t)nq * d), hi((size_t)nq * d);
  for (int64_t q = 0; q < nq; ++q) {
    for (int c = 0; c < d; ++c) {
      lo[q * d + c] = INFINITY;
      hi[q * d + c] = -INFINITY;
    }
    for (int64_t i = q * 64; i < n && i < q * 64 + 64; ++i) {
      const int64_t k = order[i];
      if (k < 0 || k >= n) return -1.0;
      for (int c = 0; c < d; ++c) {
        const double v = points[k * d + c];
        lo[q * d + c] = v < lo[q * d + c] ? v : lo[q * d + c];
        hi[q * d + c] = v > hi[q * d + c] ? v : hi[q * d + c];
      }
    }
  }
  const int per = rows / 64;
  const int64_t nr = (nq + per - 1) / per;
  int64_t step = (nr * nq + max_pairs - 1) / max_pairs;
  if (step < 1) step = 1;
  int64_t live = 0, total = 0;
  for (int64_t r = 0; r < nr; r += step) {
    double rl[3], rh[3];
    for (int c = 0; c < d; ++c) {
      rl[c] = INFINITY;
      rh[c] = -INFINITY;
    }
    for (int64_t q = r * per; q < nq && q < (r + 1) * per; ++q)
      for (int c = 0; c < d; ++c) {
        rl[c] = lo[q * d + c] < rl[c] ? lo[q * d + c] : rl[c];
        rh[c] = hi[q * d + c] > rh[c] ? hi[q * d + c] : rh[c];
      }
    for (int64_t q = 0; q < nq; ++q) {
      double d2 = 0.0;
      for (int c = 0; c < d; ++c) {
        double g = lo[q * d + c] - rh[c];
        const double g2 = rl[c] - hi[q * d + c];
        g = g2 > g ? g2 : g;
        g = g > 0.0 ? g : 0.0;
        d2 += g * g;
      }
      live += d2 <= cut2;
    }
    total += nq;
  }
  return (double)live / (double)total;
}

// ---------------------------------------------------------------------------
// Dispatch order of a segment's work items (host).  The hardware hands workgroups to
// free slots in launch order; with items of very different cost (culled blocks, the
// symmetric sums' skipped tiles) the last expensive ones keep a few CUs busy while the
// rest idle.  Running the items in decreasing estimated cost leaves only cheap items for
// the end (longest-processing-time-first list scheduling).
// ---------------------------------------------------------------------------
namespace {

struct Boxes {
  std::vector<double> lo, hi;  // (n, 3)
};

// Bounding boxes of consecutive runs of `run` points, of p(i) for i in [0, n).
template <typename P>
Boxes run_boxes(int64_t n, int64_t run, int d, P p) {
  const int64_t nb = (n + run - 1) / run;
  Boxes b;
  b.lo.assign(nb * 3, 0.0);
  b.hi.assign(nb * 3, 0.0);
  for (int64_t k = 0; k < nb; ++k) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = k * run; i < n && i < (k + 1) * run; ++i) {
      const double *v = p(i);
      for (int c = 0; c < d; ++c) {
        lo[c] = v[c] < lo[c] ? v[c] : lo[c];
        hi[c] = v[c] > hi[c] ? v[c] : hi[c];
      }
    }
    for (int c = 0; c < 3; ++c) {
      b.lo[k * 3 + c] = c < d ? lo[c] : 0.0;
      b.hi[k * 3 + c] = c < d ? hi[c] : 0.0;
    }
  }
  return b;
}

double box_gap2(const Boxes &a, int64_t i, const Boxes &b, int64_t j) {
  double d2 = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double g = std::max(std::max(a.lo[i * 3 + c] - b.hi[j * 3 + c],
                                       b.lo[j * 3 + c] - a.hi[i * 3 + c]), 0.0);
    d2 += g * g;
  }
  return d2;
}

// Largest squared distance between a point of box i of a and a point of box j of b.
double box_far2(const Boxes &a, int64_t i, const Boxes &b, int64_t j) {
  double d2 = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double g = std::max(a.hi[i * 3 + c] - b.lo[j * 3 + c], b.hi[j * 3 + c] - a.lo[i * 3 + c]);
    d2 += g * g;
  }
  return d2;
}

}  // namespace

extern "C" int lddmm_item_order(int op, int prec, int d, const lddmm_seg_t *seg, const double *y,
                                const double *x1, const int32_t *perm, int32_t *order,
                                int64_t n) {
  if (!seg || (d != 2 && d != 3)) return LDDMM_EINVAL;
  if (op != LDDMM_OP_FLOW_STEP && op != LDDMM_OP_COVECTOR_STEP && op != LDDMM_OP_KDOT &&
      op != LDDMM_OP_KNORM && op != LDDMM_OP_CURRENTS && op != LDDMM_OP_MEASURE)
    return LDDMM_EUNSUPPORTED;
  // Plan a copy.  Device pointers the plan only checks for presence get stand-ins; the
  // output pointers and rowdot keep the caller's values (they decide the energy-only form).
  lddmm_seg_t s = *seg;
  s.item_order = nullptr;
  s.n_order = 0;
  if (!s.y) s.y = QP;
  if (!s.w0) s.w0 = QP;
  switch (op) {
    case LDDMM_OP_FLOW_STEP:
      if (!s.out0) s.out0 = (double *)QP;
      break;
    case LDDMM_OP_COVECTOR_STEP:
      if (!s.w1) s.w1 = QP;
      if (!s.out0) s.out0 = (double *)QP;
      break;
    case LDDMM_OP_KDOT:
      if (s.na < 1) s.na = 1;
      for (int j = 0; j < s.na && j < 2; ++j)
        if (!s.g[j]) s.g[j] = QP;
      break;
    case LDDMM_OP_CURRENTS:
    case LDDMM_OP_MEASURE:
      if (s.nx1 > 0 && !s.x1) s.x1 = QP;
      if (s.nx1 > 0 && !s.w1) s.w1 = QP;
      if (op == LDDMM_OP_MEASURE && s.out0 && !s.w2) s.w2 = QP;
      break;
  }
  int64_t n_items = 0, n_fin = 0;
  size_t wb = 0;
  int rc = lddmm_segments_plan(op, prec, d, &s, 1, &n_items, &n_fin, &wb);
  if (rc) return rc;
  if (!order) return n_items > INT32_MAX ? LDDMM_EINVAL : (int)n_items;  // size query
  if (!y || n != n_items) return LDDMM_EINVAL;
  const bool two_sets = (op == LDDMM_OP_CURRENTS || op == LDDMM_OP_MEASURE) && s.nx1 > 0;
  if (two_sets && !x1) return LDDMM_EINVAL;
  Shape sh;
  rc = op_shape(op, d, s, &sh);
  if (rc) return rc;

  const bool f32 = prec == LDDMM_PREC_FP32C;
  const int64_t wave_rows = 64 * (f32 ? rows_per_thread<float>() : rows_per_thread<double>());
  const int64_t tile_rows = (BLOCK / 64) * wave_rows;
  const double scale = coord_scale(prec, s.kernel, s.sigma);
  const double cut2 = cull_r2(prec, s.kernel, s.sigma);
  // The pair kernel's savings that only energy-only sums take: each unordered tile pair of
  // the self sum once, and (FP64 Gaussian) the FP32 exponential on far blocks, whose pairs
  // cost about half of a near block's.
  const bool energy_only = !s.rowdot && !s.out0 && sh.NA == 0 &&
                           (op == LDDMM_OP_KNORM || op == LDDMM_OP_CURRENTS ||
                            op == LDDMM_OP_MEASURE);
  const bool mixed = energy_only && !f32 && s.kernel == LDDMM_KERNEL_GAUSS;
  const double far2 = far_r2(prec, s.sigma);

  std::vector<double> q0pts((size_t)s.ny * 3), q1pts;
  for (int64_t i = 0; i < s.ny; ++i) {
    const int64_t g = perm ? (int64_t)perm[i] : i;
    if (g < 0 || g >= s.ny) return LDDMM_EINVAL;
    for (int c = 0; c < d; ++c) q0pts[i * 3 + c] = y[g * d + c] * scale;
  }
  if (two_sets) {
    q1pts.resize((size_t)s.nx1 * 3);
    for (int64_t i = 0; i < s.nx1; ++i)
      for (int c = 0; c < d; ++c) q1pts[i * 3 + c] = x1[i * d + c] * scale;
  }
  const Boxes waves = run_boxes(s.ny, wave_rows, d, [&](int64_t i) { return &q0pts[i * 3]; });
  const Boxes rtiles = run_boxes(s.ny, tile_rows, d, [&](int64_t i) { return &q0pts[i * 3]; });
  const Boxes quarters0 = run_boxes(s.ny, 64, d, [&](int64_t i) { return &q0pts[i * 3]; });
  const Boxes tiles0 = run_boxes(s.ny, TS, d, [&](int64_t i) { return &q0pts[i * 3]; });
  Boxes quarters1, tiles1;
  if (two_sets) {
    quarters1 = run_boxes(s.nx1, 64, d, [&](int64_t i) { return &q1pts[i * 3]; });
    tiles1 = run_boxes(s.nx1, TS, d, [&](int64_t i) { return &q1pts[i * 3]; });
  }
  const int64_t nwaves = (int64_t)waves.lo.size() / 3;
  const int wpt = BLOCK / 64;  // waves per row tile

  // Cost of one (row tile, source tile) pair: its live (wave, quarter) blocks, FP64 near
  // blocks counted twice.  The pair's bounding boxes decide most pairs at once (all blocks
  // culled, or all within the cut / the near radius); the rest are tested block by block.
  auto tile_cost = [&](int64_t rt, int set, int64_t ti) -> int64_t {
    const Boxes &tb = set ? tiles1 : tiles0;
    const Boxes &qb = set ? quarters1 : quarters0;
    const int64_t nsrc = set ? s.nx1 : s.nx0;
    const int64_t q0 = ti * (TS / 64), q1 = std::min((ti + 1) * (TS / 64), (nsrc + 63) / 64);
    const double g2 = box_gap2(rtiles, rt, tb, ti);
    if (g2 > cut2) return 0;
    const double m2 = box_far2(rtiles, rt, tb, ti);
    int uniform = 0;  // cost of every block when the pair decides them all
    if (!mixed && m2 <= cut2) uniform = 1;
    if (mixed && m2 <= far2) uniform = 2;
    if (mixed && g2 > far2 && m2 <= cut2) uniform = 1;
    int64_t c = 0;
    for (int w = 0; w < wpt; ++w) {
      const int64_t wi = rt * wpt + w;
      if (wi >= nwaves) break;
      const int64_t wave_blk = (rt * tile_rows + w * wave_rows) / TS;
      if (energy_only && set == 0 && ti < wave_blk) continue;  // symmetric: skipped
      int wu = uniform;  // the same decision for this wave's box alone
      if (!wu) {
        const double wg2 = box_gap2(waves, wi, tb, ti);
        if (wg2 > cut2) continue;
        const double wm2 = box_far2(waves, wi, tb, ti);
        if (!mixed && wm2 <= cut2) wu = 1;
        if (mixed && wm2 <= far2) wu = 2;
        if (mixed && wg2 > far2 && wm2 <= cut2) wu = 1;
      }
      if (wu) {
        c += wu * (q1 - q0);
        continue;
      }
      for (int64_t q = q0; q < q1; ++q) {
        const double b2 = box_gap2(waves, wi, qb, q);
        if (b2 > cut2) continue;
        c += (mixed && !(b2 > far2)) ? 2 : 1;
      }
    }
    return c;
  };

  std::vector<int64_t> cost((size_t)n_items);
  for (int64_t it = 0; it < n_items; ++it) {
    const int64_t rt = it % s.row_tiles, cg = it / s.row_tiles;
    const int set = cg < s.chunks0 ? 0 : 1;
    const int64_t ch = set ? cg - s.chunks0 : cg;
    const int64_t nsrc = set ? s.nx1 : s.nx0;
    const int64_t t0 = ch * s.chunk_len / TS;
    const int64_t t1 = std::min((ch + 1) * (int64_t)s.chunk_len, nsrc + TS - 1) / TS;
    int64_t c = 1;  // the item's fixed part: rows, boxes, partial stores
    for (int64_t ti = t0; ti < t1 && ti * TS < nsrc; ++ti) c += tile_cost(rt, set, ti);
    cost[it] = c;
  }
  std::vector<int32_t> idx((size_t)n_items);
  for (int64_t i = 0; i < n_items; ++i) idx[i] = (int32_t)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
  std::copy(idx.begin(), idx.end(), order);
  return LDDMM_OK;
}

extern "C" int lddmm_last_hip_error(void) { return g_last_hip_error; }

extern "C" const char *lddmm_version(void) { return "lddmm-mi355x 0.1 (gfx950)"; }
