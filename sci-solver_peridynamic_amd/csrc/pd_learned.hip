// pd_learned.hip -- learned block preconditioner.
//
// Per particle p, a 16-64-64-8 MLP maps features of the assembled system
// (live bond count, mean / max bond stretch, the diagonal block D_p relative
// to m I, distance to the nearest broken bond) to L_p (lower triangular,
// positive diagonal exp(o)), and
//     M_p = C_p^{-T} (L_p L_p^T + eps I) C_p^{-1},   D_p = C_p C_p^T,
// an SPD block that is block Jacobi (D_p^{-1}) when L_p = I.  The CG kernels
// apply M_p exactly as they apply the block-Jacobi inverse, so the solve and
// its stopping rule are unchanged; the network only changes the iteration
// count.
//
// The layers run on MFMA: v_mfma_f32_32x32x16_bf16, one wave per 32 rows,
// BF16 weights and activations, FP32 accumulation.  Training minimises
// ||M_p (A e)_p - e_p||^2 over random sign probes e on rows of systems the
// simulator assembled, with a hand-written backward pass (also on MFMA) and
// Adam on FP32 master weights.
//
// MFMA operand layouts (32x32x16): lane l holds A[l % 32][8 (l / 32) + i]
// and B[8 (l / 32) + i][l % 32], i = 0..7; the 16 accumulators hold
// C[8 (j / 4) + 4 (l / 32) + j % 4][l % 32], j = 0..15.
#include <math.h>

#include "pd_internal.h"

namespace pd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kPad = 8;  // LDS row padding (bf16 elements)
constexpr int kH = PC_NH + kPad;

__device__ __forceinline__ floatx16 mfma(const bf16x8& a, const bf16x8& b, const floatx16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int crow(int j, int kh) { return 8 * (j >> 2) + 4 * kh + (j & 3); }

__device__ __forceinline__ floatx16 zero16() {
  floatx16 z;
#pragma unroll
  for (int j = 0; j < 16; ++j) z[j] = 0.0f;
  return z;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// L from the network outputs: exp of the clamped diagonal outputs
template <typename T>
__device__ __forceinline__ void make_l(const float* o, T L[6], bool clamp_flags[3]) {
  const float d0 = fminf(fmaxf(o[0], -4.0f), 4.0f);
  const float d1 = fminf(fmaxf(o[2], -4.0f), 4.0f);
  const float d2 = fminf(fmaxf(o[5], -4.0f), 4.0f);
  clamp_flags[0] = d0 != o[0];
  clamp_flags[1] = d1 != o[2];
  clamp_flags[2] = d2 != o[5];
  L[0] = exp((T)d0);  // L00
  L[1] = o[1];        // L10
  L[2] = exp((T)d1);  // L11
  L[3] = o[3];        // L20
  L[4] = o[4];        // L21
  L[5] = exp((T)d2);  // L22
}

// lower Cholesky factor of the symmetric block (d00 d01 d02 d11 d12 d22) and
// its inverse; false if the block is not positive definite
template <typename T>
__device__ __forceinline__ bool chol_inv(const T d[6], T ci[6]) {
  if (!(d[0] > T(0))) return false;
  const T c00 = sqrt(d[0]);
  const T c10 = d[1] / c00, c20 = d[2] / c00;
  const T t1 = d[3] - c10 * c10;
  if (!(t1 > T(0))) return false;
  const T c11 = sqrt(t1);
  const T c21 = (d[4] - c20 * c10) / c11;
  const T t2 = d[5] - c20 * c20 - c21 * c21;
  if (!(t2 > T(0))) return false;
  const T c22 = sqrt(t2);
  const T i00 = T(1) / c00, i11 = T(1) / c11, i22 = T(1) / c22;
  const T i10 = -(c10 * i00) / c11;
  const T i21 = -(c21 * i11) / c22;
  const T i20 = -(c20 * i00 + c21 * i10) / c22;
  ci[0] = i00;  // Ci00
  ci[1] = i10;  // Ci10
  ci[2] = i11;  // Ci11
  ci[3] = i20;  // Ci20
  ci[4] = i21;  // Ci21
  ci[5] = i22;  // Ci22
  return true;
}

// full 3x3 of a lower triangle stored (00 10 11 20 21 22)
template <typename T>
__device__ __forceinline__ void lower_full(const T l[6], T m[3][3]) {
  m[0][0] = l[0]; m[0][1] = 0;    m[0][2] = 0;
  m[1][0] = l[1]; m[1][1] = l[2]; m[1][2] = 0;
  m[2][0] = l[3]; m[2][1] = l[4]; m[2][2] = l[5];
}

// M = Ci^T S Ci with S = L L^T + eps I; out (M00 M01 M02 M11 M12 M22)
template <typename T>
__device__ __forceinline__ void learned_block(const T Lc[6], T eps, const T ci[6], T out[6]) {
  T L[3][3], C[3][3], S[3][3], SC[3][3], M[3][3];
  lower_full(Lc, L);
  lower_full(ci, C);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = i == j ? eps : T(0);
      for (int k = 0; k < 3; ++k) s += L[i][k] * L[j][k];
      S[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += S[i][k] * C[k][j];
      SC[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T s = 0;
      for (int k = 0; k < 3; ++k) s += C[k][i] * SC[k][j];
      M[i][j] = s;
    }
  out[0] = M[0][0];
  out[1] = 0.5 * (M[0][1] + M[1][0]);
  out[2] = 0.5 * (M[0][2] + M[2][0]);
  out[3] = M[1][1];
  out[4] = 0.5 * (M[1][2] + M[2][1]);
  out[5] = M[2][2];
}

// The block-Jacobi inverse of (d00 d01 d02 d11 d12 d22) by cofactors, the
// assembly's arithmetic (six distinct entries)
__device__ __forceinline__ void bj_inverse(const double d[6], double m[6]) {
  const double d00 = d[0], d01 = d[1], d02 = d[2], d11 = d[3], d12 = d[4], d22 = d[5];
  const double c00 = d11 * d22 - d12 * d12;
  const double c01 = d02 * d12 - d01 * d22;
  const double c02 = d01 * d12 - d02 * d11;
  const double c11 = d00 * d22 - d02 * d02;
  const double c12 = d01 * d02 - d00 * d12;
  const double c22 = d00 * d11 - d01 * d01;
  const double det = (d00 * c00 + d01 * c01) + d02 * c02;
  const double id = 1.0 / det;
  m[0] = c00 * id;
  m[1] = c01 * id;
  m[2] = c02 * id;
  m[3] = c11 * id;
  m[4] = c12 * id;
  m[5] = c22 * id;
}

__global__ __launch_bounds__(PD_BLOCK) void k_pc_features(PcFeatArgs a) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  float f[PC_NIN];
#pragma unroll
  for (int i = 0; i < PC_NIN; ++i) f[i] = 0.0f;
  if (a.act[p] != PD_INVISIBLE) {
    const pd_mat_dev mt = a.mat[a.body[p]];
    const int top = a.top[p], cap = a.cap[p];
    const int s = p >> 6, lane = p & 63;
    const double* dg = a.diag + (int64_t)s * 6 * 64 + lane;
    const double m = mt.mass;
    const float2 bs = a.bond_stats[p];
    f[0] = (float)(top + 1) * 0.01f;
    f[1] = bs.x * 100.0f;
    f[2] = bs.y * 100.0f;
    f[3] = (float)((dg[0] - m) / m);
    f[4] = (float)(dg[64] / m);
    f[5] = (float)(dg[128] / m);
    f[6] = (float)((dg[192] - m) / m);
    f[7] = (float)(dg[256] / m);
    f[8] = (float)((dg[320] - m) / m);
    const double4 xp = a.x[p];
    double dmin = 2.0 * mt.horizon;
    const int64_t base = a.slice_off[s];
    for (int k = top + 1; k < cap; ++k) {
      const int q = a.col[(base + k) * 64 + lane];
      const double4 xq = a.x[q];
      const double dx = xq.x - xp.x, dy = xq.y - xp.y, dz = xq.z - xp.z;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      dmin = d < dmin ? d : dmin;
    }
    f[9] = (float)(dmin / mt.horizon);
  }
  // four 16-byte stores (a row of features is 64 bytes)
  float4* o = (float4*)(a.feat + (int64_t)p * PC_NIN);
#pragma unroll
  for (int i = 0; i < PC_NIN / 4; ++i)
    o[i] = make_float4(f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
}

// One wave evaluates the MLP for 32 rows (its part of a 128-row tile); the
// outputs land in outs[w][row][0..7].  hs is the wave's activation tile.
__device__ __forceinline__ void mlp_forward(const bf16x8& x, const __bf16* w, const float* P,
                                            __bf16 (*h1)[kH], __bf16 (*h2)[kH],
                                            float (*outs)[PC_NOUT]) {
  const int l = threadIdx.x & 63;
  const int r = l & 31, kh = l >> 5;
  // layer 1: K = 16
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    // 16-byte loads: every operand row starts on a multiple of 8 elements
    const bf16x8 b = *(const bf16x8*)(w + PC_OFF_W1 + (nt * 32 + r) * PC_NIN + 8 * kh);
    const floatx16 c = mfma(x, b, zero16());
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int col = nt * 32 + r;
      const float v = c[j] + P[PC_OFF_B1 + col];
      h1[crow(j, kh)][col] = (__bf16)(v > 0.0f ? v : 0.0f);
    }
  }
  __syncthreads();
  // layer 2: K = 64
  floatx16 c0 = zero16(), c1 = zero16();
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const bf16x8 av = *(const bf16x8*)&h1[r][ks * 16 + 8 * kh];
    const bf16x8 b0 = *(const bf16x8*)(w + PC_OFF_W2 + r * PC_NH + ks * 16 + 8 * kh);
    const bf16x8 b1 = *(const bf16x8*)(w + PC_OFF_W2 + (32 + r) * PC_NH + ks * 16 + 8 * kh);
    c0 = mfma(av, b0, c0);
    c1 = mfma(av, b1, c1);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const float v0 = c0[j] + P[PC_OFF_B2 + r];
    const float v1 = c1[j] + P[PC_OFF_B2 + 32 + r];
    h2[crow(j, kh)][r] = (__bf16)(v0 > 0.0f ? v0 : 0.0f);
    h2[crow(j, kh)][32 + r] = (__bf16)(v1 > 0.0f ? v1 : 0.0f);
  }
  __syncthreads();
  // layer 3: N = 8 (of a 32-wide tile), K = 64
  floatx16 c3 = zero16();
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const bf16x8 av = *(const bf16x8*)&h2[r][ks * 16 + 8 * kh];
    bf16x8 bv;
#pragma unroll
    for (int i = 0; i < 8; ++i) bv[i] = (__bf16)0.0f;
    if (r < PC_NOUT) bv = *(const bf16x8*)(w + PC_OFF_W3 + r * PC_NH + ks * 16 + 8 * kh);
    c3 = mfma(av, bv, c3);
  }
  if (r < PC_NOUT) {
#pragma unroll
    for (int j = 0; j < 16; ++j) outs[crow(j, kh)][r] = c3[j] + P[PC_OFF_B3 + r];
  }
  __syncthreads();
}

__device__ __forceinline__ bf16x8 load_x(const float* f, const float* P) {
  const int kh = (threadIdx.x & 63) >> 5;
  bf16x8 x;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = 8 * kh + i;
    x[i] = (__bf16)((f[k] - P[PC_OFF_MEAN + k]) * P[PC_OFF_ISTD + k]);
  }
  return x;
}

// M_p of row p from the network's outputs o (the lane's row), written to
// minv (and, rounded, to minv32 when non-null); m returns the FP64 values.
__device__ __forceinline__ void learned_row(int p, const float* o, const float* P,
                                            const double* diag, double* minv, float* minv32,
                                            double m[6]) {
  const int s = p >> 6, lane = p & 63;
  const double* dg = diag + (int64_t)s * 6 * 64 + lane;
  const double d[6] = {dg[0], dg[64], dg[128], dg[192], dg[256], dg[320]};
  // the block algebra in float, as in training (the network is a BF16
  // approximation and any SPD block serves; the solve takes M as given)
  const float df[6] = {(float)d[0], (float)d[1], (float)d[2], (float)d[3], (float)d[4],
                       (float)d[5]};
  float ci[6], L[6], mf6[6];
  bool cl[3];
  if (chol_inv(df, ci)) {
    make_l(o, L, cl);
    learned_block(L, P[PC_OFF_EPS], ci, mf6);
#pragma unroll
    for (int c = 0; c < 6; ++c) m[c] = (double)mf6[c];
  } else {  // a block without a Cholesky factor keeps block Jacobi
    bj_inverse(d, m);
  }
  double* mo = minv + (int64_t)s * 6 * 64 + lane;
#pragma unroll
  for (int c = 0; c < 6; ++c) mo[c * 64] = m[c];
  if (minv32) {
    float* mf = minv32 + (int64_t)s * 6 * 64 + lane;
#pragma unroll
    for (int c = 0; c < 6; ++c) mf[c * 64] = (float)m[c];
  }
}

__global__ __launch_bounds__(PD_BLOCK) void k_pc_infer(int n, const float* feat, PcNet net,
                                                      const double* diag, const int32_t* act,
                                                      double* minv, float* minv32) {
  __shared__ __bf16 h1[4][32][kH];
  __shared__ __bf16 h2[4][32][kH];
  __shared__ float outs[4][32][PC_NOUT];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l & 31;
  const int row0 = blockIdx.x * 128 + w * 32;
  const float* P = net.p;
  // the lane's eight features straight from memory (a row past the end
  // reads the last row; its outputs are dropped below)
  const int pr = row0 + r < n ? row0 + r : n - 1;
  const bf16x8 x = load_x(feat + (int64_t)pr * PC_NIN, P);
  mlp_forward(x, net.w, P, h1[w], h2[w], outs[w]);
  if (l >= 32) return;
  const int p = row0 + l;
  if (p >= n || act[p] == PD_INVISIBLE) return;
  double m[6];
  learned_row(p, outs[w][l], P, diag, minv, minv32, m);
}

// k_pc_infer and the learned solve's start in one pass, after a fused
// assembly: a workgroup takes the 256 rows of one solver tile (each wave
// two MLP passes of 32 rows), every lane forms the M_p of its row and then
// z = M_p r (mixed: with M_p rounded to float, on the float residual),
// p = z (FP64: and the gather record's p part) and its term of the <r, z>
// tile partial, folded by the solver's tile tree into the slot the
// assembly's block-Jacobi start used (FP64: part[tile]; mixed:
// part[ntiles + tile]).
template <bool MIXED>
__global__ __launch_bounds__(PD_BLOCK) void k_pc_infer_start(PcInferStartArgs a) {
  __shared__ __bf16 h1[4][32][kH];
  __shared__ __bf16 h2[4][32][kH];
  __shared__ float outs[4][64][PC_NOUT];
  __shared__ double lds4[4];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l & 31;
  const int tile = blockIdx.x;
  const int row0 = tile * PD_BLOCK + w * 64;
  const int n = a.n;
  const float* P = a.net.p;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int pr = row0 + 32 * half + r < n ? row0 + 32 * half + r : n - 1;
    const bf16x8 x = load_x(a.feat + (int64_t)pr * PC_NIN, P);
    mlp_forward(x, a.net.w, P, h1[w], h2[w], outs[w] + 32 * half);
  }
  const int p = row0 + l;
  double t = 0.0;
  if (p < n && a.act[p] != PD_INVISIBLE) {
    double m[6];
    learned_row(p, outs[w][l], P, a.diag, a.minv, MIXED ? a.minv32 : nullptr, m);
    if constexpr (MIXED) {  // z = M32 r32 with the float blocks just written
      const float m0 = (float)m[0], m1 = (float)m[1], m2 = (float)m[2], m3 = (float)m[3],
                  m4 = (float)m[4], m5 = (float)m[5];
      const float4 h = a.r32[p];
      const float z0 = m0 * h.x + m1 * h.y + m2 * h.z;
      const float z1 = m1 * h.x + m3 * h.y + m4 * h.z;
      const float z2 = m2 * h.x + m4 * h.y + m5 * h.z;
      a.p32[p] = make_float4(z0, z1, z2, 0.0f);
      t = ((double)h.x * (double)z0 + (double)h.y * (double)z1) + (double)h.z * (double)z2;
    } else {
      const double4 rr = a.r64[p];
      const double z0 = m[0] * rr.x + m[1] * rr.y + m[2] * rr.z;
      const double z1 = m[1] * rr.x + m[3] * rr.y + m[4] * rr.z;
      const double z2 = m[2] * rr.x + m[4] * rr.y + m[5] * rr.z;
      if (a.p64) a.p64[p] = make_double4(z0, z1, z2, 0.0);  // null: matrix-free
      if (a.rec) {
        double* rp = a.rec + 6 * (int64_t)p;
        rp[3] = z0;
        rp[4] = z1;
        rp[5] = z2;
      }
      t = (rr.x * z0 + rr.y * z1) + rr.z * z2;
    }
  }
  const double v = tile_sum(t, lds4);
  if (threadIdx.x == 0) a.part[(MIXED ? (int64_t)gridDim.x : 0) + tile] = v;
}


// random-sign probe vectors
__global__ __launch_bounds__(PD_BLOCK) void k_pc_probe(int n, uint64_t seed, double4* e) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (p >= n) return;
  const uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)p);
  e[p] = make_double4((h & 1) ? 1.0 : -1.0, (h & 2) ? 1.0 : -1.0, (h & 4) ? 1.0 : -1.0, 0.0);
}

// samples of the rows p with p % stride == offset; rows that cannot be
// used (INVISIBLE, Dirichlet) are stored with D = 0 and skipped in training
__global__ __launch_bounds__(PD_BLOCK) void k_pc_samples(PcSampleArgs a) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  const uint32_t stride = a.keep;
  const uint32_t off = (uint32_t)(mix64(a.seed) % stride);
  if ((uint32_t)p % stride != off) return;
  const int64_t slot = (int64_t)(p / stride);
  if (slot >= a.capacity) return;
  float* o = a.out + slot * PC_SAMPLE;
  const bool use = a.act[p] != PD_INVISIBLE && a.bc[p] == 0;
  for (int i = 0; i < PC_NIN; ++i) o[i] = a.feat[(int64_t)p * PC_NIN + i];
  const int s = p >> 6, lane = p & 63;
  const double* dg = a.diag + (int64_t)s * 6 * 64 + lane;
  for (int c = 0; c < 6; ++c) o[PC_NIN + c] = use ? (float)dg[c * 64] : 0.0f;
  const double4 v = a.Ae[p], e = a.e[p];
  o[PC_NIN + 6] = (float)v.x;
  o[PC_NIN + 7] = (float)v.y;
  o[PC_NIN + 8] = (float)v.z;
  o[PC_NIN + 9] = (float)e.x;
  o[PC_NIN + 10] = (float)e.y;
  o[PC_NIN + 11] = (float)e.z;
}

// Loss of one sample and its gradient with respect to the network outputs:
// M = Ci^T S Ci, r = M v - e, loss = |r|^2; dS = sym(Ci G Ci^T) with
// G = 2 r v^T, dL = 2 dS L (lower part), d(exp o) = dL_dd exp(o).
__device__ void sample_grad(const float* o, const float* smp, float eps, float* dout,
                            float& loss) {
  for (int k = 0; k < PC_NOUT; ++k) dout[k] = 0.0f;
  loss = 0.0f;
  const float d[6] = {smp[PC_NIN + 0], smp[PC_NIN + 1], smp[PC_NIN + 2],
                      smp[PC_NIN + 3], smp[PC_NIN + 4], smp[PC_NIN + 5]};
  float ci[6];
  if (!chol_inv(d, ci)) return;
  double Ld[6];
  bool cl[3];
  make_l<double>(o, Ld, cl);
  const float Lc[6] = {(float)Ld[0], (float)Ld[1], (float)Ld[2], (float)Ld[3], (float)Ld[4],
                       (float)Ld[5]};
  float m6[6];
  learned_block(Lc, eps, ci, m6);
  const float M[3][3] = {{m6[0], m6[1], m6[2]}, {m6[1], m6[3], m6[4]}, {m6[2], m6[4], m6[5]}};
  const float* v = smp + PC_NIN + 6;
  const float* e = smp + PC_NIN + 9;
  float rr[3];
  for (int i = 0; i < 3; ++i) rr[i] = (M[i][0] * v[0] + M[i][1] * v[1] + M[i][2] * v[2]) - e[i];
  loss = rr[0] * rr[0] + rr[1] * rr[1] + rr[2] * rr[2];
  float G[3][3], C[3][3], T[3][3], GS[3][3], L[3][3], GL[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) G[i][j] = 2.0f * rr[i] * v[j];
  lower_full(ci, C);
  lower_full(Lc, L);
  for (int i = 0; i < 3; ++i)  // T = Ci G
    for (int j = 0; j < 3; ++j) {
      float s = 0;
      for (int k = 0; k < 3; ++k) s += C[i][k] * G[k][j];
      T[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)  // GS = T Ci^T
    for (int j = 0; j < 3; ++j) {
      float s = 0;
      for (int k = 0; k < 3; ++k) s += T[i][k] * C[j][k];
      GS[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < i; ++j) {
      const float sy = 0.5f * (GS[i][j] + GS[j][i]);
      GS[i][j] = GS[j][i] = sy;
    }
  for (int i = 0; i < 3; ++i)  // GL = 2 GS L
    for (int j = 0; j < 3; ++j) {
      float s = 0;
      for (int k = 0; k < 3; ++k) s += GS[i][k] * L[k][j];
      GL[i][j] = 2.0f * s;
    }
  dout[0] = cl[0] ? 0.0f : GL[0][0] * Lc[0];
  dout[1] = GL[1][0];
  dout[2] = cl[1] ? 0.0f : GL[1][1] * Lc[2];
  dout[3] = GL[2][0];
  dout[4] = GL[2][1];
  dout[5] = cl[2] ? 0.0f : GL[2][2] * Lc[5];
}

// C tile (32x32) of A^T-style products from LDS: A[i][k] = src_a(k, i), etc.
// Forward, loss, backward of one 32-sample chunk per wave; per-wave
// gradients (and the loss sum) go to partial[global wave].
__global__ __launch_bounds__(PD_BLOCK) void k_pc_train(PcTrainArgs a) {
  __shared__ __bf16 xs[4][32][PC_NIN + kPad];
  __shared__ __bf16 h1[4][32][kH];
  __shared__ __bf16 h2[4][32][kH];
  __shared__ __bf16 g[4][32][kH];  // dH2, then dH1
  __shared__ float outs[4][32][PC_NOUT];
  __shared__ float smp[4][32][PC_SAMPLE];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l & 31, kh = l >> 5;
  const int gw = blockIdx.x * 4 + w;
  const float* P = a.p;
  // draw the samples (with replacement)
  {
    const uint64_t h = mix64(a.seed ^ mix64((uint64_t)a.step * 0x100000001B3ull + gw * 32 + r));
    const int64_t idx = (int64_t)(h % (uint64_t)a.nsamples);
    const float* src = a.samples + idx * PC_SAMPLE;
    for (int k = kh; k < PC_SAMPLE; k += 2) smp[w][r][k] = src[k];
  }
  __syncthreads();
  const bf16x8 x = load_x(smp[w][r], P);
#pragma unroll
  for (int i = 0; i < 8; ++i) xs[w][r][8 * kh + i] = x[i];
  mlp_forward(x, a.w, P, h1[w], h2[w], outs[w]);
  // loss and output gradients (lanes 0..31, one sample each)
  const float scale = 1.0f / (3.0f * (float)a.batch);
  float loss = 0.0f;
  if (l < 32) {
    float o[PC_NOUT], dout[PC_NOUT];
    for (int k = 0; k < PC_NOUT; ++k) o[k] = outs[w][l][k];
    sample_grad(o, smp[w][l], P[PC_OFF_EPS], dout, loss);
    for (int k = 0; k < PC_NOUT; ++k) outs[w][l][k] = dout[k] * scale;
  }
  __syncthreads();
  float* part = a.partial + (int64_t)gw * (PC_NWEIGHT + 1);
  // dW3 = dO^T H2 (rows: outputs < 8), db3
  {
    floatx16 c0 = zero16(), c1 = zero16();
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 av, b0, b1;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = ks * 16 + 8 * kh + i;
        av[i] = r < PC_NOUT ? (__bf16)outs[w][k][r] : (__bf16)0.0f;
        b0[i] = h2[w][k][r];
        b1[i] = h2[w][k][32 + r];
      }
      c0 = mfma(av, b0, c0);
      c1 = mfma(av, b1, c1);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // rows 4 kh + j < 8
      const int row = crow(j, kh);
      part[PC_OFF_W3 + row * PC_NH + r] = c0[j];
      part[PC_OFF_W3 + row * PC_NH + 32 + r] = c1[j];
    }
    if (l < PC_NOUT) {
      float s = 0.0f;
      for (int k = 0; k < 32; ++k) s += outs[w][k][l];
      part[PC_OFF_B3 + l] = s;
    }
  }
  // dH2 = dO W3, masked by H2 > 0
  {
    floatx16 c0 = zero16(), c1 = zero16();
    bf16x8 av, b0, b1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = 8 * kh + i;  // output index, < 8 only for kh = 0
      av[i] = k < PC_NOUT ? (__bf16)outs[w][r][k] : (__bf16)0.0f;
      b0[i] = k < PC_NOUT ? a.w[PC_OFF_W3 + k * PC_NH + r] : (__bf16)0.0f;
      b1[i] = k < PC_NOUT ? a.w[PC_OFF_W3 + k * PC_NH + 32 + r] : (__bf16)0.0f;
    }
    c0 = mfma(av, b0, c0);
    c1 = mfma(av, b1, c1);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int row = crow(j, kh);
      g[w][row][r] = (float)h2[w][row][r] > 0.0f ? (__bf16)c0[j] : (__bf16)0.0f;
      g[w][row][32 + r] = (float)h2[w][row][32 + r] > 0.0f ? (__bf16)c1[j] : (__bf16)0.0f;
    }
  }
  __syncthreads();
  // dW2 = dH2^T H1, db2, and dH1 = dH2 W2 (masked by H1 > 0)
  floatx16 d1a = zero16(), d1b = zero16();
  {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        floatx16 c = zero16();
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
          bf16x8 av, bv;
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int k = ks * 16 + 8 * kh + i;
            av[i] = g[w][k][mt * 32 + r];
            bv[i] = h1[w][k][nt * 32 + r];
          }
          c = mfma(av, bv, c);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j)
          part[PC_OFF_W2 + (mt * 32 + crow(j, kh)) * PC_NH + nt * 32 + r] = c[j];
      }
    {
      float s = 0.0f;
      for (int k = 0; k < 32; ++k) s += (float)g[w][k][l];
      part[PC_OFF_B2 + l] = s;
    }
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      bf16x8 av, b0, b1;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = ks * 16 + 8 * kh + i;  // H2 index
        av[i] = g[w][r][k];
        b0[i] = a.w[PC_OFF_W2 + k * PC_NH + r];
        b1[i] = a.w[PC_OFF_W2 + k * PC_NH + 32 + r];
      }
      d1a = mfma(av, b0, d1a);
      d1b = mfma(av, b1, d1b);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int row = crow(j, kh);
    g[w][row][r] = (float)h1[w][row][r] > 0.0f ? (__bf16)d1a[j] : (__bf16)0.0f;
    g[w][row][32 + r] = (float)h1[w][row][32 + r] > 0.0f ? (__bf16)d1b[j] : (__bf16)0.0f;
  }
  __syncthreads();
  // dW1 = dH1^T X, db1
  {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      floatx16 c = zero16();
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        bf16x8 av, bv;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int k = ks * 16 + 8 * kh + i;
          av[i] = g[w][k][mt * 32 + r];
          bv[i] = r < PC_NIN ? xs[w][k][r] : (__bf16)0.0f;
        }
        c = mfma(av, bv, c);
      }
      if (r < PC_NIN) {
#pragma unroll
        for (int j = 0; j < 16; ++j) part[PC_OFF_W1 + (mt * 32 + crow(j, kh)) * PC_NIN + r] = c[j];
      }
    }
    float s = 0.0f;
    for (int k = 0; k < 32; ++k) s += (float)g[w][k][l];
    part[PC_OFF_B1 + l] = s;
  }
  // loss of the wave's samples
  float ls = l < 32 ? loss : 0.0f;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) ls += __shfl_xor(ls, m, 64);
  if (l == 0) part[PC_NWEIGHT] = ls;
}

// fold the per-wave gradients in wave order, Adam, bf16 refresh; thread i
// owns weight i
__global__ __launch_bounds__(PD_BLOCK) void k_pc_adam(const float* partial, int nwaves, float* p,
                                                     float* m, float* v, __bf16* w, int step,
                                                     float lr, float* loss_out) {
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i == 0 && loss_out) {
    float s = 0.0f;
    for (int k = 0; k < nwaves; ++k) s += partial[(int64_t)k * (PC_NWEIGHT + 1) + PC_NWEIGHT];
    loss_out[0] = s;
  }
  if (i >= PC_NWEIGHT) return;
  float gsum = 0.0f;
  for (int k = 0; k < nwaves; ++k) gsum += partial[(int64_t)k * (PC_NWEIGHT + 1) + i];
  const float b1 = 0.9f, b2 = 0.999f;
  const float mi = b1 * m[i] + (1.0f - b1) * gsum;
  const float vi = b2 * v[i] + (1.0f - b2) * gsum * gsum;
  m[i] = mi;
  v[i] = vi;
  const float t = (float)(step + 1);
  const float mh = mi / (1.0f - powf(b1, t));
  const float vh = vi / (1.0f - powf(b2, t));
  const float pi = p[i] - lr * mh / (sqrtf(vh) + 1e-8f);
  p[i] = pi;
  w[i] = (__bf16)pi;
}

__global__ __launch_bounds__(PD_BLOCK) void k_pc_to_bf16(const float* p, __bf16* w) {
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i < PC_NWEIGHT) w[i] = (__bf16)p[i];
}

// one workgroup: feature means and inverse standard deviations
__global__ __launch_bounds__(PD_BLOCK) void k_pc_norm(const float* samples, int64_t ns, float* p) {
  __shared__ double sm[PD_BLOCK], sq[PD_BLOCK];
  for (int k = 0; k < PC_NIN; ++k) {
    double s = 0.0, q = 0.0;
    for (int64_t i = threadIdx.x; i < ns; i += PD_BLOCK) {
      const double f = samples[i * PC_SAMPLE + k];
      s += f;
      q += f * f;
    }
    sm[threadIdx.x] = s;
    sq[threadIdx.x] = q;
    __syncthreads();
    if (threadIdx.x == 0) {
      double S = 0.0, Q = 0.0;
      for (int t = 0; t < PD_BLOCK; ++t) {
        S += sm[t];
        Q += sq[t];
      }
      const double mean = ns > 0 ? S / (double)ns : 0.0;
      const double var = ns > 0 ? Q / (double)ns - mean * mean : 0.0;
      const double sd = var > 0.0 ? sqrt(var) : 0.0;
      p[PC_OFF_MEAN + k] = (float)mean;
      p[PC_OFF_ISTD + k] = sd > 1e-6 ? (float)(1.0 / sd) : 0.0f;
    }
    __syncthreads();
  }
}

inline int blocks_of(int64_t n) { return (int)((n + PD_BLOCK - 1) / PD_BLOCK); }

}  // namespace

void launch_pc_features(const PcFeatArgs& a, hipStream_t s) {
  if (a.n) hipLaunchKernelGGL(k_pc_features, dim3(blocks_of(a.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_pc_infer(int32_t n, const float* feat, const PcNet& net, const double* diag,
                     const int32_t* act, double* minv, float* minv32, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_pc_infer, dim3((n + 127) / 128), dim3(PD_BLOCK), 0, s, n, feat, net,
                            diag, act, minv, minv32);
}
void launch_pc_infer_start(const PcInferStartArgs& a, hipStream_t s) {
  if (!a.n) return;
  if (a.mixed)
    hipLaunchKernelGGL(k_pc_infer_start<true>, dim3(blocks_of(a.n)), dim3(PD_BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL(k_pc_infer_start<false>, dim3(blocks_of(a.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_pc_probe(int32_t n, uint64_t seed, double4* e, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_pc_probe, dim3(blocks_of(n)), dim3(PD_BLOCK), 0, s, n, seed, e);
}
void launch_pc_samples(const PcSampleArgs& a, hipStream_t s) {
  if (a.n) hipLaunchKernelGGL(k_pc_samples, dim3(blocks_of(a.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_pc_train(const PcTrainArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_pc_train, dim3(a.batch / 128), dim3(PD_BLOCK), 0, s, a);
}
void launch_pc_adam(const float* partial, int32_t nwaves, float* p, float* m, float* v, __bf16* w,
                    int32_t step, float lr, float* loss_out, hipStream_t s) {
  hipLaunchKernelGGL(k_pc_adam, dim3(blocks_of(PC_NWEIGHT)), dim3(PD_BLOCK), 0, s, partial, nwaves,
                     p, m, v, w, step, lr, loss_out);
}
void launch_pc_to_bf16(const float* p, __bf16* w, hipStream_t s) {
  hipLaunchKernelGGL(k_pc_to_bf16, dim3(blocks_of(PC_NWEIGHT)), dim3(PD_BLOCK), 0, s, p, w);
}
void launch_pc_norm(const float* samples, int64_t nsamples, float* p, hipStream_t s) {
  hipLaunchKernelGGL(k_pc_norm, dim3(1), dim3(PD_BLOCK), 0, s, samples, nsamples, p);
}

}  // namespace pd
