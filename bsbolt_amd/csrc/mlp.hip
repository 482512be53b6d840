// Band predictor of the learned-band form: per-task features, an MLP
// 24 -> 64 -> 64 -> 5 (ReLU, softmax over the frame classes 24/40/80/144/plain)
// with BF16 weights and FP32 accumulation on MFMA, its backward pass and Adam.
//
// Matrix layout.  Activations are kept transposed (features x tasks) so the
// output of one v_mfma_f32_16x16x16_bf16 is, lane for lane, the B operand of the
// next layer: for a 16-task tile, lane l holds rows 4*(l/16)+v (v = 0..3) of
// task column l%16 -- the D layout of the instruction and the B layout of its
// K x N operand coincide.  Weights are the A operands (lane l: row l%16, columns
// 4*(l/16)+v of a 16x16 block), read from packed BF16 copies of the FP32 master
// parameters.  Weight gradients sum over tasks, so their operands need the
// tile transposed (lane l: feature l%16, tasks 4*(l/16)+v); that transpose goes
// through a 512-byte LDS tile per wave.
//
// Parameter vector (FP32, 6085 values): W1[64][24], b1[64], W2[64][64], b2[64],
// W3[5][64], b3[5] (row = output unit).
#include "bsx_internal.h"
#include "ext_common.h"
#include "mlp.h"

namespace bsx {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf4 pack4(float a, float b, float c, float d) {
  bf4 r;
  r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
  return r;
}

__device__ __forceinline__ f4 mfma(bf4 a, bf4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf4 ld_bf4(const uint16_t* p) {  // 8-byte aligned
  const uint2 u = *(const uint2*)p;
  bf4 r;
  r[0] = (short)(u.x & 0xffff); r[1] = (short)(u.x >> 16); r[2] = (short)(u.y & 0xffff); r[3] = (short)(u.y >> 16);
  return r;
}

// ---------------- features ----------------

// Task features (kBandFeatures = 24), each an integer scaled by a power of two,
// so the FP32 value and its BF16 rounding are reproducible anywhere.  The
// ungapped scans run over the first K = min(qlen, kFeatScan) query bases on the
// five diagonals d = -2..2 (query p against target p + d).
__device__ void band_features(const ExtArgs& A, int64_t t, float* f) {
  const int qlen = A.qlen[t], tlen = A.tlen[t], h0 = A.h0[t], eb = A.end_bonus[t];
  const int oe_ins = A.o_ins + A.e_ins, oe_del = A.o_del + A.e_del;
  static_assert(kFeatScan == 16, "the scan reads 16 query and 19 target bases");
  // The bases the scans read, with all loads issued up front: aligned dwords
  // (indices clamped to the sequence's last dword, so only mapped dwords that
  // hold sequence bases are read), realigned to byte 0.  A task without target
  // bases reads its query instead and masks every target position.
  uint32_t qb[4], tb[5];
  {
    const uintptr_t qa = (uintptr_t)(A.query + A.q_off[t]);
    const uint32_t* qp = (const uint32_t*)(qa & ~(uintptr_t)3);
    const int qsh = (int)(qa & 3), qlast = (qsh + (qlen > 0 ? qlen : 1) - 1) >> 2;
    uint32_t w[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = qp[min(k, qlast)];
#pragma unroll
    for (int k = 0; k < 4; ++k) qb[k] = __builtin_amdgcn_alignbyte(w[k + 1], w[k], (uint32_t)qsh);
    const uintptr_t ta = (uintptr_t)(tlen > 0 ? A.target + A.t_off[t] : A.query + A.q_off[t]);
    const uint32_t* tp = (const uint32_t*)(ta & ~(uintptr_t)3);
    const int tsh = (int)(ta & 3), tlast = (tsh + (tlen > 0 ? tlen : 1) - 1) >> 2;
    uint32_t v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = tp[min(k, tlast)];
#pragma unroll
    for (int k = 0; k < 5; ++k) tb[k] = __builtin_amdgcn_alignbyte(v[k + 1], v[k], (uint32_t)tsh);
  }
  auto qbase = [&](int p) { return (int)((qb[p >> 2] >> (8 * (p & 3))) & 0xffu); };
  auto tbase = [&](int r) { return (r >= 0 && r < tlen) ? (int)((tb[r >> 2] >> (8 * (r & 3))) & 0xffu) : 5; };
  const int K = qlen < kFeatScan ? qlen : kFeatScan;
  int S[5] = {0, 0, 0, 0, 0}, MX[5] = {0, 0, 0, 0, 0};
  int arg0 = 0, match0 = 0, mism0 = 0, first_mm = K, match16 = 0;
  int qT = 0, qN = 0, gc = 0, cpg = 0;
#pragma unroll
  for (int p = 0; p < kFeatScan; ++p) {
    if (p < K) {
      const uint32_t qc = (uint32_t)qbase(p);
      qT += qc == 3;
      qN += qc > 3;
      uint32_t lo = A.prof_lo[4];  // mat[k*5 + qc], k = 0..3
      int hi = A.prof_hi[4];       // mat[20 + qc]
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        lo = qc == (uint32_t)k ? A.prof_lo[k] : lo;
        hi = qc == (uint32_t)k ? A.prof_hi[k] : hi;
      }
      int tw[5];  // target bases p-2 .. p+2 (5 = outside the target)
#pragma unroll
      for (int d = 0; d < 5; ++d) tw[d] = tbase(p + d - 2);
#pragma unroll
      for (int d = 0; d < 5; ++d) {
        if (tw[d] == 5) continue;
        S[d] += tw[d] < 4 ? sbyte(lo, tw[d]) : hi;
        MX[d] = S[d] > MX[d] ? S[d] : MX[d];
      }
      if (tw[2] != 5) {
        const bool eq = (int)qc == tw[2] && qc < 4;
        match0 += eq;
        mism0 += !eq;
        match16 += (eq && p < 16);
        if (!eq && first_mm == K) first_mm = p;
        if (S[2] == MX[2] && S[2] > 0) arg0 = p + 1;
        gc += tw[2] == 1 || tw[2] == 2;
        cpg += tw[2] == 1 && tw[3] == 2;
      }
    }
  }
  int wcl = A.w[t];
  {
    int lim = (int)((double)(qlen * A.mat_max + eb - A.o_ins) / A.e_ins + 1.);
    lim = lim > 1 ? lim : 1;
    wcl = wcl < lim ? wcl : lim;
    lim = (int)((double)(qlen * A.mat_max + eb - A.o_del) / A.e_del + 1.);
    lim = lim > 1 ? lim : 1;
    wcl = wcl < lim ? wcl : lim;
  }
  const int rin = h0 > oe_ins ? (h0 - oe_ins) / A.e_ins : 0;  // columns reached by the first-row ramp
  const int rde = h0 > oe_del ? (h0 - oe_del) / A.e_del : 0;  // rows reached by the first-column ramp
  auto clampi = [](int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; };
  f[0] = (float)clampi(qlen, 0, 4096) * (1.f / 128);
  f[1] = (float)clampi(tlen, 0, 4096) * (1.f / 256);
  f[2] = (float)clampi(h0, 0, 4096) * (1.f / 64);
  f[3] = (float)clampi(wcl, 0, 4096) * (1.f / 128);
  f[4] = (float)clampi(eb, -256, 256) * (1.f / 32);
  f[5] = (float)clampi(rin, 0, 4096) * (1.f / 64);
  f[6] = (float)clampi(rde, 0, 4096) * (1.f / 64);
  f[7] = (float)clampi(rin - qlen, -4096, 4096) * (1.f / 64);
  f[8] = (float)match0 * (1.f / 64);
  f[9] = (float)MX[2] * (1.f / 64);
  f[10] = (float)arg0 * (1.f / 128);
  f[11] = (float)clampi(S[2], -256, 256) * (1.f / 64);
  f[12] = (float)first_mm * (1.f / 128);
  f[13] = (float)match16 * (1.f / 16);
  f[14] = (float)MX[3] * (1.f / 64);
  f[15] = (float)MX[4] * (1.f / 64);
  f[16] = (float)MX[1] * (1.f / 64);
  f[17] = (float)MX[0] * (1.f / 64);
  f[18] = (float)clampi(MX[2] - S[2], 0, 1024) * (1.f / 64);
  f[19] = (float)mism0 * (1.f / 64);
  f[20] = (float)qT * (1.f / 128);
  f[21] = (float)gc * (1.f / 128);
  f[22] = (float)cpg * (1.f / 32);
  f[23] = (float)qN * (1.f / 8);
}

__global__ void __launch_bounds__(256) band_features_kernel(ExtArgs A, uint16_t* __restrict__ out) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n; t += (int64_t)gridDim.x * blockDim.x) {
    float f[kBandFeatures];
    band_features(A, t, f);
    uint32_t w[kFeatPad / 2];
#pragma unroll
    for (int k = 0; k < kFeatPad / 2; ++k) {
      const uint32_t lo = 2 * k < kBandFeatures ? f2bf(f[2 * k]) : 0u;
      const uint32_t hi = 2 * k + 1 < kBandFeatures ? f2bf(f[2 * k + 1]) : 0u;
      w[k] = lo | (hi << 16);
    }
    uint4* o = (uint4*)(out + t * kFeatPad);
#pragma unroll
    for (int k = 0; k < kFeatPad / 8; ++k) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
  }
}

// ---------------- forward ----------------

// Forward pass of one 16-task tile.  xb[kt]: B operand of input block kt.
// Returns the logits in D layout (lane l: classes 4*(l/16)+v of task l%16) and
// the hidden activations (FP32, D layout) when asked for.
struct FwdOut {
  f4 h1[4], h2[4];
  f4 logit;
};

// Weights and biases from any address space: w1p/w2p/w3p the BF16 operand
// copies, b1/b2/b3 the biases (b3 read only for its kBandClasses entries
// unless B3_PADDED, which promises 16 entries, zero past kBandClasses).
template <bool B3_PADDED>
__device__ __forceinline__ void mlp_tile_forward_w(const uint16_t* w1p, const uint16_t* w2p, const uint16_t* w3p,
                                                   const float* b1, const float* b2, const float* b3, const bf4 xb[2],
                                                   FwdOut& o, int lane) {
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  bf4 h1b[4], h2b[4];
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) acc = mfma(ld_bf4(w1p + (16 * ob + r16) * kFeatPad + 16 * kt + q4), xb[kt], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float z = acc[v] + b1[16 * ob + q4 + v];
      acc[v] = z > 0.f ? z : 0.f;
    }
    o.h1[ob] = acc;
    h1b[ob] = pack4(acc[0], acc[1], acc[2], acc[3]);
  }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = mfma(ld_bf4(w2p + (16 * ob + r16) * 64 + 16 * kt + q4), h1b[kt], acc);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float z = acc[v] + b2[16 * ob + q4 + v];
      acc[v] = z > 0.f ? z : 0.f;
    }
    o.h2[ob] = acc;
    h2b[ob] = pack4(acc[0], acc[1], acc[2], acc[3]);
  }
  f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) acc = mfma(ld_bf4(w3p + r16 * 64 + 16 * kt + q4), h2b[kt], acc);
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] += B3_PADDED ? b3[q4 + v] : (q4 + v < kBandClasses ? b3[q4 + v] : 0.f);
  o.logit = acc;
}

__device__ __forceinline__ void mlp_tile_forward(const MlpDev& M, const bf4 xb[2], FwdOut& o, int lane) {
  mlp_tile_forward_w<false>(M.w1p, M.w2p, M.w3p, M.params + kOffB1, M.params + kOffB2, M.params + kOffB3, xb, o,
                            lane);
}

// argmax over the 5 logits of task l%16, valid in lanes 0..15 (ties -> lower class)
__device__ __forceinline__ int tile_argmax(const f4& lg, int lane) {
  const float z4 = __shfl(lg[0], (lane & 15) + 16, 64);
  int c = 0;
  float m = lg[0];
#pragma unroll
  for (int v = 1; v < 4; ++v)
    if (lg[v] > m) { m = lg[v]; c = v; }
  if (z4 > m) c = 4;
  return c;
}

// Forward from stored features: logits [n][5] (optional) and class [n] (optional).
__global__ void __launch_bounds__(256)
mlp_forward_kernel(MlpDev M, const uint16_t* __restrict__ feats, int64_t n, float* __restrict__ logits,
                   uint8_t* __restrict__ cls) {
  const int lane = threadIdx.x & 63;
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  const int64_t ntiles = (n + 15) / 16;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t s = tile * 16 + r16;
    bf4 xb[2];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
      if (s < n) xb[kt] = ld_bf4(feats + s * kFeatPad + 16 * kt + q4);
      else xb[kt] = bf4{0, 0, 0, 0};
    }
    FwdOut o;
    mlp_tile_forward(M, xb, o, lane);
    if (s < n && logits) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (q4 + v < kBandClasses) logits[s * kBandClasses + q4 + v] = o.logit[v];
    }
    const int c = tile_argmax(o.logit, lane);
    if (s < n && lane < 16 && cls) cls[s] = (uint8_t)c;
  }
}

// Features and forward in one pass for the learned-band dispatch: each lane
// computes the features of one task into the wave's LDS tile, then the wave
// runs the four 16-task MLP tiles.  The block copies the BF16 weights and the
// biases into LDS once, so the tiles read their A operands from LDS.
__global__ void __launch_bounds__(256)
band_predict_kernel(ExtArgs A, MlpDev M, uint8_t* __restrict__ cls) {
  __shared__ uint16_t xs[4][64 * kFeatPad];
  __shared__ __align__(16) uint16_t s_w1[kHidden * kFeatPad], s_w2[kHidden * kHidden], s_w3[16 * kHidden];
  __shared__ float s_b[2 * kHidden + 16];  // b1, b2, b3 padded with zeros to 16
  {
    const uint32_t* g1 = (const uint32_t*)M.w1p;
    const uint32_t* g2 = (const uint32_t*)M.w2p;
    const uint32_t* g3 = (const uint32_t*)M.w3p;
    for (int k = threadIdx.x; k < kHidden * kFeatPad / 2; k += blockDim.x) ((uint32_t*)s_w1)[k] = g1[k];
    for (int k = threadIdx.x; k < kHidden * kHidden / 2; k += blockDim.x) ((uint32_t*)s_w2)[k] = g2[k];
    for (int k = threadIdx.x; k < 16 * kHidden / 2; k += blockDim.x) ((uint32_t*)s_w3)[k] = g3[k];
    for (int k = threadIdx.x; k < 2 * kHidden + 16; k += blockDim.x) {
      float v = 0.f;
      if (k < kHidden) v = M.params[kOffB1 + k];
      else if (k < 2 * kHidden) v = M.params[kOffB2 + k - kHidden];
      else if (k - 2 * kHidden < kBandClasses) v = M.params[kOffB3 + k - 2 * kHidden];
      s_b[k] = v;
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  uint16_t* X = xs[wv];
  const int64_t nchunks = (A.n + 63) / 64;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t ch = wave; ch < nchunks; ch += nwaves) {
    const int64_t t = ch * 64 + lane;
    float f[kBandFeatures];
    if (t < A.n) {
      band_features(A, t, f);
    } else {
#pragma unroll
      for (int k = 0; k < kBandFeatures; ++k) f[k] = 0.f;
    }
    uint32_t* xr = (uint32_t*)(X + lane * kFeatPad);
#pragma unroll
    for (int k = 0; k < kFeatPad / 2; ++k) {
      const uint32_t lo = 2 * k < kBandFeatures ? f2bf(f[2 * k]) : 0u;
      const uint32_t hi = 2 * k + 1 < kBandFeatures ? f2bf(f[2 * k + 1]) : 0u;
      xr[k] = lo | (hi << 16);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 1
    for (int tq = 0; tq < 4; ++tq) {
      bf4 xb[2];
#pragma unroll
      for (int kt = 0; kt < 2; ++kt) xb[kt] = ld_bf4(X + (16 * tq + r16) * kFeatPad + 16 * kt + q4);
      FwdOut o;
      mlp_tile_forward_w<true>(s_w1, s_w2, s_w3, s_b, s_b + kHidden, s_b + 2 * kHidden, xb, o, lane);
      const int c = tile_argmax(o.logit, lane);
      const int64_t s = ch * 64 + 16 * tq + r16;
      if (lane < 16 && s < A.n) cls[s] = (uint8_t)c;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ---------------- backward ----------------

// D-layout tile (lane l: rows 4*(l/16)+v of column l%16) -> transposed operand
// (lane l: row l%16, columns 4*(l/16)+v), through a 16x16 BF16 LDS tile.
__device__ __forceinline__ bf4 transpose_tile(uint16_t* T, const f4& d, int lane) {
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int v = 0; v < 4; ++v) T[(q4 + v) * 16 + r16] = f2bf(d[v]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const bf4 r = ld_bf4(T + r16 * 16 + q4);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return r;
}

// Softmax cross-entropy gradients summed over the samples [0, n) (mean taken by
// the caller); each wave writes its partial sums to part[wave][kParamCount + 1]
// (parameter layout of the FP32 vector, then the summed loss).
__global__ void __launch_bounds__(256)
mlp_grad_kernel(MlpDev M, const uint16_t* __restrict__ feats, const uint8_t* __restrict__ labels, int64_t n,
                float* __restrict__ part) {
  __shared__ __align__(16) uint16_t tbuf[4][256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r16 = lane & 15, q4 = (lane >> 4) * 4;
  uint16_t* T = tbuf[wv];
  const int64_t ntiles = (n + 15) / 16;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  f4 gW1[4][2], gW2[4][4], gW3[4], gb1[4], gb2[4], gb3;
  const f4 z4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gW1[a][0] = z4; gW1[a][1] = z4; gW3[a] = z4; gb1[a] = z4; gb2[a] = z4;
#pragma unroll
    for (int b = 0; b < 4; ++b) gW2[a][b] = z4;
  }
  gb3 = z4;
  float loss = 0.f;
  for (int64_t tile = wave; tile < ntiles; tile += nwaves) {
    const int64_t s = tile * 16 + r16;
    const bool valid = s < n;
    bf4 xb[2];
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) xb[kt] = valid ? ld_bf4(feats + s * kFeatPad + 16 * kt + q4) : bf4{0, 0, 0, 0};
    FwdOut o;
    mlp_tile_forward(M, xb, o, lane);
    // softmax over classes 0..4 of sample l%16: lanes 0..15 hold classes 0..3, lanes 16..31 class 4
    const float z4v = __shfl(o.logit[0], r16 + 16, 64);
    float mx = z4v;
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = fmaxf(mx, o.logit[v]);
    float e[4], sum = __expf(z4v - mx);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      e[v] = __expf(o.logit[v] - mx);
      sum += e[v];
    }
    const float inv = 1.f / sum;
    const int y = valid ? (int)labels[s] : 0;
    const float p4 = __expf(z4v - mx) * inv;
    const float p4_here = __shfl(p4, r16, 64);  // lanes 16..31 take p4 from lane l-16
    f4 dlog = z4;
    if (valid) {
      if (lane < 16) {
#pragma unroll
        for (int v = 0; v < 4; ++v) dlog[v] = e[v] * inv - (y == v ? 1.f : 0.f);
        float py = p4;
#pragma unroll
        for (int v = 0; v < 4; ++v) py = y == v ? e[v] * inv : py;
        loss += -__logf(fmaxf(py, 1e-30f));
      } else if (lane < 32) {
        dlog[0] = p4_here - (y == 4 ? 1.f : 0.f);
      }
    }
    // dH2 = W3^T dlog, dZ2 = dH2 * [h2 > 0]
    const bf4 dlogb = pack4(dlog[0], dlog[1], dlog[2], dlog[3]);
    f4 dz2[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      f4 acc = mfma(ld_bf4(M.w3t + (16 * hb + r16) * 16 + q4), dlogb, z4);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = o.h2[hb][v] > 0.f ? acc[v] : 0.f;
      dz2[hb] = acc;
    }
    bf4 dz2b[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) dz2b[hb] = pack4(dz2[hb][0], dz2[hb][1], dz2[hb][2], dz2[hb][3]);
    // dH1 = W2^T dZ2, dZ1 = dH1 * [h1 > 0]
    f4 dz1[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      f4 acc = z4;
#pragma unroll
      for (int b = 0; b < 4; ++b) acc = mfma(ld_bf4(M.w2t + (16 * a + r16) * 64 + 16 * b + q4), dz2b[b], acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = o.h1[a][v] > 0.f ? acc[v] : 0.f;
      dz1[a] = acc;
    }
    // weight gradients: sum over the tile's samples (K = samples)
    const bf4 dlT = transpose_tile(T, dlog, lane);
    bf4 h2T[4], h1T[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) h2T[hb] = transpose_tile(T, o.h2[hb], lane);
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) gW3[hb] = mfma(dlT, h2T[hb], gW3[hb]);
#pragma unroll
    for (int a = 0; a < 4; ++a) h1T[a] = transpose_tile(T, o.h1[a], lane);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const bf4 dT = transpose_tile(T, dz2[ob], lane);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) gW2[ob][ib] = mfma(dT, h1T[ib], gW2[ob][ib]);
    }
    bf4 xT[2];  // lane l: feature 16kt + l%16 of samples 4*(l/16)+v
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t sv = tile * 16 + q4 + v;
        xT[kt][v] = sv < n ? (short)feats[sv * kFeatPad + 16 * kt + r16] : (short)0;
      }
    }
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const bf4 dT = transpose_tile(T, dz1[ob], lane);
#pragma unroll
      for (int kt = 0; kt < 2; ++kt) gW1[ob][kt] = mfma(dT, xT[kt], gW1[ob][kt]);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        gb1[a][v] += dz1[a][v];
        gb2[a][v] += dz2[a][v];
      }
      gb3[v] += dlog[v];
    }
  }
  // ---- write this wave's partial sums ----
  float* P = part + wave * (kParamCount + 1);
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * ob + q4 + v, col = 16 * kt + r16;
        if (col < kBandFeatures) P[kOffW1 + row * kBandFeatures + col] = gW1[ob][kt][v];
      }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
#pragma unroll
      for (int v = 0; v < 4; ++v) P[kOffW2 + (16 * ob + q4 + v) * 64 + 16 * ib + r16] = gW2[ob][ib][v];
#pragma unroll
  for (int hb = 0; hb < 4; ++hb)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (q4 + v < kBandClasses) P[kOffW3 + (q4 + v) * 64 + 16 * hb + r16] = gW3[hb][v];
  // bias sums: reduce over the 16 sample columns of each row group
  auto rsum = [&](float x) {
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    x += __shfl_xor(x, 4, 64);
    x += __shfl_xor(x, 8, 64);
    return x;
  };
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float s1 = rsum(gb1[a][v]), s2 = rsum(gb2[a][v]);
      if (r16 == 0) {
        P[kOffB1 + 16 * a + q4 + v] = s1;
        P[kOffB2 + 16 * a + q4 + v] = s2;
      }
    }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float s3 = rsum(gb3[v]);
    if (r16 == 0 && q4 + v < kBandClasses) P[kOffB3 + q4 + v] = s3;
  }
  const float ls = rsum(loss);
  if (lane == 0) P[kParamCount] = ls;
}

// grad[p] = sum over waves of part[w][p] * scale  (p = kParamCount: the loss)
__global__ void __launch_bounds__(256)
mlp_reduce_kernel(const float* __restrict__ part, int nwaves, float scale, float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParamCount) return;
  float s = 0.f;
  for (int w = 0; w < nwaves; ++w) s += part[(int64_t)w * (kParamCount + 1) + p];
  grad[p] = s * scale;
}

// Adam on the FP32 master parameters from the mean gradient grad[0..kParamCount).
__global__ void __launch_bounds__(256)
mlp_adam_kernel(MlpDev M, const float* __restrict__ grad, float lr, float b1, float b2, float eps, float c1,
                float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParamCount) return;
  const float g = grad[p];
  const float m = b1 * M.adam_m[p] + (1.f - b1) * g;
  const float v = b2 * M.adam_v[p] + (1.f - b2) * g * g;
  M.adam_m[p] = m;
  M.adam_v[p] = v;
  M.params[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// BF16 operand copies of the master weights: w1p[64][32], w2p[64][64],
// w3p[16][64] (row = output unit) and the transposes w2t[64][64], w3t[64][16].
__global__ void __launch_bounds__(256) mlp_pack_kernel(MlpDev M) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < 64 * kFeatPad) {
    const int r = k / kFeatPad, c = k % kFeatPad;
    M.w1p[k] = c < kBandFeatures ? f2bf(M.params[kOffW1 + r * kBandFeatures + c]) : 0;
  }
  if (k < 64 * 64) {
    const int r = k / 64, c = k % 64;
    const uint16_t b = f2bf(M.params[kOffW2 + r * 64 + c]);
    M.w2p[k] = b;
    M.w2t[c * 64 + r] = b;
  }
  if (k < 16 * 64) {
    const int r = k / 64, c = k % 64;
    const uint16_t b = r < kBandClasses ? f2bf(M.params[kOffW3 + r * 64 + c]) : 0;
    M.w3p[k] = b;
    M.w3t[c * 16 + r] = b;
  }
}

// ---------------- launchers ----------------

static int64_t grid_for(int64_t items, int per_block, int num_cu, int per_cu) {
  int64_t b = (items + per_block - 1) / per_block;
  const int64_t cap = (int64_t)num_cu * per_cu;
  if (b > cap) b = cap;
  return b < 1 ? 1 : b;
}

hipError_t mlp_launch_features(const ExtArgs& a, uint16_t* out, int num_cu, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(band_features_kernel, dim3((unsigned)grid_for(a.n, 256, num_cu, 8)), dim3(256), 0, st, a, out);
  return hipGetLastError();
}

hipError_t mlp_launch_predict_tasks(const ExtArgs& a, const MlpDev& m, uint8_t* cls, int num_cu, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(band_predict_kernel, dim3((unsigned)grid_for(a.n, 256, num_cu, 8)), dim3(256), 0, st, a, m,
                     cls);
  return hipGetLastError();
}

hipError_t mlp_launch_forward(const MlpDev& m, const uint16_t* feats, int64_t n, float* logits, uint8_t* cls,
                              int num_cu, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)grid_for((n + 15) / 16, 4, num_cu, 8)), dim3(256), 0, st,
                     m, feats, n, logits, cls);
  return hipGetLastError();
}

int mlp_grad_waves(int64_t n) {
  int64_t tiles = (n + 15) / 16;
  if (tiles > kMaxGradWaves) tiles = kMaxGradWaves;
  if (tiles < 4) tiles = 4;
  return (int)((tiles + 3) / 4 * 4);
}

hipError_t mlp_launch_grad(const MlpDev& m, const uint16_t* feats, const uint8_t* labels, int64_t n, float* part,
                           float* grad, hipStream_t st) {
  const int nw = mlp_grad_waves(n);
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)(nw / 4)), dim3(256), 0, st, m, feats, labels, n, part);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3((kParamCount + 1 + 255) / 256), dim3(256), 0, st, (const float*)part, nw,
                     n > 0 ? 1.f / (float)n : 0.f, grad);
  return hipGetLastError();
}

hipError_t mlp_launch_adam(const MlpDev& m, const float* grad, float lr, int step, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParamCount + 255) / 256), dim3(256), 0, st, m, grad, lr, b1, b2, eps, c1,
                     c2);
  return mlp_launch_pack(m, st);
}

hipError_t mlp_launch_pack(const MlpDev& m, hipStream_t st) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((64 * 64 + 255) / 256), dim3(256), 0, st, m);
  return hipGetLastError();
}

}  // namespace bsx
