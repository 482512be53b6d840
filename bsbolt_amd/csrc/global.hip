// Bit-exact batched ksw_global2 (bsbolt/External/BWA/ksw.c:504-606) for gfx950.
//
// A group of L lanes owns one task and holds the band of the current row: lane
// g keeps positions r = j - beg in [g*C, g*C+C) of H(i-1,j-1) and E(i,j) in
// VGPRs, so the group's width follows n_col = min(qlen, 2w+1), not qlen; rows
// are processed in order.  The band is fixed (beg = max(i-w,0),
// end = min(i+w+1,qlen), ksw.c:530-531).  F along a row is
//   F(j) = max_{beg<=k<j}(M_k - oe_ins + k*e_ins) - (j-1)*e_ins   (j > beg),
//   F(beg) = -inf,
// which equals the reference's sequential recurrence whenever every in-band M
// is finite -- guaranteed by the precondition |tlen - qlen| <= w together with
// the length limits checked per task.
//
// The direction byte of every in-band cell (bits 0-1: source of H, bit 2: E
// extended, bit 5: F extended; ksw.c:551-564) is written to the group's slot in
// a scratch buffer, row-major with n_col = min(qlen, 2w+1) bytes per row, and
// the group walks it back (ksw.c:588-602), writing the CIGAR into the task's
// caller-sized slot.  The walk reads the slot through an LDS window of whole
// rows: the group loads the next rows with all its lanes at once (one HBM
// latency per window, not one per step) and every step then reads LDS.
#include "bsx_internal.h"

namespace bsx {

constexpr int kNegInf = -0x40000000;  // MINUS_INF, ksw.c:490
constexpr int kGlbMaxT = 1 << 16;
// dwords of a group's LDS window on its direction bytes (16 KB per block)
template <int L>
constexpr int zwin_dw() { return L < 16 ? 128 : 256; }

// Group scans and reductions on DPP (L = 8, 16, 32 or 64 lanes; values may be
// negative, so lanes without a source keep the identity instead of 0).
template <int L>
__device__ __forceinline__ int g_incl_max(int x, int ident) {
  if (L == 8) {  // quad steps take the lane itself where there is no source (max is idempotent)
    x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x90, 0xf, 0xf, false));  // quad_perm [0,0,1,2]
    x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x40, 0xf, 0xf, false));  // quad_perm [0,0,0,1]
    const int q = __builtin_amdgcn_update_dpp(ident, x, 0xFF, 0xf, 0xf, false);  // quad_perm [3,3,3,3]
    // row_shr:4 into banks 1 and 3 only (lanes 4-7 of each 8-lane group)
    return max(x, __builtin_amdgcn_update_dpp(ident, q, 0x114, 0xf, 0xa, false));
  }
  x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x118, 0xf, 0xf, false));  // row_shr:8
  if (L >= 32) x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  if (L >= 64) x = max(x, __builtin_amdgcn_update_dpp(ident, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

// exclusive prefix max over the group (ident for its first lane)
template <int L>
__device__ __forceinline__ int g_excl_max(int x, int lig, int ident) {
  const int incl = g_incl_max<L>(x, ident);
  const int e = L <= 16 ? __builtin_amdgcn_update_dpp(ident, incl, 0x111, 0xf, 0xf, false)   // row_shr:1
                        : __builtin_amdgcn_update_dpp(ident, incl, 0x138, 0xf, 0xf, false);  // wave_shr:1
  return lig == 0 ? ident : e;
}

// every lane of the group receives the maximum
template <int L>
__device__ __forceinline__ int g_max(int x) {
  if (L <= 16) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, true));   // quad_perm [1,0,3,2]
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, true));   // quad_perm [2,3,0,1]
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xf, 0xf, true));  // row_half_mirror
    if (L == 8) return x;
    return max(x, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xf, 0xf, true));  // row_mirror
  }
  x = g_incl_max<L>(x, INT32_MIN);
  if (L == 64) return __builtin_amdgcn_readlane(x, 63);
  const int a = __builtin_amdgcn_readlane(x, 31), b = __builtin_amdgcn_readlane(x, 63);
  return (int)(threadIdx.x & 63) < 32 ? a : b;
}

// value of the previous lane of the group (lane 0: unused by the callers)
template <int L>
__device__ __forceinline__ int g_prev(int x) {
  return L <= 16 ? __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true)   // row_shr:1
                 : __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true);  // wave_shr:1
}

// value of the next lane of the group (lane L-1: unused by the callers)
template <int L>
__device__ __forceinline__ int g_next(int x) {
  return L <= 16 ? __builtin_amdgcn_update_dpp(0, x, 0x101, 0xf, 0xf, true)   // row_shl:1
                 : __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true);  // wave_shl:1
}

template <int L, int C>
__global__ void __launch_bounds__(256)
glb_kernel(GlbArgs A, GlbWork W, int bin, int64_t zbase) {
  const int lane = threadIdx.x & 63;
  const int lig = lane & (L - 1);
  const int base = lig * C;
  const int group = (blockIdx.x * blockDim.x + threadIdx.x) / L;
  uint32_t lbase = 0;
  for (int b = 0; b < bin; ++b) lbase += W.counts[b];
  const int32_t* __restrict__ list = W.list + lbase;
  const uint32_t nlist = W.counts[bin];
  uint8_t* __restrict__ z = W.z + (zbase + group) * W.zslot;  // zbase: this launch's slot region
  constexpr int ZW = zwin_dw<L>();
  __shared__ uint32_t zwin_all[256 / L][ZW];
  uint32_t* __restrict__ zwin = zwin_all[threadIdx.x / L];
  const int oe_del = A.o_del + A.e_del, oe_ins = A.o_ins + A.e_ins;
  const int e_del = A.e_del, e_ins = A.e_ins;

  // profile words by query code (code 4 for N and positions past the query),
  // for the position entering the band on each row
  __shared__ uint2 s_prof[5];
  if (threadIdx.x < 5) s_prof[threadIdx.x] = make_uint2(A.prof_lo[threadIdx.x], (uint32_t)A.prof_hi[threadIdx.x]);
  __syncthreads();
  // An empty bin returns before its first fetch (uniform over the grid): each
  // wave's fetch is a device atomic on one address, ~10 ns apiece serialised.
  if (nlist == 0) return;
  for (;;) {
    // Branch-free fetch: every lane issues the atomic (only lane 0 of the group
    // adds 1; the compiler folds it into one atomic per wave).  No lane-divergent
    // region may precede the broadcast: with `if (lig == 0)` around the atomic
    // the compiler can give the other lanes a loop of their own that reaches
    // the broadcast without lane 0's new value, and they never leave it.
    const uint32_t mine = atomicAdd(W.fetch + bin, lig == 0 ? 1u : 0u);
    const uint32_t idx = __shfl(mine, 0, L);
    if (idx >= nlist) break;
    const int task = list[idx];
    const int qlen = A.qlen[task], tlen = A.tlen[task];
    // a band wider than both sequences is the same band (beg = 0, end = qlen
    // on every row): clamping keeps i + w + 1 and the like in int range
    const int w = min(A.w[task], max(qlen, tlen));
    const uint64_t qoff = A.q_off[task], toff = A.t_off[task];
    const int ncol = qlen < 2 * w + 1 ? qlen : 2 * w + 1;
    const bool want_cigar = A.cigar != nullptr;

    // Lanes hold band positions r = j - beg (lane g: r in [g*C, g*C+C)), so a
    // group needs L*C >= n_col columns whatever qlen is.  Row i works on
    // positions [0, end-beg).  While beg stays 0 (i < w) the state moves one
    // position right each row (H(i,j-1) becomes the next row's H(i-1,j-1));
    // once beg advances with i, position r of row i+1 is column beg+1+r, so
    // H stays in place and E and the query profile move one position left, the
    // top position taking query[beg+1+L*C-1].
    constexpr int LC = L * C;
    int G[C], E[C];
    uint32_t plo[C];
    int phi[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = base + c;  // row 0: beg = 0
      const int qc = j < qlen ? (int)A.query[qoff + j] : 4;
      uint32_t lo = A.prof_lo[4];
      int hi = A.prof_hi[4];
      if (qc == 0) { lo = A.prof_lo[0]; hi = A.prof_hi[0]; }
      if (qc == 1) { lo = A.prof_lo[1]; hi = A.prof_hi[1]; }
      if (qc == 2) { lo = A.prof_lo[2]; hi = A.prof_hi[2]; }
      if (qc == 3) { lo = A.prof_lo[3]; hi = A.prof_hi[3]; }
      plo[c] = lo;
      phi[c] = hi;
      // row -1 (ksw.c:522-525)
      G[c] = j == 0 ? 0 : (j <= w ? -(A.o_ins + e_ins * j) : kNegInf);
      E[c] = kNegInf;
    }
    int hq = qlen <= w ? -(A.o_ins + e_ins * qlen) : kNegInf;  // eh[qlen].h

    // Per-row codes, two rows per lane and 2L rows per chunk: byte 0 the
    // target base of row rho, byte 1 the query base entering the top position
    // after row rho (column rho + L*C - w; 4 outside the query).  The raw bytes
    // of the next chunk are loaded one chunk ahead from clamped (in-bounds)
    // addresses, unconditionally, so nothing waits for them before the next
    // refill combines them.
    uint32_t raw[4];
    auto load_raw = [&](int cb) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int rho = cb + 2 * lig + k;
        const int jq = rho + LC - w;
        raw[2 * k] = A.target[toff + (rho < tlen ? rho : tlen - 1)];
        raw[2 * k + 1] = A.query[qoff + (jq < 0 ? 0 : (jq < qlen ? jq : qlen - 1))];
      }
    };
    auto combine = [&](int cb) -> uint32_t {
      uint32_t v = 0;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int rho = cb + 2 * lig + k;
        const int jq = rho + LC - w;
        const uint32_t tb = rho < tlen ? raw[2 * k] : 0u;
        const uint32_t qb = (jq >= 0 && jq < qlen) ? raw[2 * k + 1] : 4u;
        v |= (tb | qb << 8) << (16 * k);
      }
      return v;
    };
    load_raw(0);
    uint32_t ch = combine(0);
    load_raw(2 * L);
    int cb = 0;
    for (int i = 0; i < tlen; ++i) {
      int tr = i - cb;
      if (tr >= 2 * L) {
        cb = i;
        tr = 0;
        ch = combine(i);
        load_raw(i + 2 * L);
      }
      const uint32_t wd = (uint32_t)__shfl((int)ch, tr >> 1, L) >> (16 * (tr & 1));
      const int t = (int)(wd & 0xffu), qn = (int)((wd >> 8) & 0xffu);
      const int beg = i > w ? i - w : 0;
      const int end = i + w + 1 < qlen ? i + w + 1 : qlen;
      const int nb = end - beg;  // in-band positions [0, nb)
      const int h1p = beg == 0 ? -(A.o_del + e_del * (i + 1)) : kNegInf;

      int M[C], ex[C];
      int run = INT32_MIN;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int r = base + c;
        const bool inb = r < nb;
        const int s = t < 4 ? __builtin_amdgcn_sbfe((int)plo[c], 8 * (t & 3), 8) : phi[c];  // signed byte t
        const int mm = G[c] + s;
        M[c] = mm;
        const int key = inb ? mm - oe_ins + r * e_ins : INT32_MIN;
        ex[c] = run;
        run = max(run, key);
      }
      const int lex = g_excl_max<L>(run, lig, INT32_MIN);
      int h[C];
      int h1c = INT32_MIN;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int r = base + c;
        const bool inb = r < nb;
        const int x = max(ex[c], lex);
        // positions without an in-band predecessor get -inf (this also keeps
        // the subtraction away from INT32_MIN on out-of-band positions)
        const int f = (r == 0 || x == INT32_MIN) ? kNegInf : x - (r - 1) * e_ins;
        const int mm = M[c], e = E[c];
        int d = mm >= e ? 0 : 1;
        int hh = mm >= e ? mm : e;
        if (!(hh >= f)) { d = 2; hh = f; }
        h[c] = hh;
        const int od = mm - oe_del, ee = e - e_del;
        if (ee > od) d |= 1 << 2;
        const int ne = ee > od ? ee : od;
        const int oi = mm - oe_ins, ff = f - e_ins;
        if (ff > oi) d |= 2 << 4;
        // out-of-band positions hold E = -inf: the position end (eh[end].e,
        // ksw.c:575) and every position that enters the band later
        E[c] = inb ? ne : kNegInf;
        if (want_cigar && inb) z[(int64_t)i * ncol + r] = (uint8_t)d;
        if (r == nb - 1) h1c = hh;
      }
      if (i >= w) {  // beg(i+1) = beg(i) + 1
#pragma unroll
        for (int c = 0; c < C; ++c) G[c] = h[c];
        const int e0 = g_next<L>(E[0]);
        const uint32_t p0 = (uint32_t)g_next<L>((int)plo[0]);
        const int q0 = g_next<L>(phi[0]);
#pragma unroll
        for (int c = 0; c + 1 < C; ++c) {
          E[c] = E[c + 1];
          plo[c] = plo[c + 1];
          phi[c] = phi[c + 1];
        }
        const uint2 pq = s_prof[qn < 4 ? qn : 4];
        const bool top = lig == L - 1;
        E[C - 1] = top ? kNegInf : e0;
        plo[C - 1] = top ? pq.x : p0;
        phi[C - 1] = top ? (int)pq.y : q0;
      } else {  // beg stays 0: H(i,j-1) moves to position j, eh[0].h = h1p
        const int hprev = g_prev<L>(h[C - 1]);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int r = base + c;
          const int left = c == 0 ? hprev : h[c - 1];
          G[c] = r == 0 ? h1p : left;
        }
      }
      h1c = g_max<L>(h1c);
      if (end == qlen) hq = h1c;  // eh[end] = {h1, -inf} with end == qlen
    }
    if (lig == 0) {
      if (A.score) A.score[task] = hq;
    }
    if (want_cigar) {
      // The slot was written by every lane of this group; the walk below reads
      // it back.  All lanes of the group run the walk (the loop is uniform over
      // the group; lanes outside this task's group are not involved), so no
      // lane-divergent loop is left in front of the next fetch broadcast; only
      // lane 0 stores.  Runs of equal ops are merged in registers, which gives
      // the CIGAR push_cigar builds (ksw.c:492-502).
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      uint32_t* cg = A.cigar + (int64_t)task * A.cigar_slot;
      const int64_t cap = A.cigar_slot;
      int n = 0, state = 0, run_op = -1, run_len = 0;
      bool over = false;
      auto emit = [&]() {
        if (run_op < 0) return;
        if (n < cap) {
          if (lig == 0) cg[n] = (uint32_t)run_len << 4 | (uint32_t)run_op;
          ++n;
        } else {
          over = true;
        }
      };
      auto push = [&](int op, int len) {
        if (op == run_op) {
          run_len += len;
        } else {
          emit();
          run_op = op;
          run_len = len;
        }
      };
      int i = tlen - 1;
      int k = (i + w + 1 < qlen ? i + w + 1 : qlen) - 1;
      // LDS window: rows [wlo, whi] of the slot, from byte wbase (a dword
      // boundary) on; rows per window so that they fit with the alignment slack
      const int wrows = (4 * ZW - 3) / ncol;
      const uint32_t* __restrict__ zw = (const uint32_t*)z;
      int wlo = 1, whi = 0;
      uint32_t wbase = 0;
      // Offsets in the slot fit 32 bits: n_col <= 2048 and tlen <= 65536 are
      // checked per task.  Every lane computes the same (i, k, state), so the
      // step needs no cross-lane operation.
      while (i >= 0 && k >= 0) {
        const int kc = k - (i > w ? i - w : 0);
        if ((unsigned)kc >= (unsigned)ncol) {  // cannot happen under the preconditions; never read outside the slot
          over = true;
          break;
        }
        const uint32_t zi = (uint32_t)i * (uint32_t)ncol + (uint32_t)kc;
        int zb;
        if (wrows >= 1) {
          if (i < wlo || i > whi) {  // refill: rows [i - wrows + 1, i], every lane loading
            const int r0 = i - wrows + 1 > 0 ? i - wrows + 1 : 0;
            const uint32_t d0 = ((uint32_t)r0 * (uint32_t)ncol) >> 2;
            const uint32_t d1 = ((uint32_t)(i + 1) * (uint32_t)ncol - 1u) >> 2;
            __builtin_amdgcn_wave_barrier();
            for (int q = lig; q <= (int)(d1 - d0); q += L) zwin[q] = zw[d0 + q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            wlo = r0; whi = i; wbase = d0 << 2;
          }
          const uint32_t q = zi - wbase;  // every lane reads the same byte
          zb = (int)__builtin_amdgcn_ubfe(zwin[q >> 2], 8 * (q & 3), 8);
        } else {
          zb = z[zi];  // every lane loads the same byte
        }
        state = zb >> (state << 1) & 3;
        // 0: M, i-1 and k-1; 1: D, i-1; otherwise I, k-1 (ksw.c:596-598)
        const int op = state == 0 ? 0 : (state == 1 ? 2 : 1);
        i -= state <= 1 ? 1 : 0;
        k -= state != 1 ? 1 : 0;
        if (op != run_op) {  // the run of run_op ends (there is none before the first step)
          emit();
          run_op = op;
          run_len = 1;
        } else {
          ++run_len;
        }
      }
      if (i >= 0) push(2, i + 1);
      if (k >= 0) push(1, k + 1);
      emit();
      if (over) {
        if (lig == 0) {
          A.n_cigar[task] = -1;
          atomicAdd(&W.err[2], 1);
        }
      } else {
        // reverse in place (the walk produced the ops from the end); uniform loop
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        for (int a = 0, b = n - 1; a < b; ++a, --b) {
          const uint32_t ta = cg[a], tb = cg[b];
          if (lig == 0) { cg[a] = tb; cg[b] = ta; }
        }
        if (lig == 0) A.n_cigar[task] = n;
      }
    } else if (A.n_cigar && lig == 0) {
      A.n_cigar[task] = 0;
    }
  }
}

__device__ __forceinline__ int glb_bin_of(int ncol) {
#pragma unroll
  for (int b = 0; b < kGlbBins; ++b)
    if (ncol <= glb_bin_limit(b)) return b;
  return -1;
}

__global__ void __launch_bounds__(256)
glb_bin_count_kernel(GlbArgs A, GlbWork W) {
  __shared__ uint32_t hist[kGlbBins];
  if (threadIdx.x < kGlbBins) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ql = A.qlen[t], tl = A.tlen[t], w = A.w[t];
    const int diff = tl > ql ? tl - ql : ql - tl;
    const int64_t ncol = ql < 2 * (int64_t)w + 1 ? ql : 2 * (int64_t)w + 1;
    int b = glb_bin_of((int)ncol);  // by band width: a group holds n_col positions
    const bool ok = ql >= 1 && ql <= kGlbMaxT && tl >= 1 && tl <= kGlbMaxT && w >= 0 && b >= 0 && diff <= w &&
                    (A.cigar == nullptr || ncol * tl <= W.zslot);
    if (!ok) {
      b = 0xff;
      atomicAdd(&W.err[0], 1);
      atomicMin(&W.err[1], (int)t);
    } else {
      atomicAdd(&hist[b], 1u);
    }
    W.bin_of[t] = (uint8_t)b;
  }
  __syncthreads();
  if (threadIdx.x < kGlbBins && hist[threadIdx.x]) atomicAdd(&W.counts[threadIdx.x], hist[threadIdx.x]);
}

// Ranks within a bin are taken in LDS per 256 tasks and the block reserves
// each bin's run with one device atomic (a same-address device atomic per
// task serialises the whole batch on kGlbBins addresses).
__global__ void __launch_bounds__(256)
glb_bin_scatter_kernel(int64_t n, GlbWork W) {
  __shared__ uint32_t off[kGlbBins], cnt[kGlbBins], run[kGlbBins];
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int b = 0; b < kGlbBins; ++b) { off[b] = s; s += W.counts[b]; }
  }
  for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x; t0 < n; t0 += (int64_t)gridDim.x * blockDim.x) {
    if (threadIdx.x < kGlbBins) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t = t0 + threadIdx.x;
    const int b = t < n ? W.bin_of[t] : 0xff;
    uint32_t r = 0;
    if (b != 0xff) r = atomicAdd(&cnt[b], 1u);
    __syncthreads();
    if (threadIdx.x < kGlbBins)
      run[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(&W.cursor[threadIdx.x], cnt[threadIdx.x]) : 0u;
    __syncthreads();
    if (b != 0xff) W.list[off[b] + run[b] + r] = (int32_t)t;
  }
}

static_assert(kGlbBins <= kCountSlots && kGlbBins <= 64, "CIGAR bin counters");
__global__ void glb_reset_kernel(GlbWork W) {
  const int k = threadIdx.x;
  if (k < kGlbBins) { W.counts[k] = 0; W.cursor[k] = 0; W.fetch[k] = 0; }
  if (k == 0) { W.err[0] = 0; W.err[1] = 0x7fffffff; W.err[2] = 0; }
}


// The 8x4 and 16x4 bins (bands 25-64 wide) hold about 1.5 tasks per group in
// a CIGAR pass, so their kernels are mostly tail.  With side streams they run
// beside the 8x2/8x3 bins on a side stream, with fewer blocks and a slot
// region of their own after the main one.
constexpr int kGlbSideBlocksPerCU = 2;
constexpr bool glb_on_side(int b) { return b == 2 || b == 3; }

constexpr int64_t glb_groups_per_cu(bool side) {
  int64_t most = 0;
  for (int b = 0; b < kGlbBins; ++b) {
    if (glb_on_side(b) != side) continue;
    const int bpc = side ? kGlbSideBlocksPerCU : glb_blocks_per_cu(kGlbShape[b].lanes);
    const int64_t g = (int64_t)bpc * (256 / kGlbShape[b].lanes);
    most = g > most ? g : most;
  }
  return most;
}

// with side streams off every bin runs on the call's stream at its own block
// count, inside the main region
constexpr int64_t glb_groups_per_cu_one_stream() {
  int64_t most = 0;
  for (int b = 0; b < kGlbBins; ++b) {
    const int64_t g = (int64_t)glb_blocks_per_cu(kGlbShape[b].lanes) * (256 / kGlbShape[b].lanes);
    most = g > most ? g : most;
  }
  return most;
}
static_assert(glb_groups_per_cu(false) >= glb_groups_per_cu_one_stream(), "main slot region covers every bin");

int64_t glb_groups(int num_cu, bool side) {
  // one direction-byte slot per group: the main region covers the largest grid
  // of the call's stream (and every bin when side streams are off), the side
  // region the side-stream grids
  return (int64_t)num_cu * (glb_groups_per_cu(false) + (side ? glb_groups_per_cu(true) : 0));
}

template <int B>
static void glb_launch_bin(const GlbArgs& a, const GlbWork& w, int num_cu, hipStream_t main) {
  constexpr int L = kGlbShape[B].lanes, C = kGlbShape[B].cols;
  const bool side = glb_on_side(B) && w.ss != nullptr;
  const hipStream_t st = side ? pick_stream(w.ss, main, 1) : main;
  const int bpc = side ? kGlbSideBlocksPerCU : glb_blocks_per_cu(L);
  const int64_t zbase = side ? (int64_t)num_cu * glb_groups_per_cu(false) : 0;
  hipLaunchKernelGGL((glb_kernel<L, C>), dim3((unsigned)(num_cu * bpc)), dim3(256), 0, st, a, w, B, zbase);
}

hipError_t launch_global_impl(const GlbArgs& a, const GlbWork& w, int num_cu, hipStream_t st) {
  hipLaunchKernelGGL(glb_reset_kernel, dim3(1), dim3(64), 0, st, w);
  if (a.n <= 0) return hipGetLastError();
  int64_t blocks = (a.n + 255) / 256;
  if (blocks > (int64_t)num_cu * 16) blocks = (int64_t)num_cu * 16;
  hipLaunchKernelGGL(glb_bin_count_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, w);
  hipLaunchKernelGGL(glb_bin_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a.n, w);
  ss_fork(w.ss, st);  // side streams start after the scatter
  glb_launch_bin<3>(a, w, num_cu, st);
  glb_launch_bin<2>(a, w, num_cu, st);
  glb_launch_bin<8>(a, w, num_cu, st);
  glb_launch_bin<7>(a, w, num_cu, st);
  glb_launch_bin<6>(a, w, num_cu, st);
  glb_launch_bin<5>(a, w, num_cu, st);
  glb_launch_bin<4>(a, w, num_cu, st);
  glb_launch_bin<1>(a, w, num_cu, st);
  glb_launch_bin<0>(a, w, num_cu, st);
  ss_join(w.ss, st);  // the call's stream waits for the side streams
  return hipGetLastError();
}

}  // namespace bsx
