// Bit-exact batched ksw_extend2 (bsbolt/External/BWA/ksw.c:380-479) for gfx950.
//
// Layout: a task is owned by a group of L lanes of one wave64 (64/L tasks per
// wave); lane g holds query columns [g*C, g*C+C) of the DP state in VGPRs:
//   G[c] = eh[j].h = H(i-1, j-1)   (diagonal predecessor)
//   E[c] = eh[j].e = E(i, j)
// Rows (target bases) are processed one at a time, exactly like the reference
// row loop, so the adaptive band [beg,end), the z-drop and the early exits are
// decided on the same values.  Columns outside the active band are never
// written, which reproduces the reference's reads of stale eh[] entries when
// the band regrows (ksw.c:466-469).
//
// Within a row the only sequential dependency is F.  Because F is built from
// M (not H), F(j) = max_{beg<=k<j}(max(M_k - oe_ins, 0) + k*e_ins) - (j-1)*e_ins
// for j > beg and 0 at j = beg: an exclusive prefix-max over the row, done as a
// lane-local scan plus a group scan.  Row maximum (ties -> larger j), the new
// beg/end and the value at column end-1 are group reductions.
//
// The kernel is persistent: groups pull task indices from a per-bin counter and
// run a flattened loop (fetch-or-row per iteration) so that groups of a wave do
// not wait for one another between tasks.
#include "bsx_internal.h"

namespace bsx {

template <int L>
__device__ __forceinline__ int group_max(int x) {
#pragma unroll
  for (int d = L / 2; d >= 1; d >>= 1) x = max(x, __shfl_xor(x, d, L));
  return x;
}

template <int L>
__device__ __forceinline__ int group_min(int x) {
#pragma unroll
  for (int d = L / 2; d >= 1; d >>= 1) x = min(x, __shfl_xor(x, d, L));
  return x;
}

// Exclusive prefix max over the lanes of a group; lane 0 receives `ident`.
template <int L>
__device__ __forceinline__ int group_excl_max(int x, int lig, int ident) {
#pragma unroll
  for (int d = 1; d < L; d <<= 1) {
    int y = __shfl_up(x, d, L);
    if (lig >= d) x = max(x, y);
  }
  int e = __shfl_up(x, 1, L);
  return lig == 0 ? ident : e;
}

__device__ __forceinline__ int prof_score(uint32_t lo, int hi, int t) {
  // signed byte t of lo for t < 4, hi for t == 4
  return t < 4 ? (int)((int32_t)(lo << (24 - 8 * t)) >> 24) : hi;
}

template <int L, int C>
__global__ void __launch_bounds__(256)
ext_plain_kernel(ExtArgs A, const int32_t* __restrict__ list_all, const uint32_t* __restrict__ counts,
                 int bin, uint32_t* __restrict__ fetch) {
  const int lane = threadIdx.x & 63;
  const int lig = lane & (L - 1);
  const int base = lig * C;
  uint32_t lbase = 0;
  for (int b = 0; b < bin; ++b) lbase += counts[b];
  const int32_t* __restrict__ list = list_all + lbase;
  const uint32_t nlist = counts[bin];
  const int oe_del = A.o_del + A.e_del, oe_ins = A.o_ins + A.e_ins;
  const int e_del = A.e_del, e_ins = A.e_ins;

  int G[C], E[C];
  uint32_t plo[C];
  int phi[C];
#pragma unroll
  for (int c = 0; c < C; ++c) { G[c] = 0; E[c] = 0; plo[c] = 0; phi[c] = 0; }

  int task = 0, qlen = 0, tlen = 0, h0 = 0, w = 0, i = 0, beg = 0, end = 0;
  int best = 0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
  uint32_t ncell = 0;
  uint64_t toff = 0;
  uint32_t tch = 0;
  int tcb = 0;
  bool need = true, done = false;

  for (;;) {
    if (need && !done) {
      // Branch-free fetch: every lane of the group issues the atomic (only lane
      // 0 adds 1), so no lane-divergent region precedes the broadcast.
      const uint32_t mine = atomicAdd(fetch, lig == 0 ? 1u : 0u);
      const uint32_t idx = __shfl(mine, 0, L);
      if (idx >= nlist) {
        done = true;
      } else {
        task = list[idx];
        qlen = A.qlen[task];
        tlen = A.tlen[task];
        h0 = A.h0[task];
        w = A.w[task];
        const int eb = A.end_bonus[task];
        const uint64_t qoff = A.q_off[task];
        toff = A.t_off[task];
        // band clamp of ksw.c:399-407, in the same double arithmetic
        int lim = (int)((double)(qlen * A.mat_max + eb - A.o_ins) / e_ins + 1.);
        lim = lim > 1 ? lim : 1;
        w = w < lim ? w : lim;
        lim = (int)((double)(qlen * A.mat_max + eb - A.o_del) / e_del + 1.);
        lim = lim > 1 ? lim : 1;
        w = w < lim ? w : lim;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int j = base + c;
          const int qc = j < qlen ? (int)A.query[qoff + j] : 4;
          uint32_t lo = A.prof_lo[4];
          int hi = A.prof_hi[4];
          if (qc == 0) { lo = A.prof_lo[0]; hi = A.prof_hi[0]; }
          if (qc == 1) { lo = A.prof_lo[1]; hi = A.prof_hi[1]; }
          if (qc == 2) { lo = A.prof_lo[2]; hi = A.prof_hi[2]; }
          if (qc == 3) { lo = A.prof_lo[3]; hi = A.prof_hi[3]; }
          plo[c] = lo;
          phi[c] = hi;
          // first row (ksw.c:395-397): eh[0].h = h0, eh[j].h = max(h0-oe_ins-(j-1)e_ins, 0)
          int g = j == 0 ? h0 : h0 - oe_ins - (j - 1) * e_ins;
          G[c] = g > 0 ? g : 0;
          E[c] = 0;
        }
        i = 0; beg = 0; end = qlen;
        best = h0; best_i = best_j = -1; gbest = -1; gbest_i = -1; moff = 0;
        ncell = 0;
        tcb = -(1 << 30);
        if (tlen <= 0) {
          if (lig == 0) {
            if (A.score) A.score[task] = h0;
            if (A.qle) A.qle[task] = 0;
            if (A.tle) A.tle[task] = 0;
            if (A.gtle) A.gtle[task] = 0;
            if (A.gscore) A.gscore[task] = -1;
            if (A.max_off) A.max_off[task] = 0;
            if (A.cells) A.cells[task] = 0;
          }
        } else {
          need = false;
        }
      }
    }
    if (__all(done)) break;
    if (need || done) continue;

    // ---------------- one DP row (ksw.c:412-470) ----------------
    if (i - tcb >= 4 * L) {  // refill the group's window of 4*L target bases
      tcb = i;
      uint32_t v = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = i + 4 * lig + k;
        const uint32_t b = r < tlen ? (uint32_t)A.target[toff + r] : 0u;
        v |= b << (8 * k);
      }
      tch = v;
    }
    const int tr = i - tcb;
    const int t = (int)((__shfl(tch, tr >> 2, L) >> (8 * (tr & 3))) & 0xffu);

    if (beg < i - w) beg = i - w;
    if (end > i + w + 1) end = i + w + 1;
    if (end > qlen) end = qlen;
    int h1p = 0;
    if (beg == 0) {
      h1p = h0 - (A.o_del + e_del * (i + 1));
      h1p = h1p > 0 ? h1p : 0;
    }

    int M[C], ex[C];
    int run = -1;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = base + c;
      const bool inb = (j >= beg) & (j < end);
      const int s = prof_score(plo[c], phi[c], t);
      const int g = G[c];
      const int mm = g ? g + s : 0;
      M[c] = mm;
      int tt = mm - oe_ins;
      tt = tt > 0 ? tt : 0;
      const int key = inb ? tt + j * e_ins : -1;
      ex[c] = run;
      run = max(run, key);
    }
    const int lex = group_excl_max<L>(run, lig, -1);

    int h[C], nE[C];
    int mk = -1, h1c = -1;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = base + c;
      const bool inb = (j >= beg) & (j < end);
      const int x = max(ex[c], lex);
      const int f = x < 0 ? 0 : x - (j - 1) * e_ins;
      int hh = max(M[c], E[c]);
      hh = max(hh, f);
      h[c] = hh;
      int od = M[c] - oe_del;
      od = od > 0 ? od : 0;
      const int ne = max(E[c] - e_del, od);
      nE[c] = ne;
      if (inb) {
        mk = max(mk, (hh << 11) | j);
        if (j == end - 1) h1c = hh;
      }
    }
    const int hprev = __shfl_up(h[C - 1], 1, L);
    int lo = 0x7fffffff, hi = -1;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = base + c;
      const bool inb = (j >= beg) & (j < end);
      const int left = c == 0 ? hprev : h[c - 1];
      const int ng = j == beg ? h1p : left;
      if ((j >= beg) & (j <= end) & (j < qlen)) G[c] = ng;
      if (inb) E[c] = nE[c];
      else if (j == end) E[c] = 0;
      if (inb & ((G[c] | E[c]) != 0)) {
        lo = min(lo, j);
        hi = max(hi, j);
      }
    }
    mk = group_max<L>(mk);
    h1c = group_max<L>(h1c);
    lo = group_min<L>(lo);
    hi = group_max<L>(hi);

    const int h1 = end > beg ? h1c : h1p;
    const int jafter = end > beg ? end : beg;
    if (jafter == qlen) {  // ksw.c:450-453
      if (!(gbest > h1)) gbest_i = i;
      gbest = gbest > h1 ? gbest : h1;
    }
    ncell += end > beg ? (uint32_t)(end - beg) : 0u;
    const int m = mk >= 0 ? (mk >> 11) : 0;
    const int mj = mk >= 0 ? (mk & 2047) : -1;
    bool fin = false;
    if (m == 0) {
      fin = true;
    } else if (m > best) {
      best = m; best_i = i; best_j = mj;
      const int off = mj > i ? mj - i : i - mj;
      moff = moff > off ? moff : off;
    } else if (A.zdrop > 0) {
      const int di = i - best_i, dj = mj - best_j;
      const int drop = di > dj ? best - m - (di - dj) * e_del : best - m - (dj - di) * e_ins;
      if (drop > A.zdrop) fin = true;
    }
    if (!fin) {  // ksw.c:466-469
      const int nb = lo == 0x7fffffff ? end : lo;
      const int hz = h1 != 0 ? end : hi;
      beg = nb;
      const int ne = hz >= 0 ? hz + 2 : nb + 1;
      end = ne < qlen ? ne : qlen;
      ++i;
      if (i >= tlen) fin = true;
    }
    if (fin) {
      if (lig == 0) {
        if (A.score) A.score[task] = best;
        if (A.qle) A.qle[task] = best_j + 1;
        if (A.tle) A.tle[task] = best_i + 1;
        if (A.gtle) A.gtle[task] = gbest_i + 1;
        if (A.gscore) A.gscore[task] = gbest;
        if (A.max_off) A.max_off[task] = moff;
        if (A.cells) A.cells[task] = (int64_t)ncell;
      }
      need = true;
    }
  }
}

// ---- binning: validate every task and group task indices by query length ----

__device__ __forceinline__ int ext_bin_of(int qlen) {
#pragma unroll
  for (int b = 0; b < kExtBins; ++b)
    if (qlen <= bin_limit(b)) return b;
  return -1;
}

__global__ void __launch_bounds__(256)
ext_bin_count_kernel(ExtArgs A, ExtWork W) {
  __shared__ uint32_t hist[kExtBins];
  if (threadIdx.x < kExtBins) hist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ql = A.qlen[t], tl = A.tlen[t], h0 = A.h0[t], w = A.w[t], eb = A.end_bonus[t];
    int b = ext_bin_of(ql);
    const bool ok = ql >= 1 && tl >= 0 && h0 >= 1 && w >= 0 && b >= 0 && eb > -kScoreLimit &&
                    eb < kScoreLimit && (int64_t)h0 + (int64_t)ql * A.mat_max < kScoreLimit;
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
  if (threadIdx.x < kExtBins && hist[threadIdx.x]) atomicAdd(&W.counts[threadIdx.x], hist[threadIdx.x]);
}

__global__ void __launch_bounds__(256)
ext_bin_scatter_kernel(int64_t n, ExtWork W) {
  __shared__ uint32_t off[kExtBins];
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int b = 0; b < kExtBins; ++b) { off[b] = s; s += W.counts[b]; }
  }
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = W.bin_of[t];
    if (b == 0xff) continue;
    const uint32_t p = atomicAdd(&W.cursor[b], 1u);
    W.list[off[b] + p] = (int32_t)t;
  }
}

__global__ void ext_reset_kernel(ExtWork W) {
  const int k = threadIdx.x;
  if (k < kExtBins) { W.counts[k] = 0; W.cursor[k] = 0; W.fetch[k] = 0; }
  if (k == 0) { W.err[0] = 0; W.err[1] = 0x7fffffff; }
}

template <int B>
static void launch_bin(const ExtArgs& a, const ExtWork& wk, int num_cu, hipStream_t st, int64_t hint) {
  constexpr int L = kExtShape[B].lanes, C = kExtShape[B].cols;
  constexpr int groups_per_block = 256 / L;
  // persistent grid: enough blocks to fill the chip, fewer for small bins
  int64_t blocks = (int64_t)num_cu * 8;
  if (hint >= 0) {
    const int64_t need = (hint + groups_per_block - 1) / groups_per_block;
    if (need < blocks) blocks = need;
  }
  if (blocks < 1) return;
  hipLaunchKernelGGL((ext_plain_kernel<L, C>), dim3((unsigned)blocks), dim3(256), 0, st, a,
                     (const int32_t*)wk.list, (const uint32_t*)wk.counts, B, wk.fetch + B);
}

hipError_t launch_extend(const ExtArgs& a, const ExtWork& wk, int num_cu, hipStream_t st,
                         const int64_t* host_bin_counts) {
  hipLaunchKernelGGL(ext_reset_kernel, dim3(1), dim3(64), 0, st, wk);
  if (a.n <= 0) return hipGetLastError();
  int64_t blocks = (a.n + 255) / 256;
  if (blocks > (int64_t)num_cu * 16) blocks = (int64_t)num_cu * 16;
  hipLaunchKernelGGL(ext_bin_count_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, wk);
  hipLaunchKernelGGL(ext_bin_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a.n, wk);
  // longest queries first so the long tasks do not form the tail
#define BSX_BIN(B) launch_bin<B>(a, wk, num_cu, st, host_bin_counts ? host_bin_counts[B] : -1)
  BSX_BIN(7); BSX_BIN(6); BSX_BIN(5); BSX_BIN(4); BSX_BIN(3); BSX_BIN(2); BSX_BIN(1); BSX_BIN(0);
#undef BSX_BIN
  return hipGetLastError();
}

}  // namespace bsx
