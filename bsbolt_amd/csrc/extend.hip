// Bit-exact batched ksw_extend2 (bsbolt/External/BWA/ksw.c:380-479) for gfx950.
//
// Layout.  A task is owned by a group of L lanes (L = 16, 32 or 64; 64/L tasks
// per wave64); lane g holds query columns [g*C, g*C+C) of the DP state in VGPRs:
//   G[c] = eh[j].h = H(i-1, j-1)   (diagonal predecessor)
//   E[c] = eh[j].e = E(i, j)
// Rows (target bases) are processed one at a time exactly like the reference's
// row loop, so the adaptive band [beg,end), the z-drop and the early exits are
// decided on the same values.  Columns >= end are never written, which keeps
// the stale eh[] entries the reference reads when the band regrows
// (ksw.c:466-469).
//
// Row recurrences.  F is built from M, not H (ksw.c:433-447), so
//   F(j) = max_{beg<=k<j}(max(M_k - oe_ins, 0) + k*e_ins) - (j-1)*e_ins  (j > beg)
// and F(beg) = 0: an exclusive prefix maximum, computed as a lane-local scan
// plus a DPP scan across the group.  Keys are shifted by +e_ins so that every
// key is >= 1 and 0 is the scan identity (fused v_max_i32_dpp, bound_ctrl).
//
// Dead columns.  beg never decreases, so columns < beg are never read again.
// The kernel keeps them at G = E = 0 (zeroing them when the band cap i-w moves
// beg past non-zero columns).  A dead column then has M = 0, key = (k+1)*e_ins
// and H = 0, which can neither raise F of a live column above its true value
// (its key contributes (k-j+1)*e_ins <= 0 <= F(j)) nor change the row maximum,
// the G shift or the zero scan.  So in-band masks are needed only against
// columns >= end, and G(j) = H(i, j-1) holds for every j <= end with
// H(i, -1) = the first-column value of ksw.c:421-424.
//
// Work distribution.  Tasks are binned by query length (bin b: qlen <= L*C);
// the binning pass writes one 48-byte descriptor per task in bin order.  Block
// k of a bin's launch owns descriptors [k*T, k*T+T); its groups take them one
// at a time through an LDS counter, so no device-scope atomic is on the task
// path (same-address device atomics serialise at a few per microsecond) and the
// hardware dispatcher balances the blocks.  The next task's descriptor, query
// bases and first target word are loaded one task ahead, so a task starts
// without a memory round trip.  Groups run a flattened loop (start-task-or-row
// per iteration) so the groups of a wave never wait for one another between
// tasks.
#include "bsx_internal.h"
#include "ext_common.h"

namespace bsx {



// waves per SIMD the register budget must allow (4 keeps VGPRs <= 128)
constexpr int plain_min_waves(int C) { return C <= 6 ? 4 : C <= 8 ? 3 : C <= 16 ? 2 : 1; }

template <int L, int C, bool E1>
__global__ void __launch_bounds__(kExtBlock, plain_min_waves(C))
ext_plain_kernel(ExtKArgs A, const ExtDesc* __restrict__ desc_all, const uint32_t* __restrict__ counts,
                 int bin) {
  __shared__ uint32_t lds_next;
  const int lane = threadIdx.x & 63;
  const int lig = lane & (L - 1);
  const int base = lig * C;
  uint32_t lbase = 0;
  for (int b = 0; b < bin; ++b) lbase += counts[b];
  const uint32_t nlist = counts[bin];
  const uint32_t blk0 = blockIdx.x * (uint32_t)A.tpb;
  if (blk0 >= nlist) return;  // uniform over the block
  const uint32_t blk1 = min(blk0 + (uint32_t)A.tpb, nlist);
  if (threadIdx.x == 0) lds_next = blk0;
  __syncthreads();
  const ExtDesc* __restrict__ desc = desc_all + lbase;
  // E1: both gap extensions are 1 (BSBolt's -E 1,1), so the per-column
  // multiples of e_ins fold into immediates
  const int e_del = E1 ? 1 : A.e_del, e_ins = E1 ? 1 : A.e_ins;
  const int oe_del = A.o_del + e_del, oe_ins = A.o_ins + e_ins;

  // next descriptor position for this group; blk1 once the block's range is taken
  auto next_pos = [&]() -> uint32_t {
    // Branch-free fetch: every lane issues the atomic and only lane 0 of the
    // group adds 1.  No lane-divergent region may precede the broadcast: with
    // `if (lig == 0)` around the atomic the compiler can give the other lanes a
    // loop of their own that reaches the broadcast without lane 0's new value.
    const uint32_t mine = atomicAdd(&lds_next, lig == 0 ? 1u : 0u);
    const uint32_t p = __shfl(mine, 0, L);
    return p < blk1 ? p : blk1;
  };

  // ---- two-stage prefetch ----
  // stage 2: descriptor of the task after next (loaded two tasks ahead)
  // stage 1: descriptor fields + query/target bytes of the next task
  // Every load is unconditional with a clamped in-bounds address, so nothing
  // waits on it until the next task starts.
  bool d2_ok = false;
  int4 d2a = make_int4(0, 0, 0, 0), d2b = d2a, d2c = d2a;
  auto load_desc = [&]() {
    const uint32_t p = next_pos();
    d2_ok = p < blk1;
    const int4* d = (const int4*)(desc + (d2_ok ? p : blk0));
    d2a = d[0]; d2b = d[1]; d2c = d[2];
  };
  bool nx_ok = false;
  int nx_task = 0, nx_qlen = 0, nx_tlen = 0, nx_h0 = 0, nx_w = 0;
  uint64_t nx_toff = 0, nx_qoff = 0;
  uint32_t nx_q[C];   // raw query bytes of the owned columns (clamped index)
  uint32_t nx_tb[4];  // raw target bytes [4*lig, 4*lig+4) (clamped index)
  auto stage1 = [&]() {  // descriptor d2 -> next task, issue its data loads
    nx_ok = d2_ok;
    nx_task = d2a.x; nx_qlen = d2a.y; nx_tlen = d2a.z; nx_h0 = d2a.w;
    nx_w = d2b.x;
    const uint64_t qoff = (uint64_t)(uint32_t)d2b.z | ((uint64_t)(uint32_t)d2b.w << 32);
    nx_toff = (uint64_t)(uint32_t)d2c.x | ((uint64_t)(uint32_t)d2c.y << 32);
    nx_qoff = qoff;
    const int qmax = nx_qlen > 0 ? nx_qlen - 1 : 0;
#pragma unroll
    for (int c = 0; c < C; ++c) nx_q[c] = A.query[qoff + min(base + c, qmax)];
    // a task without target bases reads its first query base instead (always valid)
    const uint8_t* tp = nx_tlen > 0 ? A.target + nx_toff : A.query + qoff;
    const int tmax = nx_tlen > 0 ? nx_tlen - 1 : 0;
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) nx_tb[k2] = tp[min(4 * lig + k2, tmax)];
  };
  // 4*L-base target window of the running task, double-buffered.  Always
  // issued (clamped index; a task without target bases reads its first query
  // base), so the prefetched registers never merge with a not-loaded path.
  auto load_window = [&](const uint8_t* tp, int tlen_, int first, uint32_t* out) {
    const int tmax = tlen_ > 0 ? tlen_ - 1 : 0;
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) out[k2] = tp[min(first + 4 * lig + k2, tmax)];
  };

  // ---- current task ----
  int G[C], E[C];
  uint32_t plo[C];
  uint32_t qpk[(C + 3) / 4];  // query codes, 4 per register (for rows whose target base is N)
#pragma unroll
  for (int c = 0; c < C; ++c) { G[c] = 0; E[c] = 0; plo[c] = 0; nx_q[c] = 4; }
#pragma unroll
  for (int c = 0; c < (C + 3) / 4; ++c) qpk[c] = 0;
  int task = 0, qlen = 0, tlen = 0, h0 = 0, w = 0, i = 0, beg = 0, end = 0;
  int best = 0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
  uint32_t ncell = 0, tch = 0;
  uint32_t tnx[4] = {0, 0, 0, 0};  // raw bytes of the following window
  uint64_t toff = 0;
  const uint8_t* tbase = A.target;  // target bases of the running task (or a valid dummy)
  int tcb = 0;
  bool need = true, done = false;
  const int kE0 = base * e_ins;  // column j contributes j*e_ins = kE0 + c*e_ins

  load_desc();
  stage1();
  load_desc();
  for (;;) {
    if (need && !done) {
      if (!nx_ok) {
        done = true;
      } else {
        task = nx_task; qlen = nx_qlen; tlen = nx_tlen; h0 = nx_h0; w = nx_w; toff = nx_toff;
        {
          uint32_t v = 0;
#pragma unroll
          for (int k2 = 0; k2 < 4; ++k2) v |= (4 * lig + k2 < tlen ? (nx_tb[k2] & 0xffu) : 0u) << (8 * k2);
          tch = v;
        }
        tcb = 0;
#pragma unroll
        for (int c = 0; c < (C + 3) / 4; ++c) qpk[c] = 0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const int j = base + c;
          const uint32_t qc = j < qlen ? (nx_q[c] & 0xffu) : 4u;
          uint32_t lo = A.prof_lo[4];
          if (qc == 0) lo = A.prof_lo[0];
          if (qc == 1) lo = A.prof_lo[1];
          if (qc == 2) lo = A.prof_lo[2];
          if (qc == 3) lo = A.prof_lo[3];
          plo[c] = lo;
          qpk[c / 4] |= (qc > 4 ? 4u : qc) << (8 * (c % 4));
          // first row (ksw.c:395-397): eh[0].h = h0, eh[j].h = max(h0-oe_ins-(j-1)e_ins, 0)
          const int g = j == 0 ? h0 : h0 - oe_ins - (j - 1) * e_ins;
          G[c] = g > 0 ? g : 0;
          E[c] = 0;
        }
        i = 0; beg = 0; end = qlen;
        best = h0; best_i = best_j = -1; gbest = -1; gbest_i = -1; moff = 0;
        ncell = 0;
        tbase = tlen > 0 ? A.target + toff : A.query + nx_qoff;
        load_window(tbase, tlen, 4 * L, tnx);
        stage1();
        load_desc();
        if (tlen <= 0) {
          write_rec<L>(A.rec, task, lig, h0, 0, 0, 0, -1, 0, 0u);  // ksw.c:473-478 with no row
        } else {
          need = false;
        }
      }
    }
    if (__all(done)) break;
    if (need || done) continue;

    // ---------------- one DP row (ksw.c:412-470) ----------------
    if (i - tcb >= 4 * L) {  // move to the prefetched window, prefetch the one after
      tcb = i;
      uint32_t v = 0;
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v |= (i + 4 * lig + k2 < tlen ? (tnx[k2] & 0xffu) : 0u) << (8 * k2);
      tch = v;
      load_window(tbase, tlen, i + 4 * L, tnx);
    }
    const int tr = i - tcb;
    const int t = (int)((__shfl(tch, tr >> 2, L) >> (8 * (tr & 3))) & 0xffu);
    const bool tN = t > 3;

    const int cap_beg = i - w;
    if (beg < cap_beg) {  // the band cap moves beg: keep the dead-column invariant
      beg = cap_beg;
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (base + c < beg) { G[c] = 0; E[c] = 0; }
    }
    if (end > i + w + 1) end = i + w + 1;
    if (end > qlen) end = qlen;
    int h1p = 0;
    if (beg == 0) {
      h1p = h0 - (A.o_del + e_del * (i + 1));
      h1p = h1p > 0 ? h1p : 0;
    }

    // live columns of this lane: c < d (j < end).  Every other band mask of
    // the row derives from these C+1 compares (scalar mask logic):
    //   j <= end  <=>  c-1 < d;   j == end-1  <=>  lt[c] && !lt[c+1];   j == end  <=>  le && !lt[c]
    const int d = end - base;
    bool lt[C + 1];
#pragma unroll
    for (int c = 0; c <= C; ++c) lt[c] = c < d;
    const bool le0 = d >= 0;

    int M[C], ex[C];
    // substitution scores of this row; a target base N (code 4, which the
    // 2-bit reference never produces) takes the slow path, decided per wave
    const uint32_t sh8 = 8u * (uint32_t)(t & 3);
#pragma unroll
    for (int c = 0; c < C; ++c) M[c] = __builtin_amdgcn_sbfe((int)plo[c], sh8, 8u);
    if (__builtin_expect(__any(tN), 0)) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const uint32_t qc = (qpk[c / 4] >> (8 * (c % 4))) & 0xffu;
        int sN = A.prof_hi[4];
        if (qc == 0) sN = A.prof_hi[0];
        if (qc == 1) sN = A.prof_hi[1];
        if (qc == 2) sN = A.prof_hi[2];
        if (qc == 3) sN = A.prof_hi[3];
        M[c] = tN ? sN : M[c];
      }
    }
    int run = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int g = G[c];
      const int mm = g ? g + M[c] : 0;
      M[c] = mm;
      // key + e_ins = max(M - oe_ins, 0) + (j+1)*e_ins = max(M + (j+1)*e_ins - oe_ins, (j+1)*e_ins)  (>= 1)
      const int kb = kE0 + (c + 1) * e_ins;
      const int key = max(mm + (kb - oe_ins), kb);
      ex[c] = run;
      run = max(run, key);
    }
    const int lex = grp_excl_from_incl<L>(grp_incl_max<L>(run), lig);

    int h[C];
    int mk = 0, h1c = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int j = base + c;
      const int x = max(ex[c], lex);
      const int f = x - (kE0 + c * e_ins);  // F(j); 0 when no column precedes j
      const int hh = max(max(M[c], E[c]), f);
      h[c] = hh;
      const int ne = max(max(E[c] - e_del, M[c] - oe_del), 0);
      mk = lt[c] ? max(mk, (hh << 11) | j) : mk;
      E[c] = lt[c] ? ne : E[c];
      h1c = (lt[c] && !lt[c + 1]) ? hh : h1c;  // j == end - 1
    }
    // G(j) = H(i, j-1) for j <= end; column 0 receives the first-column value
    const int hsh = __builtin_amdgcn_update_dpp(0, h[C - 1], L == 16 ? 0x111 : 0x138, 0xf, 0xf, true);
    const int hprev = lig == 0 ? h1p : hsh;
    uint32_t nzb = 0;  // bit c: column c holds a non-zero G or E after the row
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int left = c == 0 ? hprev : h[c - 1];
      const bool le = c == 0 ? le0 : lt[c - 1];
      G[c] = le ? left : G[c];
      E[c] = (le && !lt[c]) ? 0 : E[c];  // j == end
      nzb |= (uint32_t)min(G[c] | E[c], 1) << c;
    }
    // only columns j < end count: keep bits c < d
    nzb &= d >= C ? 0xffffffffu : d > 0 ? (1u << d) - 1u : 0u;
    int hik = 0, lok = 0;  // last non-zero column + 1; 2048 - first non-zero column
    if (nzb) {
      hik = base + 32 - __builtin_clz(nzb);
      lok = 2048 - base - __builtin_ctz(nzb);
    }
    mk = grp_max<L>(mk, lane);
    h1c = grp_max<L>(h1c, lane);
    hik = grp_max<L>(hik, lane);
    lok = grp_max<L>(lok, lane);

    const int h1 = end > beg ? h1c : h1p;
    const int jafter = end > beg ? end : beg;
    if (jafter == qlen) {  // ksw.c:450-453
      if (!(gbest > h1)) gbest_i = i;
      gbest = gbest > h1 ? gbest : h1;
    }
    ncell += end > beg ? (uint32_t)(end - beg) : 0u;
    const int m = end > beg ? mk >> 11 : 0;
    const int mj = mk & 2047;
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
      const int nb = lok > 0 ? 2048 - lok : end;
      const int hz = h1 != 0 ? end : hik - 1;
      beg = nb;
      const int ne = hz >= 0 ? hz + 2 : nb + 1;
      end = ne < qlen ? ne : qlen;
      ++i;
      if (i >= tlen) fin = true;
    }
    if (fin) {
      write_rec<L>(A.rec, task, lig, best, best_j + 1, best_i + 1, gbest_i + 1, gbest, moff, ncell);  // ksw.c:473-478
      need = true;
    }
  }
}

// ---- binning: validate every task, group tasks by query length, write descriptors ----

__device__ __forceinline__ int ext_bin_of(int qlen) {
#pragma unroll
  for (int b = 0; b < kExtBins; ++b)
    if (qlen <= bin_limit(b)) return b;
  return -1;
}

// mode 0: validate every task, write bin_of and count per bin and row bucket
// mode 1: validate only (the learned-band form counts after its narrow pass)
// mode 2: count the valid tasks still flagged in need[] (no validation)
__global__ void __launch_bounds__(256)
ext_bin_count_kernel(ExtArgs A, ExtWork W, int mode) {
  constexpr int NB = kExtBins * kRowBuckets;
  __shared__ uint32_t hist[kExtBins], bhist[NB];
  if (threadIdx.x < kExtBins) hist[threadIdx.x] = 0;
  if (threadIdx.x < NB) bhist[threadIdx.x] = 0;
  __syncthreads();
  int bad = 0, first_bad = 0x7fffffff;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (mode == 2) {
      const int b = W.bin_of[t];
      if (b != 0xff && W.need_plain[t]) {
        atomicAdd(&hist[b], 1u);
        atomicAdd(&bhist[b * kRowBuckets + row_bucket(A.tlen[t], A.h0[t])], 1u);
      }
      continue;
    }
    const int ql = A.qlen[t], tl = A.tlen[t], h0 = A.h0[t], w = A.w[t], eb = A.end_bonus[t];
    int b = ext_bin_of(ql);
    const bool ok = ql >= 1 && tl >= 0 && h0 >= 1 && w >= 0 && b >= 0 && eb > -kScoreLimit &&
                    eb < kScoreLimit && (int64_t)h0 + (int64_t)ql * A.mat_max < kScoreLimit;
    if (!ok) {
      b = 0xff;
      ++bad;
      first_bad = min(first_bad, (int)t);
    } else if (mode == 0) {
      atomicAdd(&hist[b], 1u);
      atomicAdd(&bhist[b * kRowBuckets + row_bucket(tl, h0)], 1u);
    }
    W.bin_of[t] = (uint8_t)b;
  }
  if (bad) {
    atomicAdd(&W.err[0], bad);
    atomicMin(&W.err[1], first_bad);
  }
  __syncthreads();
  if (threadIdx.x < kExtBins && hist[threadIdx.x]) atomicAdd(&W.counts[threadIdx.x], hist[threadIdx.x]);
  if (threadIdx.x < NB && bhist[threadIdx.x]) atomicAdd(&W.bcount[threadIdx.x], bhist[threadIdx.x]);
}

// Each block handles a contiguous range of tasks: it counts them per bin and
// row bucket in LDS, reserves one range per (bin, bucket) with a single global
// atomic, and writes the descriptors of its tasks into those ranges.  A task's
// rank in its bin's longest-first order is its bucket's base plus its place in
// the bucket; lane-kernel bins spread the ranks over the fetch slices.
__global__ void __launch_bounds__(256)
ext_bin_scatter_kernel(ExtArgs A, ExtWork W, int64_t per_block, const uint8_t* __restrict__ need) {
  constexpr int NB = kExtBins * kRowBuckets;
  __shared__ uint32_t off[NB], hist[NB], cur[NB], lbase[kExtBins], lsize[kExtBins];
  const int64_t t0 = (int64_t)blockIdx.x * per_block;
  const int64_t t1 = min(A.n, t0 + per_block);
  if (threadIdx.x < NB) { hist[threadIdx.x] = 0; cur[threadIdx.x] = 0; }
  __syncthreads();
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int b = W.bin_of[t];
    if (b != 0xff && (!need || need[t])) atomicAdd(&hist[b * kRowBuckets + row_bucket(A.tlen[t], A.h0[t])], 1u);
  }
  __syncthreads();
  if (threadIdx.x < NB) {
    const int b = threadIdx.x / kRowBuckets, k = threadIdx.x % kRowBuckets;
    uint32_t s = 0;
    for (int j = 0; j < k; ++j) s += W.bcount[b * kRowBuckets + j];
    const uint32_t hcount = hist[threadIdx.x];
    off[threadIdx.x] = s + (hcount ? atomicAdd(&W.bcursor[threadIdx.x], hcount) : 0u);
  }
  if (threadIdx.x < kExtBins) {
    uint32_t s = 0;
    for (int b = 0; b < (int)threadIdx.x; ++b) s += W.counts[b];
    lbase[threadIdx.x] = s;
    lsize[threadIdx.x] = W.counts[threadIdx.x];
  }
  __syncthreads();
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int b = W.bin_of[t];
    if (b == 0xff || (need && !need[t])) continue;
    const int kb = b * kRowBuckets + row_bucket(A.tlen[t], A.h0[t]);
    const uint32_t r = off[kb] + atomicAdd(&cur[kb], 1u);
    const uint32_t p = (kLaneBinMask >> b) & 1u ? slice_begin(lsize[b], (int)(r % kExtSlices)) + r / kExtSlices : r;
    write_desc(A, t, W.desc + lbase[b] + p);
  }
}

static_assert(kExtBins * kRowBuckets <= kFetchSlots && kExtBins * kRowBuckets <= 256, "bucket counters");
__global__ void ext_reset_kernel(ExtWork W) {
  const int k = threadIdx.x;
  if (k < kCountSlots) { W.counts[k] = 0; W.cursor[k] = 0; }
  if (k < kFetchSlots) W.fetch[k] = 0;
  if (k < 2 * kLaneFrameClasses) W.split[k] = 0;
  if (k < kExtBins * kRowBuckets) { W.bcount[k] = 0; W.bcursor[k] = 0; }
  if (k < kLaneFrameClasses * kRowBuckets) { W.fcount[k] = 0; W.fcursor[k] = 0; }
  if (k == 0) { W.err[0] = 0; W.err[1] = 0x7fffffff; }
}

// Copies the packed per-task records into the caller's (nullable) arrays.
__global__ void __launch_bounds__(256)
ext_unpack_kernel(ExtArgs A, ExtWork W, const int32_t* __restrict__ rec) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (W.bin_of[t] == 0xff) continue;  // rejected: result slots stay untouched
    const int4 a = *(const int4*)(rec + t * 8);
    const int4 b = *(const int4*)(rec + t * 8 + 4);
    if (A.score) A.score[t] = a.x;
    if (A.qle) A.qle[t] = a.y;
    if (A.tle) A.tle[t] = a.z;
    if (A.gtle) A.gtle[t] = a.w;
    if (A.gscore) A.gscore[t] = b.x;
    if (A.max_off) A.max_off[t] = b.y;
    if (A.cells) A.cells[t] = (int64_t)(uint32_t)b.z;
  }
}

// one block per k.tpb descriptors of bin B; without a host-side count the
// grid covers the whole batch and surplus blocks exit at once
template <int L, int C, bool E1>
static void launch_group(const ExtKArgs& k, const ExtWork& wk, int B, int64_t n, hipStream_t st) {
  const int64_t blocks = (n + k.tpb - 1) / k.tpb;
  if (blocks < 1) return;
  time_begin(wk, B, st);
  hipLaunchKernelGGL((ext_plain_kernel<L, C, E1>), dim3((unsigned)blocks), dim3(kExtBlock), 0, st, k,
                     (const ExtDesc*)wk.desc, (const uint32_t*)wk.counts, B);
  time_end(wk, B, st);
}

template <int B, bool E1>
static void launch_bin(const ExtKArgs& k, const ExtWork& wk, int64_t n_all, int num_cu, hipStream_t st,
                       int64_t hint) {
  constexpr int L = kExtShape[B].lanes, C = kExtShape[B].cols;
  const int64_t n = hint >= 0 ? hint : n_all;
  if constexpr (L == 1) {
    launch_lane_bin<C, E1>(k, wk, B, num_cu, st, hint);
  } else if constexpr (B == kTunedBin0 || B == kTunedBin0 + 1) {
    constexpr int V = B - kTunedBin0;
    if (wk.shape_var[V] == 1)
      launch_group<kAltShape[V][1].lanes, kAltShape[V][1].cols, E1>(k, wk, B, n, st);
    else
      launch_group<L, C, E1>(k, wk, B, n, st);
  } else {
    launch_group<L, C, E1>(k, wk, B, n, st);
  }
}

template <int B, bool E1>
static void launch_bins_from(const ExtKArgs& k, const ExtWork& wk, int64_t n, int num_cu, hipStream_t st,
                             const int64_t* hc) {
  launch_bin<B, E1>(k, wk, n, num_cu, pick_stream(wk.ss, st, kBinStream[B]), hc ? hc[B] : -1);
  if constexpr (B > 0) launch_bins_from<B - 1, E1>(k, wk, n, num_cu, st, hc);
}

template <bool E1>
static void launch_bins(const ExtKArgs& k, const ExtWork& wk, int64_t n, int num_cu, hipStream_t st,
                        const int64_t* hc) {
  // longest queries first so the long tasks do not form the tail
  launch_bins_from<kExtBins - 1, E1>(k, wk, n, num_cu, st, hc);
}

hipError_t launch_extend(const ExtArgs& a, const ExtWork& wk, int num_cu, hipStream_t st,
                         const int64_t* host_bin_counts, const BandPlan* plan) {
  hipLaunchKernelGGL(ext_reset_kernel, dim3(1), dim3(kFetchSlots), 0, st, wk);
  if (a.n <= 0) return hipGetLastError();
  int64_t blocks = (a.n + 255) / 256;  // unpack pass
  if (blocks > (int64_t)num_cu * 16) blocks = (int64_t)num_cu * 16;
  // count and scatter passes: each block ends with one global atomic per
  // non-empty bin and row bucket, so fewer, larger blocks
  const int64_t sblocks = blocks < (int64_t)num_cu * 4 ? blocks : (int64_t)num_cu * 4;
  const bool band = plan && plan->mode != kBandPlain;
  ExtKArgs k{};
  k.query = a.query; k.target = a.target; k.rec = wk.rec;
  k.o_del = a.o_del; k.e_del = a.e_del; k.o_ins = a.o_ins; k.e_ins = a.e_ins; k.zdrop = a.zdrop;
  k.tpb = wk.tpb_plain;
  for (int q = 0; q < 5; ++q) { k.prof_lo[q] = a.prof_lo[q]; k.prof_hi[q] = a.prof_hi[q]; }
  for (int t = 0; t < 5; ++t) {  // the same matrix by target code
    uint32_t lo = 0;
    for (int q = 0; q < 4; ++q) {
      const uint32_t b = t < 4 ? (a.prof_lo[q] >> (8 * t)) & 0xffu : (uint32_t)a.prof_hi[q] & 0xffu;
      lo |= b << (8 * q);
    }
    k.trow_lo[t] = lo;
    k.trow_hi[t] = t < 4 ? (a.prof_lo[4] >> (8 * t)) & 0xffu : (uint32_t)a.prof_hi[4] & 0xffu;
  }
  hipLaunchKernelGGL(ext_bin_count_kernel, dim3((unsigned)sblocks), dim3(256), 0, st, a, wk, band ? 1 : 0);
  if (band) {
    ExtKArgs kn = k;
    kn.tpb = wk.tpb_narrow;
    hipError_t e = launch_narrow(a, kn, wk, num_cu, st, *plan);
    if (e != hipSuccess || plan->skip_plain) return e;
    hipLaunchKernelGGL(ext_bin_count_kernel, dim3((unsigned)sblocks), dim3(256), 0, st, a, wk, 2);
    host_bin_counts = nullptr;
  }
  const int64_t per_block = (a.n + sblocks - 1) / sblocks;
  hipLaunchKernelGGL(ext_bin_scatter_kernel, dim3((unsigned)sblocks), dim3(256), 0, st, a, wk, per_block,
                     band ? (const uint8_t*)wk.need_plain : nullptr);
  ss_fork(wk.ss, st);
  if (a.e_del == 1 && a.e_ins == 1) launch_bins<true>(k, wk, a.n, num_cu, st, host_bin_counts);
  else launch_bins<false>(k, wk, a.n, num_cu, st, host_bin_counts);
  ss_join(wk.ss, st);
  hipLaunchKernelGGL(ext_unpack_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, wk, (const int32_t*)wk.rec);
  return hipGetLastError();
}

}  // namespace bsx
