// Lane-per-task form of the bit-exact ksw_extend2 (bsbolt/External/BWA/ksw.c:380-479)
// for short queries.
//
// Each lane of a wave owns one whole task: the DP state of its query columns
// 0..Q-1 lives in that lane's VGPRs (G[c] = eh[c].h = H(i-1, c-1), E[c] =
// eh[c].e) and a row is the reference's column loop run sequentially inside the
// lane -- the same recurrences in the same order, so F needs no scan and no
// value crosses lanes.  The cost of a row is ~22 VALU instructions per column
// for 64 tasks at once, against the group kernels' fixed per-row cost of prefix
// scans and group reductions paid per 4 tasks.
//
// Band.  Columns < beg are dead: beg never decreases and the kernel keeps them
// at G = E = 0 (zeroing them when the cap i-w moves beg past non-zero columns),
// so running them is harmless -- their M, E, F and H are 0, which leaves F(beg)
// = 0, H(i, beg-1) = 0 and the row maximum unchanged.  Columns >= end keep
// their stale state (the reference reads it again when the band regrows): the
// per-column masks c < end and c <= end select the writes.
//
// Work distribution.  Waves are persistent: a bin's work list is split into
// eight slices, each with a fetch counter; a wave starts on slice blockIdx % 8
// and moves on when its slice runs dry.  Every K rows the wave stops at a sync
// point where lanes without a task install their prefetched next task, lanes
// without a next task move the prefetched descriptor into place and issue the
// query/target loads, free descriptor slots are refilled with one atomic per
// wave.  All loads a sync point consumes were issued K rows earlier, so the
// rows themselves never wait on memory.  Target bases come in 16-base windows, the next one prefetched.
#include <algorithm>

#include "bsx_internal.h"
#include "ext_common.h"

namespace bsx {

constexpr int kLaneRows = 4;  // rows between sync points

__host__ __device__ constexpr int lane_min_waves(int) { return 2; }

// 4 bytes starting at byte `sh` (0..3) of the 8-byte value hi:lo
__device__ __forceinline__ uint32_t bytes_at(uint32_t lo, uint32_t hi, int sh) {
  return __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)sh);
}

template <int Q, bool E1>
__global__ void __launch_bounds__(64, lane_min_waves(Q))
ext_lane_kernel(ExtKArgs A, const ExtDesc* __restrict__ desc_all, const uint32_t* __restrict__ counts, int bin,
                uint32_t* __restrict__ fetch) {
  constexpr int QW = Q / 4;          // packed query dwords (4 codes each)
  constexpr int NZ = (Q + 31) / 32;  // dwords of the non-zero column mask
  constexpr int K = kLaneRows;
  static_assert(Q % 4 == 0 && Q <= 64, "lane kernel query width");
  // scores of target code t against query codes 0..3 (bytes of .x) and 4 (bytes of .y);
  // codes 5..7 read as 4, like the group kernels
  __shared__ uint2 trow[8];
  const int lane = threadIdx.x;
  if (lane < 8) {
    const int t = lane < 5 ? lane : 4;
    const uint32_t n = A.trow_hi[t] & 0xffu;
    trow[lane] = make_uint2(A.trow_lo[t], n * 0x01010101u);
  }
  __syncthreads();

  uint32_t lbase = 0;
  for (int b = 0; b < bin; ++b) lbase += counts[b];
  const uint32_t nlist = counts[bin];
  const ExtDesc* __restrict__ desc = desc_all + lbase;
  uint32_t* __restrict__ fslot = fetch + bin * kExtSlices;
  const int cs0 = (int)(blockIdx.x % kExtSlices);
  int cs = cs0;
  bool exhausted = nlist == 0;
  // work-list indices reserved by this wave: current range and the next one; a
  // reservation is issued at one sync point and collected at the next
  constexpr uint32_t CH = 64;  // tasks per reservation
  uint32_t plo = 0, phi = 0, nlo = 0, nhi = 0;
  bool apend = false;
  uint32_t aret = 0;
  int acs = 0;

  const int e_del = E1 ? 1 : A.e_del, e_ins = E1 ? 1 : A.e_ins;
  const int oe_del = A.o_del + e_del, oe_ins = A.o_ins + e_ins;

  // ---- per-lane task pipeline ----
  // stage 2: raw descriptor; stage 1: next task's fields + raw query/target dwords
  bool d2_ok = false;
  int4 d2a = make_int4(0, 0, 0, 0), d2b = d2a;
  int2 d2c = make_int2(0, 0);
  bool nx_ok = false;
  int nx_task = 0, nx_qlen = 0, nx_tlen = 0, nx_h0 = 0, nx_w = 0, nx_qsh = 0, nx_tsh = 0;
  uint64_t nx_toff = 0;
  uint32_t nx_q[QW + 1], nx_t[5];
#pragma unroll
  for (int k = 0; k <= QW; ++k) nx_q[k] = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) nx_t[k] = 0;

  // current task
  bool act = false;
  int task = 0, qlen = 0, tlen = 0, h0 = 0, W = 0, i = 0, beg = 0, end = 0;
  int best = 0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
  uint32_t ncell = 0;
  uint64_t toff = 0;
  int G[Q], E[Q];
  uint32_t qpk[QW];
  uint32_t tb[4], tn[5];  // current 16-base target window; raw dwords of the next one
  int tn_sh = 0, wb = 0;
#pragma unroll
  for (int c = 0; c < Q; ++c) { G[c] = 0; E[c] = 0; }
#pragma unroll
  for (int k = 0; k < QW; ++k) qpk[k] = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) tb[k] = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) tn[k] = 0;

  // raw aligned dwords covering bytes [first, first + 16) of a sequence of
  // `len` bases at `base` (indices clamped to the sequence's last dword, so no
  // byte outside a mapped dword of the sequence is read)
  auto load_window = [&](const uint8_t* base, int len, int first, uint32_t* raw, int& sh) {
    const uintptr_t a0 = (uintptr_t)base;
    const uint32_t* p = (const uint32_t*)(a0 & ~(uintptr_t)3);
    const int s0 = (int)(a0 & 3);
    const int last = (s0 + (len > 0 ? len : 1) - 1) >> 2;
    const int d0 = (s0 + first) >> 2;
    sh = (s0 + first) & 3;
#pragma unroll
    for (int k = 0; k < 5; ++k) raw[k] = p[min(d0 + k, last)];
  };

  for (;;) {
    // ================= sync point =================
    // (1) consume: install next tasks, rotate target windows
    if (!act && nx_ok) {
      act = true;
      nx_ok = false;
      task = nx_task; qlen = nx_qlen; tlen = nx_tlen; h0 = nx_h0; W = nx_w; toff = nx_toff;
#pragma unroll
      for (int k = 0; k < QW; ++k) qpk[k] = bytes_at(nx_q[k], nx_q[k + 1], nx_qsh);
#pragma unroll
      for (int k = 0; k < 4; ++k) tb[k] = bytes_at(nx_t[k], nx_t[k + 1], nx_tsh);
      wb = 0;
      {  // first row (ksw.c:395-397): eh[0].h = h0, eh[j].h = max(h0 - oe_ins - (j-1)*e_ins, 0)
        int g = h0 - oe_ins;
        G[0] = h0;
        E[0] = 0;
#pragma unroll
        for (int c = 1; c < Q; ++c) {
          G[c] = max(g, 0);
          E[c] = 0;
          g -= e_ins;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      i = 0; beg = 0; end = qlen;
      best = h0; best_i = best_j = -1; gbest = -1; gbest_i = -1; moff = 0;
      ncell = 0;
      if (tlen <= 0) {  // ksw.c:473-478 with no row: qle = tle = gtle = 0, gscore = -1
        act = false;
        int4* r = (int4*)(A.rec + (int64_t)task * 8);
        r[0] = make_int4(h0, 0, 0, 0);
        r[1] = make_int4(-1, 0, 0, 0);
      }
    }
    bool rot = false;
    if (act && i - wb >= 16) {  // window consumed: the prefetched one takes its place
#pragma unroll
      for (int k = 0; k < 4; ++k) tb[k] = bytes_at(tn[k], tn[k + 1], tn_sh);
      wb += 16;
      rot = true;
    }
    uint32_t t4 = tb[0];  // target bases of rows i .. i+3
    {
      const int q = (i - wb) >> 2;
      t4 = q == 1 ? tb[1] : t4;
      t4 = q == 2 ? tb[2] : t4;
      t4 = q == 3 ? tb[3] : t4;
    }
    const bool fresh = act && i == 0;  // installed at this sync point
    // (2) stage 2 -> stage 1: next task fields and its data loads
    if (!nx_ok && d2_ok) {
      nx_ok = true;
      d2_ok = false;
      nx_task = d2a.x; nx_qlen = d2a.y; nx_tlen = d2a.z; nx_h0 = d2a.w;
      nx_w = d2b.x;
      const uint64_t qoff = (uint64_t)(uint32_t)d2b.z | ((uint64_t)(uint32_t)d2b.w << 32);
      nx_toff = (uint64_t)(uint32_t)d2c.x | ((uint64_t)(uint32_t)d2c.y << 32);
      {
        const uintptr_t a0 = (uintptr_t)(A.query + qoff);
        const uint32_t* p = (const uint32_t*)(a0 & ~(uintptr_t)3);
        nx_qsh = (int)(a0 & 3);
        const int last = (nx_qsh + nx_qlen - 1) >> 2;
#pragma unroll
        for (int k = 0; k <= QW; ++k) nx_q[k] = p[min(k, last)];
      }
      // a task without target bases reads its first query base instead (always valid)
      const uint8_t* tp = nx_tlen > 0 ? A.target + nx_toff : A.query + qoff;
      load_window(tp, nx_tlen > 0 ? nx_tlen : 1, 0, nx_t, nx_tsh);
    }
    // (3) descriptor slots: collect the reservation issued at the last sync
    // point, hand out reserved indices, reserve more when running low
    if (apend) {
      const uint32_t old = __builtin_amdgcn_readfirstlane(aret);
      apend = false;
      const uint32_t s0 = (uint32_t)((uint64_t)nlist * acs / kExtSlices);
      const uint32_t s1 = (uint32_t)((uint64_t)nlist * (acs + 1) / kExtSlices);
      const uint32_t size = s1 - s0;
      if (old < size) {
        const uint32_t lo = s0 + old, hi = s0 + (old + CH < size ? old + CH : size);
        if (plo == phi) { plo = lo; phi = hi; }
        else { nlo = lo; nhi = hi; }
      }
      if (old + CH >= size) {  // this slice is drained: move on
        cs = cs + 1 == kExtSlices ? 0 : cs + 1;
        if (cs == cs0) exhausted = true;
      }
    }
    {
      const uint64_t want = __ballot(!d2_ok);
      const uint32_t nw = (uint32_t)__popcll(want);
      const uint32_t n1 = phi - plo, n2 = nhi - nlo;
      if (nw > 0 && n1 + n2 > 0) {
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(want >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)want, 0u));
        if (((want >> lane) & 1ull) && rank < n1 + n2) {
          const uint32_t idx = rank < n1 ? plo + rank : nlo + (rank - n1);
          const int4* d = (const int4*)(desc + idx);
          d2a = d[0];
          d2b = d[1];
          d2c = *(const int2*)(d + 2);
          d2_ok = true;
        }
        const uint32_t taken = nw < n1 + n2 ? nw : n1 + n2;
        if (taken <= n1) {
          plo += taken;
        } else {
          plo = nlo + (taken - n1);
          phi = nhi;
          nlo = nhi = 0;
        }
      }
      if (plo == phi && nlo < nhi) { plo = nlo; phi = nhi; nlo = nhi = 0; }
      if (!apend && !exhausted && nlo == nhi && phi - plo < CH / 2) {
        if (lane == 0) aret = atomicAdd(&fslot[cs], CH);  // every lane is active here
        apend = true;
        acs = cs;
      }
    }
    // (4) next target windows
    if (fresh && tlen > 16) load_window(A.target + toff, tlen, 16, tn, tn_sh);
    if (rot && tlen > wb + 16) load_window(A.target + toff, tlen, wb + 16, tn, tn_sh);
    // (a `continue` here would make the compiler keep a second copy of the
    // column state; rows of a wave without a task are skipped by `act`)
    if (!__any(act || nx_ok || d2_ok) && exhausted && !apend && plo == phi && nlo == nhi) break;
    // scores of the first row's target base (each row reads the next row's ahead)
    uint2 srn = trow[(t4 & 0xffu) < 8u ? (t4 & 0xffu) : 4u];

    // ================= K rows =================
#pragma unroll 1
    for (int r = 0; r < K; ++r) {
      if (act) {
        const uint2 sr = srn;
        {
          const uint32_t tn1 = r + 1 < K ? (t4 >> (8 * (r + 1))) & 0xffu : 4u;  // next row's base
          srn = trow[tn1 < 8u ? tn1 : 4u];
        }
        // band (ksw.c:413-416)
        if (beg < i - W) {  // the cap moves beg: keep the dead-column invariant
          beg = i - W;
#pragma unroll
          for (int c = 0; c < Q; ++c) {
            const int keep = ~((c - beg) >> 31);  // 0 for c < beg
            G[c] &= keep;
            E[c] &= keep;
          }
        }
        if (end > i + W + 1) end = i + W + 1;
        if (end > qlen) end = qlen;
        int h1 = 0;
        if (beg == 0) {
          h1 = h0 - (A.o_del + e_del * (i + 1));
          h1 = h1 > 0 ? h1 : 0;
        }
        int f = 0, mk = 0;
        uint32_t nz[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) nz[k] = 0;
        bool le = true;  // column c <= end (c-1 < end)
        uint32_t sc = 0;
#pragma unroll
        for (int c = 0; c < Q; ++c) {  // ksw.c:427-448
          if ((c & 3) == 0) sc = __builtin_amdgcn_perm(sr.y, sr.x, qpk[c >> 2]);
          const int s = __builtin_amdgcn_sbfe((int)sc, 8 * (c & 3), 8);
          const int g = G[c], e = E[c];
          const int M = __mul24(g + s, min(g, 1));  // g ? g + s : 0 (g >= 0)
          const int h = max(max(M, e), f);
          const bool lt = c < end;
          G[c] = le ? h1 : g;
          h1 = lt ? h : h1;
          mk = lt ? max(mk, (h << 11) | c) : mk;
          const int ne = max(max(M - oe_del, e - e_del), 0);
          E[c] = lt ? ne : (le ? 0 : e);
          f = max(max(M - oe_ins, f - e_ins), 0);
          nz[c >> 5] |= (uint32_t)min(G[c] | E[c], 1) << (c & 31);
          le = lt;
          // keep the columns in order: hoisting the work of all Q columns to the
          // row start would hold Q lane masks and temporaries live at once
          __builtin_amdgcn_sched_barrier(0);
        }
        // only columns j < end count in the zero scans
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          const int d = end - 32 * k;
          nz[k] &= d >= 32 ? 0xffffffffu : d > 0 ? (1u << d) - 1u : 0u;
        }
        // ksw.c:449-453: H(i, end-1) at the last column, gscore at the query end
        if ((end > beg ? end : beg) == qlen) {
          if (!(gbest > h1)) gbest_i = i;
          gbest = gbest > h1 ? gbest : h1;
        }
        ncell += end > beg ? (uint32_t)(end - beg) : 0u;
        const int m = end > beg ? mk >> 11 : 0;
        const int mj = mk & 2047;
        // Branch-free from here to the task-end store: a branch would let the
        // compiler sink the row's state updates into the continuing path and
        // keep every column's temporaries live across the test.
        const bool better = m > best;
        const int di = i - best_i, dj = mj - best_j;
        const int drop = di > dj ? best - m - (di - dj) * e_del : best - m - (dj - di) * e_ins;
        bool fin = m == 0 || (!better && A.zdrop > 0 && drop > A.zdrop);
        const int off = mj > i ? mj - i : i - mj;
        const bool upd = m != 0 && better;
        moff = upd ? max(moff, off) : moff;
        best = upd ? m : best;
        best_i = upd ? i : best_i;
        best_j = upd ? mj : best_j;
        {  // ksw.c:466-469 (applied only when the task goes on)
          int lo = -1, hi = -1;
#pragma unroll
          for (int k = NZ - 1; k >= 0; --k) lo = nz[k] ? 32 * k + (int)__builtin_ctz(nz[k]) : lo;
#pragma unroll
          for (int k = 0; k < NZ; ++k) hi = nz[k] ? 32 * k + 31 - (int)__builtin_clz(nz[k]) : hi;
          const int nb = lo >= 0 ? lo : end;
          const int hz = h1 != 0 ? end : hi;
          const int ne2 = hz >= 0 ? hz + 2 : nb + 1;
          beg = fin ? beg : nb;
          end = fin ? end : (ne2 < qlen ? ne2 : qlen);
          i = fin ? i : i + 1;
          fin = fin || i >= tlen;
        }
        if (fin) {  // ksw.c:473-478 (a store issued a row or more before the next sync point)
          act = false;
          int4* rp = (int4*)(A.rec + (int64_t)task * 8);
          rp[0] = make_int4(best, best_j + 1, best_i + 1, gbest_i + 1);
          rp[1] = make_int4(gbest, moff, (int)ncell, 0);
        }
      }
    }
  }
}

template <int Q, bool E1>
static int lane_grid(int num_cu) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ext_lane_kernel<Q, E1>, 64, 0) != hipSuccess || nb < 1)
      nb = lane_min_waves(Q) * 4;
    per_cu = nb;
  }
  return per_cu * num_cu;
}

template <int Q, bool E1>
void launch_lane_bin(const ExtKArgs& k, const ExtWork& wk, int bin, int num_cu, hipStream_t st, int64_t hint) {
  int64_t blocks = lane_grid<Q, E1>(num_cu);
  if (hint >= 0) blocks = std::min<int64_t>(blocks, (hint + 63) / 64);
  if (blocks < 1) return;
  time_begin(wk, bin, st);
  hipLaunchKernelGGL((ext_lane_kernel<Q, E1>), dim3((unsigned)blocks), dim3(64), 0, st, k,
                     (const ExtDesc*)wk.desc, (const uint32_t*)wk.counts, bin, wk.fetch);
  time_end(wk, bin, st);
}

#define BSX_LANE_INST(Q)                                                                                         \
  template void launch_lane_bin<Q, true>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t, int64_t);     \
  template void launch_lane_bin<Q, false>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t, int64_t);
BSX_LANE_INST(32)
#undef BSX_LANE_INST

}  // namespace bsx
