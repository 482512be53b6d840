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
      const uint32_t s0 = slice_begin(nlist, acs);
      const uint32_t s1 = slice_begin(nlist, acs + 1);
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
        aret = atomicAdd(&fslot[cs], lane == 0 ? (uint32_t)CH : 0u);  // branch-free; lane 0's value is read
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
          // Keep the columns in order and each column's results materialised at
          // its end: otherwise the compiler hoists the compares of all Q columns
          // to the row start and sinks the state updates to their last use,
          // holding Q lane masks and temporaries live at once.
          asm volatile("" ::"v"(G[c]), "v"(E[c]), "v"(h1), "v"(f), "v"(mk));
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

// ---------------------------------------------------------------------------
// Lane-per-task band-following frame run (learned-band form, classes with
// frames of WF = 24 and 40 columns, queries of at most kLaneFrameQMax bases).
//
// The same frame run as the group frame kernel: frame position k holds column
// f0 + k, f0 <= beg advances toward beg by at most two columns per row, the row
// is the reference row over the frame, and the run falls back to the plain
// kernel when the band reaches past the frame with a non-zero cell (H at the
// frame's last column, or a positive first-row insertion value past it).  One
// lane owns a task: G/E of the frame live in its VGPRs and shift down one
// position per frame step.  The task's query is copied into LDS at staging
// (global_load_lds, 16-byte aligned chunks; two buffers per lane so the next
// task's copy overlaps the current task) and each row reads the frame's WF
// codes from there.
template <int WF, bool E1>
__global__ void __launch_bounds__(64, 2)
ext_lane_frame_kernel(ExtKArgs A, const ExtDesc* __restrict__ desc_all, const uint32_t* __restrict__ counts,
                      const uint32_t* __restrict__ split, int cls, uint32_t* __restrict__ fetch,
                      uint8_t* __restrict__ need_plain) {
  constexpr int FD = WF / 4;                           // frame code dwords
  constexpr int NZ = (WF + 31) / 32;
  constexpr int NCH = (15 + kLaneFrameQMax + 15) / 16;  // 16-byte query chunks per lane and buffer
  constexpr int K = kLaneRows;
  static_assert(WF % 4 == 0 && WF <= 64, "lane frame width");
  __shared__ uint2 trow[8];
  __shared__ uint32_t qbuf[2][NCH][64][4];  // [buffer][chunk][lane][dword]
  const int lane = threadIdx.x;
  if (lane < 8) {
    const int t = lane < 5 ? lane : 4;
    const uint32_t n = A.trow_hi[t] & 0xffu;
    trow[lane] = make_uint2(A.trow_lo[t], n * 0x01010101u);
  }
  __syncthreads();

  uint32_t lbase = 0;
  for (int b = 0; b < cls; ++b) lbase += counts[kExtBins + b];
  const uint32_t nlist = split[cls];  // the class's short-query tasks lead its list
  const ExtDesc* __restrict__ desc = desc_all + lbase;
  uint32_t* __restrict__ fslot = fetch + (kExtBins + cls) * kExtSlices;
  const int cs0 = (int)(blockIdx.x % kExtSlices);
  int cs = cs0;
  bool exhausted = nlist == 0;
  constexpr uint32_t CH = 64;
  uint32_t plo = 0, phi = 0, nlo = 0, nhi = 0;
  bool apend = false;
  uint32_t aret = 0;
  int acs = 0;

  const int e_del = E1 ? 1 : A.e_del, e_ins = E1 ? 1 : A.e_ins;
  const int oe_del = A.o_del + e_del, oe_ins = A.o_ins + e_ins;

  bool d2_ok = false;
  int4 d2a = make_int4(0, 0, 0, 0), d2b = d2a;
  int2 d2c = make_int2(0, 0);
  bool nx_ok = false;
  int nx_task = 0, nx_qlen = 0, nx_tlen = 0, nx_h0 = 0, nx_w = 0, nx_tsh = 0, nx_sh16 = 0;
  uint64_t nx_toff = 0;
  uint32_t nx_t[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) nx_t[k] = 0;

  bool act = false;
  int cur = 0;  // LDS buffer of the current task's query; the next task is staged in cur ^ 1
  int task = 0, qlen = 0, tlen = 0, h0 = 0, W = 0, i = 0, ba = 0, ea = 0, f0 = 0, sh16 = 0;
  int best = 0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
  uint32_t ncell = 0;
  uint64_t toff = 0;
  int G[WF], E[WF];
  uint32_t tb[4], tn[5];
  int tn_sh = 0, wb = 0;
#pragma unroll
  for (int c = 0; c < WF; ++c) { G[c] = 0; E[c] = 0; }
#pragma unroll
  for (int k = 0; k < 4; ++k) tb[k] = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) tn[k] = 0;

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
  auto ramp_at = [&](int j) {  // eh[j].h before row 0 (ksw.c:395-397), j >= 1
    const int g = h0 - oe_ins - (j - 1) * e_ins;
    return g > 0 ? g : 0;
  };

  for (;;) {
    // ================= sync point =================
    // Everything issued at the previous sync point has landed by now (K rows
    // ago); this wait makes the LDS copies of staged queries visible to the
    // rows that follow.  (vmcnt(0); expcnt and lgkmcnt unconstrained)
    __builtin_amdgcn_s_waitcnt(0x0F70);
    if (!act && nx_ok) {
      act = true;
      nx_ok = false;
      cur ^= 1;
      task = nx_task; qlen = nx_qlen; tlen = nx_tlen; h0 = nx_h0; W = nx_w; toff = nx_toff; sh16 = nx_sh16;
#pragma unroll
      for (int k = 0; k < 4; ++k) tb[k] = bytes_at(nx_t[k], nx_t[k + 1], nx_tsh);
      wb = 0;
      {  // first frame: column k, first-row values (ksw.c:395-397)
        int g = h0 - oe_ins;
        G[0] = h0;
        E[0] = 0;
#pragma unroll
        for (int c = 1; c < WF; ++c) {
          G[c] = c <= qlen ? max(g, 0) : 0;
          E[c] = 0;
          g -= e_ins;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      i = 0; ba = 0; ea = qlen; f0 = 0;
      best = h0; best_i = best_j = -1; gbest = -1; gbest_i = -1; moff = 0;
      ncell = 0;
      if (tlen <= 0) {
        act = false;
        int4* r = (int4*)(A.rec + (int64_t)task * 8);
        r[0] = make_int4(h0, 0, 0, 0);
        r[1] = make_int4(-1, 0, 0, 0);
      }
    }
    bool rot = false;
    if (act && i - wb >= 16) {
#pragma unroll
      for (int k = 0; k < 4; ++k) tb[k] = bytes_at(tn[k], tn[k + 1], tn_sh);
      wb += 16;
      rot = true;
    }
    uint32_t t4 = tb[0];
    {
      const int q = (i - wb) >> 2;
      t4 = q == 1 ? tb[1] : t4;
      t4 = q == 2 ? tb[2] : t4;
      t4 = q == 3 ? tb[3] : t4;
    }
    const bool fresh = act && i == 0;
    // stage 2 -> stage 1: next task's fields, its first target window and the
    // copy of its query into the free LDS buffer
    bool stg = false;
    uintptr_t stg_q16 = 0;
    if (!nx_ok && d2_ok) {
      nx_ok = true;
      d2_ok = false;
      nx_task = d2a.x; nx_qlen = d2a.y; nx_tlen = d2a.z; nx_h0 = d2a.w;
      nx_w = d2b.x;
      const uint64_t qoff = (uint64_t)(uint32_t)d2b.z | ((uint64_t)(uint32_t)d2b.w << 32);
      nx_toff = (uint64_t)(uint32_t)d2c.x | ((uint64_t)(uint32_t)d2c.y << 32);
      const uint8_t* tp = nx_tlen > 0 ? A.target + nx_toff : A.query + qoff;
      load_window(tp, nx_tlen > 0 ? nx_tlen : 1, 0, nx_t, nx_tsh);
      const uintptr_t qa = (uintptr_t)(A.query + qoff);
      nx_sh16 = (int)(qa & 15);
      stg_q16 = qa & ~(uintptr_t)15;
      stg = true;
    }
    // Copy the staged queries into the free LDS buffer.  16-byte aligned chunks
    // never cross a page, and a chunk past the query's last one is replaced by
    // that one, so only pages holding query bases are read.
    if (__any(stg)) {
      const int lastch = (nx_sh16 + nx_qlen - 1) >> 4;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (stg && (cur ^ 1) == b) {
#pragma unroll
          for (int k = 0; k < NCH; ++k)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void*)(stg_q16 + 16 * (uintptr_t)min(k, lastch)),
                (__attribute__((address_space(3))) void*)&qbuf[b][k][0][0], 16, 0, 0);
        }
      }
    }
    if (apend) {
      const uint32_t old = __builtin_amdgcn_readfirstlane(aret);
      apend = false;
      const uint32_t s0 = slice_begin(nlist, acs);
      const uint32_t s1 = slice_begin(nlist, acs + 1);
      const uint32_t size = s1 - s0;
      if (old < size) {
        const uint32_t lo = s0 + old, hi = s0 + (old + CH < size ? old + CH : size);
        if (plo == phi) { plo = lo; phi = hi; }
        else { nlo = lo; nhi = hi; }
      }
      if (old + CH >= size) {
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
        aret = atomicAdd(&fslot[cs], lane == 0 ? (uint32_t)CH : 0u);
        apend = true;
        acs = cs;
      }
    }
    if (fresh && tlen > 16) load_window(A.target + toff, tlen, 16, tn, tn_sh);
    if (rot && tlen > wb + 16) load_window(A.target + toff, tlen, wb + 16, tn, tn_sh);
    if (!__any(act || nx_ok || d2_ok) && exhausted && !apend && plo == phi && nlo == nhi) break;

    uint2 srn = trow[(t4 & 0xffu) < 8u ? (t4 & 0xffu) : 4u];
#pragma unroll 1
    for (int r = 0; r < K; ++r) {
      if (act) {
        const uint2 sr = srn;
        {
          const uint32_t tn1 = r + 1 < K ? (t4 >> (8 * (r + 1))) & 0xffu : 4u;
          srn = trow[tn1 < 8u ? tn1 : 4u];
        }
        // the frame's query codes (bytes sh16 + f0 .. of the lane's LDS copy)
        uint32_t cw[FD];
        {
          const int p = sh16 + f0;
          const int d0 = p >> 2, o = p & 3;
          uint32_t w[FD + 1];
#pragma unroll
          for (int m = 0; m <= FD; ++m) {
            const int d = min(d0 + m, NCH * 4 - 1);
            w[m] = qbuf[cur][d >> 2][lane][d & 3];
          }
#pragma unroll
          for (int m = 0; m < FD; ++m) cw[m] = bytes_at(w[m], w[m + 1], o);
        }
        // the reference's band of this row (ksw.c:413-416)
        if (ba < i - W) {
          ba = i - W;
#pragma unroll
          for (int c = 0; c < WF; ++c) {
            const int keep = ~((f0 + c - ba) >> 31);  // 0 for columns < ba
            G[c] &= keep;
            E[c] &= keep;
          }
        }
        if (ea > i + W + 1) ea = i + W + 1;
        if (ea > qlen) ea = qlen;
        const int bN = ba, eN = ea;
        const int eF = min(eN, f0 + WF - 1);  // band end computed in the frame
        const int kN = eF - f0;
        const bool spill = eN > eF;           // cells past the frame must be 0
        bool fail = spill && ramp_at(eF + 1) > 0;
        int h1p = 0;
        if (bN == 0) {
          h1p = h0 - (A.o_del + e_del * (i + 1));
          h1p = h1p > 0 ? h1p : 0;
        }
        int h1 = h1p, f = 0, mk = 0, hlast = 0;
        uint32_t nz[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) nz[k] = 0;
        bool le = true;
        uint32_t sc = 0;
#pragma unroll
        for (int c = 0; c < WF; ++c) {  // ksw.c:427-448 over the frame
          if ((c & 3) == 0) sc = __builtin_amdgcn_perm(sr.y, sr.x, cw[c >> 2]);
          const int s = __builtin_amdgcn_sbfe((int)sc, 8 * (c & 3), 8);
          const int g = G[c], e = E[c];
          const int M = __mul24(g + s, min(g, 1));
          const int h = max(max(M, e), f);
          const bool lt = c < kN;
          G[c] = le ? h1 : g;
          h1 = lt ? h : h1;
          mk = lt ? max(mk, (h << 11) | c) : mk;
          const int ne = max(max(M - oe_del, e - e_del), 0);
          E[c] = lt ? ne : (le ? 0 : e);
          f = max(max(M - oe_ins, f - e_ins), 0);
          nz[c >> 5] |= (uint32_t)min(G[c] | E[c], 1) << (c & 31);
          if (c == WF - 1) hlast = h;
          le = lt;
          asm volatile("" ::"v"(G[c]), "v"(E[c]), "v"(h1), "v"(f), "v"(mk));  // as in the lane kernel
          __builtin_amdgcn_sched_barrier(0);
        }
        fail = fail || (spill && hlast != 0);  // a non-zero cell past the frame
        // zero scans cover positions <= kN (end scan) and < kN, or <= kN when
        // the band spills (beg scan)
        uint32_t nzh[NZ], nzb[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
          const int dh = kN + 1 - 32 * k, db = (spill ? kN + 1 : kN) - 32 * k;
          nzh[k] = nz[k] & (dh >= 32 ? 0xffffffffu : dh > 0 ? (1u << dh) - 1u : 0u);
          nzb[k] = nz[k] & (db >= 32 ? 0xffffffffu : db > 0 ? (1u << db) - 1u : 0u);
        }
        const int h1e = eN > bN ? (spill ? 0 : h1) : h1p;  // H(i, end-1)
        if ((eN > bN ? eN : bN) == qlen) {                 // ksw.c:449-453
          if (!(gbest > h1e)) gbest_i = i;
          gbest = gbest > h1e ? gbest : h1e;
        }
        ncell += eN > bN ? (uint32_t)(eN - bN) : 0u;
        const int m = eN > bN ? mk >> 11 : 0;
        const int mj = (mk & 2047) + f0;
        const bool better = m > best;
        const int di = i - best_i, dj = mj - best_j;
        const int drop = di > dj ? best - m - (di - dj) * e_del : best - m - (dj - di) * e_ins;
        bool fin = m == 0 || (!better && A.zdrop > 0 && drop > A.zdrop);
        const bool upd = !fail && m != 0 && better;
        const int off = mj > i ? mj - i : i - mj;
        moff = upd ? max(moff, off) : moff;
        best = upd ? m : best;
        best_i = upd ? i : best_i;
        best_j = upd ? mj : best_j;
        int s = 0;
        {  // ksw.c:466-469
          int lo = -1, hi = -1;
#pragma unroll
          for (int k = NZ - 1; k >= 0; --k) lo = nzb[k] ? 32 * k + (int)__builtin_ctz(nzb[k]) : lo;
#pragma unroll
          for (int k = 0; k < NZ; ++k) hi = nzh[k] ? 32 * k + 31 - (int)__builtin_clz(nzh[k]) : hi;
          const int nb = lo >= 0 ? lo + f0 : eN;
          const int hz = hi >= 0 ? hi + f0 : -1;
          const int ne2 = hz >= nb ? hz + 2 : nb + 1;
          const bool go = !fail && !fin;
          ba = go ? nb : ba;
          ea = go ? (ne2 < qlen ? ne2 : qlen) : ea;
          i = go ? i + 1 : i;
          fin = fin || (go && i >= tlen);
          s = go ? min(ba - f0, 2) : 0;
        }
        // move the frame toward beg; the entering column holds its first-row value
        if (s >= 1) {
          ++f0;
#pragma unroll
          for (int c = 0; c < WF - 1; ++c) { G[c] = G[c + 1]; E[c] = E[c + 1]; }
          const int jin = f0 + WF - 1;
          G[WF - 1] = jin <= qlen ? ramp_at(jin) : 0;
          E[WF - 1] = 0;
        }
        if (s >= 2) {
          ++f0;
#pragma unroll
          for (int c = 0; c < WF - 1; ++c) { G[c] = G[c + 1]; E[c] = E[c + 1]; }
          const int jin = f0 + WF - 1;
          G[WF - 1] = jin <= qlen ? ramp_at(jin) : 0;
          E[WF - 1] = 0;
        }
        if (fail) {
          act = false;
          need_plain[task] = 1;
        } else if (fin) {
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

template <int WF, bool E1>
void launch_lane_frame(const ExtKArgs& k, const ExtWork& wk, int cls, int num_cu, hipStream_t st) {
  static int per_cu = 0;
  if (per_cu == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ext_lane_frame_kernel<WF, E1>, 64, 0) != hipSuccess ||
        nb < 1)
      nb = 4;
    per_cu = nb;
  }
  hipLaunchKernelGGL((ext_lane_frame_kernel<WF, E1>), dim3((unsigned)(per_cu * num_cu)), dim3(64), 0, st, k,
                     (const ExtDesc*)wk.desc, (const uint32_t*)wk.counts, (const uint32_t*)wk.split, cls, wk.fetch,
                     wk.need_plain);
}
template void launch_lane_frame<24, true>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t);
template void launch_lane_frame<24, false>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t);
template void launch_lane_frame<40, true>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t);
template void launch_lane_frame<40, false>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t);

#define BSX_LANE_INST(Q)                                                                                         \
  template void launch_lane_bin<Q, true>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t, int64_t);     \
  template void launch_lane_bin<Q, false>(const ExtKArgs&, const ExtWork&, int, int, hipStream_t, int64_t);
BSX_LANE_INST(32)
BSX_LANE_INST(48)
BSX_LANE_INST(64)
#undef BSX_LANE_INST

}  // namespace bsx
