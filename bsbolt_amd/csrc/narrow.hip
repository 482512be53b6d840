// Band-following frame run of ksw_extend2 (learned-band form).
//
// The reference computes, per row, only the columns of its adaptive band
// [beg, end) (ksw.c:412-469); for short-lived and drifting extensions that band
// is far narrower than the query.  This kernel gives a task a frame of WF = L*C
// consecutive columns [f0, f0 + WF) that follows the band: f0 <= beg always (f0
// advances toward beg by at most two columns per row), and the row is computed
// exactly as the plain kernel computes it, over the frame.  The run IS the
// reference run -- same recurrences, same beg/end, same stale state -- as long
// as no non-zero cell falls past the frame:
//   when the band end reaches past the frame (end > f0 + WF - 1 = eF), the
//   cells in [eF, end) must all be 0.  That holds, and is checked, when the
//   frame's last column eF computes H(i, eF) = 0 and the columns past it hold
//   zero state: their first-row insertion value max(h0 - oe_ins - (j-1)*e_ins, 0)
//   is 0 at j = eF + 1 (ksw.c:395-397), and any earlier row that reached past
//   the frame wrote only zeros there (this same check).
// A run that fails the check is flagged for the plain kernel; every accepted
// result equals the reference's in all six outputs and the cell count.
//
// Layout: frame position k = g*C + c (lane g, column c) holds column f0 + k.
// G (= H(i-1, j-1)) and E stay in place during a row; when the frame advances,
// G and E shift one position left and the entering column takes its first-row
// value (it was never computed, or computed as 0 where that value is 0).  Dead
// columns left of beg hold G = E = 0, so they need no masks.  F keys are taken
// relative to the frame ((k+1)*e_ins); the offset is common to the row and
// cancels.  Raw query codes are staged per task in LDS; each row selects its
// target base's score row (mat[t*5+q]) and extracts per-column scores with
// v_perm_b32.
#include "bsx_internal.h"
#include "ext_common.h"

namespace bsx {

// group shapes of the frames 24, 40, 80, 144 (lanes x columns per lane; fewer
// lanes with more columns each measured slower for the 40- and 80-column frames)
__host__ __device__ constexpr int frame_lanes(int c) { return c <= 1 ? 8 : 16; }
__host__ __device__ constexpr int frame_cols(int c) { return c == 0 ? 3 : c == 1 ? 5 : c == 2 ? 5 : 9; }
__host__ __device__ constexpr int frame_of_class(int c) { return frame_lanes(c) * frame_cols(c); }

// Measured cost of one DP row per task on MI355X (picoseconds of kernel time
// per row over a 150 bp paired-end bisulfite batch, each kernel run alone):
// the plain form's bins and the frame classes (lane-per-task frames of 24 and
// 40 columns -- upper bounds, from the rows of the runs that completed --
// and group frames of 80 and 144 columns).  A task goes to its predicted frame
// class only where that class is at least 20% cheaper per row than the plain
// bin of its query length.
constexpr int kPlainRowPs[kExtBins] = {42, 43, 57, 59, 109, 158, 600, 700, 1000, 2000, 4000};
constexpr int kFrameRowPs[kBandClasses - 1] = {29, 75, 152, 239};
static_assert(kExtBins == 11 && kBandClasses == 5, "row cost tables");
__host__ __device__ constexpr bool narrow_pays(int c, int b) {
  return 10 * kFrameRowPs[c] < 8 * kPlainRowPs[b];
}

__device__ __forceinline__ int ramp_at(int j, int h0, int oe_ins, int e_ins) {  // eh[j].h before row 0
  const int g = j == 0 ? h0 : h0 - oe_ins - (j - 1) * e_ins;
  return g > 0 ? g : 0;
}

// every lane of the group receives the bitwise OR
template <int L>
__device__ __forceinline__ int grp_or(int x, int lane) {
  if (L <= 16) {
    x |= __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, true);
    x |= __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, true);
    if (L == 4) return x;
    x |= __builtin_amdgcn_update_dpp(0, x, 0x141, 0xf, 0xf, true);
    if (L == 16) x |= __builtin_amdgcn_update_dpp(0, x, 0x140, 0xf, 0xf, true);
    return x;
  }
  x |= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
  if (L >= 64) x |= __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
  if (L == 64) return __builtin_amdgcn_readlane(x, 63);
  const int a = __builtin_amdgcn_readlane(x, 31), b = __builtin_amdgcn_readlane(x, 63);
  return lane < 32 ? a : b;
}

template <int L, int C, bool E1>
__global__ void __launch_bounds__(kExtBlock)
ext_narrow_kernel(ExtKArgs A, const ExtDesc* __restrict__ desc_all, const uint32_t* __restrict__ counts,
                  const uint32_t* __restrict__ split, int cls, uint8_t* __restrict__ need_plain) {
  constexpr int WF = L * C;
  constexpr int NG = kExtBlock / L;
  constexpr int QD = (kNarrowQMax + 6) / 4;  // aligned dwords that can cover a query
  constexpr int QR = (QD + L - 1) / L;       // query dword loads per lane
  // Query bytes of a group: byte sh + j holds query[j] (sh = the query's offset
  // in its first aligned dword); rows read bytes sh + f0 + k <= 3 + qlen + WF - 1.
  constexpr int SZW_STAGE = L * QR;
  constexpr int SZW_READ = (WF + kNarrowQMax + 3) / 4 + 1;
  constexpr int SZW = SZW_STAGE > SZW_READ ? SZW_STAGE : SZW_READ;  // dwords per group
  __shared__ uint32_t q_lds[NG * SZW];
  __shared__ uint32_t lds_next;
  const int lane = threadIdx.x & 63;
  const int lig = lane & (L - 1);
  const int base = lig * C;
  uint32_t* __restrict__ Q32 = q_lds + (threadIdx.x / L) * SZW;
  const uint8_t* __restrict__ Q8 = (const uint8_t*)Q32;
  uint32_t lbase = 0;
  for (int b = 0; b < cls; ++b) lbase += counts[kExtBins + b];
  const uint32_t nlist = counts[kExtBins + cls];
  // the short-query tasks that lead a lane-frame class's list run in the lane kernel
  const uint32_t first = cls < kLaneFrameClasses ? split[cls] : 0u;
  const uint32_t blk0 = first + blockIdx.x * (uint32_t)A.tpb;
  if (blk0 >= nlist) return;  // uniform over the block
  const uint32_t blk1 = min(blk0 + (uint32_t)A.tpb, nlist);
  if (threadIdx.x == 0) lds_next = blk0;
  __syncthreads();
  const ExtDesc* __restrict__ desc = desc_all + lbase;
  const int e_del = E1 ? 1 : A.e_del, e_ins = E1 ? 1 : A.e_ins;
  const int oe_del = A.o_del + e_del, oe_ins = A.o_ins + e_ins;

  auto next_pos = [&]() -> uint32_t {
    // Branch-free fetch: every lane issues the atomic and only lane 0 of the
    // group adds 1.  No lane-divergent region may precede the broadcast: with
    // `if (lig == 0)` around the atomic the compiler can give the other lanes a
    // loop of their own that reaches the broadcast without lane 0's new value.
    const uint32_t mine = atomicAdd(&lds_next, lig == 0 ? 1u : 0u);
    const uint32_t p = __shfl(mine, 0, L);
    return p < blk1 ? p : blk1;
  };
  // ---- prefetch: descriptor two tasks ahead, sequence words of the next task ----
  bool d2_ok = false;
  int4 d2a = make_int4(0, 0, 0, 0), d2b = d2a, d2c = d2a;
  auto load_desc = [&]() {
    const uint32_t p = next_pos();
    d2_ok = p < blk1;
    const int4* d = (const int4*)(desc + (d2_ok ? p : blk0));
    d2a = d[0]; d2b = d[1]; d2c = d[2];
  };
  bool nx_ok = false;
  int nx_task = 0, nx_qlen = 0, nx_tlen = 0, nx_h0 = 0, nx_w = 0, nx_sh = 0;
  uint64_t nx_qoff = 0, nx_toff = 0;
  uint32_t nx_qd[QR];  // aligned query dwords lig + L*r (clamped to the query's last dword)
  uint32_t nx_tb[4];   // target bytes [4*lig, 4*lig+4) (clamped)
  auto stage1 = [&]() {
    nx_ok = d2_ok;
    nx_task = d2a.x; nx_qlen = d2a.y; nx_tlen = d2a.z; nx_h0 = d2a.w;
    nx_w = d2b.x;
    nx_qoff = (uint64_t)(uint32_t)d2b.z | ((uint64_t)(uint32_t)d2b.w << 32);
    nx_toff = (uint64_t)(uint32_t)d2c.x | ((uint64_t)(uint32_t)d2c.y << 32);
    // aligned dwords never cross a page, so the bytes read around the query are mapped
    const uintptr_t qa = (uintptr_t)(A.query + nx_qoff);
    const uint32_t* qp = (const uint32_t*)(qa & ~(uintptr_t)3);
    nx_sh = (int)(qa & 3);
    const int nd = (nx_sh + (nx_qlen > 0 ? nx_qlen : 1) + 3) >> 2;
#pragma unroll
    for (int r = 0; r < QR; ++r) nx_qd[r] = qp[min(lig + L * r, nd - 1)];
    const uint8_t* tp = nx_tlen > 0 ? A.target + nx_toff : A.query + nx_qoff;
    const int tmax = nx_tlen > 0 ? nx_tlen - 1 : 0;
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) nx_tb[k2] = tp[min(4 * lig + k2, tmax)];
  };
  auto load_window = [&](const uint8_t* tp, int tlen_, int first, uint32_t* out) {
    const int tmax = tlen_ > 0 ? tlen_ - 1 : 0;
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) out[k2] = tp[min(first + 4 * lig + k2, tmax)];
  };

  int G[C], E[C];
#pragma unroll
  for (int c = 0; c < C; ++c) { G[c] = 0; E[c] = 0; }
#pragma unroll
  for (int r = 0; r < QR; ++r) nx_qd[r] = 0;
  int task = 0, qlen = 0, tlen = 0, h0 = 0, W = 0, i = 0, ba = 0, ea = 0, f0 = 0;
  int best = 0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
  uint32_t ncell = 0, tch = 0;
  uint32_t tnx[4] = {0, 0, 0, 0};
  int sh = 0;
  const uint8_t* tbase = A.target;
  int tcb = 0;
  bool need = true, done = false;

  load_desc();
  stage1();
  load_desc();
  for (;;) {
    if (need && !done) {
      if (!nx_ok) {
        done = true;
      } else {
        task = nx_task; qlen = nx_qlen; tlen = nx_tlen; h0 = nx_h0; W = nx_w;
        // raw query dwords; bytes outside the query land only in columns that are never live
#pragma unroll
        for (int r = 0; r < QR; ++r) Q32[lig + L * r] = nx_qd[r];
        sh = nx_sh;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int c = 0; c < C; ++c) {  // first frame: column j = k, first-row values (ksw.c:395-397)
          const int j = base + c;
          G[c] = j <= qlen ? ramp_at(j, h0, oe_ins, e_ins) : 0;
          E[c] = 0;
        }
        {
          uint32_t v = 0;
#pragma unroll
          for (int k2 = 0; k2 < 4; ++k2) v |= (4 * lig + k2 < tlen ? (nx_tb[k2] & 0xffu) : 0u) << (8 * k2);
          tch = v;
        }
        tcb = 0;
        i = 0; ba = 0; ea = qlen; f0 = 0;
        best = h0; best_i = best_j = -1; gbest = -1; gbest_i = -1; moff = 0;
        ncell = 0;
        tbase = tlen > 0 ? A.target + nx_toff : A.query + nx_qoff;
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

    // ---------------- one row of the narrow run ----------------
    if (i - tcb >= 4 * L) {
      tcb = i;
      uint32_t v = 0;
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v |= (i + 4 * lig + k2 < tlen ? (tnx[k2] & 0xffu) : 0u) << (8 * k2);
      tch = v;
      load_window(tbase, tlen, i + 4 * L, tnx);
    }
    const int tr = i - tcb;
    const int t = (int)((__shfl(tch, tr >> 2, L) >> (8 * (tr & 3))) & 0xffu);
    // scores of this row's target base against query codes 0..4 (mat[t*5+q])
    uint32_t s_lo = A.trow_lo[4], s_hi = A.trow_hi[4];
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
      s_lo = t == k2 ? A.trow_lo[k2] : s_lo;
      s_hi = t == k2 ? A.trow_hi[k2] : s_hi;
    }

    // the reference's band of this row (ksw.c:413-416)
    if (ba < i - W) {  // the band cap moves beg: zero the frame columns it passes (dead columns)
      ba = i - W;
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (f0 + base + c < ba) { G[c] = 0; E[c] = 0; }
    }
    if (ea > i + W + 1) ea = i + W + 1;
    if (ea > qlen) ea = qlen;
    const int bN = ba, eN = ea;
    const int eF = min(eN, f0 + WF - 1);  // band end computed in the frame
    const int kN = eF - f0;               // frame position of eF
    const bool spill = eN > eF;           // the band reaches past the frame: those cells must be 0
    bool fail = spill && ramp_at(eF + 1, h0, oe_ins, e_ins) > 0;
    int h1p = 0;
    if (bN == 0) {
      h1p = h0 - (A.o_del + e_del * (i + 1));
      h1p = h1p > 0 ? h1p : 0;
    }

    int M[C], ex[C];
    const uint8_t* __restrict__ Qr = Q8 + sh + f0;  // f0 <= beg <= qlen
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const uint32_t qc = Qr[base + c];  // query code of column f0 + k
      M[c] = (int)(int8_t)(__builtin_amdgcn_perm(s_hi, s_lo, qc) & 0xffu);
    }
    // frame positions k < kN are live; the other band masks derive from these compares
    const int dN = kN - base;
    bool lt[C + 1];
#pragma unroll
    for (int c = 0; c <= C; ++c) lt[c] = c < dN;
    const bool le0 = dN >= 0;
    int run = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int g = G[c];
      const int mm = g ? g + M[c] : 0;
      M[c] = mm;
      const int kb = (base + c + 1) * e_ins;
      const int key = max(mm + (kb - oe_ins), kb);
      ex[c] = run;
      run = max(run, key);
    }
    const int lex = grp_excl_from_incl<L>(grp_incl_max<L>(run), lig);

    int h[C];
    int mk = 0, hf = 0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = base + c;
      const int x = max(ex[c], lex);
      const int f = x - k * e_ins;
      const int hh = max(max(M[c], E[c]), f);
      h[c] = hh;
      const int ne = max(max(E[c] - e_del, M[c] - oe_del), 0);
      mk = lt[c] ? max(mk, (hh << 11) | k) : mk;
      E[c] = lt[c] ? ne : E[c];
      hf = (lt[c] && !lt[c + 1]) ? (hf | (hh << 1)) : hf;  // k == kN - 1: H(i, eF - 1)
      const bool le = c == 0 ? le0 : lt[c - 1];
      hf = (le && !lt[c] && hh != 0) ? (hf | 1) : hf;  // k == kN: H(i, eF) (must be 0 when the band spills)
    }
    const int hsh = __builtin_amdgcn_update_dpp(0, h[C - 1], L == 16 ? 0x111 : 0x138, 0xf, 0xf, true);
    uint32_t nzb = 0;  // bit c: position c holds a non-zero G or E after the row
#pragma unroll
    for (int c = 0; c < C; ++c) {
      int left = c == 0 ? (lig == 0 ? 0 : hsh) : h[c - 1];
      left = (f0 == 0 && base + c == 0) ? h1p : left;  // column 0 takes the first-column value
      const bool le = c == 0 ? le0 : lt[c - 1];
      G[c] = le ? left : G[c];
      E[c] = (le && !lt[c]) ? 0 : E[c];
      nzb |= (uint32_t)min(G[c] | E[c], 1) << c;
    }
    // zero scans (ksw.c:466-469): the end scan covers positions <= kN (position kN holds
    // H(i, eF-1) in G); the beg scan stops before the band end, which is kN unless the band spills
    auto upto = [&](int d) -> uint32_t { return d >= C ? 0xffffffffu : d > 0 ? (1u << d) - 1u : 0u; };
    const uint32_t nz_hz = nzb & upto(dN + 1);
    const uint32_t nz_nb = nzb & upto(spill ? dN + 1 : dN);
    int hik = 0, lok = 0;
    if (nz_hz) hik = base + 32 - __builtin_clz(nz_hz);
    if (nz_nb) lok = 2048 - base - __builtin_ctz(nz_nb);
    mk = grp_max<L>(mk, lane);
    hf = grp_or<L>(hf, lane);
    hik = grp_max<L>(hik, lane);
    lok = grp_max<L>(lok, lane);
    if (spill && (hf & 1)) fail = true;  // a non-zero cell past the frame

    // end-to-end column and cell count (ksw.c:449-453, 424)
    const int h1 = eN > bN ? (spill ? 0 : (hf >> 1)) : h1p;
    const int jafter = eN > bN ? eN : bN;
    if (jafter == qlen) {
      if (!(gbest > h1)) gbest_i = i;
      gbest = gbest > h1 ? gbest : h1;
    }
    ncell += eN > bN ? (uint32_t)(eN - bN) : 0u;
    const int m = eN > bN ? mk >> 11 : 0;
    const int mj = (mk & 2047) + f0;
    bool fin = false;
    if (!fail) {
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
        const int nb = lok > 0 ? 2048 - lok + f0 : eN;
        const int hz = hik > 0 ? hik - 1 + f0 : -1;
        ba = nb;
        const int ne = hz >= nb ? hz + 2 : nb + 1;
        ea = ne < qlen ? ne : qlen;
        ++i;
        if (i >= tlen) fin = true;
        // move the frame toward beg, at most two columns per row; an entering column has its
        // first-row value (it was never computed, or computed as 0 where that value is 0)
        const int s = min(ba - f0, 2);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          if (r < s) {
            ++f0;
            const int jin = f0 + WF - 1;
            const int gin = jin <= qlen ? ramp_at(jin, h0, oe_ins, e_ins) : 0;
            const int gnx = __builtin_amdgcn_update_dpp(0, G[0], L == 16 ? 0x101 : 0x130, 0xf, 0xf, true);
            const int enx = __builtin_amdgcn_update_dpp(0, E[0], L == 16 ? 0x101 : 0x130, 0xf, 0xf, true);
#pragma unroll
            for (int c = 0; c < C - 1; ++c) { G[c] = G[c + 1]; E[c] = E[c + 1]; }
            G[C - 1] = lig == L - 1 ? gin : gnx;
            E[C - 1] = lig == L - 1 ? 0 : enx;
          }
        }
      }
    }
    if (fail) {
      if (lig == 0) need_plain[task] = 1;
      need = true;
    } else if (fin) {
      write_rec<L>(A.rec, task, lig, best, best_j + 1, best_i + 1, gbest_i + 1, gbest, moff, ncell);  // ksw.c:473-478
      need = true;
    }
  }
}

// ---- routing of the learned-band form ----

__device__ __forceinline__ bool narrow_pays_dev(int c, int b) {
  bool r = false;
#pragma unroll
  for (int cc = 0; cc < kBandClasses - 1; ++cc)
#pragma unroll
    for (int q = 0; q < kExtBins; ++q) r = (cc == c && q == b) ? narrow_pays(cc, q) : r;
  return r;
}

// Frame class of every valid task (0xff: plain only) and the class counts.  A
// task goes to a frame when its class is a frame class, its query fits the
// frame kernels and (unless any_gain) the class's row cost is below the plain
// kernel's for its query length.
__global__ void __launch_bounds__(256)
ext_narrow_count_kernel(ExtArgs A, ExtWork W, BandPlan plan) {
  constexpr int NF = kLaneFrameClasses * kRowBuckets;
  __shared__ uint32_t hist[kBandClasses], front[kLaneFrameClasses], fhist[NF];
  if (threadIdx.x < kBandClasses) hist[threadIdx.x] = 0;
  if (threadIdx.x < kLaneFrameClasses) front[threadIdx.x] = 0;
  if (threadIdx.x < NF) fhist[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = W.bin_of[t];
    if (b == 0xff) continue;
    int c = plan.mode == kBandFixed ? plan.fixed : (int)plan.cls[t];
    // a predicted 40-column class whose frame does not pay for this query
    // length tries the 24-column frame (a failed frame run costs a few rows)
    if (plan.mode == kBandLearned && c == 1 && !narrow_pays_dev(1, b)) c = 0;
    int nb = 0xff;
    if (c >= 0 && c < kBandClasses - 1 && A.qlen[t] <= kNarrowQMax && (plan.any_gain || narrow_pays_dev(c, b)))
      nb = c;
    W.nbin_of[t] = (uint8_t)nb;
    W.need_plain[t] = nb == 0xff ? 1 : 0;
    if (nb != 0xff) atomicAdd(&hist[nb], 1u);
    if (nb < kLaneFrameClasses && A.qlen[t] <= kLaneFrameQMax) {
      atomicAdd(&front[nb], 1u);
      atomicAdd(&fhist[nb * kRowBuckets + row_bucket(A.tlen[t], A.h0[t])], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < kBandClasses && hist[threadIdx.x])
    atomicAdd(&W.counts[kExtBins + threadIdx.x], hist[threadIdx.x]);
  if (threadIdx.x < kLaneFrameClasses && front[threadIdx.x]) atomicAdd(&W.split[threadIdx.x], front[threadIdx.x]);
  if (threadIdx.x < NF && fhist[threadIdx.x]) atomicAdd(&W.fcount[threadIdx.x], fhist[threadIdx.x]);
}

__global__ void __launch_bounds__(256)
ext_narrow_scatter_kernel(ExtArgs A, ExtWork W, int64_t per_block) {
  // A lane-frame class's list holds its short-query tasks at the front and
  // the others at the back (back cursor, counting down from the end); the
  // other classes fill their lists from the front.  The front is the lane
  // frame kernel's list: longest run first, in row buckets, each rank placed
  // in its fetch slice as in the lane-kernel bins.
  constexpr int NF = kLaneFrameClasses * kRowBuckets;
  __shared__ uint32_t off[kBandClasses], hist[kBandClasses], cur[kBandClasses];
  __shared__ uint32_t boff[kLaneFrameClasses], bhist[kLaneFrameClasses], bcur[kLaneFrameClasses];
  __shared__ uint32_t foff[NF], fhist[NF], fcur[NF], fbase[kLaneFrameClasses], fsize[kLaneFrameClasses];
  const int64_t t0 = (int64_t)blockIdx.x * per_block;
  const int64_t t1 = min(A.n, t0 + per_block);
  if (threadIdx.x < kBandClasses) { hist[threadIdx.x] = 0; cur[threadIdx.x] = 0; }
  if (threadIdx.x < kLaneFrameClasses) { bhist[threadIdx.x] = 0; bcur[threadIdx.x] = 0; }
  if (threadIdx.x < NF) { fhist[threadIdx.x] = 0; fcur[threadIdx.x] = 0; }
  __syncthreads();
  auto front_of = [&](int64_t t, int nb) { return nb < kLaneFrameClasses && A.qlen[t] <= kLaneFrameQMax; };
  auto back_of = [&](int64_t t, int nb) { return nb < kLaneFrameClasses && A.qlen[t] > kLaneFrameQMax; };
  auto member = [&](int64_t t, int& nb) {
    if (W.bin_of[t] == 0xff) return false;
    nb = W.nbin_of[t];
    return nb != 0xff;
  };
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    int nb;
    if (!member(t, nb)) continue;
    if (back_of(t, nb)) atomicAdd(&bhist[nb], 1u);
    else if (front_of(t, nb)) atomicAdd(&fhist[nb * kRowBuckets + row_bucket(A.tlen[t], A.h0[t])], 1u);
    else atomicAdd(&hist[nb], 1u);
  }
  __syncthreads();
  if (threadIdx.x < NF) {
    const int c = threadIdx.x / kRowBuckets, k = threadIdx.x % kRowBuckets;
    uint32_t s = 0;
    for (int j = 0; j < k; ++j) s += W.fcount[c * kRowBuckets + j];
    const uint32_t hc = fhist[threadIdx.x];
    foff[threadIdx.x] = s + (hc ? atomicAdd(&W.fcursor[threadIdx.x], hc) : 0u);
  }
  if (threadIdx.x < kBandClasses) {
    uint32_t s = 0;
    for (int b = 0; b < (int)threadIdx.x; ++b) s += W.counts[kExtBins + b];
    const uint32_t hc = hist[threadIdx.x];
    off[threadIdx.x] = s + (hc ? atomicAdd(&W.cursor[kExtBins + threadIdx.x], hc) : 0u);
    if (threadIdx.x < kLaneFrameClasses) {  // last slot of the class's list minus the back cursor
      const uint32_t bc = bhist[threadIdx.x];
      boff[threadIdx.x] = s + W.counts[kExtBins + threadIdx.x] - 1u -
                          (bc ? atomicAdd(&W.split[kLaneFrameClasses + threadIdx.x], bc) : 0u);
      fbase[threadIdx.x] = s;
      fsize[threadIdx.x] = W.split[threadIdx.x];
    }
  }
  __syncthreads();
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    int nb;
    if (!member(t, nb)) continue;
    uint32_t p;
    if (back_of(t, nb)) {
      p = boff[nb] - atomicAdd(&bcur[nb], 1u);
    } else if (front_of(t, nb)) {
      const int kb = nb * kRowBuckets + row_bucket(A.tlen[t], A.h0[t]);
      const uint32_t r = foff[kb] + atomicAdd(&fcur[kb], 1u);
      p = fbase[nb] + slice_begin(fsize[nb], (int)(r % kExtSlices)) + r / kExtSlices;
    } else {
      p = off[nb] + atomicAdd(&cur[nb], 1u);
    }
    write_desc(A, t, W.desc + p);
  }
}

template <int CL, bool E1>
static void launch_narrow_class(const ExtKArgs& k, const ExtWork& wk, int64_t n, int num_cu, hipStream_t st) {
  constexpr int L = frame_lanes(CL), C = frame_cols(CL);
  const int64_t blocks = (n + k.tpb - 1) / k.tpb;
  if (blocks < 1) return;
  // classes on separate streams; a lane-frame class's lane kernel beside its group kernel
  const hipStream_t sg = pick_stream(wk.ss, st, CL == 3 ? 0 : CL + 1);
  time_begin(wk, kExtBins + CL, sg);
  if constexpr (CL < kLaneFrameClasses) launch_lane_frame<frame_of_class(CL), E1>(k, wk, CL, num_cu, sg);
  hipLaunchKernelGGL((ext_narrow_kernel<L, C, E1>), dim3((unsigned)blocks), dim3(kExtBlock), 0, sg, k,
                     (const ExtDesc*)wk.desc, (const uint32_t*)wk.counts, (const uint32_t*)wk.split, CL,
                     wk.need_plain);
  time_end(wk, kExtBins + CL, sg);
}

template <bool E1>
static void launch_narrow_classes(const ExtKArgs& k, const ExtWork& wk, int64_t n, int num_cu, hipStream_t st,
                                  const BandPlan& plan) {
  if (plan.mode == kBandFixed) {  // one class only
    switch (plan.fixed) {
      case 0: launch_narrow_class<0, E1>(k, wk, n, num_cu, st); break;
      case 1: launch_narrow_class<1, E1>(k, wk, n, num_cu, st); break;
      case 2: launch_narrow_class<2, E1>(k, wk, n, num_cu, st); break;
      case 3: launch_narrow_class<3, E1>(k, wk, n, num_cu, st); break;
      default: break;
    }
    return;
  }
  launch_narrow_class<3, E1>(k, wk, n, num_cu, st);
  launch_narrow_class<2, E1>(k, wk, n, num_cu, st);
  launch_narrow_class<1, E1>(k, wk, n, num_cu, st);
  launch_narrow_class<0, E1>(k, wk, n, num_cu, st);
}

hipError_t launch_narrow(const ExtArgs& a, const ExtKArgs& k, const ExtWork& wk, int num_cu, hipStream_t st,
                         const BandPlan& plan) {
  // count and scatter: each block ends with global atomics per class and row bucket
  int64_t blocks = (a.n + 255) / 256;
  if (blocks > (int64_t)num_cu * 4) blocks = (int64_t)num_cu * 4;
  hipLaunchKernelGGL(ext_narrow_count_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, wk, plan);
  const int64_t per_block = (a.n + blocks - 1) / blocks;
  hipLaunchKernelGGL(ext_narrow_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, wk, per_block);
  ss_fork(wk.ss, st);
  if (a.e_del == 1 && a.e_ins == 1) launch_narrow_classes<true>(k, wk, a.n, num_cu, st, plan);
  else launch_narrow_classes<false>(k, wk, a.n, num_cu, st, plan);
  ss_join(wk.ss, st);
  return hipGetLastError();
}

// ---- labels and statistics ----

// After a fixed-class run of class c: a task whose frame run completed gets
// label c.  Classes are run from 3 down to 0, so the label ends as the
// smallest frame that holds the task's band; kBandClasses-1 = none.
__global__ void __launch_bounds__(256)
band_label_merge_kernel(ExtArgs A, ExtWork W, int c, uint8_t* __restrict__ labels, int first) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < A.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (W.bin_of[t] == 0xff) {
      labels[t] = 0xff;
      continue;
    }
    uint8_t lab = first ? (uint8_t)(kBandClasses - 1) : labels[t];
    if (W.nbin_of[t] == c && W.need_plain[t] == 0) lab = (uint8_t)c;
    labels[t] = lab;
  }
}

// stats[2c] = tasks routed to class c (c = kBandClasses-1: plain only),
// stats[2c+1] = of those, rejected by the certificate
__global__ void __launch_bounds__(256)
band_stats_kernel(int64_t n, const uint8_t* __restrict__ bin_of, const uint8_t* __restrict__ nbin_of,
                  const uint8_t* __restrict__ need_plain, unsigned long long* __restrict__ stats) {
  __shared__ uint32_t h[2 * kBandClasses];
  if (threadIdx.x < 2 * kBandClasses) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    if (bin_of[t] == 0xff) continue;
    const int c = nbin_of[t];
    if (c == 0xff) {
      atomicAdd(&h[2 * (kBandClasses - 1)], 1u);
      continue;
    }
    atomicAdd(&h[2 * c], 1u);
    if (need_plain[t]) atomicAdd(&h[2 * c + 1], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 2 * kBandClasses && h[threadIdx.x])
    atomicAdd(&stats[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void band_stats_clear_kernel(unsigned long long* stats) {
  if (threadIdx.x < 2 * kBandClasses) stats[threadIdx.x] = 0;
}

hipError_t launch_band_label_merge(const ExtArgs& a, const ExtWork& wk, int c, uint8_t* labels, int first,
                                   int num_cu, hipStream_t st) {
  int64_t blocks = (a.n + 255) / 256;
  if (blocks > (int64_t)num_cu * 16) blocks = (int64_t)num_cu * 16;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL(band_label_merge_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, wk, c, labels, first);
  return hipGetLastError();
}

hipError_t launch_band_stats(int64_t n, const ExtWork& wk, unsigned long long* stats, int num_cu, hipStream_t st) {
  hipLaunchKernelGGL(band_stats_clear_kernel, dim3(1), dim3(64), 0, st, stats);
  int64_t blocks = (n + 255) / 256;
  if (blocks > (int64_t)num_cu * 16) blocks = (int64_t)num_cu * 16;
  if (blocks < 1) return hipGetLastError();
  hipLaunchKernelGGL(band_stats_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n, (const uint8_t*)wk.bin_of,
                     (const uint8_t*)wk.nbin_of, (const uint8_t*)wk.need_plain, stats);
  return hipGetLastError();
}

}  // namespace bsx
