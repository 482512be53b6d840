// Device helpers shared by the extension kernels (plain and narrow-band).
#pragma once
#include "bsx_internal.h"

namespace bsx {

// ---- DPP group primitives (identity 0, values >= 0) ----
// A group is L = 4, 8, 16, 32 or 64 consecutive lanes of the wave.

template <int L>
__device__ __forceinline__ int grp_incl_max(int x) {
  if (L == 4) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x90, 0xf, 0xf, false));  // quad_perm [0,0,1,2]
    return max(x, __builtin_amdgcn_update_dpp(0, x, 0x40, 0xf, 0xf, false));  // quad_perm [0,0,0,1]
  }
  if (L == 8) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x90, 0xf, 0xf, false));  // quad_perm [0,0,1,2]
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x40, 0xf, 0xf, false));  // quad_perm [0,0,0,1]
    const int q = __builtin_amdgcn_update_dpp(0, x, 0xFF, 0xf, 0xf, false);  // quad_perm [3,3,3,3]
    // row_shr:4 into banks 1 and 3 only (lanes 4-7 of each 8-lane group)
    return max(x, __builtin_amdgcn_update_dpp(0, q, 0x114, 0xf, 0xa, false));
  }
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true));  // row_shr:1
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true));  // row_shr:2
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true));  // row_shr:4
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true));  // row_shr:8
  if (L >= 32) x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  if (L >= 64) x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

// exclusive version of an inclusive scan result (lane 0 of the group gets 0)
template <int L>
__device__ __forceinline__ int grp_excl_from_incl(int incl, int lig) {
  if (L == 16) return __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, true);  // row_shr:1
  const int e = L <= 8 ? __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, true)   // row_shr:1
                       : __builtin_amdgcn_update_dpp(0, incl, 0x138, 0xf, 0xf, true);  // wave_shr:1
  return lig == 0 ? 0 : e;
}

// every lane of the group receives the maximum
template <int L>
__device__ __forceinline__ int grp_max(int x, int lane) {
  if (L <= 16) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xf, 0xf, true));   // quad_perm [1,0,3,2]
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xf, 0xf, true));   // quad_perm [2,3,0,1]
    if (L == 4) return x;
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xf, 0xf, true));  // row_half_mirror
    if (L == 16) x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x140, 0xf, 0xf, true));  // row_mirror
    return x;
  }
  x = grp_incl_max<L>(x);
  if (L == 64) return __builtin_amdgcn_readlane(x, 63);
  const int a = __builtin_amdgcn_readlane(x, 31), b = __builtin_amdgcn_readlane(x, 63);
  return lane < 32 ? a : b;
}

__device__ __forceinline__ int sbyte(uint32_t w, int t) {  // signed byte t of w
  return (int)((int32_t)(w << (24 - 8 * t)) >> 24);
}

// Kernel-side arguments of the extension kernels: only what the row loop and
// the task start read (the caller's output arrays are written by a separate
// unpack pass).
struct ExtKArgs {
  const uint8_t* query;
  const uint8_t* target;
  int32_t* rec;  // [n][8]: score qle tle gtle gscore max_off cells_lo cells_hi
  int o_del, e_del, o_ins, e_ins, zdrop;
  uint32_t prof_lo[5];  // by query code q: bytes t = 0..3 hold mat[t*5+q]
  int32_t prof_hi[5];   // by query code q: mat[20+q]
  uint32_t trow_lo[5];  // by target code t: bytes q = 0..3 hold mat[t*5+q]
  uint32_t trow_hi[5];  // by target code t: byte 0 holds mat[t*5+4]
  int tpb;              // task descriptors owned by one block
};

// The eight result words of a task (ksw.c:473-478 plus the cell count), one
// word per lane of the group (two per lane for L = 4).
template <int L>
__device__ __forceinline__ void write_rec(int32_t* rec, int task, int lig, int best, int qle, int tle, int gtle,
                                          int gscore, int moff, uint32_t ncell) {
  constexpr int S = L < 8 ? L : 8;
#pragma unroll
  for (int r = 0; r < 8; r += S) {
    const int f = lig + r;
    if (f < 8) {
      int v = best;
      v = f == 1 ? qle : v;
      v = f == 2 ? tle : v;
      v = f == 3 ? gtle : v;
      v = f == 4 ? gscore : v;
      v = f == 5 ? moff : v;
      v = f == 6 ? (int)ncell : v;
      v = f == 7 ? 0 : v;
      rec[(int64_t)task * 8 + f] = v;
    }
  }
}

// Band of a task clamped as ksw.c:399-407, in the same double arithmetic.
__device__ __forceinline__ int ext_clamped_w(const ExtArgs& A, int64_t t) {
  const int ql = A.qlen[t], eb = A.end_bonus[t];
  int w = A.w[t];
  int lim = (int)((double)(ql * A.mat_max + eb - A.o_ins) / A.e_ins + 1.);
  lim = lim > 1 ? lim : 1;
  w = w < lim ? w : lim;
  lim = (int)((double)(ql * A.mat_max + eb - A.o_del) / A.e_del + 1.);
  lim = lim > 1 ? lim : 1;
  return w < lim ? w : lim;
}

// 48-byte work-list descriptor of task t (clamped band included).
__device__ __forceinline__ void write_desc(const ExtArgs& A, int64_t t, ExtDesc* dst) {
  const uint64_t qo = A.q_off[t], to = A.t_off[t];
  int4* d = (int4*)dst;
  d[0] = make_int4((int)t, A.qlen[t], A.tlen[t], A.h0[t]);
  d[1] = make_int4(ext_clamped_w(A, t), 0, (int)(uint32_t)qo, (int)(uint32_t)(qo >> 32));
  d[2] = make_int4((int)(uint32_t)to, (int)(uint32_t)(to >> 32), 0, 0);
}

}  // namespace bsx
