// NM, MD:Z and XB:Z of a CIGAR: the tail of bwa_gen_cigar2
// (bsbolt/External/BWA/bwa.c:250-343) with getMethylationContext (bwa.c:180-197).
//
// One lane per task.  A task walks its CIGAR once over the unconverted
// reference window opac[rb-2, re+2) (bns_get_seq's clamps, bntseq.c:413-437,
// reversed when rb >= l_pac like the read, bwa.c:264-271) and the read as
// sequenced, appending to its MD and XB slots.  The lane first copies the
// window (2-bit, 16 bases per dword) and its read into its own LDS rows with
// independent dword loads, so the walk's byte reads never wait on HBM; a task
// too long for its rows reads both from HBM instead.
#include "bsx_internal.h"

namespace bsx {

__device__ __forceinline__ int pac_at(const uint8_t* __restrict__ pac, int64_t p) {
  return (pac[p >> 2] >> ((~p & 3) << 1)) & 3;
}

// A lane's output string: characters gather in a register and leave as one
// dword store per four (the slot is dword-aligned), or byte by byte otherwise.
struct StrOut {
  char* __restrict__ p;
  int len, cap;
  uint32_t acc;
  bool words;
  __device__ __forceinline__ bool put(int c) {
    if (len >= cap) return false;
    acc |= (uint32_t)(c & 0xff) << (8 * (len & 3));
    ++len;
    if ((len & 3) == 0) {
      if (words) {
        *(uint32_t*)(p + len - 4) = acc;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) p[len - 4 + k] = (char)(acc >> (8 * k));
      }
      acc = 0;
    }
    return true;
  }
  __device__ __forceinline__ void flush() {
    for (int k = 0; k < (len & 3); ++k) p[(len & ~3) + k] = (char)(acc >> (8 * k));
  }
  // appends the decimal digits of v >= 0 (kputw, kstring.h:62-76); false on
  // overflow.  The digits gather in nibbles of one register, most significant
  // in the low nibble, so no array is indexed at run time.
  __device__ __forceinline__ bool put_int(int v) {
    uint64_t dg = 0;
    int n = 0;
    do {
      dg = (dg << 4) | (uint64_t)(v % 10);
      v /= 10;
      ++n;
    } while (v > 0);
    if (len + n > cap) return false;
    for (int k = 0; k < n; ++k) put('0' + (int)((dg >> (4 * k)) & 15));
    return true;
  }
};

constexpr int kMdBlock = 128;
constexpr int kMdWinDw = 20;   // LDS dwords per lane for the window: 320 bases
constexpr int kMdReadDw = 48;  // ... for the read: 192 bytes

__global__ void __launch_bounds__(kMdBlock) md_kernel(MdArgs a) {
  __shared__ uint32_t s_win[kMdWinDw][kMdBlock];   // column per lane: lanes read distinct banks
  __shared__ uint32_t s_read[kMdReadDw][kMdBlock];
  const int lid = threadIdx.x;
  const uint32_t* __restrict__ opac32 = (const uint32_t*)a.opac;
  // dword stores for the strings when every slot starts on a dword
  const bool md_words = ((uintptr_t)a.md & 3) == 0 && (a.md_slot & 3) == 0;
  const bool xb_words = ((uintptr_t)a.xb & 3) == 0 && (a.xb_slot & 3) == 0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.n; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rb = a.rb[t], re = a.re[t], l_pac = a.l_pac;
    const int l_query = a.l_query[t];
    int32_t* meth = a.meth + 4 * t;
    a.nm[t] = -1;
    a.md_len[t] = 0;
    a.xb_len[t] = 0;
    meth[0] = meth[1] = meth[2] = meth[3] = 0;
    const int nc = a.n_cigar[t];
    const uint32_t* __restrict__ cig = a.cigar + t * a.cigar_slot;
    {  // a region bwa_gen_cigar2 makes a CIGAR for, whose CIGAR fits its slot
       // and consumes at most the read
      bool bad = rb < 0 || rb >= re || re > 2 * l_pac || nc < 0 || (int64_t)nc > a.cigar_slot;
      int64_t qsum = 0;
      for (int k = 0; k < nc && !bad; ++k) {
        const uint32_t op = cig[k] & 0xfu;
        if (op == 0 || op == 1) qsum += cig[k] >> 4;
        else if (op != 2) bad = true;
      }
      if (bad || qsum > l_query) {
        a.status[t] = -3;
        continue;
      }
    }
    // window opac[rb-2, re+2) as bns_get_seq clamps it (rb < re, so no swap)
    int64_t beg = rb - 2, end = re + 2;
    if (end > 2 * l_pac) end = 2 * l_pac;
    if (beg < 0) beg = 0;
    if (!(beg >= l_pac || end <= l_pac)) {  // bridges the strands: the reference stops (bwa.c:262)
      a.status[t] = 1;
      continue;
    }
    const int64_t clen = end - beg;
    const bool fwd = rb < l_pac;  // rb >= l_pac: window and read reversed (bwa.c:264-271)
    // base k of the window in the order the walk reads it: forward positions
    // beg + k; on the reverse strand the reversed reverse-complement window,
    // which is 3 - opac[2*l_pac - end + k]
    const int64_t k0 = fwd ? beg : 2 * l_pac - end;
    // stage the window and the read in this lane's LDS rows: aligned dwords
    // that hold at least one byte of the window / read (such a dword lies in
    // the allocation the byte lies in), all loads issued before any use
    const int64_t wd0 = k0 >> 4, wd1 = (k0 + clen - 1) >> 4;
    const uint64_t oq0 = a.oq_off[t];
    const int64_t rd0 = (int64_t)(oq0 >> 2), rd1 = l_query > 0 ? (int64_t)((oq0 + l_query - 1) >> 2) : rd0;
    const bool staged = wd1 - wd0 < kMdWinDw && rd1 - rd0 < kMdReadDw;
    if (staged) {  // every load in flight at once (indices past the last dword repeat it)
      const uint32_t* __restrict__ rq = (const uint32_t*)a.oquery;
      const int lw = (int)(wd1 - wd0), lr = (int)(rd1 - rd0);
      uint32_t wv[kMdWinDw], rv[kMdReadDw];
#pragma unroll
      for (int j = 0; j < kMdWinDw; ++j) wv[j] = opac32[wd0 + (j < lw ? j : lw)];
#pragma unroll
      for (int j = 0; j < kMdReadDw; ++j) rv[j] = rq[rd0 + (j < lr ? j : lr)];
#pragma unroll
      for (int j = 0; j < kMdWinDw; ++j) s_win[j][lid] = wv[j];
#pragma unroll
      for (int j = 0; j < kMdReadDw; ++j) s_read[j][lid] = rv[j];
    }
    bool oob = false;
    auto cs = [&](int64_t k) -> int {
      if (k < 0 || k >= clen) { oob = true; return 4; }
      const int64_t p = k0 + k;
      int b;
      if (staged) {
        const uint32_t w = s_win[(int)((p >> 4) - wd0)][lid];
        b = (int)((w >> (8 * ((p >> 2) & 3))) >> ((~p & 3) << 1)) & 3;  // byte (p>>2)&3 of the dword
      } else {
        b = pac_at(a.opac, p);
      }
      return fwd ? b : 3 - b;
    };
    const uint8_t* __restrict__ oq = a.oquery + oq0;
    auto oqb = [&](int x) -> int {
      const int xx = fwd ? x : l_query - 1 - x;
      if (!staged) return oq[xx];
      const uint64_t q = oq0 + (uint64_t)xx;
      return (int)(s_read[(int)((int64_t)(q >> 2) - rd0)][lid] >> (8 * (q & 3))) & 0xff;
    };
    const bool reverse = rb > l_pac;  // bwa.c:212 (strictly greater)
    const bool is_crick = (rb >= a.crick_l && rb < l_pac) || rb >= a.crick_l + l_pac;
    const int ref_base = is_crick ? 2 : 1, methyl_base = is_crick ? 0 : 3;
    // window offsets (from the base's own index) of the three context bases
    // (bwa.c:290-293): after it on the forward Watson copy, around it on the
    // others
    const int o1 = !reverse ? (!is_crick ? 2 : 1) : (!is_crick ? 2 : 3);
    const int o2 = !reverse ? (!is_crick ? 3 : 2) : (!is_crick ? 1 : 2);
    const int o3 = !reverse ? (!is_crick ? 4 : 3) : (!is_crick ? 0 : 1);
    const char* i2b = fwd ? "ACGTN" : "TGCAN";
    StrOut md{a.md + t * a.md_slot, 0, (int)a.md_slot, 0u, md_words};
    StrOut xb{a.xb + t * a.xb_slot, 0, (int)a.xb_slot, 0u, xb_words};
    int x = 0, y = 0, u = 0, n_mm = 0, n_gap = 0, meth_pos = 0;
    int cgm = 0, cgu = 0, chm = 0, chu = 0;
    int st = 0;
    // One step is one base of an M op or one whole I/D op, so a wave's lanes
    // stay in step across their different CIGAR layouts.  Each base has one
    // shape (the lanes diverge less): the context of a C (G on the Crick copy)
    // is read where the reference reads it, one XB and one MD emission site,
    // counters by select.
    int k = 0, pos = 0;  // op index, bases of the current M op walked
    uint32_t cw = nc > 0 ? cig[0] : 0u;
    while (k < nc && st == 0) {
      const int op = (int)(cw & 0xfu), len = (int)(cw >> 4);
      bool next = true;
      if (op == 0) {
        if (pos < len) {
          const int c = cs(y + pos + 2);
          const int o = oqb(x + pos);
          const bool isref = c == ref_base;
          const bool ev = isref && (o == ref_base || o == methyl_base);  // methylation call
          const bool mm = isref ? !ev : o != c;                           // MD mismatch
          if (isref) {
            const int b1 = cs(y + pos + o1), b2 = cs(y + pos + o2), b3 = cs(y + pos + o3);
            if (ev) {
              const bool meth = o == ref_base;
              const bool cg = b1 == 1 && b2 == 2, chg = !cg && b1 == 1 && b3 == 2;
              const char m = (char)((cg ? 'X' : chg ? 'Y' : 'Z') + (meth ? 0 : 32));
              cgm += meth && cg; chm += meth && !cg; cgu += !meth && cg; chu += !meth && !cg;
              if ((meth_pos > 0 && !xb.put_int(meth_pos)) || !xb.put(m)) st = -1;
            }
          }
          if (mm && (!md.put_int(u) || !md.put(i2b[c]))) st = -1;
          n_mm += mm;
          u = mm ? 0 : u + 1;
          meth_pos = ev ? 0 : meth_pos + 1;
          if (oob) st = -2;
          ++pos;
          next = pos >= len;
        }
        if (next) { x += len; y += len; }
      } else if (op == 2) {
        if (k > 0 && k < nc - 1) {  // not for a leading or trailing D
          if (!md.put_int(u) || !md.put('^')) st = -1;
          for (int i = 0; i < len && st == 0; ++i)
            if (!md.put(i2b[cs(y + i + 2)])) st = -1;
          if (st == 0 && oob) st = -2;
          u = 0; n_gap += len;
        }
        y += len;
      } else if (op == 1) {
        x += len; n_gap += len; meth_pos += len;
      }
      if (next) {
        ++k;
        pos = 0;
        cw = k < nc ? cig[k] : 0u;
      }
    }
    if (st == 0 && !md.put_int(u)) st = -1;
    if (st == 0 && meth_pos > 0 && !xb.put_int(meth_pos)) st = -1;
    a.status[t] = st;
    if (st != 0) continue;
    md.flush();
    xb.flush();
    a.nm[t] = n_mm + n_gap;
    a.md_len[t] = md.len;
    a.xb_len[t] = xb.len;
    meth[0] = cgm; meth[1] = cgu; meth[2] = chm; meth[3] = chu;
  }
}

hipError_t launch_md(const MdArgs& a, int num_cu, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  int64_t blocks = (a.n + kMdBlock - 1) / kMdBlock;
  if (blocks > (int64_t)num_cu * 64) blocks = (int64_t)num_cu * 64;
  hipLaunchKernelGGL(md_kernel, dim3((unsigned)blocks), dim3(kMdBlock), 0, st, a);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The tail of mem_reg2aln (bwamem.c:1237-1263): the SAM CIGAR of a region from
// the CIGAR bwa_gen_cigar2 made for it.  The position is bns_depos of rb (or
// of re - 1 on the reverse strand); a leading deletion is squeezed out and
// moves the position, or else (only if the first op is not one) a trailing
// deletion is dropped; then the unaligned read ends become clip ops
// (len << 4 | 3), 5' first, swapped on the reverse strand.  One lane per task:
// a handful of words each.
__global__ void __launch_bounds__(256)
reg2aln_cigar_kernel(Reg2alnArgs a) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t rb = a.rb[t], re = a.re[t];
    const int qb = a.qb[t], qe = a.qe[t], lq = a.l_seq[t], nc = a.n_cigar[t];
    const uint32_t* __restrict__ c = a.cigar + t * a.cigar_slot;
    if (nc < 0 || nc > a.cigar_slot || rb < 0 || re <= rb || re > 2 * a.l_pac || qb < 0 || qe < qb || qe > lq) {
      a.out_n[t] = -1;
      a.pos[t] = -1;
      a.is_rev[t] = 0;
      continue;
    }
    const int64_t p0 = rb < a.l_pac ? rb : re - 1;  // bns_depos (bntseq.h)
    const bool rev = p0 >= a.l_pac;
    int64_t pos = rev ? 2 * a.l_pac - 1 - p0 : p0;
    int first = 0, last = nc;
    if (nc > 0) {
      if ((c[0] & 0xf) == 2) {
        pos += c[0] >> 4;
        first = 1;
      } else if ((c[nc - 1] & 0xf) == 2) {
        last = nc - 1;
      }
    }
    int clip5 = 0, clip3 = 0;
    if (qb != 0 || qe != lq) {
      clip5 = rev ? lq - qe : qb;
      clip3 = rev ? qb : lq - qe;
    }
    const int n_out = (clip5 ? 1 : 0) + (last - first) + (clip3 ? 1 : 0);
    a.pos[t] = pos;
    a.is_rev[t] = rev ? 1 : 0;
    if (n_out > a.out_slot) {
      a.out_n[t] = -2;
      continue;
    }
    uint32_t* __restrict__ o = a.out + t * a.out_slot;
    int k = 0;
    if (clip5) o[k++] = (uint32_t)clip5 << 4 | 3u;
    for (int j = first; j < last; ++j) o[k++] = c[j];
    if (clip3) o[k++] = (uint32_t)clip3 << 4 | 3u;
    a.out_n[t] = k;
  }
}

hipError_t launch_reg2aln_cigar(const Reg2alnArgs& a, int num_cu, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  int64_t blocks = (a.n + 255) / 256;
  if (blocks > (int64_t)num_cu * 16) blocks = (int64_t)num_cu * 16;
  hipLaunchKernelGGL(reg2aln_cigar_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace bsx
