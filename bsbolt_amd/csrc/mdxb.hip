// NM, MD:Z and XB:Z of a CIGAR: the tail of bwa_gen_cigar2
// (bsbolt/External/BWA/bwa.c:250-343) with getMethylationContext (bwa.c:180-197).
//
// One lane per task.  A task walks its CIGAR once over the unconverted
// reference window opac[rb-2, re+2) (bns_get_seq's clamps, bntseq.c:413-437,
// reversed when rb >= l_pac like the read, bwa.c:264-271) and the read as
// sequenced, appending to its MD and XB slots.  The window is read straight
// from the 2-bit opac (four bases per byte, first base in the top bits), so a
// task touches only the bytes of its window, its read, its CIGAR and its
// output slots.  The work per task is a few hundred byte operations; the
// kernel is bound by the latency of those gathers, not by arithmetic.
#include "bsx_internal.h"

namespace bsx {

__device__ __forceinline__ int pac_at(const uint8_t* __restrict__ pac, int64_t p) {
  return (pac[p >> 2] >> ((~p & 3) << 1)) & 3;
}

// appends the decimal digits of v >= 0 (kputw, kstring.h:62-76); false on overflow
__device__ __forceinline__ bool put_int(char* __restrict__ s, int& l, int cap, int v) {
  char d[11];
  int n = 0;
  if (v == 0) d[n++] = '0';
  for (; v > 0; v /= 10) d[n++] = (char)('0' + v % 10);
  if (l + n > cap) return false;
  while (n > 0) s[l++] = d[--n];
  return true;
}

__device__ __forceinline__ char meth_context(int b1, int b2, int b3, bool meth, int& cg, int& ch) {
  if (b1 == 1 && b2 == 2) { ++cg; return meth ? 'X' : 'x'; }
  if (b1 == 1 && b3 == 2) { ++ch; return meth ? 'Y' : 'y'; }
  ++ch;
  return meth ? 'Z' : 'z';
}

__global__ void __launch_bounds__(256) md_kernel(MdArgs a) {
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
    bool oob = false;
    auto cs = [&](int64_t k) -> int {
      if (k < 0 || k >= clen) { oob = true; return 4; }
      const int b = pac_at(a.opac, k0 + k);
      return fwd ? b : 3 - b;
    };
    const uint8_t* __restrict__ oq = a.oquery + a.oq_off[t];
    auto oqb = [&](int x) -> int { return fwd ? oq[x] : oq[l_query - 1 - x]; };
    const bool reverse = rb > l_pac;  // bwa.c:212 (strictly greater)
    const bool is_crick = (rb >= a.crick_l && rb < l_pac) || rb >= a.crick_l + l_pac;
    const int ref_base = is_crick ? 2 : 1, methyl_base = is_crick ? 0 : 3;
    const char* i2b = fwd ? "ACGTN" : "TGCAN";
    char* __restrict__ md = a.md + t * a.md_slot;
    char* __restrict__ xb = a.xb + t * a.xb_slot;
    const int md_cap = (int)a.md_slot, xb_cap = (int)a.xb_slot;
    int ml = 0, xl = 0, x = 0, y = 0, u = 0, n_mm = 0, n_gap = 0, meth_pos = 0;
    int cgm = 0, cgu = 0, chm = 0, chu = 0;
    int st = 0;
    for (int k = 0; k < nc && st == 0; ++k) {
      const uint32_t cw = cig[k];
      const int op = (int)(cw & 0xfu), len = (int)(cw >> 4);
      if (op == 0) {
        for (int i = 0; i < len && st == 0; ++i) {
          const int c = cs(y + i + 2);
          const int o = oqb(x + i);
          if (c == ref_base) {
            int b1, b2, b3;
            if (!reverse && !is_crick) { b1 = cs(y + i + 2); b2 = cs(y + i + 3); b3 = cs(y + i + 4); }
            else if (!reverse) { b1 = cs(y + i + 1); b2 = cs(y + i + 2); b3 = cs(y + i + 3); }
            else if (!is_crick) { b1 = cs(y + i + 2); b2 = cs(y + i + 1); b3 = cs(y + i); }
            else { b1 = cs(y + i + 3); b2 = cs(y + i + 2); b3 = cs(y + i + 1); }
            if (o == ref_base || o == methyl_base) {
              const char m = o == ref_base ? meth_context(b1, b2, b3, true, cgm, chm)
                                           : meth_context(b1, b2, b3, false, cgu, chu);
              if (meth_pos > 0 && !put_int(xb, xl, xb_cap, meth_pos)) st = -1;
              if (st == 0 && xl + 1 > xb_cap) st = -1;
              if (st == 0) xb[xl++] = m;
              meth_pos = 0;
              ++u;
            } else {
              if (!put_int(md, ml, md_cap, u) || ml + 1 > md_cap) st = -1;
              else md[ml++] = i2b[c];
              ++n_mm; ++meth_pos; u = 0;
            }
          } else if (o != c) {
            if (!put_int(md, ml, md_cap, u) || ml + 1 > md_cap) st = -1;
            else md[ml++] = i2b[c];
            ++n_mm; ++meth_pos; u = 0;
          } else {
            ++u; ++meth_pos;
          }
          if (oob) st = -2;
        }
        x += len; y += len;
      } else if (op == 2) {
        if (k > 0 && k < nc - 1) {  // not for a leading or trailing D
          if (!put_int(md, ml, md_cap, u) || ml + 1 + len > md_cap) st = -1;
          else {
            md[ml++] = '^';
            for (int i = 0; i < len; ++i) md[ml++] = i2b[cs(y + i + 2)];
            if (oob) st = -2;
          }
          u = 0; n_gap += len;
        }
        y += len;
      } else if (op == 1) {
        x += len; n_gap += len; meth_pos += len;
      }
    }
    if (st == 0 && !put_int(md, ml, md_cap, u)) st = -1;
    if (st == 0 && meth_pos > 0 && !put_int(xb, xl, xb_cap, meth_pos)) st = -1;
    a.status[t] = st;
    if (st != 0) continue;
    a.nm[t] = n_mm + n_gap;
    a.md_len[t] = ml;
    a.xb_len[t] = xl;
    meth[0] = cgm; meth[1] = cgu; meth[2] = chm; meth[3] = chu;
  }
}

hipError_t launch_md(const MdArgs& a, int num_cu, hipStream_t st) {
  if (a.n <= 0) return hipSuccess;
  int64_t blocks = (a.n + 255) / 256;
  if (blocks > (int64_t)num_cu * 32) blocks = (int64_t)num_cu * 32;
  hipLaunchKernelGGL(md_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace bsx
