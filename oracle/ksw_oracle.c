/*
 * ksw_oracle.c -- CPU restatement of BSBolt's seed-extend DP, used ONLY as the
 * test oracle (tests/, the smoke check and the CPU-baseline leg of bench.py).
 * Nothing in the product library links or calls this file.
 *
 * Every function restates the algorithm of the forked BWA-MEM 0.7.17 shipped in
 * BSBolt (bsbolt/External/BWA), citing the lines it follows:
 *   orc_fill_scmat      bwa.c:169-178          (bwa_fill_scmat)
 *   orc_extend2         ksw.c:380-479          (ksw_extend2)
 *   orc_global2         ksw.c:486-606          (ksw_global2, push_cigar)
 *   orc_cal_max_gap     bwamem.c:625-632       (cal_max_gap)
 *   orc_get_seq         bntseq.c:236-237,413-437 (_get_pac, bns_get_seq)
 *   orc_global_band     bwa.c:231-241          (band choice in bwa_gen_cigar2)
 *   orc_bs_convert      bs_helpers.cpp:18-31   (bsConversion)
 *   orc_nt4             bntseq.c:48-65         (nst_nt4_table)
 *   orc_gen_md          bwa.c:180-197,250-343  (NM, MD and XB:Z of bwa_gen_cigar2,
 *                                               getMethylationContext)
 *   orc_reg2aln_cigar   bwamem.c:1237-1263     (SAM CIGAR and position in mem_reg2aln:
 *                                               deletion squeeze, clipping)
 * Parity of each function is pinned by tests against golden vectors produced by
 * the compiled reference (oracle/_ref) and by closed-form cases.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

#define ORC_NEG_INF (-0x40000000) /* ksw.c:490 */

/* bwa.c:169-178: +a on the 4x4 diagonal, -b off it, -1 in row/column 4 (N). */
void orc_fill_scmat(int a, int b, int8_t mat[25])
{
	for (int r = 0; r < 5; ++r)
		for (int c = 0; c < 5; ++c)
			mat[r * 5 + c] = (r == 4 || c == 4) ? -1 : (r == c ? a : -b);
}

/* The band clamp of ksw.c:399-407, kept in double arithmetic exactly as there. */
static int clamp_band(int w, int qlen, int m, const int8_t *mat, int end_bonus,
                      int o_del, int e_del, int o_ins, int e_ins)
{
	int best = 0;
	for (int k = 0; k < m * m; ++k) if (mat[k] > best) best = mat[k];
	int lim_ins = (int)((double)(qlen * best + end_bonus - o_ins) / e_ins + 1.);
	if (lim_ins < 1) lim_ins = 1;
	if (w > lim_ins) w = lim_ins;
	int lim_del = (int)((double)(qlen * best + end_bonus - o_del) / e_del + 1.);
	if (lim_del < 1) lim_del = 1;
	if (w > lim_del) w = lim_del;
	return w;
}

/*
 * ksw.c:380-479.  Row i walks the target, column j the query.  Two arrays carry
 * the state between rows and are NOT cleared outside the active band (the
 * reference reads such stale entries when the band regrows):
 *   hd[j]  H(i-1, j-1)   (the diagonal predecessor of cell (i, j))
 *   ev[j]  E(i, j)
 * `cells` (nullable) receives the inner-loop trip count, sum over executed rows
 * of (end - beg), the work unit of the measurements.
 */
int orc_extend2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                int m, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins,
                int w, int end_bonus, int zdrop, int h0,
                int *qle, int *tle, int *gtle, int *gscore_out, int *max_off_out,
                int64_t *cells)
{
	const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
	int32_t *hd = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
	int32_t *ev = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
	int8_t *prof = (int8_t *)malloc((size_t)qlen * m + 1);
	int64_t ncell = 0;

	/* query profile, ksw.c:390-393 */
	for (int k = 0; k < m; ++k)
		for (int j = 0; j < qlen; ++j)
			prof[k * qlen + j] = mat[k * m + query[j]];

	/* row -1: the insertion ramp out of the anchor, ksw.c:395-397 */
	hd[0] = h0;
	hd[1] = h0 > oe_ins ? h0 - oe_ins : 0;
	for (int j = 2; j <= qlen; ++j) {
		if (!(hd[j - 1] > e_ins)) break;
		hd[j] = hd[j - 1] - e_ins;
	}

	w = clamp_band(w, qlen, m, mat, end_bonus, o_del, e_del, o_ins, e_ins);

	int best = h0, best_i = -1, best_j = -1; /* ksw.c:409 */
	int gbest = -1, gbest_i = -1, moff = 0;
	int lo = 0, hi = qlen;
	for (int i = 0; i < tlen; ++i) {
		const int8_t *srow = prof + (size_t)target[i] * qlen;
		/* band for this row, ksw.c:417-419 */
		if (lo < i - w) lo = i - w;
		if (hi > i + w + 1) hi = i + w + 1;
		if (hi > qlen) hi = qlen;
		/* H(i, lo-1): the deletion ramp in column -1, else 0 (ksw.c:421-424) */
		int hleft = 0;
		if (lo == 0) {
			hleft = h0 - (o_del + e_del * (i + 1));
			if (hleft < 0) hleft = 0;
		}
		int f = 0, rmax = 0, rarg = -1, j = lo;
		for (; j < hi; ++j) { /* ksw.c:425-448 */
			int diag = hd[j], e = ev[j];
			hd[j] = hleft;
			int mm = diag ? diag + srow[j] : 0; /* a zero predecessor never extends */
			int h = mm > e ? mm : e;
			if (f > h) h = f;
			hleft = h;
			if (h >= rmax) { rmax = h; rarg = j; } /* ties move to the larger j */
			int open_d = mm - oe_del; if (open_d < 0) open_d = 0;
			e -= e_del; if (open_d > e) e = open_d;
			ev[j] = e;
			int open_i = mm - oe_ins; if (open_i < 0) open_i = 0;
			f -= e_ins; if (open_i > f) f = open_i;
		}
		ncell += hi > lo ? hi - lo : 0;
		hd[hi] = hleft; ev[hi] = 0; /* ksw.c:449 */
		if (j == qlen) { /* end-to-end column, ksw.c:450-453 */
			if (!(gbest > hleft)) { gbest_i = i; gbest = hleft; }
		}
		if (rmax == 0) break; /* ksw.c:454 */
		if (rmax > best) { /* ksw.c:455-457 */
			best = rmax; best_i = i; best_j = rarg;
			int off = rarg > i ? rarg - i : i - rarg;
			if (off > moff) moff = off;
		} else if (zdrop > 0) { /* ksw.c:458-464 */
			int di = i - best_i, dj = rarg - best_j;
			int drop = di > dj ? best - rmax - (di - dj) * e_del
			                   : best - rmax - (dj - di) * e_ins;
			if (drop > zdrop) break;
		}
		/* shrink/grow the band from the zero pattern just written, ksw.c:466-469 */
		int k = lo;
		while (k < hi && hd[k] == 0 && ev[k] == 0) ++k;
		lo = k;
		k = hi;
		while (k >= lo && hd[k] == 0 && ev[k] == 0) --k;
		hi = k + 2 < qlen ? k + 2 : qlen;
	}
	free(hd); free(ev); free(prof);
	if (qle) *qle = best_j + 1;
	if (tle) *tle = best_i + 1;
	if (gtle) *gtle = gbest_i + 1;
	if (gscore_out) *gscore_out = gbest;
	if (max_off_out) *max_off_out = moff;
	if (cells) *cells = ncell;
	return best;
}

/* ksw.c:492-502: append (op,len) to a BAM CIGAR, merging equal ops. */
static uint32_t *cigar_push(uint32_t *cig, int *n, int *cap, int op, int len)
{
	if (*n > 0 && (int)(cig[*n - 1] & 0xf) == op) {
		cig[*n - 1] += (uint32_t)len << 4;
		return cig;
	}
	if (*n == *cap) {
		*cap = *cap ? *cap * 2 : 4;
		cig = (uint32_t *)realloc(cig, (size_t)*cap * 4);
	}
	cig[(*n)++] = (uint32_t)len << 4 | (uint32_t)op;
	return cig;
}

/*
 * ksw.c:504-606.  Banded global DP with a fixed band and a direction byte per
 * cell (bits 0-1: source of H, bit 2: E extended, bit 5: F extended), then a
 * backtrack from the last cell.  CIGAR ops: 0=M, 1=I, 2=D.  *cigar is malloc'd
 * (caller frees); n_cigar/cigar may be NULL for a score-only call.
 */
int orc_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                int m, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins,
                int w, int *n_cigar, uint32_t **cigar)
{
	const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
	const int want_cigar = n_cigar && cigar;
	if (n_cigar) *n_cigar = 0;
	const int ncol = qlen < 2 * w + 1 ? qlen : 2 * w + 1;
	uint8_t *dir = want_cigar ? (uint8_t *)malloc((size_t)ncol * tlen + 1) : 0;
	int32_t *hd = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
	int32_t *ev = (int32_t *)calloc((size_t)qlen + 1, sizeof(int32_t));
	int8_t *prof = (int8_t *)malloc((size_t)qlen * m + 1);
	for (int k = 0; k < m; ++k)
		for (int j = 0; j < qlen; ++j)
			prof[k * qlen + j] = mat[k * m + query[j]];
	/* row -1, ksw.c:522-525 */
	hd[0] = 0; ev[0] = ORC_NEG_INF;
	int j = 1;
	for (; j <= qlen && j <= w; ++j) { hd[j] = -(o_ins + e_ins * j); ev[j] = ORC_NEG_INF; }
	for (; j <= qlen; ++j) hd[j] = ev[j] = ORC_NEG_INF;
	for (int i = 0; i < tlen; ++i) {
		const int8_t *srow = prof + (size_t)target[i] * qlen;
		int lo = i > w ? i - w : 0;
		int hi = i + w + 1 < qlen ? i + w + 1 : qlen;
		int32_t hleft = lo == 0 ? -(o_del + e_del * (i + 1)) : ORC_NEG_INF;
		int32_t f = ORC_NEG_INF;
		uint8_t *drow = want_cigar ? dir + (size_t)i * ncol : 0;
		for (j = lo; j < hi; ++j) { /* ksw.c:546-565 */
			int32_t mm = hd[j] + srow[j], e = ev[j];
			hd[j] = hleft;
			uint8_t d = mm >= e ? 0 : 1;
			int32_t h = mm >= e ? mm : e;
			if (!(h >= f)) { d = 2; h = f; }
			hleft = h;
			int32_t open_d = mm - oe_del;
			e -= e_del;
			if (e > open_d) d |= 1 << 2; else e = open_d;
			ev[j] = e;
			int32_t open_i = mm - oe_ins;
			f -= e_ins;
			if (f > open_i) d |= 2 << 4; else f = open_i;
			if (drow) drow[j - lo] = d;
		}
		hd[hi] = hleft; ev[hi] = ORC_NEG_INF;
	}
	int score = hd[qlen];
	if (want_cigar) { /* backtrack, ksw.c:588-602 */
		int n = 0, cap = 0, state = 0;
		uint32_t *cg = 0;
		int i = tlen - 1;
		int k = (i + w + 1 < qlen ? i + w + 1 : qlen) - 1;
		while (i >= 0 && k >= 0) {
			int lo = i > w ? i - w : 0;
			state = dir[(size_t)i * ncol + (k - lo)] >> (state << 1) & 3;
			if (state == 0) { cg = cigar_push(cg, &n, &cap, 0, 1); --i; --k; }
			else if (state == 1) { cg = cigar_push(cg, &n, &cap, 2, 1); --i; }
			else { cg = cigar_push(cg, &n, &cap, 1, 1); --k; }
		}
		if (i >= 0) cg = cigar_push(cg, &n, &cap, 2, i + 1);
		if (k >= 0) cg = cigar_push(cg, &n, &cap, 1, k + 1);
		for (int a = 0, b = n - 1; a < b; ++a, --b) { uint32_t t = cg[a]; cg[a] = cg[b]; cg[b] = t; }
		*n_cigar = n; *cigar = cg;
	}
	free(dir); free(hd); free(ev); free(prof);
	return score;
}

void orc_free(void *p) { free(p); }

/* bwamem.c:625-632: longest gap worth considering on a flank of length qlen. */
int orc_cal_max_gap(int a, int o_del, int e_del, int o_ins, int e_ins, int w, int qlen)
{
	int ld = (int)((double)(qlen * a - o_del) / e_del + 1.);
	int li = (int)((double)(qlen * a - o_ins) / e_ins + 1.);
	int l = ld > li ? ld : li;
	if (l < 1) l = 1;
	return l < 2 * w ? l : 2 * w;
}

/* bwa.c:231-241: band of the global (CIGAR) alignment.  mat0 = mat[0] (= a). */
int orc_global_band(int mat0, int o_del, int e_del, int o_ins, int e_ins, int w_,
                    int l_query, int rlen)
{
	int half = (l_query + 1) >> 1;
	int li = (int)((double)(half * mat0 - o_ins) / e_ins + 1.);
	int ld = (int)((double)(half * mat0 - o_del) / e_del + 1.);
	int gap = li > ld ? li : ld;
	if (gap < 1) gap = 1;
	int diff = rlen - l_query; if (diff < 0) diff = -diff;
	int w = (gap + diff + 1) >> 1;
	if (w > w_) w = w_;
	if (w < diff + 3) w = diff + 3;
	return w;
}

/*
 * bntseq.c:413-437: fetch [beg,end) of the doubled reference from the 2-bit
 * forward-only pac (4 bases per byte, first base in the top bits).  Positions
 * >= l_pac are the reverse-complement strand.  Returns the length written (0 if
 * the interval bridges the two strands), after the same swap/clamp steps.
 */
int64_t orc_get_seq(int64_t l_pac, const uint8_t *pac, int64_t beg, int64_t end, uint8_t *out)
{
	if (end < beg) { int64_t t = beg; beg = end; end = t; }
	if (end > 2 * l_pac) end = 2 * l_pac;
	if (beg < 0) beg = 0;
	if (!(beg >= l_pac || end <= l_pac)) return 0;
	int64_t n = 0;
	if (beg >= l_pac) {
		for (int64_t k = 2 * l_pac - 1 - beg; k > 2 * l_pac - 1 - end; --k)
			out[n++] = 3 - (pac[k >> 2] >> ((~k & 3) << 1) & 3);
	} else {
		for (int64_t k = beg; k < end; ++k)
			out[n++] = pac[k >> 2] >> ((~k & 3) << 1) & 3;
	}
	return n;
}

/* Decimal digits of a non-negative int appended to buf (kputw, kstring.h:62-76). */
static int put_int(char *buf, int l, int cap, int v)
{
	char d[12];
	int n = 0;
	if (v == 0) d[n++] = '0';
	for (; v > 0; v /= 10) d[n++] = (char)('0' + v % 10);
	if (l + n > cap) return -1;
	while (n > 0) buf[l++] = d[--n];
	return l;
}

/* bwa.c:180-197: methylation context of a C (or G on the Crick copy): CpG when
 * the next base is G, CHG when the one after is, CHH otherwise; upper case
 * when the read kept the base (methylated). */
static char meth_context(int b1, int b2, int b3, int meth, int *cg, int *ch)
{
	if (b1 == 1 && b2 == 2) { ++*cg; return meth ? 'X' : 'x'; }
	if (b1 == 1 && b3 == 2) { ++*ch; return meth ? 'Y' : 'y'; }
	++*ch;
	return meth ? 'Z' : 'z';
}

/* bwa.c:250-343: the NM, MD and XB:Z strings bwa_gen_cigar2 appends after a
 * CIGAR it produced for the region [rb, re) of the doubled reference.  opac is
 * the unconverted reference in the same 2-bit layout as pac (both copies,
 * l_pac bases); oquery the read as sequenced (before the in-silico conversion),
 * in read order; cigar the ksw_global2 CIGAR (computed on the reversed
 * sequences when rb >= l_pac, bwa.c:216-221).  The context window is opac
 * [rb-2, re+2) fetched by bns_get_seq (clamped at the reference ends exactly as
 * there) and reversed when rb >= l_pac, as is the read.  Writes md[] (MD:Z
 * value), xb[] (XB:Z value), *NM and counts = {CpG methylated, CpG
 * unmethylated, CH methylated, CH unmethylated}.  Returns 0; 1 when the window
 * bridges the strands (the reference stops there: no MD, NM stays -1); -1 when
 * md_cap or xb_cap is too small; -2 when the walk would index the window out of
 * range (possible only within two bases of the reference ends, where the
 * reference reads past its buffer); -3 when [rb, re) is not a region the
 * reference makes a CIGAR for (rb < 0, rb >= re, re > 2*l_pac) or the CIGAR has
 * an op other than M/I/D or consumes more than l_query read bases. */
int orc_gen_md(int64_t l_pac, const uint8_t *opac, int64_t rb, int64_t re, int64_t crick_l, int l_query,
               const uint8_t *oquery, int n_cigar, const uint32_t *cigar, char *md, int md_cap, int *md_len,
               char *xb, int xb_cap, int *xb_len, int *NM, int counts[4])
{
	*NM = -1; *md_len = 0; *xb_len = 0;
	counts[0] = counts[1] = counts[2] = counts[3] = 0;
	{  /* a CIGAR of M/I/D ops consuming at most the read (ksw_global2 makes no other) */
		int64_t qsum = 0;
		for (int k = 0; k < n_cigar; ++k) {
			const int op = cigar[k] & 0xf;
			if (op == 0 || op == 1) qsum += cigar[k] >> 4;
			else if (op != 2) return -3;
		}
		if (n_cigar < 0 || qsum > l_query || rb < 0 || rb >= re || re > 2 * l_pac) return -3;
	}
	const int reverse = rb > l_pac ? 1 : 0;  /* bwa.c:212: strictly greater */
	const int is_crick = (rb >= crick_l && rb < l_pac) || rb >= crick_l + l_pac;
	const int ref_base = is_crick ? 2 : 1, methyl_base = is_crick ? 0 : 3;
	uint8_t *cs = (uint8_t *)malloc((size_t)(re - rb + 4));
	int64_t clen = orc_get_seq(l_pac, opac, rb - 2, re + 2, cs);
	if (!clen) { free(cs); return 1; }
	uint8_t *oq = (uint8_t *)malloc((size_t)(l_query > 0 ? l_query : 1));
	for (int i = 0; i < l_query; ++i) oq[i] = oquery[i];
	if (rb >= l_pac) {
		for (int64_t i = 0; i < clen >> 1; ++i) { uint8_t t = cs[i]; cs[i] = cs[clen - 1 - i]; cs[clen - 1 - i] = t; }
		for (int i = 0; i < l_query >> 1; ++i) { uint8_t t = oq[i]; oq[i] = oq[l_query - 1 - i]; oq[l_query - 1 - i] = t; }
	}
	const char *i2b = rb < l_pac ? "ACGTN" : "TGCAN";
	int ret = 0, ml = 0, xl = 0, x = 0, y = 0, u = 0, n_mm = 0, n_gap = 0, meth_pos = 0;
	int cgm = 0, cgu = 0, chm = 0, chu = 0;
#define CS(k) ((k) >= 0 && (k) < clen ? (int)cs[k] : (ret = -2, 4))
#define PUTW(v) do { ml = put_int(md, ml, md_cap, (v)); if (ml < 0) { ret = -1; goto out; } } while (0)
#define PUTC(c) do { if (ml + 1 > md_cap) { ret = -1; goto out; } md[ml++] = (char)(c); } while (0)
#define XPUTW(v) do { xl = put_int(xb, xl, xb_cap, (v)); if (xl < 0) { ret = -1; goto out; } } while (0)
#define XPUTC(c) do { if (xl + 1 > xb_cap) { ret = -1; goto out; } xb[xl++] = (char)(c); } while (0)
	for (int k = 0; k < n_cigar; ++k) {
		const int op = cigar[k] & 0xf, len = (int)(cigar[k] >> 4);
		if (op == 0) {
			for (int i = 0; i < len; ++i) {
				const int c = CS(y + i + 2);
				if (c == ref_base) {
					int b1, b2, b3;
					if (!reverse && !is_crick) b1 = CS(y + i + 2), b2 = CS(y + i + 3), b3 = CS(y + i + 4);
					else if (!reverse && is_crick) b1 = CS(y + i + 1), b2 = CS(y + i + 2), b3 = CS(y + i + 3);
					else if (reverse && !is_crick) b1 = CS(y + i + 2), b2 = CS(y + i + 1), b3 = CS(y + i);
					else b1 = CS(y + i + 3), b2 = CS(y + i + 2), b3 = CS(y + i + 1);
					const int o = oq[x + i];
					if (o == ref_base || o == methyl_base) {
						const char m = o == ref_base ? meth_context(b1, b2, b3, 1, &cgm, &chm)
						                             : meth_context(b1, b2, b3, 0, &cgu, &chu);
						if (meth_pos > 0) XPUTW(meth_pos);
						XPUTC(m);
						meth_pos = 0, ++u;
					} else {
						PUTW(u); PUTC(i2b[c]);
						++n_mm; ++meth_pos; u = 0;
					}
				} else if (oq[x + i] != c) {
					PUTW(u); PUTC(i2b[c]);
					++n_mm; ++meth_pos; u = 0;
				} else {
					++u; ++meth_pos;
				}
				if (ret == -2) goto out;
			}
			x += len; y += len;
		} else if (op == 2) {
			if (k > 0 && k < n_cigar - 1) {  /* not for a leading or trailing D */
				PUTW(u); PUTC('^');
				for (int i = 0; i < len; ++i) { const int c = CS(y + i + 2); PUTC(i2b[c]); }
				if (ret == -2) goto out;
				u = 0; n_gap += len;
			}
			y += len;
		} else if (op == 1) {
			x += len; n_gap += len; meth_pos += len;
		}
	}
	PUTW(u);
	if (meth_pos > 0) XPUTW(meth_pos);
	*NM = n_mm + n_gap;
	*md_len = ml; *xb_len = xl;
	counts[0] = cgm; counts[1] = cgu; counts[2] = chm; counts[3] = chu;
out:
#undef CS
#undef PUTW
#undef PUTC
#undef XPUTW
#undef XPUTC
	free(cs); free(oq);
	return ret;
}

/* bs_helpers.cpp:18-31: in-silico bisulfite conversion of ASCII bases.
 * pattern 0: C->T (Watson/read 1), pattern 1: G->A (Crick/read 2). */
void orc_bs_convert(char *seq, int pattern, int len)
{
	char from = pattern ? 'G' : 'C', to = pattern ? 'A' : 'T';
	for (int i = 0; i < len; ++i) if (seq[i] == from) seq[i] = to;
}

/* bntseq.c:48-65: ASCII -> 0..3 for ACGT (either case), '-' -> 5, else 4. */
int orc_nt4(int c)
{
	switch (c) {
	case 'A': case 'a': return 0;
	case 'C': case 'c': return 1;
	case 'G': case 'g': return 2;
	case 'T': case 't': return 3;
	case '-': return 5;
	default: return 4;
	}
}

/* ---------------- batch drivers (CPU baseline and bulk checks) ----------------
 * Tasks are held in the same structure-of-arrays layout as the product ABI:
 * concatenated 1-byte codes plus per-task offsets/lengths.  Threads pull chunks
 * of `chunk` tasks from a shared counter (dynamic scheduling, as kt_for does in
 * kthread.c:25-58).  `fn` selects the per-task routine so the same driver can
 * time this restatement or the compiled reference.
 */
typedef int (*orc_ext_fn)(int, const uint8_t *, int, const uint8_t *, int, const int8_t *,
                          int, int, int, int, int, int, int, int,
                          int *, int *, int *, int *, int *);

typedef struct {
	const int8_t *mat; int o_del, e_del, o_ins, e_ins, zdrop;
	int64_t n;
	const uint8_t *query; const uint64_t *q_off; const int32_t *qlen;
	const uint8_t *target; const uint64_t *t_off; const int32_t *tlen;
	const int32_t *h0, *w, *end_bonus;
	int32_t *score, *qle, *tle, *gtle, *gscore, *max_off; int64_t *cells;
	orc_ext_fn fn;
	int64_t next; int chunk;
	pthread_mutex_t mu;
} ext_job;

static void run_one(ext_job *J, int64_t t)
{
	int a, b, c, d, e; int64_t cl = 0;
	int s;
	if (J->fn)
		s = J->fn(J->qlen[t], J->query + J->q_off[t], J->tlen[t], J->target + J->t_off[t], 5, J->mat,
		          J->o_del, J->e_del, J->o_ins, J->e_ins, J->w[t], J->end_bonus[t], J->zdrop,
		          J->h0[t], &a, &b, &c, &d, &e);
	else
		s = orc_extend2(J->qlen[t], J->query + J->q_off[t], J->tlen[t], J->target + J->t_off[t], 5, J->mat,
		                J->o_del, J->e_del, J->o_ins, J->e_ins, J->w[t], J->end_bonus[t], J->zdrop,
		                J->h0[t], &a, &b, &c, &d, &e, &cl);
	if (J->score) J->score[t] = s;
	if (J->qle) J->qle[t] = a;
	if (J->tle) J->tle[t] = b;
	if (J->gtle) J->gtle[t] = c;
	if (J->gscore) J->gscore[t] = d;
	if (J->max_off) J->max_off[t] = e;
	if (J->cells) J->cells[t] = cl;
}

static void *ext_worker(void *arg)
{
	ext_job *J = (ext_job *)arg;
	for (;;) {
		pthread_mutex_lock(&J->mu);
		int64_t b = J->next; J->next += J->chunk;
		pthread_mutex_unlock(&J->mu);
		if (b >= J->n) break;
		int64_t e = b + J->chunk < J->n ? b + J->chunk : J->n;
		for (int64_t t = b; t < e; ++t) run_one(J, t);
	}
	return 0;
}

static double now_s(void)
{
	struct timespec ts; clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* Runs the batch on `nthreads` threads; returns wall seconds. fn NULL = orc_extend2. */
double orc_extend2_batch(void *fn, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int zdrop,
                         int64_t n, const uint8_t *query, const uint64_t *q_off, const int32_t *qlen,
                         const uint8_t *target, const uint64_t *t_off, const int32_t *tlen,
                         const int32_t *h0, const int32_t *w, const int32_t *end_bonus,
                         int32_t *score, int32_t *qle, int32_t *tle, int32_t *gtle, int32_t *gscore,
                         int32_t *max_off, int64_t *cells, int nthreads)
{
	ext_job J;
	memset(&J, 0, sizeof J);
	J.mat = mat; J.o_del = o_del; J.e_del = e_del; J.o_ins = o_ins; J.e_ins = e_ins; J.zdrop = zdrop;
	J.n = n; J.query = query; J.q_off = q_off; J.qlen = qlen; J.target = target; J.t_off = t_off; J.tlen = tlen;
	J.h0 = h0; J.w = w; J.end_bonus = end_bonus;
	J.score = score; J.qle = qle; J.tle = tle; J.gtle = gtle; J.gscore = gscore; J.max_off = max_off; J.cells = cells;
	J.fn = (orc_ext_fn)fn; J.chunk = 4096; J.next = 0;
	pthread_mutex_init(&J.mu, 0);
	if (nthreads < 1) nthreads = 1;
	if (nthreads > 256) nthreads = 256;
	pthread_t th[256];
	double t0 = now_s();
	for (int k = 1; k < nthreads; ++k) pthread_create(&th[k], 0, ext_worker, &J);
	ext_worker(&J);
	for (int k = 1; k < nthreads; ++k) pthread_join(th[k], 0);
	double dt = now_s() - t0;
	pthread_mutex_destroy(&J.mu);
	return dt;
}

/* Global (CIGAR) batch: per-task band w; CIGARs go to fixed slots of
 * `slot` uint32 per task (n_cigar[t] = -1 if the CIGAR did not fit).  Threads
 * pull dynamic chunks of 4096 tasks as the extension driver does; `fn` NULL
 * runs orc_global2, otherwise a routine with ksw_global2's signature
 * (ksw.h:84), e.g. the compiled reference.  Returns the wall-clock seconds. */
typedef int (*orc_glb_fn)(int, const uint8_t *, int, const uint8_t *, int, const int8_t *,
                          int, int, int, int, int, int *, uint32_t **);

typedef struct {
	const int8_t *mat; int o_del, e_del, o_ins, e_ins;
	int64_t n, slot;
	const uint8_t *query; const uint64_t *q_off; const int32_t *qlen;
	const uint8_t *target; const uint64_t *t_off; const int32_t *tlen; const int32_t *w;
	int32_t *score, *n_cigar; uint32_t *cigar;
	orc_glb_fn fn;
	int64_t next;
	pthread_mutex_t mu;
} glb_job;

static void *glb_worker(void *arg)
{
	glb_job *J = (glb_job *)arg;
	for (;;) {
		pthread_mutex_lock(&J->mu);
		int64_t b = J->next; J->next += 4096;
		pthread_mutex_unlock(&J->mu);
		if (b >= J->n) break;
		int64_t e = b + 4096 < J->n ? b + 4096 : J->n;
		for (int64_t t = b; t < e; ++t) {
			int nc = 0; uint32_t *cg = 0;
			int *pnc = J->cigar ? &nc : 0; uint32_t **pcg = J->cigar ? &cg : 0;
			int s = J->fn ? J->fn(J->qlen[t], J->query + J->q_off[t], J->tlen[t], J->target + J->t_off[t], 5,
			                      J->mat, J->o_del, J->e_del, J->o_ins, J->e_ins, J->w[t], pnc, pcg)
			              : orc_global2(J->qlen[t], J->query + J->q_off[t], J->tlen[t], J->target + J->t_off[t], 5,
			                            J->mat, J->o_del, J->e_del, J->o_ins, J->e_ins, J->w[t], pnc, pcg);
			J->score[t] = s;
			if (J->cigar) {
				if (nc <= J->slot) { memcpy(J->cigar + t * J->slot, cg, (size_t)nc * 4); J->n_cigar[t] = nc; }
				else J->n_cigar[t] = -1;
				free(cg);
			}
		}
	}
	return 0;
}

double orc_global2_batch(void *fn, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins,
                         int64_t n, const uint8_t *query, const uint64_t *q_off, const int32_t *qlen,
                         const uint8_t *target, const uint64_t *t_off, const int32_t *tlen,
                         const int32_t *w, int32_t *score, int32_t *n_cigar, uint32_t *cigar, int64_t slot,
                         int nthreads)
{
	glb_job J;
	memset(&J, 0, sizeof J);
	J.mat = mat; J.o_del = o_del; J.e_del = e_del; J.o_ins = o_ins; J.e_ins = e_ins;
	J.n = n; J.slot = slot; J.query = query; J.q_off = q_off; J.qlen = qlen;
	J.target = target; J.t_off = t_off; J.tlen = tlen; J.w = w;
	J.score = score; J.n_cigar = n_cigar; J.cigar = cigar; J.fn = (orc_glb_fn)fn; J.next = 0;
	pthread_mutex_init(&J.mu, 0);
	if (nthreads < 1) nthreads = 1;
	if (nthreads > 256) nthreads = 256;
	pthread_t th[256];
	double t0 = now_s();
	for (int k = 1; k < nthreads; ++k) pthread_create(&th[k], 0, glb_worker, &J);
	glb_worker(&J);
	for (int k = 1; k < nthreads; ++k) pthread_join(th[k], 0);
	double dt = now_s() - t0;
	pthread_mutex_destroy(&J.mu);
	return dt;
}

/* orc_gen_md over a batch on nthreads host threads (dynamic chunks of 4096
 * tasks); strings go to fixed slots md[t*md_slot], xb[t*xb_slot].  Returns the
 * wall-clock seconds. */
typedef struct {
	int64_t n, l_pac, crick_l, cigar_slot, md_slot, xb_slot;
	const uint8_t *opac, *oquery;
	const int64_t *rb, *re;
	const uint64_t *oq_off;
	const int32_t *l_query, *n_cigar;
	const uint32_t *cigar;
	char *md, *xb;
	int32_t *md_len, *xb_len, *nm, *meth, *status;
	int64_t next;
	pthread_mutex_t mu;
} md_job;

static void *md_worker(void *arg)
{
	md_job *J = (md_job *)arg;
	for (;;) {
		pthread_mutex_lock(&J->mu);
		const int64_t a = J->next;
		J->next += 4096;
		pthread_mutex_unlock(&J->mu);
		if (a >= J->n) break;
		const int64_t b = a + 4096 < J->n ? a + 4096 : J->n;
		for (int64_t t = a; t < b; ++t)
			J->status[t] = orc_gen_md(J->l_pac, J->opac, J->rb[t], J->re[t], J->crick_l, J->l_query[t],
			                          J->oquery + J->oq_off[t], J->n_cigar[t], J->cigar + t * J->cigar_slot,
			                          J->md + t * J->md_slot, (int)J->md_slot, &J->md_len[t],
			                          J->xb + t * J->xb_slot, (int)J->xb_slot, &J->xb_len[t], &J->nm[t],
			                          J->meth + 4 * t);
	}
	return 0;
}

double orc_gen_md_batch(int64_t n, int64_t l_pac, const uint8_t *opac, int64_t crick_l, const int64_t *rb,
                        const int64_t *re, const uint8_t *oquery, const uint64_t *oq_off, const int32_t *l_query,
                        const int32_t *n_cigar, const uint32_t *cigar, int64_t cigar_slot, char *md, int64_t md_slot,
                        int32_t *md_len, char *xb, int64_t xb_slot, int32_t *xb_len, int32_t *nm, int32_t *meth,
                        int32_t *status, int nthreads)
{
	md_job J;
	memset(&J, 0, sizeof J);
	J.n = n; J.l_pac = l_pac; J.crick_l = crick_l; J.cigar_slot = cigar_slot; J.md_slot = md_slot; J.xb_slot = xb_slot;
	J.opac = opac; J.oquery = oquery; J.rb = rb; J.re = re; J.oq_off = oq_off; J.l_query = l_query;
	J.n_cigar = n_cigar; J.cigar = cigar; J.md = md; J.xb = xb; J.md_len = md_len; J.xb_len = xb_len;
	J.nm = nm; J.meth = meth; J.status = status;
	pthread_mutex_init(&J.mu, 0);
	if (nthreads < 1) nthreads = 1;
	if (nthreads > 256) nthreads = 256;
	pthread_t th[256];
	double t0 = now_s();
	for (int k = 1; k < nthreads; ++k) pthread_create(&th[k], 0, md_worker, &J);
	md_worker(&J);
	for (int k = 1; k < nthreads; ++k) pthread_join(th[k], 0);
	double dt = now_s() - t0;
	pthread_mutex_destroy(&J.mu);
	return dt;
}

/* bwamem.c:1237-1263: the tail of mem_reg2aln.  From the CIGAR bwa_gen_cigar2
 * returned for region [rb, re) x [qb, qe) of a read of l_seq bases: position
 * = bns_depos(rb < l_pac ? rb : re - 1); a leading D is dropped and moves the
 * position, else a trailing D is dropped; then clip ops (len<<4|3) for the
 * unaligned read ends, 5' end first (qb forward, l_seq - qe reverse).  Writes
 * *pos, *is_rev and out[]; returns the op count, -1 for inputs outside the
 * reference's (n_cigar < 0, rb < 0, re <= rb, re > 2*l_pac, qb < 0, qe < qb,
 * qe > l_seq), -2 when out_cap is too small. */
int orc_reg2aln_cigar(int64_t l_pac, int64_t rb, int64_t re, int qb, int qe, int l_seq, int n_cigar,
                      const uint32_t *cigar, uint32_t *out, int out_cap, int64_t *pos, int *is_rev)
{
	int64_t p;
	int n = n_cigar, i, k = 0, clip5 = 0, clip3 = 0, skip_first = 0;
	*pos = -1; *is_rev = 0;
	if (n_cigar < 0 || rb < 0 || re <= rb || re > 2 * l_pac || qb < 0 || qe < qb || qe > l_seq) return -1;
	p = rb < l_pac ? rb : re - 1;            /* bns_depos */
	*is_rev = p >= l_pac;
	*pos = *is_rev ? (l_pac << 1) - 1 - p : p;
	if (n > 0) {                              /* squeeze out a leading or a trailing deletion */
		if ((cigar[0] & 0xf) == 2) { *pos += cigar[0] >> 4; skip_first = 1; --n; }
		else if ((cigar[n - 1] & 0xf) == 2) --n;
	}
	if (qb != 0 || qe != l_seq) {             /* clipping */
		clip5 = *is_rev ? l_seq - qe : qb;
		clip3 = *is_rev ? qb : l_seq - qe;
	}
	if ((clip5 > 0) + n + (clip3 > 0) > out_cap) return -2;
	if (clip5) out[k++] = (uint32_t)clip5 << 4 | 3;
	for (i = 0; i < n; ++i) out[k++] = cigar[skip_first + i];
	if (clip3) out[k++] = (uint32_t)clip3 << 4 | 3;
	return k;
}
