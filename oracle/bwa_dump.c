/*
 * bwa_dump.c -- call recorder linked into the reference aligner built by
 * oracle/Makefile (test infrastructure: it only writes golden vectors).
 *
 * Linked with -Wl,--wrap=ksw_extend2,--wrap=ksw_global2,--wrap=bwa_gen_cigar2,
 * so every cross-object call the reference aligner makes to
 *   ksw_extend2    (ksw.h:108; callers bwamem.c:735,763)
 *   ksw_global2    (ksw.h:84;  caller bwa.c:248)
 *   bwa_gen_cigar2 (bwa.h:49;  callers bwamem.c:432,1228)
 * lands here first; the wrapper runs the reference's own function
 * (__real_*) unchanged and appends one record with its inputs and outputs to
 * the file named by $BSX_DUMP.  Records are little-endian, tagged by one byte:
 *
 *  'E' ksw_extend2: i32 qlen, tlen, m, o_del, e_del, o_ins, e_ins, w,
 *      end_bonus, zdrop, h0; i8 mat[m*m]; u8 query[qlen]; u8 target[tlen];
 *      i32 score, qle, tle, gtle, gscore, max_off
 *  'G' ksw_global2: i32 qlen, tlen, m, o_del, e_del, o_ins, e_ins, w;
 *      i8 mat[m*m]; u8 query[qlen]; u8 target[tlen]; i32 score, n_cigar
 *      (-1 when the caller asked for the score only); u32 cigar[n_cigar]
 *  'C' bwa_gen_cigar2: i32 w_, l_query; i64 l_pac, rb, re, crick_l;
 *      i32 want_cigar (1 when n_cigar is non-NULL, i.e. the mem_reg2aln call);
 *      u8 query[l_query]; i32 score, n_cigar, NM; u32 cigar[n_cigar];
 *      i32 tail_len; char tail[tail_len] (the MD:Z and XB:Z text bwa.c:250-343
 *      appends after the CIGAR; empty when NM < 0); i32 cg_meth, cg_unmeth,
 *      ch_meth, ch_unmeth
 *
 * The aligner is run with one thread (-t 1), so records come in call order.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int __real_ksw_extend2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int zdrop, int h0,
                       int *qle, int *tle, int *gtle, int *gscore, int *max_off);
int __real_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int *n_cigar, uint32_t **cigar);
uint32_t *__real_bwa_gen_cigar2(const int8_t mat[25], int o_del, int e_del, int o_ins, int e_ins, int w_,
                                int64_t l_pac, const uint8_t *pac, const uint8_t *opac, int l_query,
                                uint8_t *query, int64_t rb, int64_t re, int *score, int *n_cigar, int *NM,
                                int64_t crick_l, uint8_t *oquery, int *cg_meth, int *cg_unmeth, int *ch_meth,
                                int *ch_unmeth);

static FILE *dump_file(void) {
	static FILE *f = NULL;
	static int tried = 0;
	if (!tried) {
		const char *p = getenv("BSX_DUMP");
		tried = 1;
		if (p && *p) f = fopen(p, "wb");
		if (p && *p && !f) { fprintf(stderr, "bwa_dump: cannot open %s\n", p); exit(1); }
	}
	return f;
}

static void put(FILE *f, const void *p, size_t n) {
	if (n && fwrite(p, 1, n, f) != n) { fprintf(stderr, "bwa_dump: write failed\n"); exit(1); }
}

static void put_i32(FILE *f, int32_t v) { put(f, &v, 4); }
static void put_i64(FILE *f, int64_t v) { put(f, &v, 8); }

int __wrap_ksw_extend2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int zdrop, int h0,
                       int *qle, int *tle, int *gtle, int *gscore, int *max_off) {
	int o[5] = {0, 0, 0, 0, 0};
	int s = __real_ksw_extend2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop,
	                           h0, &o[0], &o[1], &o[2], &o[3], &o[4]);
	FILE *f = dump_file();
	if (f) {
		const int32_t in[11] = {qlen, tlen, m, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0};
		put(f, "E", 1);
		put(f, in, sizeof in);
		put(f, mat, (size_t)m * m);
		put(f, query, (size_t)qlen);
		put(f, target, (size_t)tlen);
		put_i32(f, s);
		put(f, o, sizeof o);
	}
	if (qle) *qle = o[0];
	if (tle) *tle = o[1];
	if (gtle) *gtle = o[2];
	if (gscore) *gscore = o[3];
	if (max_off) *max_off = o[4];
	return s;
}

int __wrap_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int *n_cigar, uint32_t **cigar) {
	int s = __real_ksw_global2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, n_cigar, cigar);
	FILE *f = dump_file();
	if (f) {
		const int32_t in[8] = {qlen, tlen, m, o_del, e_del, o_ins, e_ins, w};
		const int want = n_cigar && cigar;
		put(f, "G", 1);
		put(f, in, sizeof in);
		put(f, mat, (size_t)m * m);
		put(f, query, (size_t)qlen);
		put(f, target, (size_t)tlen);
		put_i32(f, s);
		put_i32(f, want ? *n_cigar : -1);
		if (want) put(f, *cigar, 4 * (size_t)*n_cigar);
	}
	return s;
}

uint32_t *__wrap_bwa_gen_cigar2(const int8_t mat[25], int o_del, int e_del, int o_ins, int e_ins, int w_,
                                int64_t l_pac, const uint8_t *pac, const uint8_t *opac, int l_query,
                                uint8_t *query, int64_t rb, int64_t re, int *score, int *n_cigar, int *NM,
                                int64_t crick_l, uint8_t *oquery, int *cg_meth, int *cg_unmeth, int *ch_meth,
                                int *ch_unmeth) {
	uint8_t *qcopy = NULL;
	FILE *f = dump_file();
	if (f && l_query > 0) {  /* the query as passed: bwa_gen_cigar2 may reverse it in place */
		qcopy = (uint8_t *)malloc((size_t)l_query);
		memcpy(qcopy, query, (size_t)l_query);
	}
	uint32_t *c = __real_bwa_gen_cigar2(mat, o_del, e_del, o_ins, e_ins, w_, l_pac, pac, opac, l_query, query, rb,
	                                    re, score, n_cigar, NM, crick_l, oquery, cg_meth, cg_unmeth, ch_meth,
	                                    ch_unmeth);
	if (f) {
		const int want = n_cigar != NULL;
		const int nc = want ? *n_cigar : 0;
		put(f, "C", 1);
		put_i32(f, w_);
		put_i32(f, l_query);
		put_i64(f, l_pac);
		put_i64(f, rb);
		put_i64(f, re);
		put_i64(f, crick_l);
		put_i32(f, want);
		put(f, qcopy, l_query > 0 ? (size_t)l_query : 0);
		put_i32(f, score ? *score : 0);
		put_i32(f, nc);
		put_i32(f, NM ? *NM : -1);
		if (c && nc > 0) put(f, c, 4 * (size_t)nc);
		else if (nc > 0) { uint32_t z = 0; for (int i = 0; i < nc; ++i) put(f, &z, 4); }
		/* bwa.c:250-343 appends the MD:Z and XB:Z text after the CIGAR words and
		 * sets NM; where it does not (NM stays -1) the buffer ends at the CIGAR */
		const char *tail = (c && want && NM && *NM >= 0) ? (const char *)(c + nc) : "";
		const int32_t tl = (int32_t)strlen(tail);
		put_i32(f, tl);
		put(f, tail, (size_t)tl);
		put_i32(f, cg_meth ? *cg_meth : 0);
		put_i32(f, cg_unmeth ? *cg_unmeth : 0);
		put_i32(f, ch_meth ? *ch_meth : 0);
		put_i32(f, ch_unmeth ? *ch_unmeth : 0);
	}
	free(qcopy);
	return c;
}
