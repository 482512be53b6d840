/*
 * bwa_bsx.c -- links BSBolt's reference aligner to libbsx's scalar drop-ins.
 *
 * oracle/Makefile compiles bsbolt/External/BWA from its own sources and links
 * those unmodified objects with this file and libbsx.so under
 * -Wl,--wrap=ksw_extend2,--wrap=ksw_global2: every call mem_chain2aln makes to
 * ksw_extend2 (bwamem.c:735,763) and bwa_gen_cigar2 makes to ksw_global2
 * (bwa.c:248) goes to bsx_ksw_extend2 / bsx_ksw_global2 on the GPU.  Where a
 * drop-in returns INT32_MIN (an input it does not take, see include/bsx.h)
 * the call goes to the reference's own function instead, as a maintainer's
 * integration must do.  At exit the counts of GPU and reference calls go to
 * stderr as one line:  bsx: ksw_extend2 gpu=N cpu=N ksw_global2 gpu=N cpu=N
 *
 * tests/test_gpu_bwa_link.py runs this aligner and the unlinked one on the same
 * reads and compares their SAM output.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../include/bsx.h"

int __real_ksw_extend2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int zdrop, int h0,
                       int *qle, int *tle, int *gtle, int *gscore, int *max_off);
int __real_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int *n_cigar, uint32_t **cigar);

static long n_ext_gpu, n_ext_cpu, n_glb_gpu, n_glb_cpu;

static void report(void) {
	fprintf(stderr, "bsx: ksw_extend2 gpu=%ld cpu=%ld ksw_global2 gpu=%ld cpu=%ld\n", n_ext_gpu, n_ext_cpu,
	        n_glb_gpu, n_glb_cpu);
}

static void once(void) {
	static int done = 0;
	if (!done) { done = 1; atexit(report); }
}

int __wrap_ksw_extend2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int zdrop, int h0,
                       int *qle, int *tle, int *gtle, int *gscore, int *max_off) {
	once();
	int s = bsx_ksw_extend2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0,
	                        qle, tle, gtle, gscore, max_off);
	if (s != INT32_MIN) { ++n_ext_gpu; return s; }
	++n_ext_cpu;
	return __real_ksw_extend2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop,
	                          h0, qle, tle, gtle, gscore, max_off);
}

int __wrap_ksw_global2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int m, const int8_t *mat,
                       int o_del, int e_del, int o_ins, int e_ins, int w, int *n_cigar, uint32_t **cigar) {
	once();
	int s = bsx_ksw_global2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, n_cigar, cigar);
	if (s != INT32_MIN) { ++n_glb_gpu; return s; }
	++n_glb_cpu;
	return __real_ksw_global2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w, n_cigar, cigar);
}
