/*
 * band_oracle.c -- CPU statement of the learned-band form's integer parts, used
 * ONLY as the test oracle (tests/ and the smoke check).  Nothing in the product
 * library links or calls this file.
 *
 *   orc_narrow_extend   the reference ksw_extend2 run (ksw.c:380-479) watched
 *                       through a band-following frame of WF columns: the frame
 *                       start f0 <= beg advances toward beg by at most two
 *                       columns per row, and the run is rejected when the band
 *                       reaches past the frame (end > f0+WF-1 = eF) unless
 *                       H(i, eF) = 0 and the first-row insertion value at eF+1
 *                       (ksw.c:395-397) is 0 -- the statement the GPU frame
 *                       kernel and its labels are checked against
 *   orc_band_labels     smallest frame class whose run is accepted
 *   orc_band_features   the 24 per-task predictor features
 *
 * The run computes the reference itself, so an accepted result is the
 * reference's by construction.  It also checks the fact the GPU kernel relies
 * on: every column entering the frame holds its first-row value (returns -2
 * otherwise, which the tests treat as a failure of the argument).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static int clampw(int w, int qlen, const int8_t *mat, int eb, int od, int ed, int oi, int ei)
{
	int best = 0;
	for (int k = 0; k < 25; ++k) if (mat[k] > best) best = mat[k];
	int li = (int)((double)(qlen * best + eb - oi) / ei + 1.);
	if (li < 1) li = 1;
	if (w > li) w = li;
	int ld = (int)((double)(qlen * best + eb - od) / ed + 1.);
	if (ld < 1) ld = 1;
	if (w > ld) w = ld;
	return w;
}

static int ramp(int j, int h0, int oe_ins, int e_ins)
{
	const int g = j == 0 ? h0 : h0 - oe_ins - (j - 1) * e_ins;
	return g > 0 ? g : 0;
}

/* Returns 1 when the frame run is accepted (out[0..5] = score qle tle gtle
 * gscore max_off, *cells = cell count), 0 when rejected (*why = 1: the band
 * spills a non-zero cell past the frame, 2: first-row insertion value past the
 * frame), -1 on bad input, -2 if an entering column's state is not its
 * first-row value. */
int orc_narrow_extend(int qlen, const uint8_t *query, int tlen, const uint8_t *target, const int8_t *mat,
                      int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int zdrop, int h0,
                      int wf, int *out, int64_t *cells_out, int *why)
{
	if (qlen < 1 || h0 < 1 || tlen < 0 || wf < 2) return -1;
	const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
	w = clampw(w, qlen, mat, end_bonus, o_del, e_del, o_ins, e_ins);
	int32_t *hd = calloc(qlen + 2, 4), *ev = calloc(qlen + 2, 4);
	int8_t *prof = malloc((size_t)qlen * 5 + 1);
	for (int k = 0; k < 5; ++k)
		for (int j = 0; j < qlen; ++j) prof[k * qlen + j] = mat[k * 5 + query[j]];
	hd[0] = h0;
	hd[1] = h0 > oe_ins ? h0 - oe_ins : 0;
	for (int j = 2; j <= qlen && hd[j - 1] > e_ins; ++j) hd[j] = hd[j - 1] - e_ins;
	int best = h0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
	int beg = 0, end = qlen, f0 = 0, ok = 1;
	int64_t cells = 0;
	*why = 0;
	for (int i = 0; i < tlen; ++i) {
		const int8_t *srow = prof + target[i] * qlen;
		if (beg < i - w) beg = i - w;
		if (end > i + w + 1) end = i + w + 1;
		if (end > qlen) end = qlen;
		const int eF = end < f0 + wf - 1 ? end : f0 + wf - 1;
		const int spill = end > eF;
		if (spill && ramp(eF + 1, h0, oe_ins, e_ins) > 0) { ok = 0; *why = 2; break; }
		int h1 = 0;
		if (beg == 0) { h1 = h0 - (o_del + e_del * (i + 1)); if (h1 < 0) h1 = 0; }
		int f = 0, m = 0, mj = -1, h_eF = 0;
		for (int j = beg; j < end; ++j) {
			const int M0 = hd[j];
			int e = ev[j];
			hd[j] = h1;
			const int M = M0 ? M0 + srow[j] : 0;
			int h = M > e ? M : e;
			h = h > f ? h : f;
			h1 = h;
			if (j == eF) h_eF = h;
			mj = m > h ? mj : j;
			m = m > h ? m : h;
			int t = M - oe_del; t = t > 0 ? t : 0;
			e -= e_del; e = e > t ? e : t;
			ev[j] = e;
			t = M - oe_ins; t = t > 0 ? t : 0;
			f -= e_ins; f = f > t ? f : t;
		}
		if (spill && h_eF != 0) { ok = 0; *why = 1; break; }
		hd[end] = h1;
		ev[end] = 0;
		cells += end > beg ? end - beg : 0;
		const int jafter = end > beg ? end : beg;
		if (jafter == qlen) { gbest_i = gbest > h1 ? gbest_i : i; gbest = gbest > h1 ? gbest : h1; }
		if (m == 0) break;
		if (m > best) {
			best = m; best_i = i; best_j = mj;
			const int off = mj > i ? mj - i : i - mj;
			moff = moff > off ? moff : off;
		} else if (zdrop > 0) {
			const int di = i - best_i, dj = mj - best_j;
			const int drop = di > dj ? best - m - (di - dj) * e_del : best - m - (dj - di) * e_ins;
			if (drop > zdrop) break;
		}
		int k = beg;
		while (k < end && hd[k] == 0 && ev[k] == 0) ++k;
		beg = k;
		k = end;
		while (k >= beg && hd[k] == 0 && ev[k] == 0) --k;
		end = k + 2 < qlen ? k + 2 : qlen;
		/* the frame advances toward beg; the entering columns must hold their first-row value */
		int s = beg - f0;
		if (s > 2) s = 2;
		for (int r = 0; r < s; ++r) {
			++f0;
			const int jin = f0 + wf - 1;
			if (jin <= qlen && (hd[jin] != ramp(jin, h0, oe_ins, e_ins) || ev[jin] != 0)) ok = -2;
		}
		if (ok == -2) break;
	}
	free(hd);
	free(ev);
	free(prof);
	out[0] = best; out[1] = best_j + 1; out[2] = best_i + 1; out[3] = gbest_i + 1; out[4] = gbest; out[5] = moff;
	*cells_out = cells;
	return ok;
}

/* Frame classes 0..3 = 24, 40, 80, 144 columns; the label of a task is the
 * smallest class whose frame run is accepted, 4 if none (or qlen > qmax). */
void orc_band_labels(int64_t n, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int zdrop,
                     const uint8_t *query, const uint64_t *q_off, const int32_t *qlen, const uint8_t *target,
                     const uint64_t *t_off, const int32_t *tlen, const int32_t *h0, const int32_t *w,
                     const int32_t *end_bonus, int qmax, uint8_t *labels)
{
	static const int frames[4] = {24, 40, 80, 144};
	for (int64_t t = 0; t < n; ++t) {
		int lab = 4;
		if (qlen[t] <= qmax) {
			for (int c = 0; c < 4; ++c) {
				int out[6], why;
				int64_t cl;
				const int r = orc_narrow_extend(qlen[t], query + q_off[t], tlen[t], target + t_off[t], mat, o_del,
				                                e_del, o_ins, e_ins, w[t], end_bonus[t], zdrop, h0[t], frames[c], out,
				                                &cl, &why);
				if (r == 1) { lab = c; break; }
			}
		}
		labels[t] = (uint8_t)lab;
	}
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

/* The 24 predictor features of one task, each an integer times a power of two
 * (so exactly representable in FP32).  Ungapped scans cover the first
 * K = min(qlen, scan) query bases on the diagonals d = -2..2 (query p against
 * target p+d, pairs outside the target skipped): running score S_d and its
 * running maximum MX_d (starting at 0) with the substitution matrix. */
void orc_band_features(int qlen, const uint8_t *q, int tlen, const uint8_t *tg, int h0, int w, int eb,
                       const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int scan, float *f)
{
	const int oe_ins = o_ins + e_ins, oe_del = o_del + e_del;
	const int K = qlen < scan ? qlen : scan;
	int S[5] = {0}, MX[5] = {0};
	int arg0 = 0, match0 = 0, mism0 = 0, first_mm = K, match16 = 0, qT = 0, qN = 0, gc = 0, cpg = 0;
	for (int p = 0; p < K; ++p) {
		const int qc = q[p];
		qT += qc == 3;
		qN += qc > 3;
		for (int d = -2; d <= 2; ++d) {
			const int r = p + d;
			if (r < 0 || r >= tlen) continue;
			S[d + 2] += mat[tg[r] * 5 + qc];
			if (S[d + 2] > MX[d + 2]) MX[d + 2] = S[d + 2];
		}
		if (p < tlen) {
			const int tc = tg[p];
			const int eq = qc == tc && qc < 4;
			match0 += eq;
			mism0 += !eq;
			match16 += eq && p < 16;
			if (!eq && first_mm == K) first_mm = p;
			if (S[2] == MX[2] && S[2] > 0) arg0 = p + 1;
			gc += tc == 1 || tc == 2;
			cpg += tc == 1 && p + 1 < tlen && tg[p + 1] == 2;
		}
	}
	const int wcl = clampw(w, qlen, mat, eb, o_del, e_del, o_ins, e_ins);
	const int rin = h0 > oe_ins ? (h0 - oe_ins) / e_ins : 0;
	const int rde = h0 > oe_del ? (h0 - oe_del) / e_del : 0;
	f[0] = clampi(qlen, 0, 4096) / 128.f;
	f[1] = clampi(tlen, 0, 4096) / 256.f;
	f[2] = clampi(h0, 0, 4096) / 64.f;
	f[3] = clampi(wcl, 0, 4096) / 128.f;
	f[4] = clampi(eb, -256, 256) / 32.f;
	f[5] = clampi(rin, 0, 4096) / 64.f;
	f[6] = clampi(rde, 0, 4096) / 64.f;
	f[7] = clampi(rin - qlen, -4096, 4096) / 64.f;
	f[8] = match0 / 64.f;
	f[9] = MX[2] / 64.f;
	f[10] = arg0 / 128.f;
	f[11] = clampi(S[2], -256, 256) / 64.f;
	f[12] = first_mm / 128.f;
	f[13] = match16 / 16.f;
	f[14] = MX[3] / 64.f;
	f[15] = MX[4] / 64.f;
	f[16] = MX[1] / 64.f;
	f[17] = MX[0] / 64.f;
	f[18] = clampi(MX[2] - S[2], 0, 1024) / 64.f;
	f[19] = mism0 / 64.f;
	f[20] = qT / 128.f;
	f[21] = gc / 128.f;
	f[22] = cpg / 32.f;
	f[23] = qN / 8.f;
}

void orc_band_features_batch(int64_t n, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins,
                             const uint8_t *query, const uint64_t *q_off, const int32_t *qlen,
                             const uint8_t *target, const uint64_t *t_off, const int32_t *tlen, const int32_t *h0,
                             const int32_t *w, const int32_t *end_bonus, int scan, float *f /* [n][24] */)
{
	for (int64_t t = 0; t < n; ++t)
		orc_band_features(qlen[t], query + q_off[t], tlen[t], target + t_off[t], h0[t], w[t], end_bonus[t], mat,
		                  o_del, e_del, o_ins, e_ins, scan, f + t * 24);
}
