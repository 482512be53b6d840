/*
 * band_oracle.c -- CPU statement of the learned-band form's integer parts, used
 * ONLY as the test oracle (tests/ and the smoke check).  Nothing in the product
 * library links or calls this file.
 *
 *   orc_narrow_extend   the narrow-band run of ksw_extend2 (ksw.c:380-479 with
 *                       the band capped at wn) and the exact certificate that
 *                       its six outputs and cell count equal the run at the
 *                       task's own band: the statement the GPU narrow kernel
 *                       and its labels are checked against
 *   orc_band_features   the 24 per-task predictor features
 *
 * The narrow run is the restatement of ksw.c:380-479 (orc_extend2) with the cap
 * i-wn <= j <= i+wn applied to beg/end; the certificate checks, per row,
 *   left : the column that leaves the narrow band is not the first non-zero
 *          column the reference's zero scan would keep (ksw.c:466-467);
 *   ramp : the first column past the narrow end has no first-row insertion
 *          value (ksw.c:395-397) while the wide band reaches past it;
 *   F    : no insertion value leaves the narrow band to the right;
 * and replays the wide band's cell count and end-to-end column (ksw.c:449-453).
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

/* Returns 1 when the certificate accepts (out[0..5] = score qle tle gtle gscore
 * max_off, *cells = cells of the wide run), 0 when it rejects (*why = 1 left,
 * 2 ramp, 3 F), -1 on bad input.  wn >= the clamped band runs the wide band
 * itself (always accepted). */
int orc_narrow_extend(int qlen, const uint8_t *query, int tlen, const uint8_t *target, const int8_t *mat,
                      int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus, int zdrop, int h0,
                      int wn, int *out, int64_t *cells_out, int *why)
{
	if (qlen < 1 || h0 < 1 || tlen < 0 || wn < 0) return -1;
	const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;
	const int W = clampw(w, qlen, mat, end_bonus, o_del, e_del, o_ins, e_ins);
	if (wn > W) wn = W;
	int32_t *hd = calloc(qlen + 2, 4), *ev = calloc(qlen + 2, 4);
	int8_t *prof = malloc((size_t)qlen * 5 + 1);
	for (int k = 0; k < 5; ++k)
		for (int j = 0; j < qlen; ++j) prof[k * qlen + j] = mat[k * 5 + query[j]];
	hd[0] = h0;
	hd[1] = h0 > oe_ins ? h0 - oe_ins : 0;
	for (int j = 2; j <= qlen && hd[j - 1] > e_ins; ++j) hd[j] = hd[j - 1] - e_ins;
	int best = h0, best_i = -1, best_j = -1, gbest = -1, gbest_i = -1, moff = 0;
	int ba = 0, ea = qlen, ok = 1;
	int64_t cells = 0;
	*why = 0;
	for (int i = 0; i < tlen; ++i) {
		const int8_t *srow = prof + target[i] * qlen;
		int bW = ba < i - W ? i - W : ba, eW = ea;
		if (eW > i + W + 1) eW = i + W + 1;
		if (eW > qlen) eW = qlen;
		int bN = ba < i - wn ? i - wn : ba, eN = ea;
		if (eN > i + wn + 1) eN = i + wn + 1;
		if (eN > qlen) eN = qlen;
		if (bN > bW) { /* left: column i-1-wn leaves the narrow band */
			const int j = i - 1 - wn;
			if (j >= ba && j < bN && (hd[j] != 0 || ev[j] != 0)) { ok = 0; *why = 1; break; }
		}
		if (eW > eN) { /* ramp past the narrow end */
			const int j = eN;
			const int g = j == 0 ? h0 : h0 - oe_ins - (j - 1) * e_ins;
			if (g > 0) { ok = 0; *why = 2; break; }
		}
		int h1 = 0;
		if (bN == 0) { h1 = h0 - (o_del + e_del * (i + 1)); if (h1 < 0) h1 = 0; }
		int f = 0, m = 0, mj = -1;
		for (int j = bN; j < eN; ++j) {
			const int M0 = hd[j];
			int e = ev[j];
			hd[j] = h1;
			const int M = M0 ? M0 + srow[j] : 0;
			int h = M > e ? M : e;
			h = h > f ? h : f;
			h1 = h;
			mj = m > h ? mj : j;
			m = m > h ? m : h;
			int t = M - oe_del; t = t > 0 ? t : 0;
			e -= e_del; e = e > t ? e : t;
			ev[j] = e;
			t = M - oe_ins; t = t > 0 ? t : 0;
			f -= e_ins; f = f > t ? f : t;
		}
		if (eW > eN && eN > bN && f != 0) { ok = 0; *why = 3; break; } /* F leaves the band */
		hd[eN] = h1;
		ev[eN] = 0;
		/* the wide run's cell count and end-to-end column */
		cells += eW > bW ? eW - bW : 0;
		int h1pW = 0;
		if (bW == 0) { h1pW = h0 - (o_del + e_del * (i + 1)); if (h1pW < 0) h1pW = 0; }
		const int h1W = eW > bW ? ((eW == eN && eN > bN) ? h1 : 0) : h1pW;
		const int jW = eW > bW ? eW : bW;
		if (jW == qlen) { gbest_i = gbest > h1W ? gbest_i : i; gbest = gbest > h1W ? gbest : h1W; }
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
		int k = bN;
		while (k < eN && hd[k] == 0 && ev[k] == 0) ++k;
		ba = k;
		k = eN;
		while (k >= ba && hd[k] == 0 && ev[k] == 0) --k;
		ea = k + 2 < qlen ? k + 2 : qlen;
	}
	free(hd);
	free(ev);
	free(prof);
	out[0] = best; out[1] = best_j + 1; out[2] = best_i + 1; out[3] = gbest_i + 1; out[4] = gbest; out[5] = moff;
	*cells_out = cells;
	return ok;
}

/* Band classes 0..3 = w 8, 16, 32, 64; the label of a task is the smallest
 * class whose narrow run the certificate accepts (or whose width is not below
 * the task's clamped band), 4 if none. */
void orc_band_labels(int64_t n, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins, int zdrop,
                     const uint8_t *query, const uint64_t *q_off, const int32_t *qlen, const uint8_t *target,
                     const uint64_t *t_off, const int32_t *tlen, const int32_t *h0, const int32_t *w,
                     const int32_t *end_bonus, int qmax, uint8_t *labels)
{
	for (int64_t t = 0; t < n; ++t) {
		int lab = 4;
		if (qlen[t] <= qmax) {
			for (int c = 0; c < 4; ++c) {
				int out[6], why;
				int64_t cl;
				const int wn = 8 << c;
				const int r = orc_narrow_extend(qlen[t], query + q_off[t], tlen[t], target + t_off[t], mat, o_del,
				                                e_del, o_ins, e_ins, w[t], end_bonus[t], zdrop, h0[t], wn, out,
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
