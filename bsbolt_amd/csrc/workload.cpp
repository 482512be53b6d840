// workload.cpp -- host-side producer of extension/CIGAR task batches (libbsxwl.so).
//
// It builds the inputs the reference aligner hands to ksw_extend2/ksw_global2:
//  * a synthetic genome and its bisulfite index layout: the forward-only 2-bit
//    pac holds the Watson copy converted C->T followed by the Crick copy
//    converted G->A (bntseq.c:239-300, bwtindex.c:256-321); positions >= l_pac
//    of the doubled reference are reverse complements (bntseq.c:413-437);
//  * simulated directional bisulfite reads (methylation-dependent C->T on the
//    sampled strand, substitutions, indels), converted in silico the way the
//    aligner does: read 1 C->T, read 2 G->A (bs_helpers.cpp:18-31, bwa.c:103,116);
//  * seeds: maximal exact matches (>= min_seed_len) between the converted read
//    and the converted reference along the read's true alignment, standing in
//    for the SMEMs of the FM-index (which is upstream of the seed-extend path);
//  * mem_chain2aln (bwamem.c:636-790) as a round-based state machine over many
//    reads: each round emits at most one ksw_extend2 task per chain (left flank,
//    its w<<1 retry, right flank with h0 = left score, its retry), the caller
//    runs the round on the GPU and feeds the results back.  The seed
//    containment test (bwamem.c:675-710) and the local/to-end choice
//    (bwamem.c:740-746,768-774) run on the results exactly as in the reference.
//  * CIGAR tasks as mem_reg2aln/bwa_gen_cigar2 build them (bwamem.c:1197-1233,
//    bwa.c:199-248), also round based (the band-doubling retry needs scores).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

struct Rng {  // xoshiro256** seeded by splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; ++i) {
      seed += 0x9e3779b97f4a7c15ULL;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  int64_t below(int64_t n) { return (int64_t)(uni() * (double)n); }
  double normal() {
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int geometric(double p) {  // support 1, 2, ...
    double u = uni();
    if (u < 1e-300) u = 1e-300;
    return 1 + (int)std::floor(std::log(u) / std::log(1.0 - p));
  }
};

inline int pac_get(const uint8_t* pac, int64_t l) { return pac[l >> 2] >> ((~l & 3) << 1) & 3; }
inline void pac_set(uint8_t* pac, int64_t l, int c) { pac[l >> 2] |= c << ((~l & 3) << 1); }

// Base of the doubled reference at position k (bntseq.c:427-434).
inline int ref_base(const uint8_t* pac, int64_t l_pac, int64_t k) {
  return k >= l_pac ? 3 - pac_get(pac, 2 * l_pac - 1 - k) : pac_get(pac, k);
}

struct Opt {  // the subset of mem_opt_t the seed-extend path reads (bwamem.h:26-62)
  int a, b, o_del, e_del, o_ins, e_ins, pen_clip5, pen_clip3, w, zdrop, min_seed_len;
};

// bwamem.c:625-632
inline int cal_max_gap(const Opt& o, int qlen) {
  int ld = (int)((double)(qlen * o.a - o.o_del) / o.e_del + 1.);
  int li = (int)((double)(qlen * o.a - o.o_ins) / o.e_ins + 1.);
  int l = ld > li ? ld : li;
  l = l > 1 ? l : 1;
  return l < o.w << 1 ? l : o.w << 1;
}

struct Seed { int64_t rbeg; int32_t qbeg, len, score; };
struct Region {  // mem_alnreg_t fields the path produces (bwamem.h:64-82)
  int64_t rb, re;
  int32_t qb, qe, score, truesc, w, seedcov, seedlen0;
};

struct Chain {
  int read = 0;
  int wait_for = -1;  // previous chain of the same read: chains of a read run in order
  std::vector<Seed> seeds;
  int64_t rmax0 = 0, rmax1 = 0;
  std::vector<uint8_t> rseq;
  std::vector<uint64_t> srt;
  int k = 0;        // current position in srt (descending)
  int phase = 0;    // 0 pick seed, 1 left, 2 right, 3 done
  int tryi = 0, prev = -1, sc0 = 0, aw0 = 0, aw1 = 0;
  Region cur{};
  int last_out[6]{};
};

struct Read { std::vector<uint8_t> seq, oseq; };  // converted as aligned; as sequenced

}  // namespace

struct wl_ctx {
  Opt opt{};
  int64_t glen = 0, l_pac = 0;
  std::vector<uint8_t> pac;
  // per read: converted query codes and regions
  std::vector<Read> reads;
  std::vector<std::vector<Region>> regions;
  std::vector<Chain> chains;
  std::vector<int> active;        // chain ids with a pending task
  std::vector<int> emitted;       // chain id of each task of the current round
  int64_t n_calls = 0;
};

extern "C" {

// ---------------- genome and index ----------------

// Genome of `len` bases (codes 0..3): a first-order Markov chain with
// P(A)=P(T)=p_at, P(C)=P(G)=0.5-p_at, and P(G|C) lowered so that CpG makes up
// `cpg` of all dinucleotides; the removed mass goes to A/C/T in proportion.
void wl_genome(uint64_t seed, int64_t len, double p_at, double cpg, uint8_t* out) {
  Rng r(seed);
  const double pc = 0.5 - p_at;
  double base[4] = {p_at, pc, pc, p_at};
  double afterC[4];
  const double pg = cpg / pc;
  const double rest = 1.0 - pc;  // mass of A, C, T
  for (int k = 0; k < 4; ++k) afterC[k] = k == 2 ? pg : base[k] * (1.0 - pg) / rest;
  int prev = 0;
  for (int64_t i = 0; i < len; ++i) {
    const double* p = prev == 1 ? afterC : base;
    double u = r.uni();
    int c = 0;
    while (c < 3 && u >= p[c]) { u -= p[c]; ++c; }
    out[i] = (uint8_t)c;
    prev = c;
  }
}

// Converted forward-only pac: [genome C->T] ++ [genome G->A] (bntseq.c:262-269),
// 2 bits per base, first base in the top bits.  `pac` must hold (2*len+3)/4 zeroed bytes.
void wl_bs_pac(const uint8_t* g, int64_t len, uint8_t* pac) {
  for (int64_t i = 0; i < len; ++i) {
    const int c = g[i];
    pac_set(pac, i, c == 1 ? 3 : c);
    pac_set(pac, len + i, c == 2 ? 0 : c);
  }
}

wl_ctx* wl_create(const int32_t* opt10, const uint8_t* pac, int64_t glen) {
  wl_ctx* c = new wl_ctx();
  c->opt = Opt{opt10[0], opt10[1], opt10[2], opt10[3], opt10[4], opt10[5], opt10[6], opt10[7], opt10[8],
               opt10[9], opt10[10]};
  c->glen = glen;
  c->l_pac = 2 * glen;
  c->pac.assign(pac, pac + (c->l_pac + 3) / 4);
  return c;
}

void wl_destroy(wl_ctx* c) { delete c; }

// ---------------- undirectional libraries (-z) ----------------

// Conversion type of a read (pair) of an undirectional library
// (bs_helpers.cpp:41-62): 0 = C->T, 1 = G->A, 2 = align both patterns.  The
// C and G shares of the read (for a pair, read 2 counted with C and G swapped,
// as an Illumina pair's second read is the other strand) are compared in
// single precision, as the reference does.
int wl_assess_conversion(const char* s1, const char* s2, int paired, float sub) {
  float observed = (float)strlen(s1);
  float c = 0, g = 0;
  for (const char* p = s1; *p; ++p) {
    c += *p == 'C';
    g += *p == 'G';
  }
  if (paired) {
    observed += (float)strlen(s2);
    for (const char* p = s2; *p; ++p) {
      g += *p == 'C';
      c += *p == 'G';
    }
  }
  const float cp = c / observed, gp = g / observed;
  float d = cp - gp;
  if (d < 0) d = d * -1.0f;
  if (cp > sub && gp > sub) return 2;
  if (cp == gp) return 2;
  if (d < 0.02) return 2;
  return cp < gp ? 0 : 1;
}

// The read records bseq_read (bwa.c:73-133) makes for one read, or one pair
// (s2 non-NULL), as (first, read group, substitution pattern) triples in
// order; returns their count (1, 2 or 4).  Directional libraries: read 1 C->T,
// read 2 G->A.  Undirectional: the pattern wl_assess_conversion picks, read 2
// the other one; when it picks both, the pattern-0 records come first, then a
// second read group with the patterns swapped.
int wl_bseq_records(const char* s1, const char* s2, int undirectional, float sub, int32_t* out) {
  int pattern = 0, both = 0, n = 0;
  if (undirectional) {
    const int t = s2 ? wl_assess_conversion(s1, s2, 1, sub) : wl_assess_conversion(s1, s1, 0, sub);
    if (t == 2) both = 1;
    else pattern = t;
  }
  auto put = [&](int first, int group, int pat) {
    out[3 * n] = first; out[3 * n + 1] = group; out[3 * n + 2] = pat; ++n;
  };
  put(0, 0, pattern);
  if (s2) put(1, 0, pattern ? 0 : 1);
  if (both) {
    put(0, 1, 1);
    if (s2) put(1, 1, 0);
  }
  return n;
}

// ---------------- reads ----------------

// Simulates n fragments.  cfg (doubles): [0] read_len, [1] paired (0/1),
// [2] insert mean, [3] insert sd, [4] CpG methylation p, [5] CH methylation p,
// [6] substitution rate, [7] indel rate, [8] indel length mode (0: uniform 1-3,
// 1: geometric p=0.25 capped at 30), [9] trimmed min length (0 = fixed length),
// [10] RRBS mode (fragments between MspI C^CGG sites, 40-500 bp),
// [11] duplicate fraction, [12] probability that a read also carries a shadow
// chain, [13] parameter of the geometric excess length of its seed.
// Reads are stored in the context (converted as the aligner converts them) and
// their seeds are appended to (seed_read, seed_chain, seed_rbeg, seed_qbeg,
// seed_len) up to max_seeds.  Returns the number of reads; *n_seeds receives
// the seed count.
//
// Shadow chains: in the 3-letter converted alphabet, re-seeding
// (bwamem.c:118-166) finds short seeds at unrelated loci, and mem_chain_flt
// keeps the first chain shadowed by the best one for extension (kept=1,
// bwamem.c:345-376).  Such a chain is modelled as one seed of length
// min_seed_len + Geometric(p) - 1 taken from the read at a uniform offset and
// placed at a uniform locus of a strand copy the read's conversion can match;
// its flanks are then unrelated sequence, as for a real spurious hit.
int64_t wl_simulate(wl_ctx* c, const uint8_t* g, uint64_t seed, int64_t n_frag, const double* cfg,
                    int32_t* seed_read, int32_t* seed_chain, int64_t* seed_rbeg, int32_t* seed_qbeg,
                    int32_t* seed_len, int64_t max_seeds, int64_t* n_seeds) {
  Rng r(seed);
  const int L = (int)cfg[0];
  const bool pe = cfg[1] != 0;
  const double imu = cfg[2], isd = cfg[3], p_cg = cfg[4], p_ch = cfg[5], sub = cfg[6], ind = cfg[7];
  const int lmode = (int)cfg[8], tmin = (int)cfg[9];
  const bool rrbs = cfg[10] != 0;
  const double dup = cfg[11], p_shadow = cfg[12], p_slen = cfg[13];
  const int64_t G = c->glen;
  const uint8_t comp[5] = {3, 2, 1, 0, 4};
  int64_t ns = 0;
  std::vector<int64_t> sites;  // MspI C^CGG cut positions (top strand), RRBS mode
  if (rrbs) {
    for (int64_t i = 0; i + 4 <= G; ++i)
      if (g[i] == 1 && g[i + 1] == 1 && g[i + 2] == 2 && g[i + 3] == 2) sites.push_back(i + 1);
    if (sites.size() < 2) { *n_seeds = 0; return -1; }
  }
  std::vector<uint8_t> frag, src, rd;
  std::vector<int32_t> fmap, rmap;
  int64_t last_p = 0, last_F = 0;
  int last_strand = 0;
  for (int64_t f = 0; f < n_frag; ++f) {
    int64_t p, F;
    int strand;
    if (f > 0 && dup > 0 && r.uni() < dup) {  // duplicate fragment
      p = last_p; F = last_F; strand = last_strand;
    } else if (rrbs) {
      // a fragment between two adjacent sites, 40..500 bp (RRBSIndex.py:101-140 size selection)
      int64_t tries = 0;
      do {
        const int64_t s = r.below((int64_t)sites.size() - 1);
        p = sites[s]; F = sites[s + 1] - sites[s];
      } while ((F < 40 || F > 500) && ++tries < 1000);
      strand = (int)(r.next() & 1);
    } else {
      F = L + 10;
      if (pe) {
        F = (int64_t)std::llround(imu + isd * r.normal());
        if (F < L + 10) F = L + 10;
      }
      p = r.below(G - F);
      strand = (int)(r.next() & 1);
    }
    last_p = p; last_F = F; last_strand = strand;
    // fragment on its own strand, 5'->3'
    frag.resize(F);
    for (int64_t k = 0; k < F; ++k)
      frag[k] = strand == 0 ? g[p + k] : comp[g[p + F - 1 - k]];
    // methylation-dependent conversion on the sampled strand
    for (int64_t k = 0; k < F; ++k) {
      if (frag[k] != 1) continue;
      const bool cg = k + 1 < F && frag[k + 1] == 2;
      if (r.uni() >= (cg ? p_cg : p_ch)) frag[k] = 3;
    }
    const int nreads = pe ? 2 : 1;
    for (int m = 0; m < nreads; ++m) {
      // source sequence in read orientation and the frag index of each source base
      src.resize(F);
      fmap.resize(F);
      for (int64_t k = 0; k < F; ++k) {
        if (m == 0) { src[k] = frag[k]; fmap[k] = (int32_t)k; }
        else { src[k] = comp[frag[F - 1 - k]]; fmap[k] = (int32_t)(F - 1 - k); }
      }
      int rl = L;
      if (rrbs) rl = std::min<int64_t>(F, L);
      if (tmin > 0) rl = tmin + (int)r.below(L - tmin + 1);
      rd.clear();
      rmap.clear();
      for (int64_t k = 0; k < F && (int)rd.size() < rl; ++k) {
        const double u = r.uni();
        if (u < sub) {
          rd.push_back((uint8_t)((src[k] + 1 + r.below(3)) & 3));
          rmap.push_back(fmap[k]);
        } else if (u < sub + ind * 0.5) {  // deletion
          const int len = lmode == 0 ? 1 + (int)r.below(3) : std::min(30, r.geometric(0.25));
          k += len - 1;
        } else if (u < sub + ind) {  // insertion after this base
          rd.push_back(src[k]);
          rmap.push_back(fmap[k]);
          const int len = lmode == 0 ? 1 + (int)r.below(3) : std::min(30, r.geometric(0.25));
          for (int q = 0; q < len && (int)rd.size() < rl; ++q) {
            rd.push_back((uint8_t)r.below(4));
            rmap.push_back(-1);
          }
        } else {
          rd.push_back(src[k]);
          rmap.push_back(fmap[k]);
        }
      }
      std::vector<uint8_t> ord(rd);  // the read as sequenced (bseq1_t.oseq)
      // aligner-side conversion: read 1 C->T, read 2 G->A
      for (auto& b : rd) {
        if (m == 0 && b == 1) b = 3;
        if (m == 1 && b == 2) b = 0;
      }
      // doubled-reference position of each read base on the copy it aligns to
      // (read 1: Watson -> W_C2T forward, Crick -> reverse of the G->A copy;
      //  read 2: Watson -> reverse of the C->T copy, Crick -> G->A copy forward)
      const int64_t l_pac = c->l_pac;
      auto dpos = [&](int32_t fi) -> int64_t {
        const int64_t x = strand == 0 ? p + fi : p + F - 1 - fi;  // top-strand coordinate
        if (m == 0) return strand == 0 ? x : 2 * l_pac - 1 - (G + x);
        return strand == 0 ? 2 * l_pac - 1 - x : G + x;
      };
      const int read_id = (int)c->reads.size();
      c->reads.push_back(Read{rd, std::move(ord)});
      // seeds: maximal runs along the true path where consecutive read bases map
      // to consecutive reference positions and the converted bases agree
      int run_q = -1;
      int64_t run_r = 0;
      const int n = (int)rd.size();
      for (int q = 0; q <= n; ++q) {
        bool ok = false;
        int64_t rp = 0;
        if (q < n && rmap[q] >= 0) {
          rp = dpos(rmap[q]);
          ok = ref_base(c->pac.data(), l_pac, rp) == rd[q];
        }
        const bool cont = ok && run_q >= 0 && rp == run_r + (q - run_q);
        if (!cont) {
          if (run_q >= 0 && q - run_q >= c->opt.min_seed_len && ns < max_seeds) {
            seed_read[ns] = read_id; seed_chain[ns] = 0; seed_rbeg[ns] = run_r; seed_qbeg[ns] = run_q;
            seed_len[ns] = q - run_q; ++ns;
          }
          run_q = -1;
          if (ok) { run_q = q; run_r = rp; }
        }
      }
      if (p_shadow > 0 && n >= c->opt.min_seed_len && r.uni() < p_shadow && ns < max_seeds) {
        int sl = c->opt.min_seed_len + r.geometric(p_slen) - 1;
        if (sl > n) sl = n;
        const int qb = (int)r.below(n - sl + 1);
        // read 1 (C->T) matches W_C2T forward [0,G) or the reverse of the G->A copy [2G,3G);
        // read 2 (G->A) matches the G->A copy forward [G,2G) or the reverse of the C->T copy [3G,4G)
        const int64_t lo = (m == 0 ? 0 : G) + ((r.next() & 1) ? 2 * G : 0);
        const int64_t rb = lo + r.below(G - sl);
        seed_read[ns] = read_id; seed_chain[ns] = 1; seed_rbeg[ns] = rb; seed_qbeg[ns] = qb;
        seed_len[ns] = sl; ++ns;
      }
    }
  }
  *n_seeds = ns;
  return (int64_t)c->reads.size();
}

int32_t wl_read_len(const wl_ctx* c, int64_t i) { return (int32_t)c->reads[i].seq.size(); }
void wl_read_seq(const wl_ctx* c, int64_t i, uint8_t* out) {
  memcpy(out, c->reads[i].seq.data(), c->reads[i].seq.size());
}
// the read as sequenced, before the in-silico conversion (bwa.c:95-133 keeps it as oseq)
void wl_read_oseq(const wl_ctx* c, int64_t i, uint8_t* out) {
  memcpy(out, c->reads[i].oseq.data(), c->reads[i].oseq.size());
}

// ---------------- mem_chain2aln rounds ----------------

// Chains: consecutive seeds with the same (read, chain) key form one chain (the
// true-locus seeds are colinear, as mem_chain groups them, bwamem.c:255-329);
// a read's chains are extended in input order, best chain first, sharing the
// read's region list as mem_chain2aln's av does.
int64_t wl_chains_init(wl_ctx* c, int64_t n_seeds, const int32_t* seed_read, const int32_t* seed_chain,
                       const int64_t* seed_rbeg, const int32_t* seed_qbeg, const int32_t* seed_len) {
  c->chains.clear();
  c->regions.assign(c->reads.size(), {});
  const Opt& o = c->opt;
  const int64_t l_pac = c->l_pac;
  for (int64_t s = 0; s < n_seeds;) {
    Chain ch;
    ch.read = seed_read[s];
    const int key = seed_chain[s];
    if (!c->chains.empty() && c->chains.back().read == ch.read) ch.wait_for = (int)c->chains.size() - 1;
    while (s < n_seeds && seed_read[s] == ch.read && seed_chain[s] == key) {
      ch.seeds.push_back(Seed{seed_rbeg[s], seed_qbeg[s], seed_len[s], seed_len[s]});
      ++s;
    }
    const int l_query = (int)c->reads[ch.read].seq.size();
    // span (bwamem.c:646-661)
    int64_t r0 = l_pac << 1, r1 = 0;
    for (const Seed& t : ch.seeds) {
      const int64_t b = t.rbeg - (t.qbeg + cal_max_gap(o, t.qbeg));
      const int64_t e = t.rbeg + t.len + ((l_query - t.qbeg - t.len) + cal_max_gap(o, l_query - t.qbeg - t.len));
      r0 = std::min(r0, b);
      r1 = std::max(r1, e);
    }
    r0 = std::max<int64_t>(r0, 0);
    r1 = std::min<int64_t>(r1, l_pac << 1);
    if (r0 < l_pac && l_pac < r1) {
      if (ch.seeds[0].rbeg < l_pac) r1 = l_pac;
      else r0 = l_pac;
    }
    // bns_fetch_seq (bntseq.c:439-461): clamp to the contig holding seeds[0]
    const int64_t mid = ch.seeds[0].rbeg;
    const bool is_rev = mid >= l_pac;
    const int64_t fpos = is_rev ? 2 * l_pac - 1 - mid : mid;
    int64_t far_b = fpos < c->glen ? 0 : c->glen, far_e = far_b + c->glen;  // two contigs of glen
    if (is_rev) {
      const int64_t t = far_b;
      far_b = 2 * l_pac - far_e;
      far_e = 2 * l_pac - t;
    }
    r0 = std::max(r0, far_b);
    r1 = std::min(r1, far_e);
    ch.rmax0 = r0;
    ch.rmax1 = r1;
    ch.rseq.resize(r1 - r0);
    for (int64_t k = r0; k < r1; ++k) ch.rseq[k - r0] = (uint8_t)ref_base(c->pac.data(), l_pac, k);
    // seeds by score, ascending (ks_introsort_64 on score<<32|i, bwamem.c:665-669)
    for (size_t i = 0; i < ch.seeds.size(); ++i)
      ch.srt.push_back((uint64_t)ch.seeds[i].score << 32 | i);
    std::sort(ch.srt.begin(), ch.srt.end());
    ch.k = (int)ch.seeds.size() - 1;
    ch.phase = 0;
    c->chains.push_back(std::move(ch));
  }
  return (int64_t)c->chains.size();
}

// Advances chain `ch` until it needs an extension or is done.
static void advance(wl_ctx* c, Chain& ch) {
  const Opt& o = c->opt;
  const int l_query = (int)c->reads[ch.read].seq.size();
  auto& av = c->regions[ch.read];
  while (ch.phase == 0) {
    if (ch.k < 0) { ch.phase = 3; return; }
    const Seed& s = ch.seeds[(uint32_t)ch.srt[ch.k]];
    size_t i;
    for (i = 0; i < av.size(); ++i) {  // containment test, bwamem.c:675-692
      const Region& p = av[i];
      if (s.rbeg < p.rb || s.rbeg + s.len > p.re || s.qbeg < p.qb || s.qbeg + s.len > p.qe) continue;
      if (s.len - p.seedlen0 > .1 * l_query) continue;
      int64_t qd = s.qbeg - p.qb, rd = s.rbeg - p.rb;
      int max_gap = cal_max_gap(o, (int)(qd < rd ? qd : rd));
      int w = max_gap < p.w ? max_gap : p.w;
      if (qd - rd < w && rd - qd < w) break;
      qd = p.qe - (s.qbeg + s.len); rd = p.re - (s.rbeg + s.len);
      max_gap = cal_max_gap(o, (int)(qd < rd ? qd : rd));
      w = max_gap < p.w ? max_gap : p.w;
      if (qd - rd < w && rd - qd < w) break;
    }
    if (i < av.size()) {  // bwamem.c:693-710
      size_t j;
      for (j = ch.k + 1; j < ch.seeds.size(); ++j) {
        if (ch.srt[j] == 0) continue;
        const Seed& t = ch.seeds[(uint32_t)ch.srt[j]];
        if (t.len < s.len * .95) continue;
        if (s.qbeg <= t.qbeg && s.qbeg + s.len - t.qbeg >= s.len >> 2 && t.qbeg - s.qbeg != t.rbeg - s.rbeg) break;
        if (t.qbeg <= s.qbeg && t.qbeg + t.len - s.qbeg >= s.len >> 2 && s.qbeg - t.qbeg != s.rbeg - t.rbeg) break;
      }
      if (j == ch.seeds.size()) {
        ch.srt[ch.k] = 0;
        --ch.k;
        continue;
      }
    }
    // start a new region (bwamem.c:712-716)
    ch.cur = Region{};
    ch.cur.w = o.w;
    ch.aw0 = ch.aw1 = o.w;
    ch.cur.score = ch.cur.truesc = -1;
    ch.tryi = 0;
    if (s.qbeg) {
      ch.phase = 1;
    } else {
      ch.cur.score = ch.cur.truesc = s.len * o.a;
      ch.cur.qb = 0;
      ch.cur.rb = s.rbeg;
      ch.phase = 2;
    }
    ch.prev = ch.cur.score;
    ch.sc0 = ch.cur.score;
    if (ch.phase == 2 && s.qbeg + s.len == l_query) {
      // no right flank either: finish the region now
      ch.cur.qe = l_query;
      ch.cur.re = s.rbeg + s.len;
      ch.phase = 4;
    }
  }
}

static void finish_region(wl_ctx* c, Chain& ch) {
  const Seed& s = ch.seeds[(uint32_t)ch.srt[ch.k]];
  Region& a = ch.cur;
  a.seedcov = 0;
  for (const Seed& t : ch.seeds)  // bwamem.c:778-783
    if (t.qbeg >= a.qb && t.qbeg + t.len <= a.qe && t.rbeg >= a.rb && t.rbeg + t.len <= a.re) a.seedcov += t.len;
  a.w = ch.aw0 > ch.aw1 ? ch.aw0 : ch.aw1;
  a.seedlen0 = s.len;
  c->regions[ch.read].push_back(a);
  --ch.k;
  ch.phase = 0;
}

// Emits the next round: for every unfinished chain, the one ksw_extend2 call it
// is waiting for.  Returns the task count; the caller reads them with
// wl_round_sizes / wl_round_tasks.
int64_t wl_round_emit(wl_ctx* c) {
  c->emitted.clear();
  for (size_t id = 0; id < c->chains.size(); ++id) {
    Chain& ch = c->chains[id];
    if (ch.wait_for >= 0 && c->chains[ch.wait_for].phase != 3) continue;
    for (;;) {
      if (ch.phase == 0) advance(c, ch);
      if (ch.phase == 4) { finish_region(c, ch); continue; }
      break;
    }
    if (ch.phase == 1 || ch.phase == 2) c->emitted.push_back((int)id);
  }
  return (int64_t)c->emitted.size();
}

// Sequence lengths of the current round's tasks (for sizing buffers).
void wl_round_sizes(const wl_ctx* c, int64_t* qtot, int64_t* ttot) {
  int64_t qs = 0, ts = 0;
  for (int id : c->emitted) {
    const Chain& ch = c->chains[id];
    const Seed& s = ch.seeds[(uint32_t)ch.srt[ch.k]];
    const int l_query = (int)c->reads[ch.read].seq.size();
    if (ch.phase == 1) { qs += s.qbeg; ts += s.rbeg - ch.rmax0; }
    else { qs += l_query - (s.qbeg + s.len); ts += ch.rmax1 - ch.rmax0 - (s.rbeg + s.len - ch.rmax0); }
  }
  *qtot = qs;
  *ttot = ts;
}

// Writes the round's tasks in the batch layout of include/bsx.h.
void wl_round_tasks(const wl_ctx* c, uint8_t* query, uint64_t* q_off, int32_t* qlen, uint8_t* target,
                    uint64_t* t_off, int32_t* tlen, int32_t* h0, int32_t* w, int32_t* end_bonus,
                    int32_t* chain_id, int32_t* side) {
  const Opt& o = c->opt;
  uint64_t qo = 0, to = 0;
  for (size_t n = 0; n < c->emitted.size(); ++n) {
    const int id = c->emitted[n];
    const Chain& ch = c->chains[id];
    const Seed& s = ch.seeds[(uint32_t)ch.srt[ch.k]];
    const std::vector<uint8_t>& q = c->reads[ch.read].seq;
    const int l_query = (int)q.size();
    q_off[n] = qo; t_off[n] = to; chain_id[n] = id;
    if (ch.phase == 1) {  // left flank: both sequences reversed (bwamem.c:722-726)
      const int64_t tmp = s.rbeg - ch.rmax0;
      for (int i = 0; i < s.qbeg; ++i) query[qo + i] = q[s.qbeg - 1 - i];
      for (int64_t i = 0; i < tmp; ++i) target[to + i] = ch.rseq[tmp - 1 - i];
      qlen[n] = s.qbeg; tlen[n] = (int32_t)tmp;
      h0[n] = s.len * o.a;
      w[n] = o.w << ch.tryi;
      end_bonus[n] = o.pen_clip5;
      side[n] = 0;
    } else {  // right flank (bwamem.c:748-763)
      const int qe = s.qbeg + s.len;
      const int64_t re = s.rbeg + s.len - ch.rmax0;
      for (int i = 0; i < l_query - qe; ++i) query[qo + i] = q[qe + i];
      const int64_t tl = ch.rmax1 - ch.rmax0 - re;
      for (int64_t i = 0; i < tl; ++i) target[to + i] = ch.rseq[re + i];
      qlen[n] = l_query - qe; tlen[n] = (int32_t)tl;
      h0[n] = ch.sc0;
      w[n] = o.w << ch.tryi;
      end_bonus[n] = o.pen_clip3;
      side[n] = 1;
    }
    qo += qlen[n];
    to += tlen[n];
  }
}

// Feeds the round's results back (arrays indexed like the emitted tasks).
void wl_round_consume(wl_ctx* c, const int32_t* score, const int32_t* qle, const int32_t* tle,
                      const int32_t* gtle, const int32_t* gscore, const int32_t* max_off) {
  const Opt& o = c->opt;
  for (size_t n = 0; n < c->emitted.size(); ++n) {
    Chain& ch = c->chains[c->emitted[n]];
    const Seed& s = ch.seeds[(uint32_t)ch.srt[ch.k]];
    const int l_query = (int)c->reads[ch.read].seq.size();
    Region& a = ch.cur;
    ++c->n_calls;
    a.score = score[n];
    const int aw = o.w << ch.tryi;
    const bool stop = a.score == ch.prev || max_off[n] < (aw >> 1) + (aw >> 2) || ch.tryi + 1 >= 2;
    if (ch.phase == 1) {
      ch.aw0 = aw;
      if (!stop) { ch.prev = a.score; ++ch.tryi; continue; }
      if (gscore[n] <= 0 || gscore[n] <= a.score - o.pen_clip5) {  // bwamem.c:740-746
        a.qb = s.qbeg - qle[n]; a.rb = s.rbeg - tle[n]; a.truesc = a.score;
      } else {
        a.qb = 0; a.rb = s.rbeg - gtle[n]; a.truesc = gscore[n];
      }
      if (s.qbeg + s.len != l_query) {
        ch.phase = 2; ch.tryi = 0; ch.sc0 = a.score; ch.prev = a.score;
      } else {
        a.qe = l_query; a.re = s.rbeg + s.len;
        finish_region(c, ch);
      }
    } else {
      ch.aw1 = aw;
      if (!stop) { ch.prev = a.score; ++ch.tryi; continue; }
      const int qe = s.qbeg + s.len;
      const int64_t re = s.rbeg + s.len - ch.rmax0;
      if (gscore[n] <= 0 || gscore[n] <= a.score - o.pen_clip3) {  // bwamem.c:768-774
        a.qe = qe + qle[n]; a.re = ch.rmax0 + re + tle[n]; a.truesc += a.score - ch.sc0;
      } else {
        a.qe = l_query; a.re = ch.rmax0 + re + gtle[n]; a.truesc += gscore[n] - ch.sc0;
      }
      finish_region(c, ch);
    }
  }
}

int64_t wl_n_calls(const wl_ctx* c) { return c->n_calls; }

// Regions of read i: returns the count; writes up to `cap` entries as
// (qb, qe, rb, re, score, truesc, w, seedcov, seedlen0) int64 nonuples.
int64_t wl_regions(const wl_ctx* c, int64_t i, int64_t* out, int64_t cap) {
  const auto& v = c->regions[i];
  for (size_t k = 0; k < v.size() && (int64_t)k < cap; ++k) {
    const Region& a = v[k];
    int64_t* o = out + 9 * k;
    o[0] = a.qb; o[1] = a.qe; o[2] = a.rb; o[3] = a.re; o[4] = a.score; o[5] = a.truesc;
    o[6] = a.w; o[7] = a.seedcov; o[8] = a.seedlen0;
  }
  return (int64_t)v.size();
}

int64_t wl_n_reads(const wl_ctx* c) { return (int64_t)c->reads.size(); }

// Reference window [beg, end) of the doubled reference (bns_get_seq semantics,
// bntseq.c:413-437); returns its length (0 when it bridges the strands).
int64_t wl_get_seq(const wl_ctx* c, int64_t beg, int64_t end, uint8_t* out) {
  const int64_t l_pac = c->l_pac;
  if (end < beg) std::swap(beg, end);
  if (end > 2 * l_pac) end = 2 * l_pac;
  if (beg < 0) beg = 0;
  if (!(beg >= l_pac || end <= l_pac)) return 0;
  for (int64_t k = beg; k < end; ++k) out[k - beg] = (uint8_t)ref_base(c->pac.data(), l_pac, k);
  return end - beg;
}

}  // extern "C"
