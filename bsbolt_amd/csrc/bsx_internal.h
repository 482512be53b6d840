// Internal declarations shared by the kernels and the C-ABI layer of libbsx.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bsx {

// Extension work is split into bins by query length; each bin runs a kernel
// instantiation with L lanes per task and C query columns per lane
// (qlen <= L*C).  Index = bin id.  L = 1 is the lane-per-task kernel (one task
// per lane, its whole query in registers, ~22 VALU instructions per column);
// L >= 16 are the group kernels, whose row cost per task ~ (fixed + per-column
// work * C) * L / 64 wave instructions carries a fixed part (prefix scans,
// group reductions) that the lane kernel does not have.
constexpr int kExtBins = 11;
struct BinShape { int lanes, cols; };
constexpr BinShape kExtShape[kExtBins] = {
    {4, 4}, {1, 32}, {1, 48}, {1, 64}, {16, 6}, {16, 8}, {32, 6}, {32, 8}, {64, 8}, {64, 16}, {64, 32}};
constexpr int kExtBlock = 64;             // threads per extension block: one wave, so a block
                                          // never holds wave slots idle while a sibling finishes
constexpr int kExtTasksPerBlock = 32;     // task descriptors owned by one plain-kernel block
constexpr int kNarrowTasksPerBlock = 16;  // ... by one narrow-kernel block
constexpr int bin_limit(int b) { return kExtShape[b].lanes * kExtShape[b].cols; }

// The CIGAR (ksw_global2) kernel bins tasks the same way with its own shapes.
constexpr int kGlbBins = 9;
constexpr BinShape kGlbShape[kGlbBins] = {
    {8, 2}, {8, 3}, {8, 4}, {16, 4}, {32, 4}, {64, 4}, {64, 8}, {64, 16}, {64, 32}};
constexpr int kExtSlices = 8;  // fetch-counter slots per bin (CIGAR bins, lane-kernel bins and classes)
constexpr int glb_bin_limit(int b) { return kGlbShape[b].lanes * kGlbShape[b].cols; }
// First entry of fetch slice s of a list of n entries: a balanced partition,
// the first n % kExtSlices slices hold one entry more.
__host__ __device__ inline uint32_t slice_begin(uint32_t n, int s) {
  const uint32_t q = n / kExtSlices, r = n % kExtSlices;
  return (uint32_t)s * q + ((uint32_t)s < r ? (uint32_t)s : r);
}

// Plain bins list their tasks longest run first, so that a kernel's last
// tasks are short ones and its waves drain together.  The rows a run takes
// follow min(tlen, 2*h0) closely (the run ends when the target does, or once
// the score built on h0 is spent or the z-drop fires); the key is bucketed in
// kRowBuckets steps of 32 rows, bucket 0 = longest.  Lane-kernel bins place
// rank r of that order at slot r / kExtSlices of fetch slice r % kExtSlices,
// so that every slice is itself longest first.  The order changes no result.
constexpr int kRowBuckets = 8;
__host__ __device__ inline int row_bucket(int tlen, int h0) {
  const int est = tlen < 2 * h0 ? tlen : 2 * h0;
  const int k = est >> 5;
  return kRowBuckets - 1 - (k < kRowBuckets - 1 ? (k > 0 ? k : 0) : kRowBuckets - 1);
}
constexpr uint32_t lane_bin_mask() {
  uint32_t m = 0;
  for (int b = 0; b < kExtBins; ++b)
    if (kExtShape[b].lanes == 1) m |= 1u << b;
  return m;
}
constexpr uint32_t kLaneBinMask = lane_bin_mask();

// Learned-band form: classes 0..3 run the band-following frame kernel with
// frames of 24, 40, 80, 144 columns; class 4 is the plain kernel.  Count slots: plain bins
// [0, kExtBins), narrow classes [kExtBins, kExtBins + kBandClasses - 1).
constexpr int kBandClasses = 5;
constexpr int kCountSlots = 16;
static_assert(kExtBins + kBandClasses - 1 <= kCountSlots, "count slots for bins and narrow classes");
constexpr int kNarrowQMax = 256;  // longest query the narrow kernels take
// Frame classes run by the lane-per-task frame kernel (one task per lane) for
// queries of at most kLaneFrameQMax bases; longer queries of these classes run
// the group frame kernel.  Both compute the same frame run.
constexpr int kLaneFrameClasses = 2;
constexpr int kLaneFrameQMax = 128;
constexpr int kFetchSlots = kCountSlots * kExtSlices;
enum BandMode { kBandPlain = 0, kBandLearned = 1, kBandFixed = 2 };
struct BandPlan {
  int mode;            // BandMode
  int fixed;           // class for kBandFixed
  const uint8_t* cls;  // [n] predicted class for kBandLearned
  bool any_gain;       // route a task to its class even where the frame is not narrower than its bin
  bool skip_plain;     // narrow pass only (labelling): no plain pass, no result unpack
};

// Scores are tracked as (h << 11 | column) keys in 32-bit lanes.
constexpr int kScoreLimit = 1 << 20;

// Kernel-side view of an extension batch (all pointers are device memory).
struct ExtArgs {
  int o_del, e_del, o_ins, e_ins, zdrop, mat_max;
  // Query profile by query code qc (0..4): bytes k=0..3 of prof_lo[qc] hold
  // mat[k*5+qc] (target code k), prof_hi[qc] holds mat[20+qc] (target N).
  uint32_t prof_lo[5];
  int32_t prof_hi[5];
  int64_t n;
  const uint8_t* query;
  const uint64_t* q_off;
  const int32_t* qlen;
  const uint8_t* target;
  const uint64_t* t_off;
  const int32_t* tlen;
  const int32_t* h0;
  const int32_t* w;
  const int32_t* end_bonus;
  int32_t *score, *qle, *tle, *gtle, *gscore, *max_off;
  int64_t* cells;
};

// One task of a bin's work list (written by the binning pass, in bin order).
struct ExtDesc {
  int32_t task, qlen, tlen, h0;
  int32_t w, pad0;  // w already clamped as ksw.c:399-407
  uint32_t qoff_lo, qoff_hi, toff_lo, toff_hi, pad1, pad2;
};
static_assert(sizeof(ExtDesc) == 48, "descriptor layout");

// Side streams of a context: independent launches (the plain bins, the frame
// classes) are spread over the call's stream and these, between a fork and a
// join, so that one launch's tail overlaps the others' work.  All streams keep
// the default priority.
struct StreamSet {
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr, join[kSide] = {};
};
// launches of the plain bins: stream 0 = the call's stream, k > 0 = side[k-1]
constexpr int kBinStream[11] = {2, 1, 3, 1, 0, 2, 3, 0, 0, 0, 0};
inline hipStream_t pick_stream(const StreamSet* ss, hipStream_t main, int k) {
  return (ss && k > 0 && ss->side[k - 1]) ? ss->side[k - 1] : main;
}
inline void ss_fork(const StreamSet* ss, hipStream_t main) {
  if (!ss || !ss->fork) return;
  (void)hipEventRecord(ss->fork, main);
  for (int k = 0; k < StreamSet::kSide; ++k) (void)hipStreamWaitEvent(ss->side[k], ss->fork, 0);
}
inline void ss_join(const StreamSet* ss, hipStream_t main) {
  if (!ss || !ss->fork) return;
  for (int k = 0; k < StreamSet::kSide; ++k) {
    (void)hipEventRecord(ss->join[k], ss->side[k]);
    (void)hipStreamWaitEvent(main, ss->join[k], 0);
  }
}
// Per-call device workspace for the extension path.
struct ExtWork {
  uint8_t* bin_of;     // [n] bin id, 0xff = rejected task
  ExtDesc* desc;       // [n] task descriptors grouped by bin
  int32_t* rec;        // [n][8] packed per-task results
  uint32_t* counts;    // [kCountSlots] tasks per bin / narrow class
  uint32_t* cursor;    // [kCountSlots] scatter cursors
  uint32_t* fetch;     // [kFetchSlots] persistent-kernel fetch counters, kExtSlices per count slot
  uint32_t* split;     // [2 * kLaneFrameClasses]: per lane-frame class, tasks at the front of its
                       // list (query <= kLaneFrameQMax) and the back-of-list cursor
  int32_t* err;        // [0] rejected count, [1] lowest rejected index
  uint8_t* nbin_of;    // [n] narrow class of the task, 0xff = plain only (learned-band form)
  uint8_t* need_plain; // [n] 1 = the plain kernel computes the task (learned-band form)
  uint32_t* bcount;    // [kExtBins * kRowBuckets] tasks per plain bin and row bucket
  uint32_t* bcursor;   // [kExtBins * kRowBuckets] scatter cursors of the same
  uint32_t* fcount;    // [kLaneFrameClasses * kRowBuckets] front tasks per lane-frame class and row bucket
  uint32_t* fcursor;   // [kLaneFrameClasses * kRowBuckets] scatter cursors of the same
  // optional per-kernel timing: events [2*slot], [2*slot+1] around each kernel
  // launch (slots: plain bins, then narrow classes, then the predictor); *tmask
  // gets bit `slot` set for every pair recorded
  hipEvent_t* tev = nullptr;
  uint32_t* tmask = nullptr;
  const StreamSet* ss = nullptr;  // side streams (nullable: everything on the call's stream)
  // tuning: descriptors per block of the plain group kernels and of the frame
  // group kernels, and the shape variant of the bins with alternatives
  int tpb_plain = kExtTasksPerBlock, tpb_narrow = kNarrowTasksPerBlock;
  int shape_var[2] = {0, 0};
};
// Bins with alternative group shapes (same columns per task, so the binning is
// unchanged; the tuner picks one per device): index 0 = bin 4 (queries of
// 65-96 bases), 1 = bin 5 (97-128).  Variant 0 is kExtShape's shape.
constexpr int kTunedBin0 = 4;
constexpr BinShape kAltShape[2][2] = {{{16, 6}, {32, 3}}, {{16, 8}, {32, 4}}};
static_assert(kExtShape[kTunedBin0].lanes == kAltShape[0][0].lanes &&
              kExtShape[kTunedBin0].cols == kAltShape[0][0].cols &&
              kExtShape[kTunedBin0 + 1].lanes == kAltShape[1][0].lanes &&
              kExtShape[kTunedBin0 + 1].cols == kAltShape[1][0].cols &&
              kAltShape[0][1].lanes * kAltShape[0][1].cols == bin_limit(kTunedBin0) &&
              kAltShape[1][1].lanes * kAltShape[1][1].cols == bin_limit(kTunedBin0 + 1), "alternative shapes");
constexpr int kTimedSlots = 16;  // kExtBins + 4 narrow classes + predictor
static_assert(kExtBins + kBandClasses <= kTimedSlots, "timing slots");
inline void time_begin(const ExtWork& wk, int slot, hipStream_t st) {
  if (wk.tev) (void)hipEventRecord(wk.tev[2 * slot], st);
}
inline void time_end(const ExtWork& wk, int slot, hipStream_t st) {
  if (wk.tev) {
    (void)hipEventRecord(wk.tev[2 * slot + 1], st);
    *wk.tmask |= 1u << slot;
  }
}

struct ExtKArgs;
// lane-per-task kernel of bin `bin` (query width Q); grid = resident waves, or
// fewer when the host knows the bin's task count (hint >= 0)
template <int Q, bool E1>
void launch_lane_bin(const ExtKArgs& k, const ExtWork& wk, int bin, int num_cu, hipStream_t st, int64_t hint);
// lane-per-task frame kernel (frame of WF columns) of lane-frame class `cls`
template <int WF, bool E1>
void launch_lane_frame(const ExtKArgs& k, const ExtWork& wk, int cls, int num_cu, hipStream_t st);
hipError_t launch_extend(const ExtArgs& a, const ExtWork& wk, int num_cu, hipStream_t st,
                         const int64_t* host_bin_counts, const BandPlan* plan);
hipError_t launch_narrow(const ExtArgs& a, const ExtKArgs& k, const ExtWork& wk, int num_cu, hipStream_t st,
                         const BandPlan& plan);

struct GlbArgs {
  int o_del, e_del, o_ins, e_ins;
  uint32_t prof_lo[5];
  int32_t prof_hi[5];
  int64_t n;
  const uint8_t* query;
  const uint64_t* q_off;
  const int32_t* qlen;
  const uint8_t* target;
  const uint64_t* t_off;
  const int32_t* tlen;
  const int32_t* w;
  int32_t* score;
  int32_t* n_cigar;
  uint32_t* cigar;
  int64_t cigar_slot;
};

struct GlbWork {
  uint8_t* bin_of;
  int32_t* list;
  uint32_t* counts;
  uint32_t* cursor;
  uint32_t* fetch;
  int32_t* err;  // [0] rejected, [1] first rejected, [2] CIGAR overflows
  uint8_t* z;    // per-group direction-byte slots
  int64_t zslot; // bytes per slot
  const StreamSet* ss;  // side streams for the bins with few tasks per group (nullptr: one stream)
};

// NM / MD:Z / XB:Z of CIGARs (bwa_gen_cigar2's tail, bwa.c:250-343); fields as
// bsx_md_tasks / bsx_md_results in include/bsx.h.
struct MdArgs {
  int64_t n, l_pac, crick_l;
  const uint8_t* opac;
  const int64_t *rb, *re;
  const uint8_t* oquery;
  const uint64_t* oq_off;
  const int32_t* l_query;
  const int32_t* n_cigar;
  const uint32_t* cigar;
  int64_t cigar_slot;
  char* md;
  int64_t md_slot;
  int32_t* md_len;
  char* xb;
  int64_t xb_slot;
  int32_t *xb_len, *nm, *meth, *status;
};
hipError_t launch_md(const MdArgs& a, int num_cu, hipStream_t st);

// SAM CIGAR and position of regions (mem_reg2aln's tail, bwamem.c:1237-1263);
// fields as bsx_reg2aln_tasks / bsx_reg2aln_results in include/bsx.h.
struct Reg2alnArgs {
  int64_t n, l_pac;
  const int64_t *rb, *re;
  const int32_t *qb, *qe, *l_seq, *n_cigar;
  const uint32_t* cigar;
  int64_t cigar_slot;
  uint32_t* out;
  int64_t out_slot;
  int32_t* out_n;
  int64_t* pos;
  int32_t* is_rev;
};
hipError_t launch_reg2aln_cigar(const Reg2alnArgs& a, int num_cu, hipStream_t st);

// Blocks per CU of the global kernel by lanes per task (8-lane shapes hold
// twice the groups per block, so fewer blocks keep the scratch the same
// size); the scratch holds one slot per group.
constexpr int glb_blocks_per_cu(int lanes) { return lanes < 16 ? 4 : 7; }
int64_t glb_groups(int num_cu, bool side);  // scratch slots, with or without the side-stream region
hipError_t launch_global_impl(const GlbArgs& a, const GlbWork& w, int num_cu, hipStream_t st);

}  // namespace bsx
