// C ABI of libbsx.so (declared in include/bsx.h): contexts, argument checks,
// host<->device staging, and dispatch to the extension / global kernels.
#include <hip/hip_runtime.h>

#include <climits>
#include <dlfcn.h>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bsx.h"
#include "bsx_internal.h"
#include "mlp.h"

using namespace bsx;

// counts[kCountSlots], cursor[kCountSlots], err[4], fetch[kFetchSlots], split[2 * kLaneFrameClasses],
// bcount[kExtBins * kRowBuckets], bcursor[kExtBins * kRowBuckets],
// fcount[kLaneFrameClasses * kRowBuckets], fcursor[kLaneFrameClasses * kRowBuckets]
static constexpr size_t kSmallBytes = 4 * (2 * kCountSlots + 4 + kFetchSlots + 2 * kLaneFrameClasses +
                                           2 * kExtBins * kRowBuckets + 2 * kLaneFrameClasses * kRowBuckets);

// band model buffer: params, adam_m, adam_v (FP32), then the BF16 operand copies
static constexpr size_t kAlign = 256;
static constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
static constexpr size_t kParamBytes = align_up(4 * (size_t)kParamCount);
static constexpr size_t kModelBytes = 3 * kParamBytes + align_up(2 * 64 * kFeatPad) + 2 * align_up(2 * 64 * 64) +
                                      2 * align_up(2 * 16 * 64);

namespace {

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

}  // namespace

// Device workspace of the batch entry points.  A context has its own; a
// captured graph owns another, so later calls on the context can grow theirs
// without moving memory the graph refers to.
struct Work {
  DevBuf bin_of, list, desc, rec, small;  // small: counts/cursor/err/fetch
  DevBuf nbin_of, need_plain, cls;        // learned-band form
  DevBuf zbuf;                            // direction-byte slots of the global kernel
  void release() {
    for (DevBuf* b : {&bin_of, &list, &desc, &rec, &small, &nbin_of, &need_plain, &cls, &zbuf}) b->release();
  }
};

struct bsx_ctx {
  int device = 0;
  int num_cu = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  StreamSet ss;  // side streams of the extension launches
  bool ss_on = true;
  std::string err;
  Work wown;           // workspace of the eager calls
  Work* ws = &wown;    // workspace in use (a graph's while it is captured)
  // learned-band form
  DevBuf stats, model, part, grad;
  int band_policy = BSX_BAND_PLAIN;
  int band_class = 0;
  bool model_ready = false;
  int adam_step = 0;
  int64_t last_n = 0;  // tasks of the last extension call (for bsx_band_stats)
  // per-kernel timing of the extension calls (bsx_set_kernel_timing)
  std::vector<hipEvent_t> kt_ev;   // [calls][2 * kTimedSlots]
  std::vector<uint32_t> kt_mask;   // [calls] slots recorded
  int kt_calls = 0, kt_used = 0;
  int64_t zslot = 1 << 16;     // bytes per slot: largest min(qlen,2w+1)*tlen accepted
  // staging for the host-memory entry points
  DevBuf h_query, h_target, h_qoff, h_toff, h_i32, h_res;
  int32_t* err_host = nullptr;  // pinned: [0]=rejected count, [1]=first rejected index
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool pending_check = false;
  int pending_kind = 0;  // 1 = extension, 2 = global
  void* scratch = nullptr;  // this device's scratch array (allocation-free entry point)
  // tuning table (bsx_get_tuning layout): tpb_plain, tpb_narrow, shape variant of
  // the 65-96 and 97-128 bins; where it came from; the device's key in the cache
  int32_t tune[BSX_TUNE_FIELDS] = {kExtTasksPerBlock, kNarrowTasksPerBlock, 0, 0};
  int tune_source = BSX_TUNE_DEFAULT;
  char tune_key[64] = {0};
};

static int fail(bsx_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

static int hip_fail(bsx_ctx* c, hipError_t e, const char* what) {
  return fail(c, BSX_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(c, expr, what)                      \
  do {                                             \
    hipError_t _e = (expr);                        \
    if (_e != hipSuccess) return hip_fail(c, _e, what); \
  } while (0)

static int check_scoring(bsx_ctx* c, const bsx_scoring* sc) {
  if (!sc) return fail(c, BSX_EINVAL, "scoring is NULL");
  const int lim = 1 << 16;
  if (sc->e_del < 1 || sc->e_ins < 1 || sc->e_del > lim || sc->e_ins > lim)
    return fail(c, BSX_EINVAL, "gap extension penalties must be in [1, 65536] (got %d, %d)", sc->e_del,
                sc->e_ins);
  if (sc->o_del < 0 || sc->o_ins < 0 || sc->o_del > lim || sc->o_ins > lim)
    return fail(c, BSX_EINVAL, "gap open penalties must be in [0, 65536] (got %d, %d)", sc->o_del,
                sc->o_ins);
  return BSX_OK;
}

template <class Args>
static void fill_profile(Args& a, const bsx_scoring* sc) {
  for (int qc = 0; qc < 5; ++qc) {
    uint32_t lo = 0;
    for (int t = 0; t < 4; ++t) lo |= (uint32_t)(uint8_t)sc->mat[t * 5 + qc] << (8 * t);
    a.prof_lo[qc] = lo;
    a.prof_hi[qc] = sc->mat[20 + qc];
  }
}

static int mat_max(const bsx_scoring* sc) {
  int m = 0;  // ksw.c:400-401 starts the maximum at 0
  for (int k = 0; k < 25; ++k) m = sc->mat[k] > m ? sc->mat[k] : m;
  return m;
}

// ---- allocation-free extension entry point ----
// Workspace of bsx_extend_batch_device_noalloc: one file-scope device array
// per GPU, sized for BSX_NOALLOC_MAX_TASKS tasks and held until the process
// ends.  Host code never uses the array's name as an address (on the host it
// is a stand-in, not device memory): bsx_create looks up each device's
// address once with hipGetSymbolAddress and every use goes through it.
static constexpr size_t kNoallocN = (size_t)BSX_NOALLOC_MAX_TASKS + 1;
static constexpr size_t kScratchBytes = 4 * align_up(kNoallocN) +        // bin_of, nbin_of, need_plain, cls
                                        align_up(sizeof(ExtDesc) * kNoallocN) + align_up(32 * kNoallocN) +
                                        align_up(kSmallBytes);
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Address of this device's scratch array (the device must be current);
// nullptr if the lookup failed.
static void* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

static bool tune_valid(const int32_t* t) {
  return t[0] >= 16 && t[0] <= 4096 && t[1] >= 16 && t[1] <= 4096 && (t[2] == 0 || t[2] == 1) &&
         (t[3] == 0 || t[3] == 1);
}

// The autotuner's cache (written by `python -m bsbolt_amd.tune`): a JSON object
// keyed by "<gfx arch>/<CU count>", each entry an object holding
// "bsx_table": [tpb_plain, tpb_narrow, shape variant 65-96, shape variant
// 97-128] beside the tuner's timings.  The file is $BSX_TUNE, else tune.json
// next to libbsx.so; BSX_TUNE=off keeps the defaults.  An entry that does not
// parse, or a table outside the valid ranges, leaves the defaults.
static void load_tuning(bsx_ctx* c) {
  const char* env = getenv("BSX_TUNE");
  if (env && !strcmp(env, "off")) return;
  std::string path;
  if (env && *env) {
    path = env;
  } else {
    Dl_info info;
    if (!dladdr((void*)&load_tuning, &info) || !info.dli_fname) return;
    path = info.dli_fname;
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tune.json";
  }
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return;
  std::string js;
  char buf[4096];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, f)) > 0) js.append(buf, k);
  fclose(f);
  const std::string key = std::string("\"") + c->tune_key + "\"";
  const size_t p = js.find(key);
  if (p == std::string::npos) return;
  // the entry's table: the first "bsx_table": [n, n, n, n] after the device key
  // (the tuner writes each device's entry as one object holding it)
  const size_t q = js.find("\"bsx_table\"", p);
  const size_t br = q == std::string::npos ? q : js.find('[', q);
  int32_t t[BSX_TUNE_FIELDS];
  if (br == std::string::npos ||
      sscanf(js.c_str() + br + 1, " %d , %d , %d , %d", &t[0], &t[1], &t[2], &t[3]) != BSX_TUNE_FIELDS ||
      !tune_valid(t))
    return;
  memcpy(c->tune, t, sizeof t);
  c->tune_source = BSX_TUNE_CACHE;
}

extern "C" {

const char* bsx_version(void) { return "bsx 0.1 (gfx950)"; }

void bsx_default_scoring(bsx_scoring* sc) {
  // bwa_fill_scmat(a=1, b=4) (bwa.c:169-178); BSBolt defaults -O 6,6 -E 1,1 -d 100
  for (int r = 0; r < 5; ++r)
    for (int q = 0; q < 5; ++q) sc->mat[r * 5 + q] = (r == 4 || q == 4) ? -1 : (r == q ? 1 : -4);
  sc->o_del = sc->o_ins = 6;
  sc->e_del = sc->e_ins = 1;
  sc->zdrop = 100;
}

bsx_ctx* bsx_create(int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  bsx_ctx* c = new bsx_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return nullptr; }
  c->num_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) { delete c; return nullptr; }
  c->stream = c->own;
  {
    bool ok = hipEventCreateWithFlags(&c->ss.fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < StreamSet::kSide; ++k)
      ok = hipStreamCreateWithFlags(&c->ss.side[k], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&c->ss.join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {  // run everything on the call's stream
      for (int k = 0; k < StreamSet::kSide; ++k) {
        if (c->ss.side[k]) (void)hipStreamDestroy(c->ss.side[k]);
        if (c->ss.join[k]) (void)hipEventDestroy(c->ss.join[k]);
      }
      if (c->ss.fork) (void)hipEventDestroy(c->ss.fork);
      c->ss = StreamSet{};
    }
  }
  if (hipHostMalloc((void**)&c->err_host, 4 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipStreamDestroy(c->own);
    delete c;
    return nullptr;
  }
  c->err_host[0] = 0;
  c->err_host[1] = INT_MAX;
  c->err_host[2] = 0;
  c->scratch = scratch_address(device);  // once per device; nullptr leaves only the no-alloc call failing
  // tuning table: the cache entry for this device, else the defaults
  {
    char arch[64];
    snprintf(arch, sizeof arch, "%s", prop.gcnArchName);
    for (char* p = arch; *p; ++p)
      if (*p == ':') { *p = 0; break; }
    snprintf(c->tune_key, sizeof c->tune_key, "%s/%d", arch, c->num_cu);
    load_tuning(c);
  }
  return c;
}

void bsx_destroy(bsx_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->wown.release();
  for (DevBuf* b : {&c->h_query, &c->h_target, &c->h_qoff, &c->h_toff, &c->h_i32, &c->h_res, &c->stats, &c->model,
                    &c->part, &c->grad})
    b->release();
  for (hipEvent_t e : c->kt_ev) (void)hipEventDestroy(e);
  if (c->err_host) (void)hipHostFree(c->err_host);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (int k = 0; k < StreamSet::kSide; ++k) {
    if (c->ss.side[k]) {
      (void)hipStreamSynchronize(c->ss.side[k]);
      (void)hipStreamDestroy(c->ss.side[k]);
    }
    if (c->ss.join[k]) (void)hipEventDestroy(c->ss.join[k]);
  }
  if (c->ss.fork) (void)hipEventDestroy(c->ss.fork);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

const char* bsx_last_error(const bsx_ctx* c) { return c ? c->err.c_str() : "no context"; }

int bsx_set_stream(bsx_ctx* c, void* s) {
  if (!c) return BSX_EINVAL;
  c->stream = s ? (hipStream_t)s : c->own;
  return BSX_OK;
}

int bsx_sync(bsx_ctx* c) {
  if (!c) return BSX_EINVAL;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (c->pending_check) {
    c->pending_check = false;
    if (c->err_host[0] > 0)
      return fail(c, BSX_EINVAL, "%d task(s) rejected (precondition violated); first is task %d",
                  c->err_host[0], c->err_host[1]);
    if (c->pending_kind == 2 && c->err_host[2] > 0)
      return fail(c, BSX_ERANGE, "%d CIGAR(s) did not fit their slot", c->err_host[2]);
  }
  return BSX_OK;
}

static MlpDev model_view(bsx_ctx* c) {
  char* b = (char*)c->model.p;
  MlpDev m;
  m.params = (float*)b;
  m.adam_m = (float*)(b + kParamBytes);
  m.adam_v = (float*)(b + 2 * kParamBytes);
  size_t o = 3 * kParamBytes;
  m.w1p = (uint16_t*)(b + o); o += align_up(2 * 64 * kFeatPad);
  m.w2p = (uint16_t*)(b + o); o += align_up(2 * 64 * 64);
  m.w2t = (uint16_t*)(b + o); o += align_up(2 * 64 * 64);
  m.w3p = (uint16_t*)(b + o); o += align_up(2 * 16 * 64);
  m.w3t = (uint16_t*)(b + o);
  return m;
}

static ExtWork ext_work(bsx_ctx* c) {
  ExtWork wk;
  wk.bin_of = (uint8_t*)c->ws->bin_of.p;
  wk.desc = (ExtDesc*)c->ws->desc.p;
  wk.rec = (int32_t*)c->ws->rec.p;
  wk.counts = (uint32_t*)c->ws->small.p;
  wk.cursor = wk.counts + kCountSlots;
  wk.err = (int32_t*)(wk.cursor + kCountSlots);
  wk.fetch = (uint32_t*)(wk.err + 4);
  wk.split = wk.fetch + kFetchSlots;
  wk.bcount = wk.split + 2 * kLaneFrameClasses;
  wk.bcursor = wk.bcount + kExtBins * kRowBuckets;
  wk.fcount = wk.bcursor + kExtBins * kRowBuckets;
  wk.fcursor = wk.fcount + kLaneFrameClasses * kRowBuckets;
  wk.nbin_of = (uint8_t*)c->ws->nbin_of.p;
  wk.need_plain = (uint8_t*)c->ws->need_plain.p;
  wk.ss = c->ss_on ? &c->ss : nullptr;
  wk.tpb_plain = c->tune[0];
  wk.tpb_narrow = c->tune[1];
  wk.shape_var[0] = c->tune[2];
  wk.shape_var[1] = c->tune[3];
  return wk;
}

int bsx_set_tuning(bsx_ctx* c, const int32_t* table, int n) {
  if (!c) return BSX_EINVAL;
  if (!table) {  // back to the defaults
    const int32_t d[BSX_TUNE_FIELDS] = {kExtTasksPerBlock, kNarrowTasksPerBlock, 0, 0};
    memcpy(c->tune, d, sizeof d);
    c->tune_source = BSX_TUNE_DEFAULT;
    return BSX_OK;
  }
  if (n != BSX_TUNE_FIELDS || !tune_valid(table))
    return fail(c, BSX_EINVAL, "tuning table: %d fields, tpb in [16, 4096], shape variants 0 or 1", BSX_TUNE_FIELDS);
  memcpy(c->tune, table, sizeof c->tune);
  c->tune_source = BSX_TUNE_CALLER;
  return BSX_OK;
}

int bsx_get_tuning(const bsx_ctx* c, int32_t* table, int n, char* key, int key_cap) {
  if (!c || !table || n < BSX_TUNE_FIELDS) return BSX_EINVAL;
  memcpy(table, c->tune, sizeof c->tune);
  if (key && key_cap > 0) snprintf(key, (size_t)key_cap, "%s", c->tune_key);
  return c->tune_source;
}

static int ext_args(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t, ExtArgs& a) {
  int rc = check_scoring(c, sc);
  if (rc) return rc;
  if (t->n < 0 || t->n > INT_MAX) return fail(c, BSX_EINVAL, "n must be in [0, 2^31-1] (got %lld)", (long long)t->n);
  if (t->n > 0 && (!t->query || !t->q_off || !t->qlen || !t->target || !t->t_off || !t->tlen || !t->h0 ||
                   !t->w || !t->end_bonus))
    return fail(c, BSX_EINVAL, "task arrays must not be NULL");
  a = ExtArgs{};
  a.o_del = sc->o_del; a.e_del = sc->e_del; a.o_ins = sc->o_ins; a.e_ins = sc->e_ins;
  a.zdrop = sc->zdrop;
  a.mat_max = mat_max(sc);
  fill_profile(a, sc);
  a.n = t->n;
  a.query = t->query; a.q_off = t->q_off; a.qlen = t->qlen;
  a.target = t->target; a.t_off = t->t_off; a.tlen = t->tlen;
  a.h0 = t->h0; a.w = t->w; a.end_bonus = t->end_bonus;
  return BSX_OK;
}

static int ext_buffers(bsx_ctx* c, size_t n) {
  HIPCHK(c, c->ws->bin_of.ensure(n + 1), "hipMalloc(bin_of)");
  HIPCHK(c, c->ws->desc.ensure(sizeof(ExtDesc) * (n + 1)), "hipMalloc(desc)");
  HIPCHK(c, c->ws->rec.ensure(32 * (n + 1)), "hipMalloc(rec)");
  HIPCHK(c, c->ws->small.ensure(kSmallBytes), "hipMalloc(work)");
  HIPCHK(c, c->ws->nbin_of.ensure(n + 1), "hipMalloc(nbin_of)");
  HIPCHK(c, c->ws->need_plain.ensure(n + 1), "hipMalloc(need_plain)");
  return BSX_OK;
}

// Queues one extension batch on the context's stream with the workspace `wk`
// (ExtWork arrays sized for the batch) and `cls` ([n] predicted classes, used
// under the learned policy).
static int extend_queue(bsx_ctx* c, const ExtArgs& a, ExtWork wk, uint8_t* cls) {
  const size_t n = (size_t)a.n;
  if (c->kt_used < c->kt_calls) {  // per-kernel timing of this call
    wk.tev = &c->kt_ev[(size_t)c->kt_used * 2 * kTimedSlots];
    wk.tmask = &c->kt_mask[c->kt_used];
    *wk.tmask = 0;
    ++c->kt_used;
  }
  BandPlan plan{kBandPlain, 0, nullptr, false, false};
  if (c->band_policy == BSX_BAND_LEARNED) {
    if (!c->model_ready) return fail(c, BSX_EINVAL, "learned band policy without a band model (bsx_band_model_set)");
    time_begin(wk, kTimedSlots - 1, c->stream);
    HIPCHK(c, mlp_launch_predict_tasks(a, model_view(c), cls, c->num_cu, c->stream), "band predict");
    time_end(wk, kTimedSlots - 1, c->stream);
    plan = BandPlan{kBandLearned, 0, cls, false, false};
  } else if (c->band_policy == BSX_BAND_FIXED) {
    plan = BandPlan{kBandFixed, c->band_class, nullptr, true, false};
  }
  HIPCHK(c, launch_extend(a, wk, c->num_cu, c->stream, nullptr, &plan), "extension launch");
  HIPCHK(c, hipMemcpyAsync(c->err_host, wk.err, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
         "hipMemcpyAsync(err)");
  c->last_n = plan.mode == kBandPlain ? 0 : (int64_t)n;
  c->pending_check = true;
  c->pending_kind = 1;
  return BSX_OK;
}

int bsx_extend_batch_device(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t,
                            bsx_ext_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  ExtArgs a;
  int rc = ext_args(c, sc, t, a);
  if (rc) return rc;
  a.score = r->score; a.qle = r->qle; a.tle = r->tle; a.gtle = r->gtle; a.gscore = r->gscore;
  a.max_off = r->max_off; a.cells = r->cells;
  (void)hipSetDevice(c->device);
  const size_t n = (size_t)t->n;
  rc = ext_buffers(c, n);
  if (rc) return rc;
  if (c->band_policy == BSX_BAND_LEARNED) HIPCHK(c, c->ws->cls.ensure(n + 1), "hipMalloc(cls)");
  return extend_queue(c, a, ext_work(c), (uint8_t*)c->ws->cls.p);
}

int bsx_extend_batch_device_noalloc(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t,
                                    bsx_ext_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  if (t->n > BSX_NOALLOC_MAX_TASKS)
    return fail(c, BSX_EINVAL, "bsx_extend_batch_device_noalloc takes at most %d tasks (got %lld)",
                BSX_NOALLOC_MAX_TASKS, (long long)t->n);
  if (!c->scratch) return fail(c, BSX_EDEVICE, "no scratch array on device %d (hipGetSymbolAddress failed)", c->device);
  ExtArgs a;
  int rc = ext_args(c, sc, t, a);
  if (rc) return rc;
  a.score = r->score; a.qle = r->qle; a.tle = r->tle; a.gtle = r->gtle; a.gscore = r->gscore;
  a.max_off = r->max_off; a.cells = r->cells;
  (void)hipSetDevice(c->device);
  // carve the workspace out of the scratch array (layout of kScratchBytes);
  // every counter in it is cleared by the first kernel of the launch sequence
  char* b = (char*)c->scratch;
  ExtWork wk = ext_work(c);
  wk.bin_of = (uint8_t*)b; b += align_up(kNoallocN);
  wk.nbin_of = (uint8_t*)b; b += align_up(kNoallocN);
  wk.need_plain = (uint8_t*)b; b += align_up(kNoallocN);
  uint8_t* cls = (uint8_t*)b; b += align_up(kNoallocN);
  wk.desc = (ExtDesc*)b; b += align_up(sizeof(ExtDesc) * kNoallocN);
  wk.rec = (int32_t*)b; b += align_up(32 * kNoallocN);
  wk.counts = (uint32_t*)b;
  wk.cursor = wk.counts + kCountSlots;
  wk.err = (int32_t*)(wk.cursor + kCountSlots);
  wk.fetch = (uint32_t*)(wk.err + 4);
  wk.split = wk.fetch + kFetchSlots;
  wk.bcount = wk.split + 2 * kLaneFrameClasses;
  wk.bcursor = wk.bcount + kExtBins * kRowBuckets;
  wk.fcount = wk.bcursor + kExtBins * kRowBuckets;
  wk.fcursor = wk.fcount + kLaneFrameClasses * kRowBuckets;
  return extend_queue(c, a, wk, cls);
}

int bsx_set_side_streams(bsx_ctx* c, int enable) {
  if (!c || (enable != 0 && enable != 1)) return fail(c, BSX_EINVAL, "enable must be 0 or 1");
  c->ss_on = enable != 0;
  return BSX_OK;
}

int bsx_set_kernel_timing(bsx_ctx* c, int max_calls) {
  if (!c || max_calls < 0 || max_calls > 4096) return fail(c, BSX_EINVAL, "max_calls must be in [0, 4096]");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  const size_t need = (size_t)max_calls * 2 * kTimedSlots;
  while (c->kt_ev.size() < need) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e), "hipEventCreate");
    c->kt_ev.push_back(e);
  }
  c->kt_mask.assign((size_t)max_calls, 0u);
  c->kt_calls = max_calls;
  c->kt_used = 0;
  return BSX_OK;
}

int bsx_kernel_times(bsx_ctx* c, float* ms, int n_slots, int* calls) {
  if (!c || !ms || n_slots < 1) return fail(c, BSX_EINVAL, "bad argument");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (int s = 0; s < n_slots; ++s) ms[s] = 0.f;
  for (int k = 0; k < c->kt_used; ++k)
    for (int s = 0; s < kTimedSlots && s < n_slots; ++s) {
      if (!(c->kt_mask[k] >> s & 1u)) continue;
      float t = 0.f;
      const hipEvent_t* e = &c->kt_ev[(size_t)k * 2 * kTimedSlots + 2 * s];
      HIPCHK(c, hipEventElapsedTime(&t, e[0], e[1]), "hipEventElapsedTime");
      ms[s] += t;
    }
  if (calls) *calls = c->kt_used;
  c->kt_used = 0;
  return BSX_OK;
}

int bsx_ext_bin_limits(int32_t* limits, int n) {
  if (!limits) return kExtBins;
  for (int b = 0; b < kExtBins && b < n; ++b) limits[b] = bin_limit(b);
  return kExtBins;
}

int bsx_set_band_policy(bsx_ctx* c, int policy, int band_class) {
  if (!c) return BSX_EINVAL;
  if (policy != BSX_BAND_PLAIN && policy != BSX_BAND_LEARNED && policy != BSX_BAND_FIXED)
    return fail(c, BSX_EINVAL, "unknown band policy %d", policy);
  if (policy == BSX_BAND_FIXED && (band_class < 0 || band_class >= kBandClasses - 1))
    return fail(c, BSX_EINVAL, "fixed band class must be in [0, %d] (got %d)", kBandClasses - 2, band_class);
  c->band_policy = policy;
  c->band_class = band_class;
  return BSX_OK;
}

int bsx_band_stats(bsx_ctx* c, int64_t* out) {
  if (!c || !out) return fail(c, BSX_EINVAL, "NULL argument");
  if (c->last_n <= 0) {
    for (int k = 0; k < 2 * kBandClasses; ++k) out[k] = 0;
    return BSX_OK;
  }
  (void)hipSetDevice(c->device);
  HIPCHK(c, c->stats.ensure(2 * kBandClasses * sizeof(unsigned long long)), "hipMalloc(stats)");
  ExtWork wk = ext_work(c);
  HIPCHK(c, launch_band_stats(c->last_n, wk, (unsigned long long*)c->stats.p, c->num_cu, c->stream), "band stats");
  unsigned long long h[2 * kBandClasses];
  HIPCHK(c, hipMemcpyAsync(h, c->stats.p, sizeof h, hipMemcpyDeviceToHost, c->stream), "D2H stats");
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  for (int k = 0; k < 2 * kBandClasses; ++k) out[k] = (int64_t)h[k];
  return BSX_OK;
}

int bsx_band_model_set(bsx_ctx* c, const float* params, int64_t n) {
  if (!c || !params) return fail(c, BSX_EINVAL, "NULL argument");
  if (n != kParamCount) return fail(c, BSX_EINVAL, "band model has %d parameters (got %lld)", kParamCount, (long long)n);
  for (int64_t k = 0; k < n; ++k)
    if (!(params[k] > -1e30f && params[k] < 1e30f)) return fail(c, BSX_EINVAL, "parameter %lld is not finite", (long long)k);
  (void)hipSetDevice(c->device);
  HIPCHK(c, c->model.ensure(kModelBytes), "hipMalloc(model)");
  MlpDev m = model_view(c);
  HIPCHK(c, hipMemcpyAsync(m.params, params, 4 * (size_t)kParamCount, hipMemcpyHostToDevice, c->stream), "H2D model");
  std::vector<float> zero(kParamCount, 0.f);
  HIPCHK(c, hipMemcpyAsync(m.adam_m, zero.data(), 4 * (size_t)kParamCount, hipMemcpyHostToDevice, c->stream), "H2D adam");
  HIPCHK(c, hipMemcpyAsync(m.adam_v, zero.data(), 4 * (size_t)kParamCount, hipMemcpyHostToDevice, c->stream), "H2D adam");
  HIPCHK(c, mlp_launch_pack(m, c->stream), "model pack");
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  c->model_ready = true;
  c->adam_step = 0;
  return BSX_OK;
}

int bsx_band_model_get(bsx_ctx* c, float* params, int64_t n) {
  if (!c || !params) return fail(c, BSX_EINVAL, "NULL argument");
  if (n != kParamCount) return fail(c, BSX_EINVAL, "band model has %d parameters (got %lld)", kParamCount, (long long)n);
  if (!c->model_ready) return fail(c, BSX_EINVAL, "no band model");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipMemcpyAsync(params, model_view(c).params, 4 * (size_t)kParamCount, hipMemcpyDeviceToHost, c->stream),
         "D2H model");
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return BSX_OK;
}

int bsx_band_features_device(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t, uint16_t* feats) {
  if (!c || !t || (t->n > 0 && !feats)) return fail(c, BSX_EINVAL, "NULL argument");
  ExtArgs a;
  int rc = ext_args(c, sc, t, a);
  if (rc) return rc;
  (void)hipSetDevice(c->device);
  HIPCHK(c, mlp_launch_features(a, feats, c->num_cu, c->stream), "band features");
  return BSX_OK;
}

int bsx_band_labels_device(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t, uint8_t* labels) {
  if (!c || !t || (t->n > 0 && !labels)) return fail(c, BSX_EINVAL, "NULL argument");
  ExtArgs a;
  int rc = ext_args(c, sc, t, a);
  if (rc) return rc;
  if (a.n == 0) return BSX_OK;
  (void)hipSetDevice(c->device);
  rc = ext_buffers(c, (size_t)a.n);
  if (rc) return rc;
  ExtWork wk = ext_work(c);
  for (int cl = kBandClasses - 2; cl >= 0; --cl) {
    BandPlan plan{kBandFixed, cl, nullptr, true, true};
    HIPCHK(c, launch_extend(a, wk, c->num_cu, c->stream, nullptr, &plan), "narrow launch");
    HIPCHK(c, launch_band_label_merge(a, wk, cl, labels, cl == kBandClasses - 2, c->num_cu, c->stream), "label merge");
  }
  HIPCHK(c, hipMemcpyAsync(c->err_host, wk.err, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
         "hipMemcpyAsync(err)");
  c->last_n = 0;
  c->pending_check = true;
  c->pending_kind = 1;
  return BSX_OK;
}

int bsx_band_forward_device(bsx_ctx* c, const uint16_t* feats, int64_t n, float* logits, uint8_t* cls) {
  if (!c || (n > 0 && !feats) || n < 0) return fail(c, BSX_EINVAL, "bad argument");
  if (!c->model_ready) return fail(c, BSX_EINVAL, "no band model");
  (void)hipSetDevice(c->device);
  HIPCHK(c, mlp_launch_forward(model_view(c), feats, n, logits, cls, c->num_cu, c->stream), "band forward");
  return BSX_OK;
}

int bsx_band_predict_device(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t, uint8_t* cls) {
  if (!c || !t || (t->n > 0 && !cls)) return fail(c, BSX_EINVAL, "NULL argument");
  if (!c->model_ready) return fail(c, BSX_EINVAL, "no band model");
  ExtArgs a;
  int rc = ext_args(c, sc, t, a);
  if (rc) return rc;
  (void)hipSetDevice(c->device);
  HIPCHK(c, mlp_launch_predict_tasks(a, model_view(c), cls, c->num_cu, c->stream), "band predict");
  return BSX_OK;
}

static int band_grad(bsx_ctx* c, const uint16_t* feats, const uint8_t* labels, int64_t n, float* grad) {
  HIPCHK(c, c->part.ensure(4 * (size_t)(kParamCount + 1) * (size_t)mlp_grad_waves(n)), "hipMalloc(part)");
  HIPCHK(c, mlp_launch_grad(model_view(c), feats, labels, n, (float*)c->part.p, grad, c->stream), "band grad");
  return BSX_OK;
}

int bsx_band_grad_device(bsx_ctx* c, const uint16_t* feats, const uint8_t* labels, int64_t n, float* grad) {
  if (!c || !grad || n < 1 || !feats || !labels) return fail(c, BSX_EINVAL, "bad argument");
  if (!c->model_ready) return fail(c, BSX_EINVAL, "no band model");
  (void)hipSetDevice(c->device);
  return band_grad(c, feats, labels, n, grad);
}

int bsx_band_train_device(bsx_ctx* c, const uint16_t* feats, const uint8_t* labels, int64_t n, int64_t batch,
                          int steps, float lr, float* losses) {
  if (!c || n < 1 || !feats || !labels || batch < 1 || steps < 0 || !(lr > 0.f))
    return fail(c, BSX_EINVAL, "bad argument");
  if (!c->model_ready) return fail(c, BSX_EINVAL, "no band model");
  (void)hipSetDevice(c->device);
  HIPCHK(c, c->grad.ensure(4 * (size_t)(kParamCount + 1)), "hipMalloc(grad)");
  float* g = (float*)c->grad.p;
  int64_t start = 0;
  for (int s = 0; s < steps; ++s) {
    if (start >= n) start = 0;
    const int64_t nb = batch < n - start ? batch : n - start;
    int rc = band_grad(c, feats + start * kFeatPad, labels + start, nb, g);
    if (rc) return rc;
    HIPCHK(c, mlp_launch_adam(model_view(c), g, lr, ++c->adam_step, c->stream), "adam");
    if (losses)
      HIPCHK(c, hipMemcpyAsync(losses + s, g + kParamCount, 4, hipMemcpyDeviceToDevice, c->stream), "loss copy");
    start += nb;
  }
  return BSX_OK;
}

int bsx_extend_batch(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t, bsx_ext_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  int rc = check_scoring(c, sc);
  if (rc) return rc;
  if (t->n < 0 || t->n > INT_MAX) return fail(c, BSX_EINVAL, "n must be in [0, 2^31-1]");
  const size_t n = (size_t)t->n;
  if (n == 0) return BSX_OK;
  if (!t->query || !t->q_off || !t->qlen || !t->target || !t->t_off || !t->tlen || !t->h0 || !t->w ||
      !t->end_bonus)
    return fail(c, BSX_EINVAL, "task arrays must not be NULL");
  // extent of the sequence buffers actually referenced
  uint64_t qext = 0, text = 0;
  for (size_t k = 0; k < n; ++k) {
    if (t->qlen[k] > 0 && t->q_off[k] + (uint64_t)t->qlen[k] > qext) qext = t->q_off[k] + t->qlen[k];
    if (t->tlen[k] > 0 && t->t_off[k] + (uint64_t)t->tlen[k] > text) text = t->t_off[k] + t->tlen[k];
  }
  (void)hipSetDevice(c->device);
  HIPCHK(c, c->h_query.ensure(qext + 1), "hipMalloc(query)");
  HIPCHK(c, c->h_target.ensure(text + 1), "hipMalloc(target)");
  HIPCHK(c, c->h_qoff.ensure(8 * n), "hipMalloc(q_off)");
  HIPCHK(c, c->h_toff.ensure(8 * n), "hipMalloc(t_off)");
  HIPCHK(c, c->h_i32.ensure(5 * 4 * n), "hipMalloc(task ints)");
  HIPCHK(c, c->h_res.ensure(8 * 4 * n + 8 * n), "hipMalloc(results)");
  hipStream_t st = c->stream;
  uint8_t* dq = (uint8_t*)c->h_query.p;
  uint8_t* dt = (uint8_t*)c->h_target.p;
  int32_t* di = (int32_t*)c->h_i32.p;
  if (qext) HIPCHK(c, hipMemcpyAsync(dq, t->query, qext, hipMemcpyHostToDevice, st), "H2D query");
  if (text) HIPCHK(c, hipMemcpyAsync(dt, t->target, text, hipMemcpyHostToDevice, st), "H2D target");
  HIPCHK(c, hipMemcpyAsync(c->h_qoff.p, t->q_off, 8 * n, hipMemcpyHostToDevice, st), "H2D q_off");
  HIPCHK(c, hipMemcpyAsync(c->h_toff.p, t->t_off, 8 * n, hipMemcpyHostToDevice, st), "H2D t_off");
  const int32_t* src[5] = {t->qlen, t->tlen, t->h0, t->w, t->end_bonus};
  for (int k = 0; k < 5; ++k)
    HIPCHK(c, hipMemcpyAsync(di + k * n, src[k], 4 * n, hipMemcpyHostToDevice, st), "H2D task ints");
  bsx_ext_tasks dtk{(int64_t)n, dq, (const uint64_t*)c->h_qoff.p, di, dt, (const uint64_t*)c->h_toff.p,
                    di + n, di + 2 * n, di + 3 * n, di + 4 * n};
  int32_t* dr = (int32_t*)c->h_res.p;
  int32_t* outs[6] = {r->score, r->qle, r->tle, r->gtle, r->gscore, r->max_off};
  bsx_ext_results drs{};
  int32_t** dslots[6] = {&drs.score, &drs.qle, &drs.tle, &drs.gtle, &drs.gscore, &drs.max_off};
  for (int k = 0; k < 6; ++k) *dslots[k] = outs[k] ? dr + k * n : nullptr;
  drs.cells = r->cells ? (int64_t*)(dr + 8 * n) : nullptr;
  rc = bsx_extend_batch_device(c, sc, &dtk, &drs);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(st), "hipStreamSynchronize");
  c->pending_check = false;
  const int rejected = c->err_host[0];
  std::vector<uint8_t> ok;
  if (rejected > 0) {  // copy back only the tasks that ran; rejected slots stay untouched
    ok.resize(n);
    HIPCHK(c, hipMemcpy(ok.data(), c->ws->bin_of.p, n, hipMemcpyDeviceToHost), "D2H bin_of");
  }
  std::vector<int32_t> tmp;
  std::vector<int64_t> tmp64;
  for (int k = 0; k < 6; ++k) {
    if (!outs[k]) continue;
    if (rejected == 0) {
      HIPCHK(c, hipMemcpy(outs[k], dr + k * n, 4 * n, hipMemcpyDeviceToHost), "D2H results");
    } else {
      tmp.resize(n);
      HIPCHK(c, hipMemcpy(tmp.data(), dr + k * n, 4 * n, hipMemcpyDeviceToHost), "D2H results");
      for (size_t q = 0; q < n; ++q)
        if (ok[q] != 0xff) outs[k][q] = tmp[q];
    }
  }
  if (r->cells) {
    if (rejected == 0) {
      HIPCHK(c, hipMemcpy(r->cells, drs.cells, 8 * n, hipMemcpyDeviceToHost), "D2H cells");
    } else {
      tmp64.resize(n);
      HIPCHK(c, hipMemcpy(tmp64.data(), drs.cells, 8 * n, hipMemcpyDeviceToHost), "D2H cells");
      for (size_t q = 0; q < n; ++q)
        if (ok[q] != 0xff) r->cells[q] = tmp64[q];
    }
  }
  if (rejected > 0)
    return fail(c, BSX_EINVAL, "%d task(s) rejected (precondition violated); first is task %d", rejected,
                c->err_host[1]);
  return BSX_OK;
}

void* bsx_dev_alloc(bsx_ctx* c, size_t bytes) {
  if (!c) return nullptr;
  (void)hipSetDevice(c->device);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    hip_fail(c, e, "hipMalloc");
    return nullptr;
  }
  return p;
}

int bsx_dev_free(bsx_ctx* c, void* p) {
  if (!c) return BSX_EINVAL;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  if (p) HIPCHK(c, hipFree(p), "hipFree");
  return BSX_OK;
}

int bsx_memcpy(bsx_ctx* c, void* dst, const void* src, size_t bytes, int kind) {
  if (!c || (bytes && (!dst || !src))) return fail(c, BSX_EINVAL, "bsx_memcpy: NULL argument");
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                  : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
  if (kind < 1 || kind > 3) return fail(c, BSX_EINVAL, "bsx_memcpy: kind must be 1, 2 or 3");
  (void)hipSetDevice(c->device);
  if (bytes) HIPCHK(c, hipMemcpyAsync(dst, src, bytes, k, c->stream), "hipMemcpyAsync");
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return BSX_OK;
}

int bsx_timer_start(bsx_ctx* c) {
  if (!c) return BSX_EINVAL;
  (void)hipSetDevice(c->device);
  if (!c->ev0) HIPCHK(c, hipEventCreate(&c->ev0), "hipEventCreate");
  if (!c->ev1) HIPCHK(c, hipEventCreate(&c->ev1), "hipEventCreate");
  HIPCHK(c, hipEventRecord(c->ev0, c->stream), "hipEventRecord");
  return BSX_OK;
}

int bsx_timer_stop(bsx_ctx* c, float* ms) {
  if (!c || !ms || !c->ev0) return fail(c, BSX_EINVAL, "bsx_timer_stop without bsx_timer_start");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipEventRecord(c->ev1, c->stream), "hipEventRecord");
  HIPCHK(c, hipEventSynchronize(c->ev1), "hipEventSynchronize");
  HIPCHK(c, hipEventElapsedTime(ms, c->ev0, c->ev1), "hipEventElapsedTime");
  return BSX_OK;
}

// Largest direction slot the host-memory CIGAR call sets by itself: 256 KiB,
// a task of up to 256 Ki band cells (e.g. a 500-base read with a 201-column
// band); the scratch is #CU * 192 slots, 12 GB on 256 CUs at this size.
static constexpr int64_t kZslotHostMax = 1 << 18;

void bsx_set_global_zslot(bsx_ctx* c, int64_t bytes) {
  if (c && bytes >= 64) c->zslot = bytes & ~(int64_t)3;  // whole dwords: the backtrack reads dword-wise
}

int bsx_md_batch_device(bsx_ctx* c, const bsx_md_tasks* t, bsx_md_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  if (t->n < 0 || t->n > INT_MAX) return fail(c, BSX_EINVAL, "n must be in [0, 2^31-1]");
  if (t->n == 0) return BSX_OK;
  if (!t->opac || !t->rb || !t->re || !t->oquery || !t->oq_off || !t->l_query || !t->n_cigar || !t->cigar)
    return fail(c, BSX_EINVAL, "task arrays must not be NULL");
  if (!r->md || !r->md_len || !r->xb || !r->xb_len || !r->nm || !r->meth || !r->status)
    return fail(c, BSX_EINVAL, "result arrays must not be NULL");
  if (t->l_pac < 1 || t->crick_l < 0 || t->crick_l > t->l_pac || t->cigar_slot < 1 || r->md_slot < 1 ||
      r->xb_slot < 1 || r->md_slot > INT_MAX || r->xb_slot > INT_MAX)
    return fail(c, BSX_EINVAL, "need l_pac >= 1, 0 <= crick_l <= l_pac, slots >= 1");
  (void)hipSetDevice(c->device);
  MdArgs a{};
  a.n = t->n; a.l_pac = t->l_pac; a.crick_l = t->crick_l; a.opac = t->opac;
  a.rb = t->rb; a.re = t->re; a.oquery = t->oquery; a.oq_off = t->oq_off; a.l_query = t->l_query;
  a.n_cigar = t->n_cigar; a.cigar = t->cigar; a.cigar_slot = t->cigar_slot;
  a.md = r->md; a.md_slot = r->md_slot; a.md_len = r->md_len;
  a.xb = r->xb; a.xb_slot = r->xb_slot; a.xb_len = r->xb_len;
  a.nm = r->nm; a.meth = r->meth; a.status = r->status;
  HIPCHK(c, launch_md(a, c->num_cu, c->stream), "md launch");
  return BSX_OK;
}

int bsx_reg2aln_cigar_device(bsx_ctx* c, const bsx_reg2aln_tasks* t, bsx_reg2aln_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  if (t->n < 0 || t->n > INT_MAX) return fail(c, BSX_EINVAL, "n must be in [0, 2^31-1]");
  if (t->n == 0) return BSX_OK;
  if (!t->rb || !t->re || !t->qb || !t->qe || !t->l_seq || !t->n_cigar || !t->cigar || !r->cigar || !r->n_cigar ||
      !r->pos || !r->is_rev)
    return fail(c, BSX_EINVAL, "arrays must not be NULL");
  if (t->l_pac < 1 || t->cigar_slot < 1 || r->cigar_slot < 1)
    return fail(c, BSX_EINVAL, "need l_pac >= 1 and slots >= 1");
  (void)hipSetDevice(c->device);
  Reg2alnArgs a{};
  a.n = t->n; a.l_pac = t->l_pac; a.rb = t->rb; a.re = t->re; a.qb = t->qb; a.qe = t->qe; a.l_seq = t->l_seq;
  a.n_cigar = t->n_cigar; a.cigar = t->cigar; a.cigar_slot = t->cigar_slot;
  a.out = r->cigar; a.out_slot = r->cigar_slot; a.out_n = r->n_cigar; a.pos = r->pos; a.is_rev = r->is_rev;
  HIPCHK(c, launch_reg2aln_cigar(a, c->num_cu, c->stream), "reg2aln cigar launch");
  return BSX_OK;
}

int bsx_global_batch_device(bsx_ctx* c, const bsx_scoring* sc, const bsx_glb_tasks* t,
                            bsx_glb_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  int rc = check_scoring(c, sc);
  if (rc) return rc;
  if (t->n < 0 || t->n > INT_MAX) return fail(c, BSX_EINVAL, "n must be in [0, 2^31-1]");
  if (t->n > 0 && (!t->query || !t->q_off || !t->qlen || !t->target || !t->t_off || !t->tlen || !t->w))
    return fail(c, BSX_EINVAL, "task arrays must not be NULL");
  if (r->cigar && (!r->n_cigar || r->cigar_slot < 1))
    return fail(c, BSX_EINVAL, "a CIGAR request needs n_cigar and cigar_slot >= 1");
  (void)hipSetDevice(c->device);
  GlbArgs a{};
  a.o_del = sc->o_del; a.e_del = sc->e_del; a.o_ins = sc->o_ins; a.e_ins = sc->e_ins;
  fill_profile(a, sc);
  a.n = t->n;
  a.query = t->query; a.q_off = t->q_off; a.qlen = t->qlen;
  a.target = t->target; a.t_off = t->t_off; a.tlen = t->tlen; a.w = t->w;
  a.score = r->score; a.n_cigar = r->n_cigar; a.cigar = r->cigar; a.cigar_slot = r->cigar_slot;
  const size_t n = (size_t)t->n;
  HIPCHK(c, c->ws->bin_of.ensure(n + 1), "hipMalloc(bin_of)");
  HIPCHK(c, c->ws->list.ensure(4 * n + 4), "hipMalloc(list)");
  HIPCHK(c, c->ws->small.ensure(kSmallBytes), "hipMalloc(work)");
  GlbWork wk;
  wk.bin_of = (uint8_t*)c->ws->bin_of.p;
  wk.list = (int32_t*)c->ws->list.p;
  wk.counts = (uint32_t*)c->ws->small.p;
  wk.cursor = wk.counts + kCountSlots;
  wk.err = (int32_t*)(wk.cursor + kCountSlots);
  wk.fetch = (uint32_t*)(wk.err + 4);
  wk.zslot = c->zslot;
  wk.ss = c->ss_on ? &c->ss : nullptr;
  wk.z = nullptr;
  if (r->cigar) {
    HIPCHK(c, c->ws->zbuf.ensure((size_t)glb_groups(c->num_cu, c->ss_on) * (size_t)c->zslot), "hipMalloc(z)");
    wk.z = (uint8_t*)c->ws->zbuf.p;
  }
  HIPCHK(c, launch_global_impl(a, wk, c->num_cu, c->stream), "global launch");
  HIPCHK(c, hipMemcpyAsync(c->err_host, wk.err, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream),
         "hipMemcpyAsync(err)");
  c->pending_check = true;
  c->pending_kind = 2;
  return BSX_OK;
}

int bsx_global_batch(bsx_ctx* c, const bsx_scoring* sc, const bsx_glb_tasks* t, bsx_glb_results* r) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  int rc = check_scoring(c, sc);
  if (rc) return rc;
  if (t->n < 0 || t->n > INT_MAX) return fail(c, BSX_EINVAL, "n must be in [0, 2^31-1]");
  const size_t n = (size_t)t->n;
  if (n == 0) return BSX_OK;
  if (!t->query || !t->q_off || !t->qlen || !t->target || !t->t_off || !t->tlen || !t->w)
    return fail(c, BSX_EINVAL, "task arrays must not be NULL");
  if (r->cigar && (!r->n_cigar || r->cigar_slot < 1))
    return fail(c, BSX_EINVAL, "a CIGAR request needs n_cigar and cigar_slot >= 1");
  uint64_t qext = 0, text = 0;
  for (size_t k = 0; k < n; ++k) {
    if (t->qlen[k] > 0 && t->q_off[k] + (uint64_t)t->qlen[k] > qext) qext = t->q_off[k] + t->qlen[k];
    if (t->tlen[k] > 0 && t->t_off[k] + (uint64_t)t->tlen[k] > text) text = t->t_off[k] + t->tlen[k];
  }
  (void)hipSetDevice(c->device);
  const size_t slot = r->cigar ? (size_t)r->cigar_slot : 0;
  HIPCHK(c, c->h_query.ensure(qext + 1), "hipMalloc(query)");
  HIPCHK(c, c->h_target.ensure(text + 1), "hipMalloc(target)");
  HIPCHK(c, c->h_qoff.ensure(8 * n), "hipMalloc(q_off)");
  HIPCHK(c, c->h_toff.ensure(8 * n), "hipMalloc(t_off)");
  HIPCHK(c, c->h_i32.ensure(3 * 4 * n), "hipMalloc(task ints)");
  HIPCHK(c, c->h_res.ensure(2 * 4 * n + 4 * n * slot), "hipMalloc(results)");
  hipStream_t st = c->stream;
  uint8_t* dq = (uint8_t*)c->h_query.p;
  uint8_t* dt = (uint8_t*)c->h_target.p;
  int32_t* di = (int32_t*)c->h_i32.p;
  if (qext) HIPCHK(c, hipMemcpyAsync(dq, t->query, qext, hipMemcpyHostToDevice, st), "H2D query");
  if (text) HIPCHK(c, hipMemcpyAsync(dt, t->target, text, hipMemcpyHostToDevice, st), "H2D target");
  HIPCHK(c, hipMemcpyAsync(c->h_qoff.p, t->q_off, 8 * n, hipMemcpyHostToDevice, st), "H2D q_off");
  HIPCHK(c, hipMemcpyAsync(c->h_toff.p, t->t_off, 8 * n, hipMemcpyHostToDevice, st), "H2D t_off");
  const int32_t* src[3] = {t->qlen, t->tlen, t->w};
  for (int k = 0; k < 3; ++k)
    HIPCHK(c, hipMemcpyAsync(di + k * n, src[k], 4 * n, hipMemcpyHostToDevice, st), "H2D task ints");
  bsx_glb_tasks dtk{(int64_t)n, dq, (const uint64_t*)c->h_qoff.p, di, dt, (const uint64_t*)c->h_toff.p,
                    di + n, di + 2 * n};
  int32_t* dr = (int32_t*)c->h_res.p;
  bsx_glb_results drs{r->score ? dr : nullptr, r->n_cigar ? dr + n : nullptr,
                      r->cigar ? (uint32_t*)(dr + 2 * n) : nullptr, r->cigar_slot};
  // The host sees every task's size, so the direction slot is sized for this
  // batch: raised (for this call only) to the largest min(qlen, 2w+1) * tlen,
  // within a scratch budget of kZslotHostMax bytes per slot.
  const int64_t zslot0 = c->zslot;
  if (r->cigar) {
    int64_t need = 0;
    for (size_t k = 0; k < n; ++k) {
      const int64_t ql = t->qlen[k], tl = t->tlen[k], w = t->w[k] < 0 ? 0 : t->w[k];
      const int64_t ncol = ql < 2 * w + 1 ? ql : 2 * w + 1;
      if (ql > 0 && tl > 0 && ncol * tl > need) need = ncol * tl;
    }
    if (need > c->zslot && need <= kZslotHostMax) c->zslot = (need + 3) & ~(int64_t)3;
  }
  rc = bsx_global_batch_device(c, sc, &dtk, &drs);
  c->zslot = zslot0;
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(st), "hipStreamSynchronize");
  c->pending_check = false;
  const int rejected = c->err_host[0], overflow = c->err_host[2];
  std::vector<uint8_t> ok(n, 0);
  if (rejected > 0) HIPCHK(c, hipMemcpy(ok.data(), c->ws->bin_of.p, n, hipMemcpyDeviceToHost), "D2H bin_of");
  std::vector<int32_t> sco(n), ncg(n);
  std::vector<uint32_t> cg(n * slot);
  if (r->score) HIPCHK(c, hipMemcpy(sco.data(), drs.score, 4 * n, hipMemcpyDeviceToHost), "D2H score");
  if (r->n_cigar) HIPCHK(c, hipMemcpy(ncg.data(), drs.n_cigar, 4 * n, hipMemcpyDeviceToHost), "D2H n_cigar");
  if (r->cigar) HIPCHK(c, hipMemcpy(cg.data(), drs.cigar, 4 * n * slot, hipMemcpyDeviceToHost), "D2H cigar");
  for (size_t q = 0; q < n; ++q) {
    if (ok[q] == 0xff) continue;
    if (r->score) r->score[q] = sco[q];
    if (r->n_cigar) r->n_cigar[q] = ncg[q];
    if (r->cigar && ncg[q] > 0) memcpy(r->cigar + q * slot, cg.data() + q * slot, 4 * (size_t)ncg[q]);
  }
  if (rejected > 0)
    return fail(c, BSX_EINVAL, "%d task(s) rejected (precondition violated); first is task %d", rejected,
                c->err_host[1]);
  if (overflow > 0) return fail(c, BSX_ERANGE, "%d CIGAR(s) did not fit their slot", overflow);
  return BSX_OK;
}

}  // extern "C"

struct bsx_graph {
  bsx_ctx* ctx = nullptr;
  Work work;                  // the graph's own workspace
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int kind = 0;               // 1 = extension, 2 = global
};

// Captures one device-memory batch call on the context's stream into a graph
// that uses its own workspace; `sized` grows that workspace beforehand, since
// no allocation may happen while the stream is captured.
template <class Sized, class Call>
static int capture(bsx_ctx* c, int kind, Sized sized, Call call, bsx_graph** out) {
  if (!c || !out) return fail(c, BSX_EINVAL, "NULL argument");
  *out = nullptr;
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  bsx_graph* g = new bsx_graph();
  g->ctx = c;
  g->kind = kind;
  const int kt_calls = c->kt_calls;
  c->kt_calls = 0;  // no per-kernel timing events inside a graph
  c->ws = &g->work;
  int rc = sized();
  if (rc == BSX_OK) {
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
      rc = hip_fail(c, e, "hipStreamBeginCapture");
    } else {
      rc = call();
      hipGraph_t gr = nullptr;
      e = hipStreamEndCapture(c->stream, &gr);
      g->graph = gr;
      if (rc == BSX_OK && e != hipSuccess) rc = hip_fail(c, e, "hipStreamEndCapture");
      if (rc == BSX_OK) {
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(c, e, "hipGraphInstantiate");
      }
    }
  }
  c->ws = &c->wown;
  c->kt_calls = kt_calls;
  c->pending_check = false;
  if (rc != BSX_OK) {
    bsx_graph_destroy(g);
    return rc;
  }
  *out = g;
  return BSX_OK;
}

extern "C" {

int bsx_graph_capture_extend(bsx_ctx* c, const bsx_scoring* sc, const bsx_ext_tasks* t, bsx_ext_results* r,
                             bsx_graph** out) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  auto sized = [&]() -> int {
    const size_t n = t->n > 0 ? (size_t)t->n : 0;
    int rc = ext_buffers(c, n);
    if (rc) return rc;
    if (c->band_policy == BSX_BAND_LEARNED) HIPCHK(c, c->ws->cls.ensure(n + 1), "hipMalloc(cls)");
    return BSX_OK;
  };
  return capture(c, 1, sized, [&]() { return bsx_extend_batch_device(c, sc, t, r); }, out);
}

int bsx_graph_capture_global(bsx_ctx* c, const bsx_scoring* sc, const bsx_glb_tasks* t, bsx_glb_results* r,
                             bsx_graph** out) {
  if (!c || !t || !r) return fail(c, BSX_EINVAL, "NULL argument");
  auto sized = [&]() -> int {
    const size_t n = t->n > 0 ? (size_t)t->n : 0;
    HIPCHK(c, c->ws->bin_of.ensure(n + 1), "hipMalloc(bin_of)");
    HIPCHK(c, c->ws->list.ensure(4 * n + 4), "hipMalloc(list)");
    HIPCHK(c, c->ws->small.ensure(kSmallBytes), "hipMalloc(work)");
    if (r->cigar) HIPCHK(c, c->ws->zbuf.ensure((size_t)glb_groups(c->num_cu, c->ss_on) * (size_t)c->zslot), "hipMalloc(z)");
    return BSX_OK;
  };
  return capture(c, 2, sized, [&]() { return bsx_global_batch_device(c, sc, t, r); }, out);
}

int bsx_graph_launch(bsx_ctx* c, bsx_graph* g) {
  if (!c || !g || g->ctx != c || !g->exec) return fail(c, BSX_EINVAL, "graph does not belong to this context");
  (void)hipSetDevice(c->device);
  HIPCHK(c, hipGraphLaunch(g->exec, c->stream), "hipGraphLaunch");
  c->pending_check = true;
  c->pending_kind = g->kind;
  return BSX_OK;
}

void bsx_graph_destroy(bsx_graph* g) {
  if (!g) return;
  if (g->ctx) {
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->work.release();
  delete g;
}

// Contexts of the scalar drop-in.  A call borrows one from a process-wide
// free list and returns it, so the number of contexts is the peak number of
// concurrent callers, not the number of threads that ever called: the
// reference's kt_for (kthread.c:49-58) starts fresh worker threads for every
// batch of reads, and a context per thread would grow without bound.  The
// contexts live until the process ends.
static std::mutex g_dropin_mu;
static std::vector<bsx_ctx*> g_dropin_free;
static int64_t g_dropin_made = 0;

static bsx_ctx* dropin_acquire() {
  {
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    if (!g_dropin_free.empty()) {
      bsx_ctx* c = g_dropin_free.back();
      g_dropin_free.pop_back();
      return c;
    }
  }
  bsx_ctx* c = bsx_create(0);
  if (c) {
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    ++g_dropin_made;
  }
  return c;
}

static void dropin_release(bsx_ctx* c) {
  std::lock_guard<std::mutex> lk(g_dropin_mu);
  g_dropin_free.push_back(c);
}

int64_t bsx_dropin_contexts(void) {
  std::lock_guard<std::mutex> lk(g_dropin_mu);
  return g_dropin_made;
}

int bsx_ksw_extend2(int qlen, const uint8_t* query, int tlen, const uint8_t* target, int m,
                    const int8_t* mat, int o_del, int e_del, int o_ins, int e_ins, int w, int end_bonus,
                    int zdrop, int h0, int* qle, int* tle, int* gtle, int* gscore, int* max_off) {
  if (m != 5 || !query || !mat || (tlen > 0 && !target)) return INT32_MIN;
  bsx_ctx* ctx = dropin_acquire();
  if (!ctx) return INT32_MIN;
  bsx_scoring sc;
  memcpy(sc.mat, mat, 25);
  sc.o_del = o_del; sc.e_del = e_del; sc.o_ins = o_ins; sc.e_ins = e_ins; sc.zdrop = zdrop;
  const uint64_t zero = 0;
  static const uint8_t dummy = 0;
  bsx_ext_tasks t{1, query, &zero, &qlen, tlen > 0 ? target : &dummy, &zero, &tlen, &h0, &w, &end_bonus};
  int32_t o[6];
  bsx_ext_results r{&o[0], &o[1], &o[2], &o[3], &o[4], &o[5], nullptr};
  const int rc = bsx_extend_batch(ctx, &sc, &t, &r);
  dropin_release(ctx);
  if (rc != BSX_OK) return INT32_MIN;
  if (qle) *qle = o[1];
  if (tle) *tle = o[2];
  if (gtle) *gtle = o[3];
  if (gscore) *gscore = o[4];
  if (max_off) *max_off = o[5];
  return o[0];
}

int bsx_ksw_global2(int qlen, const uint8_t* query, int tlen, const uint8_t* target, int m, const int8_t* mat,
                    int o_del, int e_del, int o_ins, int e_ins, int w, int* n_cigar, uint32_t** cigar) {
  if (n_cigar) *n_cigar = 0;  // ksw.c:510
  if (m != 5 || !query || !target || !mat || qlen < 1 || tlen < 1) return INT32_MIN;
  const bool want = n_cigar && cigar;
  bsx_ctx* ctx = dropin_acquire();
  if (!ctx) return INT32_MIN;
  bsx_scoring sc;
  memcpy(sc.mat, mat, 25);
  sc.o_del = o_del; sc.e_del = e_del; sc.o_ins = o_ins; sc.e_ins = e_ins; sc.zdrop = 0;
  const uint64_t zero = 0;
  bsx_glb_tasks t{1, query, &zero, &qlen, target, &zero, &tlen, &w};
  const int64_t slot = (int64_t)qlen + tlen;  // a CIGAR has at most qlen + tlen ops
  std::vector<uint32_t> buf(want ? (size_t)slot : 0);
  int32_t score = 0, nc = 0;
  bsx_glb_results r{&score, &nc, want ? buf.data() : nullptr, want ? slot : 0};
  const int rc = bsx_global_batch(ctx, &sc, &t, &r);
  dropin_release(ctx);
  if (rc != BSX_OK || nc < 0) return INT32_MIN;
  if (want) {  // a fresh malloc'd CIGAR the caller frees, as push_cigar leaves it (ksw.c:492-502,602)
    uint32_t* c = nc > 0 ? (uint32_t*)malloc(4 * (size_t)nc) : nullptr;
    if (nc > 0 && !c) return INT32_MIN;
    if (nc > 0) memcpy(c, buf.data(), 4 * (size_t)nc);
    *n_cigar = nc;
    *cigar = c;
  }
  return score;
}

}  // extern "C"
