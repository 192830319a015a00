// C ABI of the signed triangle counter (include/stc.h): context, model and
// tune-table loading, and the host orchestration of upload -> CSR build ->
// binning -> counting.  Replaces the reference's operator entry
// gunrock::signed_tc::run (Signed_TC/GPU Implementations/GlobalMEM/
// signed_tc.hxx:527-542) and the main()s of the OpenMP programs
// (Signed_TC/Multicore Implementations/tc_merge_sign_enu.cpp:46-154).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/stc.h"
#include "../../include/stc_train.h"
#include "internal.h"

using namespace stc;

struct stc_ctx {
  int device = 0;
  hipStream_t st = nullptr;  // every kernel and the small copies
  hipStream_t up = nullptr;  // staged host -> device copies of edge lists
  int num_cus = 256;
  std::string err;
  int model_status = STC_OK;
  DispatchParams hp{};
  DispatchParams *dp = nullptr;
  BuildBuffers bb;
  CountBuffers cb;
  BuildResult br;
  DevCsr g;
  bool prepared = false;        // a single graph's CSR is on the device
  bool batch_ready = false;     // a batch's union CSR is on the device
  uint32_t batch_k = 0;         // graphs in that batch
  bool kstats_pending = false;  // per-kernel statistics of the last run not read yet
  stc_stats stats{};
  stc_classes classes{};
  bool have_classes = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t marks[kMarks] = {};
  DevBuf batch_counts, batch_vbase;
  UploadRing ring;  // pinned staging for host edge lists (allocated on first use)
  bool fused_pull = false;  // last count was one graph's: with workgroup-pull edges, one
                            // launch counted both pull slots (its time in the last)
  bool batch_launch = false;  // last count was a batch's: one launch counted every
                              // item (its time, edges' bytes in the last slot)
  // Device address of the library's static scratch array on this context's
  // device (hipGetSymbolAddress at stc_create); null with scratch_err set when
  // that call failed.
  void *scratch = nullptr;
  hipError_t scratch_err = hipSuccess;
  // strategy per bucket of the last count (STC_TABLE_* codes), before the
  // per-edge demotions of lists that do not fit the LDS structures
  uint8_t last_table[kBuckets] = {};
  bool have_table = false;
};

namespace {

int fail(stc_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(stc_ctx *c, hipError_t e, const char *where) {
  (void)hipGetLastError();
  std::string m = std::string(where) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(c, STC_E_OOM, m);
  return fail(c, STC_E_HIP, m);
}

// Pinned staging for an upload of `lines` host lines: true when the lines go
// through the staging ring.  A host that refuses the pinned buffers (a locked-
// memory limit, say) gets the direct pageable copies instead of an error.
bool use_staging(stc_ctx *c, uint64_t lines) {
  if (lines < kStageMinLines) return false;
  if (c->ring.ensure() == hipSuccess) return true;
  c->ring.release();
  (void)hipGetLastError();
  return false;
}

// Column words are (rank << 1) | neg with 0xFFFFFFFF as padding, and the
// bitmap probe relies on 2^32 - 2 rank exceeding twice the largest rank window
// (2^19): both hold for ranks below 2^31 - 2^20.
constexpr uint64_t kMaxVertices = (1ull << 31) - (1ull << 20);
// Edge indices plus a load group (256) or a workgroup stride stay below 2^32.
constexpr uint64_t kMaxLines = (1ull << 32) - (1ull << 20);

// Dispatcher graph features of what is counted: k graphs (k = 1 for one
// graph) with n vertices and m oriented edges in total.  Sizes are per
// graph, so a batch of small graphs reads as small graphs, and log2 k tells
// the batched kernel's context apart.
void graph_features(uint32_t n, uint64_t m, uint32_t k, float out[3]) {
  const double kk = k ? (double)k : 1.0;
  out[0] = (float)std::log2((double)n / kk + 1.0);
  out[1] = (float)std::log2((double)m / kk + 1.0);
  out[2] = (float)std::log2(kk);
}

void default_tune(TuneParams &t) {
  for (int i = 0; i < kBucketSide; i++) t.merge_wave_min_log2[i] = 5;
  t.hash_max_log2 = 10;
  t.bitmap_span_log2 = 17;
  t.wave_window_log2 = 14;
  t.hash_block = 256;
  // untuned fixed rule: merge where both lists hold at most 8 entries, pull
  // where the suffix is the shorter list, the reference's 4x rule, hub bitmap
  t.fixed_merge_log2 = 3;
  t.fixed_pull_log2 = 0;
  t.fixed_skew_log2 = 2;
  t.fixed_hub_log2 = 9;
  t.pull_block_chunk_log2 = 11;
  t.pull_wave_chunk_log2 = 8;
  t.pull_wave_max_log2 = 6;
  t.batch_pull_wave_max_log2 = 6;
  t.wave_pop = 8;
  t.lane_item_log2 = 9;
}

std::string lib_dir() {
  Dl_info info;
  if (dladdr((void *)&default_tune, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    size_t s = p.find_last_of('/');
    if (s != std::string::npos) return p.substr(0, s);
  }
  return ".";
}

uint64_t fnv1a(const std::string &s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ull;
  }
  return h;
}

// Tune table: "stc-tune v1" then "key values..." lines.
bool parse_tune(std::istream &f, TuneParams &t) {
  std::string line;
  if (!std::getline(f, line) || line.rfind("stc-tune v1", 0) != 0) return false;
  TuneParams x = t;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string k;
    if (!(is >> k) || k[0] == '#') continue;
    if (k == "merge_wave_min_log2") {
      for (int i = 0; i < kBucketSide; i++) {
        int v = 0;
        if (!(is >> v) || v < 0 || v > 16) return false;
        x.merge_wave_min_log2[i] = (int8_t)v;
      }
    } else if (k == "hash_max_log2") {
      if (!(is >> x.hash_max_log2) || x.hash_max_log2 < 4 || x.hash_max_log2 > 12) return false;
    } else if (k == "bitmap_span_log2") {
      if (!(is >> x.bitmap_span_log2) || x.bitmap_span_log2 < 10 || x.bitmap_span_log2 > 19)
        return false;
    } else if (k == "wave_window_log2") {
      if (!(is >> x.wave_window_log2) || x.wave_window_log2 < 10 || x.wave_window_log2 > 16)
        return false;
    } else if (k == "fixed_merge_log2") {
      if (!(is >> x.fixed_merge_log2) || x.fixed_merge_log2 < -1 || x.fixed_merge_log2 > 15)
        return false;
    } else if (k == "fixed_pull_log2") {
      if (!(is >> x.fixed_pull_log2) || x.fixed_pull_log2 < -15 || x.fixed_pull_log2 > 16)
        return false;
    } else if (k == "fixed_skew_log2") {
      if (!(is >> x.fixed_skew_log2) || x.fixed_skew_log2 < 1 || x.fixed_skew_log2 > 16)
        return false;
    } else if (k == "fixed_hub_log2") {
      if (!(is >> x.fixed_hub_log2) || x.fixed_hub_log2 < 0 || x.fixed_hub_log2 > 16)
        return false;
    } else if (k == "pull_block_chunk_log2") {
      if (!(is >> x.pull_block_chunk_log2) || x.pull_block_chunk_log2 < 8 ||
          x.pull_block_chunk_log2 > 14)
        return false;
    } else if (k == "pull_wave_chunk_log2") {
      if (!(is >> x.pull_wave_chunk_log2) || x.pull_wave_chunk_log2 < 5 ||
          x.pull_wave_chunk_log2 > 12)
        return false;
    } else if (k == "pull_wave_max_log2") {
      if (!(is >> x.pull_wave_max_log2) || x.pull_wave_max_log2 < 6 ||
          x.pull_wave_max_log2 > 9)
        return false;
    } else if (k == "batch_pull_wave_max_log2") {
      if (!(is >> x.batch_pull_wave_max_log2) || x.batch_pull_wave_max_log2 < 6 ||
          x.batch_pull_wave_max_log2 > 9)
        return false;
    } else if (k == "wave_pop") {
      if (!(is >> x.wave_pop) || x.wave_pop < 1 || x.wave_pop > 8) return false;
    } else if (k == "lane_item_log2") {
      if (!(is >> x.lane_item_log2) || x.lane_item_log2 < 9 || x.lane_item_log2 > 12)
        return false;
    } else {
      return false;  // an unknown key: a table for another build
    }
  }
  t = x;
  return true;
}

bool load_tune(const std::string &path, TuneParams &t) {
  std::ifstream f(path);
  if (!f) return false;
  return parse_tune(f, t);
}

// Dispatcher model: "stc-dispatch-mlp v1", dims line, arrays by name, then
// "check <fnv1a-64 of every preceding line>".  Any mismatch rejects the file.
bool load_model(const std::string &path, MlpModel &m, std::string &why) {
  std::ifstream f(path);
  if (!f) {
    why = "cannot open " + path;
    return false;
  }
  std::string all, line;
  std::vector<std::string> lines;
  while (std::getline(f, line)) lines.push_back(line);
  if (lines.size() < 3 || lines[0] != "stc-dispatch-mlp v1") {
    why = "bad header";
    return false;
  }
  std::string body;
  size_t i = 0;
  for (; i < lines.size(); i++) {
    if (lines[i].rfind("check ", 0) == 0) break;
    body += lines[i];
    body += '\n';
  }
  if (i == lines.size()) {
    why = "missing check line";
    return false;
  }
  unsigned long long want = std::strtoull(lines[i].c_str() + 6, nullptr, 16);
  if (fnv1a(body) != want) {
    why = "checksum mismatch";
    return false;
  }
  MlpModel x{};
  auto read_arr = [&](const std::string &name, float *dst, int cnt) -> bool {
    for (auto &l : lines) {
      std::istringstream is(l);
      std::string k;
      is >> k;
      if (k != name) continue;
      for (int j = 0; j < cnt; j++)
        if (!(is >> dst[j]) || !std::isfinite(dst[j])) return false;
      return true;
    }
    return false;
  };
  bool ok = true;
  {
    std::istringstream is(lines[1]);
    std::string k;
    int a = 0, b = 0, c = 0;
    is >> k >> a >> b >> c;
    ok = k == "dims" && a == kMlpIn && b == kMlpHidden && c == kMlpOut;
  }
  ok = ok && read_arr("mean", x.mean, kMlpIn) && read_arr("inv_std", x.inv_std, kMlpIn) &&
       read_arr("w1", &x.w1[0][0], kMlpHidden * kMlpIn) && read_arr("b1", x.b1, kMlpHidden) &&
       read_arr("w2", &x.w2[0][0], kMlpOut * kMlpHidden) && read_arr("b2", x.b2, kMlpOut);
  if (!ok) {
    why = "bad dims or arrays";
    return false;
  }
  m = x;
  return true;
}

std::string fmt_arr(const char *name, const float *v, int n) {
  std::string out = name;
  char buf[32];
  for (int i = 0; i < n; i++) {
    std::snprintf(buf, sizeof buf, " %.9g", (double)v[i]);
    out += buf;
  }
  out += '\n';
  return out;
}

bool write_model(const std::string &path, const MlpModel &m, const std::string &note) {
  std::string body = "stc-dispatch-mlp v1\n";
  body += "dims " + std::to_string(kMlpIn) + " " + std::to_string(kMlpHidden) + " " +
          std::to_string(kMlpOut) + "\n";
  body += fmt_arr("mean", m.mean, kMlpIn);
  body += fmt_arr("inv_std", m.inv_std, kMlpIn);
  body += fmt_arr("w1", &m.w1[0][0], kMlpHidden * kMlpIn);
  body += fmt_arr("b1", m.b1, kMlpHidden);
  body += fmt_arr("w2", &m.w2[0][0], kMlpOut * kMlpHidden);
  body += fmt_arr("b2", m.b2, kMlpOut);
  body += "arch gfx950\n";
  body += "note " + note + "\n";
  char chk[64];
  std::snprintf(chk, sizeof chk, "check %016llx\n", (unsigned long long)fnv1a(body));
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return false;
  f << body << chk;
  return (bool)f;
}

float median(std::vector<float> v) {
  if (v.empty()) return 0.f;
  std::sort(v.begin(), v.end());
  size_t k = v.size();
  return (k & 1) ? v[k / 2] : 0.5f * (v[k / 2 - 1] + v[k / 2]);
}

// Binning time and per-kernel times of the run whose marks were recorded.
hipError_t read_marks(stc_ctx *c, float &ms_bin, float kt[kKernels]) {
  hipError_t e = hipEventElapsedTime(&ms_bin, c->marks[0], c->marks[1]);
  for (int k = 0; k < kKernels && e == hipSuccess; k++)
    e = hipEventElapsedTime(&kt[k], c->marks[k + 1], c->marks[k + 2]);
  return e;
}

// The device table (Strategy values) as STC_TABLE_* codes.
void keep_table(stc_ctx *c, const CountResult &res) {
  for (int b = 0; b < kBuckets; b++) {
    const uint8_t s = res.table[b];
    c->last_table[b] = s == S_MERGE_LANE || s == S_MERGE_WAVE ? STC_TABLE_MERGE
                       : s == S_HASH                          ? STC_TABLE_HASH
                       : s == S_BITMAP                        ? STC_TABLE_BITMAP
                       : s == S_PULL                          ? STC_TABLE_PULL
                                                              : STC_TABLE_SKIP;
  }
  c->have_table = true;
}

int check_graph(stc_ctx *c, const stc_edges *g) {
  if (!g) return fail(c, STC_E_ARG, "null graph");
  if (g->m && (!g->u || !g->v || !g->sign)) return fail(c, STC_E_ARG, "null edge arrays");
  if (g->n > kMaxVertices) return fail(c, STC_E_TOO_LARGE, "n must be <= 2^31 - 2^20");
  if (g->m > kMaxLines) return fail(c, STC_E_TOO_LARGE, "at most 2^32 - 2^20 lines");
  return STC_OK;
}

}  // namespace

extern "C" {

int stc_create(int device, const char *model_dir, stc_ctx **out) {
  if (!out) return STC_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return STC_E_NODEV;
  }
  if (device < 0 || device >= ndev) return STC_E_NODEV;
  auto *c = new stc_ctx();
  c->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->num_cus = prop.multiProcessorCount;
    e = hipStreamCreate(&c->st);
    if (e == hipSuccess) e = hipStreamCreate(&c->up);
  }
  for (int i = 0; i < 4 && e == hipSuccess; i++) e = hipEventCreate(&c->ev[i]);
  for (int i = 0; i < kMarks && e == hipSuccess; i++) e = hipEventCreate(&c->marks[i]);
  if (e == hipSuccess) e = hipMalloc(&c->dp, sizeof(DispatchParams));
  if (e == hipSuccess) e = init_kernel_attributes();
  if (e == hipSuccess) {
    // the static scratch's address on this device (the allocation-free count
    // reports the error if this fails; nothing else needs it)
    c->scratch_err = static_scratch_address(&c->scratch);
    if (c->scratch_err != hipSuccess) {
      c->scratch = nullptr;
      (void)hipGetLastError();
    }
  }
  if (e != hipSuccess) {
    int rc = hip_fail(c, e, "stc_create");
    stc_destroy(c);
    return rc;
  }
  default_tune(c->hp.tune);
  std::string dir = model_dir ? std::string(model_dir) : lib_dir() + "/models";
  load_tune(dir + "/tune.txt", c->hp.tune);
  std::string why;
  if (load_model(dir + "/dispatch_mlp.txt", c->hp.model, why)) {
    c->hp.have_model = 1;
    c->model_status = STC_OK;
  } else {
    c->hp.have_model = 0;
    c->model_status = STC_E_MODEL;
    c->err = "dispatcher model unavailable (" + why + "); fixed rule in use";
  }
  *out = c;
  return STC_OK;
}

void stc_destroy(stc_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->up) (void)hipStreamSynchronize(c->up);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto &ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : c->marks)
    if (ev) (void)hipEventDestroy(ev);
  if (c->dp) (void)hipFree(c->dp);
  c->ring.release();
  if (c->up) (void)hipStreamDestroy(c->up);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char *stc_last_error(const stc_ctx *c) { return c ? c->err.c_str() : "null context"; }
int stc_model_status(const stc_ctx *c) { return c ? c->model_status : STC_E_ARG; }

int stc_prepare(stc_ctx *c, const stc_edges *g, int g_is_device) {
  if (!c) return STC_E_ARG;
  int rc = check_graph(c, g);
  if (rc) return rc;
  (void)hipSetDevice(c->device);
  c->prepared = false;
  c->batch_ready = false;
  BuildInput in;
  in.u = g->u;
  in.v = g->v;
  in.s = g->sign;
  in.m_in = g->m;
  in.n = g->n;
  hipError_t e = hipSuccess;
  std::string why;
  float ms_up = 0.f;
  bool canon_done = false;  // lines canonicalised chunk by chunk during the upload
  if (!g_is_device) {
    const uint64_t mm = g->m ? g->m : 1;
    if ((e = c->bb.in_u.ensure(mm * 4)) != hipSuccess ||
        (e = c->bb.in_v.ensure(mm * 4)) != hipSuccess ||
        (e = c->bb.in_s.ensure(mm)) != hipSuccess)
      return hip_fail(c, e, "upload alloc");
    in.u = c->bb.in_u.as<uint32_t>();
    in.v = c->bb.in_v.as<uint32_t>();
    in.s = c->bb.in_s.as<int8_t>();
  }
  e = build_begin(in, c->bb, why, c->st);
  if (e == hipErrorInvalidValue && !why.empty()) return fail(c, STC_E_ARG, why);
  if (e != hipSuccess) return hip_fail(c, e, "csr build");
  if (!g_is_device) {
    const bool staged = use_staging(c, g->m);
    e = hipStreamSynchronize(c->st);
    const auto t0 = std::chrono::steady_clock::now();
    if (staged) {
      // copies on the copy stream; each chunk is canonicalised (and its first
      // sort digit counted) on the compute stream while the next one copies
      if (e == hipSuccess)
        e = c->ring.upload({EdgePiece{g->u, g->v, g->sign, g->m}}, c->bb.in_u.as<uint32_t>(),
                           c->bb.in_v.as<uint32_t>(), c->bb.in_s.as<int8_t>(), c->up, c->st,
                           [&](uint64_t l0, uint64_t l1) {
                             return build_canon(in, l0, l1, c->bb, c->st);
                           });
      canon_done = true;
    } else if (g->m) {
      if (e == hipSuccess)
        e = hipMemcpyAsync(c->bb.in_u.p, g->u, g->m * 4, hipMemcpyHostToDevice, c->st);
      if (e == hipSuccess)
        e = hipMemcpyAsync(c->bb.in_v.p, g->v, g->m * 4, hipMemcpyHostToDevice, c->st);
      if (e == hipSuccess)
        e = hipMemcpyAsync(c->bb.in_s.p, g->sign, g->m, hipMemcpyHostToDevice, c->st);
      if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    }
    ms_up = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (e != hipSuccess) return hip_fail(c, e, "upload");
  }
  e = hipEventRecord(c->ev[0], c->st);
  if (e == hipSuccess && !canon_done) e = build_canon(in, 0, in.m_in, c->bb, c->st);
  if (e == hipSuccess) e = build_finish(in, c->bb, c->br, why, c->st);
  if (e == hipErrorInvalidValue && !why.empty()) return fail(c, STC_E_ARG, why);
  if (e != hipSuccess && !why.empty()) return fail(c, STC_E_HIP, why);
  if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->st);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
  float ms_build = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms_build, c->ev[0], c->ev[1]);
  if (e != hipSuccess) return hip_fail(c, e, "csr build");
  c->g.n = c->br.n;
  c->g.m = c->br.m;
  c->g.off = c->bb.off.as<uint32_t>();
  c->g.col = c->bb.col.as<uint32_t>();
  c->g.src = c->bb.src.as<uint32_t>();
  c->g.tin_off = c->bb.tin_off.as<uint32_t>();
  c->g.tin = c->bb.tin.as<InEdge>();
  c->g.dcls = c->bb.dcls.as<uint8_t>();
  c->g.ebk = c->bb.ebk.as<uint16_t>();
  c->g.smask = c->bb.smask.as<uint32_t>();
  c->prepared = true;
  c->batch_ready = false;
  std::memset(&c->stats, 0, sizeof(c->stats));
  c->kstats_pending = false;
  c->stats.n = c->br.n;
  c->stats.m_in = c->br.m_in;
  c->stats.m_oriented = c->br.m;
  c->stats.max_out_degree = c->br.max_out;
  c->stats.b_alg = c->br.b_alg;
  c->stats.ms_upload = ms_up;
  c->stats.ms_build = ms_build;
  return STC_OK;
}

}  // extern "C"

// Dispatch parameters to the device for the next count: the host's copy,
// plus whether every class-11 target list of the prepared graph fits the
// current bitmap window (binning then needs no per-edge pass for pull).
static hipError_t upload_params(stc_ctx *c, bool batch) {
  c->hp.pull_c11_fits = c->br.max_span_c11 <= (1u << c->hp.tune.bitmap_span_log2) ? 1u : 0u;
  c->hp.wave_pull_max_log2 =
      (uint32_t)(batch ? c->hp.tune.batch_pull_wave_max_log2 : c->hp.tune.pull_wave_max_log2);
  return hipMemcpyAsync(c->dp, &c->hp, sizeof(DispatchParams), hipMemcpyHostToDevice, c->st);
}

// Binning + counting of the prepared graph.  cbuf: the buffers to count
// with (null: the context's own, grown on first use; else borrowed views of
// the static scratch, for which no per-kernel statistics are kept).
static int count_core(stc_ctx *c, int mode, const uint8_t *table, int repeat, bool classes,
                      stc_counts *out, CountResult *res_out, CountBuffers *cbuf = nullptr) {
  if (!c->prepared) return fail(c, STC_E_ARG, "no prepared graph");
  (void)hipSetDevice(c->device);
  CountBuffers &cb = cbuf ? *cbuf : c->cb;
  c->hp.mode = mode;
  c->hp.tune.allow_pull = 1;
  if (table) std::memcpy(c->hp.host_table, table, kBuckets);
  graph_features(c->g.n, c->g.m, 1, c->hp.graph_feat);
  hipError_t e = upload_params(c, false);
  if (e != hipSuccess) return hip_fail(c, e, "params upload");
  // The event marks between the count's kernels (binning and per-kernel
  // times) are recorded in the first half of the runs only; with repeats,
  // the count's time is the median of the other runs, which have nothing
  // between their kernels.
  std::vector<float> times, marked_times, bins, kt[kKernels];
  const int nmarked = repeat > 1 ? repeat / 2 : 1;
  for (int r = 0; r < repeat; r++) {
    const bool marked = r < nmarked;
    e = hipEventRecord(c->ev[0], c->st);
    if (e == hipSuccess)
      e = enqueue_count(c->g, c->hp, c->dp, cb, classes, c->num_cus, c->st,
                        marked ? c->marks : nullptr);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->st);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
    float ms = 0.f, msb = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    if (marked) {
      if (e == hipSuccess) e = hipEventElapsedTime(&msb, c->marks[0], c->marks[1]);
      for (int k = 0; k < kKernels && e == hipSuccess; k++) {
        float t = 0.f;
        e = hipEventElapsedTime(&t, c->marks[k + 1], c->marks[k + 2]);
        kt[k].push_back(t);
      }
      bins.push_back(msb);
    }
    if (e != hipSuccess) return hip_fail(c, e, "count");
    (marked && repeat > 1 ? marked_times : times).push_back(ms);
  }
  CountResult res;
  e = read_count(cb, classes, res, c->st);
  if (e != hipSuccess) return hip_fail(c, e, "count readback");
  keep_table(c, res);
  if (classes) {
    c->classes = {res.t[0], res.t[1], res.t[2], res.t[3]};
    c->have_classes = true;
    out->balanced = res.t[3] + res.t[1];
    out->unbalanced = res.t[2] + res.t[0];
  } else {
    c->have_classes = false;
    out->balanced = res.t[0];
    out->unbalanced = res.t[1];
  }
  c->fused_pull = true;
  c->batch_launch = false;
  c->stats.ms_count = median(times);
  c->stats.ms_count_min = *std::min_element(times.begin(), times.end());
  c->stats.ms_bin = median(bins);
  for (int k = 0; k < kKernels; k++) c->stats.ms_kernel[k] = median(kt[k]);
  // per-kernel statistics are computed later from the binning buffers, which
  // a static-scratch count does not keep: its per-kernel sizes read zero
  c->kstats_pending = cbuf == nullptr;
  if (cbuf) {
    c->stats.edges_merge = c->stats.edges_hash = c->stats.edges_bitmap = 0;
    c->stats.edges_pull = c->stats.edges_skipped = 0;
    for (int k = 0; k < kKernels; k++) c->stats.bytes_kernel[k] = c->stats.alg_bytes_kernel[k] = 0;
  }
  c->stats.dispatch_used = (mode == STC_LEARNED && !c->hp.have_model) ? STC_FIXED : mode;
  if (res_out) *res_out = res;
  return STC_OK;
}

extern "C" {

int stc_count_prepared(stc_ctx *c, const stc_opts *opts, stc_counts *out) {
  if (!c || !out) return STC_E_ARG;
  int dispatch = opts ? opts->dispatch : STC_LEARNED;
  int repeat = opts && opts->repeat > 0 ? opts->repeat : 1;
  bool classes = opts && opts->classes;
  if (dispatch < STC_LEARNED || dispatch > STC_FORCE_PULL)
    return fail(c, STC_E_ARG, "bad dispatch");
  return count_core(c, dispatch, nullptr, repeat, classes, out, nullptr);
}

int stc_count_prepared_static(stc_ctx *c, const stc_opts *opts, stc_counts *out) {
  if (!c || !out) return STC_E_ARG;
  int dispatch = opts ? opts->dispatch : STC_LEARNED;
  int repeat = opts && opts->repeat > 0 ? opts->repeat : 1;
  bool classes = opts && opts->classes;
  if (dispatch < STC_LEARNED || dispatch > STC_FORCE_PULL)
    return fail(c, STC_E_ARG, "bad dispatch");
  if (!c->prepared) return fail(c, STC_E_ARG, "no prepared graph");
  if (!c->scratch)
    return hip_fail(c, c->scratch_err != hipSuccess ? c->scratch_err : hipErrorInvalidValue,
                    "static scratch address");
  (void)hipSetDevice(c->device);
  CountBuffers sb;
  size_t need = 0;
  if (!carve_count_buffers(c->scratch, kStaticScratchBytes, c->g.n, c->g.m, c->num_cus, sb, &need))
    return fail(c, STC_E_TOO_LARGE,
                "graph needs " + std::to_string(need) + " bytes of counting scratch; the static "
                "scratch holds " + std::to_string(kStaticScratchBytes) + " (use stc_count_prepared)");
  // the scratch keeps the last call's contents: the vertex flags must start at zero
  hipError_t e = clear_vflags(sb, c->g.n, c->st);
  if (e != hipSuccess) return hip_fail(c, e, "static scratch clear");
  return count_core(c, dispatch, nullptr, repeat, classes, out, nullptr, &sb);
}

int stc_count(stc_ctx *c, const stc_edges *g, const stc_opts *opts, stc_counts *out) {
  if (!c || !out) return STC_E_ARG;
  int rc = stc_prepare(c, g, 0);
  if (rc) return rc;
  return stc_count_prepared(c, opts, out);
}

int stc_count_device(stc_ctx *c, const stc_edges *g, const stc_opts *opts, stc_counts *out) {
  if (!c || !out) return STC_E_ARG;
  int rc = stc_prepare(c, g, 1);
  if (rc) return rc;
  return stc_count_prepared(c, opts, out);
}

int stc_last_classes(const stc_ctx *c, stc_classes *out) {
  if (!c || !out) return STC_E_ARG;
  if (!c->have_classes) return STC_E_ARG;
  *out = c->classes;
  return STC_OK;
}

int stc_get_stats(stc_ctx *c, stc_stats *out) {
  if (!c || !out) return STC_E_ARG;
  if (c->kstats_pending) {
    // Per-kernel edge / word counts of the last run, by one pass over its
    // bucket bytes and strategy table (kept out of the timed counting path).
    (void)hipSetDevice(c->device);
    CountResult res;
    hipError_t e = enqueue_stats(c->g, c->dp, c->cb, c->num_cus, c->batch_ready, c->st);
    if (e == hipSuccess) e = read_count(c->cb, false, res, c->st);
    if (e != hipSuccess) return hip_fail(c, e, "count statistics");
    // kernel order: merge-lane, merge-wave, wave-hash, workgroup-hash,
    // wave-bitmap, workgroup-bitmap, wave-pull, workgroup-pull; slot 9 counts
    // the edges that cannot close a triangle
    c->stats.edges_skipped = res.kern_edges[9];
    c->stats.edges_merge = res.kern_edges[0] + res.kern_edges[1];
    c->stats.edges_hash = res.kern_edges[2] + res.kern_edges[3];
    c->stats.edges_bitmap = res.kern_edges[4] + res.kern_edges[5];
    c->stats.edges_pull = res.kern_edges[6] + res.kern_edges[7];
    for (int k = 0; k < kKernels; k++) {
      c->stats.bytes_kernel[k] = 4.0 * (double)res.kern_words[k];
      c->stats.alg_bytes_kernel[k] = (double)res.kern_bytes[k];
    }
    if (c->batch_launch) {  // one launch counted every item of the batch
      for (int k = 0; k < kKernels - 1; k++) {
        c->stats.bytes_kernel[kKernels - 1] += c->stats.bytes_kernel[k];
        c->stats.alg_bytes_kernel[kKernels - 1] += c->stats.alg_bytes_kernel[k];
        c->stats.bytes_kernel[k] = c->stats.alg_bytes_kernel[k] = 0.0;
      }
    } else if (c->fused_pull && res.kern_edges[7] > 0) {  // one launch counted both pull slots
      c->stats.bytes_kernel[7] += c->stats.bytes_kernel[6];
      c->stats.bytes_kernel[6] = 0.0;
      c->stats.alg_bytes_kernel[7] += c->stats.alg_bytes_kernel[6];
      c->stats.alg_bytes_kernel[6] = 0.0;
    }
    c->kstats_pending = false;
  }
  *out = c->stats;
  return STC_OK;
}

float stc_run(stc_ctx *c, const stc_edges *g, uint64_t *total_balanced,
              uint64_t *total_unbalanced) {
  stc_opts o{STC_LEARNED, 1, 0};
  stc_counts r{0, 0};
  int rc = stc_count(c, g, &o, &r);
  if (rc) return (float)rc;
  if (total_balanced) *total_balanced = r.balanced;
  if (total_unbalanced) *total_unbalanced = r.unbalanced;
  return c->stats.ms_count;
}

}  // extern "C"

// Batched counting (reference: each graph would be one run of
// tc_merge_sign_enu.cpp:46-154).  The k graphs become one disjoint union:
// line ids are shifted into [vbase[g], vbase[g+1]) on the device, one CSR is
// built with ranks grouped per graph, and the counting phase -- binning plus
// ONE persistent kernel that adds each item's triangles to its graph's
// counters -- is captured once in a hipGraph and replayed.  The first
// (eager) run allocates every buffer, so the captured sequence allocates
// nothing; each replay must reproduce the eager per-graph counts.
// prepare_batch validates, uploads, shifts and builds the union CSR.
static int prepare_batch(stc_ctx *c, const stc_edges *graphs, uint32_t k) {
  if (!c || !graphs || k == 0) return STC_E_ARG;
  (void)hipSetDevice(c->device);
  std::vector<uint32_t> vbase(k + 1, 0);
  std::vector<uint64_t> ebase(k + 1, 0);
  bool contiguous = true;
  for (uint32_t i = 0; i < k; i++) {
    int rc = check_graph(c, &graphs[i]);
    if (rc) return rc;
    uint64_t vb = (uint64_t)vbase[i] + graphs[i].n;
    if (vb > kMaxVertices) return fail(c, STC_E_TOO_LARGE, "batch has > 2^31 - 2^20 vertices");
    vbase[i + 1] = (uint32_t)vb;
    ebase[i + 1] = ebase[i] + graphs[i].m;
    if (i > 0 && graphs[i].m &&
        (graphs[i].u != graphs[i - 1].u + graphs[i - 1].m ||
         graphs[i].v != graphs[i - 1].v + graphs[i - 1].m ||
         graphs[i].sign != graphs[i - 1].sign + graphs[i - 1].m))
      contiguous = false;
  }
  const uint64_t m = ebase[k];
  const uint32_t n = vbase[k];
  if (m > kMaxLines) return fail(c, STC_E_TOO_LARGE, "batch has > 2^32 - 2^20 lines");
  c->prepared = false;
  c->batch_ready = false;
  hipError_t e = hipSuccess;
  const uint64_t mm = m ? m : 1;
  if ((e = c->bb.in_u.ensure(mm * 4)) != hipSuccess || (e = c->bb.in_v.ensure(mm * 4)) != hipSuccess ||
      (e = c->bb.in_s.ensure(mm)) != hipSuccess ||
      (e = c->batch_vbase.ensure((k + 1) * 4 + (k + 1) * 8 + 64)) != hipSuccess ||
      (e = c->batch_counts.ensure((uint64_t)k * 16 + 64)) != hipSuccess)
    return hip_fail(c, e, "batch alloc");
  uint32_t *du = c->bb.in_u.as<uint32_t>(), *dv = c->bb.in_v.as<uint32_t>();
  int8_t *ds = c->bb.in_s.as<int8_t>();
  uint64_t *d_ebase = c->batch_vbase.as<uint64_t>();
  uint32_t *d_vbase = (uint32_t *)(d_ebase + (k + 1));
  BuildInput in;
  in.u = du;
  in.v = dv;
  in.s = ds;
  in.m_in = m;
  in.n = n;
  in.ebase = d_ebase;
  in.vbase = d_vbase;
  in.k = k;
  std::string why;
  float ms_up = 0.f;
  const bool staged = use_staging(c, m);
  // the graph ranges first: the per-chunk canonicalisation reads them
  e = hipMemcpyAsync(d_ebase, ebase.data(), (k + 1) * 8, hipMemcpyHostToDevice, c->st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_vbase, vbase.data(), (k + 1) * 4, hipMemcpyHostToDevice, c->st);
  if (e == hipSuccess) e = build_begin(in, c->bb, why, c->st);
  if (e == hipErrorInvalidValue && !why.empty()) return fail(c, STC_E_ARG, why);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return hip_fail(c, e, "batch upload");
  const auto t0 = std::chrono::steady_clock::now();
  // find the first graph with lines for the contiguous fast path
  uint32_t first = 0;
  while (first < k && graphs[first].m == 0) first++;
  bool canon_done = false;  // lines canonicalised chunk by chunk during the upload
  if (staged) {
    // the graphs back to back through the pinned staging buffers, each chunk
    // shifted into the union's ids and canonicalised while the next copies
    std::vector<EdgePiece> pieces;
    pieces.reserve(k);
    for (uint32_t i = 0; i < k; i++)
      if (graphs[i].m) pieces.push_back(EdgePiece{graphs[i].u, graphs[i].v, graphs[i].sign, graphs[i].m});
    e = c->ring.upload(pieces, du, dv, ds, c->up, c->st, [&](uint64_t l0, uint64_t l1) {
      return build_canon(in, l0, l1, c->bb, c->st);
    });
    canon_done = true;
  } else if (m && contiguous && first < k) {
    e = hipMemcpyAsync(du, graphs[first].u, m * 4, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(dv, graphs[first].v, m * 4, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, graphs[first].sign, m, hipMemcpyHostToDevice, c->st);
  } else {
    for (uint32_t i = 0; i < k && e == hipSuccess; i++) {
      const uint64_t mi = graphs[i].m, o = ebase[i];
      if (!mi) continue;
      e = hipMemcpyAsync(du + o, graphs[i].u, mi * 4, hipMemcpyHostToDevice, c->st);
      if (e == hipSuccess) e = hipMemcpyAsync(dv + o, graphs[i].v, mi * 4, hipMemcpyHostToDevice, c->st);
      if (e == hipSuccess) e = hipMemcpyAsync(ds + o, graphs[i].sign, mi, hipMemcpyHostToDevice, c->st);
    }
  }
  if (e == hipSuccess && !staged) e = hipStreamSynchronize(c->st);
  ms_up = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (e != hipSuccess) return hip_fail(c, e, "batch upload");
  // the union CSR (ids shifted per graph by the canonicalisation)
  e = hipEventRecord(c->ev[0], c->st);
  if (e == hipSuccess && !canon_done) e = build_canon(in, 0, m, c->bb, c->st);
  if (e == hipSuccess) e = build_finish(in, c->bb, c->br, why, c->st);
  if (e == hipErrorInvalidValue && !why.empty()) return fail(c, STC_E_ARG, why);
  if (e != hipSuccess && !why.empty()) return fail(c, STC_E_HIP, why);
  if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->st);
  if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
  float ms_build = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms_build, c->ev[0], c->ev[1]);
  if (e != hipSuccess) return hip_fail(c, e, "batch csr build");
  c->g.n = c->br.n;
  c->g.m = c->br.m;
  c->g.off = c->bb.off.as<uint32_t>();
  c->g.col = c->bb.col.as<uint32_t>();
  c->g.src = c->bb.src.as<uint32_t>();
  c->g.tin_off = c->bb.tin_off.as<uint32_t>();
  c->g.tin = c->bb.tin.as<InEdge>();
  c->g.dcls = c->bb.dcls.as<uint8_t>();
  c->g.ebk = c->bb.ebk.as<uint16_t>();
  c->g.smask = c->bb.smask.as<uint32_t>();
  c->batch_ready = true;
  c->batch_k = k;
  std::memset(&c->stats, 0, sizeof(c->stats));
  c->kstats_pending = false;
  c->stats.n = c->br.n;
  c->stats.m_in = c->br.m_in;
  c->stats.m_oriented = c->br.m;
  c->stats.max_out_degree = c->br.max_out;
  c->stats.b_alg = c->br.b_alg;
  c->stats.ms_upload = ms_up;
  c->stats.ms_build = ms_build;
  return STC_OK;
}

// Eager batch runs of the prepared union CSR under dispatch `mode` (host
// table for D_TABLE): per-graph counts land in the batch counter buffer;
// ms_count is the median of the runs.  Used by the training hooks.
static int batch_table_core(stc_ctx *c, int mode, const uint8_t *table, int repeat,
                            stc_counts *total, CountResult *res_out) {
  const uint32_t k = c->batch_k;
  c->hp.mode = mode;
  c->hp.tune.allow_pull = 1;
  if (table) std::memcpy(c->hp.host_table, table, kBuckets);
  graph_features(c->g.n, c->g.m, k, c->hp.graph_feat);
  hipError_t e = upload_params(c, true);
  unsigned long long *d_pg = c->batch_counts.as<unsigned long long>();
  const uint32_t *graph_of = c->bb.graph_of.as<uint32_t>();
  std::vector<float> times, bins, kt[kKernels];
  for (int r = 0; r < repeat && e == hipSuccess; r++) {
    e = hipEventRecord(c->ev[0], c->st);
    if (e == hipSuccess)
      e = enqueue_batch(c->g, c->hp, c->dp, c->cb, graph_of, d_pg, k, c->num_cus, c->st,
                        c->marks);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->st);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
    float ms = 0.f, msb = 0.f, t[kKernels] = {};
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    if (e == hipSuccess) e = read_marks(c, msb, t);
    times.push_back(ms);
    bins.push_back(msb);
    for (int j = 0; j < kKernels; j++) kt[j].push_back(t[j]);
  }
  std::vector<unsigned long long> got(2ull * k);
  if (e == hipSuccess) e = hipMemcpyAsync(got.data(), d_pg, 16ull * k, hipMemcpyDeviceToHost, c->st);
  CountResult res;
  if (e == hipSuccess) e = read_count(c->cb, false, res, c->st);
  if (e != hipSuccess) return hip_fail(c, e, "batch count");
  keep_table(c, res);
  total->balanced = total->unbalanced = 0;
  for (uint32_t i = 0; i < k; i++) {
    total->balanced += got[2 * i];
    total->unbalanced += got[2 * i + 1];
  }
  c->stats.ms_count = median(times);
  c->stats.ms_count_min = *std::min_element(times.begin(), times.end());
  c->stats.ms_bin = median(bins);
  for (int j = 0; j < kKernels; j++) c->stats.ms_kernel[j] = median(kt[j]);
  c->fused_pull = false;
  c->batch_launch = true;
  c->kstats_pending = true;
  if (res_out) *res_out = res;
  return STC_OK;
}

extern "C" int stc_count_per_vertex(stc_ctx *c, const stc_opts *opts, uint64_t *cnt,
                                    stc_counts *out) {
  if (!c || !cnt) return STC_E_ARG;
  if (!c->prepared) return fail(c, STC_E_ARG, "no prepared graph");
  (void)hipSetDevice(c->device);
  int dispatch = opts ? opts->dispatch : STC_LEARNED;
  int repeat = opts && opts->repeat > 0 ? opts->repeat : 1;
  if (dispatch < STC_LEARNED || dispatch > STC_FORCE_PULL)
    return fail(c, STC_E_ARG, "bad dispatch");
  const uint32_t n = c->g.n;
  const uint64_t bytes = 32ull * (n ? n : 1);
  hipError_t e = c->cb.pv.ensure(bytes);
  if (e == hipSuccess) e = c->cb.pv_ids.ensure(bytes);
  if (e != hipSuccess) return hip_fail(c, e, "per-vertex alloc");
  stc_counts tot{0, 0};
  c->g.pv = c->cb.pv.as<unsigned long long>();
  int rc = count_core(c, dispatch, nullptr, repeat, true, &tot, nullptr);
  c->g.pv = nullptr;
  if (rc) return rc;
  if (n) {
    e = enqueue_pv_to_ids(c->cb.pv.as<unsigned long long>(), c->bb.perm.as<uint32_t>(), n,
                          c->cb.pv_ids.as<unsigned long long>(), c->num_cus, c->st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(cnt, c->cb.pv_ids.p, 32ull * n, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return hip_fail(c, e, "per-vertex readback");
  }
  if (out) *out = tot;
  return STC_OK;
}

extern "C" int stc_prepare_batch(stc_ctx *c, const stc_edges *graphs, uint32_t k) {
  return prepare_batch(c, graphs, k);
}

// Counting phase of the prepared batch: an eager run (which allocates every
// buffer and gives the counts), then binning plus the persistent kernel
// captured once in a hipGraph and replayed opts.repeat times; every replay
// must reproduce the eager per-graph counts.
extern "C" int stc_count_batch_prepared(stc_ctx *c, const stc_opts *opts, stc_counts *out) {
  if (!c || !out) return STC_E_ARG;
  if (!c->batch_ready) return fail(c, STC_E_ARG, "no prepared batch");
  (void)hipSetDevice(c->device);
  const uint32_t k = c->batch_k;
  int dispatch = opts ? opts->dispatch : STC_LEARNED;
  int repeat = opts && opts->repeat > 0 ? opts->repeat : 1;
  if (dispatch < STC_LEARNED || dispatch > STC_FORCE_PULL)
    return fail(c, STC_E_ARG, "bad dispatch");
  if (opts && opts->classes) return fail(c, STC_E_ARG, "batched counting reports two classes");
  c->hp.mode = dispatch;
  c->hp.tune.allow_pull = 1;
  graph_features(c->g.n, c->g.m, k, c->hp.graph_feat);
  hipError_t e = upload_params(c, true);
  unsigned long long *d_pg = c->batch_counts.as<unsigned long long>();
  const uint32_t *graph_of = c->bb.graph_of.as<uint32_t>();
  // eager run (allocates every buffer; its counts are what each replay must
  // give), then with repeats a second, warm eager run whose event marks give
  // the binning and batched-launch times
  float msb = 0.f, kt[kKernels] = {};
  std::vector<unsigned long long> eager(2ull * k), got(2ull * k);
  for (int r = 0; r < (repeat > 1 ? 2 : 1) && e == hipSuccess; r++) {
    e = enqueue_batch(c->g, c->hp, c->dp, c->cb, graph_of, d_pg, k, c->num_cus, c->st, c->marks);
    if (e == hipSuccess)
      e = hipMemcpyAsync(eager.data(), d_pg, 16ull * k, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e == hipSuccess) e = read_marks(c, msb, kt);
  }
  if (e != hipSuccess) return hip_fail(c, e, "batch count");
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  e = hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed);
  if (e == hipSuccess) {
    hipError_t e2 =
        enqueue_batch(c->g, c->hp, c->dp, c->cb, graph_of, d_pg, k, c->num_cus, c->st, nullptr);
    e = hipStreamEndCapture(c->st, &graph);
    if (e2 != hipSuccess) e = e2;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  std::vector<float> times;
  for (int r = 0; r < repeat && e == hipSuccess; r++) {
    e = hipEventRecord(c->ev[0], c->st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, c->st);
    if (e == hipSuccess) e = hipEventRecord(c->ev[1], c->st);
    if (e == hipSuccess) e = hipEventSynchronize(c->ev[1]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    times.push_back(ms);
    if (e == hipSuccess)
      e = hipMemcpyAsync(got.data(), d_pg, 16ull * k, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e == hipSuccess && got != eager) {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      return fail(c, STC_E_HIP, "graph replay differs from the eager batch count");
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  CountResult res;
  if (e == hipSuccess) e = read_count(c->cb, false, res, c->st);
  if (e != hipSuccess) return hip_fail(c, e, "batch graph");
  keep_table(c, res);
  for (uint32_t i = 0; i < k; i++) {
    out[i].balanced = got[2 * i];
    out[i].unbalanced = got[2 * i + 1];
  }
  c->stats.ms_count = median(times);
  c->stats.ms_count_min = *std::min_element(times.begin(), times.end());
  c->stats.ms_bin = msb;
  for (int j = 0; j < kKernels; j++) c->stats.ms_kernel[j] = kt[j];
  c->fused_pull = false;
  c->batch_launch = true;
  c->kstats_pending = true;
  c->stats.dispatch_used = (dispatch == STC_LEARNED && !c->hp.have_model) ? STC_FIXED : dispatch;
  c->have_classes = false;
  return STC_OK;
}

extern "C" int stc_count_batch(stc_ctx *c, const stc_edges *graphs, uint32_t k,
                               const stc_opts *opts, stc_counts *out) {
  if (!c || (!graphs && k) || (!out && k)) return STC_E_ARG;
  if (k == 0) return STC_OK;
  int dispatch = opts ? opts->dispatch : STC_LEARNED;
  if (dispatch < STC_LEARNED || dispatch > STC_FORCE_PULL)
    return fail(c, STC_E_ARG, "bad dispatch");
  if (opts && opts->classes) return fail(c, STC_E_ARG, "batched counting reports two classes");
  int rc = prepare_batch(c, graphs, k);
  if (rc) return rc;
  return stc_count_batch_prepared(c, opts, out);
}

// ---- training / tuning hooks (include/stc_train.h) -------------------------
// Both hooks act on whatever is prepared: one graph (stc_prepare) or a batch
// (stc_prepare_batch, counted by the batched kernel, counts summed).
extern "C" int stc_bucket_histogram(stc_ctx *c, uint32_t out[STC_BUCKETS]) {
  if (!c || !out) return STC_E_ARG;
  uint8_t skip[kBuckets];
  std::memset(skip, 0, sizeof skip);
  stc_counts cnt{0, 0};
  CountResult res;
  int rc = c->batch_ready ? batch_table_core(c, D_TABLE, skip, 1, &cnt, &res)
                          : count_core(c, D_TABLE, skip, 1, false, &cnt, &res);
  if (rc) return rc;
  for (int k = 0; k < kBuckets; k++) out[k] = res.bucket_cnt[k];
  return STC_OK;
}

extern "C" int stc_count_with_table(stc_ctx *c, const uint8_t table[STC_BUCKETS], int repeat,
                                    stc_counts *out) {
  if (!c || !table || !out) return STC_E_ARG;
  for (int k = 0; k < kBuckets; k++)
    if (table[k] > STC_TABLE_PULL) return fail(c, STC_E_ARG, "bad table entry");
  if (c->batch_ready) return batch_table_core(c, D_TABLE, table, repeat > 0 ? repeat : 1, out, nullptr);
  return count_core(c, D_TABLE, table, repeat > 0 ? repeat : 1, false, out, nullptr);
}

extern "C" void stc_bucket_features(int bucket, uint32_t count, uint32_t n, uint64_t m,
                                    uint32_t k, float out[STC_MLP_FEATURES]) {
  float la = (float)(bucket / kBucketSide), lb = (float)(bucket % kBucketSide);
  out[0] = la;
  out[1] = lb;
  out[2] = la - lb;
  out[3] = std::log2((float)count + 1.0f);
  graph_features(n, m, k, out + 4);
}

extern "C" int stc_train_dispatcher(stc_ctx *c, const float *feat, const float *cost,
                                    int nsamp, int iters, float lr, uint64_t seed,
                                    const char *out_path) {
  if (!c || !feat || !cost || nsamp <= 0 || iters <= 0 || !out_path) return STC_E_ARG;
  (void)hipSetDevice(c->device);
  MlpModel m{};
  // feature normalisation
  for (int k = 0; k < kMlpIn; k++) {
    double mu = 0, var = 0;
    for (int i = 0; i < nsamp; i++) mu += feat[i * kMlpIn + k];
    mu /= nsamp;
    for (int i = 0; i < nsamp; i++) {
      double d = feat[i * kMlpIn + k] - mu;
      var += d * d;
    }
    double sd = std::sqrt(var / nsamp);
    m.mean[k] = (float)mu;
    m.inv_std[k] = (float)(1.0 / (sd > 1e-3 ? sd : 1e-3));
  }
  std::vector<float> x((size_t)nsamp * kMlpIn), r((size_t)nsamp * kMlpOut), w(nsamp);
  double wsum = 0;
  for (int i = 0; i < nsamp; i++) {
    for (int k = 0; k < kMlpIn; k++)
      x[(size_t)i * kMlpIn + k] = (feat[i * kMlpIn + k] - m.mean[k]) * m.inv_std[k];
    float lo = cost[i * kMlpOut], hi = lo;
    for (int s = 1; s < kMlpOut; s++) {
      lo = std::min(lo, cost[i * kMlpOut + s]);
      hi = std::max(hi, cost[i * kMlpOut + s]);
    }
    for (int s = 0; s < kMlpOut; s++)
      r[(size_t)i * kMlpOut + s] = (cost[i * kMlpOut + s] - lo) / (hi - lo + 1e-9f);
    // the relative time at stake: each cost is a whole count of the sample's
    // graph, so lo is about that graph's best time; clamping at 1 (a 2x
    // slowdown) keeps a few extreme buckets from drowning the small graphs
    w[i] = std::min(1.f, (hi - lo) / (lo > 0.f ? lo : 1.f));
    wsum += w[i];
  }
  for (int i = 0; i < nsamp; i++) w[i] = (float)(w[i] / (wsum > 0 ? wsum : 1.0));
  // deterministic initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in))
  const int np = mlp_param_count();
  std::vector<float> params(np, 0.f);
  uint64_t st = seed ^ 0x6d6c70ull;
  auto uni = [&](float a) {
    st += 0x9e3779b97f4a7c15ULL;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return a * (2.0f * (float)((double)(z >> 11) * 0x1.0p-53) - 1.0f);
  };
  const int nw1 = kMlpHidden * kMlpIn, nw2 = kMlpOut * kMlpHidden;
  for (int i = 0; i < nw1; i++) params[i] = uni(1.0f / std::sqrt((float)kMlpIn));
  for (int i = 0; i < nw2; i++) params[nw1 + kMlpHidden + i] = uni(1.0f / std::sqrt((float)kMlpHidden));
  hipError_t e = train_mlp_device(x.data(), r.data(), w.data(), nsamp, iters, lr, params.data(), c->st);
  if (e != hipSuccess) return hip_fail(c, e, "train");
  mlp_unpack(params.data(), m);
  char note[160];
  std::snprintf(note, sizeof note, "samples %d iters %d lr %g seed %llu", nsamp, iters, (double)lr,
                (unsigned long long)seed);
  if (!write_model(out_path, m, note)) return fail(c, STC_E_IO, "cannot write model");
  c->hp.model = m;
  c->hp.have_model = 1;
  c->model_status = STC_OK;
  return STC_OK;
}

extern "C" int stc_get_csr(stc_ctx *c, uint32_t *off, uint32_t *col, uint32_t *perm,
                           uint32_t *tin_off, uint32_t *tin, uint8_t *dcls, uint32_t *smask) {
  if (!c) return STC_E_ARG;
  if (!c->prepared && !c->batch_ready) return fail(c, STC_E_ARG, "nothing prepared");
  (void)hipSetDevice(c->device);
  const uint64_t n = c->g.n, m = c->g.m;
  struct Part {
    void *dst;
    const void *src;
    uint64_t bytes;
  } parts[] = {{off, c->g.off, (n + 1) * 4},        {col, c->g.col, m * 4},
               {perm, c->bb.perm.p, n * 4},         {tin_off, c->g.tin_off, (n + 1) * 4},
               {tin, c->g.tin, m * 8},              {dcls, c->g.dcls, n},
               {smask, c->g.smask, n * 4}};
  hipError_t e = hipSuccess;
  for (const Part &pt : parts)
    if (pt.dst && pt.bytes && e == hipSuccess)
      e = hipMemcpyAsync(pt.dst, pt.src, pt.bytes, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) return hip_fail(c, e, "csr readback");
  return STC_OK;
}

extern "C" int stc_last_table(const stc_ctx *c, uint8_t out[STC_BUCKETS]) {
  if (!c || !out) return STC_E_ARG;
  if (!c->have_table) return STC_E_ARG;
  std::memcpy(out, c->last_table, kBuckets);
  return STC_OK;
}

extern "C" int stc_load_models(stc_ctx *c, const char *dir) {
  if (!c) return STC_E_ARG;
  std::string d = dir ? std::string(dir) : lib_dir() + "/models";
  // a directory without a (valid) tune.txt keeps the context's tune values
  TuneParams t;
  default_tune(t);
  if (load_tune(d + "/tune.txt", t)) c->hp.tune = t;
  std::string why;
  if (load_model(d + "/dispatch_mlp.txt", c->hp.model, why)) {
    c->hp.have_model = 1;
    c->model_status = STC_OK;
    return STC_OK;
  }
  c->hp.have_model = 0;
  c->model_status = STC_E_MODEL;
  c->err = "dispatcher model unavailable (" + why + "); fixed rule in use";
  return STC_E_MODEL;
}

extern "C" int stc_set_tune(stc_ctx *c, const char *text) {
  if (!c || !text) return STC_E_ARG;
  std::istringstream is(text);
  TuneParams t = c->hp.tune;
  if (!parse_tune(is, t)) return fail(c, STC_E_ARG, "bad tune text");
  c->hp.tune = t;
  return STC_OK;
}
