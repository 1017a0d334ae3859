// batch.hip -- batched dense assembly: many independent P0 x P0 assemblies (each with
// its own grid, wavenumber and output buffer) in one persistent launch.
//
// Work list: the 64 x 64 pair tiles of every item ({item, tile row, tile column}),
// upper-triangle tiles only for symmetric operators. A persistent kernel of
// `blocks` workgroups takes tiles from a device counter, so items of very uneven
// size (320 to 20480 cells) share the GPU without per-item launches. The adjacent
// (Duffy) pairs of all items follow in a second launch over the concatenated pair
// lists, and, in mixed precision, the FP64 near-pair lists in a third. The launch
// sequence (counter reset by a kernel, no memset node) can be captured once into a
// hipGraph and replayed.
//
// The launch shape (workgroups per CU or one per tile, trial chunk CH, tile order) is
// chosen by batch_autotune: timed candidates, best kept in a JSON-lines cache keyed by
// the device and the batch signature.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "core.hpp"
#include "kernels.hpp"
#include "kernels_mixed.hpp"

namespace bempp_amd {

#define HIP_CHECK_B(expr)                                                                         \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess) throw Error(std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

struct Batch {
  int device = -1;
  int cus = 0;
  std::string arch;
  Kind kind;
  Op op;
  DType dtype;
  bool mixed = false;
  bool symmetric = false;
  int Q = 0;
  int n = 0;
  std::vector<const Grid*> grids;
  std::vector<uint64_t> versions;
  std::vector<int> ncells;
  // device arrays
  dev::RegularArgs* items = nullptr;
  int* centred = nullptr;
  dev::BatchTile* tiles[2] = {nullptr, nullptr};  // tile orders: item-major, largest items first
  int ntiles = 0;
  int ngroup[3] = {0, 0, 0};  // tiles of the items of phase tier 0, 1, 2 (in this order)
  int* counter = nullptr;
  dev::SingularArgs* sitems = nullptr;
  int* small = nullptr;
  int* sfirst = nullptr;
  int npairs = 0;
  int2* near = nullptr;
  int* nfirst = nullptr;
  int nnear = 0;
  // launch shape
  int blocks_per_cu = 2;  // 0: one workgroup per tile
  int chunk = 8;
  int order = 0;
  int waves = 4;          // wavefronts per 64 x 64 tile (2, 4 or 8)
  // graph
  bool use_graph = false;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_shape[4] = {-1, -1, -1, -1};
  // profiling
  bool profiling = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool recorded = false;
};

namespace {

template <typename T>
T* upload_b(const std::vector<T>& h) {
  T* d = nullptr;
  HIP_CHECK_B(hipMalloc((void**)&d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) HIP_CHECK_B(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

void free_arrays(Batch& b) {
  void* ptrs[] = {b.items, b.centred, b.tiles[0], b.tiles[1], b.counter, b.sitems, b.small, b.sfirst, b.near, b.nfirst};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}

void drop_graph(Batch& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
  b.graph_shape[0] = b.graph_shape[1] = b.graph_shape[2] = b.graph_shape[3] = -1;
}

// ---- launches ----
template <int Q, int OP, bool HELM, int CH, int NW>
void launch_tiles_t(const Batch& b, int blocks, hipStream_t s) {
  constexpr int kThreads = 64 * NW;
  // group g = phase tier of the items: 0 not centred (range-reduced phase), 1 and 2
  // centred small-argument tiers (Helmholtz only)
  const int sym = b.symmetric ? 1 : 0;
  int first = 0;
  for (int g = 0; g < 3; ++g) {
    const int nt = b.ngroup[g];
    const dev::BatchTile* tg = b.tiles[b.order] + first;
    first += nt;
    if (nt == 0) continue;
    const int grid = blocks > 0 ? std::min(nt, blocks) : nt;
    int* ctr = b.counter + g;
    if (b.mixed) {
      if constexpr (Q == 0) {
        throw Error("mixed-precision batches need a rule compiled for its point count");
      } else if (!HELM) {
        dev::batched_p0_mixed_persistent<Q, OP, HELM, 1, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      } else if (g == 2) {
        dev::batched_p0_mixed_persistent<Q, OP, HELM, 2, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      } else if (g == 1) {
        dev::batched_p0_mixed_persistent<Q, OP, HELM, 1, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      } else {
        dev::batched_p0_mixed_persistent<Q, OP, HELM, 0, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      }
    } else {
      if (!HELM)
        dev::batched_p0_persistent<Q, OP, HELM, 0, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      else if (g == 2)
        dev::batched_p0_persistent<Q, OP, HELM, 2, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      else if (g == 1)
        dev::batched_p0_persistent<Q, OP, HELM, 1, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
      else
        dev::batched_p0_persistent<Q, OP, HELM, 0, CH, NW><<<grid, kThreads, 0, s>>>(b.items, tg, nt, ctr, sym);
    }
  }
}

// tile shapes: (trial chunk, wavefronts per tile) = (4, 4), (8, 4), (8, 2), (8, 8)
template <int Q, int OP, bool HELM>
void launch_tiles_c(const Batch& b, int blocks, hipStream_t s) {
  if (b.waves == 2) launch_tiles_t<Q, OP, HELM, 8, 2>(b, blocks, s);
  else if (b.waves == 8) launch_tiles_t<Q, OP, HELM, 8, 8>(b, blocks, s);
  else if (b.chunk == 4) launch_tiles_t<Q, OP, HELM, 4, 4>(b, blocks, s);
  else launch_tiles_t<Q, OP, HELM, 8, 4>(b, blocks, s);
}

template <int OP, bool HELM>
void launch_tiles_q(const Batch& b, int blocks, hipStream_t s) {
  switch (b.Q) {
    case 1: launch_tiles_c<1, OP, HELM>(b, blocks, s); break;
    case 3: launch_tiles_c<3, OP, HELM>(b, blocks, s); break;
    case 6: launch_tiles_c<6, OP, HELM>(b, blocks, s); break;
    case 7: launch_tiles_c<7, OP, HELM>(b, blocks, s); break;
    default: launch_tiles_c<0, OP, HELM>(b, blocks, s); break;  // point count at run time (FP64 only)
  }
}

template <int OP, bool HELM>
void launch_near_q(const Batch& b, hipStream_t s) {
  const dim3 g((unsigned)((b.nnear + 255) / 256));
  switch (b.Q) {
    case 1: dev::batched_near_kernel<1, OP, HELM><<<g, 256, 0, s>>>(b.items, b.centred, b.near, b.nfirst, b.n); break;
    case 3: dev::batched_near_kernel<3, OP, HELM><<<g, 256, 0, s>>>(b.items, b.centred, b.near, b.nfirst, b.n); break;
    case 6: dev::batched_near_kernel<6, OP, HELM><<<g, 256, 0, s>>>(b.items, b.centred, b.near, b.nfirst, b.n); break;
    case 7: dev::batched_near_kernel<7, OP, HELM><<<g, 256, 0, s>>>(b.items, b.centred, b.near, b.nfirst, b.n); break;
    default: throw Error("unsupported regular quadrature point count");
  }
}

template <int OP, bool HELM>
void launch_all_t(const Batch& b, hipStream_t s, bool events) {
  const int blocks = b.blocks_per_cu * b.cus;  // 0: one workgroup per tile
  if (events) HIP_CHECK_B(hipEventRecord(b.ev[0], s));
  if (b.ntiles > 0) {
    dev::reset_counter_kernel<int><<<1, 64, 0, s>>>(b.counter, 3);
    launch_tiles_q<OP, HELM>(b, blocks, s);
  }
  if (b.mixed && b.nnear > 0) launch_near_q<OP, HELM>(b, s);
  if (events) {
    HIP_CHECK_B(hipEventRecord(b.ev[1], s));
    HIP_CHECK_B(hipEventRecord(b.ev[2], s));
  }
  if (b.npairs > 0) {
    const dim3 g((unsigned)((b.npairs + 3) / 4));
    dev::batched_singular_kernel<OP, HELM><<<g, 256, 0, s>>>(b.sitems, b.small, b.sfirst, b.n);
  }
  if (events) HIP_CHECK_B(hipEventRecord(b.ev[3], s));
  HIP_CHECK_B(hipGetLastError());
}

template <bool HELM>
void launch_all_h(const Batch& b, hipStream_t s, bool events) {
  switch (b.op) {
    case kSL: launch_all_t<kSL, HELM>(b, s, events); break;
    case kDL: launch_all_t<kDL, HELM>(b, s, events); break;
    case kADL: launch_all_t<kADL, HELM>(b, s, events); break;
    case kHS: launch_all_t<kHS, HELM>(b, s, events); break;
    default: throw Error("operator not supported by the batched assembler");
  }
}

void launch_all(const Batch& b, hipStream_t s, bool events) {
  if (b.kind == kHelmholtz) launch_all_h<true>(b, s, events);
  else launch_all_h<false>(b, s, events);
}

void check_versions(const Batch& b) {
  for (int i = 0; i < b.n; ++i)
    if (b.grids[i]->version != b.versions[i])
      throw Error("a grid of the batch changed after the batch was built; build a new batch");
}

hipStream_t own_stream(Batch& b) {
  if (!b.stream) HIP_CHECK_B(hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking));
  return b.stream;
}

}  // namespace

Batch* batch_new(const std::vector<const Assembler*>& asms, const std::vector<const FunctionSpace*>& spaces,
                 const std::vector<void*>& outputs) {
  const size_t n = asms.size();
  if (n == 0) throw Error("a batch needs at least one item");
  if (spaces.size() != n || outputs.size() != n) throw Error("assemblers, spaces and outputs differ in length");
  if (n > (1u << 20)) throw Error("too many batch items");
  for (size_t i = 0; i < n; ++i) {
    if (!asms[i] || !spaces[i]) throw Error("null assembler or space in the batch");
    if (!outputs[i]) throw Error("null output pointer in the batch");
  }
  const Assembler& a0 = *asms[0];
  std::unique_ptr<Batch> b(new Batch());
  HIP_CHECK_B(hipGetDevice(&b->device));
  hipDeviceProp_t prop;
  HIP_CHECK_B(hipGetDeviceProperties(&prop, b->device));
  b->cus = prop.multiProcessorCount;
  b->arch = prop.gcnArchName;
  b->kind = a0.kind;
  b->op = a0.op;
  b->dtype = a0.dtype;
  b->mixed = a0.options.precision == kMixed;
  b->symmetric = a0.op == kSL || a0.op == kHS;
  b->n = (int)n;
  std::vector<dev::RegularArgs> R(n);
  std::vector<int> centred(n), small(n), sfirst(n + 1, 0), nfirst(n + 1, 0);
  std::vector<dev::SingularArgs> S(n);
  std::vector<int2> near_all;
  std::vector<std::pair<long long, int>> size_order;
  for (size_t i = 0; i < n; ++i) {
    const Assembler& a = *asms[i];
    const FunctionSpace& s = *spaces[i];
    if (a.kind != a0.kind || a.op != a0.op || a.dtype != a0.dtype || a.options.precision != a0.options.precision)
      throw Error("batch items must share kernel, operator, dtype and precision");
    if (s.nb != 1) throw Error("the batched assembler supports P0 (DP0) spaces only");
    Prepared P = prepare(a, s, s, outputs[i], nullptr);
    if (P.device != b->device) throw Error("batch items on different devices");
    if (i == 0) b->Q = P.Q;
    b->mixed = b->mixed && P.mixed;  // mixed precision runs the rules compiled for their point count
    if (P.Q != b->Q) throw Error("batch items must share the regular quadrature degree");
    R[i] = P.R;
    R[i].mode = dev::kStore;
    centred[i] = P.tier;  // 0: not centred; 1, 2: phase tiers (nonzero = centred)
    small[i] = P.small_adjacent ? 1 : 0;
    S[i] = P.S;
    S[i].mode = dev::kAccumulate;
    sfirst[i + 1] = sfirst[i] + P.npairs;
    if (b->mixed && P.nnear > 0) {
      std::vector<int2> h(P.nnear);
      HIP_CHECK_B(hipMemcpy(h.data(), P.near, h.size() * sizeof(int2), hipMemcpyDeviceToHost));
      near_all.insert(near_all.end(), h.begin(), h.end());
    }
    nfirst[i + 1] = (int)near_all.size();
    b->grids.push_back(s.grid);
    b->versions.push_back(s.grid->version);
    b->ncells.push_back(R[i].n_test);
    size_order.push_back({-(long long)R[i].n_test, (int)i});
    if ((long long)sfirst[i + 1] > (1LL << 30)) throw Error("too many adjacent pairs in the batch");
  }
  // tile lists in two orders, each grouped by phase mode (non-centred items first)
  std::stable_sort(size_order.begin(), size_order.end());
  std::vector<dev::BatchTile> t0, t1;
  auto add_tiles = [&](std::vector<dev::BatchTile>& t, int item) {
    const int nt = (b->ncells[item] + dev::kTile - 1) / dev::kTile;
    for (int bb = 0; bb < nt; ++bb)
      for (int aa = 0; aa < (b->symmetric ? bb + 1 : nt); ++aa) t.push_back({item, aa, bb, 0});
  };
  for (int g = 0; g < 3; ++g) {
    const size_t before = t0.size();
    for (int i = 0; i < b->n; ++i)
      if (centred[i] == g) add_tiles(t0, i);
    for (auto& so : size_order)
      if (centred[so.second] == g) add_tiles(t1, so.second);
    b->ngroup[g] = (int)(t0.size() - before);
  }
  if (t0.size() > (size_t)INT32_MAX - 65536) throw Error("too many tiles in the batch");
  b->ntiles = (int)t0.size();
  b->npairs = sfirst[n];
  b->nnear = (int)near_all.size();
  try {
    b->items = upload_b(R);
    b->centred = upload_b(centred);
    b->tiles[0] = upload_b(t0);
    b->tiles[1] = upload_b(t1);
    b->counter = upload_b(std::vector<int>(3, 0));
    b->sitems = upload_b(S);
    b->small = upload_b(small);
    b->sfirst = upload_b(sfirst);
    b->near = upload_b(near_all);
    b->nfirst = upload_b(nfirst);
  } catch (...) {
    free_arrays(*b);
    throw;
  }
  return b.release();
}

void batch_free(Batch* b) {
  if (!b) return;
  int cur = 0;
  const bool have = hipGetDevice(&cur) == hipSuccess;
  if (have) (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(*b);
  free_arrays(*b);
  for (auto e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (have) (void)hipSetDevice(cur);
  delete b;
}

void batch_assemble(Batch& b, void* stream) {
  check_versions(b);
  int cur = 0;
  HIP_CHECK_B(hipGetDevice(&cur));
  if (cur != b.device) throw Error("the batch was built on another device");
  hipStream_t s = (hipStream_t)stream;  // NULL: the null stream
  if (b.profiling)
    for (auto& e : b.ev)
      if (!e) HIP_CHECK_B(hipEventCreate(&e));
  if (!b.use_graph) {
    launch_all(b, s, b.profiling);
    b.recorded = b.profiling;
    return;
  }
  const int shape[4] = {b.blocks_per_cu, b.chunk, b.order, b.waves};
  if (!b.exec || !std::equal(shape, shape + 4, b.graph_shape)) {
    drop_graph(b);
    // capture on the batch's own stream (kernel nodes only), replay on any stream
    hipStream_t cs = own_stream(b);
    HIP_CHECK_B(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      launch_all(b, cs, false);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIP_CHECK_B(hipStreamEndCapture(cs, &b.graph));
    HIP_CHECK_B(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    std::copy(shape, shape + 4, b.graph_shape);
  }
  if (b.profiling) HIP_CHECK_B(hipEventRecord(b.ev[0], s));
  HIP_CHECK_B(hipGraphLaunch(b.exec, s));
  if (b.profiling) HIP_CHECK_B(hipEventRecord(b.ev[3], s));
  b.recorded = false;  // per-part times are not available for a graph replay
}

void batch_set_graph(Batch& b, bool enable) {
  b.use_graph = enable;
  if (!enable) drop_graph(b);
}

void batch_set_profiling(Batch& b, bool enable) { b.profiling = enable; }

void batch_set_launch(Batch& b, int blocks_per_cu, int chunk, int order) {
  if (blocks_per_cu < 0 || blocks_per_cu > 64) throw Error("blocks per CU must be in [0, 64] (0: one per tile)");
  if (chunk != 4 && chunk != 8) throw Error("trial chunk must be 4 or 8");
  if (chunk == 4 && b.waves != 4) throw Error("trial chunk 4 needs 4 wavefronts per tile");
  if (order != 0 && order != 1) throw Error("tile order must be 0 (item-major) or 1 (largest items first)");
  b.blocks_per_cu = blocks_per_cu;
  b.chunk = chunk;
  b.order = order;
}

void batch_set_tile_waves(Batch& b, int waves) {
  if (waves != 2 && waves != 4 && waves != 8) throw Error("wavefronts per tile must be 2, 4 or 8");
  if (waves != 4 && b.chunk != 8) throw Error("2 or 8 wavefronts per tile need trial chunk 8");
  b.waves = waves;
}

int batch_get_tile_waves(const Batch& b) { return b.waves; }

void batch_get_launch(const Batch& b, int* blocks_per_cu, int* chunk, int* order) {
  *blocks_per_cu = b.blocks_per_cu;
  *chunk = b.chunk;
  *order = b.order;
}

void batch_counts(const Batch& b, long long* out) {
  out[0] = b.n;
  out[1] = b.ntiles;
  out[2] = b.npairs;
  out[3] = b.nnear;
}

// ms[0]: tiles (+ near pairs), ms[1]: adjacent pairs of the last eager launch; in graph
// mode ms[0] is the whole replay and ms[1] is -1.
void batch_kernel_times(const Batch& b, float* ms) {
  ms[0] = ms[1] = -1.0f;
  if (!b.profiling || !b.ev[3]) return;
  HIP_CHECK_B(hipEventSynchronize(b.ev[3]));
  if (b.recorded) {
    HIP_CHECK_B(hipEventElapsedTime(&ms[0], b.ev[0], b.ev[1]));
    HIP_CHECK_B(hipEventElapsedTime(&ms[1], b.ev[2], b.ev[3]));
  } else {
    HIP_CHECK_B(hipEventElapsedTime(&ms[0], b.ev[0], b.ev[3]));
  }
}

namespace {

std::string signature(const Batch& b) {
  long long cells = 0;
  for (int c : b.ncells) cells += c;
  // sizes enter as powers of two so that similar batches share a tuned shape
  auto lg = [](long long v) { return v > 0 ? (int)std::floor(std::log2((double)v)) : -1; };
  std::ostringstream s;
  s << b.arch << "/cu" << b.cus << "/k" << (int)b.kind << "/op" << (int)b.op << "/dt" << (int)b.dtype << "/mx"
    << (b.mixed ? 1 : 0) << "/q" << b.Q << "/n" << lg(b.n) << "/c" << lg(cells) << "/t" << lg(b.ntiles);
  return s.str();
}

bool cache_lookup(const char* path, const std::string& key, int* bpc, int* ch, int* ord, int* nw, double* ms) {
  if (!path || !*path) return false;
  std::ifstream f(path);
  std::string line;
  bool found = false;
  const std::string needle = "\"key\": \"" + key + "\"";
  while (std::getline(f, line)) {
    if (line.find(needle) == std::string::npos) continue;
    int a = 0, c = 0, o = 0;
    double m = 0.0;
    const char* p = strstr(line.c_str(), "\"blocks_per_cu\":");
    const char* q = strstr(line.c_str(), "\"chunk\":");
    const char* r = strstr(line.c_str(), "\"order\":");
    const char* t = strstr(line.c_str(), "\"ms\":");
    if (p && q && r && t && sscanf(p, "\"blocks_per_cu\": %d", &a) == 1 && sscanf(q, "\"chunk\": %d", &c) == 1 &&
        sscanf(r, "\"order\": %d", &o) == 1 && sscanf(t, "\"ms\": %lf", &m) == 1) {
      const char* w = strstr(line.c_str(), "\"waves\":");
      int wv = 4;  // records written before the tile split was tuned
      if (w && sscanf(w, "\"waves\": %d", &wv) != 1) wv = 4;
      *bpc = a;
      *ch = c;
      *ord = o;
      *nw = wv;
      *ms = m;
      found = true;  // the last record of a key wins
    }
  }
  return found;
}

void cache_store(const char* path, const std::string& key, int bpc, int ch, int ord, int nw, double ms) {
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  if (!f) throw Error(std::string("cannot write the autotune cache ") + path);
  f << "{\"key\": \"" << key << "\", \"blocks_per_cu\": " << bpc << ", \"chunk\": " << ch << ", \"order\": " << ord
    << ", \"waves\": " << nw << ", \"ms\": " << ms << "}\n";
}

double time_shape(Batch& b, int reps) {
  hipStream_t s = own_stream(b);
  hipEvent_t e0, e1;
  HIP_CHECK_B(hipEventCreate(&e0));
  HIP_CHECK_B(hipEventCreate(&e1));
  float best = 1e30f;
  try {
    launch_all(b, s, false);  // warm-up (first-launch code loading)
    for (int r = 0; r < reps; ++r) {
      HIP_CHECK_B(hipEventRecord(e0, s));
      launch_all(b, s, false);
      HIP_CHECK_B(hipEventRecord(e1, s));
      HIP_CHECK_B(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIP_CHECK_B(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  HIP_CHECK_B(hipEventDestroy(e0));
  HIP_CHECK_B(hipEventDestroy(e1));
  return best;
}

}  // namespace

// Coordinate search over the launch shape: trial chunk, then workgroups per CU
// (including one workgroup per tile), then tile order, then the tile split (2, 4 or 8
// wavefronts per 64 x 64 tile; the workgroup size); each candidate is the best of `reps`
// timed launches. The batch's outputs are overwritten. Returns the best time.
double batch_autotune(Batch& b, const char* cache_path, int reps, int* from_cache) {
  check_versions(b);
  if (reps < 1) reps = 1;
  const std::string key = signature(b);
  int bpc, ch, ord, nw;
  double ms;
  *from_cache = 0;
  auto set_shape = [&](int p, int c, int o, int w) {
    b.waves = 4;  // a valid intermediate state for any chunk
    batch_set_launch(b, p, c, o);
    batch_set_tile_waves(b, w);
  };
  if (cache_lookup(cache_path, key, &bpc, &ch, &ord, &nw, &ms)) {
    set_shape(bpc, ch, ord, nw);
    *from_cache = 1;
    return ms;
  }
  int best[4] = {2, 8, 0, 4};
  set_shape(best[0], best[1], best[2], best[3]);
  double best_ms = time_shape(b, reps);
  auto consider = [&](int p, int c, int o, int w) {
    set_shape(p, c, o, w);
    const double t = time_shape(b, reps);
    if (t < best_ms) {
      best_ms = t;
      best[0] = p;
      best[1] = c;
      best[2] = o;
      best[3] = w;
    }
  };
  consider(best[0], 4, best[2], 4);
  const int bpcs[] = {1, 3, 4, 6, 8, 0};
  const int c_best = best[1];
  for (int p : bpcs) consider(p, c_best, best[2], 4);
  consider(best[0], best[1], 1 - best[2], 4);
  for (int w : {2, 8}) consider(best[0], 8, best[2], w);
  set_shape(best[0], best[1], best[2], best[3]);
  HIP_CHECK_B(hipStreamSynchronize(own_stream(b)));
  cache_store(cache_path, key, best[0], best[1], best[2], best[3], best_ms);
  return best_ms;
}

}  // namespace bempp_amd
