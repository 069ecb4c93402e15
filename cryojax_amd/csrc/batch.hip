// Grouped persistent image formation: up to CRYOMI_MAX_GROUPS (volume, box,
// pose table) groups in one launch, and its HIP-graph capture (gfx950).
//
// Workgroups of NTH threads take images from eight queues, one per XCD
// (atomic ticket counters in the workspace, cleared by a one-wave kernel
// enqueued just before, so the launches can be captured and replayed):
// workgroup b serves queue b mod 8, the XCD it is dispatched to. Each group's
// images are put in the plane-ordered schedule of order.hpp first (one
// order_groups_kernel launch), and queue x serves the schedule positions
// p = x (mod 8) of every group, so each XCD's L2 sees slices whose planes are
// a few degrees apart, as in the single-box launch. Tickets run over the
// groups in descending box order, so the largest images are formed first and
// the small ones fill the tail. Each image is formed by the register-FFT
// plan of its box (regfft.hpp) with NTH threads; the twiddle table in LDS is
// refilled when a workgroup moves to another box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/cryomi_batch.h"
#include "common.hpp"
#include "order.hpp"
#include "regfft.hpp"

namespace cryomi {

constexpr int kQueues = 8;  // one ticket queue per XCD

struct GroupDesc {
  const float2* grid;
  const cryomi_image_params* params;
  float* out;
  int box;
  unsigned flags;
  int first;     // the group's first entry in the order table (images of earlier groups)
  int n_images;
};

struct GroupTable {
  GroupDesc g[CRYOMI_MAX_GROUPS];
  int n_groups;
  int total;
  // queue_first[x][q]: tickets of queue x before group q (queue x serves schedule positions
  // p = x (mod kQueues) of each group); queue_first[x][n_groups] is queue x's ticket count
  int queue_first[kQueues][CRYOMI_MAX_GROUPS + 1];
  int chunk_first[CRYOMI_MAX_GROUPS + 1];  // order_groups_kernel blocks before group q
};

constexpr int kMaxBoxTw = 256;  // twiddle table entries (largest supported box)
constexpr size_t kQueueBytes = 256;
static_assert(kQueues * sizeof(unsigned) <= kQueueBytes, "queue header");

template <int NTH>
constexpr int groups_lds_elems() {
  constexpr int a = RegPlan<64, NTH>::LDS, b = RegPlan<128, NTH>::LDS;
  constexpr int c = RegPlan<192, NTH>::LDS, d = RegPlan<256, NTH>::LDS;
  constexpr int ab = a > b ? a : b, cd = c > d ? c : d;
  return kMaxBoxTw + (ab > cd ? ab : cd);
}

__global__ void queue_reset_kernel(unsigned* counter) {
  if (threadIdx.x < kQueues) counter[threadIdx.x] = 0u;
}

// The plane-ordered schedule of every group: block b orders chunk
// b - chunk_first[q] of group q into order + first_q.
__global__ __launch_bounds__(kOrderThreads) void order_groups_kernel(GroupTable t, int* __restrict__ order) {
  int q = 0;  // the last group whose first block is <= blockIdx.x
  while (q + 1 < t.n_groups && (int)blockIdx.x >= t.chunk_first[q + 1]) ++q;
  const GroupDesc& g = t.g[q];
  order_chunk(g.params, g.n_images, ((int)blockIdx.x - t.chunk_first[q]) * kOrderChunk, order + g.first);
}

// The next image of a workgroup's queue and its group, broadcast through LDS
// (group = n_groups once the queue is empty). Tickets a workgroup draws only
// increase, so with the groups in descending box order its boxes never
// increase either: the kernel runs one loop per box, one after the other, and
// each plan's registers are live only inside its own loop.
struct Ticket {
  int group;
  int img;
};

// `qf` is the workgroup's row of GroupTable::queue_first and `groups` the group
// table, both in LDS.
__device__ __forceinline__ Ticket next_ticket(unsigned* counter, const GroupDesc* groups, const int* qf, int n_groups,
                                              const int* __restrict__ order, int queue, int* s_group, int* s_img,
                                              int tid) {
  if (tid == 0) {
    const int ticket = (int)atomicAdd(counter + queue, 1u);
    int q = n_groups, img = 0;
    if (ticket < qf[n_groups]) {
      q = 0;  // the last group whose first ticket in this queue is <= ticket
      while (q + 1 < n_groups && ticket >= qf[q + 1]) ++q;
      const int p = queue + kQueues * (ticket - qf[q]);  // schedule position in group q
      img = order[groups[q].first + p];
    }
    *s_group = q;
    *s_img = img;
  }
  __syncthreads();
  Ticket tk{__builtin_amdgcn_readfirstlane(*s_group), __builtin_amdgcn_readfirstlane(*s_img)};
  __syncthreads();  // thread 0 may draw again right away (the next image can be empty of barriers)
  return tk;
}

template <int N, int NTH, bool BRICKED>
__device__ __forceinline__ void box_loop(Ticket& tk, int n_groups, const GroupDesc* groups, const int* qf,
                                         unsigned* counter, const int* order, int queue, int* s_group, int* s_img,
                                         float2* tw, float2* lds, float2* slot, int tid) {
  if (tk.group >= n_groups || groups[tk.group].box != N) return;
  fill_twiddles(tw, N, tid, NTH);
  __syncthreads();
#pragma unroll 1
  while (tk.group < n_groups) {
    const GroupDesc g = groups[tk.group];
    if (g.box != N) break;
    const ImageConsts k = make_consts(g.params[tk.img], N, true, g.flags);
    reg_image<N, NTH, false, BRICKED>(g.grid, nullptr, k, g.flags, tw, lds, slot,
                                      reinterpret_cast<float2*>(g.out + (size_t)tk.img * N * N), tid);
    tk = next_ticket(counter, groups, qf, n_groups, order, queue, s_group, s_img, tid);
  }
}

template <int NTH, bool BRICKED>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(NTH == 256 ? 3 : 2, 8))) void groups_kernel(
    GroupTable t, unsigned* __restrict__ counter, const int* __restrict__ order, float2* __restrict__ scratch,
    size_t slot_elems) {
  extern __shared__ float2 smem[];
  // the group table lives in LDS (in SGPRs it would crowd the plans' registers)
  __shared__ GroupDesc s_groups[CRYOMI_MAX_GROUPS];
  __shared__ int s_qf[CRYOMI_MAX_GROUPS + 1];
  __shared__ int s_group, s_img;
  float2* tw = smem;
  float2* lds = smem + kMaxBoxTw;
  float2* slot = scratch + (size_t)blockIdx.x * slot_elems;
  const int tid = threadIdx.x;
  const int queue = (int)(blockIdx.x % kQueues);  // the XCD this workgroup is dispatched to
  const int n_groups = t.n_groups;
  if (tid < n_groups) s_groups[tid] = t.g[tid];
  if (tid <= n_groups) s_qf[tid] = t.queue_first[queue][tid];
  __syncthreads();
  Ticket tk = next_ticket(counter, s_groups, s_qf, n_groups, order, queue, &s_group, &s_img, tid);
  box_loop<256, NTH, BRICKED>(tk, n_groups, s_groups, s_qf, counter, order, queue, &s_group, &s_img, tw, lds, slot,
                              tid);
  box_loop<192, NTH, BRICKED>(tk, n_groups, s_groups, s_qf, counter, order, queue, &s_group, &s_img, tw, lds, slot,
                              tid);
  box_loop<128, NTH, BRICKED>(tk, n_groups, s_groups, s_qf, counter, order, queue, &s_group, &s_img, tw, lds, slot,
                              tid);
  box_loop<64, NTH, BRICKED>(tk, n_groups, s_groups, s_qf, counter, order, queue, &s_group, &s_img, tw, lds, slot,
                             tid);
}

template <int NTH, bool BRICKED>
static const void* groups_kernel_ptr() {
  return reinterpret_cast<const void*>(&groups_kernel<NTH, BRICKED>);
}

static const void* pick_groups_kernel(int nth, bool bricked) {
  if (nth == 512) return bricked ? groups_kernel_ptr<512, true>() : groups_kernel_ptr<512, false>();
  return bricked ? groups_kernel_ptr<256, true>() : groups_kernel_ptr<256, false>();
}

static size_t groups_lds_bytes(int nth) {
  return (size_t)(nth == 512 ? groups_lds_elems<512>() : groups_lds_elems<256>()) * sizeof(float2);
}

// Validated launch description of one grouped call.
struct GroupsPlan {
  GroupTable table;
  int nth = 256;
  bool bricked = false;
  int blocks = 0;
  size_t slot_elems = 0;
  size_t lds = 0;
  const void* kernel = nullptr;
};

static bool group_box_ok(int box) { return box == 64 || box == 128 || box == 192 || box == 256; }

// Checks the descriptions and sizes the launch (no HIP calls beyond device
// queries, so it may run before a capture begins).
static int plan_groups(const cryomi_volume_group* groups, int n_groups, GroupsPlan& p) {
  CRYOMI_REQUIRE(groups != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_groups: NULL groups");
  CRYOMI_REQUIRE(n_groups >= 1 && n_groups <= CRYOMI_MAX_GROUPS, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_groups: n_groups must be in [1, CRYOMI_MAX_GROUPS]");
  const unsigned known = CRYOMI_APPLY_CTF | CRYOMI_APPLY_TRANSLATION | CRYOMI_GRID_BRICKED | CRYOMI_GATHER_FILL;
  const bool bricked = (groups[0].flags & CRYOMI_GRID_BRICKED) != 0;
  int order[CRYOMI_MAX_GROUPS];
  int max_box = 0;
  long long total = 0;
  for (int i = 0; i < n_groups; ++i) {
    const cryomi_volume_group& g = groups[i];
    CRYOMI_REQUIRE(group_box_ok(g.box), CRYOMI_ERR_UNSUPPORTED,
                   "cryomi_simulate_groups: group box must be 64, 128, 192 or 256 (other boxes: cryomi_simulate)");
    CRYOMI_REQUIRE(g.n_images >= 0, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_groups: n_images < 0");
    CRYOMI_REQUIRE((g.flags & ~known) == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_simulate_groups: unknown or unsupported flag bits (CRYOMI_OUTPUT_FOURIER is not grouped)");
    CRYOMI_REQUIRE(((g.flags & CRYOMI_GRID_BRICKED) != 0) == bricked, CRYOMI_ERR_UNSUPPORTED,
                   "cryomi_simulate_groups: every group must use the same grid layout");
    CRYOMI_REQUIRE(g.n_images == 0 || (g.fourier_voxel_grid && g.params && g.out), CRYOMI_ERR_INVALID_ARGUMENT,
                   "cryomi_simulate_groups: NULL grid, params or output pointer");
    order[i] = i;
    total += g.n_images;
    if (g.n_images > 0) max_box = std::max(max_box, g.box);
  }
  CRYOMI_REQUIRE(total < (1LL << 31), CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_groups: too many images");
  // largest box first; stable, so equal boxes keep the caller's order
  std::stable_sort(order, order + n_groups, [&](int a, int b) { return groups[a].box > groups[b].box; });
  GroupTable& t = p.table;
  t = GroupTable{};
  int first = 0;
  for (int q = 0; q < n_groups; ++q) {
    const cryomi_volume_group& g = groups[order[q]];
    t.g[q] = GroupDesc{static_cast<const float2*>(g.fourier_voxel_grid), g.params, static_cast<float*>(g.out), g.box,
                       g.flags, first, g.n_images};
    first += g.n_images;
    for (int x = 0; x < kQueues; ++x) {
      const int c = g.n_images > x ? (g.n_images - x + kQueues - 1) / kQueues : 0;  // positions p = x (mod 8)
      t.queue_first[x][q + 1] = t.queue_first[x][q] + c;
    }
    t.chunk_first[q + 1] = t.chunk_first[q] + (g.n_images + kOrderChunk - 1) / kOrderChunk;
  }
  t.n_groups = n_groups;
  t.total = (int)total;
  p.bricked = bricked;
  p.nth = tuned_groups_threads();
  p.kernel = pick_groups_kernel(p.nth, bricked);
  p.lds = groups_lds_bytes(p.nth);
  p.slot_elems = max_box > 0 ? (size_t)max_box * (max_box / 2 + 1) : 0;
  int per_cu = tuned_groups_wg_per_cu();
  if (per_cu <= 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, p.kernel, p.nth, p.lds) != hipSuccess || nb <= 0) nb = 1;
    per_cu = nb;
  }
  const long long g = (long long)device_cu_count() * per_cu;
  p.blocks = (int)std::min<long long>(g, std::max<long long>(total, 1));
  return CRYOMI_OK;
}

// the order table after the queue header, rounded up to 256 bytes
static size_t order_table_bytes(const GroupsPlan& p) {
  return ((size_t)p.table.total * sizeof(int) + 255) / 256 * 256;
}

static size_t plan_workspace_bytes(const GroupsPlan& p) {
  return kQueueBytes + order_table_bytes(p) + (size_t)p.blocks * p.slot_elems * sizeof(float2);
}

static int set_lds_limit(const GroupsPlan& p) {
  if (p.lds > (size_t)65536) {
    const hipError_t e = hipFuncSetAttribute(p.kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
    if (e != hipSuccess) {
      set_error(std::string("groups_kernel: LDS request too large: ") + hipGetErrorString(e));
      return CRYOMI_ERR_UNSUPPORTED;
    }
  }
  return CRYOMI_OK;
}

static int enqueue_groups(const GroupsPlan& p, void* workspace, hipStream_t stream) {
  if (p.table.total == 0) return CRYOMI_OK;
  unsigned* counter = static_cast<unsigned*>(workspace);
  int* order = reinterpret_cast<int*>(static_cast<char*>(workspace) + kQueueBytes);
  float2* scratch = reinterpret_cast<float2*>(static_cast<char*>(workspace) + kQueueBytes + order_table_bytes(p));
  hipLaunchKernelGGL(queue_reset_kernel, dim3(1), dim3(64), 0, stream, counter);
  int st = check_launch("queue_reset_kernel");
  if (st) return st;
  hipLaunchKernelGGL(order_groups_kernel, dim3((unsigned)p.table.chunk_first[p.table.n_groups]), dim3(kOrderThreads),
                     0, stream, p.table, order);
  if ((st = check_launch("order_groups_kernel"))) return st;
  const dim3 grid((unsigned)p.blocks), block((unsigned)p.nth);
  if (p.nth == 512) {
    if (p.bricked)
      hipLaunchKernelGGL((groups_kernel<512, true>), grid, block, p.lds, stream, p.table, counter, order, scratch,
                         p.slot_elems);
    else
      hipLaunchKernelGGL((groups_kernel<512, false>), grid, block, p.lds, stream, p.table, counter, order, scratch,
                         p.slot_elems);
  } else {
    if (p.bricked)
      hipLaunchKernelGGL((groups_kernel<256, true>), grid, block, p.lds, stream, p.table, counter, order, scratch,
                         p.slot_elems);
    else
      hipLaunchKernelGGL((groups_kernel<256, false>), grid, block, p.lds, stream, p.table, counter, order, scratch,
                         p.slot_elems);
  }
  return check_launch("groups_kernel");
}

}  // namespace cryomi

using namespace cryomi;

struct cryomi_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

size_t cryomi_simulate_groups_workspace_bytes(const cryomi_volume_group* groups, int n_groups) {
  GroupsPlan p;
  if (plan_groups(groups, n_groups, p) != CRYOMI_OK) return 0;
  return plan_workspace_bytes(p);
}

int cryomi_simulate_groups(const cryomi_volume_group* groups, int n_groups, void* workspace,
                           size_t workspace_bytes, ihipStream_t* stream) {
  clear_error();
  GroupsPlan p;
  int st = plan_groups(groups, n_groups, p);
  if (st) return st;
  if (p.table.total == 0) return CRYOMI_OK;
  CRYOMI_REQUIRE(workspace && workspace_bytes >= plan_workspace_bytes(p), CRYOMI_ERR_WORKSPACE,
                 "cryomi_simulate_groups: workspace smaller than cryomi_simulate_groups_workspace_bytes()");
  CRYOMI_REQUIRE(reinterpret_cast<uintptr_t>(workspace) % 256 == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_groups: workspace must be 256-byte aligned");
  if ((st = set_lds_limit(p))) return st;
  return enqueue_groups(p, workspace, stream);
}

int cryomi_simulate_groups_graph_create(const cryomi_volume_group* groups, int n_groups, void* workspace,
                                        size_t workspace_bytes, cryomi_graph** graph) {
  clear_error();
  CRYOMI_REQUIRE(graph != nullptr, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_simulate_groups_graph_create: NULL graph");
  *graph = nullptr;
  GroupsPlan p;
  int st = plan_groups(groups, n_groups, p);
  if (st) return st;
  CRYOMI_REQUIRE(p.table.total == 0 || (workspace && workspace_bytes >= plan_workspace_bytes(p)),
                 CRYOMI_ERR_WORKSPACE,
                 "cryomi_simulate_groups_graph_create: workspace smaller than cryomi_simulate_groups_workspace_bytes()");
  CRYOMI_REQUIRE(reinterpret_cast<uintptr_t>(workspace) % 256 == 0, CRYOMI_ERR_INVALID_ARGUMENT,
                 "cryomi_simulate_groups_graph_create: workspace must be 256-byte aligned");
  if ((st = set_lds_limit(p))) return st;
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("cryomi_simulate_groups_graph_create: ") + hipGetErrorString(e));
    return CRYOMI_ERR_LAUNCH;
  }
  cryomi_graph* g = new cryomi_graph;
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    st = enqueue_groups(p, workspace, s);
    const hipError_t e2 = hipStreamEndCapture(s, &g->graph);
    if (!st && e2 != hipSuccess) e = e2;
  }
  if (!st && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(s);
  if (st || e != hipSuccess) {
    if (!st) set_error(std::string("cryomi_simulate_groups_graph_create: ") + hipGetErrorString(e));
    cryomi_graph_destroy(g);
    return st ? st : CRYOMI_ERR_LAUNCH;
  }
  *graph = g;
  return CRYOMI_OK;
}

int cryomi_graph_launch(cryomi_graph* graph, ihipStream_t* stream) {
  clear_error();
  CRYOMI_REQUIRE(graph && graph->exec, CRYOMI_ERR_INVALID_ARGUMENT, "cryomi_graph_launch: NULL graph");
  const hipError_t e = hipGraphLaunch(graph->exec, stream);
  if (e != hipSuccess) {
    set_error(std::string("cryomi_graph_launch: ") + hipGetErrorString(e));
    return CRYOMI_ERR_LAUNCH;
  }
  return CRYOMI_OK;
}

void cryomi_graph_destroy(cryomi_graph* graph) {
  if (!graph) return;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
}

}  // extern "C"
