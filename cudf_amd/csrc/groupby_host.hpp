// Host-side pieces shared by the plain groupby driver (groupby.hip) and the
// batched / served executions (batch.hip): error reporting, type helpers,
// the aggregation plan (record layout, ops, outputs), device buffers,
// profiling, LDS geometry, and the kernels both launch.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "groupby_device.hpp"
#include "mgb_groupby.h"

namespace mgb {

constexpr int kLdsBlock = 1024;     // threads of an LDS aggregation workgroup
constexpr int kScanItems = 16;      // u32 scan: items per thread (x256 threads)
constexpr int kScanTile = 256 * kScanItems;
constexpr int kMaxRecordWords = 64;
constexpr int64_t kMaxRows = 0x7fffffff;  // cudf::size_type: rows of one call

// Writes `tmpl` (header + identities) into every record; record `capacity`
// gets `special_hdr` as header. Also zeroes up to two ranges of 8-byte
// words (the call's counters), saving their separate clears.
struct FillParams {
  uint64_t* table;
  int64_t num_records;
  int32_t words;
  int32_t pad;
  uint64_t special_hdr;
  uint64_t tmpl[kMaxRecordWords];
  unsigned long long* zero[2];
  int64_t zero_words[2];
};

// ------------------------------------------------------------ profiling

struct ProfRecord {
  std::string kernel;
  hipEvent_t start, stop;
};
extern std::mutex g_prof_mu;
extern bool g_prof_enabled;
extern std::vector<ProfRecord> g_prof;

// Brackets one kernel launch with events when profiling is on.
struct ProfScope {
  hipEvent_t start = nullptr, stop = nullptr;
  hipStream_t stream;
  const char* name;
  ProfScope(const char* kernel, hipStream_t s) : stream{s}, name{kernel}
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_prof_enabled) return;
    if (hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
      start = stop = nullptr;
      return;
    }
    (void)hipEventRecord(start, stream);
  }
  ~ProfScope()
  {
    if (!start) return;
    (void)hipEventRecord(stop, stream);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(ProfRecord{name, start, stop});
  }
};

// ------------------------------------------------------------ call memory

inline size_t align_up(size_t x, size_t a);

// Device memory of one call. Every temporary comes from the caller's
// allocator for temporaries (the role of cudf's current device resource,
// cpp/src/groupby/hash/groupby.cu:404,478,484) and goes back to it,
// stream-ordered on the call's stream, when the call ends; or, for the
// entry point that allocates nothing, it is carved from a fixed region
// (bump allocation; the block on top can grow in place).
struct CallMem {
  const mgb_allocator* alloc = nullptr;
  uint8_t* region = nullptr;
  size_t region_bytes = 0, top = 0;
  hipStream_t stream = nullptr;
  std::vector<std::pair<void*, size_t>> live;
  size_t held = 0, peak = 0;  // bytes of the live temporaries, and their peak

  void* get(size_t bytes)
  {
    bytes = align_up(std::max<size_t>(bytes, 1), 256);
    void* p = nullptr;
    if (region) {
      if (top + bytes > region_bytes) return nullptr;
      p = region + top;
      top += bytes;
    } else {
      p = alloc->allocate(alloc->ctx, bytes, (void*)stream);
      if (!p) return nullptr;
    }
    live.push_back({p, bytes});
    held += bytes;
    peak = std::max(peak, held);
    return p;
  }
  void put(void* p)
  {
    for (size_t i = 0; i < live.size(); ++i) {
      if (live[i].first != p) continue;
      const size_t bytes = live[i].second;
      if (region) {
        if (static_cast<uint8_t*>(p) + bytes == region + top) top -= bytes;  // the top block: reused at once
      } else {
        alloc->deallocate(alloc->ctx, p, bytes, (void*)stream);
      }
      held -= bytes;
      live.erase(live.begin() + (std::ptrdiff_t)i);
      return;
    }
  }
  void release_all()
  {
    while (!live.empty()) put(live.back().first);
  }
  ~CallMem() { release_all(); }
};

// A temporary that grows on demand (its contents are not kept when it grows).
struct DeviceBuffer {
  CallMem* mem = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
  hipError_t reserve(size_t need, hipStream_t)
  {
    if (need <= bytes) return hipSuccess;
    if (ptr) mem->put(ptr);
    ptr = mem->get(need);
    bytes = ptr ? align_up(need, 256) : 0;
    return ptr ? hipSuccess : hipErrorOutOfMemory;
  }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
constexpr size_t align_up_c(size_t x, size_t a) { return (x + a - 1) / a * a; }

inline int64_t bitmask_bytes(int64_t n) { return (int64_t)align_up((size_t)((n + 31) / 32) * 4, 64); }

// ------------------------------------------------------------ plan

struct ValKey {
  const void* data;
  const uint32_t* mask;
  int64_t offset, size;
  int32_t type;
  bool operator<(const ValKey& o) const
  {
    return std::tie(data, mask, offset, size, type) < std::tie(o.data, o.mask, o.offset, o.size, o.type);
  }
};

struct PlanCol {
  ValCol col;               // op_begin/op_end filled when batching
  int type;
  bool nullable;
  std::vector<std::pair<int, int>> ops;  // (op kind, byte offset)
  std::vector<std::pair<int, int>> p2;   // second-pass fields (OP_M2, OP_ARG*_IDX), byte offset
  int count_off = -1;       // the row-count field
  int nulls_off = -1;       // OP_NULLS field when the column has nulls
  int op_offset(int kind) const
  {
    for (auto& o : ops)
      if (o.first == kind) return o.second;
    return -1;
  }
  int p2_offset(int kind) const
  {
    for (auto& o : p2)
      if (o.first == kind) return o.second;
    return -1;
  }
};

struct Plan {
  std::vector<KeyCol> keys;
  int key_mode = KEY_INT;
  bool skip_null_keys = false;
  bool compare_key_nulls = false;
  bool key_has_nulls = false;
  std::vector<PlanCol> cols;
  int rows_off = -1;
  int stride = 8;
  bool has_pass2 = false;   // VARIANCE / STD / ARGMIN / ARGMAX: a second pass over the rows
  bool has_sort = false;    // MEDIAN / QUANTILE / NUNIQUE / NTH_ELEMENT: sorts of the rows by group
  struct Out {
    int col;        // index into cols
    int kind;       // aggregation kind
    int out_type;
    OutSpec spec;
    mgb_aggregation agg{};  // the request's parameters (sort-path aggregations)
  };
  std::vector<Out> outs;
};

struct Allocs {
  const mgb_allocator* alloc;
  hipStream_t stream;
  std::vector<std::pair<void*, size_t>> live;
  void* get(size_t bytes)
  {
    void* p = alloc->allocate(alloc->ctx, bytes, (void*)stream);
    if (p) live.push_back({p, bytes});
    return p;
  }
  // n buffers (sizes > 0) in one request when the allocator takes one; false
  // when any allocation failed (those made are in `live`)
  bool get_many(int n, const size_t* bytes, void** out)
  {
    if (n == 0) return true;
    if (alloc->allocate_many) {
      if (alloc->allocate_many(alloc->ctx, n, bytes, out, (void*)stream) != 0) return false;
      for (int i = 0; i < n; ++i) live.push_back({out[i], bytes[i]});
      return true;
    }
    for (int i = 0; i < n; ++i)
      if (!(out[i] = get(bytes[i]))) return false;
    return true;
  }
  void release_all()
  {
    for (auto& l : live) alloc->deallocate(alloc->ctx, l.first, l.second, (void*)stream);
    live.clear();
  }
};

struct LdsGeom {
  int slots = 0, cap = 0;
  size_t bytes = 0;
  bool ok = false;
};

// ------------------------------------------------------------ errors

extern thread_local std::string g_last_error;

// Records `msg` as the calling thread's last error and returns `s`.
mgb_status fail(mgb_status s, std::string msg);

#define MGB_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(MGB_ERROR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------ types

int type_width(int t);
bool is_float_t(int t);
bool is_unsigned_t(int t);
bool is_signed_int_t(int t);
bool is_integral_t(int t);
bool is_timestamp_t(int t);
bool is_duration_t(int t);
int value_class(int t);
int32_t target_type(int32_t t, int32_t kind);

// ------------------------------------------------------------ plan and launch geometry

int device_cu_count();
mgb_status build_plan(const mgb_column* keys, int32_t num_keys, int32_t null_policy, const mgb_request* reqs,
                      int32_t nreq, Plan& plan);
void record_template(const Plan& plan, uint64_t header, uint64_t* words);
bool single_launch_ops(const Plan& plan, AggParams& ap);
size_t lds_limit();
// Raises a kernel's dynamic LDS limit to `bytes` (once per kernel, device and size).
hipError_t set_dynamic_lds(const void* fn, size_t bytes);
int lds_stride_of(int stride);
LdsGeom lds_geometry(int stride);

// ------------------------------------------------------------ compact partitioned execution (lean.hip)

// Whether the compact kernels are compiled for key column widths (k0, and k1
// for a second key column, else 0) and value widths (absent columns 0, last).
bool lean_supported(int k0, int k1, const int (&w)[kLeanCols]);
// The compact scatter over `grid` chunks, and the compact aggregation over
// `grid` work items (lp: the LDS parameters in device memory).
mgb_status lean_scatter(int k0, int k1, const int (&w)[kLeanCols], const LeanScatterArgs& a, unsigned grid,
                        size_t lds, hipStream_t stream);
mgb_status lean_aggregate(int k0, int k1, const int (&w)[kLeanCols], const LeanAggArgs& a, const LdsParams* lp,
                          unsigned grid, size_t lds, hipStream_t stream);

// ------------------------------------------------------------ kernels shared with batch.hip

__global__ __launch_bounds__(kBlock) void fill_table_kernel(const FillParams p);
__global__ __launch_bounds__(kBlock) void fill_u64_kernel(unsigned long long* p, int64_t n, unsigned long long v);
__global__ __launch_bounds__(kBlock) void pack_validity_kernel(const uint8_t* valid, int64_t n, uint32_t* mask,
                                                               int64_t words_alloc, unsigned long long* null_count);
__global__ __launch_bounds__(256) void scan_reduce_u32_kernel(const uint32_t* in, int64_t n, uint32_t* bsum);
__global__ __launch_bounds__(1024) void scan_bsums_u32_kernel(uint32_t* bsum, int64_t n, unsigned long long* total);
__global__ __launch_bounds__(256) void scan_apply_u32_kernel(const uint32_t* in, int64_t n, const uint32_t* bsum,
                                                            uint32_t* out);

}  // namespace mgb
