// A bounded, stream-ordered device memory pool (mgb_pool_* in mgb_groupby.h).
//
// The pool takes chunks from the device (hipMalloc) as it needs them, never
// holds more than its limit, and keeps its chunks until it is destroyed or
// trimmed. Blocks are carved first-fit in 256-byte units. A block returned on
// a stream stays tied to that stream: allocations on the same stream may take
// it at once (stream order puts them after the work that used it), other
// streams only after the pool has waited for that stream, which it does only
// when it would otherwise exceed its limit. It serves the temporaries of the
// groupby calls of a caller that has no memory resource of its own (the
// Python binding); the role of rmm's pool resource in front of cudf.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "groupby_host.hpp"
#include "mgb_groupby.h"

using mgb::fail;

namespace {

struct Span {
  size_t len;
  hipStream_t owner;  // the stream that returned it; nullptr with `free_any`
  bool free_any;      // usable from any stream
};

struct Chunk {
  uint8_t* base = nullptr;
  size_t size = 0;
  std::map<size_t, Span> free;  // offset -> span
};

}  // namespace

struct mgb_pool {
  std::mutex mu;
  int64_t limit = 0;
  size_t held = 0, in_use = 0, peak = 0;
  std::vector<Chunk> chunks;
  std::map<void*, std::pair<size_t, size_t>> live;  // ptr -> (chunk index, bytes)
  int64_t waits = 0;

  static constexpr size_t kUnit = 256;
  static constexpr size_t kMinChunk = 64ull << 20;

  bool usable(const Span& s, hipStream_t stream) const { return s.free_any || s.owner == stream; }

  void* take(size_t n, hipStream_t stream)
  {
    for (size_t c = 0; c < chunks.size(); ++c) {
      Chunk& ch = chunks[c];
      for (auto it = ch.free.begin(); it != ch.free.end(); ++it) {
        if (it->second.len < n || !usable(it->second, stream)) continue;
        const size_t off = it->first;
        const Span s = it->second;
        ch.free.erase(it);
        if (s.len > n) ch.free.emplace(off + n, Span{s.len - n, s.owner, s.free_any});
        void* p = ch.base + off;
        live.emplace(p, std::make_pair(c, n));
        in_use += n;
        peak = std::max(peak, in_use);
        return p;
      }
    }
    return nullptr;
  }

  // Merges `off` with its neighbours where both sides are usable alike.
  static void coalesce(Chunk& ch, std::map<size_t, Span>::iterator it)
  {
    auto same = [](const Span& a, const Span& b) {
      return (a.free_any && b.free_any) || (!a.free_any && !b.free_any && a.owner == b.owner);
    };
    auto next = std::next(it);
    if (next != ch.free.end() && it->first + it->second.len == next->first && same(it->second, next->second)) {
      it->second.len += next->second.len;
      ch.free.erase(next);
    }
    if (it != ch.free.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second.len == it->first && same(prev->second, it->second)) {
        prev->second.len += it->second.len;
        ch.free.erase(it);
      }
    }
  }

  bool add_chunk(size_t sz)
  {
    void* p = nullptr;
    if (sz == 0 || hipMalloc(&p, sz) != hipSuccess) return false;
    Chunk ch;
    ch.base = static_cast<uint8_t*>(p);
    ch.size = sz;
    ch.free.emplace(0, Span{sz, nullptr, true});
    chunks.push_back(std::move(ch));
    held += sz;
    return true;
  }

  // Frees chunks that hold no block (hipFree waits for the device).
  void drop_idle()
  {
    for (size_t i = 0; i < chunks.size();) {
      Chunk& ch = chunks[i];
      if (ch.free.size() == 1 && ch.free.begin()->second.len == ch.size) {
        (void)hipFree(ch.base);
        held -= ch.size;
        chunks.erase(chunks.begin() + (std::ptrdiff_t)i);
        // chunk indices of live blocks above i shift down
        for (auto& l : live)
          if (l.second.first > i) --l.second.first;
      } else {
        ++i;
      }
    }
  }

  // Waits for every stream that returned blocks, so that any stream may take
  // them; then merges what became adjacent.
  void reclaim()
  {
    std::set<hipStream_t> owners;
    for (auto& ch : chunks)
      for (auto& f : ch.free)
        if (!f.second.free_any) owners.insert(f.second.owner);
    for (hipStream_t s : owners) {
      (void)hipStreamSynchronize(s);
      ++waits;
    }
    for (auto& ch : chunks) {
      for (auto& f : ch.free) f.second = Span{f.second.len, nullptr, true};
      for (auto it = ch.free.begin(); it != ch.free.end();) {
        auto next = std::next(it);
        if (next != ch.free.end() && it->first + it->second.len == next->first) {
          it->second.len += next->second.len;
          ch.free.erase(next);
        } else {
          ++it;
        }
      }
    }
  }

  void* allocate(size_t bytes, hipStream_t stream)
  {
    const size_t n = mgb::align_up(std::max<size_t>(bytes, 1), kUnit);
    std::lock_guard<std::mutex> lk(mu);
    if (void* p = take(n, stream)) return p;
    // a new chunk: the request, and at least half what is held (up to 4 GiB),
    // so that chunks stay few; near the limit, exactly the request
    size_t sz = std::max(mgb::align_up(n, kMinChunk), std::min<size_t>(held / 2, 4ull << 30));
    if ((int64_t)(held + sz) > limit) sz = n;
    if ((int64_t)(held + sz) <= limit && add_chunk(sz)) return take(n, stream);
    // at the limit: blocks other streams returned, then idle chunks
    reclaim();
    if (void* p = take(n, stream)) return p;
    drop_idle();
    if ((int64_t)(held + n) <= limit && add_chunk(n)) return take(n, stream);
    return nullptr;
  }

  void deallocate(void* p, hipStream_t stream)
  {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu);
    auto it = live.find(p);
    if (it == live.end()) return;
    Chunk& ch = chunks[it->second.first];
    const size_t n = it->second.second;
    in_use -= n;
    live.erase(it);
    auto ins = ch.free.emplace(static_cast<size_t>(static_cast<uint8_t*>(p) - ch.base), Span{n, stream, false}).first;
    coalesce(ch, ins);
  }

  ~mgb_pool()
  {
    for (auto& ch : chunks) (void)hipFree(ch.base);
  }
};

namespace {

void* pool_alloc(void* ctx, size_t bytes, void* stream)
{
  return static_cast<mgb_pool*>(ctx)->allocate(bytes, static_cast<hipStream_t>(stream));
}

void pool_free(void* ctx, void* p, size_t, void* stream)
{
  static_cast<mgb_pool*>(ctx)->deallocate(p, static_cast<hipStream_t>(stream));
}

}  // namespace

extern "C" mgb_status mgb_pool_create(int64_t limit_bytes, mgb_pool** out)
{
  mgb::g_last_error.clear();
  if (!out) return fail(MGB_ERROR_LOGIC, "out is NULL");
  if (limit_bytes <= 0) return fail(MGB_ERROR_LOGIC, "a pool needs a positive limit");
  mgb_pool* p = new mgb_pool();
  p->limit = limit_bytes;
  *out = p;
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_pool_allocator(mgb_pool* pool, mgb_allocator* out)
{
  mgb::g_last_error.clear();
  if (!pool || !out) return fail(MGB_ERROR_LOGIC, "NULL argument");
  *out = mgb_allocator{};
  out->allocate = pool_alloc;
  out->deallocate = pool_free;
  out->ctx = pool;
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_pool_get_stats(mgb_pool* pool, mgb_pool_stats* out)
{
  mgb::g_last_error.clear();
  if (!pool || !out) return fail(MGB_ERROR_LOGIC, "NULL argument");
  std::lock_guard<std::mutex> lk(pool->mu);
  out->limit_bytes = pool->limit;
  out->held_bytes = (int64_t)pool->held;
  out->in_use_bytes = (int64_t)pool->in_use;
  out->peak_in_use_bytes = (int64_t)pool->peak;
  out->chunks = (int64_t)pool->chunks.size();
  out->stream_waits = pool->waits;
  return MGB_SUCCESS;
}

extern "C" mgb_status mgb_pool_trim(mgb_pool* pool)
{
  mgb::g_last_error.clear();
  if (!pool) return fail(MGB_ERROR_LOGIC, "pool is NULL");
  std::lock_guard<std::mutex> lk(pool->mu);
  pool->reclaim();
  pool->drop_idle();
  return MGB_SUCCESS;
}

extern "C" void mgb_pool_destroy(mgb_pool* pool)
{
  if (!pool) return;
  {
    std::lock_guard<std::mutex> lk(pool->mu);
    pool->reclaim();  // the last users' work has completed before the chunks go
  }
  delete pool;
}
