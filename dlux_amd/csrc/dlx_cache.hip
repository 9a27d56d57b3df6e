// Device-resident cache of transfer matrices.
//
// The transfer matrices of an MFT depend only on geometry and per-instance
// (wavelength, pixel scales, shift) -- not on the field or the weights -- so
// repeated propagations (phase-retrieval iterations, a serving loop over the
// same wavelength grid) reuse them instead of regenerating N_out x N_in
// complex exponentials per call.  An entry holds the matrices of one batch of
// instances, contiguous, in complex64 or split4 form.
//
// Stream safety: an entry records an event after its generation; a user on any
// stream waits on it (hipStreamWaitEvent) before the matrices are read.  Each
// stream that used an entry keeps its own "last use" event on the entry, and
// eviction synchronises on every one of them before the block is reused, so a
// block read by work still queued on one stream is never handed to another.
// An evicted entry's block becomes a spare without the host waiting: it keeps
// the entry's events, and the stream that next takes the block waits on them
// on the device (hipStreamWaitEvent) before writing it; only a release to
// hipFree synchronises on them.
// An entry is pinned from its acquire to its release: eviction skips pinned
// entries (a call that holds two entries, such as the gradient's forward and
// adjoint matrices, can never have the first evicted to make room for the
// second), and an acquire that finds no room falls back to the workspace.
// The cache is bypassed while the stream is being captured into a graph
// (allocation and event records are not capturable); callers then generate the
// matrices into their workspace.
#include <cstring>
#include <mutex>
#include <vector>

#include "dlx_internal.h"
#include "dlx_split.h"

namespace dlx {
namespace {

// Streams that may use one entry before it is evicted: a user on a further
// stream first orders itself after the oldest user (hipStreamWaitEvent) and
// takes over its slot.
constexpr int kMaxUsers = 16;

struct MatEntry {
  std::vector<double> key;
  int dev = 0;
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t ready = nullptr;
  // last use per stream (stream handle, event recorded after its last reader)
  std::vector<std::pair<hipStream_t, hipEvent_t>> users;
  size_t oldest = 0;         // next slot to take over when users is full
  int pins = 0;              // acquires not yet released
  bool pending = false;      // raw entry whose contents are not enqueued yet
  uint64_t tick = 0;
};

struct MatCache {
  std::mutex mu;
  std::vector<MatEntry*> entries;
  // raw keys seen once (hash -> tick): a raw block is cached only on its second
  // sighting, so one-off batches (a serving loop's) do not allocate or evict
  std::vector<std::pair<uint64_t, uint64_t>> seen;
  size_t budget = (size_t)16 << 30;  // 16 GiB by default (DLUX_AMD_CACHE_MB)
  size_t used = 0;                    // bytes of live entries
  // blocks of evicted entries, kept for reuse: hipFree synchronises the whole
  // device, which would stall every stream of a serving loop on each eviction
  struct Spare {
    void* ptr;
    size_t bytes;
    int dev;
    std::vector<hipEvent_t> after;   // the evicted entry's ready and last-use events
  };
  std::vector<Spare> spare;
  size_t spare_bytes = 0;
  uint64_t tick = 0, hits = 0, misses = 0;
  bool init = false;
};

// Per-thread bypass (dlx_cache_bypass): calls from this thread generate their
// matrices into their workspace and never allocate, evict or wait.
thread_local bool t_bypass = false;

MatCache& cache() {
  static MatCache c;
  if (!c.init) {
    c.init = true;
    if (const char* v = getenv("DLUX_AMD_CACHE_MB")) c.budget = (size_t)atoll(v) << 20;
  }
  return c;
}

// Retire an entry.  As a spare (release false) its block keeps the entry's
// events, which the next taker's stream waits on; released (true), the host
// waits for its last use on every stream that read it (only that entry's
// work, not the device) and frees it.
void free_entry(MatCache& c, MatEntry* e, bool release) {
  if (release) {
    for (auto& u : e->users) (void)hipEventSynchronize(u.second);
    (void)hipEventSynchronize(e->ready);
    (void)hipFree(e->ptr);
    (void)hipEventDestroy(e->ready);
    for (auto& u : e->users) (void)hipEventDestroy(u.second);
  } else {
    MatCache::Spare b{e->ptr, e->bytes, e->dev, {e->ready}};
    for (auto& u : e->users) b.after.push_back(u.second);
    c.spare.push_back(std::move(b));
    c.spare_bytes += e->bytes;
  }
  delete e;
}

void release_spares(MatCache& c) {
  for (auto& b : c.spare) {
    for (hipEvent_t ev : b.after) (void)hipEventSynchronize(ev);
    (void)hipFree(b.ptr);
    for (hipEvent_t ev : b.after) (void)hipEventDestroy(ev);
  }
  c.spare.clear();
  c.spare_bytes = 0;
}

// Evict least-recently-used unpinned entries until `need` more bytes fit the
// budget (live entries plus spare blocks).  Returns false when the pinned
// entries alone leave no room.
bool make_room(MatCache& c, size_t need) {
  while (c.used + need > c.budget) {
    size_t lru = c.entries.size();
    for (size_t i = 0; i < c.entries.size(); ++i)
      if (c.entries[i]->pins == 0 &&
          (lru == c.entries.size() || c.entries[i]->tick < c.entries[lru]->tick))
        lru = i;
    if (lru == c.entries.size()) return false;
    MatEntry* e = c.entries[lru];
    c.used -= e->bytes;
    c.entries.erase(c.entries.begin() + lru);
    free_entry(c, e, false);
  }
  return true;
}

MatEntry* new_entry(std::vector<double>&& key, int dev, void* mem, size_t bytes,
                    uint64_t tick) {
  MatEntry* e = new MatEntry();
  e->key = std::move(key);
  e->dev = dev;
  e->ptr = mem;
  e->bytes = bytes;
  e->tick = tick;
  e->pins = 1;
  if (hipEventCreateWithFlags(&e->ready, hipEventDisableTiming) != hipSuccess) {
    delete e;
    return nullptr;
  }
  return e;
}

// A block of at least `need` bytes on device `dev` for work on stream s: the
// smallest spare there that is no more than twice as large (s then waits, on
// the device, for the work that last read it), else a new allocation (spares
// are released first when the budget would be exceeded).  Returns its size in
// *bytes; NULL on failure.
void* take_block(MatCache& c, size_t need, int dev, hipStream_t s, size_t* bytes) {
  size_t best = c.spare.size();
  for (size_t i = 0; i < c.spare.size(); ++i)
    if (c.spare[i].dev == dev && c.spare[i].bytes >= need && c.spare[i].bytes <= 2 * need &&
        (best == c.spare.size() || c.spare[i].bytes < c.spare[best].bytes))
      best = i;
  if (best < c.spare.size()) {
    MatCache::Spare b = std::move(c.spare[best]);
    c.spare.erase(c.spare.begin() + best);
    c.spare_bytes -= b.bytes;
    bool ordered = true;
    for (hipEvent_t ev : b.after)
      if (hipStreamWaitEvent(s, ev, 0) != hipSuccess) ordered = false;
    if (!ordered) {   // no device-side order: wait on the host instead
      (void)hipGetLastError();
      for (hipEvent_t ev : b.after) (void)hipEventSynchronize(ev);
    }
    for (hipEvent_t ev : b.after) (void)hipEventDestroy(ev);
    *bytes = b.bytes;
    return b.ptr;
  }
  if (c.used + c.spare_bytes + need > c.budget) release_spares(c);
  void* mem = nullptr;
  if (hipMalloc(&mem, need) != hipSuccess) {
    (void)hipGetLastError();
    release_spares(c);
    if (hipMalloc(&mem, need) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  *bytes = need;
  return mem;
}

}  // namespace

int cache_acquire(const dlx_mft_geom& g, int nb, const double* wl, const double* dx,
                  const double* pout, const double* shift, bool split, hipStream_t s,
                  void** ptr, void** handle, bool adjoint) {
  *ptr = nullptr;
  *handle = nullptr;
  MatCache& c = cache();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return DLX_OK;
  }
  if (cs != hipStreamCaptureStatusNone || c.budget == 0 || t_bypass) return DLX_OK;
  int dev = 0;
  DLX_HIP(hipGetDevice(&dev));

  std::vector<double> key;
  key.reserve(9 + 4 * (size_t)nb);
  key.insert(key.end(), {(double)dev, (double)g.n_in, (double)g.n_out, (double)g.inverse,
                         (double)g.has_focal, g.has_focal ? g.focal_length : 0.0,
                         g.has_focal ? g.focal_shift : 0.0, (split ? 1.0 : 0.0) + (adjoint ? 2.0 : 0.0),
                         (double)nb});
  key.insert(key.end(), wl, wl + nb);
  key.insert(key.end(), dx, dx + nb);
  key.insert(key.end(), pout, pout + nb);
  for (int i = 0; i < nb; ++i) key.push_back(shift ? shift[i] : 0.0);

  std::lock_guard<std::mutex> lock(c.mu);
  for (MatEntry* e : c.entries) {
    if (e->key.size() == key.size() &&
        std::memcmp(e->key.data(), key.data(), key.size() * sizeof(double)) == 0) {
      e->tick = ++c.tick;
      ++c.hits;
      DLX_HIP(hipStreamWaitEvent(s, e->ready, 0));
      ++e->pins;
      *ptr = e->ptr;
      *handle = e;
      return DLX_OK;
    }
  }
  ++c.misses;
  const size_t per = adjoint ? mat_bytes(g.n_in, g.n_out, split)
                             : mat_bytes(g.n_out, g.n_in, split);
  const size_t need = align_up(per * (size_t)nb);
  if (need > c.budget) return DLX_OK;  // larger than the whole cache: caller generates
  if (!make_room(c, need)) return DLX_OK;  // pinned entries fill it: caller generates
  size_t bytes = 0;
  void* mem = take_block(c, need, dev, s, &bytes);
  if (!mem) return DLX_OK;  // no memory for the cache: caller generates into its workspace
  MatEntry* e = new_entry(std::move(key), dev, mem, bytes, ++c.tick);
  if (!e) {
    (void)hipFree(mem);
    set_error("hipEventCreate failed in the matrix cache");
    return DLX_ERR_HIP;
  }
  if (int st = gen_dft(g, nb, wl, dx, pout, shift, nullptr, mem, split, s, adjoint)) {
    free_entry(c, e, true);
    return st;
  }
  DLX_HIP(hipEventRecord(e->ready, s));
  c.entries.push_back(e);
  c.used += bytes;
  *ptr = mem;
  *handle = e;
  return DLX_OK;
}

uint64_t hash_key(const std::vector<double>& key) {
  uint64_t h = 1469598103934665603ull;   // FNV-1a over the key bytes
  const uint8_t* p = (const uint8_t*)key.data();
  for (size_t i = 0; i < key.size() * sizeof(double); ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

int cache_acquire_raw(const std::vector<double>& key_in, size_t bytes, hipStream_t s,
                      void** ptr, void** handle, bool* fresh) {
  *ptr = nullptr;
  *handle = nullptr;
  *fresh = true;   // bypass and miss: the caller generates the contents
  MatCache& c = cache();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return DLX_OK;
  }
  if (cs != hipStreamCaptureStatusNone || c.budget == 0 || t_bypass) return DLX_OK;
  int dev = 0;
  DLX_HIP(hipGetDevice(&dev));
  std::vector<double> key;
  key.reserve(key_in.size() + 3);
  // tag raw entries so they never compare equal to an MFT batch key
  key.insert(key.end(), {-1.0, (double)dev, (double)bytes});
  key.insert(key.end(), key_in.begin(), key_in.end());

  std::lock_guard<std::mutex> lock(c.mu);
  for (MatEntry* e : c.entries) {
    if (e->key.size() == key.size() &&
        std::memcmp(e->key.data(), key.data(), key.size() * sizeof(double)) == 0) {
      if (e->pending) return DLX_OK;   // being filled by another caller: bypass
      e->tick = ++c.tick;
      ++c.hits;
      *fresh = false;
      DLX_HIP(hipStreamWaitEvent(s, e->ready, 0));
      ++e->pins;
      *ptr = e->ptr;
      *handle = e;
      return DLX_OK;
    }
  }
  ++c.misses;
  const size_t need = align_up(bytes);
  if (need > c.budget) return DLX_OK;
  const uint64_t h = hash_key(key);
  bool again = false;
  for (auto& kv : c.seen)
    if (kv.first == h) { again = true; kv.second = ++c.tick; break; }
  if (!again) {
    if (c.seen.size() >= 256) {
      size_t old = 0;
      for (size_t i = 1; i < c.seen.size(); ++i)
        if (c.seen[i].second < c.seen[old].second) old = i;
      c.seen.erase(c.seen.begin() + old);
    }
    c.seen.push_back({h, ++c.tick});
    return DLX_OK;   // first sighting: the caller generates into its workspace
  }
  if (!make_room(c, need)) return DLX_OK;
  size_t got = 0;
  void* mem = take_block(c, need, dev, s, &got);
  if (!mem) return DLX_OK;
  MatEntry* e = new_entry(std::move(key), dev, mem, got, ++c.tick);
  if (!e) {
    (void)hipFree(mem);
    set_error("hipEventCreate failed in the matrix cache");
    return DLX_ERR_HIP;
  }
  e->pending = true;
  c.entries.push_back(e);
  c.used += got;
  *ptr = mem;
  *handle = e;
  *fresh = true;
  return DLX_OK;
}

int cache_ready(void* handle, hipStream_t s) {
  if (!handle) return DLX_OK;
  MatCache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  MatEntry* e = (MatEntry*)handle;
  DLX_HIP(hipEventRecord(e->ready, s));
  e->pending = false;
  return DLX_OK;
}

int cache_release(void* handle, hipStream_t s) {
  if (!handle) return DLX_OK;
  MatCache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  MatEntry* e = (MatEntry*)handle;
  if (e->pins > 0) --e->pins;
  // this stream's last use: re-record its own event, or take a new slot (or the
  // oldest user's, after ordering this stream behind that user's last use)
  hipEvent_t ev = nullptr;
  for (auto& u : e->users)
    if (u.first == s) ev = u.second;
  if (!ev && (int)e->users.size() < kMaxUsers) {
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      // no event: keep the ordering by syncing the stream before the entry can go
      (void)hipStreamSynchronize(s);
      ev = nullptr;
    } else {
      e->users.push_back({s, ev});
    }
  } else if (!ev) {
    auto& u = e->users[e->oldest];
    e->oldest = (e->oldest + 1) % e->users.size();
    DLX_HIP(hipStreamWaitEvent(s, u.second, 0));
    u.first = s;
    ev = u.second;
  }
  if (ev) DLX_HIP(hipEventRecord(ev, s));
  if (e->pending) {
    // acquired for filling but never made ready (the caller failed before
    // enqueueing its contents): the block holds nothing valid, drop the entry
    for (size_t i = 0; i < c.entries.size(); ++i)
      if (c.entries[i] == e) {
        c.entries.erase(c.entries.begin() + i);
        c.used -= e->bytes;
        free_entry(c, e, false);
        break;
      }
  }
  return DLX_OK;
}

}  // namespace dlx

using namespace dlx;

extern "C" int dlx_cache_clear(void) {
  clear_error();
  MatCache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  // entries pinned by a call in progress (another thread) stay until released
  std::vector<MatEntry*> keep;
  for (MatEntry* e : c.entries) {
    if (e->pins > 0) {
      keep.push_back(e);
    } else {
      c.used -= e->bytes;
      free_entry(c, e, true);
    }
  }
  c.entries.swap(keep);
  release_spares(c);
  return DLX_OK;
}

extern "C" int dlx_cache_set_budget(size_t bytes) {
  clear_error();
  MatCache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  c.budget = bytes;
  make_room(c, 0);
  if (c.used + c.spare_bytes > c.budget) release_spares(c);
  return DLX_OK;
}

extern "C" int dlx_cache_bypass(int32_t on) {
  clear_error();
  t_bypass = on != 0;
  return DLX_OK;
}

extern "C" int dlx_cache_stats(uint64_t* hits, uint64_t* misses, size_t* bytes,
                               int32_t* entries) {
  clear_error();
  MatCache& c = cache();
  std::lock_guard<std::mutex> lock(c.mu);
  if (hits) *hits = c.hits;
  if (misses) *misses = c.misses;
  if (bytes) *bytes = c.used;
  if (entries) *entries = (int32_t)c.entries.size();
  return DLX_OK;
}
