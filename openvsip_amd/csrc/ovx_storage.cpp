// ovx_storage.cpp -- a block's data in host memory and in HBM, each copy
// valid or stale, moved lazily on access.
//
// Follows the reference's storage_manager (src/ovxx/storage/manager.hpp):
//   * sync(where) allocates the location and copies the data there if that
//     copy is stale (manager.hpp:320-360);
//   * ptr(where) for writing syncs, then invalidates every other location;
//     the const ptr(where) only syncs (manager.hpp:247-279);
//   * admit / release hand user storage between the application and the
//     block (manager.hpp:207-243): released, the host array is the
//     application's; admitted with update, its contents become the block's.
// The reference's device location (OpenCL, src/ovxx/opencl/storage.hpp) is
// commented out of the manager; this one is a HIP allocation in HBM, so a
// block's data stays on the GPU between calls and crosses the host link only
// when the host touches it.  A write-only access (no copy of contents that
// will be overwritten) matches what DDA's dda::out asks of a block
// (src/ovxx/dda.hpp, sync policy `out`).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <memory>
#include <new>

#include "../../include/ovxhip.h"
#include "ovx_plan.hpp"

using ovx::fail;
using ovx::hip_fail;

struct ovx_storage_s {
  size_t bytes = 0;
  void *host = nullptr;       // user array or pinned library allocation
  bool user = false;          // `host` belongs to the application
  bool released = false;      // user storage handed back to the application
  unsigned valid = 0;         // bit OVX_LOC_HOST / OVX_LOC_DEVICE
  ovx::DevBuf dev;
  hipEvent_t last_copy = nullptr; // recorded after every copy
  hipStream_t copy_stream = nullptr;
  bool copy_pending = false;  // last_copy recorded and not yet waited for
  uint64_t h2d = 0, d2h = 0;

  ~ovx_storage_s()
  {
    if (copy_pending) (void)hipEventSynchronize(last_copy);
    if (last_copy) (void)hipEventDestroy(last_copy);
    if (host && !user) (void)hipHostFree(host);
  }
};

namespace {

unsigned bit(int where) { return 1u << where; }

// The host must not touch the host array while a copy out of it may still
// be reading it, nor read it before a copy into it has landed.
int wait_copies(ovx_storage s)
{
  if (!s->copy_pending) return OVX_OK;
  hipError_t e = hipEventSynchronize(s->last_copy);
  if (e != hipSuccess) return hip_fail(e, "storage copy wait");
  s->copy_pending = false;
  return OVX_OK;
}

int record_copy(ovx_storage s, hipStream_t stream)
{
  hipError_t e = hipSuccess;
  if (!s->last_copy) e = hipEventCreateWithFlags(&s->last_copy, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(s->last_copy, stream);
  if (e != hipSuccess) return hip_fail(e, "storage copy event");
  s->copy_stream = stream;
  s->copy_pending = true;
  return OVX_OK;
}

int ensure_host(ovx_storage s)
{
  if (s->host || !s->bytes) return OVX_OK;
  hipError_t e = hipHostMalloc(&s->host, s->bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    s->host = nullptr;
    return e == hipErrorOutOfMemory ? fail(OVX_ERR_NOMEM, "pinned host allocation of %zu bytes", s->bytes)
                                    : hip_fail(e, "pinned host allocation");
  }
  return OVX_OK;
}

// storage_manager::sync(where): allocate, then copy from the valid location.
int sync(ovx_storage s, int where, hipStream_t stream)
{
  if (where == OVX_LOC_HOST) {
    int st = ensure_host(s);
    if (st) return st;
    if (!(s->valid & bit(OVX_LOC_HOST))) {
      hipError_t e = hipMemcpyAsync(s->host, s->dev.p, s->bytes, hipMemcpyDeviceToHost, stream);
      if (e != hipSuccess) return hip_fail(e, "storage D2H copy");
      s->d2h += s->bytes;
      if ((st = record_copy(s, stream))) return st;
      s->valid |= bit(OVX_LOC_HOST);
    }
    // the host pointer is about to be used by the host
    return wait_copies(s);
  }
  int st = s->dev.alloc(s->bytes);
  if (st) return st;
  if (!(s->valid & bit(OVX_LOC_DEVICE))) {
    hipError_t e = hipMemcpyAsync(s->dev.p, s->host, s->bytes, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return hip_fail(e, "storage H2D copy");
    s->h2d += s->bytes;
    if ((st = record_copy(s, stream))) return st;
    s->valid |= bit(OVX_LOC_DEVICE);
  } else if (s->copy_pending && stream != s->copy_stream) {
    // a copy queued on another stream must land before work on this one
    hipError_t e = hipStreamWaitEvent(stream, s->last_copy, 0);
    if (e != hipSuccess) return hip_fail(e, "storage stream wait");
  }
  return OVX_OK;
}

} // namespace

extern "C" int ovx_storage_create(ovx_storage *out, size_t bytes, void *user_host)
{
  if (!out) return fail(OVX_ERR_INVALID, "null storage pointer");
  *out = nullptr;
  std::unique_ptr<ovx_storage_s> s(new (std::nothrow) ovx_storage_s());
  if (!s) return fail(OVX_ERR_NOMEM, "host allocation");
  s->bytes = bytes;
  if (user_host) {
    // user storage, admitted: the application's array holds the data
    s->host = user_host;
    s->user = true;
    s->valid = bit(OVX_LOC_HOST);
  } else {
    // fresh storage: contents undefined in every location, nothing to copy
    s->valid = bit(OVX_LOC_HOST) | bit(OVX_LOC_DEVICE);
  }
  *out = s.release();
  return OVX_OK;
}

extern "C" int ovx_storage_destroy(ovx_storage s)
{
  if (!s) return fail(OVX_ERR_INVALID, "null storage");
  delete s;
  return OVX_OK;
}

extern "C" int ovx_storage_ptr(ovx_storage s, int where, int access, void **ptr, void *stream)
{
  if (!s || !ptr) return fail(OVX_ERR_INVALID, "null argument");
  *ptr = nullptr;
  if (where != OVX_LOC_HOST && where != OVX_LOC_DEVICE) return fail(OVX_ERR_INVALID, "location %d", where);
  if (access != OVX_ACCESS_READ && access != OVX_ACCESS_WRITE && access != OVX_ACCESS_READ_WRITE)
    return fail(OVX_ERR_INVALID, "access mode %d", access);
  if (s->released) return fail(OVX_ERR_STATE, "user storage is released to the application");
  if (!s->bytes) return OVX_OK;
  hipStream_t st_ = (hipStream_t)stream;
  int st;
  if (access & OVX_ACCESS_READ) {
    if ((st = sync(s, where, st_))) return st;
  } else if (where == OVX_LOC_HOST) {
    // write-only: no copy, but an earlier copy may still read the array
    if ((st = ensure_host(s)) || (st = wait_copies(s))) return st;
  } else {
    if ((st = s->dev.alloc(s->bytes))) return st;
    if (s->copy_pending && st_ != s->copy_stream) {
      hipError_t e = hipStreamWaitEvent(st_, s->last_copy, 0);
      if (e != hipSuccess) return hip_fail(e, "storage stream wait");
    }
  }
  if (access & OVX_ACCESS_WRITE) s->valid = bit(where); // every other copy is stale
  *ptr = where == OVX_LOC_HOST ? s->host : s->dev.p;
  return OVX_OK;
}

extern "C" int ovx_storage_release(ovx_storage s, void *stream)
{
  if (!s) return fail(OVX_ERR_INVALID, "null storage");
  if (!s->user) return fail(OVX_ERR_STATE, "storage has no user array to release");
  if (s->released) return OVX_OK;
  if (s->bytes) {
    int st = sync(s, OVX_LOC_HOST, (hipStream_t)stream);
    if (st) return st;
  }
  s->released = true;
  return OVX_OK;
}

extern "C" int ovx_storage_admit(ovx_storage s, int update)
{
  if (!s) return fail(OVX_ERR_INVALID, "null storage");
  if (!s->user) return fail(OVX_ERR_STATE, "storage has no user array to admit");
  if (!s->released) return OVX_OK;
  // with update the application may have changed the array: the device copy
  // is stale; without, every copy stays as it was
  if (update) s->valid = bit(OVX_LOC_HOST);
  s->released = false;
  return OVX_OK;
}

extern "C" int ovx_storage_info(ovx_storage s, unsigned *valid, uint64_t *h2d_bytes, uint64_t *d2h_bytes)
{
  if (!s) return fail(OVX_ERR_INVALID, "null storage");
  if (valid) *valid = s->valid;
  if (h2d_bytes) *h2d_bytes = s->h2d;
  if (d2h_bytes) *d2h_bytes = s->d2h;
  return OVX_OK;
}
