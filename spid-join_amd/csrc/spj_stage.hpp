// spj_stage.hpp -- host <-> device copies of caller memory through a ring of
// page-locked chunks (the host I/O path: LoadColumn / SEND / RECV,
// src/query.cc:213-317, src/queryHostFunc.cc:1058-1126).
//
// Caller memory is pageable: a direct copy of it runs at about half the link
// rate, and into freshly allocated memory (a result buffer) the page faults
// of first touch serialise in the copying thread.  Here host threads copy
// between caller memory and page-locked chunks (faulting pages in parallel)
// while the DMA engine moves the previous chunk, so both sides overlap.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace spj {

// Copies n bytes of host memory to the device on stream st.  Returns after
// the last chunk's copy is enqueued; src may be reused then (every byte has
// been read), dst holds the data once st reaches that point.
hipError_t stage_h2d(void *dst_dev, const void *src_host, size_t n, hipStream_t st);

// Copies n bytes of device memory (ready in stream order on st) to host
// memory; returns when dst holds them.
hipError_t stage_d2h(void *dst_host, const void *src_dev, size_t n, hipStream_t st);

// A result buffer the caller releases with free(): aligned_alloc, and for
// large buffers 2 MiB alignment with transparent huge pages requested, so
// that first touch (the copy into it) takes one fault per 2 MiB.
void *alloc_result_buffer(size_t bytes);

}  // namespace spj
