// ingest.h -- host -> HBM streaming of input slices (ingest.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>

namespace parhip {

#define RC_(expr)                 \
	do {                          \
		int rc_ = (expr);         \
		if (rc_) return rc_;      \
	} while (0)

// Fill `count` slices starting at slice `first` into the pinned buffer `dst`.
using FillFn = std::function<int(void* dst, uint64_t first, uint64_t count)>;
// Enqueue device work on stream `s` that reads `count` slices (device copy at
// `src`) starting at slice `first`.
using ConsumeFn = std::function<int(const void* src, uint64_t first, uint64_t count, hipStream_t s)>;

// Stream n slices of slice_bytes through the calling thread's pinned ring;
// returns once everything is enqueued (the caller synchronises `compute`).
int stream_slices(uint64_t n, size_t slice_bytes, const FillFn& fill, const ConsumeFn& consume,
                  hipStream_t compute);

// Run fn(0) .. fn(nparts-1) on the host copy threads and the caller.
void parallel_parts(size_t nparts, const std::function<void(size_t)>& fn);

// Bytes per staged chunk (parhip_set_stream_chunk, default 64 MiB).
size_t ingest_chunk_bytes();

// Free this thread's staging ring (pinned + device slots).
void ingest_release();

}  // namespace parhip
