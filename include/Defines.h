// Defines.h — type ABI of the graph library (include/Defines.h:4-17 of the reference), for
// 64-lane wavefronts.
#pragma once
#include <cstdint>

typedef uint8_t byteT;
typedef int32_t vidT;   // vertex id
typedef int64_t eidT;   // edge id / row pointer
typedef int32_t idxT;
typedef uint64_t embIdxT;

// The reference's WARP_SIZE (32) becomes the gfx950 wavefront width.
constexpr int WAVE_SIZE = 64;

// WorkPartition (include/ops/DeviceIterator.cuh:47-54): how edge ranges are split across
// groups. DYNAMIC is declared there and rejected by a static_assert (:141); here it is
// implemented: groups claim chunks of edges from a device-wide atomic cursor.
enum WorkPartition { DYNAMIC = -1, INTERLEAVED = 0, CHUNKED = 1 };
static constexpr WorkPartition interleavedPolicy = WorkPartition::INTERLEAVED;
static constexpr WorkPartition chunkedPolicy = WorkPartition::CHUNKED;
static constexpr WorkPartition dynamicPolicy = WorkPartition::DYNAMIC;
