// Device BVH builder: the reference's binned-SAH build and 4-wide collapse
// (bvh.hpp:39-52,126-247) as gfx950 kernels, producing the same nodes and
// the same primitive order as the host builder (bvh_build.cpp), word for word.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ptd {
namespace bvhg {

// Most boxes one device build takes: node slots (4 per box) stay below
// 2^32 and the scratch (about 270 bytes a box) within one GPU's memory.
constexpr uint32_t kMaxDeviceBoxes = 1u << 28;

// Device scratch a build over n boxes needs (bytes).
size_t scratch_bytes(uint32_t n);

// Flags of the boxes a build was given: with a NaN coordinate std::min and
// std::max depend on the order they are taken in, and -0.0 == +0.0 keeps
// whichever comes first, while the device's reductions pick one; either way
// the device build stops before building.
constexpr uint32_t BOX_NAN = 1u, BOX_NEG_ZERO = 2u;

// Builds the tree over n boxes. d_boxes: n boxes on the device, 6 floats
// each (min xyz, max xyz: the aabb layout, aabb.hpp:9). Outputs on the
// device: d_order (n words) and d_nodes (128-byte 4-wide nodes, room for
// max(n, 1)); *n_nodes is the number written. *box_flags: the BOX_* flags
// of the boxes; when it is not 0 nothing is built and *n_nodes is 0.
// 1 <= n <= kMaxDeviceBoxes. d_scratch: at least scratch_bytes(n) bytes of
// device memory, 256-byte aligned. Synchronises `stream` before returning.
hipError_t build(const float* d_boxes, uint32_t n, uint32_t* d_order, void* d_nodes, uint32_t* n_nodes,
                 uint32_t* box_flags, void* d_scratch, size_t scratch_size, hipStream_t stream);

}  // namespace bvhg
}  // namespace ptd
