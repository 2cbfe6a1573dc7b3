// The denoiser on the device: the reference's edge-avoiding a-trous filter
// steered by each pixel's noise (denoise.hpp:55-139), fed from the render's
// own color and aux outputs as the CLI's --denoise step feeds it
// (main.cpp:172-186).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ptd {
namespace dn {

// Per pixel, in HBM: the lighting (color over the floored albedo) with its
// luminance in w, ping-ponged across the passes; its variance, likewise; the
// unit first-hit normal (zero for the sky) and the albedo as given.
struct DenoiseBufs {
    float4* light[2];
    float* var[2];
    float4* normal;
    float4* albedo;
};

// rgb: 3 floats a pixel; aux: 7 (albedo, normal, mean squared luminance);
// out: 3. All device addresses.
hipError_t launch_denoise(int width, int height, int spp, const float* rgb, const float* aux, float* out,
                          const DenoiseBufs& b, hipStream_t stream);

}  // namespace dn
}  // namespace ptd
