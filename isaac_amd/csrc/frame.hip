// The frame message of the reference's video stream, built from a frame in
// device memory: serverSendFrame without libjpeg (ISAAC_JPEG off,
// lib/isaac/isaac_communicator.hpp:217-287) sends
//
//   {"payload": "data:image/raw-rgba;base64,<base64 of width*height*4 bytes>"}
//
// where the base64 text is boost's base64_from_binary<transform_width<.,6,8>>
// over the bytes (isaac_communicator.hpp:251-263): the standard alphabet, most
// significant bits first, ceil(8n/6) characters with the last partial 6-bit
// group zero-filled and no '=' padding.  The encoding runs on the GPU (each
// thread turns 12 bytes into 16 characters); the header and footer are added on
// the host.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "isaac_hip.h"

namespace isaac_hip
{
    __device__ inline uint32_t b64_char(uint32_t v)
    {
        // A-Z a-z 0-9 + /
        return v < 26u ? 65u + v : v < 52u ? 97u + (v - 26u) : v < 62u ? 48u + (v - 52u) : (v == 62u ? 43u : 47u);
    }

    // 3 bytes (b0 most significant) -> 4 characters packed little-endian into a word
    __device__ inline uint32_t b64_quad(uint32_t b0, uint32_t b1, uint32_t b2)
    {
        const uint32_t t = (b0 << 16) | (b1 << 8) | b2;
        return b64_char(t >> 18) | (b64_char((t >> 12) & 63u) << 8) | (b64_char((t >> 6) & 63u) << 16)
            | (b64_char(t & 63u) << 24);
    }

    // full groups of 12 bytes (input 4-byte aligned, output 16-byte aligned)
    __global__ __launch_bounds__(256) void base64_kernel(const uint32_t* in, int64_t groups, uint4* out)
    {
        for(int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < groups; g += int64_t(gridDim.x) * blockDim.x)
        {
            const uint32_t w0 = in[3 * g], w1 = in[3 * g + 1], w2 = in[3 * g + 2];
            // bytes in memory order: w0 = b0 | b1 << 8 | b2 << 16 | b3 << 24, ...
            auto byte = [](uint32_t w, int k) { return (w >> (8 * k)) & 0xffu; };
            out[g] = make_uint4(b64_quad(byte(w0, 0), byte(w0, 1), byte(w0, 2)), b64_quad(byte(w0, 3), byte(w1, 0), byte(w1, 1)),
                                b64_quad(byte(w1, 2), byte(w1, 3), byte(w2, 0)), b64_quad(byte(w2, 1), byte(w2, 2), byte(w2, 3)));
        }
    }

    // the last n % 12 bytes (one thread): whole triples, then 1 or 2 bytes -> 2 or 3 characters
    __global__ void base64_tail_kernel(const uint8_t* in, int64_t rem, char* out)
    {
        if(threadIdx.x != 0 || blockIdx.x != 0)
            return;
        int64_t i = 0, o = 0;
        for(; i + 3 <= rem; i += 3, o += 4)
        {
            const uint32_t q = b64_quad(in[i], in[i + 1], in[i + 2]);
            for(int k = 0; k < 4; k++)
                out[o + k] = char((q >> (8 * k)) & 0xffu);
        }
        if(rem - i == 1)
        {
            const uint32_t q = b64_quad(in[i], 0u, 0u);
            out[o] = char(q & 0xffu);
            out[o + 1] = char((q >> 8) & 0xffu);
        }
        else if(rem - i == 2)
        {
            const uint32_t q = b64_quad(in[i], in[i + 1], 0u);
            out[o] = char(q & 0xffu);
            out[o + 1] = char((q >> 8) & 0xffu);
            out[o + 2] = char((q >> 16) & 0xffu);
        }
    }

    int64_t base64_length(int64_t n)
    {
        return (8 * n + 5) / 6;
    }
} // namespace isaac_hip

using namespace isaac_hip;

extern "C"
{
    int isaac_hip_base64(const uint8_t* in, int64_t n, char* out, void* stream)
    {
        if(n < 0 || (n > 0 && (!in || !out)))
            return ISAAC_HIP_ERR_ARG;
        if(n == 0)
            return 0;
        if((reinterpret_cast<uintptr_t>(in) & 3u) || (reinterpret_cast<uintptr_t>(out) & 15u))
            return ISAAC_HIP_ERR_ARG;
        const hipStream_t s = (hipStream_t) stream;
        const int64_t groups = n / 12;
        if(groups > 0)
        {
            int64_t blocks = (groups + 255) / 256;
            if(blocks > 8192)
                blocks = 8192;
            hipLaunchKernelGGL(base64_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, reinterpret_cast<const uint32_t*>(in), groups,
                               reinterpret_cast<uint4*>(out));
        }
        if(n % 12)
            hipLaunchKernelGGL(base64_tail_kernel, dim3(1), dim3(64), 0, s, in + 12 * groups, n % 12, out + 16 * groups);
        return hipGetLastError() == hipSuccess ? 0 : ISAAC_HIP_ERR_DEVICE;
    }

    int isaac_hip_frame_message(
        const uint8_t* rgba,
        int32_t width,
        int32_t height,
        char* scratch,
        char* msg,
        int64_t capacity,
        int64_t* length,
        void* stream)
    {
        if(width < 0 || height < 0 || !msg || !length)
            return ISAAC_HIP_ERR_ARG;
        static const char header[] = "{\"payload\": \"data:image/raw-rgba;base64,";
        static const char footer[] = "\"}";
        const int64_t n = int64_t(width) * height * 4;
        const int64_t hl = int64_t(sizeof(header)) - 1, fl = int64_t(sizeof(footer)) - 1, pl = base64_length(n);
        *length = hl + pl + fl;
        if(capacity < *length + 1)
            return ISAAC_HIP_ERR_ARG;
        int rc = isaac_hip_base64(rgba, n, scratch, stream);
        if(rc)
            return rc;
        std::memcpy(msg, header, size_t(hl));
        if(pl > 0)
        {
            const hipStream_t s = (hipStream_t) stream;
            if(hipMemcpyAsync(msg + hl, scratch, size_t(pl), hipMemcpyDeviceToHost, s) != hipSuccess
               || hipStreamSynchronize(s) != hipSuccess)
                return ISAAC_HIP_ERR_DEVICE;
        }
        std::memcpy(msg + hl + pl, footer, size_t(fl) + 1);
        return 0;
    }
}
