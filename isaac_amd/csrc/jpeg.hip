// JPEG video frames on the GPU: the frame encoding of serverSendFrame
// (lib/isaac/isaac_communicator.hpp:221-287), which hands the RGBX frame to
// libjpeg-turbo (JCS_EXT_RGBX, jpeg_set_defaults, jpeg_set_quality(q, FALSE)) and
// sends {"payload": "data:image/jpeg;base64,..."}.  The same baseline JPEG,
// byte for byte, as libjpeg-turbo's published algorithm produces it:
//
//   blocks   one thread per 8x8 block of the scan (MCUs of Y00 Y01 Y10 Y11 Cb Cr,
//            4:2:0): RGB -> YCbCr in 16-bit fixed point, the 2x2 chroma box with
//            alternating bias 1, 2, edges replicated (right edge to whole blocks,
//            bottom edge to whole MCU rows), the accurate integer DCT (13-bit
//            constants, 2 extra bits in the row pass), quantisation of DCT/8 by
//            rounding division, zigzag order; blocks past the image are dummies
//            (zero AC, the DC of the last real block before them in the MCU)
//   sizes    one thread per block: its Huffman-coded length in bits (DC difference
//            to the previous block of its component, AC run lengths)
//   scan     exclusive prefix sum of the lengths (bit offsets)
//   emit     one thread per block: its codes OR-ed into a zeroed big-endian word
//            stream at its offset (words shared with a neighbour: atomicOr)
//   stuff    1-bit padding to a byte, then 0x00 after every 0xFF byte (per-chunk
//            counts, their prefix sum, an order-keeping copy)
//
// The host adds the markers (SOI, JFIF APP0, DQT per table, SOF0 — SOF1 when a
// table needs 16-bit entries —, DHT per table, SOS, EOI).  Tables: T.81 Annex K.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <vector>

#include "isaac_hip.h"

namespace isaac_hip
{
#define JPEG_TABLE(type, name, n) static const type name[n]
#include "jpeg_tables.inc"
#undef JPEG_TABLE

    // Huffman code (value, length) of every DC category and AC symbol, both tables
    struct JpegHuff
    {
        uint16_t dcCode[2][16];
        uint8_t dcLen[2][16];
        uint16_t acCode[2][256];
        uint8_t acLen[2][256];
    };

    struct JpegGeom
    {
        int32_t W, H;
        int32_t mcux, mcuy; // MCUs across and down
        int32_t ybw, ybh; // Y size in blocks
        int32_t cbw, cbh; // chroma size in blocks
        int32_t Hg; // rows after completing the last 2-row conversion group
        int64_t nblocks; // 6 per MCU
        uint32_t q[2][64]; // quantisation divisors (natural order), 8 * table entry
        uint8_t zz[64]; // natural index of the k-th zigzag coefficient
    };

    __device__ inline int32_t clampi(int32_t v, int32_t lo, int32_t hi)
    {
        return v < lo ? lo : (v > hi ? hi : v);
    }

    // RGB of frame pixel (x, y), both clamped into the image (edge replication)
    __device__ inline void jpeg_rgb(const uint32_t* px, const JpegGeom& G, int x, int y, int& r, int& g, int& b)
    {
        const uint32_t p = px[int64_t(clampi(y, 0, G.H - 1)) * G.W + clampi(x, 0, G.W - 1)];
        r = int(p & 0xffu);
        g = int((p >> 8) & 0xffu);
        b = int((p >> 16) & 0xffu);
    }

    // jccolor.c fixed point: FIX(x) = round(x * 2^16)
    constexpr int32_t jfix(double x)
    {
        return int32_t(x * 65536.0 + 0.5);
    }

    __device__ inline int32_t ycc_y(int r, int g, int b)
    {
        return (jfix(0.29900) * r + jfix(0.58700) * g + jfix(0.11400) * b + (1 << 15)) >> 16;
    }
    __device__ inline int32_t ycc_cb(int r, int g, int b)
    {
        return (-jfix(0.16874) * r - jfix(0.33126) * g + jfix(0.5) * b + (128 << 16) + (1 << 15) - 1) >> 16;
    }
    __device__ inline int32_t ycc_cr(int r, int g, int b)
    {
        return (jfix(0.5) * r - jfix(0.41869) * g - jfix(0.08131) * b + (128 << 16) + (1 << 15) - 1) >> 16;
    }

    // sample (sx, sy) of component ci (0 Y, 1 Cb, 2 Cr) as the coefficient controller sees it
    __device__ inline int32_t jpeg_sample(const uint32_t* px, const JpegGeom& G, int ci, int sx, int sy)
    {
        int r, g, b;
        if(ci == 0)
        {
            jpeg_rgb(px, G, sx, sy, r, g, b);
            return ycc_y(r, g, b);
        }
        // downsampled rows past the last conversion group repeat its last downsampled row
        const int cy = min(sy, G.Hg / 2 - 1);
        int32_t s = 0;
#pragma unroll
        for(int dy = 0; dy < 2; dy++)
#pragma unroll
            for(int dx = 0; dx < 2; dx++)
            {
                jpeg_rgb(px, G, 2 * sx + dx, 2 * cy + dy, r, g, b);
                s += ci == 1 ? ycc_cb(r, g, b) : ycc_cr(r, g, b);
            }
        return (s + 1 + (sx & 1)) >> 2;
    }

    // jfdctint.c: LL&M integer DCT, constants round(x * 2^13)
    constexpr int32_t dfix(double x)
    {
        return int32_t(x * 8192.0 + 0.5);
    }

    template<bool FIRST>
    __device__ inline void fdct_pass(int32_t* d, int stride)
    {
        constexpr int CB = 13, P1 = 2, SH = FIRST ? CB - P1 : CB + P1;
        auto descale = [](int32_t x, int n) { return (x + (1 << (n - 1))) >> n; };
        const int32_t t0 = d[0] + d[7 * stride], t7 = d[0] - d[7 * stride];
        const int32_t t1 = d[stride] + d[6 * stride], t6 = d[stride] - d[6 * stride];
        const int32_t t2 = d[2 * stride] + d[5 * stride], t5 = d[2 * stride] - d[5 * stride];
        const int32_t t3 = d[3 * stride] + d[4 * stride], t4 = d[3 * stride] - d[4 * stride];
        const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
        int32_t o[8];
        if(FIRST)
        {
            o[0] = (t10 + t11) * (1 << P1);
            o[4] = (t10 - t11) * (1 << P1);
        }
        else
        {
            o[0] = descale(t10 + t11, P1);
            o[4] = descale(t10 - t11, P1);
        }
        const int32_t z1 = (t12 + t13) * dfix(0.541196100);
        o[2] = descale(z1 + t13 * dfix(0.765366865), SH);
        o[6] = descale(z1 + t12 * (-dfix(1.847759065)), SH);
        int32_t y1 = t4 + t7, y2 = t5 + t6, y3 = t4 + t6, y4 = t5 + t7;
        const int32_t z5 = (y3 + y4) * dfix(1.175875602);
        const int32_t u4 = t4 * dfix(0.298631336), u5 = t5 * dfix(2.053119869), u6 = t6 * dfix(3.072711026), u7 = t7 * dfix(1.501321110);
        y1 = y1 * (-dfix(0.899976223));
        y2 = y2 * (-dfix(2.562915447));
        y3 = y3 * (-dfix(1.961570560)) + z5;
        y4 = y4 * (-dfix(0.390180644)) + z5;
        o[7] = descale(u4 + y1 + y3, SH);
        o[5] = descale(u5 + y2 + y4, SH);
        o[3] = descale(u6 + y2 + y3, SH);
        o[1] = descale(u7 + y1 + y4, SH);
#pragma unroll
        for(int k = 0; k < 8; k++)
            d[k * stride] = o[k];
    }

    // block index -> (MCU, position): component, block coordinates in that component
    __device__ inline void jpeg_block_at(const JpegGeom& G, int64_t bi, int& ci, int& bx, int& by, int& pos)
    {
        const int64_t m = bi / 6;
        pos = int(bi - m * 6);
        const int mx = int(m % G.mcux), my = int(m / G.mcux);
        if(pos < 4)
        {
            ci = 0;
            bx = 2 * mx + (pos & 1);
            by = 2 * my + (pos >> 1);
        }
        else
        {
            ci = pos - 3;
            bx = mx;
            by = my;
        }
    }

    // quantised zigzag coefficients of a real block
    __device__ inline void jpeg_real_block(const uint32_t* px, const JpegGeom& G, int ci, int bx, int by, int16_t* out)
    {
        int32_t d[64];
        for(int v = 0; v < 8; v++)
#pragma unroll
            for(int u = 0; u < 8; u++)
                d[v * 8 + u] = jpeg_sample(px, G, ci, bx * 8 + u, by * 8 + v) - 128;
#pragma unroll
        for(int v = 0; v < 8; v++)
            fdct_pass<true>(d + v * 8, 1);
#pragma unroll
        for(int u = 0; u < 8; u++)
            fdct_pass<false>(d + u, 8);
        const uint32_t* q = G.q[ci == 0 ? 0 : 1];
#pragma unroll
        for(int k = 0; k < 64; k++)
        {
            const int n = G.zz[k];
            const int32_t qv = int32_t(q[n]), c = d[n];
            const int32_t a = c < 0 ? -c : c, r = (a + (qv >> 1)) / qv;
            out[k] = int16_t(c < 0 ? -r : r);
        }
    }

    __global__ __launch_bounds__(64) void jpeg_blocks_kernel(const uint32_t* px, const JpegGeom G, int16_t* coef)
    {
        const int64_t bi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
        if(bi >= G.nblocks)
            return;
        int ci, bx, by, pos;
        jpeg_block_at(G, bi, ci, bx, by, pos);
        int16_t* out = coef + bi * 64;
        const int bw = ci == 0 ? G.ybw : G.cbw, bh = ci == 0 ? G.ybh : G.cbh;
        if(bx < bw && by < bh)
        {
            jpeg_real_block(px, G, ci, bx, by, out);
            return;
        }
        // dummy block (only Y blocks of the MCU's right column or bottom row): the DC of
        // the block before it in the MCU (jccoefct.c), i.e. of the last real one: for
        // (1,1) in a real row (1,0), else (0,1) when that column is real, else (0,0)
        const int mx0 = bx & ~1, my0 = by & ~1;
        const bool colReal = mx0 + 1 < bw, rowReal = my0 + 1 < bh;
        int sbx = mx0, sby = my0;
        if(pos == 3 && rowReal)
            sby = my0 + 1;
        else if((pos == 2 || pos == 3) && colReal)
            sbx = mx0 + 1;
        int16_t src[64];
        jpeg_real_block(px, G, 0, sbx, sby, src);
        out[0] = src[0];
        for(int k = 1; k < 64; k++)
            out[k] = 0;
    }

    __device__ inline int jpeg_nbits(int32_t v)
    {
        const uint32_t a = uint32_t(v < 0 ? -v : v);
        return a ? 32 - __clz(a) : 0;
    }

    // the previous block of the same component in scan order (-1: none)
    __device__ inline int64_t jpeg_prev(int64_t bi)
    {
        const int pos = int(bi % 6);
        if(pos > 0 && pos < 4)
            return bi - 1;
        if(pos == 0)
            return bi >= 6 ? bi - 3 : -1; // (1,1) of the previous MCU
        return bi >= 6 ? bi - 6 : -1;
    }

    template<bool EMIT>
    __device__ inline uint32_t jpeg_code_block(
        const int16_t* coef, int64_t bi, const JpegHuff& Hf, uint64_t off, uint32_t* words)
    {
        const int16_t* c = coef + bi * 64;
        const int pos = int(bi % 6), tbl = pos < 4 ? 0 : 1;
        const int64_t pb = jpeg_prev(bi);
        const int32_t diff = int32_t(c[0]) - (pb >= 0 ? int32_t(coef[pb * 64]) : 0);
        uint32_t bits = 0;
        // local word accumulator: codes OR-ed into big-endian words
        uint64_t curW = ~0ull;
        uint32_t cur = 0;
        auto put = [&](uint32_t v, int len) {
            if(!EMIT || len == 0)
            {
                bits += uint32_t(len);
                return;
            }
            const uint64_t o = off + bits;
            bits += uint32_t(len);
            v &= (len == 32 ? 0xffffffffu : ((1u << len) - 1u));
            const uint64_t w = o >> 5;
            const int s = int(o & 31);
            if(w != curW)
            {
                if(curW != ~0ull && cur)
                    atomicOr(words + curW, cur);
                curW = w;
                cur = 0;
            }
            if(s + len <= 32)
                cur |= v << (32 - s - len);
            else
            {
                cur |= v >> (s + len - 32);
                atomicOr(words + curW, cur);
                curW = w + 1;
                cur = v << (64 - s - len);
            }
        };
        const int nb = jpeg_nbits(diff);
        put(Hf.dcCode[tbl][nb], Hf.dcLen[tbl][nb]);
        put(uint32_t(diff < 0 ? diff - 1 : diff), nb);
        int run = 0;
        for(int k = 1; k < 64; k++)
        {
            const int32_t v = c[k];
            if(v == 0)
            {
                run++;
                continue;
            }
            while(run > 15)
            {
                put(Hf.acCode[tbl][0xF0], Hf.acLen[tbl][0xF0]);
                run -= 16;
            }
            const int n = jpeg_nbits(v);
            const int sym = (run << 4) | n;
            put(Hf.acCode[tbl][sym], Hf.acLen[tbl][sym]);
            put(uint32_t(v < 0 ? v - 1 : v), n);
            run = 0;
        }
        if(run)
            put(Hf.acCode[tbl][0], Hf.acLen[tbl][0]);
        if(EMIT && curW != ~0ull && cur)
            atomicOr(words + curW, cur);
        return bits;
    }

    __global__ __launch_bounds__(256) void jpeg_sizes_kernel(const int16_t* coef, int64_t n, const JpegHuff Hf, uint32_t* bits)
    {
        const int64_t bi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
        if(bi < n)
            bits[bi] = jpeg_code_block<false>(coef, bi, Hf, 0, nullptr);
    }

    __global__ __launch_bounds__(256) void jpeg_emit_kernel(
        const int16_t* coef, int64_t n, const JpegHuff Hf, const uint64_t* offs, uint32_t* words)
    {
        const int64_t bi = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
        if(bi < n)
            (void) jpeg_code_block<true>(coef, bi, Hf, offs[bi], words);
    }

    // exclusive scan of 32-bit counts into 64-bit offsets: one workgroup per 1024
    // entries, then the chunk sums in one workgroup, then the chunk bases added
    constexpr int kJScan = 1024;

    __device__ inline uint64_t block_scan_u64(uint64_t v, uint64_t* sh)
    {
        const int t = threadIdx.x;
        sh[t] = v;
        __syncthreads();
        for(int o = 1; o < kJScan; o <<= 1)
        {
            const uint64_t a = t >= o ? sh[t - o] : 0ull;
            __syncthreads();
            sh[t] += a;
            __syncthreads();
        }
        return sh[t]; // inclusive
    }

    __global__ __launch_bounds__(kJScan) void jscan_chunks_kernel(const uint32_t* in, int64_t n, uint64_t* out, uint64_t* sums)
    {
        __shared__ uint64_t sh[kJScan];
        const int64_t i = int64_t(blockIdx.x) * kJScan + threadIdx.x;
        const uint64_t v = i < n ? in[i] : 0u;
        const uint64_t s = block_scan_u64(v, sh);
        if(i < n)
            out[i] = s - v;
        if(threadIdx.x == kJScan - 1)
            sums[blockIdx.x] = s;
    }

    __global__ __launch_bounds__(kJScan) void jscan_sums_kernel(uint64_t* sums, int nchunks, uint64_t* total)
    {
        __shared__ uint64_t sh[kJScan];
        uint64_t carry = 0;
        for(int base = 0; base < nchunks; base += kJScan)
        {
            const int i = base + int(threadIdx.x);
            const uint64_t v = i < nchunks ? sums[i] : 0ull;
            const uint64_t s = block_scan_u64(v, sh);
            if(i < nchunks)
                sums[i] = carry + s - v;
            __syncthreads();
            carry += sh[kJScan - 1];
            __syncthreads();
        }
        if(threadIdx.x == 0)
            *total = carry;
    }

    __global__ __launch_bounds__(256) void jscan_add_kernel(uint64_t* out, int64_t n, const uint64_t* sums)
    {
        const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
        if(i < n)
            out[i] += sums[i / kJScan];
    }

    __global__ __launch_bounds__(256) void jfill_kernel(uint32_t* p, int64_t n)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
            p[i] = 0u;
    }

    // bits [total, round up to 8) set to 1 (jchuff's flush padding)
    __global__ void jpad_kernel(uint32_t* words, const uint64_t* total)
    {
        if(threadIdx.x != 0 || blockIdx.x != 0)
            return;
        const uint64_t t = *total, e = (t + 7) & ~7ull;
        for(uint64_t b = t; b < e; b++)
            atomicOr(words + (b >> 5), 1u << (31 - int(b & 31)));
    }

    __device__ inline uint8_t jbyte(const uint32_t* words, int64_t i)
    {
        return uint8_t((words[i >> 2] >> (24 - 8 * (i & 3))) & 0xffu);
    }

    // 0xFF bytes per chunk of kJStuff bytes of the entropy-coded data
    constexpr int kJStuff = 1024;

    __global__ __launch_bounds__(256) void jff_count_kernel(const uint32_t* words, const uint64_t* total, uint32_t* counts, int64_t nchunks)
    {
        const int64_t nbytes = int64_t((*total + 7) >> 3);
        for(int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x)
        {
            __shared__ uint32_t red[256];
            uint32_t c = 0;
            for(int64_t i = ch * kJStuff + threadIdx.x; i < min(nbytes, (ch + 1) * kJStuff); i += blockDim.x)
                c += jbyte(words, i) == 0xffu ? 1u : 0u;
            red[threadIdx.x] = c;
            __syncthreads();
            for(int o = 128; o > 0; o >>= 1)
            {
                if(int(threadIdx.x) < o)
                    red[threadIdx.x] += red[threadIdx.x + o];
                __syncthreads();
            }
            if(threadIdx.x == 0)
                counts[ch] = red[0];
            __syncthreads();
        }
    }

    // one workgroup per chunk, bytes in order: byte i goes to i + (0xFF bytes before i)
    __global__ __launch_bounds__(256) void jff_copy_kernel(const uint32_t* words, const uint64_t* total, const uint64_t* base, uint8_t* out, int64_t nchunks)
    {
        __shared__ uint32_t sh[256];
        const int64_t nbytes = int64_t((*total + 7) >> 3);
        for(int64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x)
        {
            // kJStuff / 256 = 4 bytes per thread, consecutive
            const int64_t i0 = ch * kJStuff + int64_t(threadIdx.x) * 4;
            uint8_t b[4];
            uint32_t ff = 0;
#pragma unroll
            for(int k = 0; k < 4; k++)
            {
                b[k] = i0 + k < nbytes ? jbyte(words, i0 + k) : 0u;
                ff += (i0 + k < nbytes && b[k] == 0xffu) ? 1u : 0u;
            }
            sh[threadIdx.x] = ff;
            __syncthreads();
            for(int o = 1; o < 256; o <<= 1)
            {
                const uint32_t a = int(threadIdx.x) >= o ? sh[threadIdx.x - o] : 0u;
                __syncthreads();
                sh[threadIdx.x] += a;
                __syncthreads();
            }
            uint64_t dst = uint64_t(i0) + base[ch] + (sh[threadIdx.x] - ff);
#pragma unroll
            for(int k = 0; k < 4; k++)
            {
                if(i0 + k >= nbytes)
                    break;
                out[dst++] = b[k];
                if(b[k] == 0xffu)
                    out[dst++] = 0u;
            }
            __syncthreads();
        }
    }

    void jpeg_huff_tables(JpegHuff& H)
    {
        std::memset(&H, 0, sizeof(H));
        const uint8_t* bits[4] = {kJpegBits_dc0, kJpegBits_dc1, kJpegBits_ac0, kJpegBits_ac1};
        const uint8_t* vals[4] = {kJpegVals_dc0, kJpegVals_dc1, kJpegVals_ac0, kJpegVals_ac1};
        for(int t = 0; t < 4; t++)
        {
            uint32_t code = 0;
            int k = 0;
            for(int len = 1; len <= 16; len++)
            {
                for(int i = 0; i < bits[t][len - 1]; i++, k++, code++)
                {
                    const int v = vals[t][k];
                    if(t < 2)
                    {
                        H.dcCode[t][v] = uint16_t(code);
                        H.dcLen[t][v] = uint8_t(len);
                    }
                    else
                    {
                        H.acCode[t - 2][v] = uint16_t(code);
                        H.acLen[t - 2][v] = uint8_t(len);
                    }
                }
                code <<= 1;
            }
        }
    }

    // jpeg_set_quality (jcparam.c): natural-order table entries for quality q
    void jpeg_quality_tables(int quality, int out[2][64])
    {
        const int q = std::min(std::max(quality, 1), 100);
        const int scale = q < 50 ? 5000 / q : 200 - 2 * q;
        for(int i = 0; i < 64; i++)
        {
            out[0][i] = int(std::min<long>(std::max<long>((long(kJpegQuantLuma[i]) * scale + 50) / 100, 1), 32767));
            out[1][i] = int(std::min<long>(std::max<long>((long(kJpegQuantChroma[i]) * scale + 50) / 100, 1), 32767));
        }
    }

    struct JpegLayout
    {
        int64_t nblocks, nchunksScan, maxBytes, nchunksStuff;
        size_t coef, bits, offs, sums, total, words, ffc, ffb, sums2, out, end;
    };

    // workspace carve-up (256-byte aligned pieces); maxBytes bounds the coded data
    JpegLayout jpeg_layout(int32_t width, int32_t height)
    {
        JpegLayout L{};
        const int64_t mcux = (width + 15) / 16, mcuy = (height + 15) / 16;
        L.nblocks = mcux * mcuy * 6;
        L.nchunksScan = (L.nblocks + kJScan - 1) / kJScan;
        // per block at most 11 + 16 DC bits and 63 * (16 + 16) AC bits
        L.maxBytes = L.nblocks * ((27 + 63 * 32 + 7) / 8) + 8;
        L.nchunksStuff = (L.maxBytes + kJStuff - 1) / kJStuff;
        auto a = [](size_t v) { return (v + 255) & ~size_t(255); };
        size_t o = 0;
        L.coef = o, o += a(size_t(L.nblocks) * 64 * 2);
        L.bits = o, o += a(size_t(L.nblocks) * 4);
        L.offs = o, o += a(size_t(L.nblocks) * 8);
        L.sums = o, o += a(size_t(L.nchunksScan) * 8 + 8);
        L.total = o, o += a(8);
        L.words = o, o += a(size_t(L.maxBytes + 8));
        L.ffc = o, o += a(size_t(L.nchunksStuff) * 4);
        L.ffb = o, o += a(size_t(L.nchunksStuff) * 8 + 8);
        L.sums2 = o, o += a(size_t((L.nchunksStuff + kJScan - 1) / kJScan) * 8 + 8);
        L.out = o, o += a(size_t(L.maxBytes) * 2 + 16);
        L.end = o;
        return L;
    }

    // the marker bytes before the entropy-coded data, built in place (at most
    // 20 + 2 * 133 + 19 + 4 * 21 + 348 + 14 = 751 bytes)
    struct JpegHeaders
    {
        uint8_t b[1024];
        size_t n = 0;
        void put(std::initializer_list<uint8_t> v)
        {
            for(uint8_t x : v)
                b[n++] = x;
        }
        void put(const uint8_t* p, size_t k)
        {
            std::memcpy(b + n, p, k);
            n += k;
        }
        void push_back(uint8_t x)
        {
            b[n++] = x;
        }
    };

    // the markers before the entropy-coded data (jcmarker.c order)
    void jpeg_headers(int32_t width, int32_t height, const int qt[2][64], JpegHeaders& h)
    {
        h.n = 0;
        h.put({0xFF, 0xD8, 0xFF, 0xE0, 0x00, 0x10, 'J', 'F', 'I', 'F', 0x00, 0x01, 0x01, 0x00, 0x00, 0x01, 0x00, 0x01, 0x00, 0x00});
        bool wide = false;
        for(int t = 0; t < 2; t++)
        {
            bool w16 = false;
            for(int i = 0; i < 64; i++)
                w16 = w16 || qt[t][i] > 255;
            wide = wide || w16;
            h.put({0xFF, 0xDB, 0x00, uint8_t(w16 ? 0x83 : 0x43), uint8_t((w16 ? 0x10 : 0x00) | t)});
            for(int k = 0; k < 64; k++)
            {
                const int v = qt[t][kJpegZigzag[k]];
                if(w16)
                    h.push_back(uint8_t(v >> 8));
                h.push_back(uint8_t(v & 0xff));
            }
        }
        h.put({0xFF, uint8_t(wide ? 0xC1 : 0xC0), 0x00, 0x11, 0x08, uint8_t(height >> 8), uint8_t(height & 0xff),
               uint8_t(width >> 8), uint8_t(width & 0xff), 0x03, 0x01, 0x22, 0x00, 0x02, 0x11, 0x01, 0x03, 0x11, 0x01});
        const uint8_t* bits[4] = {kJpegBits_dc0, kJpegBits_ac0, kJpegBits_dc1, kJpegBits_ac1};
        const uint8_t* vals[4] = {kJpegVals_dc0, kJpegVals_ac0, kJpegVals_dc1, kJpegVals_ac1};
        const uint8_t cls[4] = {0x00, 0x10, 0x01, 0x11};
        for(int t = 0; t < 4; t++)
        {
            int n = 0;
            for(int i = 0; i < 16; i++)
                n += bits[t][i];
            const int len = 3 + 16 + n;
            h.put({0xFF, 0xC4, uint8_t(len >> 8), uint8_t(len & 0xff), cls[t]});
            h.put(bits[t], 16);
            h.put(vals[t], size_t(n));
        }
        h.put({0xFF, 0xDA, 0x00, 0x0C, 0x03, 0x01, 0x00, 0x02, 0x11, 0x03, 0x11, 0x00, 0x3F, 0x00});
    }

    // the encoder over a workspace laid out by jpeg_layout (L), arguments checked
    int encode(const uint8_t* rgbx, int32_t width, int32_t height, int32_t quality, uint8_t* ws, const JpegLayout& L,
               uint8_t* jpeg, int64_t capacity, int64_t* length, hipStream_t s)
    {
        int16_t* coef = reinterpret_cast<int16_t*>(ws + L.coef);
        uint32_t* bits = reinterpret_cast<uint32_t*>(ws + L.bits);
        uint64_t* offs = reinterpret_cast<uint64_t*>(ws + L.offs);
        uint64_t* sums = reinterpret_cast<uint64_t*>(ws + L.sums);
        uint64_t* total = reinterpret_cast<uint64_t*>(ws + L.total);
        uint32_t* words = reinterpret_cast<uint32_t*>(ws + L.words);
        uint32_t* ffc = reinterpret_cast<uint32_t*>(ws + L.ffc);
        uint64_t* ffb = reinterpret_cast<uint64_t*>(ws + L.ffb);
        uint64_t* sums2 = reinterpret_cast<uint64_t*>(ws + L.sums2);
        const int64_t n2 = (L.nchunksStuff + kJScan - 1) / kJScan;
        uint8_t* out = ws + L.out;

        int qt[2][64];
        jpeg_quality_tables(quality, qt);
        JpegGeom G{};
        G.W = width;
        G.H = height;
        G.mcux = (width + 15) / 16;
        G.mcuy = (height + 15) / 16;
        G.ybw = (width + 7) / 8;
        G.ybh = (height + 7) / 8;
        G.cbw = ((width + 1) / 2 + 7) / 8;
        G.cbh = ((height + 1) / 2 + 7) / 8;
        G.Hg = height + (height & 1);
        G.nblocks = L.nblocks;
        for(int t = 0; t < 2; t++)
            for(int i = 0; i < 64; i++)
                G.q[t][i] = uint32_t(qt[t][i]) * 8u;
        for(int k = 0; k < 64; k++)
            G.zz[k] = kJpegZigzag[k];
        JpegHuff Hf;
        jpeg_huff_tables(Hf);

        const unsigned nb64 = unsigned((L.nblocks + 63) / 64), nb256 = unsigned((L.nblocks + 255) / 256);
        const int64_t nwords = (L.maxBytes + 8) / 4;
        hipLaunchKernelGGL(jfill_kernel, dim3(unsigned(std::min<int64_t>((nwords + 255) / 256, 4096))), dim3(256), 0, s, words, nwords);
        hipLaunchKernelGGL(jpeg_blocks_kernel, dim3(nb64), dim3(64), 0, s, reinterpret_cast<const uint32_t*>(rgbx), G, coef);
        hipLaunchKernelGGL(jpeg_sizes_kernel, dim3(nb256), dim3(256), 0, s, coef, L.nblocks, Hf, bits);
        hipLaunchKernelGGL(jscan_chunks_kernel, dim3(unsigned(L.nchunksScan)), dim3(kJScan), 0, s, bits, L.nblocks, offs, sums);
        hipLaunchKernelGGL(jscan_sums_kernel, dim3(1), dim3(kJScan), 0, s, sums, int(L.nchunksScan), total);
        hipLaunchKernelGGL(jscan_add_kernel, dim3(nb256), dim3(256), 0, s, offs, L.nblocks, sums);
        hipLaunchKernelGGL(jpeg_emit_kernel, dim3(nb256), dim3(256), 0, s, coef, L.nblocks, Hf, offs, words);
        hipLaunchKernelGGL(jpad_kernel, dim3(1), dim3(64), 0, s, words, total);
        const unsigned gch = unsigned(std::min<int64_t>(L.nchunksStuff, 4096));
        hipLaunchKernelGGL(jff_count_kernel, dim3(gch), dim3(256), 0, s, words, total, ffc, L.nchunksStuff);
        hipLaunchKernelGGL(jscan_chunks_kernel, dim3(unsigned(n2)), dim3(kJScan), 0, s, ffc, L.nchunksStuff, ffb, sums2);
        hipLaunchKernelGGL(jscan_sums_kernel, dim3(1), dim3(kJScan), 0, s, sums2, int(n2), sums2 + n2);
        hipLaunchKernelGGL(jscan_add_kernel, dim3(unsigned((L.nchunksStuff + 255) / 256)), dim3(256), 0, s, ffb, L.nchunksStuff, sums2);
        hipLaunchKernelGGL(jff_copy_kernel, dim3(gch), dim3(256), 0, s, words, total, ffb, out, L.nchunksStuff);
        if(hipGetLastError() != hipSuccess)
            return ISAAC_HIP_ERR_DEVICE;
        // sizes: coded bits, 0xFF count
        uint64_t tbits = 0, nff = 0;
        if(hipMemcpyAsync(&tbits, total, 8, hipMemcpyDeviceToHost, s) != hipSuccess
           || hipMemcpyAsync(&nff, sums2 + n2, 8, hipMemcpyDeviceToHost, s) != hipSuccess
           || hipStreamSynchronize(s) != hipSuccess)
            return ISAAC_HIP_ERR_DEVICE;
        const int64_t coded = int64_t((tbits + 7) / 8 + nff);
        JpegHeaders hdr;
        jpeg_headers(width, height, qt, hdr);
        *length = int64_t(hdr.n) + coded + 2;
        if(capacity < *length)
            return ISAAC_HIP_ERR_ARG;
        std::memcpy(jpeg, hdr.b, hdr.n);
        if(coded > 0
           && (hipMemcpyAsync(jpeg + hdr.n, out, size_t(coded), hipMemcpyDeviceToHost, s) != hipSuccess
               || hipStreamSynchronize(s) != hipSuccess))
            return ISAAC_HIP_ERR_DEVICE;
        jpeg[*length - 2] = 0xFF;
        jpeg[*length - 1] = 0xD9;
        return 0;
    }

    // The workspace of isaac_hip_jpeg_encode_noalloc: a device array that every GPU
    // holds from the library's load to the end of the process.  Host code uses it only
    // through the address hipGetSymbolAddress returns for the current device (the
    // array's name in host code is a stand-in, not a device address).
    static_assert(ISAAC_HIP_JPEG_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");
    static __device__ __align__(256) uint8_t scratch[ISAAC_HIP_JPEG_NOALLOC_SCRATCH_BYTES];

    constexpr int kMaxScratchDevices = 64;
    std::mutex scratchLock;
    uint8_t* scratchAddr[kMaxScratchDevices] = {};

    // the scratch's address on the current device, looked up once per device
    int scratch_on_current_device(uint8_t** out)
    {
        int dev = -1;
        if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
            return ISAAC_HIP_ERR_DEVICE;
        std::lock_guard<std::mutex> guard(scratchLock);
        if(!scratchAddr[dev])
        {
            void* p = nullptr;
            if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
                return ISAAC_HIP_ERR_DEVICE;
            scratchAddr[dev] = static_cast<uint8_t*>(p);
        }
        *out = scratchAddr[dev];
        return 0;
    }
} // namespace isaac_hip

using namespace isaac_hip;

extern "C"
{
    int64_t isaac_hip_jpeg_workspace_bytes(int32_t width, int32_t height)
    {
        if(width < 1 || height < 1 || width > 65535 || height > 65535)
            return -1;
        return int64_t(jpeg_layout(width, height).end);
    }

    int isaac_hip_jpeg_encode(
        const uint8_t* rgbx,
        int32_t width,
        int32_t height,
        int32_t quality,
        void* workspace,
        int64_t workspace_bytes,
        uint8_t* jpeg,
        int64_t capacity,
        int64_t* length,
        void* stream)
    {
        if(!rgbx || !workspace || !jpeg || !length || width < 1 || height < 1 || width > 65535 || height > 65535)
            return ISAAC_HIP_ERR_ARG;
        const JpegLayout L = jpeg_layout(width, height);
        if(workspace_bytes < int64_t(L.end) || (reinterpret_cast<uintptr_t>(rgbx) & 3u) || (reinterpret_cast<uintptr_t>(workspace) & 255u))
            return ISAAC_HIP_ERR_ARG;
        return encode(rgbx, width, height, quality, static_cast<uint8_t*>(workspace), L, jpeg, capacity, length,
                      (hipStream_t) stream);
    }

    int isaac_hip_jpeg_encode_noalloc(
        const uint8_t* rgbx,
        int32_t width,
        int32_t height,
        int32_t quality,
        uint8_t* jpeg,
        int64_t capacity,
        int64_t* length,
        void* stream)
    {
        if(!rgbx || !jpeg || !length || width < 1 || height < 1 || width > 65535 || height > 65535
           || (reinterpret_cast<uintptr_t>(rgbx) & 3u))
            return ISAAC_HIP_ERR_ARG;
        const JpegLayout L = jpeg_layout(width, height);
        if(L.end > size_t(ISAAC_HIP_JPEG_NOALLOC_SCRATCH_BYTES))
            return ISAAC_HIP_ERR_ARG;
        const hipStream_t s = (hipStream_t) stream;
        int dev = -1, sdev = -1;
        if(hipGetDevice(&dev) != hipSuccess)
            return ISAAC_HIP_ERR_DEVICE;
        if(s && (hipStreamGetDevice(s, &sdev) != hipSuccess || sdev != dev))
            return ISAAC_HIP_ERR_ARG; // the scratch of another GPU than the stream's
        uint8_t* ws = nullptr;
        const int rc = scratch_on_current_device(&ws);
        if(rc)
            return rc;
        return encode(rgbx, width, height, quality, ws, L, jpeg, capacity, length, s);
    }

    int isaac_hip_jpeg_frame_message(
        const uint8_t* rgbx,
        int32_t width,
        int32_t height,
        int32_t quality,
        void* workspace,
        int64_t workspace_bytes,
        char* msg,
        int64_t capacity,
        int64_t* length,
        void* stream)
    {
        if(!msg || !length || width < 1 || height < 1)
            return ISAAC_HIP_ERR_ARG;
        const JpegLayout L = jpeg_layout(width, height);
        const size_t cap = size_t(L.maxBytes) * 2 + 4096;
        std::unique_ptr<uint8_t[]> jpg(new uint8_t[cap]); // no zero fill: the encoder writes what is read
        int64_t n = 0;
        int rc = isaac_hip_jpeg_encode(rgbx, width, height, quality, workspace, workspace_bytes, jpg.get(), int64_t(cap), &n, stream);
        if(rc)
            return rc;
        static const char header[] = "{\"payload\": \"data:image/jpeg;base64,";
        static const char footer[] = "\"}";
        const int64_t hl = int64_t(sizeof(header)) - 1, fl = int64_t(sizeof(footer)) - 1, pl = (8 * n + 5) / 6;
        *length = hl + pl + fl;
        if(capacity < *length + 1)
            return ISAAC_HIP_ERR_ARG;
        std::memcpy(msg, header, size_t(hl));
        // base64 as boost's iterators produce it: 6 bits at a time, zero-filled tail, no '='
        static const char alphabet[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        char* o = msg + hl;
        uint32_t acc = 0;
        int nb = 0;
        for(int64_t i = 0; i < n; i++)
        {
            acc = (acc << 8) | jpg[size_t(i)];
            nb += 8;
            while(nb >= 6)
            {
                nb -= 6;
                *o++ = alphabet[(acc >> nb) & 63u];
            }
        }
        if(nb > 0)
            *o++ = alphabet[(acc << (6 - nb)) & 63u];
        std::memcpy(o, footer, size_t(fl) + 1);
        return 0;
    }
}
