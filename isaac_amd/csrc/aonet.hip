// Learned occlusion pass: U-Net kernels on gfx950 (forward, backward, Adam).
//
// Convolutions are implicit GEMMs on the matrix cores: the forward layers on
// v_mfma_f32_16x16x32_f16 (below, "forward convolution"), backward data and the
// weight gradient on v_mfma_f32_16x16x16_f16.  For one 16x16x16 step, lane l of
// the wave supplies
//   A (16 x 16, rows = output channels):  row l % 16, k = 4*(l / 16) .. +3
//   B (16 x 16, cols = pixels):           col l % 16, k = 4*(l / 16) .. +3
// and receives D rows 4*(l / 16) .. +3 of column l % 16.  With k = tap*Cin + ci
// and Cin a multiple of 4, the four k of a lane are four consecutive channels of
// one tap: one 8-byte load each for A (weights) and B (input patch in LDS).
#include "aonet.hpp"

#include <cfloat>

namespace isaac_hip
{
    typedef _Float16 h4 __attribute__((ext_vector_type(4)));
    typedef float f4 __attribute__((ext_vector_type(4)));

    __device__ inline f4 mfma16(h4 a, h4 b, f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }

    // ------------------------------------------------------------------ convolution
    // stage rows y0-1 .. y0+rows of the 66-pixel-wide window at x0-1 (zero outside
    // the image) into LDS as [row][66][cin]; channels < ca come upsampled from `up`
    __device__ inline void stage_patch(
        _Float16* patch,
        const _Float16* __restrict__ x,
        const _Float16* __restrict__ up,
        int ca,
        int cin,
        int n,
        int h,
        int w,
        int y0,
        int x0,
        int rows)
    {
        const int cs = cin - ca;
        const int units = rows * 66 * cin / 4;
        for(int u = threadIdx.x; u < units; u += blockDim.x)
        {
            const int e = u * 4;
            const int pix = e / cin, c = e - pix * cin;
            const int r = pix / 66, col = pix - r * 66;
            const int yy = y0 + r - 1, xx = x0 + col - 1;
            h4 v = {0, 0, 0, 0};
            if(yy >= 0 && yy < h && xx >= 0 && xx < w)
            {
                if(c < ca)
                    v = *reinterpret_cast<const h4*>(up + ((size_t(n) * (h / 2) + yy / 2) * (w / 2) + xx / 2) * ca + c);
                else
                    v = *reinterpret_cast<const h4*>(x + ((size_t(n) * h + yy) * w + xx) * cs + (c - ca));
            }
            *reinterpret_cast<h4*>(patch + e) = v;
        }
    }

    // Backward data: dx (FP32, +=) = conv(dz, flipped transposed weights).
    // Workgroup = ROWS image rows x 64 pixels; four waves split the 64-pixel row
    // segment into 16-pixel columns, and each wave computes its column in all ROWS
    // rows for all output channels, so every weight fragment it loads feeds ROWS
    // MFMAs.  The (ROWS + 2) x 66 x cin input patch is staged in LDS once.
    template<int COUT, int ROWS>
    __global__ __launch_bounds__(256) void conv3x3_bwd_kernel(const ConvArgs a)
    {
        extern __shared__ _Float16 patch[];
        const int x0 = blockIdx.x * 64, y0 = blockIdx.y * ROWS, n = blockIdx.z;
        const int cin = a.cin;
        stage_patch(patch, a.x, a.up, a.ca, cin, n, a.h, a.w_, y0, x0, ROWS + 2);
        __syncthreads();
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int pl = wave * 16 + (lane & 15);
        const int kq = 4 * (lane >> 4);
        f4 acc[ROWS][COUT / 16];
#pragma unroll
        for(int r = 0; r < ROWS; r++)
#pragma unroll
            for(int cb = 0; cb < COUT / 16; cb++)
                acc[r][cb] = f4{0.f, 0.f, 0.f, 0.f};
        const _Float16* wrow = a.w + size_t(lane & 15) * a.kp;
        for(int kc = 0; kc < a.kp; kc += 16)
        {
            const int kk = kc + kq;
            const int tap = kk / cin, ci = kk - tap * cin;
            h4 w[COUT / 16];
#pragma unroll
            for(int cb = 0; cb < COUT / 16; cb++)
                w[cb] = *reinterpret_cast<const h4*>(wrow + size_t(cb * 16) * a.kp + kk);
            const bool valid = tap < 9;
            const int ky = valid ? tap / 3 : 0, kx = valid ? tap - ky * 3 : 0;
#pragma unroll
            for(int r = 0; r < ROWS; r++)
            {
                h4 b = {0, 0, 0, 0};
                if(valid)
                    b = *reinterpret_cast<const h4*>(patch + ((r + ky) * 66 + pl + kx) * cin + ci);
#pragma unroll
                for(int cb = 0; cb < COUT / 16; cb++)
                    acc[r][cb] = mfma16(w[cb], b, acc[r][cb]);
            }
        }
        const int x = x0 + pl;
        if(x >= a.w_)
            return;
#pragma unroll
        for(int r = 0; r < ROWS; r++)
        {
            const int y = y0 + r;
            if(y >= a.h)
                break;
            const size_t p = (size_t(n) * a.h + y) * a.w_ + x;
#pragma unroll
            for(int cb = 0; cb < COUT / 16; cb++)
            {
                const int co = cb * 16 + 4 * (lane >> 4);
                float4* dst = reinterpret_cast<float4*>(reinterpret_cast<float*>(a.y) + p * COUT + co);
                float4 d = *dst;
                d.x += acc[r][cb][0];
                d.y += acc[r][cb][1];
                d.z += acc[r][cb][2];
                d.w += acc[r][cb][3];
                *dst = d;
            }
        }
    }

    template<int COUT, int ROWS>
    hipError_t conv_bwd_launch(const ConvArgs& a, hipStream_t s)
    {
        const dim3 grid(unsigned((a.w_ + 63) / 64), unsigned((a.h + ROWS - 1) / ROWS), unsigned(a.n));
        const size_t lds = size_t((ROWS + 2) * 66) * a.cin * sizeof(_Float16);
        hipLaunchKernelGGL((conv3x3_bwd_kernel<COUT, ROWS>), grid, dim3(256), lds, s, a);
        return hipGetLastError();
    }

    template<int COUT>
    hipError_t conv_bwd_rows(const ConvArgs& a, size_t max_lds, hipStream_t s)
    {
        // four rows per workgroup when the patch fits and the grid still covers the
        // chip several times over; else one
        const size_t lds4 = size_t(6 * 66) * a.cin * sizeof(_Float16);
        const int64_t groups4 = int64_t((a.w_ + 63) / 64) * ((a.h + 3) / 4) * a.n;
        if(lds4 <= max_lds && groups4 >= 1024)
            return conv_bwd_launch<COUT, 4>(a, s);
        return conv_bwd_launch<COUT, 1>(a, s);
    }

    hipError_t launch_conv3x3_bwd(const ConvArgs& a, int cout, size_t max_lds, hipStream_t s)
    {
        if(a.cin % 4 != 0 || a.ca % 4 != 0 || a.ca > a.cin || (a.ca > 0 && !a.up) || a.kp % 16 != 0 || a.kp < 9 * a.cin
           || size_t(3 * 66) * a.cin * 2 > max_lds || (a.up && (a.h % 2 || a.w_ % 2)))
            return hipErrorInvalidValue;
        if(a.n == 0 || a.h == 0 || a.w_ == 0)
            return hipSuccess;
        switch(cout)
        {
        case 16: return conv_bwd_rows<16>(a, max_lds, s);
        case 32: return conv_bwd_rows<32>(a, max_lds, s);
        case 48: return conv_bwd_rows<48>(a, max_lds, s);
        case 64: return conv_bwd_rows<64>(a, max_lds, s);
        case 96: return conv_bwd_rows<96>(a, max_lds, s);
        case 128: return conv_bwd_rows<128>(a, max_lds, s);
        default: return hipErrorInvalidValue;
        }
    }

    // ------------------------------------------------------------------ forward convolution
    // Forward layers run on v_mfma_f32_16x16x32_f16 (gfx950): lane l supplies
    //   A row l % 16 (output channel), k = 8*(l / 16) .. +7
    //   B col l % 16 (pixel),          k = 8*(l / 16) .. +7
    // and receives D rows 4*(l / 16) .. +3 of column l % 16.  Cin is a multiple of 8,
    // so a lane's eight k are eight consecutive channels of one tap: one 16-byte load
    // each.  Layer shapes are template parameters, so the tap / channel of every
    // K-chunk is a compile-time expression of the lane's group.
    typedef _Float16 h8 __attribute__((ext_vector_type(8)));

    __device__ inline f4 mfma32(h8 a, h8 b, f4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }

    // LDS pixel stride (halves) of a patch with CIN channels: an odd number of
    // 16-byte units, so 16 consecutive pixels of one 16-byte read hit 16 different
    // bank groups
    template<int CIN>
    constexpr int patch_stride()
    {
        return ((CIN / 8) % 2) ? CIN : CIN + 8;
    }

    // Patch staging in two halves, global loads into registers and LDS stores, so
    // a persistent workgroup loads tile t+1 while it computes tile t.  A thread
    // owns one 16-byte channel unit of one patch column (of RU = 66*CIN/8 per row)
    // and walks rows; when a row has fewer than 256 units, G = 256/RU thread groups
    // split the rows.
    template<int CIN, int CA, int PROWS>
    struct PatchStager
    {
        static constexpr int CS = CIN - CA, U = CIN / 8, PS = patch_stride<CIN>(), RU = 66 * U;
        static constexpr int G = RU >= 256 ? 1 : 256 / RU; // row groups
        static constexpr int CI = (RU + 255) / 256; // unit iterations per thread
        static constexpr int RPG = (PROWS + G - 1) / G; // rows per group
        int grp;
        int col[CI], ch[CI];
        bool uok[CI];
        h8 v[CI][RPG];

        __device__ inline void init()
        {
            const int t = int(threadIdx.x);
            grp = RU >= 256 ? 0 : t / RU;
            const int u0 = RU >= 256 ? t : t - grp * RU;
#pragma unroll
            for(int ci = 0; ci < CI; ci++)
            {
                const int u = u0 + ci * 256;
                uok[ci] = (RU >= 256) ? (u < RU) : (grp < G);
                col[ci] = u / U;
                ch[ci] = (u - col[ci] * U) * 8;
            }
        }

        __device__ inline void load(const _Float16* __restrict__ x, const _Float16* __restrict__ up, int n, int h, int w, int y0, int x0)
        {
#pragma unroll
            for(int ci = 0; ci < CI; ci++)
            {
                const int xx = x0 + col[ci] - 1;
                const bool xin = uok[ci] && xx >= 0 && xx < w;
                const bool fromUp = CA > 0 && ch[ci] < CA;
                const _Float16* base;
                size_t rs;
                if(fromUp)
                {
                    base = up + (size_t(n) * size_t(h / 2) * size_t(w / 2) + size_t(xx >= 0 ? xx / 2 : 0)) * CA + ch[ci];
                    rs = size_t(w / 2) * CA;
                }
                else
                {
                    base = x + (size_t(n) * size_t(h) * size_t(w) + size_t(xx >= 0 ? xx : 0)) * CS + (ch[ci] - CA);
                    rs = size_t(w) * CS;
                }
#pragma unroll
                for(int k = 0; k < RPG; k++)
                {
                    const int r = grp + k * G;
                    const int yy = y0 + r - 1;
                    v[ci][k] = h8{0, 0, 0, 0, 0, 0, 0, 0};
                    if(xin && r < PROWS && yy >= 0 && yy < h)
                        v[ci][k] = *reinterpret_cast<const h8*>(base + size_t(fromUp ? yy / 2 : yy) * rs);
                }
            }
        }

        __device__ inline void store(_Float16* patch) const
        {
#pragma unroll
            for(int ci = 0; ci < CI; ci++)
#pragma unroll
                for(int k = 0; k < RPG; k++)
                {
                    const int r = grp + k * G;
                    if(uok[ci] && r < PROWS)
                        *reinterpret_cast<h8*>(patch + (r * 66 + col[ci]) * PS + ch[ci]) = v[ci][k];
                }
        }
    };

    // LDS of a forward layer: the weights when all of them fit beside a one-row
    // patch in the 160 KiB of a CU (one row of KP + 8 halves per output channel: an
    // odd number of 16-byte units, so the 16 rows of a fragment read hit 16 bank
    // groups), then the (ROWS + 2) x 66-pixel input patch
    template<int CIN, int COUT>
    struct FwdLds
    {
        static constexpr int KP = (9 * CIN + 31) / 32 * 32;
        static constexpr int WS = KP + 8;
        static constexpr size_t patch(int rows)
        {
            return size_t((rows + 2) * 66) * patch_stride<CIN>() * sizeof(_Float16);
        }
        static constexpr size_t weights = size_t(COUT) * WS * sizeof(_Float16);
        static constexpr bool WLDS = weights + patch(1) <= size_t(160) * 1024;
        static constexpr size_t bytes(int rows)
        {
            return patch(rows) + (WLDS ? weights : 0);
        }
    };

    // Persistent forward layer: workgroup b computes tiles b, b + grid, ... of ROWS
    // rows x 64 pixels.  The four waves split the 16-channel output blocks first
    // (COUT 64: one block per wave; COUT 32: two blocks x two 32-pixel halves;
    // COUT 16: four 16-pixel columns): a wave owns one block in NPC 16-pixel columns
    // x ROWS rows, so each weight fragment feeds NPC * ROWS MFMAs.  Weights come from
    // LDS (loaded once per workgroup) or, when they do not fit, from global memory
    // PD chunks ahead of the MFMAs.
    template<int CIN, int COUT, int CA, int ROWS>
    __global__ __launch_bounds__(256) void conv_fwd_kernel(const ConvArgs a, int tilesX, int tilesY, int numTiles)
    {
        using L = FwdLds<CIN, COUT>;
        constexpr int K = 9 * CIN, KP = L::KP, NCH = KP / 32, NCB = COUT / 16;
        constexpr int PS = patch_stride<CIN>();
        constexpr int WCB = NCB < 4 ? NCB : 4; // waves along output-channel blocks
        constexpr int NPC = WCB; // 16-pixel columns per wave
        static_assert(NCB == WCB, "one output-channel block per wave");
        constexpr bool WLDS = L::WLDS;
        constexpr int PD = WLDS ? 1 : (NCH < 8 ? NCH : 8);
        extern __shared__ _Float16 lds[];
        _Float16* wl = lds;
        _Float16* patch = lds + (WLDS ? COUT * L::WS : 0);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int cb = wave % WCB, pc0 = (wave / WCB) * NPC;
        const int g = lane >> 4;
        const int tpi = tilesX * tilesY;
        auto coords = [&](int t, int& n, int& y0, int& x0) {
            n = t / tpi;
            const int r = t - n * tpi;
            const int by = r / tilesX;
            y0 = by * ROWS;
            x0 = (r - by * tilesX) * 64;
        };
        PatchStager<CIN, CA, ROWS + 2> st;
        st.init();
        int t = int(blockIdx.x), n, y0, x0;
        coords(t, n, y0, x0);
        if constexpr(WLDS)
        {
            // all weights into LDS once (requested together with the first patch)
            constexpr int WU = COUT * KP / 8, WIT = (WU + 255) / 256;
            h8 wv[WIT];
#pragma unroll
            for(int it = 0; it < WIT; it++)
            {
                const int u = int(threadIdx.x) + it * 256;
                if(WU % 256 == 0 || u < WU)
                    wv[it] = *reinterpret_cast<const h8*>(a.w + size_t(u) * 8);
            }
            st.load(a.x, a.up, n, a.h, a.w_, y0, x0);
#pragma unroll
            for(int it = 0; it < WIT; it++)
            {
                const int u = int(threadIdx.x) + it * 256;
                if(WU % 256 == 0 || u < WU)
                {
                    const int row = u / (KP / 8), k = (u - row * (KP / 8)) * 8;
                    *reinterpret_cast<h8*>(wl + row * L::WS + k) = wv[it];
                }
            }
        }
        else
            st.load(a.x, a.up, n, a.h, a.w_, y0, x0);
        const _Float16* wrow = a.w + size_t(cb * 16 + (lane & 15)) * KP + 8 * g;
        _Float16* y = reinterpret_cast<_Float16*>(a.y);
        const int co = cb * 16 + 4 * g;
        float bias[4];
#pragma unroll
        for(int e = 0; e < 4; e++)
            bias[e] = a.bias[co + e];
        for(; t < numTiles; t += int(gridDim.x))
        {
            st.store(patch);
            __syncthreads();
            const int cn = n, cy0 = y0, cx0 = x0;
            if(t + int(gridDim.x) < numTiles)
            {
                coords(t + int(gridDim.x), n, y0, x0);
                st.load(a.x, a.up, n, a.h, a.w_, y0, x0);
            }
            h8 wq[PD];
            if constexpr(!WLDS)
            {
#pragma unroll
                for(int p = 0; p < PD; p++)
                    wq[p] = *reinterpret_cast<const h8*>(wrow + 32 * p);
            }
            f4 acc[ROWS][NPC];
#pragma unroll
            for(int r = 0; r < ROWS; r++)
#pragma unroll
                for(int q = 0; q < NPC; q++)
                    acc[r][q] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for(int j = 0; j < NCH; j++)
            {
                h8 wa;
                if constexpr(WLDS)
                    wa = *reinterpret_cast<const h8*>(wl + (cb * 16 + (lane & 15)) * L::WS + 32 * j + 8 * g);
                else
                {
                    wa = wq[j % PD];
                    if(j + PD < NCH)
                        wq[j % PD] = *reinterpret_cast<const h8*>(wrow + 32 * (j + PD));
                }
                const int k0 = 32 * j + 8 * g;
                const int tap = k0 / CIN, ci = k0 - tap * CIN;
                const bool valid = k0 < K;
                const int ky = valid ? tap / 3 : 0, kx = valid ? tap - 3 * (tap / 3) : 0;
                const _Float16* pb = patch + (ky * 66 + pc0 * 16 + (lane & 15) + kx) * PS + ci;
#pragma unroll
                for(int r = 0; r < ROWS; r++)
#pragma unroll
                    for(int q = 0; q < NPC; q++)
                    {
                        h8 b = {0, 0, 0, 0, 0, 0, 0, 0};
                        if(valid)
                            b = *reinterpret_cast<const h8*>(pb + (r * 66 + q * 16) * PS);
                        acc[r][q] = mfma32(wa, b, acc[r][q]);
                    }
            }
            // epilogue per 16-pixel column q: bias + ReLU rounded to FP16 (the
            // activation), then optionally its 2x2 average pool (rows r-1, r of this
            // wave; x and x+1 in lanes l, l^1) and the 1x1 head (COUT 16: channels of
            // lane groups 1..3 gathered into group 0)
#pragma unroll
            for(int q = 0; q < NPC; q++)
            {
                const int x = cx0 + (pc0 + q) * 16 + (lane & 15);
                const bool xin = x < a.w_;
                h4 prev;
#pragma unroll
                for(int r = 0; r < ROWS; r++)
                {
                    const int yy = cy0 + r;
                    const bool in = xin && yy < a.h;
                    const size_t p = (size_t(cn) * a.h + yy) * a.w_ + x;
                    h4 o;
#pragma unroll
                    for(int e = 0; e < 4; e++)
                    {
                        const float v = acc[r][q][e] + bias[e];
                        o[e] = _Float16(v > 0.0f ? v : 0.0f);
                    }
                    if(in && y)
                        *reinterpret_cast<h4*>(y + p * COUT + co) = o;
                    if((ROWS % 2) == 0 && a.pool && (r & 1))
                    {
                        const size_t p2 = (size_t(cn) * (a.h / 2) + yy / 2) * (a.w_ / 2) + x / 2;
                        h4 m;
#pragma unroll
                        for(int e = 0; e < 4; e++)
                        {
                            const float v00 = float(prev[e]), v10 = float(o[e]);
                            const float v01 = __shfl_xor(v00, 1), v11 = __shfl_xor(v10, 1);
                            m[e] = _Float16(((v00 + v01) + (v10 + v11)) * 0.25f);
                        }
                        if(in && (lane & 1) == 0)
                            *reinterpret_cast<h4*>(a.pool + p2 * COUT + co) = m;
                    }
                    if constexpr(COUT == 16)
                    {
                        if(a.head)
                        {
                            // channels 4g'..4g'+3 of this pixel from lane l + 16 g'
                            typedef uint32_t u2 __attribute__((ext_vector_type(2)));
                            const u2 bits = __builtin_bit_cast(u2, o);
                            h4 chv[4];
                            chv[0] = o;
#pragma unroll
                            for(int gg = 1; gg < 4; gg++)
                            {
                                u2 tt;
                                tt.x = __shfl_down(bits.x, 16 * gg);
                                tt.y = __shfl_down(bits.y, 16 * gg);
                                chv[gg] = __builtin_bit_cast(h4, tt);
                            }
                            if(g == 0 && x < a.imgW && yy < a.imgH)
                            {
                                float sum = a.headW[16];
#pragma unroll
                                for(int c4 = 0; c4 < 4; c4++)
#pragma unroll
                                    for(int e = 0; e < 4; e++)
                                        sum += float(chv[c4][e]) * a.headW[c4 * 4 + e];
                                a.head[(size_t(cn) * a.imgH + yy) * a.imgW + x] = sum < 0.0f ? 0.0f : (sum > 1.0f ? 1.0f : sum);
                            }
                        }
                    }
                    prev = o;
                }
            }
            __syncthreads();
        }
    }

    // compute units of the current device (cached per device index)
    static int device_cus()
    {
        static int cus[16] = {};
        int dev = 0;
        if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16)
            return 256;
        if(cus[dev] == 0)
        {
            int v = 0;
            if(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
                v = 256;
            cus[dev] = v;
        }
        return cus[dev];
    }

    template<int CIN, int COUT, int CA, int ROWS>
    hipError_t conv_fwd_launch(const ConvArgs& a, hipStream_t s)
    {
        const int tilesX = (a.w_ + 63) / 64, tilesY = (a.h + ROWS - 1) / ROWS;
        const int64_t numTiles = int64_t(tilesX) * tilesY * a.n;
        if(numTiles > int64_t(0x7fffffff))
            return hipErrorInvalidValue;
        const size_t lds = FwdLds<CIN, COUT>::bytes(ROWS);
        static int perCu = 0; // resident workgroups per CU (same for every gfx950 device)
        if(perCu == 0)
        {
            int v = 0;
            if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, conv_fwd_kernel<CIN, COUT, CA, ROWS>, 256, lds) != hipSuccess || v <= 0)
                v = 1;
            perCu = v;
        }
        const int64_t grid = std::min<int64_t>(numTiles, int64_t(device_cus()) * perCu);
        hipLaunchKernelGGL((conv_fwd_kernel<CIN, COUT, CA, ROWS>), dim3(unsigned(grid)), dim3(256), lds, s, a, tilesX, tilesY,
                           int(numTiles));
        return hipGetLastError();
    }

    // rows per tile: the most (of 8, 4, 2, 1) whose LDS leaves room for two
    // workgroups per CU and whose tiles cover every CU twice; else the most that fit
    // with a tile per CU; else the fewest that fit (a fused pool needs an even count).
    // A wave keeps at most 8 accumulator tiles (rows x its 16-pixel columns).
    template<int CIN, int COUT, int CA>
    hipError_t conv_fwd_rows(const ConvArgs& a, size_t max_lds, hipStream_t s)
    {
        using L = FwdLds<CIN, COUT>;
        constexpr int NPC = COUT / 16 < 4 ? COUT / 16 : 4;
        constexpr auto fits = [](int rows) { return L::bytes(rows) <= size_t(160) * 1024 && rows * NPC <= 8; };
        if(a.cin != CIN || a.ca != CA || a.kp != L::KP || (a.head && COUT != 16))
            return hipErrorInvalidValue;
        const bool pool = a.pool != nullptr;
        const int64_t cus = device_cus();
        const auto tiles = [&](int rows) { return int64_t((a.w_ + 63) / 64) * ((a.h + rows - 1) / rows) * a.n; };
        const auto ok = [&](int rows, size_t lim, int64_t minTiles) {
            return L::bytes(rows) <= lim && tiles(rows) >= minTiles && (!pool || rows % 2 == 0);
        };
        for(int pass = 0; pass < 3; pass++)
        {
            const size_t lim = pass == 0 ? max_lds / 2 : max_lds;
            const int64_t minTiles = pass == 0 ? 2 * cus : (pass == 1 ? cus : 0);
            if constexpr(fits(8))
                if(pass < 2 && ok(8, lim, minTiles))
                    return conv_fwd_launch<CIN, COUT, CA, 8>(a, s);
            if constexpr(fits(4))
                if(pass < 2 && ok(4, lim, minTiles))
                    return conv_fwd_launch<CIN, COUT, CA, 4>(a, s);
            if constexpr(fits(2))
                if(ok(2, lim, minTiles) && (pass < 2 || !ok(1, lim, 0)))
                    return conv_fwd_launch<CIN, COUT, CA, 2>(a, s);
            if constexpr(fits(1))
                if(ok(1, lim, minTiles))
                    return conv_fwd_launch<CIN, COUT, CA, 1>(a, s);
        }
        return hipErrorInvalidValue;
    }

    hipError_t launch_conv_fwd(const ConvArgs& a, int cout, size_t max_lds, hipStream_t s)
    {
        if((a.ca > 0 && !a.up) || ((a.up || a.pool) && (a.h % 2 || a.w_ % 2))
           || (a.head && (!a.headW || a.imgW > a.w_ || a.imgH > a.h)))
            return hipErrorInvalidValue;
        if(a.n == 0 || a.h == 0 || a.w_ == 0)
            return hipSuccess;
        const int cin = a.cin, ca = a.ca;
        if(cin == 8 && cout == 16 && ca == 0)
            return conv_fwd_rows<8, 16, 0>(a, max_lds, s);
        if(cin == 16 && cout == 16 && ca == 0)
            return conv_fwd_rows<16, 16, 0>(a, max_lds, s);
        if(cin == 16 && cout == 32 && ca == 0)
            return conv_fwd_rows<16, 32, 0>(a, max_lds, s);
        if(cin == 32 && cout == 32 && ca == 0)
            return conv_fwd_rows<32, 32, 0>(a, max_lds, s);
        if(cin == 32 && cout == 64 && ca == 0)
            return conv_fwd_rows<32, 64, 0>(a, max_lds, s);
        if(cin == 64 && cout == 64 && ca == 0)
            return conv_fwd_rows<64, 64, 0>(a, max_lds, s);
        if(cin == 128 && cout == 64 && ca == 64)
            return conv_fwd_rows<128, 64, 64>(a, max_lds, s);
        if(cin == 96 && cout == 32 && ca == 64)
            return conv_fwd_rows<96, 32, 64>(a, max_lds, s);
        if(cin == 48 && cout == 16 && ca == 32)
            return conv_fwd_rows<48, 16, 32>(a, max_lds, s);
        return hipErrorInvalidValue;
    }

    // ------------------------------------------------------------------ weight gradient
    // GEMM dW (cout x kp) = dZ^T (cout x P) . col(X) (P x kp) over the pixels P.
    // Workgroup (tile group g, pixel split q): waves own up to 4 16x16 tiles of dW each
    // (tile t = (cb, kb), t = g*16 + wave*4 + j) and accumulate over the 64-pixel row
    // segments [q*segs, (q+1)*segs); one FP32 atomic per tile element at the end.
    __global__ __launch_bounds__(256) void conv_dw_kernel(
        const _Float16* __restrict__ x,
        const _Float16* __restrict__ up,
        int ca,
        const _Float16* __restrict__ dz,
        float* __restrict__ dw,
        int n,
        int h,
        int w,
        int cin,
        int cout,
        int kdim,
        int segsPerBlock)
    {
        extern __shared__ _Float16 lds[];
        _Float16* patch = lds; // [3][66][cin]
        _Float16* dzt = lds + 3 * 66 * cin; // [64][cout]
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int ncb = cout / 16, nkb = (kdim + 15) / 16, ntiles = ncb * nkb;
        const int segW = (w + 63) / 64;
        const int64_t nseg = int64_t(n) * h * segW;
        int tcb[4], tkb[4];
        bool own[4];
        for(int j = 0; j < 4; j++)
        {
            const int t = blockIdx.x * 16 + wave * 4 + j;
            own[j] = t < ntiles;
            tcb[j] = own[j] ? t / nkb : 0;
            tkb[j] = own[j] ? t - tcb[j] * nkb : 0;
        }
        f4 acc[4];
        for(int j = 0; j < 4; j++)
            acc[j] = f4{0.f, 0.f, 0.f, 0.f};
        const int64_t s0 = int64_t(blockIdx.y) * segsPerBlock;
        const int64_t s1 = s0 + segsPerBlock < nseg ? s0 + segsPerBlock : nseg;
        for(int64_t sg = s0; sg < s1; sg++)
        {
            const int xs = int(sg % segW) * 64;
            const int64_t row = sg / segW;
            const int y = int(row % h), img = int(row / h);
            __syncthreads();
            stage_patch(patch, x, up, ca, cin, img, h, w, y, xs, 3);
            const int dunits = 64 * cout / 4;
            for(int u = threadIdx.x; u < dunits; u += 256)
            {
                const int e = u * 4;
                const int pix = e / cout, c = e - pix * cout;
                h4 v = {0, 0, 0, 0};
                if(xs + pix < w)
                    v = *reinterpret_cast<const h4*>(dz + ((size_t(img) * h + y) * w + xs + pix) * cout + c);
                *reinterpret_cast<h4*>(dzt + e) = v;
            }
            __syncthreads();
            for(int sub = 0; sub < 4; sub++)
            {
                const int pb = sub * 16 + 4 * (lane >> 4);
                for(int j = 0; j < 4; j++)
                {
                    if(!own[j])
                        continue;
                    const int co = tcb[j] * 16 + (lane & 15);
                    h4 av;
                    for(int e = 0; e < 4; e++)
                        av[e] = dzt[(pb + e) * cout + co];
                    const int k = tkb[j] * 16 + (lane & 15);
                    const int tap = k / cin, ci = k - tap * cin;
                    h4 bv = {0, 0, 0, 0};
                    if(tap < 9)
                    {
                        const int ky = tap / 3, kx = tap - ky * 3;
                        for(int e = 0; e < 4; e++)
                            bv[e] = patch[(ky * 66 + pb + e + kx) * cin + ci];
                    }
                    acc[j] = mfma16(av, bv, acc[j]);
                }
            }
        }
        for(int j = 0; j < 4; j++)
        {
            if(!own[j])
                continue;
            const int k = tkb[j] * 16 + (lane & 15);
            if(k >= kdim)
                continue;
            for(int r = 0; r < 4; r++)
            {
                const int co = tcb[j] * 16 + 4 * (lane >> 4) + r;
                if(acc[j][r] != 0.0f)
                    atomicAdd(dw + size_t(co) * kdim + k, acc[j][r]);
            }
        }
    }

    hipError_t launch_conv3x3_dw(
        const _Float16* x,
        const _Float16* up,
        int ca,
        const _Float16* dz,
        float* dw,
        int n,
        int h,
        int w,
        int cin,
        int cout,
        hipStream_t s)
    {
        if(cin % 4 != 0 || cout % 16 != 0 || ca % 4 != 0 || ca > cin || (ca > 0 && !up) || (up && (h % 2 || w % 2)))
            return hipErrorInvalidValue;
        const int kdim = 9 * cin;
        const int64_t nseg = int64_t(n) * h * ((w + 63) / 64);
        if(nseg == 0)
            return hipSuccess;
        const int ntiles = (cout / 16) * ((kdim + 15) / 16);
        const int groups = (ntiles + 15) / 16;
        // about 2048 pixels per workgroup, at least 256 workgroups in total when possible
        int64_t segs = 32;
        while(segs > 1 && groups * ((nseg + segs - 1) / segs) < 256)
            segs /= 2;
        const dim3 grid(unsigned(groups), unsigned((nseg + segs - 1) / segs));
        const size_t lds = (size_t(3 * 66) * cin + size_t(64) * cout) * sizeof(_Float16);
        if(lds > 64 * 1024)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(conv_dw_kernel, grid, dim3(256), lds, s, x, up, ca, dz, dw, n, h, w, cin, cout, kdim, int(segs));
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ elementwise / reshaping
    __global__ __launch_bounds__(256) void bias_grad_kernel(const _Float16* __restrict__ dz, float* db, int64_t npx, int c)
    {
        // thread t: channel t % c, pixel lane t / c; grid-stride over pixels
        const int lanesPerBlock = 256 / c;
        const int ch = threadIdx.x % c, pl = threadIdx.x / c;
        float sum = 0.0f;
        if(pl < lanesPerBlock)
            for(int64_t p = int64_t(blockIdx.x) * lanesPerBlock + pl; p < npx; p += int64_t(gridDim.x) * lanesPerBlock)
                sum += float(dz[p * c + ch]);
        __shared__ float red[256];
        red[threadIdx.x] = pl < lanesPerBlock ? sum : 0.0f;
        __syncthreads();
        if(threadIdx.x < c)
        {
            float t = 0.0f;
            for(int i = 0; i < lanesPerBlock; i++)
                t += red[i * c + threadIdx.x];
            atomicAdd(db + threadIdx.x, t);
        }
    }

    hipError_t launch_bias_grad(const _Float16* dz, float* db, int64_t npx, int c, hipStream_t s)
    {
        if(c > 256 || 256 % c != 0)
            return hipErrorInvalidValue;
        if(npx == 0)
            return hipSuccess;
        const int lanes = 256 / c;
        int64_t blocks = (npx + lanes * 64 - 1) / (lanes * 64);
        if(blocks > 1024)
            blocks = 1024;
        hipLaunchKernelGGL(bias_grad_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, dz, db, npx, c);
        return hipGetLastError();
    }

    __global__ __launch_bounds__(256) void relu_bwd_kernel(const float* dy, const _Float16* y, _Float16* dz, int64_t count)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
            dz[i] = _Float16(float(y[i]) > 0.0f ? dy[i] : 0.0f);
    }

    static unsigned grid_for(int64_t count)
    {
        int64_t b = (count + 255) / 256;
        return unsigned(b > 8192 ? 8192 : (b < 1 ? 1 : b));
    }

    hipError_t launch_relu_bwd(const float* dy, const _Float16* y, _Float16* dz, int64_t count, hipStream_t s)
    {
        if(count == 0)
            return hipSuccess;
        hipLaunchKernelGGL(relu_bwd_kernel, dim3(grid_for(count)), dim3(256), 0, s, dy, y, dz, count);
        return hipGetLastError();
    }

    // dx (n, h, w, c) += dy(n, h/2, w/2, c) / 4; one thread per dx element
    __global__ __launch_bounds__(256) void pool2_bwd_kernel(const float* dy, float* dx, int n, int h, int w, int c)
    {
        const int64_t count = int64_t(n) * h * w * c;
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        {
            const int ch = int(i % c);
            const int64_t q = i / c;
            const int xx = int(q % w), yy = int((q / w) % h);
            const int64_t img = q / (int64_t(w) * h);
            dx[i] += 0.25f * dy[((img * (h / 2) + yy / 2) * (w / 2) + xx / 2) * c + ch];
        }
    }

    hipError_t launch_pool2_bwd(const float* dy, float* dx, int n, int h, int w, int c, hipStream_t s)
    {
        const int64_t count = int64_t(n) * h * w * c;
        if(count == 0)
            return hipSuccess;
        hipLaunchKernelGGL(pool2_bwd_kernel, dim3(grid_for(count)), dim3(256), 0, s, dy, dx, n, h, w, c);
        return hipGetLastError();
    }

    // da (n, h/2, w/2, ca) += sum of the 4 upsampled positions; dskip (n, h, w, cs) += dy[.., ca:]
    __global__ __launch_bounds__(256) void upcat_bwd_a_kernel(const float* dy, float* da, int n, int h, int w, int ca, int cs)
    {
        const int ct = ca + cs, ho = h / 2, wo = w / 2;
        const int64_t count = int64_t(n) * ho * wo * ca;
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        {
            const int ch = int(i % ca);
            const int64_t q = i / ca;
            const int xo = int(q % wo), yo = int((q / wo) % ho);
            const int64_t img = q / (int64_t(wo) * ho);
            const float* b = dy + ((img * h + 2 * yo) * w + 2 * xo) * ct + ch;
            da[i] += (b[0] + b[ct]) + (b[size_t(w) * ct] + b[size_t(w) * ct + ct]);
        }
    }

    __global__ __launch_bounds__(256) void upcat_bwd_skip_kernel(const float* dy, float* dskip, int n, int h, int w, int ca, int cs)
    {
        const int ct = ca + cs;
        const int64_t count = int64_t(n) * h * w * cs;
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        {
            const int ch = int(i % cs);
            const int64_t q = i / cs;
            dskip[i] += dy[q * ct + ca + ch];
        }
    }

    hipError_t launch_upcat_bwd(const float* dy, float* da, float* dskip, int n, int h, int w, int ca, int cs, hipStream_t s)
    {
        const int64_t ca_count = int64_t(n) * (h / 2) * (w / 2) * ca;
        const int64_t cs_count = int64_t(n) * h * w * cs;
        if(ca_count)
            hipLaunchKernelGGL(upcat_bwd_a_kernel, dim3(grid_for(ca_count)), dim3(256), 0, s, dy, da, n, h, w, ca, cs);
        if(cs_count)
            hipLaunchKernelGGL(upcat_bwd_skip_kernel, dim3(grid_for(cs_count)), dim3(256), 0, s, dy, dskip, n, h, w, ca, cs);
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ features and head
    __global__ __launch_bounds__(256) void features_kernel(const FeatureArgs a)
    {
        const int64_t count = int64_t(a.n) * a.h * a.w;
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        {
            const int xx = int(i % a.w), yy = int((i / a.w) % a.h);
            const int item = int(i / (int64_t(a.w) * a.h));
            const int gx = xx + (a.origin ? a.origin[2 * item] : 0), gy = yy + (a.origin ? a.origin[2 * item + 1] : 0);
            float d = FLT_MAX, nx = 0.f, ny = 0.f, nz = 0.f, t = 0.f;
            const float invScale = a.invScaleItem ? a.invScaleItem[item] : a.invScale;
            if(gx >= 0 && gx < a.imgW && gy >= 0 && gy < a.imgH)
            {
                const size_t p = size_t(item) * size_t(a.imgStride) + size_t(gy) * a.imgW + gx;
                d = a.depth[p];
                nx = a.normal[3 * p + 0];
                ny = a.normal[3 * p + 1];
                nz = a.normal[3 * p + 2];
                if(a.ao)
                    t = a.ao[p];
            }
            const bool bg = !(d < 3.0e38f);
            h4 lo, hi;
            lo[0] = _Float16(bg ? 0.0f : d * invScale);
            lo[1] = _Float16(bg ? 1.0f : 0.0f);
            lo[2] = _Float16(nx);
            lo[3] = _Float16(ny);
            hi[0] = _Float16(nz);
            hi[1] = hi[2] = hi[3] = _Float16(0.0f);
            *reinterpret_cast<h4*>(a.x + i * 8) = lo;
            *reinterpret_cast<h4*>(a.x + i * 8 + 4) = hi;
            if(a.target)
                a.target[i] = t;
        }
    }

    hipError_t launch_features(const FeatureArgs& a, hipStream_t s)
    {
        const int64_t count = int64_t(a.n) * a.h * a.w;
        if(count == 0)
            return hipSuccess;
        hipLaunchKernelGGL(features_kernel, dim3(grid_for(count)), dim3(256), 0, s, a);
        return hipGetLastError();
    }

    __device__ inline float head_dot(const _Float16* d0, const float* wh, int64_t p)
    {
        const h4* v = reinterpret_cast<const h4*>(d0 + p * 16);
        float s = wh[16];
#pragma unroll
        for(int q = 0; q < 4; q++)
        {
            const h4 u = v[q];
#pragma unroll
            for(int r = 0; r < 4; r++)
                s += float(u[r]) * wh[q * 4 + r];
        }
        return s;
    }

    __global__ __launch_bounds__(256) void head_train_kernel(
        const _Float16* d0,
        const float* wh,
        const float* target,
        float* loss,
        float* gwh,
        float* dd0,
        float* out,
        int64_t npx)
    {
        float gl = 0.0f, gw[17];
        for(int c = 0; c < 17; c++)
            gw[c] = 0.0f;
        for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < npx; p += int64_t(gridDim.x) * blockDim.x)
        {
            const float o = head_dot(d0, wh, p);
            if(out)
                out[p] = o;
            const float e = o - target[p];
            const float dout = 2.0f * e;
            gl += e * e;
            const h4* v = reinterpret_cast<const h4*>(d0 + p * 16);
            for(int q = 0; q < 4; q++)
            {
                const h4 u = v[q];
                float4 o;
                o.x = dout * wh[q * 4 + 0];
                o.y = dout * wh[q * 4 + 1];
                o.z = dout * wh[q * 4 + 2];
                o.w = dout * wh[q * 4 + 3];
                reinterpret_cast<float4*>(dd0 + p * 16)[q] = o;
                for(int r = 0; r < 4; r++)
                    gw[q * 4 + r] += dout * float(u[r]);
            }
            gw[16] += dout;
        }
        __shared__ float red[18][256];
        red[0][threadIdx.x] = gl;
        for(int c = 0; c < 17; c++)
            red[c + 1][threadIdx.x] = gw[c];
        __syncthreads();
        for(int off = 128; off > 0; off >>= 1)
        {
            if(int(threadIdx.x) < off)
                for(int c = 0; c < 18; c++)
                    red[c][threadIdx.x] += red[c][threadIdx.x + off];
            __syncthreads();
        }
        if(threadIdx.x == 0)
            atomicAdd(loss, red[0][0]);
        if(threadIdx.x >= 1 && threadIdx.x < 18)
            atomicAdd(gwh + threadIdx.x - 1, red[threadIdx.x][0]);
    }

    hipError_t launch_head_train(
        const _Float16* d0,
        const float* wh,
        const float* target,
        float* loss,
        float* gwh,
        float* dd0,
        float* out,
        int64_t npx,
        hipStream_t s)
    {
        if(npx == 0)
            return hipSuccess;
        int64_t blocks = (npx + 255) / 256;
        if(blocks > 2048)
            blocks = 2048;
        hipLaunchKernelGGL(head_train_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, d0, wh, target, loss, gwh, dd0, out, npx);
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ optimiser
    __global__ __launch_bounds__(256) void adam_kernel(
        float* w,
        const float* g,
        float* m,
        float* v,
        int64_t count,
        float lr,
        float b1,
        float b2,
        float eps,
        float c1,
        float c2,
        float gscale)
    {
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        {
            const float gi = g[i] * gscale;
            const float mi = b1 * m[i] + (1.0f - b1) * gi;
            const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
            m[i] = mi;
            v[i] = vi;
            w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        }
    }

    hipError_t launch_adam(
        float* w,
        const float* g,
        float* m,
        float* v,
        int64_t count,
        float lr,
        float b1,
        float b2,
        float eps,
        float c1,
        float c2,
        float gscale,
        hipStream_t s)
    {
        if(count == 0)
            return hipSuccess;
        hipLaunchKernelGGL(adam_kernel, dim3(grid_for(count)), dim3(256), 0, s, w, g, m, v, count, lr, b1, b2, eps, c1, c2, gscale);
        return hipGetLastError();
    }

    __global__ __launch_bounds__(256) void pack_weights_kernel(const float* master, _Float16* wf, _Float16* wb, int cin, int cout, int kp)
    {
        const int K = 9 * cin;
        const int64_t nf = int64_t(cout) * kp;
        const int64_t nb = int64_t(cin) * 9 * cout;
        for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nf + nb; i += int64_t(gridDim.x) * blockDim.x)
        {
            if(i < nf)
            {
                const int co = int(i / kp), k = int(i - int64_t(co) * kp);
                wf[i] = _Float16(k < K ? master[int64_t(co) * K + k] : 0.0f);
            }
            else
            {
                const int64_t j = i - nf; // [ci][t][co]
                const int co = int(j % cout);
                const int t = int((j / cout) % 9);
                const int ci = int(j / (int64_t(cout) * 9));
                wb[j] = _Float16(master[(int64_t(co) * 9 + (8 - t)) * cin + ci]);
            }
        }
    }

    hipError_t launch_pack_weights(const float* master, _Float16* wf, _Float16* wb, int cin, int cout, int kp, hipStream_t s)
    {
        const int64_t count = int64_t(cout) * kp + int64_t(cin) * 9 * cout;
        hipLaunchKernelGGL(pack_weights_kernel, dim3(grid_for(count)), dim3(256), 0, s, master, wf, wb, cin, cout, kp);
        return hipGetLastError();
    }
} // namespace isaac_hip
